"""Host pieces of the multi-view inverse renderer (eleven/inverse.py).

* The block texture parameterization: block_sum is the adjoint of block_up
  (<block_up(c), g> = <c, block_sum(g)> for any c, g), block_mean inverts
  block_up, and k = 1 is the identity -- so the gradient the optimizer
  gives a block value is exactly the derivative of the loss through the
  repeated texels.
* survey_views: the 8 cameras of the terrain survey look down at the
  centres of their 4 x 2 cells of a [-ex, ex] x [-ez, ez] terrain: each
  camera's optical axis (the camera model's centre ray, rotation about x
  by the pitch) meets the ground plane at the cell's aim point.
"""
import math

import numpy as np
import pytest

import eleven
from eleven import inverse


@pytest.mark.parametrize("k", [1, 2, 8, 32])
def test_block_parameterization_adjoint(k):
    rng = np.random.default_rng(k)
    c = rng.normal(size=(64 // k, 128 // k, 3))
    g = rng.normal(size=(64, 128, 3))
    up = inverse.block_up(c, k)
    assert up.shape == g.shape
    np.testing.assert_allclose((up * g).sum(), (c * inverse.block_sum(g, k)).sum(), rtol=1e-12)
    np.testing.assert_allclose(inverse.block_mean(up, k), c, rtol=1e-12)
    if k == 1:
        assert inverse.block_up(c, 1) is c and inverse.block_sum(g, 1) is g


def test_survey_views_aim_at_their_cells():
    base = eleven.make_camera(512, 512, focal_length=0.035, sensor_width=0.035)
    ex, ez = 51.2, 25.6
    cams = inverse.survey_views(base, (ex, ez), columns=4, pitch=60.0, height=22.0)
    assert len(cams) == 8
    aims = [(-ex + (2 * i + 1) * ex / 4, zc) for zc in (-0.75 * ez, 0.5 * ez) for i in range(4)]
    for cam, (xa, za) in zip(cams, aims):
        p = math.radians(cam.rotation.x)
        # centre ray (0, 0, f) rotated about x by the pitch: (0, -f sin p, f cos p)
        d = np.array([0.0, -math.sin(p), math.cos(p)])
        t = cam.position.y / -d[1]
        hit = np.array([cam.position.x, cam.position.y, cam.position.z]) + t * d
        np.testing.assert_allclose(hit, (xa, 0.0, za), atol=1e-4)
        assert cam.x_res == 512 and cam.y_res == 512


def test_param_materials_equal_with_params():
    """The reused material array, rewritten in place, holds byte for byte
    the structs with_params builds (both round the parameters to float32)."""
    import ctypes as C
    from eleven.autograd import ParamMaterials, with_params
    base = [eleven.make_material() for _ in range(5)]
    for i, m in enumerate(base):
        m.clearcoat, m.sheen, m.eta = 0.1 * i, 0.05 * i, 1.1 + 0.1 * i
    pm = ParamMaterials(base)
    rng = np.random.default_rng(3)
    for _ in range(4):
        p = rng.random((len(base), 6)) * 1.7 - 0.2
        a = pm.set(p)
        b = (eleven.el_material * len(base))(*with_params(base, p))
        assert C.string_at(C.addressof(a), C.sizeof(a)) == C.string_at(C.addressof(b), C.sizeof(b))
