"""CPU pins of the oracle's path-replay adjoint (or_adjoint, FP64).

The adjoint differentiates the estimator with detached sampling: directions,
lobe choices, pdfs and MIS weights are those of the current parameters, the
BRDF values follow the differentiated ones.  The oracle can render exactly
that estimator (`eval_materials`: DisneyEval uses the perturbed materials,
everything else the scene's), with the same Philox random numbers, so the
adjoint must equal its central differences per realization, for every
parameter kind, to FP64 accuracy.  For albedo, specular and albedo texels the
sampling does not depend on the parameter at all, so the same holds for
central differences of the plain renderer (checked too).

Loss: sum(w * I) with fixed random weights w; I = mean over accepted samples.
"""
import ctypes as C

import numpy as np
import pytest

import eleven
import oracle_py as orc
from eleven import scenegen

SPP, DEPTH, SEED = 4, 3, 5


@pytest.fixture(scope="module")
def small(tmp_path_factory):
    rng = np.random.default_rng(7)
    tex = (rng.uniform(0.1, 0.9, size=(8, 8, 3)) * 255).astype(np.uint8)
    d = scenegen.sphere_field(tmp_path_factory.mktemp("adj") / "s", n_spheres=8, res=(16, 16), seed=3,
                              ground_texture=tex, extent=4.0)
    scene = eleven.SceneData.load(d)
    # interior metallic values: d(m^2.2)/dm and central differences are
    # ill-defined at the generator's Bernoulli 0 / 1
    mats = scene.view.materials
    for i in range(scene.view.material_count):
        mats[i].metallic = 0.2 + 0.6 * ((i * 37) % 10) / 10.0
    return scene, orc.RefBVH(scene.vertices())


def _mats_copy(scene):
    ms = scene.materials()
    arr = (eleven.el_material * len(ms))()
    for i, m in enumerate(ms):
        arr[i] = m
    return arr


def _loss(scene, rb, w, eval_materials=None):
    s, _, cnt, _, _ = orc.render(scene, rb, SPP, depth=DEPTH, seed=SEED, precision=1, shadow_mode=1, threads=4,
                                 eval_materials=eval_materials)
    img = s / np.maximum(cnt, 1)[..., None]
    return float((w * img).sum()), img


FIELDS = [("albedo", "x"), ("albedo", "y"), ("albedo", "z"), ("roughness", None), ("metallic", None),
          ("specular", None)]


def _get(arr, m, k):
    name, comp = FIELDS[k]
    v = getattr(arr[m], name)
    return v if comp is None else getattr(v, comp)


def _bump(arr, m, k, h):
    """Adds h to a material field; returns the field's new value (FP32-rounded)."""
    name, comp = FIELDS[k]
    if comp is None:
        setattr(arr[m], name, getattr(arr[m], name) + h)
    else:
        v = getattr(arr[m], name)
        setattr(v, comp, getattr(v, comp) + h)
    return _get(arr, m, k)


def test_oracle_adjoint_equals_detached_central_differences(small):
    scene, rb = small
    H, W = scene.camera.y_res, scene.camera.x_res
    w = np.random.default_rng(1).normal(size=(H, W, 3)).astype(np.float32)
    dmat, dtex, img = orc.adjoint(scene, rb, w, SPP, depth=DEPTH, seed=SEED, precision=1, shadow_mode=1,
                                  threads=4, grad_texture=0)
    L0, img0 = _loss(scene, rb, w)
    np.testing.assert_allclose(img, img0, rtol=1e-12, atol=1e-14)
    h = 1e-5
    checked = 0
    mats = scene.materials()
    for m in range(len(mats)):
        for k in range(6):
            if k < 3 and mats[m].albedo_tex >= 0:
                assert dmat[m, k] == 0.0
                continue
            ap, am = _mats_copy(scene), _mats_copy(scene)
            vp, vm = _bump(ap, m, k, h), _bump(am, m, k, -h)   # FP32 fields: the real step
            fd = (_loss(scene, rb, w, ap)[0] - _loss(scene, rb, w, am)[0]) / (vp - vm)
            assert abs(fd - dmat[m, k]) <= 1e-6 * max(1.0, abs(fd)), (m, FIELDS[k], fd, dmat[m, k])
            checked += 1
    assert checked >= 40
    assert np.abs(dmat).max() > 1e-3


def test_oracle_adjoint_albedo_texels_match_central_differences(small):
    scene, rb = small
    H, W = scene.camera.y_res, scene.camera.x_res
    w = np.random.default_rng(2).normal(size=(H, W, 3)).astype(np.float32)
    _, dtex, _ = orc.adjoint(scene, rb, w, SPP, depth=DEPTH, seed=SEED, precision=1, shadow_mode=1, threads=4,
                             grad_texture=0)
    tex = scene.texture(0)          # the loaded texels (writable view of the scene's memory)
    flat = np.argsort(-np.abs(dtex).reshape(-1))[:12]
    h = 1e-4
    for f in flat:
        y, x, c = np.unravel_index(f, dtex.shape)
        old = tex[y, x, c]
        tex[y, x, c] = old + h
        lp, _ = _loss(scene, rb, w)
        tex[y, x, c] = old - h
        lm, _ = _loss(scene, rb, w)
        tex[y, x, c] = old
        # texels are FP32: the +-h steps are rounded to the texel's ulp
        fd = (lp - lm) / float(np.float32(old + h) - np.float32(old - h))
        assert abs(fd - dtex[y, x, c]) <= 1e-5 * max(1.0, abs(fd)), ((y, x, c), fd, dtex[y, x, c])
    assert np.abs(dtex).max() > 0


def test_albedo_specular_gradient_is_the_plain_estimator_derivative(small):
    """Sampling does not depend on albedo or specular: the adjoint is the
    derivative of the renderer itself (no detaching involved)."""
    scene, rb = small
    H, W = scene.camera.y_res, scene.camera.x_res
    w = np.random.default_rng(3).normal(size=(H, W, 3)).astype(np.float32)
    dmat, _, _ = orc.adjoint(scene, rb, w, SPP, depth=DEPTH, seed=SEED, precision=1, shadow_mode=1, threads=4)
    view_mats = scene.view.materials
    h = 1e-5
    for m in (1, 2, 3):
        for k in (0, 1, 2, 5):
            v0 = _get(view_mats, m, k)
            vp = _bump(view_mats, m, k, h)
            lp, _ = _loss(scene, rb, w)
            vm = _bump(view_mats, m, k, -2 * h)
            lm, _ = _loss(scene, rb, w)
            _bump(view_mats, m, k, v0 - vm)
            assert _get(view_mats, m, k) == v0
            fd = (lp - lm) / (vp - vm)
            assert abs(fd - dmat[m, k]) <= 1e-6 * max(1.0, abs(fd)), (m, FIELDS[k], fd, dmat[m, k])
