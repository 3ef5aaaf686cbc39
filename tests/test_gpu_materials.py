"""Parity on a scene with every texture-map kind and the remaining Disney
lobes (render_scenes.maps_scene): bilinear albedo map, normal map with
corner tangents, roughness / metallic / emission maps, clearcoat, sheen
(+ tint), subsurface, anisotropic and specular tint -- inputs the sphere
fields and the heightfield do not exercise.

1. plain == wavefront: identical random numbers, counts and tallies; images
   to FP32 summation order.
2. GPU vs the oracle with common random numbers (the checks of
   test_gpu_render.py).
3. The GPU adjoint vs the FP32 oracle adjoint on the same paths, including
   the texels of the bilinear albedo map (1e-3 of the gradient norm).
"""
import numpy as np
import pytest

import eleven
import oracle_py as orc
from render_scenes import maps_scene

SPP, DEPTH, SEED = 8, 5, 11


@pytest.fixture(scope="module")
def maps():
    s = maps_scene()
    return s, orc.RefBVH(s.vertices())


def _render(ctx, scene, mode, flags=0):
    ctx.upload(scene)
    ctx.render(SPP, depth=DEPTH, seed=SEED, mode=mode, flags=flags)
    img, cnt, rays, verts = ctx.read()
    return img[..., :3].astype(np.float64), cnt, rays, verts


@pytest.mark.gpu
def test_maps_plain_equals_wavefront(gpu_ctx, maps):
    print("[materials] plain vs wavefront", flush=True)
    s, _ = maps
    a, ca, ra, va = _render(gpu_ctx, s, eleven.MODE_PLAIN)
    b, cb, rb, vb = _render(gpu_ctx, s, eleven.MODE_WAVEFRONT)
    assert np.array_equal(ca, cb) and ra == rb and va == vb
    np.testing.assert_allclose(a, b, rtol=1e-4, atol=1e-6)
    assert np.isfinite(a).all() and a.max() > 0


@pytest.mark.gpu
def test_maps_gpu_matches_oracle_common_random_numbers(gpu_ctx, maps):
    print("[materials] GPU vs oracle, common random numbers", flush=True)
    s, rb = maps
    g, gc, grays, gverts = _render(gpu_ctx, s, eleven.MODE_WAVEFRONT, flags=eleven.FLAG_NO_CULL)
    osum, _, ocnt, orays, overts = orc.render(s, rb, SPP, depth=DEPTH, seed=SEED, precision=0, shadow_mode=1)
    om = osum / np.maximum(ocnt, 1)[..., None]
    assert abs(grays - orays) <= 0.005 * orays and abs(gverts - overts) <= 0.005 * overts
    assert np.mean(gc == ocnt) > 0.99
    rel = np.abs(g - om) / (np.abs(om) + 1e-3)
    frac = np.mean(rel.max(-1) < 1e-3)
    print(f"  pixels within 1e-3: {frac:.4f}, mean rel diff {rel.mean():.2e}", flush=True)
    assert frac > 0.995
    assert rel.mean() < 5e-4


@pytest.mark.gpu
@pytest.mark.parametrize("flags", [0, eleven.FLAG_RECORD_PATHS], ids=["traced", "recorded"])
def test_maps_adjoint_matches_oracle_adjoint(gpu_ctx, maps, flags):
    """The replay traces its paths again (traced) or reads the vertices the
    recording wavefront render kept (recorded)."""
    print(f"[materials] adjoint vs oracle adjoint, flags {flags}", flush=True)
    s, rb = maps
    H, W = s.camera.y_res, s.camera.x_res
    gpu_ctx.upload(s)
    gpu_ctx.clear()
    gpu_ctx.render(SPP, depth=DEPTH, seed=SEED, mode=eleven.MODE_WAVEFRONT, flags=flags)
    w = np.random.default_rng(3).normal(size=(H, W, 3)).astype(np.float32)
    dmat, dtex = gpu_ctx.adjoint(w, grad_texture=0)
    if flags:
        assert gpu_ctx.stats().adjoint_rays == 0
    om, ot, _ = orc.adjoint(s, rb, w, SPP, depth=DEPTH, seed=SEED, precision=0, shadow_mode=1, grad_texture=0)
    rm = np.linalg.norm(dmat - om) / np.linalg.norm(om)
    rt = np.linalg.norm(dtex - ot) / np.linalg.norm(ot)
    print(f"  material rel err {rm:.2e}, texel rel err {rt:.2e}", flush=True)
    assert np.abs(om).max() > 0 and np.abs(ot).max() > 0
    assert rm < 1e-3 and rt < 1e-3
