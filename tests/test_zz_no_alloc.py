"""el_trace_closest_noalloc, the entry point that allocates nothing.

Its scratch (tallies and the traversal stack's spill area) is a static device
array of the library, so the call makes no device allocation.  Checked
through the C ABI against el_trace_closest on the same rays: closest triangle
index, t and barycentrics equal bit for bit, on two calls in a row, the
second with a smaller input (the shared array keeps the first call's
contents, which must not leak into the second: its outputs beyond n stay
untouched and its tallies start from zero).  Two scenes of different tree
shapes: a 64x64 sphere field and a 65,536-triangle heightfield.  The spill
area bounds the resident threads (16 MiB / the scene's spill slots per
thread), so these runs use fewer threads than el_trace_closest does.
"""
import numpy as np
import pytest

import eleven
import raygen
from eleven import scenegen


def _check(ctx, scene, n_rand, seed):
    ctx.upload(scene)
    rays = np.concatenate([raygen.random_rays(scene.vertices(), n_rand, seed=seed),
                           raygen.camera_rays(scene.camera, spp=1, seed=seed + 1),
                           raygen.tie_rays(scene.vertices(), 2048, seed=seed + 2)]).astype(np.float32)
    n = rays.shape[0]
    tri_ref, t_ref, uv_ref = ctx.trace_closest(rays)
    dr = ctx.alloc(rays.nbytes)
    dtri, dt, duv = ctx.alloc(4 * n), ctx.alloc(4 * n), ctx.alloc(8 * n)
    try:
        ctx.htod(dr, rays)
        for k, m in enumerate((n, n // 3 + 17)):
            print(f"[no-alloc] {scene.tri_count} tris, call {k + 1}: {m} of {n} rays", flush=True)
            ctx.htod(dtri, np.full(n, -7, np.int32))            # sentinels past m must survive
            ctx.htod(dt, np.full(n, -1.0, np.float32))
            ctx.trace_closest_noalloc(dr, m, dtri, dt, duv)
            tri = np.zeros(n, np.int32)
            t = np.zeros(n, np.float32)
            uv = np.zeros((n, 2), np.float32)
            ctx.dtoh(tri, dtri)
            ctx.dtoh(t, dt)
            ctx.dtoh(uv, duv)
            assert np.array_equal(tri[:m], tri_ref[:m])
            assert np.array_equal(t[:m].view(np.uint32), t_ref[:m].view(np.uint32))
            hit = tri_ref[:m] >= 0
            assert np.array_equal(uv[:m][hit].view(np.uint32), uv_ref[:m][hit].view(np.uint32))
            assert np.all(tri[m:] == -7) and np.all(t[m:] == -1.0)
            assert hit.mean() > 0.2
    finally:
        for p in (dr, dtri, dt, duv):
            ctx.free(p)


@pytest.mark.gpu
def test_noalloc_closest_hit_equals_el_trace_closest(gpu_ctx, scene_root):
    sf = eleven.SceneData.load(scenegen.sphere_field(scene_root / "noalloc_sf", res=(64, 64), seed=0))
    _check(gpu_ctx, sf, 1 << 17, 5)
    hf = eleven.SceneData.load(scenegen.heightfield(scene_root / "noalloc_hf", cells=(256, 128), extent=(25.6, 12.8),
                                                    res=(96, 54), n_mats=4, tex_size=16, sky=(32, 16)))
    _check(gpu_ctx, hf, 1 << 17, 9)


@pytest.mark.gpu
def test_noalloc_rejects_bad_arguments(gpu_ctx):
    ctx = eleven.Context(0)
    with pytest.raises(eleven.ElevenError) as e:
        ctx.trace_closest_noalloc(None, 4, None, None)
    assert e.value.code == eleven.EL_ERR_ARG
    d = ctx.alloc(64)
    try:
        with pytest.raises(eleven.ElevenError) as e:       # no scene uploaded
            ctx.trace_closest_noalloc(d, 1, d, d)
        assert e.value.code == eleven.EL_ERR_STATE
    finally:
        ctx.free(d)
        ctx.close()
