"""Batched rendering (el_batch_*): many scenes through one wavefront loop.

1. Scene i of a batch draws exactly the random numbers of el_render with its
   seed and rng_scene = i, so its image equals a standalone wavefront render
   of that scene: accepted-sample counts equal, radiance within rtol 1e-4
   (FP32 contraction differences between the one-scene and many-scene
   instantiations of the shading code, as for plain vs wavefront) -- with
   one generation for the whole batch and with a small queue whose
   generations cut across scene boundaries.
2. The hipGraph-captured launch sequence, replayed several times with eager
   renders and synchronisations between the replays, reproduces the eager
   result bit for bit every time.
3. The autotuner returns a valid setting, logs its trials, and the batch
   rendered with that setting equals the default one (traversal results and
   per-pixel accumulation order do not depend on the launch shape).
"""
import numpy as np
import pytest

import eleven
from eleven import scenegen

DEPTH = 4
SPECS = [((40, 24), 3, 2, 11), ((64, 64), 5, 12, 12), ((17, 33), 2, 1, 13), ((48, 48), 4, 6, 14)]


@pytest.fixture(scope="module")
def batch_scenes():
    out = []
    for (res, spp, nsph, seed) in SPECS:
        sky = scenegen.sky_hdr(64, 32, sun_dir=(0.3, 0.7, 0.2), peak=30.0, sigma_deg=8.0)
        out.append((scenegen.sphere_field_arrays(n_spheres=nsph, res=res, seed=seed, hdri=sky), spp, seed))
    return out


def _standalone(ctx, scene, spp, seed, rng_scene):
    ctx.upload(scene)
    ctx.render(spp, depth=DEPTH, seed=seed, mode=eleven.MODE_WAVEFRONT, rng_scene=rng_scene)
    img, cnt, _, _ = ctx.read()
    return img, cnt


def _batch(scenes):
    b = eleven.Batch(0)
    for s, spp, seed in scenes:
        b.add(s, spp, seed)
    return b


@pytest.mark.gpu
@pytest.mark.parametrize("queue", [0, 2048])
def test_batch_equals_standalone_renders(gpu_ctx, batch_scenes, queue):
    print(f"[batch] batch vs standalone, queue {queue}", flush=True)
    b = _batch(batch_scenes)
    st = b.render(depth=DEPTH, queue_cap=queue)
    if queue:
        assert st.extend_launches > DEPTH   # several generations
    total_rays = 0
    for i, (s, spp, seed) in enumerate(batch_scenes):
        img, cnt = b.read(i)
        ref, rcnt = _standalone(gpu_ctx, s, spp, seed, i)
        total_rays += gpu_ctx.stats().rays
        np.testing.assert_array_equal(cnt, rcnt)
        np.testing.assert_allclose(img, ref, rtol=1e-4, atol=1e-6)
        assert cnt.max() == spp
    assert st.rays == total_rays
    b.close()


@pytest.mark.gpu
def test_batch_graph_replays_match_eager(batch_scenes):
    print("[batch] hipGraph replays vs eager", flush=True)
    b = _batch(batch_scenes)
    b.render(depth=DEPTH, queue_cap=4096)
    eager = [b.read(i) for i in range(len(batch_scenes))]
    for rep in range(3):
        b.render(depth=DEPTH, queue_cap=4096, graph=True)
        for i in range(len(batch_scenes)):
            img, cnt = b.read(i)
            np.testing.assert_array_equal(cnt, eager[i][1])
            np.testing.assert_array_equal(img, eager[i][0])
        b.render(depth=DEPTH, queue_cap=4096)   # eager work between replays
        np.testing.assert_array_equal(b.read(0)[0], eager[0][0])
    b.close()


@pytest.mark.gpu
def test_batch_autotuner(batch_scenes):
    print("[batch] autotuner", flush=True)
    b = _batch(batch_scenes)
    b.render(depth=DEPTH)
    ref = [b.read(i) for i in range(len(batch_scenes))]
    cfg = b.tune(depth=DEPTH, trial_spp=1)
    log = b.log()
    print(log, flush=True)
    assert "chosen" in log and log.count(" ms") >= 5
    assert cfg["bvh_width"] in (0, 4, 8) and cfg["block"] in (0, 64, 128, 256)
    b.render(depth=DEPTH, **cfg)
    for i in range(len(batch_scenes)):
        img, cnt = b.read(i)
        np.testing.assert_array_equal(cnt, ref[i][1])
        np.testing.assert_allclose(img, ref[i][0], rtol=1e-6, atol=1e-8)
    b.close()


@pytest.mark.gpu
def test_batch_errors():
    b = eleven.Batch(0)
    with pytest.raises(eleven.ElevenError):
        b.render(depth=DEPTH)                 # empty batch
    s = scenegen.sphere_field_arrays(n_spheres=1, res=(8, 8), seed=1)
    with pytest.raises(eleven.ElevenError):
        b.add(s, 0, 1)                        # zero spp
    b.add(s, 1, 1)
    with pytest.raises(eleven.ElevenError):
        b.render(depth=DEPTH, bvh_width=6)
    with pytest.raises(eleven.ElevenError):
        b.read(5)
    b.close()


@pytest.mark.gpu
def test_batch_culling_is_exact(batch_scenes):
    b = _batch(batch_scenes)
    b.render(depth=6)
    culled = [b.read(i) for i in range(len(batch_scenes))]
    r_cull = b.stats().rays
    b.render(depth=6, cull=False)
    r_all = b.stats().rays
    for i in range(len(batch_scenes)):
        img, cnt = b.read(i)
        np.testing.assert_array_equal(cnt, culled[i][1])
        np.testing.assert_array_equal(img, culled[i][0])
    assert r_cull <= r_all
    b.close()


@pytest.mark.gpu
def test_batch_with_texture_maps_and_lobes_equals_standalone(gpu_ctx, scene_root):
    """A batch mixing the scene with every texture-map kind and Disney lobe
    (render_scenes.maps_scene), a textured sphere field loaded from files
    (8-bit textures) and an untextured one: each scene's image equals its
    standalone render (counts exactly; radiance to the FP32 contraction
    differences of the one-scene and many-scene shading code, which the
    maps and lobes amplify on a few pixels: rtol 1e-3, mean 1e-5)."""
    from render_scenes import maps_scene
    print("[batch] textures, maps and lobes", flush=True)
    rng = np.random.default_rng(9)
    tex = (rng.uniform(0, 1, size=(32, 32, 3)) * 255).astype(np.uint8)
    d = scenegen.sphere_field(scene_root / "batch_tex", n_spheres=10, res=(40, 32), seed=4, ground_texture=tex)
    scenes = [(maps_scene(res=(48, 36)), 3, 21), (eleven.SceneData.load(d), 4, 22),
              (scenegen.sphere_field_arrays(n_spheres=5, res=(24, 24), seed=23), 2, 23)]
    b = _batch(scenes)
    b.render(depth=DEPTH)
    for i, (s, spp, seed) in enumerate(scenes):
        img, cnt = b.read(i)
        ref, rcnt = _standalone(gpu_ctx, s, spp, seed, i)
        np.testing.assert_array_equal(cnt, rcnt)
        np.testing.assert_allclose(img, ref, rtol=1e-3, atol=1e-6)
        assert np.mean(np.abs(img - ref) / (np.abs(ref) + 1e-3)) < 1e-5
        assert np.isfinite(img).all() and img[..., :3].max() > 0
    b.close()


@pytest.mark.gpu
def test_batch_overlapped_shadows_equal_serial(batch_scenes):
    """Overlapping each bounce's shadow launches with the next extension on a
    second stream (the default; here every generation is small) gives the
    serial launch order's images and counts bit for bit, eager and replayed
    from a captured graph whose fork and join edges come from the events."""
    print("[batch] overlapped vs serial shadows, eager and graph", flush=True)
    b = _batch(batch_scenes)
    b.render(depth=DEPTH, queue_cap=4096, serial_shadows=True)
    ref = [b.read(i) for i in range(len(batch_scenes))]
    for graph in (False, True, True):
        b.render(depth=DEPTH, queue_cap=4096, graph=graph)
        for i in range(len(batch_scenes)):
            img, cnt = b.read(i)
            np.testing.assert_array_equal(cnt, ref[i][1])
            np.testing.assert_array_equal(img.view(np.uint32), ref[i][0].view(np.uint32))
    b.close()


@pytest.mark.gpu
def test_batch_overlapped_shadows_with_spilled_stacks(batch_scenes):
    """A batch holding a scene whose traversal stacks spill past the 16 LDS
    entries (stack bound 28): the shadow launches beside an extension spill
    to their own area, so overlapped and serial launch orders agree bit for
    bit, eager and replayed from the captured graph."""
    from render_scenes import deep_stack_scene
    print("[batch] overlapped vs serial shadows with spilling stacks", flush=True)
    scenes = [(deep_stack_scene(), 8, 41)] + list(batch_scenes[:2])
    b = _batch(scenes)
    b.render(depth=DEPTH, serial_shadows=True)
    ref = [b.read(i) for i in range(len(scenes))]
    for graph in (False, True, True):
        b.render(depth=DEPTH, graph=graph)
        for i in range(len(scenes)):
            img, cnt = b.read(i)
            np.testing.assert_array_equal(cnt, ref[i][1])
            np.testing.assert_array_equal(img.view(np.uint32), ref[i][0].view(np.uint32))
    b.close()
