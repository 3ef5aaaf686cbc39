"""Image parity of the HIP renderers (plain megakernel and wavefront).

1. plain == wavefront: same estimator and random numbers, so per-pixel
   accepted-sample counts and the ray/vertex tallies are identical and the
   images agree to FP32 summation order (rtol 1e-4).
2. GPU vs oracle with common random numbers: the oracle replays the same
   Philox streams, the same HDRI pixel choice and the same shadow rule, so
   almost every pixel agrees to 1e-3 relative (paths may diverge where FP32
   transcendental results differ in the last bit).
3. GPU vs the FP64 reference image (tests/golden/ref_*.npz: the oracle with
   the reference's own semantics -- binarySearch HDRI sampling, the
   shading-position point-light test -- at 4096 spp).  The GPU renders 16
   independent batches, whose spread gives the per-pixel variance of its
   mean.  Passing means: measured relMSE = mean((I - I_ref)^2 / (I_ref^2 +
   1e-2)) is consistent with the relMSE the two variances predict (ratio in
   [0.75, 1.33]), no 8x8 block mean is off by more than 5 standard errors,
   and at most 0.2 % of pixel channels exceed |z| = 5.
"""
import numpy as np
import pytest

import eleven
import oracle_py as orc
from render_scenes import DEPTH, SCENES, deep_stack_scene, make_scene
from pathlib import Path

GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def scenes(scene_root):
    return {k: make_scene(k, scene_root) for k in SCENES}


@pytest.fixture(scope="module")
def refbvhs(scenes):
    return {k: orc.RefBVH(s.vertices()) for k, s in scenes.items()}


def _render(ctx, scene, spp, seed, mode, depth=5, flags=0):
    ctx.upload(scene)
    st = ctx.render(spp, depth=depth, seed=seed, mode=mode, flags=flags)
    img, cnt, rays, verts = ctx.read()
    return img[..., :3].astype(np.float64), cnt, rays, verts, st


@pytest.mark.gpu
@pytest.mark.parametrize("kind", SCENES)
def test_plain_equals_wavefront(gpu_ctx, scenes, kind):
    print(f"[render] plain vs wavefront: {kind}", flush=True)
    s = scenes[kind]
    a, ca, ra, va, _ = _render(gpu_ctx, s, 16, 3, eleven.MODE_PLAIN)
    b, cb, rb, vb, _ = _render(gpu_ctx, s, 16, 3, eleven.MODE_WAVEFRONT)
    assert np.array_equal(ca, cb), "accepted-sample counts differ"
    assert ra == rb and va == vb, f"tallies differ: rays {ra} vs {rb}, vertices {va} vs {vb}"
    np.testing.assert_allclose(a, b, rtol=1e-4, atol=1e-6)
    assert ca.min() > 0


@pytest.mark.gpu
@pytest.mark.parametrize("kind", SCENES)
def test_gpu_matches_oracle_common_random_numbers(gpu_ctx, scenes, refbvhs, kind):
    print(f"[render] GPU vs oracle, common random numbers: {kind}", flush=True)
    s = scenes[kind]
    spp, seed, depth = 8, 5, 5
    # faithful mode (every path traced to the end, as the oracle does), so the
    # ray and vertex tallies are comparable too
    g, gc, grays, gverts, _ = _render(gpu_ctx, s, spp, seed, eleven.MODE_WAVEFRONT, depth=depth,
                                      flags=eleven.FLAG_NO_CULL)
    osum, _, ocnt, orays, overts = orc.render(s, refbvhs[kind], spp, depth=depth, seed=seed, precision=0,
                                                shadow_mode=1)
    om = osum / np.maximum(ocnt, 1)[..., None]
    assert abs(grays - orays) <= 0.005 * orays and abs(gverts - overts) <= 0.005 * overts
    assert np.mean(gc == ocnt) > 0.99
    rel = np.abs(g - om) / (np.abs(om) + 1e-3)
    frac = np.mean(rel.max(-1) < 1e-3)
    print(f"  pixels within 1e-3: {frac:.4f}, mean rel diff {rel.mean():.2e}", flush=True)
    assert frac > 0.995
    assert rel.mean() < 5e-4


@pytest.mark.gpu
@pytest.mark.parametrize("kind", SCENES)
def test_gpu_vs_fp64_reference_image(gpu_ctx, scenes, kind):
    print(f"[render] GPU vs FP64 reference image (relMSE consistency): {kind}", flush=True)
    ref = np.load(GOLDEN / f"ref_{kind}.npz")
    I_ref, var_ref, n_ref = ref["mean"].astype(np.float64), ref["var"].astype(np.float64), float(ref["spp"])
    s = scenes[kind]
    gpu_ctx.upload(s)
    batches, spp = 16, 64
    means = []
    for b in range(batches):
        gpu_ctx.clear()
        gpu_ctx.render(spp, depth=DEPTH, seed=1000 + b, mode=eleven.MODE_WAVEFRONT)
        img, cnt, _, _ = gpu_ctx.read()
        means.append(img[..., :3].astype(np.float64))
    M = np.stack(means)
    g = M.mean(0)
    var_g_mean = M.var(0, ddof=1) / batches           # variance of the GPU mean
    var_r_mean = var_ref / n_ref
    w = 1.0 / (I_ref ** 2 + 1e-2)
    measured = np.mean((g - I_ref) ** 2 * w)
    predicted = np.mean((var_g_mean + var_r_mean) * w)
    ratio = measured / predicted
    se = np.sqrt(var_g_mean + var_r_mean)
    ok = se > 1e-9
    assert np.all(np.abs((g - I_ref)[~ok]) < 1e-5), "deterministic pixels differ"
    z = (g - I_ref)[ok] / se[ok]
    H, W = I_ref.shape[:2]
    bz = []
    for y in range(0, H - 7, 8):
        for x in range(0, W - 7, 8):
            d = (g - I_ref)[y:y + 8, x:x + 8].mean()
            # channels of one pixel share a path: bound their covariance by full correlation
            sd = np.sqrt(var_g_mean + var_r_mean)[y:y + 8, x:x + 8].sum(-1)
            v = (sd ** 2).sum() / (64 * 3) ** 2
            bz.append(d / np.sqrt(v) if v > 0 else 0.0)
    bz = np.abs(np.array(bz))
    print(f"  relMSE measured {measured:.3e} predicted {predicted:.3e} ratio {ratio:.3f}; "
          f"|z|>5 {np.mean(np.abs(z) > 5):.4f}; max block |z| {bz.max():.2f}", flush=True)
    assert 0.75 < ratio < 1.33
    assert bz.max() < 5
    assert np.mean(np.abs(z) > 5) < 0.002


@pytest.mark.gpu
def test_aov_normal_pass(gpu_ctx, scenes):
    print("[render] AOV passes", flush=True)
    s = scenes["spheres"]
    gpu_ctx.upload(s)
    gpu_ctx.render(4, seed=1, flags=eleven.FLAG_AOV)
    n = gpu_ctx.read_aov(eleven.PASS_NORMAL)[..., :3]
    # bottom rows of the frame see the ground plane first (vn 0 1 0)
    bottom = n[-3:, :, :].reshape(-1, 3)
    ground = np.abs(bottom[:, 1] - 1) < 1e-4
    assert ground.mean() > 0.5
    assert np.all(np.abs(bottom[ground, 0]) < 1e-5) and np.all(np.abs(bottom[ground, 2]) < 1e-5)
    with pytest.raises(eleven.ElevenError):
        gpu_ctx.read_aov(eleven.PASS_BEAUTY)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [eleven.MODE_WAVEFRONT, eleven.MODE_PLAIN])
def test_aov_passes_match_oracle(gpu_ctx, scenes, refbvhs, mode):
    """NORMAL / TANGENT / BITANGENT passes (kernel.cu:435-440, 451-476): the
    mean over accepted samples of the first hit's interpolated normal, its
    interpolated tangent (the loader's MikkTSpace tangents, not renormalised)
    and tangentsSign * N x T, against the oracle with common random numbers
    on a loaded scene (tangents from the file loader)."""
    print(f"[render] AOV passes vs oracle, mode {mode}", flush=True)
    s = scenes["spheres_sky_tex"]
    spp, seed = 8, 31
    gpu_ctx.upload(s)
    gpu_ctx.render(spp, depth=DEPTH, seed=seed, mode=mode, flags=eleven.FLAG_AOV | eleven.FLAG_NO_CULL)
    _, cnt, _, _ = gpu_ctx.read()
    H, W = s.camera.y_res, s.camera.x_res
    aov = np.zeros((H, W, 9))
    _, _, oc, _, _ = orc.render(s, refbvhs["spheres_sky_tex"], spp, depth=DEPTH, seed=seed, precision=0,
                                shadow_mode=1, aov=aov)
    same = cnt == oc
    assert same.mean() > 0.99
    om = aov / np.maximum(oc, 1)[..., None]
    for k, pid in enumerate((eleven.PASS_NORMAL, eleven.PASS_TANGENT, eleven.PASS_BITANGENT)):
        g = gpu_ctx.read_aov(pid)[..., :3].astype(np.float64)
        o = om[..., 3 * k:3 * k + 3]
        close = np.all(np.abs(g - o) <= 1e-4 + 1e-4 * np.abs(o), axis=-1)[same]
        print(f"  pass {pid}: pixels within 1e-4 {close.mean():.4f}", flush=True)
        assert close.mean() > 0.99
        assert np.abs(o).max() > 0.1                  # the pass is not empty


@pytest.mark.gpu
def test_error_paths(gpu_ctx, scenes):
    print("[render] error paths", flush=True)
    ctx = eleven.Context(0)
    with pytest.raises(eleven.ElevenError) as e:
        ctx.render(1)
    assert e.value.code == eleven.EL_ERR_STATE
    ctx.upload(scenes["spheres"])
    with pytest.raises(eleven.ElevenError) as e:
        ctx.render(1, bvh_width=6)
    assert e.value.code == eleven.EL_ERR_ARG
    with pytest.raises(eleven.ElevenError) as e:
        ctx.render(1, block=100)
    assert e.value.code == eleven.EL_ERR_ARG
    with pytest.raises(eleven.ElevenError) as e:
        ctx.render(1, depth=65)
    assert e.value.code == eleven.EL_ERR_RANGE
    st = ctx.render(0)               # spp = 0 is a no-op
    img, cnt, rays, _ = ctx.read()
    assert cnt.sum() == 0 and rays == 0
    ctx.close()


@pytest.mark.gpu
def test_progressive_accumulation(gpu_ctx, scenes):
    """Two renders of 4 spp equal one render of 8 spp (sample indices continue)."""
    print("[render] progressive accumulation", flush=True)
    s = scenes["spheres"]
    gpu_ctx.upload(s)
    gpu_ctx.render(4, seed=9)
    gpu_ctx.render(4, seed=9)
    a, ca, ra, _ = gpu_ctx.read()
    gpu_ctx.clear()
    gpu_ctx.render(8, seed=9)
    b, cb, rb, _ = gpu_ctx.read()
    assert np.array_equal(ca, cb) and ra == rb
    np.testing.assert_allclose(a, b, rtol=1e-5, atol=1e-6)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", SCENES)
def test_zero_throughput_culling_is_exact(gpu_ctx, scenes, kind):
    """A path whose throughput is exactly zero stops unless a later bounce
    could make it NaN (which rejects the sample): the image and the
    accepted-sample counts are bit-identical to tracing every path to the
    end as the reference does, for both renderers; only the traced-ray
    count drops."""
    print(f"[render] culling exactness: {kind}", flush=True)
    s = scenes[kind]
    for mode in (eleven.MODE_PLAIN, eleven.MODE_WAVEFRONT):
        a, ca, ra, _, _ = _render(gpu_ctx, s, 8, 7, mode, depth=8)
        b, cb, rb, _, _ = _render(gpu_ctx, s, 8, 7, mode, depth=8, flags=eleven.FLAG_NO_CULL)
        assert np.array_equal(ca, cb)
        assert np.array_equal(a, b)
        assert ra <= rb


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["spheres_sky_tex", "terrain_bokeh"])
def test_8bit_texture_copy_is_exact(gpu_ctx, scenes, kind):
    """Textures loaded from 8-bit files are read from a 4-byte-per-texel copy
    through the loader's byte -> value tables.  Writing the same values back
    with set_texture moves the texture to its FP32 copy: the images of the
    two storages must be identical bit for bit."""
    print(f"[render] 8-bit texture copy vs FP32: {kind}", flush=True)
    scene = scenes[kind]
    assert scene.view.texture_count > 0
    texels = sum(scene.view.textures[i].width * scene.view.textures[i].height
                 for i in range(scene.view.texture_count))
    for mode in (eleven.MODE_WAVEFRONT, eleven.MODE_PLAIN):
        a, ca, ra, _, _ = _render(gpu_ctx, scene, 4, 17, mode)
        assert gpu_ctx.stats().texels_8bit == texels   # every texture came from an 8-bit file
        for i in range(scene.view.texture_count):
            gpu_ctx.set_texture(i, np.ascontiguousarray(scene.texture(i), np.float32))
        assert gpu_ctx.stats().texels_8bit == 0
        gpu_ctx.clear()
        gpu_ctx.render(4, depth=5, seed=17, mode=mode)
        b, cb, rb, _ = gpu_ctx.read()
        assert np.array_equal(ca, cb) and ra == rb
        assert np.array_equal(a, b[..., :3].astype(np.float64))


@pytest.mark.gpu
def test_generations_split_equals_one_pass(gpu_ctx, scenes):
    """A queue smaller than the frame splits it into generations whose
    boundaries fall inside a sample pass (here 1000 paths over a 96x54 frame,
    padded to 8x8 tiles); each pixel still adds its samples in sample order,
    so the image is bit-identical to one generation."""
    print("[render] generations vs one pass", flush=True)
    s = scenes["terrain_bokeh"]
    a, ca, ra, va, _ = _render(gpu_ctx, s, 4, 5, eleven.MODE_WAVEFRONT, depth=6)
    b, cb, rb, vb, _ = _render(gpu_ctx, s, 4, 5, eleven.MODE_WAVEFRONT, depth=6)
    assert np.array_equal(a, b) and np.array_equal(ca, cb)   # deterministic
    gpu_ctx.upload(s)
    gpu_ctx.render(4, depth=6, seed=5, queue_cap=1000)
    c, cc, rc, vc = gpu_ctx.read()
    assert np.array_equal(cc, ca) and rc == ra and vc == va
    assert np.array_equal(c[..., :3].astype(np.float64), a)


@pytest.mark.gpu
@pytest.mark.parametrize("res", [(1, 1), (13, 7), (64, 8)])
def test_empty_scene_sees_the_environment(gpu_ctx, res):
    """No triangles: every camera ray escapes, so each sample's radiance is
    the constant environment (0.5, up to the lookup's filter weights) and
    every pixel's mean is 0.5; one BVH query per path, no vertices; plain and
    wavefront agree exactly.  Ragged frames (not multiples of the 8x8 tile)
    and a 1x1 frame included."""
    print(f"[render] empty scene {res}", flush=True)
    scene = eleven.SceneData.from_arrays(np.zeros((0, 9), np.float32), camera=dict(x_res=res[0], y_res=res[1]),
                                         hdri=np.full((1, 1, 3), 0.5, np.float32))
    out = []
    for mode in (eleven.MODE_PLAIN, eleven.MODE_WAVEFRONT):
        gpu_ctx.upload(scene)
        gpu_ctx.render(3, depth=4, seed=1, mode=mode)
        img, cnt, rays, verts = gpu_ctx.read()
        assert img.shape[:2] == (res[1], res[0])
        assert (cnt == 3).all()
        assert rays == res[0] * res[1] * 3 and verts == 0
        np.testing.assert_allclose(img[..., :3], 0.5, rtol=0, atol=1e-6)
        out.append(img)
    assert np.array_equal(out[0], out[1])


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [eleven.MODE_WAVEFRONT, eleven.MODE_PLAIN])
def test_bvh8_renders_the_bvh4_image(gpu_ctx, scenes, mode):
    """The closest hit does not depend on the tree (min t, ties to the lower
    index), so both widths trace the same paths: identical counts, tallies
    and pixels."""
    print(f"[render] BVH8 vs BVH4, mode {mode}", flush=True)
    s = scenes["terrain_bokeh"]
    out = []
    for width in (4, 8):
        gpu_ctx.upload(s)
        gpu_ctx.render(4, depth=6, seed=21, mode=mode, bvh_width=width)
        out.append(gpu_ctx.read())
    (a, ca, ra, va), (b, cb, rb, vb) = out
    assert gpu_ctx.stats().bvh_width == 8
    assert np.array_equal(ca, cb) and ra == rb and va == vb
    assert np.array_equal(a, b)


@pytest.mark.gpu
def test_overlapped_shadows_equal_serial_with_spilled_stacks(gpu_ctx):
    """As below on a scene whose traversal stacks spill to the global area:
    the overlapped shadow launches spill to their own area, not the one the
    extension running beside them uses (both index it by grid thread)."""
    print("[render] overlapped vs serial shadows, spilling stacks", flush=True)
    s = deep_stack_scene()
    gpu_ctx.upload(s)
    out = []
    for serial in (0, eleven.FLAG_SERIAL_SHADOWS, 0):
        gpu_ctx.clear()
        st = gpu_ctx.render(16, depth=DEPTH, seed=9, flags=serial)
        img, cnt, _, _ = gpu_ctx.read()
        out.append((img, cnt, st.rays, st.vertices, st.stack_bound))
    assert out[0][4] > 20, out[0][4]   # stacks deeper than the 16 LDS entries (28)
    for o in out[1:]:
        assert np.array_equal(out[0][0].view(np.uint32), o[0].view(np.uint32))
        assert np.array_equal(out[0][1], o[1])
        assert out[0][2:4] == o[2:4]


@pytest.mark.gpu
@pytest.mark.parametrize("kind", SCENES)
def test_overlapped_shadows_equal_serial(gpu_ctx, scenes, kind):
    """The shadow launches of a bounce overlapping the next extension on a
    second stream (the default for small generations) leave the image, the
    per-pixel counts and the tallies bit-identical to serial launches
    (EL_FLAG_SERIAL_SHADOWS): the launches share no queue counters and the
    radiance is summed in the same order.  Also with recording (the shadow
    launches set visibility bits) and culling off."""
    print(f"[render] overlapped vs serial shadows: {kind}", flush=True)
    s = scenes[kind]
    for extra in (0, eleven.FLAG_NO_CULL, eleven.FLAG_RECORD_PATHS):
        out = []
        for serial in (0, eleven.FLAG_SERIAL_SHADOWS):
            gpu_ctx.upload(s)
            st = gpu_ctx.render(16, depth=DEPTH, seed=5, flags=serial | extra)
            img, cnt, rays, verts = gpu_ctx.read()
            out.append((img, cnt, st.rays, st.vertices))
        assert np.array_equal(out[0][0].view(np.uint32), out[1][0].view(np.uint32))
        assert np.array_equal(out[0][1], out[1][1])
        assert out[0][2:] == out[1][2:]


@pytest.mark.gpu
def test_kernel_times_only_on_request(gpu_ctx, scenes):
    """Per-launch timing (EL_FLAG_KERNEL_TIMES) fills the per-kind times and
    leaves the image bit for bit as it is without it."""
    print("[render] per-launch kernel times on request", flush=True)
    s = scenes["spheres"]
    gpu_ctx.upload(s)
    out = []
    for flags in (0, eleven.FLAG_KERNEL_TIMES):
        gpu_ctx.clear()
        st = gpu_ctx.render(8, depth=DEPTH, seed=3, flags=flags)
        out.append((gpu_ctx.read()[0], st.extend_ms, st.shade_ms, st.shadow_ms, st.render_ms))
    assert out[0][1] == 0.0 and out[0][2] == 0.0 and out[0][3] == 0.0 and out[0][4] > 0.0
    assert out[1][1] > 0.0 and out[1][2] > 0.0 and out[1][3] > 0.0
    assert np.array_equal(out[0][0].view(np.uint32), out[1][0].view(np.uint32))


@pytest.mark.gpu
@pytest.mark.parametrize("depth", [1, 2, 3])
def test_deferred_light_terms_at_shallow_depths(gpu_ctx, scenes, depth):
    """Point-light terms of overlapped bounces wait in a per-path buffer for
    the next shade or, on a path's last bounce, for finalize; at depths 1-3
    most or all paths end with such a term pending.  Images, counts and
    tallies equal the serial launch order (the launch adds the term) bit for
    bit; the plain megakernel's counts exactly and its image as in
    test_plain_equals_wavefront (its FP32 shading rounds a few values
    differently, 1-2 ulp)."""
    print(f"[render] deferred point-light terms, depth {depth}", flush=True)
    s = scenes["terrain_bokeh"]   # 4 point lights: every hit draws a light query
    gpu_ctx.upload(s)
    out = []
    for mode, flags in ((eleven.MODE_WAVEFRONT, 0), (eleven.MODE_WAVEFRONT, eleven.FLAG_SERIAL_SHADOWS),
                        (eleven.MODE_PLAIN, 0)):
        gpu_ctx.clear()
        st = gpu_ctx.render(8, depth=depth, seed=11, mode=mode, flags=flags)
        img, cnt, _, _ = gpu_ctx.read()
        out.append((img, cnt, st.rays, st.vertices))
    assert np.array_equal(out[0][0].view(np.uint32), out[1][0].view(np.uint32))
    assert np.array_equal(out[0][1], out[1][1])
    assert out[0][2:] == out[1][2:]
    assert np.array_equal(out[0][1], out[2][1])
    np.testing.assert_allclose(out[0][0], out[2][0], rtol=1e-4, atol=1e-6)
