"""Parity of the HIP path-replay adjoint (el_adjoint, k_adjoint).

1. GPU vs the oracle's adjoint (or_adjoint) on the same scene, seed, spp and
   depth -- the same paths, replayed on both sides.  The FP32-arithmetic
   oracle must agree to 1e-3 of the gradient norm.  Against the FP64 oracle
   (the reference gradient) the comparison is made on the pixels whose FP32
   and FP64 primal images agree: about 0.15 % of paths take another discrete
   branch in FP64 (a grazing hit, the convexity test of the shading position,
   a lobe threshold), which moves the whole-image gradient by ~2 %; on the
   remaining pixels the two must agree to 1e-2 of the norm and per parameter
   for every parameter above 1 % of the largest, which covers all six
   material parameter kinds and the texels.  The unmasked difference is
   bounded at 5e-2.
2. GPU vs central differences of the GPU renderer with common random numbers
   (same seed): an L2 loss against a target rendered with parameters
   perturbed by N(0, 0.1^2), 16 parameters sampled with seed 2 among albedo,
   specular and albedo texels, where sampling does not depend on the
   parameter and the adjoint is the derivative of the renderer itself:
   agreement within 1e-2 relative.
3. The torch.autograd binding returns exactly el_adjoint's gradients.
"""
import numpy as np
import pytest

import eleven
import oracle_py as orc
from eleven import scenegen

SPP, DEPTH, SEED = 8, 4, 5


def _scene(root, name, n_spheres, res, tex_size, seed):
    rng = np.random.default_rng(seed + 100)
    tex = (rng.uniform(0.1, 0.9, size=(tex_size, tex_size, 3)) * 255).astype(np.uint8)
    d = scenegen.sphere_field(root / name, n_spheres=n_spheres, res=res, seed=seed, ground_texture=tex,
                              extent=6.0)
    s = eleven.SceneData.load(d)
    mats = s.view.materials
    for i in range(s.view.material_count):       # interior metallic values (see test_oracle_adjoint)
        mats[i].metallic = 0.2 + 0.6 * ((i * 37) % 10) / 10.0
    return s


@pytest.fixture(scope="module")
def small(scene_root):
    s = _scene(scene_root, "adj_small", 12, (32, 32), 16, 3)
    return s, orc.RefBVH(s.vertices())


@pytest.mark.gpu
def test_adjoint_matches_oracle_adjoint(gpu_ctx, small):
    print("[adjoint] GPU vs oracle adjoint", flush=True)
    scene, rb = small
    H, W = scene.camera.y_res, scene.camera.x_res
    gpu_ctx.upload(scene)
    gpu_ctx.clear()
    gpu_ctx.render(SPP, depth=DEPTH, seed=SEED, mode=eleven.MODE_WAVEFRONT)
    w = np.random.default_rng(1).normal(size=(H, W, 3)).astype(np.float32)
    dmat, dtex = gpu_ctx.adjoint(w, grad_texture=0)
    assert np.abs(dmat).max() > 0 and np.abs(dtex).max() > 0
    imgs = {}
    for prec, tol in ((0, 1e-3), (1, 5e-2)):
        om, ot, imgs[prec] = orc.adjoint(scene, rb, w, SPP, depth=DEPTH, seed=SEED, precision=prec, shadow_mode=1,
                                         grad_texture=0)
        rm = np.linalg.norm(dmat - om) / np.linalg.norm(om)
        rt = np.linalg.norm(dtex - ot) / np.linalg.norm(ot)
        print(f"[adjoint] precision {prec}: material rel err {rm:.2e}, texel rel err {rt:.2e}", flush=True)
        assert rm < tol and rt < tol, (prec, rm, rt)
    # FP32 rounding moves a pixel's estimate by < 1e-4 relative; a different
    # branch moves it by far more
    same = np.all(np.abs(imgs[0] - imgs[1]) <= 1e-3 * (np.abs(imgs[1]) + 1e-3), axis=-1)
    print(f"[adjoint] pixels with the same FP32 / FP64 paths: {same.sum()} of {same.size}", flush=True)
    assert same.mean() > 0.95
    wm = w * same[..., None]
    dmat, dtex = gpu_ctx.adjoint(wm, grad_texture=0)
    om, ot, _ = orc.adjoint(scene, rb, wm, SPP, depth=DEPTH, seed=SEED, precision=1, shadow_mode=1, grad_texture=0)
    rm = np.linalg.norm(dmat - om) / np.linalg.norm(om)
    rt = np.linalg.norm(dtex - ot) / np.linalg.norm(ot)
    big = np.abs(om) > 0.01 * np.abs(om).max()
    rel = np.abs(dmat - om)[big] / np.abs(om)[big]
    print(f"[adjoint] FP64 on agreeing pixels: material rel err {rm:.2e}, texel rel err {rt:.2e}, "
          f"worst parameter {rel.max():.2e}", flush=True)
    assert rm < 1e-2 and rt < 1e-2 and rel.max() < 1e-2, (rm, rt, rel.max())
    for k in range(6):   # every parameter kind is exercised
        assert np.abs(om[:, k]).max() > 0


def _loss(img, target):
    return float(((img.astype(np.float64) - target) ** 2).sum())


@pytest.mark.gpu
def test_adjoint_matches_central_differences(gpu_ctx, scene_root):
    print("[adjoint] GPU vs central differences (CRN)", flush=True)
    from eleven.autograd import Renderer, material_params
    scene = _scene(scene_root, "adj_fd", 16, (64, 64), 32, 4)
    gpu_ctx.upload(scene)
    r = Renderer(gpu_ctx, scene, spp=16, depth=DEPTH, seed=SEED, grad_texture=0)
    p0 = material_params(scene)
    t0 = scene.texture(0).copy()
    rng = np.random.default_rng(0)
    pt = p0 + rng.normal(0, 0.1, p0.shape).astype(np.float32)
    pt[:, :3] = np.clip(pt[:, :3], 0, 1)
    pt[:, 3] = np.clip(pt[:, 3], 0.05, 1)
    pt[:, 4:] = np.clip(pt[:, 4:], 0, 1)
    target = r.render(pt, np.clip(t0 * 1.2, 0, 1)).astype(np.float64)
    img = r.render(p0, t0)
    dmat, dtex = gpu_ctx.adjoint(2 * (img - target), grad_texture=0)

    sel = np.random.default_rng(2)
    mats = scene.materials()
    cand = [(m, k) for m in range(len(mats)) for k in (0, 1, 2, 5)
            if not (k < 3 and mats[m].albedo_tex >= 0) and abs(dmat[m, k]) > 1e-3 * np.abs(dmat).max()]
    picks = [cand[i] for i in sel.choice(len(cand), 8, replace=False)]
    tex_idx = np.flatnonzero(np.abs(dtex) > 1e-2 * np.abs(dtex).max())
    tpicks = [np.unravel_index(i, dtex.shape) for i in sel.choice(tex_idx, 8, replace=False)]
    h = 1e-3
    worst = 0.0
    for (m, k) in picks:
        pp, pm = p0.copy(), p0.copy()
        pp[m, k] += h
        pm[m, k] -= h
        fd = (_loss(r.render(pp, t0), target) - _loss(r.render(pm, t0), target)) / float(pp[m, k] - pm[m, k])
        err = abs(fd - dmat[m, k]) / max(abs(fd), 1e-12)
        print(f"[adjoint] material {m} param {k}: fd {fd:.6e} adjoint {dmat[m, k]:.6e} rel {err:.2e}", flush=True)
        worst = max(worst, err)
    for (y, x, c) in tpicks:
        tp, tm = t0.copy(), t0.copy()
        tp[y, x, c] += h
        tm[y, x, c] -= h
        fd = (_loss(r.render(p0, tp), target) - _loss(r.render(p0, tm), target)) / float(tp[y, x, c] - tm[y, x, c])
        err = abs(fd - dtex[y, x, c]) / max(abs(fd), 1e-12)
        print(f"[adjoint] texel ({y},{x},{c}): fd {fd:.6e} adjoint {dtex[y, x, c]:.6e} rel {err:.2e}", flush=True)
        worst = max(worst, err)
    assert worst < 1e-2, worst


@pytest.mark.gpu
def test_autograd_function_returns_el_adjoint_gradients(gpu_ctx, small):
    import torch
    from eleven.autograd import Renderer, material_params
    scene, _ = small
    gpu_ctx.upload(scene)
    r = Renderer(gpu_ctx, scene, spp=4, depth=3, seed=9, grad_texture=0)
    p = torch.tensor(material_params(scene), requires_grad=True)
    t = torch.tensor(scene.texture(0).copy(), requires_grad=True)
    img = r(p, t)
    target = torch.zeros_like(img)
    loss = ((img - target) ** 2).sum()
    loss.backward()
    dmat, dtex = gpu_ctx.adjoint((2 * img).detach().numpy(), grad_texture=0)
    np.testing.assert_allclose(p.grad.numpy(), dmat, rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(t.grad.numpy(), dtex, rtol=1e-5, atol=1e-7)


@pytest.mark.gpu
def test_adjoint_error_paths(gpu_ctx, small):
    scene, _ = small
    gpu_ctx.upload(scene)
    gpu_ctx.clear()
    H, W = scene.camera.y_res, scene.camera.x_res
    with pytest.raises(eleven.ElevenError):
        gpu_ctx.adjoint(np.zeros((H, W, 3), np.float32))          # nothing rendered
    gpu_ctx.render(1, depth=3, seed=1)
    gpu_ctx.render(1, depth=3, seed=2)
    with pytest.raises(eleven.ElevenError):
        gpu_ctx.adjoint(np.zeros((H, W, 3), np.float32))          # mixed seeds
    gpu_ctx.clear()
    gpu_ctx.render(1, depth=17, seed=1)
    with pytest.raises(eleven.ElevenError):
        gpu_ctx.adjoint(np.zeros((H, W, 3), np.float32))          # deeper than the replay records


@pytest.mark.gpu
@pytest.mark.parametrize("flags,queue_cap", [(0, 0), (eleven.FLAG_NO_CULL, 0), (0, 1024)],
                         ids=["cull", "no_cull", "generations"])
def test_recorded_vertices_replay_without_tracing(gpu_ctx, small, flags, queue_cap):
    """A wavefront render with EL_FLAG_RECORD_PATHS also keeps every vertex's
    closest hit and both shadow answers.  Its image and sample counts equal
    the render without the flag bit for bit (the shadow queries it adds have
    zero contribution), and el_adjoint then replays the recorded vertices,
    tracing nothing, to the gradients of the traced replay.  Without culling
    the recording holds vertices past the point where the replay stops; with
    a small queue the recording spans several generations."""
    print(f"[adjoint] recorded vertices, flags {flags}, queue {queue_cap}", flush=True)
    scene, _ = small
    H, W = scene.camera.y_res, scene.camera.x_res
    gpu_ctx.upload(scene)
    w = np.random.default_rng(6).normal(size=(H, W, 3)).astype(np.float32)
    gpu_ctx.clear()
    gpu_ctx.render(SPP, depth=DEPTH, seed=SEED, flags=flags)
    img0, n0 = gpu_ctx.read()[:2]
    m1, t1 = gpu_ctx.adjoint(w, grad_texture=0)
    assert gpu_ctx.stats().adjoint_rays > 0
    gpu_ctx.clear()
    gpu_ctx.render(SPP, depth=DEPTH, seed=SEED, flags=flags | eleven.FLAG_RECORD_PATHS, queue_cap=queue_cap)
    img1, n1 = gpu_ctx.read()[:2]
    assert np.array_equal(n0, n1) and np.array_equal(img0, img1)
    m2, t2 = gpu_ctx.adjoint(w, grad_texture=0)
    assert gpu_ctx.stats().adjoint_rays == 0
    assert np.abs(m1).max() > 0 and np.abs(t1).max() > 0
    # the recorded radiance sums its terms in another order than the walk
    assert np.linalg.norm(m2 - m1) <= 1e-4 * np.linalg.norm(m1)
    assert np.linalg.norm(t2 - t1) <= 1e-4 * np.linalg.norm(t1)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [eleven.MODE_WAVEFRONT, eleven.MODE_PLAIN])
def test_adjoint_from_recorded_paths_equals_replay_adjoint(gpu_ctx, small, mode):
    """With EL_FLAG_RECORD_PATHS the adjoint takes each path's radiance from
    the forward render instead of walking the path a first time: the same
    gradients up to the FP32 summation order of the two radiance sums.
    Without it each path is traced once (the primal walk, which keeps its
    vertices per thread) and replayed from those vertices; with it the plain
    renderer's record traces only the paths whose adjoint weight is nonzero,
    the wavefront one's nothing."""
    print(f"[adjoint] recorded vs two-walk, mode {mode}", flush=True)
    scene, _ = small
    H, W = scene.camera.y_res, scene.camera.x_res
    gpu_ctx.upload(scene)
    w = np.random.default_rng(4).normal(size=(H, W, 3)).astype(np.float32)
    gpu_ctx.clear()
    gpu_ctx.render(SPP, depth=DEPTH, seed=SEED, mode=mode)
    m1, t1 = gpu_ctx.adjoint(w, grad_texture=0)
    r1 = gpu_ctx.stats().adjoint_rays
    gpu_ctx.clear()
    gpu_ctx.render(SPP, depth=DEPTH, seed=SEED, mode=mode, flags=eleven.FLAG_RECORD_PATHS)
    m2, t2 = gpu_ctx.adjoint(w, grad_texture=0)
    r2 = gpu_ctx.stats().adjoint_rays
    assert r1 > 0 and r2 <= r1           # at most one traced walk per path
    assert np.linalg.norm(m2 - m1) <= 1e-4 * np.linalg.norm(m1)
    assert np.linalg.norm(t2 - t1) <= 1e-4 * np.linalg.norm(t1)
    gpu_ctx.render(1, depth=DEPTH, seed=SEED, mode=mode)   # the record no longer covers the accumulation
    m3, _ = gpu_ctx.adjoint(w, grad_texture=0)
    assert np.abs(m3).max() > 0
    with pytest.raises(eleven.ElevenError):
        gpu_ctx.render(1, depth=DEPTH, seed=SEED, mode=mode, flags=eleven.FLAG_RECORD_PATHS)
