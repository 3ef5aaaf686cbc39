"""Multi-view rendering and inverse rendering (el_set_camera, eleven.inverse).

1. el_set_camera on an uploaded scene renders exactly what an upload of the
   same scene with that camera renders (same resolution and a new one):
   counts and radiance bit for bit.
2. The summed multi-view gradient of eleven.inverse.MultiView (correlated
   mode: residual and derivative from the same paths) equals central
   differences of the multi-view L2 loss with common random numbers (fixed
   seeds), within 1e-2 relative, for parameters sampled with seed 2.
3. EL_ADJ_ACCUMULATE sums the gradients of successive el_adjoint calls on
   the device: the sum over views equals the sum of the separate calls.
4. A short Adam run with decorrelated gradients from perturbed parameters
   lowers the loss and the parameter error weighted by the initial gradient
   magnitude (the error along the directions the images see: Adam moves
   every parameter by about lr per step, so ones the views hardly see drift
   with the gradient noise).
"""
import numpy as np
import pytest

import eleven
from eleven import inverse, scenegen
from eleven.autograd import material_params

DEPTH = 4


def _scene(root, name, seed=3, res=(48, 40), n_spheres=10, tex=16):
    rng = np.random.default_rng(seed + 100)
    texture = (rng.uniform(0.1, 0.9, size=(tex, tex, 3)) * 255).astype(np.uint8)
    d = scenegen.sphere_field(root / name, n_spheres=n_spheres, res=res, seed=seed, ground_texture=texture,
                              extent=6.0)
    s = eleven.SceneData.load(d)
    mats = s.view.materials
    for i in range(s.view.material_count):    # interior metallic values: differentiable
        mats[i].metallic = 0.2 + 0.6 * ((i * 37) % 10) / 10.0
    return s


def _render(ctx, spp=4, seed=9):
    ctx.render(spp, depth=DEPTH, seed=seed, mode=eleven.MODE_WAVEFRONT)
    img, cnt, rays, _ = ctx.read()
    return img, cnt, rays


@pytest.mark.gpu
@pytest.mark.parametrize("res", [None, (37, 29)])
def test_set_camera_equals_upload_with_that_camera(gpu_ctx, scene_root, res):
    print(f"[inverse] set_camera vs upload, res {res}", flush=True)
    a = _scene(scene_root, "cam_a")
    base = a.camera
    xr, yr = res if res else (base.x_res, base.y_res)
    cam = eleven.make_camera(xr, yr, position=(base.position.x + 1.5, base.position.y + 0.5, base.position.z),
                             rotation=(base.rotation.x + 4.0, base.rotation.y + 3.0, base.rotation.z),
                             focal_length=base.focal_length, sensor_width=base.sensor_width)
    gpu_ctx.upload(a)
    gpu_ctx.set_camera(cam)
    img1, cnt1, rays1 = _render(gpu_ctx)
    assert img1.shape[:2] == (yr, xr)
    b = _scene(scene_root, "cam_b")
    b.view.camera = cam
    gpu_ctx.upload(b)
    img2, cnt2, rays2 = _render(gpu_ctx)
    assert rays1 == rays2
    assert np.array_equal(cnt1, cnt2)
    assert np.array_equal(img1, img2)


def _views(scene, n=3):
    return inverse.orbit_views(scene.camera, n, x_span=1.5, pitches=(scene.camera.rotation.x,
                                                                     scene.camera.rotation.x + 5.0))


@pytest.mark.gpu
def test_multiview_gradient_matches_central_differences(gpu_ctx, scene_root):
    print("[inverse] multi-view gradient vs central differences", flush=True)
    scene = _scene(scene_root, "mv_fd", res=(40, 32))
    gpu_ctx.upload(scene)
    mv = inverse.MultiView(gpu_ctx, scene, _views(scene), spp=8, depth=DEPTH, grad_texture=0)
    p0 = material_params(scene)
    t0 = scene.texture(0).copy()
    rng = np.random.default_rng(0)
    pt = np.clip(p0 + rng.normal(0, 0.1, p0.shape).astype(np.float32), inverse.PARAM_LO, inverse.PARAM_HI)
    targets = [im.astype(np.float64) for im in mv.images(pt, np.clip(t0 * 1.2, 0, 1), seed=50)]
    loss0, dmat, dtex = mv.loss_and_grad(p0, t0, targets, seed=5, decorrelate=False)
    assert loss0 > 0 and np.abs(dmat).max() > 0 and np.abs(dtex).max() > 0

    def loss(p, t):
        return mv.loss_and_grad(p, t, targets, seed=5, decorrelate=False)[0]

    mats = scene.materials()
    sel = np.random.default_rng(2)
    cand = [(m, k) for m in range(len(mats)) for k in (0, 1, 2, 5)
            if not (k < 3 and mats[m].albedo_tex >= 0) and abs(dmat[m, k]) > 1e-3 * np.abs(dmat).max()]
    picks = [cand[i] for i in sel.choice(len(cand), 6, replace=False)]
    h = 1e-3
    worst = 0.0
    for (m, k) in picks:
        pp, pm = p0.copy(), p0.copy()
        pp[m, k] += h
        pm[m, k] -= h
        fd = (loss(pp, t0) - loss(pm, t0)) / float(pp[m, k] - pm[m, k])
        err = abs(fd - dmat[m, k]) / max(abs(fd), 1e-12)
        print(f"[inverse] material {m} param {k}: fd {fd:.6e} adjoint {dmat[m, k]:.6e} rel {err:.2e}", flush=True)
        worst = max(worst, err)
    idx = np.flatnonzero(np.abs(dtex) > 1e-2 * np.abs(dtex).max())
    for i in sel.choice(idx, 4, replace=False):
        y, x, c = np.unravel_index(i, dtex.shape)
        tp, tm = t0.copy(), t0.copy()
        tp[y, x, c] += h
        tm[y, x, c] -= h
        fd = (loss(p0, tp) - loss(p0, tm)) / float(tp[y, x, c] - tm[y, x, c])
        err = abs(fd - dtex[y, x, c]) / max(abs(fd), 1e-12)
        print(f"[inverse] texel ({y},{x},{c}): fd {fd:.6e} adjoint {dtex[y, x, c]:.6e} rel {err:.2e}", flush=True)
        worst = max(worst, err)
    assert worst < 1e-2, worst


@pytest.mark.gpu
def test_short_inverse_run_reduces_parameter_error(gpu_ctx, scene_root):
    print("[inverse] short Adam run", flush=True)
    scene = _scene(scene_root, "mv_opt", res=(48, 40))
    gpu_ctx.upload(scene)
    mv = inverse.MultiView(gpu_ctx, scene, _views(scene, 4), spp=8, depth=DEPTH, grad_texture=0)
    truth = material_params(scene)
    tex = scene.texture(0).copy()
    targets = [im.astype(np.float64) for im in mv.images(truth, tex, seed=777, spp=64)]
    rng = np.random.default_rng(3)
    p0 = np.clip(truth + rng.normal(0, 0.1, truth.shape).astype(np.float32), inverse.PARAM_LO, inverse.PARAM_HI)
    _, g0, _ = mv.loss_and_grad(p0, None, targets, seed=1)
    res = inverse.optimize(mv, p0, None, targets, steps=40, lr=1e-2)
    mask = np.ones(truth.shape, bool)
    for m, mat in enumerate(scene.materials()):
        if mat.albedo_tex >= 0:
            mask[m, :3] = False   # the texture replaces the scalar albedo
    e0 = inverse.param_rmse(p0, truth, mask, weights=g0)
    e1 = inverse.param_rmse(res.params, truth, mask, weights=g0)
    l0, l1 = np.mean(res.losses[:5]), np.mean(res.losses[-5:])
    print(f"[inverse] loss {l0:.4g} -> {l1:.4g}, gradient-weighted parameter RMSE {e0:.4f} -> {e1:.4f}, "
          f"{res.grads_per_s:.1f} grads/s", flush=True)
    assert l1 < l0
    assert e1 < 0.8 * e0


@pytest.mark.gpu
def test_adjoint_accumulate_equals_sum_of_calls(gpu_ctx, scene_root):
    print("[inverse] EL_ADJ_ACCUMULATE", flush=True)
    scene = _scene(scene_root, "acc", res=(40, 32))
    gpu_ctx.upload(scene)
    views = _views(scene)
    rng = np.random.default_rng(7)
    parts = []
    for v, cam in enumerate(views):
        gpu_ctx.set_camera(cam)
        gpu_ctx.render(4, depth=DEPTH, seed=20 + v, flags=eleven.FLAG_RECORD_PATHS)
        w = rng.normal(size=(cam.y_res, cam.x_res, 3)).astype(np.float32)
        parts.append(gpu_ctx.adjoint(w, grad_texture=0) + (w,))
    total_m = sum(p[0].astype(np.float64) for p in parts)
    total_t = sum(p[1].astype(np.float64) for p in parts)
    for v, cam in enumerate(views):
        gpu_ctx.set_camera(cam)
        gpu_ctx.render(4, depth=DEPTH, seed=20 + v, flags=eleven.FLAG_RECORD_PATHS)
        dm, dt = gpu_ctx.adjoint(parts[v][2], grad_texture=0, accumulate=v > 0, read=v == len(views) - 1)
        if v < len(views) - 1:
            assert dm is None and dt is None
    assert np.linalg.norm(dm - total_m) <= 1e-5 * np.linalg.norm(total_m)
    assert np.linalg.norm(dt - total_t) <= 1e-5 * np.linalg.norm(total_t)
    with pytest.raises(eleven.ElevenError):   # another texture than the running sums'
        gpu_ctx.adjoint(parts[0][2], grad_texture=-1, accumulate=True, read=False)


@pytest.mark.gpu
def test_device_residual_equals_host_residual(gpu_ctx, scene_root):
    """el_residual_l2 + EL_ADJ_DEVICE_RESIDUAL give the gradients of an
    adjoint fed 2 (I - target) computed on the host from read() in FP64 (the
    same FP32 residual; the gradients differ only by the order of the
    atomic sums) and the same loss;
    the residual survives el_clear and a second render (the decorrelated
    step), and an adjoint asking for it without one fails."""
    print("[inverse] device residual", flush=True)
    scene = _scene(scene_root, "res", res=(40, 32))
    gpu_ctx.upload(scene)
    cam = _views(scene)[0]
    gpu_ctx.set_camera(cam)
    with pytest.raises(eleven.ElevenError):   # no residual yet
        gpu_ctx.adjoint(None, grad_texture=0)
    target = np.random.default_rng(5).uniform(0, 1, size=(cam.y_res, cam.x_res, 3)).astype(np.float32)
    gpu_ctx.render(4, depth=DEPTH, seed=31)
    img = gpu_ctx.read()[0][..., :3]
    diff = img.astype(np.float64) - target
    loss_dev = gpu_ctx.residual_l2(target)
    assert abs(loss_dev - float(np.sum(diff * diff))) <= 1e-9 * float(np.sum(diff * diff))
    gpu_ctx.clear()
    gpu_ctx.render(4, depth=DEPTH, seed=32, flags=eleven.FLAG_RECORD_PATHS)
    dm_dev, dt_dev = gpu_ctx.adjoint(None, grad_texture=0)
    dm_host, dt_host = gpu_ctx.adjoint((2.0 * diff).astype(np.float32), grad_texture=0)
    assert np.linalg.norm(dm_dev - dm_host) <= 1e-6 * np.linalg.norm(dm_host)
    assert np.linalg.norm(dt_dev - dt_host) <= 1e-6 * np.linalg.norm(dt_host)
    assert np.abs(dm_host).max() > 0
    with pytest.raises(eleven.ElevenError):   # the host image replaced the residual
        gpu_ctx.adjoint(None, grad_texture=0)
