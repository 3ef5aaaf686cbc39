#!/usr/bin/env python3
"""Throughput of the path-tracing hot path on one MI355X.

Workload (one "step"): the 1,048,576-triangle heightfield scene at
1920x1080, 64 samples per pixel, depth 8, defocus aperture 0.05, 16 textured
materials, 8 point lights and a 2048x1024 HDR sky, rendered from scratch
(accumulation cleared) by the wavefront renderer.  Inputs (BVH, triangles,
textures, HDRI tables) are resident in HBM before the timed region.

metric: Mrays/s counting every BVH query of the estimator (camera +
extension + HDRI shadow + point-light shadow), the same logical count for
the plain and wavefront renderers.

Extra fields: `plain` (the one-thread-per-pixel megakernel on the same
frame), `roofline` of the dominant kernel (closest-hit extension),
`cpu_baseline` (the CPU oracle, a restatement of the reference estimator,
on a bounded crop of the same frame using all host cores) and `adjoint`
(the metric's second part: gradients per second of forward + path-replay
adjoint on the 64-sphere gradient-check scene, with the FP64 CPU adjoint
timed beside it) and `batch` (32 scenes of uneven size through one tuned,
graph-captured wavefront loop, against the plain megakernel scene by
scene).
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "tfg-pathtracer_amd"))

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def ensure_built():
    import eleven
    if not eleven.LIB_PATH.exists():
        import __graft_entry__
        __graft_entry__.build()


def measured_traffic(kernel="k_wf_extend", per="launch"):
    """HBM bytes per launch (or per frame) of `kernel` on this workload at the
    default settings, from the committed FETCH_SIZE / WRITE_SIZE counter
    summary (rocprofv3 --pmc, scaled by a 1 GiB copy).  None if absent."""
    f = ROOT / "profiles" / "hbm_traffic_heightfield_1080p_64spp.md"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        parts = line.split()
        if len(parts) == 3 and parts[0] == f"traffic_bytes_per_{per}" and parts[1] == kernel:
            return int(parts[2])
    return None


def copy_bandwidth(gib=2, reps=5):
    """Achievable device bandwidth on this GPU: the better of a device-to-device
    copy and an elementwise copy kernel over `gib` GiB (bytes read + written
    per second), timed with events on torch's current stream; and the same
    copy over buffers that stay cache-resident between repetitions: 1 MiB
    read + 1 MiB written (L2: 4 MB per XCD) and the traversal's working sets
    (the heightfield's 22.4 MB node array, nodes + triangles 72.7 MB: the
    256 MB MALL).  Returns {"copy_gbs": HBM copy, "resident_gbs": {MiB: GB/s}}."""
    import torch

    def timed(op, nbytes, n):
        op()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(n):
            op()
        e1.record()
        e1.synchronize()
        return nbytes * n / (e0.elapsed_time(e1) * 1e-3) / 1e9

    n = (gib << 30) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda:0")
    y = torch.empty_like(x)
    best = max(timed(lambda: y.copy_(x), 2 * n * 4, reps), timed(lambda: torch.mul(x, 1.0, out=y), 2 * n * 4, reps))
    del x, y
    resident = {}
    for mib in (1, 22.4, 72.7):
        m = int(mib * (1 << 20)) // 4
        a = torch.ones(m, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        k = max(200, int(20e9 / (8 * m)))
        resident[str(mib)] = round(timed(lambda: b.copy_(a), 2 * m * 4, k), 1)
        del a, b
    torch.cuda.empty_cache()
    return {"copy_gbs": best, "resident_gbs": resident}


def cpu_threads():
    """Host threads for the CPU baselines: every core this process may run on."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


_ORACLE = None


def native_oracle():
    """The CPU oracle compiled for this host (-O3 -march=native, one
    std::thread per core) into a temporary directory, for the CPU baselines
    only; FP contraction stays off, so it computes what the oracle the tests
    use computes."""
    global _ORACLE
    if _ORACLE is None:
        sys.path.insert(0, str(ROOT / "tests"))
        import tempfile
        import oracle_py as orc
        so = Path(tempfile.gettempdir()) / f"eleven_oracle_native_{os.getpid()}" / "liboracle.so"
        so.parent.mkdir(parents=True, exist_ok=True)
        subprocess.run(["g++", "-O3", "-march=native", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
                        "-shared", "-o", str(so), str(ROOT / "oracle" / "oracle.cpp"), "-lpthread"], check=True)
        orc.ORACLE_SO = so
        _ORACLE = orc
    return _ORACLE


def cpu_render(scene, depth, window, budget_s, threads, seed=0):
    """The oracle's FP32 estimator (the reference's arithmetic) over the
    reference-faithful depth-18 / 14-bin SAH binary BVH on `window`, its
    sample count chosen from a 1-spp probe so the timed part takes about
    budget_s.  Returns (Mrays/s, spp, rays, seconds, tree build seconds)."""
    orc = native_oracle()
    t0 = time.perf_counter()
    ref = orc.RefBVH(scene.vertices())
    build_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    orc.render(scene, ref, 1, depth=depth, seed=seed, window=window, threads=threads)
    probe = max(time.perf_counter() - t0, 1e-3)
    spp = max(1, min(4096, int(budget_s / probe)))
    t0 = time.perf_counter()
    r = orc.render(scene, ref, spp, depth=depth, seed=seed, s0=1, window=window, threads=threads)
    dt = time.perf_counter() - t0
    return r[3] / dt / 1e6, spp, r[3], dt, build_s


def cpu_baseline(scene, budget_s=10.0, threads=None):
    """CPU baseline of the headline workload: the oracle's FP32 estimator
    (the reference's arithmetic) on a centred crop of the same frame."""
    threads = threads or cpu_threads()
    cam = scene.camera
    w, h = max(1, cam.x_res // 8), max(1, cam.y_res // 8)
    x0, y0 = cam.x_res // 2 - w // 2, cam.y_res // 2 - h // 2
    mr, spp, rays, dt, build_s = cpu_render(scene, 8, (x0, x0 + w, y0, y0 + h), budget_s, threads)
    return {"value": round(mr, 3), "unit": "Mrays/s", "cores": threads, "kind": "port", "cpu": cpu_model(),
            "build": "g++ -O3 -march=native, one std::thread per core",
            "sample": f"{w}x{h} centre crop of the same frame, {spp} spp, depth 8, {rays} rays in {dt:.1f} s "
                      f"(FP32 restatement of the reference estimator over the reference-faithful depth-18 / "
                      f"14-bin SAH binary BVH; tree build {build_s:.1f} s not timed)"}


def c0_bench(with_cpu=True, reps=20):
    """c0, the CPU-runnable forward configuration: the 64-sphere field
    (61,442 tris) at 256x256, 16 spp, depth 5, constant environment 0.5, 2
    point lights.  Wavefront and plain renders of the frame (rays credited
    as for the heightfield: the faithful query count), their relMSE against
    the 16,384-spp FP64-accumulated oracle reference
    (tests/golden/ref_c0_sphere_field_256.npz) beside the relMSE the sample
    variance predicts at 16 spp, the traversal alone on 2^20 random rays +
    the frame's 256x256x16 camera rays (GPU closest hit vs the
    reference-faithful CPU traversal, BVH.hpp:120-157), and the CPU
    render baseline of the whole frame."""
    import eleven
    import numpy as np
    from eleven import scenegen
    sys.path.insert(0, str(ROOT / "tests"))
    import raygen
    d = scenegen.cached("sphere_field", n_spheres=64, res=(256, 256), seed=0)
    scene = eleven.SceneData.load(d)
    ctx = eleven.Context(0)
    ctx.upload(scene)
    spp, depth = 16, 5
    ctx.clear()
    logical = ctx.render(spp, depth=depth, seed=0, flags=eleven.FLAG_NO_CULL).rays
    out = {"metric": "Mrays/s", "unit": "Mrays/s",
           "config": "sphere field 61,442 tris, 256x256, 16 spp, depth 5, constant environment 0.5, 2 point lights",
           "rays_per_frame": logical}
    ref = np.load(ROOT / "tests" / "golden" / "ref_c0_sphere_field_256.npz")
    R, V = ref["mean"].astype(np.float64), ref["var"].astype(np.float64)
    w = 1.0 / (R ** 2 + 1e-2)
    for name, mode in (("wavefront", eleven.MODE_WAVEFRONT), ("plain", eleven.MODE_PLAIN)):
        ctx.clear()
        ctx.render(spp, depth=depth, seed=0, mode=mode)                 # warm
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.clear()
            ctx.render(spp, depth=depth, seed=0, mode=mode)
        wall = (time.perf_counter() - t0) / reps
        img, cnt, _, _ = ctx.read()
        n = np.maximum(cnt, 1)[..., None].astype(np.float64)
        rel = float(np.mean((img[..., :3] - R) ** 2 * w))
        pred = float(np.mean((V / n + V / ref["count"][..., None]) * w))
        out[name] = {"value": round(logical / wall / 1e6, 2), "ms_per_frame": round(wall * 1e3, 3),
                     "relmse": rel, "relmse_predicted": pred}
    out["value"] = out["wavefront"]["value"]
    out["wavefront_vs_plain"] = round(out["wavefront"]["value"] / out["plain"]["value"], 3)
    # traversal alone: the same ray batch on the GPU and the CPU
    rays = np.concatenate([raygen.random_rays(scene.vertices(), 1 << 20, seed=1),
                           raygen.camera_rays(scene.camera, spp=16, seed=2)])
    n = rays.shape[0]
    dr, dtri, dt = ctx.alloc(rays.nbytes), ctx.alloc(4 * n), ctx.alloc(4 * n)
    ctx.htod(dr, rays)
    ms = ctx.trace_bench(dr, n, dtri, dt, 20)
    for p_ in (dr, dtri, dt):
        ctx.free(p_)
    out["traversal"] = {"rays": n, "gpu_mrays": round(n / (ms * 1e-3) / 1e6, 1), "gpu_ms": round(ms, 4),
                        "batch": "2^20 random rays (origins in the scene box grown 20 %, uniform directions) + "
                                 "the 256x256x16 camera rays"}
    ctx.close()
    if with_cpu:
        threads = cpu_threads()
        orc = native_oracle()
        rb = orc.RefBVH(scene.vertices())
        rb.closest(rays[:4096], threads=threads)
        t0 = time.perf_counter()
        rb.closest(rays, threads=threads)
        dt = time.perf_counter() - t0
        out["traversal"]["cpu_baseline"] = {
            "value": round(n / dt / 1e6, 2), "unit": "Mrays/s", "cores": threads, "kind": "port", "cpu": cpu_model(),
            "sample": f"the same {n} rays through the reference-faithful depth-18 / 14-bin SAH binary BVH "
                      f"(44 B nodes, BVH.hpp:120-157, closest by t) in {dt:.2f} s"}
        mr, cspp, crays, cdt, _ = cpu_render(scene, depth, (0, 256, 0, 256), 10.0, threads)
        out["cpu_baseline"] = {"value": round(mr, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
                               "cpu": cpu_model(), "build": "g++ -O3 -march=native, one std::thread per core",
                               "sample": f"the full 256x256 frame at {cspp} spp, depth 5, {crays} rays in "
                                         f"{cdt:.1f} s (FP32 estimator over the reference-faithful BVH)"}
    return out


def crop_relmse(ctx, scene, args):
    """relMSE of the last timed frame on the three fixed 64x64 crops against
    their 16,384-spp FP64-accumulated oracle references
    (tests/golden/ref_c1_heightfield_crops.npz), beside the relMSE the
    per-sample variance predicts at the frame's sample count."""
    import numpy as np
    ref = np.load(ROOT / "tests" / "golden" / "ref_c1_heightfield_crops.npz")
    img, cnt, _, _ = ctx.read()
    H = scene.camera.y_res
    meas, pred = [], []
    for k, (x0, x1, y0, y1) in enumerate(ref["windows"]):
        g = img[H - y1:H - y0, x0:x1, :3].astype(np.float64)
        n = np.maximum(cnt[H - y1:H - y0, x0:x1], 1)[..., None].astype(np.float64)
        R, V = ref["mean"][k].astype(np.float64), ref["var"][k].astype(np.float64)
        w = 1.0 / (R ** 2 + 1e-2)
        meas.append(float(np.mean((g - R) ** 2 * w)))
        pred.append(float(np.mean((V / n + V / ref["count"][k][..., None]) * w)))
    return {"measured": [round(x, 6) for x in meas], "predicted": [round(x, 6) for x in pred],
            "windows": ref["windows"].tolist(), "reference": f"{int(ref['spp'])}-spp oracle (FP32 shading, FP64 sums)"}


def adjoint_bench(steps=5, with_cpu=True):
    """Gradients per second of the inverse-rendering step on the gradient-check
    scene: 64-sphere field (61,442 tris) at 128x128, 32 spp, depth 4; the
    gradient covers 65 materials x 6 parameters + a 256x256 RGB albedo texture
    (196,608 texels).  One gradient = forward render (wavefront) + adjoint
    (path replay) of an L2 loss against a target rendered with perturbed
    parameters; parameter and texture uploads included.  The headline is the
    O(pixels) replay (nothing of a path kept between forward and adjoint);
    `recorded` is the same step with EL_FLAG_RECORD_PATHS (every vertex of
    every path kept by the forward render), reported beside it."""
    import eleven
    from eleven import scenegen
    from eleven.autograd import Renderer, material_params
    import numpy as np
    d = scenegen.cached("sphere_field", n_spheres=64, res=(128, 128), seed=0, ground_tex_size=256)
    scene = eleven.SceneData.load(d)
    ctx = eleven.Context(0)
    ctx.upload(scene)
    p0 = material_params(scene)
    t0 = scene.texture(0).copy()
    rng = np.random.default_rng(0)
    pt = np.clip(p0 + rng.normal(0, 0.1, p0.shape).astype(np.float32), 0.0, 1.0)
    pt[:, 3] = np.maximum(pt[:, 3], 0.05)
    legs = {}
    for record in (False, True):
        r = Renderer(ctx, scene, spp=32, depth=4, seed=1, grad_texture=0, record=record)
        target = r.render(pt, t0)
        img = r.render(p0, t0)
        ctx.adjoint(2 * (img - target), grad_texture=0)      # warm-up
        fwd_ms = adj_ms = 0.0
        t_start = time.perf_counter()
        for _ in range(steps):
            # forward with the new parameters and texture, the L2 residual formed
            # on the device (el_residual_l2), then the adjoint of that residual
            r.render(p0, t0, read=False)
            st = ctx.stats()
            fwd_ms += st.render_ms
            ctx.residual_l2(target)
            dmat, dtex = ctx.adjoint(None, grad_texture=0)
            adj_ms += ctx.stats().adjoint_ms
        wall = time.perf_counter() - t_start
        sa = ctx.stats()
        legs[record] = {"value": round(steps / wall, 3), "unit": "grads/s", "ms_per_grad": round(wall / steps * 1e3, 2),
                        "forward_ms": round(fwd_ms / steps, 3), "adjoint_ms": round(adj_ms / steps, 3),
                        "adjoint_rays": int(sa.adjoint_rays),
                        "adjoint_mem_mb": round(sa.adjoint_bytes / 2 ** 20, 1),
                        "record_mem_mb": round(st.record_bytes / 2 ** 20, 1),
                        "grad_norm": float(np.linalg.norm(dmat))}
    img = ctx.read()[0][..., :3]
    out = {"metric": "adjoint grads/s", "value": legs[False]["value"], "unit": "grads/s",
           "config": "sphere field 61,442 tris, 128x128, 32 spp, depth 4, gradient of 65 materials x 6 "
                     "parameters + 256x256 RGB albedo texture (L2 loss)",
           "method": "path replay, O(pixels): the adjoint re-walks every path from its RNG counters (traced "
                     "once, replayed from the thread's own vertex slots); nothing of a path kept after the forward",
           **{k: v for k, v in legs[False].items() if k not in ("value", "unit", "record_mem_mb")},
           "recorded": {**legs[True], "method": "EL_FLAG_RECORD_PATHS: the forward keeps every path's radiance "
                                                "and vertices (O(pixels x spp x depth)); the adjoint traces nothing"}}
    ctx.close()
    if with_cpu:
        orc = native_oracle()
        threads = cpu_threads()
        rb = orc.RefBVH(scene.vertices())
        w = (2 * (img - target)).astype(np.float32)
        t0c = time.perf_counter()
        n = 0
        while n < 50 and time.perf_counter() - t0c < 8.0:
            orc.adjoint(scene, rb, w, 32, depth=4, seed=1, precision=1, grad_texture=0, threads=threads)
            n += 1
        dt = time.perf_counter() - t0c
        out["cpu_baseline"] = {"value": round(n / dt, 4), "unit": "grads/s", "cores": threads, "kind": "port",
                               "cpu": cpu_model(), "build": "g++ -O3 -march=native, one std::thread per core",
                               "sample": f"FP64 oracle adjoint (primal walk + replay) of the same loss and "
                                         f"configuration, {n} gradients in {dt:.2f} s"}
    return out


TEX_BLOCKS = ((0, 64), (80, 32), (150, 16))   # c4 texture: coarse to fine (first step, texel block)
INVERSE_EPS_REL = 0.3                          # c4 parameters' Adam epsilon / RMS of the first gradient
INVERSE_TARGET_SPP = 64                        # samples per pixel of the c4 target images


def inverse_bench(steps=200, views=8, res=512, spp=16, depth=6, eps_rel=INVERSE_EPS_REL, target_spp=INVERSE_TARGET_SPP):
    """Multi-view inverse rendering at scale: the 1,048,576-triangle
    heightfield at 512x512 seen from 8 views, 16 spp forward + 16 spp replay
    per view, depth 6; Adam (lr 1e-2) on 16 materials x 6 parameters and the
    1024x1024 albedo map of material 0 (the other 15 materials have scalar
    albedo) from a seeded start: parameters perturbed by N(0, 0.1^2), texture
    flat 0.5.  Targets: the true parameters and texture rendered at 64 spp.
    The views survey the terrain (eleven.inverse.survey_views: a 4 x 2 grid
    of cameras 22 units up, pitched 60 degrees, each over one column of the
    4 x 4 material tiles and half the depth), so every material is seen by
    62k-143k pixels and material 0's tile by 118k (the default frame's orbit
    of 8 views saw material 0 in 9k pixels of one view and each material in
    at most 12 of 16 views' frames).  A 512-pixel view resolves ~16 texels of
    the 1024^2 map per pixel, so the texture is optimized coarse to fine as
    block values (64-texel blocks from step 0, 32 from 80, 16 from 150;
    eleven.inverse.optimize).  One gradient = all 8 views (forward + adjoint
    each); grads/s over the Adam steps, parameter and texel RMSE before and
    after (texels also as 32x32-texel block means).  The material
    parameters' Adam epsilon is 0.3 x the RMS of the first gradient, so
    parameters the images barely see (specular of rough dielectrics) do not
    random-walk on gradient noise.  Gradients are
    decorrelated (eleven.inverse): each view is rendered twice per step,
    once for the residual and once for the paths the adjoint replays."""
    import eleven
    import numpy as np
    from eleven import inverse, scenegen
    d = scenegen.cached("heightfield", res=(res, res), textured=1)
    scene = eleven.SceneData.load(d)
    ctx = eleven.Context(0)
    ctx.upload(scene)
    cams = inverse.survey_views(scene.camera, res=(res, res))[:views]
    mv = inverse.MultiView(ctx, scene, cams, spp=spp, depth=depth, grad_texture=0)
    truth = inverse.material_params(scene)
    tex_truth = scene.texture(0).copy()
    targets = [im.astype(np.float64) for im in mv.images(truth, tex_truth, seed=777, spp=target_spp)]
    rng = np.random.default_rng(3)
    p0 = np.clip(truth + rng.normal(0, 0.1, truth.shape).astype(np.float32), inverse.PARAM_LO, inverse.PARAM_HI)
    t0 = np.full_like(tex_truth, 0.5)
    _, g0, _ = mv.loss_and_grad(p0, t0, targets, seed=1)
    res_ = inverse.optimize(mv, p0, t0, targets, steps=steps, lr=1e-2, log=log, tex_blocks=TEX_BLOCKS,
                            eps_rel=eps_rel)
    mask = np.ones(truth.shape, bool)
    mask[0, :3] = False   # material 0's scalar albedo is replaced by its texture

    def blocks(t, k=32):   # the texture at a resolution the 512x512 views resolve
        h, w, c = t.shape
        return t.reshape(h // k, k, w // k, k, c).mean(axis=(1, 3))

    out = {"metric": "inverse grads/s", "value": round(res_.grads_per_s, 3), "unit": "grads/s",
           "config": f"heightfield 1,048,576 tris, {res}x{res}, {views} survey views, {spp} spp forward + {spp} "
                     f"spp replay, depth {depth}, Adam lr 1e-2, 16 materials x 6 parameters + 1024x1024 albedo "
                     f"(texel blocks {TEX_BLOCKS})",
           "steps": steps, "adam_eps_rel": eps_rel, "target_spp": target_spp,
           "ms_per_grad": round(res_.seconds / steps * 1e3, 2),
           "loss_first5": float(np.mean(res_.losses[:5])), "loss_last5": float(np.mean(res_.losses[-5:])),
           "param_rmse": {"init": round(inverse.param_rmse(p0, truth, mask), 5),
                          "final": round(inverse.param_rmse(res_.params, truth, mask), 5)},
           "param_rmse_by_kind": {name: {"init": round(inverse.param_rmse(p0[:, cols], truth[:, cols], mask[:, cols]), 5),
                                         "final": round(inverse.param_rmse(res_.params[:, cols], truth[:, cols],
                                                                           mask[:, cols]), 5)}
                                  for name, cols in (("albedo", slice(0, 3)), ("roughness", slice(3, 4)),
                                                     ("metallic", slice(4, 5)), ("specular", slice(5, 6)))},
           "param_rmse_gradient_weighted": {"init": round(inverse.param_rmse(p0, truth, mask, g0), 5),
                                            "final": round(inverse.param_rmse(res_.params, truth, mask, g0), 5)},
           "specular": {"truth": [round(float(x), 3) for x in truth[:, 5]],
                        "final": [round(float(x), 3) for x in res_.params[:, 5]]},
           "metallic": {"truth": [round(float(x), 3) for x in truth[:, 4]],
                        "final": [round(float(x), 3) for x in res_.params[:, 4]]},
           "texel_rmse": {"init": round(inverse.param_rmse(t0, tex_truth), 5),
                          "final": round(inverse.param_rmse(res_.texture, tex_truth), 5)},
           "texture_rmse_32x32_block_means": {"init": round(inverse.param_rmse(blocks(t0), blocks(tex_truth)), 5),
                                              "final": round(inverse.param_rmse(blocks(res_.texture),
                                                                                blocks(tex_truth)), 5)}}
    ctx.close()
    return out


def batch_bench(with_plain=True, depth=6):
    """Batched forward: 32 seeded sphere-field scenes (triangles log-uniform
    1k-500k, square resolutions 128-1024, spp 4-64, depth 6, each with its
    own camera and 512x256 HDR sky) through one graph-captured wavefront loop
    whose launch parameters come from the library's autotuner; baseline =
    the plain megakernel run one scene at a time with the same random
    numbers (rng_scene = scene index), so the images agree to FP32 rounding
    and the two rates are at equal relMSE."""
    import eleven
    import numpy as np
    from eleven import scenegen
    t0 = time.perf_counter()
    work = scenegen.batch_workload(32, seed=0)
    b = eleven.Batch(0)
    for s, spp, seed in work:
        b.add(s, spp, seed)
    log(f"batch: 32 scenes, {sum(s.tri_count for s, _, _ in work)} tris, built in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    cfg = b.tune(depth=depth, trial_spp=2)
    tune_s = time.perf_counter() - t0
    sf = b.render(depth=depth, cull=False, **cfg)   # faithful: the reference's query count
    logical, nocull_ms = sf.rays, sf.render_ms
    b.render(depth=depth, graph=True, **cfg)        # capture + warm
    t0 = time.perf_counter()
    st = b.render(depth=depth, graph=True, **cfg)
    wall = time.perf_counter() - t0
    rays = logical
    out = {"metric": "batched Mrays/s", "value": round(rays / wall / 1e6, 2), "unit": "Mrays/s",
           "config": "32 sphere-field scenes, 2,650,624 tris total (1k-500k each, log-uniform), 128-1024 px square, "
                     "4-64 spp, depth 6, own camera and 512x256 HDR sky each; one hipGraph-captured wavefront loop",
           "rays": rays, "rays_traced": st.rays, "ms": round(wall * 1e3, 2), "device_ms": round(st.render_ms, 2),
           "no_cull": {"value": round(logical / (nocull_ms * 1e-3) / 1e6, 2), "device_ms": round(nocull_ms, 2)},
           "tuned": cfg, "tune_s": round(tune_s, 2), "tuner_log": b.log().strip().splitlines()}
    if with_plain:
        ctx = eleven.Context(0)
        prays, pms, rel_mse, nrel = 0, 0.0, 0.0, 0
        for i, (s, spp, seed) in enumerate(work):
            ctx.upload(s)
            t1 = time.perf_counter()
            sp = ctx.render(spp, depth=depth, seed=seed, mode=eleven.MODE_PLAIN, rng_scene=i,
                            flags=eleven.FLAG_NO_CULL)
            pms += (time.perf_counter() - t1) * 1e3
            prays += sp.rays
            if i < 4:   # same random numbers: the images agree up to FP32 rounding
                a = b.read(i)[0][..., :3].astype(np.float64)
                pimg = ctx.read()[0][..., :3].astype(np.float64)
                rel_mse += float(np.mean((a - pimg) ** 2 / (pimg ** 2 + 1e-2)))
                nrel += 1
        ctx.close()
        pv = prays / (pms * 1e-3) / 1e6
        out["plain"] = {"value": round(pv, 2), "unit": "Mrays/s", "ms": round(pms, 2), "rays": prays,
                        "scenes": "one at a time, plain megakernel (every path traced to the end, as the "
                                  "reference), same seeds and RNG streams",
                        "relmse_batch_vs_plain_first4": rel_mse / max(1, nrel)}
        out["speedup_vs_plain"] = round(out["value"] / pv, 3)
    b.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mode", choices=["wavefront", "plain"], default="wavefront")
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--depth", type=int, default=8)
    ap.add_argument("--bvh", type=int, default=0, help="BVH width 4 or 8 (0: library default)")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--queue", type=int, default=0)
    ap.add_argument("--refill", type=int, default=0)
    ap.add_argument("--res", default="1920x1080")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-count", action="store_true")
    ap.add_argument("--save", default="")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-c0", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-inverse", action="store_true")
    ap.add_argument("--inverse-steps", type=int, default=200)
    ap.add_argument("--inverse-eps-rel", type=float, default=INVERSE_EPS_REL)
    ap.add_argument("--inverse-target-spp", type=int, default=INVERSE_TARGET_SPP)
    ap.add_argument("--copy-bw", action="store_true",
                    help="only measure the achievable copy bandwidth (JSON line); the roofline leg runs this "
                         "in a child process started before this process touches the GPU")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (3 launches): the reference "
                         "kernel for scaling FETCH_SIZE / WRITE_SIZE counter readings")
    args = ap.parse_args()

    if args.copy_bw:
        print(json.dumps(copy_bandwidth()))
        return

    if args.copy_calib:
        import torch
        n = args.copy_calib * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(3):
            torch.mul(x, 1.0, out=y)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calib_bytes_read": n * 4, "copy_calib_bytes_written": n * 4, "launches": 3}))
        return

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU sharding of this path is "
              f"a design note, not a measured configuration.", file=sys.stderr)
        sys.exit(2)

    # torch and the library each load a HIP runtime: the copy kernel runs in
    # its own process, before this one initialises the GPU
    copy_gbs, resident = None, {}
    if args.mode == "wavefront" and not args.no_count:
        r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--copy-bw"], capture_output=True, text=True,
                           timeout=300)
        if r.returncode != 0:
            raise RuntimeError(f"copy bandwidth run failed: {r.stderr[-2000:]}")
        cb = json.loads(r.stdout.strip().splitlines()[-1])
        copy_gbs, resident = cb["copy_gbs"], cb["resident_gbs"]
        log(f"copy kernel: {copy_gbs:.0f} GB/s; cache-resident copies {resident}")

    ensure_built()
    import eleven
    from eleven import scenegen

    res = tuple(int(x) for x in args.res.split("x"))
    t0 = time.perf_counter()
    d = scenegen.cached("heightfield", res=res)
    log(f"scene {d} ready in {time.perf_counter() - t0:.1f} s")
    scene = eleven.SceneData.load(d)
    ctx = eleven.Context(0)
    t0 = time.perf_counter()
    ctx.upload(scene)
    st0 = ctx.stats()
    log(f"upload (BVH build + copy) {time.perf_counter() - t0:.2f} s, {scene.tri_count} tris")
    mode = eleven.MODE_WAVEFRONT if args.mode == "wavefront" else eleven.MODE_PLAIN
    kw = dict(depth=args.depth, mode=mode, bvh_width=args.bvh, block=args.block, queue_cap=args.queue,
              refill=args.refill)

    for w in range(args.warmup):
        ctx.clear()
        st = ctx.render(args.spp, seed=1000 + w, **kw)
        log(f"warmup {w}: {st.render_ms:.1f} ms, {st.rays / st.render_ms / 1e3:.1f} Mrays/s")

    # The rays credited are the reference estimator's queries for the frame:
    # camera + extension + both shadow queries of every path, dead or not.
    # A faithful render (every path traced to the end, as the reference
    # does) of the same frame counts them; the timed renders stop paths
    # whose throughput is exactly zero (same image, fewer rays traced).
    ctx.clear()
    sf = ctx.render(args.spp, seed=0, flags=eleven.FLAG_NO_CULL, **kw)
    logical = sf.rays
    log(f"faithful (no culling): {sf.render_ms:.1f} ms, {logical} rays")
    traced = 0
    ext_ms = 0.0
    ext_launches = ext_rays = 0
    t0 = time.perf_counter()
    for k in range(args.steps):
        ctx.clear()
        # per-launch HIP events on each launch's stream (the roofline's
        # average extension time; ~0.1 ms of a 100 ms frame)
        st = ctx.render(args.spp, seed=0, flags=eleven.FLAG_KERNEL_TIMES, **kw)
        traced += st.rays
        ext_ms += st.extend_ms
        ext_launches += st.extend_launches
        ext_rays += st.extend_rays
        log(f"step {k}: {st.render_ms:.1f} ms, {st.rays} rays traced")
    wall = time.perf_counter() - t0
    rays = logical * args.steps
    value = rays / wall / 1e6
    st = ctx.stats()
    crops = crop_relmse(ctx, scene, args) if (args.res, args.depth) == ("1920x1080", 8) else None
    width = st.bvh_width
    if args.save:
        from eleven import imageio
        img = ctx.read()[0]
        imageio.write_png(args.save, imageio.to_ldr(img))

    out = {"metric": "Mrays/s all bounces (wavefront vs plain, equal relMSE); adjoint grads/s",
           "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 2), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"heightfield {scene.tri_count} tris, {res[0]}x{res[1]}, {args.spp} spp, "
                                  f"depth {args.depth}, defocus f/0.05, 16 textured materials, 8 point lights, "
                                  f"2048x1024 HDR sky",
                      "renderer": args.mode, "bvh_width": width, "rays_per_step": logical,
                      "rays_traced_per_step": traced // max(1, args.steps),
                      "rays_credited": "reference estimator's queries (camera + extension + HDRI and light shadow "
                                       "for every path), counted by a faithful render of the same frame",
                      "no_cull": {"value": round(logical / (sf.render_ms * 1e-3) / 1e6, 2),
                                  "ms_per_step": round(sf.render_ms, 2)}}}

    # roofline of the dominant kernel: closest-hit extension (wavefront only)
    if mode == eleven.MODE_WAVEFRONT and not args.no_count and ext_launches:
        ctx.clear()
        sc = ctx.render(args.spp, seed=0, flags=eleven.FLAG_COUNT_TRAVERSAL, **kw)
        per_frame_launches = sc.extend_launches
        s_node, s_tri = sc.node_stride, sc.tri_stride   # bytes per node / triangle record as stored
        avg_ms = ext_ms / ext_launches
        rays_per_launch = ext_rays / ext_launches
        # algorithmic HBM bytes of one launch: its queue entries read (32 B)
        # and hit records written (16 B) per ray, plus the node array and the
        # triangle stream read once
        algo = rays_per_launch * (32 + 16) + st.node_bytes + st.tri_bytes
        achieved = algo / (avg_ms * 1e-3) / 1e9
        # the bytes the traversal visits (every node fetch and triangle
        # test), served mostly by the caches: against cache-resident copies
        visit = (sc.node_fetches * s_node + sc.tri_tests * s_tri) / max(1, per_frame_launches)
        visit_gbs = visit / (avg_ms * 1e-3) / 1e9
        ws_mib = (st.node_bytes + st.tri_bytes) / 2 ** 20
        defaults = (args.block, args.queue, args.refill, args.spp, args.depth, args.res) == (0, 0, 0, 64, 8, "1920x1080")
        traffic = measured_traffic() if defaults and width == 4 else None
        out["roofline"] = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                           "traffic_over_algorithmic": round(traffic / algo, 2) if traffic else None,
                           "traffic_gbs": round(traffic / (avg_ms * 1e-3) / 1e9, 1) if traffic else None,
                           "traffic_frac": round(traffic / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
                           "copy_gbs": round(copy_gbs, 1) if copy_gbs else None,
                           "traffic_source": "profiles/hbm_traffic_heightfield_1080p_64spp.md" if traffic else None,
                           "kernel": "k_wf_extend (closest-hit traversal)",
                           "avg_launch_ms": round(avg_ms, 4), "launches_per_step": per_frame_launches,
                           "rays_per_launch": int(rays_per_launch), "algo_bytes_per_launch": int(algo),
                           "cache": {"visit_bytes_per_launch": int(visit), "achieved_gbs": round(visit_gbs, 1),
                                     "resident_copy_gbs": resident,
                                     "working_set_mib": round(ws_mib, 1),
                                     "frac_of_working_set_copy": round(visit_gbs / resident["72.7"], 4)
                                     if resident else None,
                                     "node_fetches_per_step": sc.node_fetches, "tri_tests_per_step": sc.tri_tests,
                                     "bytes_per_node": s_node, "bytes_per_tri": s_tri}}
        out["config"]["extend_ms_per_step"] = round(ext_ms / args.steps, 2)
        out["config"]["shade_ms_per_step"] = round(st.shade_ms, 2)
        out["config"]["shadow_ms_per_step"] = round(st.shadow_ms, 2)
        # the shading kernel (memory-bound): the bytes each queue entry needs
        # -- a hit: its ray 32 + hit record 16 + throughput 16 + shading record
        # 128 + HDRI sample entry 32 + texel 4 + radiance read/write 32, and
        # writes the next ray 32 + throughput 16 + two shadow entries 96; an
        # escape: ray 32 + hit 16 + throughput 16 + environment texel 12 +
        # radiance 32 -- against the counted traffic of the same frame
        hits, escapes = st.vertices, max(0, st.extend_rays - st.vertices)
        shade_algo = hits * 404 + escapes * 108
        shade_traffic = measured_traffic("k_wf_shade", "frame") if defaults and width == 4 else None
        out["shade"] = {"kernel": "k_wf_shade", "ms_per_step": round(st.shade_ms, 3),
                        "hits": hits, "escapes": escapes, "algo_bytes_per_step": int(shade_algo),
                        "achieved_gbs": round(shade_algo / (st.shade_ms * 1e-3) / 1e9, 1),
                        "frac_of_copy": round(shade_algo / (st.shade_ms * 1e-3) / 1e9 / copy_gbs, 3)
                        if copy_gbs else None,
                        "traffic_per_step": shade_traffic,
                        "traffic_over_algorithmic": round(shade_traffic / shade_algo, 3) if shade_traffic else None}

    if not args.no_plain and mode == eleven.MODE_WAVEFRONT:
        # the reference's structure: one thread per pixel, every path traced to the end
        ctx.clear()
        t1 = time.perf_counter()
        sp = ctx.render(args.spp, seed=0, **{**kw, "mode": eleven.MODE_PLAIN, "flags": eleven.FLAG_NO_CULL})
        pw = time.perf_counter() - t1
        # the same megakernel with the exact zero-throughput stop
        ctx.clear()
        t1 = time.perf_counter()
        spc = ctx.render(args.spp, seed=0, **{**kw, "mode": eleven.MODE_PLAIN})
        pwc = time.perf_counter() - t1
        out["plain"] = {"value": round(sp.rays / pw / 1e6, 2), "ms_per_step": round(pw * 1e3, 2),
                        "rays_per_step": sp.rays, "wavefront_speedup": round(value / (sp.rays / pw / 1e6), 3),
                        "with_cull": {"value": round(sp.rays / pwc / 1e6, 2), "ms_per_step": round(pwc * 1e3, 2),
                                      "rays_traced": spc.rays}}

    if crops:
        out["config"]["crop_relmse"] = crops
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(scene)
    ctx.close()
    if not args.no_c0:
        log("c0 ...")
        out["c0"] = c0_bench(with_cpu=not args.no_cpu)
    if not args.no_adjoint:
        log("adjoint ...")
        out["adjoint"] = adjoint_bench(with_cpu=not args.no_cpu)
    if not args.no_batch:
        log("batch ...")
        out["batch"] = batch_bench(with_plain=not args.no_plain)
    if not args.no_inverse:
        log("inverse ...")
        out["inverse"] = inverse_bench(steps=args.inverse_steps, eps_rel=args.inverse_eps_rel,
                                       target_spp=args.inverse_target_spp)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
