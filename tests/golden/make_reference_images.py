"""Writes FP64-accumulated reference images produced by the CPU oracle with
the reference's own semantics (HDRI binarySearch sampling, HDRI.hpp:130-162;
point-light visibility by the closest hit's shading position,
kernel.cu:193-197) and the reference's FP32 shading arithmetic, summed in
FP64, at high sample counts.

The shading precision matters at the 1e-3 level: evaluated in FP64 the same
estimator is 0.17 % brighter on the 64-sphere field (the oracle at 128 spp
with common random numbers, FP32 vs FP64 shading; some paths take another
discrete branch -- a grazing hit, the smooth-shading position test, a lobe
threshold -- and those branches do not cancel).  The reference renders in
FP32, so its images are the FP32 estimator's.

Small render-parity scenes (tests/render_scenes.py), 4096 spp:
    tests/golden/ref_<scene>.npz
Workload references at 16384 spp (the scene sizes of the forward and
adjoint workloads):
    ref_c0_sphere_field_256.npz   64-sphere field (61,442 tris), 256x256,
                                  depth 5, full frame
    ref_c3_sphere_field_128.npz   the same field with its 256x256 ground
                                  texture at 128x128, depth 4, full frame
    ref_c1_heightfield_crops.npz  the 1,048,576-tri heightfield at 1920x1080,
                                  depth 8, defocus: three fixed 64x64 crops

Each file holds `mean` (float32) and `var` (per-pixel sample variance of one
clamped sample, float16 for the workload references) [H, W, 3], `count`
(accepted samples) and `spp`; crop files also `windows` (x0, x1, y0, y1 in the
kernel's pixel coordinates, output rows H-y1 .. H-y0-1).  Run from the
repository root:
    python tests/golden/make_reference_images.py [--spp 4096]
    python tests/golden/make_reference_images.py --configs c0,c3,c1 [--spp 16384]
The scenes are regenerated from their seeds (tests/render_scenes.py,
eleven.scenegen); partial sums are checkpointed under /tmp so an
interrupted run resumes.
"""
import argparse
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent / "tfg-pathtracer_amd"))
sys.path.insert(0, str(HERE.parent))

import oracle_py as orc          # noqa: E402
import render_scenes             # noqa: E402

SEED = 12345                     # RNG key of every reference (the GPU tests use other keys)

# Fixed 64x64 crops of the 1920x1080 heightfield frame (kernel x0, x1, y0, y1;
# kernel row y is output row H-1-y): foreground terrain, the centre of the
# frame, and terrain below the horizon on the right.
C1_WINDOWS = [(448, 512, 180, 244), (928, 992, 508, 572), (1472, 1536, 760, 824)]


def c_scene(kind):
    import eleven
    from eleven import scenegen
    if kind == "c0":
        d = scenegen.cached("sphere_field", n_spheres=64, res=(256, 256), seed=0)
        return eleven.SceneData.load(d), 5, None
    if kind == "c3":
        d = scenegen.cached("sphere_field", n_spheres=64, res=(128, 128), seed=0, ground_tex_size=256)
        return eleven.SceneData.load(d), 4, None
    if kind == "c1":
        d = scenegen.cached("heightfield", res=(1920, 1080))
        return eleven.SceneData.load(d), 8, C1_WINDOWS
    raise ValueError(kind)


FILES = {"c0": "ref_c0_sphere_field_256.npz", "c3": "ref_c3_sphere_field_128.npz",
         "c1": "ref_c1_heightfield_crops.npz"}


def stats(osum, osq, cnt):
    n = np.maximum(cnt, 1)[..., None].astype(np.float64)
    mean = osum / n
    var = np.maximum(osq / n - mean ** 2, 0) * n / np.maximum(n - 1, 1)
    return mean, var


def small_scenes(args):
    root = Path("/tmp/eleven_golden_scenes")
    for kind in args.scenes.split(","):
        s = render_scenes.make_scene(kind, root)
        ref = orc.RefBVH(s.vertices())
        t0 = time.time()
        osum, osq, cnt, rays, _ = orc.render(s, ref, args.spp, depth=render_scenes.DEPTH, seed=SEED, precision=0,
                                             shadow_mode=0)
        mean, var = stats(osum, osq, cnt)
        np.savez_compressed(Path(args.out) / f"ref_{kind}.npz", mean=mean.astype(np.float32),
                            var=var.astype(np.float32), count=cnt.astype(np.uint32), spp=args.spp)
        print(f"{kind}: {args.spp} spp, {rays} rays, {time.time() - t0:.1f} s", flush=True)


def workload(kind, args):
    scene, depth, windows = c_scene(kind)
    cam = scene.view.camera
    W, H = cam.x_res, cam.y_res
    ref = orc.RefBVH(scene.vertices())
    ck = Path(f"/tmp/eleven_golden_fp32_{kind}_{args.spp}.npz")
    osum, osq = np.zeros((H, W, 3)), np.zeros((H, W, 3))
    cnt = np.zeros((H, W), np.uint64)
    done, rays = 0, 0
    if ck.exists():
        z = np.load(ck)
        osum, osq, cnt, done, rays = z["sum"], z["sumsq"], z["count"], int(z["done"]), int(z["rays"])
    t0 = time.time()
    while done < args.spp:
        k = min(args.chunk, args.spp - done)
        for win in (windows or [None]):
            s, ss, c, r, _ = orc.render(scene, ref, k, depth=depth, seed=SEED, s0=done, precision=0, shadow_mode=0,
                                        threads=args.threads, window=win)
            osum += s
            osq += ss
            cnt += c
            rays += r
        done += k
        np.savez(ck, sum=osum, sumsq=osq, count=cnt, done=done, rays=rays)
        print(f"{kind}: {done}/{args.spp} spp, {rays} rays, {time.time() - t0:.0f} s", flush=True)
    mean, var = stats(osum, osq, cnt)
    out = {"spp": args.spp, "depth": depth, "seed": SEED, "rays": rays}
    if windows:
        rows = [(H - y1, H - y0) for (_, _, y0, y1) in windows]
        out["mean"] = np.stack([mean[r0:r1, x0:x1] for (r0, r1), (x0, x1, _, _) in zip(rows, windows)])
        out["var"] = np.stack([var[r0:r1, x0:x1] for (r0, r1), (x0, x1, _, _) in zip(rows, windows)])
        out["count"] = np.stack([cnt[r0:r1, x0:x1] for (r0, r1), (x0, x1, _, _) in zip(rows, windows)])
        out["windows"] = np.array(windows, np.int32)
    else:
        out["mean"], out["var"], out["count"] = mean, var, cnt
    out["mean"] = out["mean"].astype(np.float32)
    out["var"] = out["var"].astype(np.float16)
    out["count"] = out["count"].astype(np.uint32)
    np.savez_compressed(Path(args.out) / FILES[kind], **out)
    print(f"{kind}: written {FILES[kind]}", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--spp", type=int, default=0)
    ap.add_argument("--scenes", default=",".join(render_scenes.SCENES))
    ap.add_argument("--configs", default="", help="workload references to write: c0,c3,c1")
    ap.add_argument("--chunk", type=int, default=512, help="samples per checkpoint")
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--out", default=str(HERE))
    args = ap.parse_args()
    if args.configs:
        args.spp = args.spp or 16384
        for kind in args.configs.split(","):
            workload(kind, args)
    else:
        args.spp = args.spp or 4096
        small_scenes(args)


if __name__ == "__main__":
    main()
