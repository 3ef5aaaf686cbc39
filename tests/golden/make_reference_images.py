"""Writes tests/golden/ref_<scene>.npz: FP64 reference images of the small
render-parity scenes, produced by the CPU oracle with the reference's own
semantics (HDRI binarySearch sampling, HDRI.hpp:130-162; point-light
visibility by the closest hit's shading position, kernel.cu:193-197; FP64
shading and accumulation) at high sample counts.

Each file holds `mean` and `var` (per-pixel sample variance) as float32 [H, W, 3]
plus `spp`.  Run from the repository root:
    python tests/golden/make_reference_images.py [--spp 4096]
The scenes are regenerated from their seeds by tests/render_scenes.py.
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--spp", type=int, default=4096)
    ap.add_argument("--scenes", default=",".join(render_scenes.SCENES))
    ap.add_argument("--out", default=str(HERE))
    args = ap.parse_args()
    root = Path("/tmp/eleven_golden_scenes")
    for kind in args.scenes.split(","):
        s = render_scenes.make_scene(kind, root)
        ref = orc.RefBVH(s.vertices())
        t0 = time.time()
        osum, osq, cnt, rays, _ = orc.render(s, ref, args.spp, depth=render_scenes.DEPTH, seed=12345, precision=1,
                                             shadow_mode=0)
        n = np.maximum(cnt, 1)[..., None].astype(np.float64)
        mean = osum / n
        var = np.maximum(osq / n - mean ** 2, 0) * n / np.maximum(n - 1, 1)
        np.savez_compressed(Path(args.out) / f"ref_{kind}.npz", mean=mean.astype(np.float32),
                            var=var.astype(np.float32), count=cnt.astype(np.uint32), spp=args.spp)
        print(f"{kind}: {args.spp} spp, {rays} rays, {time.time() - t0:.1f} s", flush=True)


if __name__ == "__main__":
    main()
