"""The headless CLI (`tfg-pathtracer_amd/bin/eleven <scene> <spp> <out>`,
the reference's render-and-save flow, main.cpp:111-194): its BMP output is
the library's render of the same scene, seed and depth, tone-mapped as the
reference saves it (clamp, fastPow(x, 1/2.2) * 255 truncated)."""
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

import eleven
from eleven import imageio, scenegen

CLI = Path(__file__).resolve().parents[1] / "tfg-pathtracer_amd" / "bin" / "eleven"


def _read_bmp24(path):
    b = Path(path).read_bytes()
    assert b[:2] == b"BM"
    off, w, h, bpp = struct.unpack_from("<I", b, 10)[0], *struct.unpack_from("<ii", b, 18), struct.unpack_from("<H", b, 28)[0]
    assert bpp == 24
    stride = (w * 3 + 3) & ~3
    rows = np.frombuffer(b, np.uint8, stride * h, off).reshape(h, stride)[:, :3 * w].reshape(h, w, 3)
    return rows[::-1, :, ::-1]   # bottom-up BGR -> top-down RGB


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["wavefront", "plain"])
def test_cli_writes_the_library_render(gpu_ctx, scene_root, mode):
    print(f"[cli] {mode}", flush=True)
    assert CLI.exists(), "bin/eleven is not built"
    d = scenegen.sphere_field(scene_root / "cli", n_spheres=12, res=(48, 32), seed=5)
    out = scene_root / f"cli_{mode}.bmp"
    r = subprocess.run([str(CLI), str(d), "4", str(out), "--mode", mode, "--depth", "4", "--seed", "3"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Mrays/s" in r.stdout and "saved" in r.stdout
    got = _read_bmp24(out)
    scene = eleven.SceneData.load(d)
    gpu_ctx.upload(scene)
    gpu_ctx.render(4, depth=4, seed=3, mode=eleven.MODE_PLAIN if mode == "plain" else eleven.MODE_WAVEFRONT)
    want = imageio.to_ldr(gpu_ctx.read()[0])[..., :3]
    assert got.shape == want.shape
    assert np.array_equal(got, want)
    bad = subprocess.run([str(CLI), str(scene_root / "missing"), "1", str(out)], capture_output=True, text=True,
                         timeout=60)
    assert bad.returncode != 0 and "failed" in bad.stderr
