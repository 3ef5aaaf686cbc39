"""The host code under AddressSanitizer and UndefinedBehaviorSanitizer.

tests/cpp/host_sanitize.cpp is built with g++ -fsanitize=address,undefined
(-fno-sanitize-recover=all: any finding aborts) together with the library's
host sources -- the scene loader, MikkTSpace tangents, BVH build and
quantization, HDRI sampler tables and the scene C ABI -- and run on three
generated scenes (a small sphere field with an 8-bit .bmp ground map, a
heightfield with 8-bit texture maps and an RGBE .hdr sky, a scene of mixed
polygon faces) and on damaged copies of the first (files truncated, bytes
overwritten, numbers made huge, negative or malformed).  Every load must
either succeed or report an error, the BVHs must be permutations of the
triangles, the HDRI lookup must equal the reference search, and the
sanitizers must stay silent.
"""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

from eleven import scenegen

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "tfg-pathtracer_amd" / "csrc"
HOST_SRC = ["scene_loader.cpp", "tangent_space.cpp", "bvh.cpp", "hdri_table.cpp", "scene_capi.cpp"]


def _polygon_scene(path):
    """Quads, a pentagon and negative (relative) indices: the loader fans
    polygons and resolves relative indices."""
    path.mkdir(parents=True, exist_ok=True)
    obj = ["o poly\n", "v 0 0 0\n", "v 1 0 0\n", "v 1 1 0\n", "v 0 1 0\n", "v 0.5 1.5 0\n",
           "vt 0 0\n", "vt 1 0\n", "vt 1 1\n", "vt 0 1\n", "vt 0.5 1.5\n",
           "vn 0 0 1\n", "usemtl m\n",
           "f 1/1/1 2/2/1 3/3/1 4/4/1\n", "f 1/1/1 2/2/1 3/3/1 5/5/1 4/4/1\n", "f -5/-5/-1 -4/-4/-1 -3/-3/-1\n",
           "f 1 2 3\n", "f 1//1 2//1 4//1\n"]
    (path / "scene.obj").write_text("".join(obj))
    scenegen.write_mtl(path / "scene.mtl", [{"name": "m"}])
    scenegen.write_json(path / "scene.json", {"xRes": 8, "yRes": 8}, {"color": {"r": 1, "g": 1, "b": 1}}, [])
    return path


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_code_under_asan_and_ubsan(tmp_path):
    exe = tmp_path / "host_sanitize"
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-ffp-contract=off", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", f"-I{CSRC}",
           str(ROOT / "tests" / "cpp" / "host_sanitize.cpp")] + [str(CSRC / f) for f in HOST_SRC] + \
          ["-o", str(exe), "-lpthread"]
    b = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    if b.returncode != 0 and "asan" in b.stderr.lower() and "cannot find" in b.stderr.lower():
        pytest.skip("this toolchain has no sanitizer runtime")
    assert b.returncode == 0, b.stderr[-4000:]
    scenes = [   # the first is the one damaged: small, so its ~200 loads stay quick
        scenegen.sphere_field(tmp_path / "sf", n_spheres=3, res=(8, 8), seed=3, seg=8, rings=6, ground_tex_size=16),
        scenegen.heightfield(tmp_path / "hf", cells=(32, 16), extent=(6.4, 3.2), res=(16, 9), n_mats=3,
                             tex_size=32, sky=(64, 32)),
        _polygon_scene(tmp_path / "poly"),
    ]
    # verify_asan_link_order=0: the ASan runtime need not be the first
    # library of the process (the environment may preload others)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([str(exe)] + [str(s) for s in scenes] + ["--", str(tmp_path / "scratch")],
                       capture_output=True, text=True, timeout=900, env=env)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-6000:]
    assert "bad=0" in r.stdout
    assert "ERROR: AddressSanitizer" not in out and "runtime error" not in out
