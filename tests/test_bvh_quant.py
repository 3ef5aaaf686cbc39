"""Host check of the quantized wide-node encoding (bvh.cpp quantizeNode):
decoded planes are exact FP32 values and quantized boxes contain the FP32
boxes they replace, so the device slab test stays conservative."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "tfg-pathtracer_amd" / "csrc"


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_quantized_nodes_exact_and_conservative(tmp_path):
    exe = tmp_path / "qnode_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", f"-I{CSRC}",
                    str(ROOT / "tests" / "cpp" / "qnode_check.cpp"), str(CSRC / "bvh.cpp"),
                    "-o", str(exe), "-lpthread"], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "bad4=0 bad8=0" in r.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_wide_bvh_covers_every_triangle(tmp_path):
    """Stored wide nodes decoded as the traversal decodes them: leaves cover
    the leaf order exactly once, every triangle lies in each ancestor child
    box, empty slots never hit, and the triangle stream matches the order."""
    exe = tmp_path / "wide_bvh_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", f"-I{CSRC}",
                    str(ROOT / "tests" / "cpp" / "wide_bvh_check.cpp"), str(CSRC / "bvh.cpp"),
                    "-o", str(exe), "-lpthread"], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "bad=0" in r.stdout
