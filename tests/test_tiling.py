"""Host-side box tiling of the batched products (mlem_amd/csrc/panels.cpp):
compiles tests/tiling_check.cpp against panels.cpp with g++ and runs it (CPU
only). The program checks the invariants the GPU kernels rely on and emulates
the forward and backward products over the tiling against dense products."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "mlem_amd" / "csrc"


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_box_tiling_host_checks(tmp_path):
    exe = tmp_path / "tiling_check"
    subprocess.run(["g++", "-O2", "-std=c++17", f"-I{CSRC}", f"-I{ROOT / 'include'}",
                    str(ROOT / "tests" / "tiling_check.cpp"), str(CSRC / "panels.cpp"),
                    "-o", str(exe), "-lpthread"], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    lines = [l for l in out.splitlines() if l.startswith("ok ")]
    assert len(lines) == 6, out
