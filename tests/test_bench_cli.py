"""bench.py's command-line contract that needs no GPU: one GPU only."""
from __future__ import annotations

import subprocess
import sys

from conftest import ROOT


def test_more_than_one_gpu_is_refused_with_a_message():
    p = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr and "--gpus 2" in p.stderr
    assert p.stdout.strip() == ""
