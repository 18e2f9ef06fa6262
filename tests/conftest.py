"""Shared fixtures. Tests marked `gpu` need an MI355X; the rest run on CPU."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD MI355X (gfx950) GPU")
    # Build the native libraries once if this tree has none yet (a fresh
    # checkout); a tree that has them uses what is there.
    lib = ROOT / "mlem_amd" / "lib"
    need = [lib / "libmlem.so", lib / "libmlem_synth.so", ROOT / "mlem_amd" / "bin" / "mlem_cli"]
    if not all(p.exists() for p in need):
        from mlem_amd.build import build_all
        build_all(verbose=False)
    if not (ROOT / "oracle" / "lib" / "liboracle.so").exists():
        from oracle import oracle as _o
        _o.build()


@pytest.fixture(scope="session")
def tiny_problem():
    from mlem_amd import synth
    return synth.make_problem("tiny")


@pytest.fixture(scope="session")
def golden_tiny():
    import json
    ref = json.loads((GOLDEN / "tiny3x4.json").read_text())
    rowidx = np.array(ref["rowidx"], np.uint64)
    el = np.zeros(ref["nnz"], dtype=[("column", "<u4"), ("value", "<f4")])
    el["column"] = [c for c, _ in ref["elements"]]
    el["value"] = [v for _, v in ref["elements"]]
    return ref, rowidx, el


def rel_linf(a, b, floor_frac=1e-6):
    """max_j |a_j - b_j| / max(|b_j|, floor_frac * max|b|)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    floor = floor_frac * max(float(np.max(np.abs(b))), 1e-300)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)))
