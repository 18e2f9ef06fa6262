"""The C-ABI library loads and exports every entry point include/mlem.h
declares; GPU entry points fail cleanly (no crash, no CPU fallback) when no
GPU is visible."""
from __future__ import annotations

import ctypes
import re

import numpy as np
import pytest

from conftest import ROOT
import mlem_amd
from mlem_amd import MlemError
from mlem_amd._native import LIBMLEM, LIBSYNTH, MLEM_EDEVICE, SIGNATURES, lib


def _declared(header: str) -> set[str]:
    text = (ROOT / "include" / header).read_text()
    return set(re.findall(r"MLEM_API\s+[\w\s\*]+?\b(mlem_\w+)\s*\(", text))


def test_header_declares_what_bindings_bind():
    declared = _declared("mlem.h")
    assert declared == set(SIGNATURES), declared ^ set(SIGNATURES)


def test_library_exports_every_declared_symbol():
    so = ctypes.CDLL(str(LIBMLEM))
    for name in sorted(_declared("mlem.h")):
        assert hasattr(so, name), name


def test_synth_library_loads():
    so = ctypes.CDLL(str(LIBSYNTH))
    text = (ROOT / "mlem_amd" / "csrc" / "synth.h").read_text()
    for name in re.findall(r"\b(mlem_synth_\w+)\s*\(", text):
        assert hasattr(so, name), name


def test_version():
    assert lib().mlem_version().decode().startswith("mlem-amd")


def _gpu_visible() -> bool:
    import os
    return os.path.exists("/dev/kfd")


@pytest.mark.skipif(_gpu_visible(), reason="checks the no-GPU error path")
def test_no_gpu_error(golden_tiny):
    ref, rowidx, el = golden_tiny
    with pytest.raises(MlemError) as e:
        mlem_amd.Csr4Matrix.from_arrays(3, 4, rowidx, el)
    assert e.value.code == MLEM_EDEVICE
    assert "no GPU" in str(e.value)
    with pytest.raises(MlemError) as e:
        mlem_amd.calibration_copy(1 << 20)
    assert e.value.code == MLEM_EDEVICE


def test_bad_host_rowidx_rejected_before_device(golden_tiny):
    ref, rowidx, el = golden_tiny
    bad = rowidx.copy()
    bad[-1] = 5
    with pytest.raises(MlemError) as e:
        mlem_amd.Csr4Matrix.from_arrays(3, 4, bad, el)
    assert "nnz=6" in str(e.value)
    bad[-1] = 7
    with pytest.raises(MlemError) as e:
        mlem_amd.Csr4Matrix.from_arrays(3, 4, bad, el)
    assert "non-decreasing" in str(e.value)
