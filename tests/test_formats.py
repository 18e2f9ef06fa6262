"""csr4 / raw-vector formats and the C ABI's host-only paths (no GPU).

Reference behaviour: Csr4Matrix::Csr4Matrix (helper_files_common/
csr4matrix.cpp:32-147) and Vector<T> (helper_files_common/vector.hpp:33-87).
"""
from __future__ import annotations

import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

from conftest import GOLDEN, ROOT
import mlem_amd
from mlem_amd import MlemError, synth
from mlem_amd._native import MLEM_EFORMAT, MLEM_EIO

CLI = ROOT / "mlem_amd" / "bin" / "mlem_cli"
GOLD = (GOLDEN / "tiny3x4.csr4").read_bytes()


def test_probe_golden():
    info = mlem_amd.probe(str(GOLDEN / "tiny3x4.csr4"))
    assert info == {"rows": 3, "cols": 4, "nnz": 6, "grid": (2, 2, 1), "symflags": 0}


def test_writer_matches_hand_built_bytes(tmp_path, golden_tiny):
    ref, rowidx, el = golden_tiny
    m = synth.SynthMatrix(3, 4, 6, (2, 2, 1), rowidx, el)
    p = tmp_path / "w.csr4"
    m.write_csr4(str(p))
    assert p.read_bytes() == GOLD


def test_synth_file_roundtrip(tmp_path, tiny_problem):
    m = tiny_problem.matrix
    p = tmp_path / "t.csr4"
    m.write_csr4(str(p))
    info = mlem_amd.probe(str(p))
    assert (info["rows"], info["cols"], info["nnz"], info["grid"]) == (m.rows, m.cols, m.nnz, m.grid)
    raw = p.read_bytes()
    (scb,) = struct.unpack_from("<I", raw, 24)
    off = 28 + scb
    rowidx = np.frombuffer(raw, np.uint64, m.rows, off)
    el = np.frombuffer(raw, np.uint64, m.nnz, off + 8 * m.rows)
    assert np.array_equal(rowidx, m.rowidx)
    assert np.array_equal(el, m.elements.view(np.uint64))


def _bad(tmp_path, data: bytes) -> Path:
    p = tmp_path / "bad.csr4"
    p.write_bytes(data)
    return p


@pytest.mark.parametrize("mutate,code,msg", [
    (lambda b: b[:20], MLEM_EFORMAT, "File too short"),
    (lambda b: b"XCSR" + b[4:], MLEM_EFORMAT, "is not a CSR matrix"),
    (lambda b: b[:4] + b"\x03" + b[5:], MLEM_EFORMAT, "has wrong CSR version (need v4)"),
    (lambda b: b[:-1], MLEM_EFORMAT, "is too short"),
    (lambda b: b[:24] + struct.pack("<I", 56) + b[28:], MLEM_EFORMAT,
     "Cannot find scanner geometry configuration"),
    # scanner block claims 64 bytes but its contents parse as 60
    (lambda b: b[:24] + struct.pack("<I", 64) + b[28:88] + b"\0\0\0\0" + b[88:], MLEM_EFORMAT,
     "Error reading scanner geometry configuration"),
    # 1000 rings: the gap list runs past the declared block
    (lambda b: b[:28 + 32] + struct.pack("<I", 1000) + b[28 + 36:], MLEM_EFORMAT,
     "Error reading scanner geometry configuration"),
    # row END offsets not non-decreasing
    (lambda b: b[:88] + struct.pack("<3Q", 2, 1, 6) + b[112:], MLEM_EFORMAT, "non-decreasing"),
    # last END offset != nnz
    (lambda b: b[:88] + struct.pack("<3Q", 2, 3, 5) + b[112:], MLEM_EFORMAT, "nnz=6"),
])
def test_header_errors(tmp_path, mutate, code, msg):
    p = _bad(tmp_path, mutate(GOLD))
    with pytest.raises(MlemError) as e:
        mlem_amd.probe(str(p))
    assert e.value.code == code
    assert msg in str(e.value)


def test_missing_file():
    with pytest.raises(MlemError) as e:
        mlem_amd.probe("/nonexistent/file.csr4")
    assert e.value.code == MLEM_EIO
    assert "Cannot open file /nonexistent/file.csr4" in str(e.value)


def test_64bit_size_check(tmp_path):
    # nnz large enough that the reference's uint32 size check would wrap
    # (csr4matrix.cpp:85-86) and accept a short file; here it is rejected.
    nnz = (1 << 32) // 8 + 6
    data = GOLD[:16] + struct.pack("<Q", nnz) + GOLD[24:]
    p = _bad(tmp_path, data)
    with pytest.raises(MlemError) as e:
        mlem_amd.probe(str(p))
    assert "is too short" in str(e.value)


def test_cli_argument_errors():
    r = subprocess.run([str(CLI), "a", "b", "c"], capture_output=True, text=True)
    assert r.returncode == 1 and "wrong number of command line parameters" in r.stderr
    r = subprocess.run([str(CLI), "a", "b", "c", "x1"], capture_output=True, text=True)
    assert r.returncode == 1 and "invalid iteration count" in r.stderr


def test_cli_missing_input(tmp_path):
    r = subprocess.run([str(CLI), str(GOLDEN / "tiny3x4.csr4"), str(tmp_path / "none.sino"),
                        str(tmp_path / "o.img"), "2"], capture_output=True, text=True)
    assert r.returncode == 1 and "Cannot open file" in r.stderr
    assert "Matrix file: " in r.stdout and "Iterations: 2" in r.stdout


def test_cli_bad_matrix(tmp_path):
    p = _bad(tmp_path, b"XCSR" + GOLD[4:])
    r = subprocess.run([str(CLI), str(p), str(GOLDEN / "tiny3x4.sino"), str(tmp_path / "o.img"),
                        "2"], capture_output=True, text=True)
    assert r.returncode == 1 and "is not a CSR matrix" in r.stderr


def test_raw_sinogram_size_check(tmp_path):
    # a sinogram whose byte size is not a multiple of 4 cannot be opened
    # (vector.hpp:43-44)
    s = tmp_path / "odd.sino"
    s.write_bytes(b"\0" * 7)
    r = subprocess.run([str(CLI), str(GOLDEN / "tiny3x4.csr4"), str(s), str(tmp_path / "o.img"),
                        "1"], capture_output=True, text=True)
    assert r.returncode == 1 and "Cannot open file" in r.stderr
