"""Out-of-core (streamed) matrices on the GPU against the oracle.

The matrix stays in pinned host memory in row blocks; every projection streams
its part of each block to the device. EXACT results must equal the fp32
restatement of seqcsr4mlem bit for bit for any block split (each column's sum
continues from block to block in ascending row order); FAST and ATOMIC results
(ATOMIC iterations scatter each block while it is on the device) are held to
FP64 like the in-core kernels.
"""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, rel_linf
import mlem_amd as M
from mlem_amd import synth
from mlem_amd._native import MlemError
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

TOL_PROJ = 2e-6
TOL_RECON = 1e-4


@pytest.fixture(scope="module")
def small_streamed():
    p = synth.make_problem("small")
    m = p.matrix
    A = M.Csr4Matrix.from_arrays(m.rows, m.cols, m.rowidx, m.elements, grid=m.grid,
                                 stream_block_bytes=16 << 20)
    yield p, A
    A.close()


def test_blocks_and_fixed_launch(small_streamed):
    p, A = small_streamed
    assert A.stream_blocks() >= 5
    # iterations make one pass: each block's CSC is built on the device
    assert A.stream_passes() == 1
    with pytest.raises(MlemError):
        A.set_launch(("dict512", 32), ("dict512", 32))
    with pytest.raises(MlemError):
        A.autotune()


def test_exact_steps_and_iterates_bitwise(small_streamed):
    p, A = small_streamed
    m = p.matrix
    assert np.array_equal(M.calcColumnSums(A), orc.column_sums(m.rowidx, m.elements, m.cols))
    rng = np.random.default_rng(3)
    x = rng.uniform(0.1, 3.0, m.cols).astype(np.float32)
    r = rng.uniform(0.1, 3.0, m.rows).astype(np.float32)
    assert np.array_equal(M.calcFwProj(A, x, "exact"), orc.fwproj(m.rowidx, m.elements, x))
    assert np.array_equal(M.calcBkProj(A, r, "exact"), orc.bkproj(m.rowidx, m.elements, r, m.cols))
    xg = M.mlem(A, p.y, None, 5, mode="exact")
    assert np.array_equal(xg, orc.mlem(m.rowidx, m.elements, m.cols, p.y, 5)[0])


@pytest.mark.parametrize("mode", ["fast", "atomic"])
def test_fast_and_atomic_vs_fp64(small_streamed, mode):
    p, A = small_streamed
    m = p.matrix
    rng = np.random.default_rng(4)
    x = rng.uniform(0.1, 3.0, m.cols).astype(np.float32)
    r = rng.uniform(0.1, 3.0, m.rows).astype(np.float32)
    assert rel_linf(M.calcFwProj(A, x, "fast"), orc.f64_fwproj(m.rowidx, m.elements, x)) < TOL_PROJ
    assert rel_linf(M.calcBkProj(A, r, mode),
                    orc.f64_bkproj(m.rowidx, m.elements, r, m.cols)) < TOL_PROJ
    K = 10
    xg, ll, tot = M.mlem(A, p.y, None, K, mode=mode, stats=True)
    x64, ll64, _ = orc.f64_mlem(m.rowidx, m.elements, m.cols, p.y.astype(np.float64), K)
    assert rel_linf(xg, x64) < TOL_RECON
    assert rel_linf(ll, ll64) < 1e-6
    ysum = float(p.y.astype(np.float64).sum())
    assert np.all(np.abs(tot[1:] - ysum) <= 1e-5 * ysum)


@pytest.mark.parametrize("resident_mib", [24, 1 << 12])
def test_resident_blocks_exact(resident_mib):
    # some blocks (or all) kept in device memory, the rest streamed: the same
    # bit-exact results
    p = synth.make_problem("small")
    m = p.matrix
    A = M.Csr4Matrix.from_arrays(m.rows, m.cols, m.rowidx, m.elements,
                                 stream_block_bytes=8 << 20,
                                 stream_resident_bytes=resident_mib << 20)
    nb, nr = A.stream_blocks(), A.resident_blocks()
    assert 0 < nr <= nb and (nr < nb) == (resident_mib == 24)
    assert np.array_equal(M.mlem(A, p.y, None, 4, mode="exact"),
                          orc.mlem(m.rowidx, m.elements, m.cols, p.y, 4)[0])
    x64, ll64, _ = orc.f64_mlem(m.rowidx, m.elements, m.cols, p.y.astype(np.float64), 4)
    for mode in ("fast", "atomic"):  # atomic: one fused pass per block
        xf, ll, _ = M.mlem(A, p.y, None, 4, mode=mode, stats=True)
        assert rel_linf(xf, x64) < TOL_RECON and rel_linf(ll, ll64) < 1e-6, mode
    A.close()


def test_recon_incremental_matches_one_shot(small_streamed):
    p, A = small_streamed
    R = M.Recon(A, p.y, mode="exact")
    R.run(2)
    R.run(3)
    assert np.array_equal(R.image(), M.mlem(A, p.y, None, 5, mode="exact"))
    R.close()


def test_timed_run(small_streamed):
    # per-kernel HIP-event timing covers every block's copy and kernels; the
    # timed run makes the two passes (each projection timed), the plain run
    # one pass with each block's CSC built on the device: the same bits
    p, A = small_streamed
    R = M.Recon(A, p.y, mode="exact")
    kf, kb, span = R.run_timed(3)
    assert np.all(kf > 0) and np.all(kb > 0) and span >= kf.sum()
    assert np.array_equal(R.image(), M.mlem(A, p.y, None, 3, mode="exact"))
    R.close()


def test_cli_streamed_checkpoint_restart(tmp_path, tiny_problem):
    # the 5th argument (checkpoint interval) with a streamed matrix: restart
    # from img.chkpt / iter.chkpt gives the uninterrupted result
    p = tiny_problem
    m = p.matrix
    mtx, sino = tmp_path / "m.csr4", tmp_path / "s.sino"
    m.write_csr4(str(mtx))
    p.y.astype(np.int32).tofile(sino)
    env = dict(os.environ, MLEM_MODE="exact", MLEM_STREAM="0")
    cli = str(ROOT / "mlem_amd" / "bin" / "mlem_cli")
    xo3 = orc.mlem(m.rowidx, m.elements, m.cols, p.y, 3)[0]
    xo3.tofile(tmp_path / "img.chkpt")
    np.int32([3]).tofile(tmp_path / "iter.chkpt")
    out = tmp_path / "restart.img"
    r = subprocess.run([cli, str(mtx), str(sino), str(out), "7", "2"], cwd=tmp_path, env=env,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "I got Checkpoint at iteration 3" in r.stdout
    assert np.array_equal(np.fromfile(out, np.float32),
                          orc.mlem(m.rowidx, m.elements, m.cols, p.y, 7)[0])


def test_batched_and_unrolled_refuse_streamed(small_streamed):
    p, A = small_streamed
    with pytest.raises(MlemError):
        M.Batch(A, np.stack([p.y, p.y]), np.array([2, 2], np.int32), block_width=4)
    with pytest.raises(MlemError):
        M.Unrolled(A, p.y[None, :].astype(np.float32), None, 3)


def _ragged(seed=17):
    # rows of 0..200 elements, empty rows and one 3000-element row
    rng = np.random.default_rng(seed)
    cols = 4000
    lens = list(rng.integers(0, 200, 700)) + [0, 0, 3000, 1, 0]
    rows = [np.sort(rng.choice(cols, L, replace=False)) for L in lens]
    ends = np.cumsum(lens).astype(np.uint64)
    el = np.zeros(int(ends[-1]), dtype=M.ELEMENT_DTYPE)
    el["column"] = np.concatenate(rows).astype(np.uint32)
    el["value"] = rng.uniform(0.1, 1.7, el.size).astype(np.float32)
    return len(lens), cols, ends, el, rng


@pytest.mark.parametrize("block_bytes", [4 * 4001 + 16 + 8 * 3000, 100_000, 1 << 20, 0])
def test_ragged_block_splits_exact(block_bytes):
    rows, cols, ends, el, rng = _ragged()
    A = M.Csr4Matrix.from_arrays(rows, cols, ends, el, stream_block_bytes=block_bytes)
    if block_bytes == 4 * 4001 + 16 + 8 * 3000:
        assert A.stream_blocks() >= 20   # about 3000 elements per block
    x = rng.uniform(0.1, 3.0, cols).astype(np.float32)
    r = rng.uniform(0.1, 3.0, rows).astype(np.float32)
    assert np.array_equal(M.calcFwProj(A, x, "exact"), orc.fwproj(ends, el, x))
    assert np.array_equal(M.calcBkProj(A, r, "exact"), orc.bkproj(ends, el, r, cols))
    y = rng.integers(0, 40, rows).astype(np.int32)
    assert np.array_equal(M.mlem(A, y, None, 4, mode="exact"), orc.mlem(ends, el, cols, y, 4)[0])
    xf = M.mlem(A, y, None, 4, mode="fast")
    assert rel_linf(xf, orc.f64_mlem(ends, el, cols, y.astype(np.float64), 4)[0]) < TOL_RECON
    A.close()


def test_block_too_small_and_bad_columns():
    rows, cols, ends, el, _ = _ragged()
    with pytest.raises(MlemError) as e:
        M.Csr4Matrix.from_arrays(rows, cols, ends, el, stream_block_bytes=20_000)
    assert e.value.code == -1   # MLEM_EINVAL: the 3000-element row does not fit
    bad = el.copy()
    bad["column"][5] = cols
    with pytest.raises(MlemError) as e:
        M.Csr4Matrix.from_arrays(rows, cols, ends, bad, stream_block_bytes=1 << 20)
    assert e.value.code == -3   # MLEM_EFORMAT


def test_file_and_cli_streamed(tmp_path, tiny_problem):
    p = tiny_problem
    m = p.matrix
    path = str(tmp_path / "tiny.csr4")
    m.write_csr4(path)
    A = M.Csr4Matrix(path, stream_block_bytes=64 << 10)
    assert A.stream_blocks() > 1
    assert np.array_equal(M.mlem(A, p.y, None, 6, mode="exact"),
                          orc.mlem(m.rowidx, m.elements, m.cols, p.y, 6)[0])
    A.close()
    sino = tmp_path / "tiny.sino"
    p.y.astype(np.int32).tofile(sino)
    outs = []
    for stream in ("", "0"):
        out = tmp_path / f"img{stream or 'hbm'}.raw"
        env = dict(os.environ, MLEM_MODE="exact", MLEM_STREAM=stream)
        subprocess.run([str(ROOT / "mlem_amd" / "bin" / "mlem_cli"), path, str(sino), str(out),
                        "6"], check=True, env=env, capture_output=True, timeout=120)
        outs.append(np.fromfile(out, np.float32))
    assert np.array_equal(outs[0], outs[1])
    assert np.array_equal(outs[1], orc.mlem(m.rowidx, m.elements, m.cols, p.y, 6)[0])
