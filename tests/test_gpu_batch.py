"""Batched (many-sinogram) reconstruction on the GPU against the oracle.

EXACT mode: every frame's image after k iterations must be bit-identical to k
iterations of the fp32 restatement of seqcsr4mlem. FAST mode (the matrix's
two-dimensional tiling): every frame within 1e-4 relative of k iterations of
FP64 MLEM, deterministic run to run. Both: the per-frame FP64 log-likelihood
history within 1e-6 of the FP64 oracle's.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rel_linf
import mlem_amd as M
from mlem_amd import synth
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def make_frames(p, frames, lo=2e4, hi=2e6):
    """Per-frame sinograms: the workload's phantom with per-ellipsoid scales
    and log-uniform total counts."""
    w = synth.WORKLOADS[p.name]
    ys = []
    counts = synth.frame_counts(frames, lo, hi)
    for f, sc in enumerate(synth.frame_scales(frames, w["ellipsoids"])):
        x = synth.make_phantom(w["grid"], w["ellipsoids"], 2000 + w["key"], w["semi"], scale=sc)
        ybar = synth.project(p.matrix, x)
        ys.append(synth.poisson(ybar, counts[f], 5000 + f))
    return np.stack(ys)


@pytest.fixture(scope="module")
def tiny_batch(tiny_problem):
    A = M.Csr4Matrix.from_arrays(tiny_problem.matrix.rows, tiny_problem.matrix.cols,
                                 tiny_problem.matrix.rowidx, tiny_problem.matrix.elements,
                                 grid=tiny_problem.matrix.grid)
    Y = make_frames(tiny_problem, 11)
    yield tiny_problem, A, Y
    A.close()


@pytest.mark.parametrize("width", [1, 4, 8, 32, 0])
def test_batch_budgets_bitwise(tiny_batch, width):
    p, A, Y = tiny_batch
    m = p.matrix
    budgets = np.array([1, 2, 3, 5, 8, 13, 4, 0, 7, 9, 6], np.int32)
    B = M.Batch(A, Y, budgets, block_width=width, mode="exact")
    n = B.run(100)
    assert n == budgets.max()
    prog = B.progress()
    assert prog["active"] == 0 and prog["frame_iterations"] == budgets.sum()
    for f in range(Y.shape[0]):
        it, ll = B.frame_info(f)
        assert it == budgets[f]
        xo = orc.mlem(m.rowidx, m.elements, m.cols, Y[f], int(budgets[f]))[0]
        assert np.array_equal(B.image(f), xo), f
        if it:
            _, ll64, _ = orc.f64_mlem(m.rowidx, m.elements, m.cols, Y[f].astype(np.float64), it)
            assert rel_linf(ll, ll64) < 1e-6
    B.close()


@pytest.mark.parametrize("width", [1, 8])
def test_batch_early_stop(tiny_batch, width):
    p, A, Y = tiny_batch
    m = p.matrix
    tol = 2e-4
    B = M.Batch(A, Y, np.full(Y.shape[0], 60, np.int32), rel_tol=tol, block_width=width,
                mode="exact")
    B.run(1000)
    for f in range(Y.shape[0]):
        it, ll = B.frame_info(f)
        assert 2 <= it <= 60
        rel = np.abs(np.diff(ll)) / np.abs(ll[1:])
        if it < 60:
            assert rel[-1] <= tol and np.all(rel[:-1] > tol)   # stopped at the first small change
        xo = orc.mlem(m.rowidx, m.elements, m.cols, Y[f], it)[0]
        assert np.array_equal(B.image(f), xo), f
    B.close()


def test_batch_graph_matches_direct(tiny_batch):
    p, A, Y = tiny_batch
    budgets = np.array([3, 6, 2, 5, 4, 1, 6, 6, 3, 2, 5], np.int32)
    a = M.Batch(A, Y, budgets, block_width=4, use_graph=True, mode="exact")
    b = M.Batch(A, Y, budgets, block_width=4, use_graph=False, mode="exact")
    a.run(10)
    b.run(10)
    for f in range(Y.shape[0]):
        assert np.array_equal(a.image(f), b.image(f))
    a.close()
    b.close()


def test_batch_incremental_runs(tiny_batch):
    p, A, Y = tiny_batch
    m = p.matrix
    B = M.Batch(A, Y[:5], None, block_width=4, mode="exact")
    assert B.run(3) == 3
    assert B.run(4) == 4
    for f in range(5):
        assert B.frame_info(f)[0] == 7
        assert np.array_equal(B.image(f), orc.mlem(m.rowidx, m.elements, m.cols, Y[f], 7)[0])
    B.close()


@pytest.mark.parametrize("width", [4, 0])
def test_batch_small_workload(width):
    # 32^3 grid, 262,144 lines: 6 frames, budgets up to 12; width 4 (two
    # blocks) or autotuned (blocks and lone frames by measured cost)
    p = synth.make_problem("small")
    m = p.matrix
    A = M.Csr4Matrix.from_arrays(m.rows, m.cols, m.rowidx, m.elements, grid=m.grid)
    Y = make_frames(p, 6, 1e5, 1e7)
    budgets = np.array([12, 3, 7, 1, 10, 5], np.int32)
    B = M.Batch(A, Y, budgets, block_width=width, mode="exact")
    B.run(50)
    for f in range(6):
        xo = orc.mlem(m.rowidx, m.elements, m.cols, Y[f], int(budgets[f]))[0]
        assert np.array_equal(B.image(f), xo), f
    B.close()
    A.close()


def test_batch_argument_errors(tiny_batch):
    p, A, Y = tiny_batch
    with pytest.raises(ValueError):
        M.Batch(A, Y[:, :-1], mode="exact")
    with pytest.raises(M.MlemError):
        M.Batch(A, Y, block_width=12, mode="exact")
    with pytest.raises(M.MlemError):
        M.Batch(A, Y, np.full(Y.shape[0], -1, np.int32), mode="exact")
    with pytest.raises(ValueError):   # one budget per frame
        M.Batch(A, Y, np.full(Y.shape[0] - 1, 3, np.int32))
    with pytest.raises(KeyError):
        M.Batch(A, Y, mode="atomic-ish")


def cheapest_blocks(ms, frames):
    """Blocks (widths 4..32, or 1: a frame alone) holding `frames` frames at
    the least summed measured cost, widest first, with their cost."""
    best, pick = [0.0] + [None] * frames, [0] * (frames + 1)
    for a in range(1, frames + 1):
        for w in (4, 8, 16, 32):
            c = best[max(0, a - w)] + ms[w]
            if best[a] is None or c < best[a]:
                best[a], pick[a] = c, w
        if best[a - 1] + ms[1] < best[a]:
            best[a], pick[a] = best[a - 1] + ms[1], 1
    ws, a = [], frames
    while a > 0:
        ws.append(pick[a])
        a = max(0, a - pick[a])
    return sorted(ws, reverse=True), best[frames]


def test_batch_autotuned_widths(tiny_batch):
    p, A, Y = tiny_batch
    B = M.Batch(A, Y, np.full(Y.shape[0], 3, np.int32), block_width=0, mode="exact")
    ms = B.tuning()
    assert set(ms) == {1, 4, 8, 16, 32} and all(v > 0 for v in ms.values())
    # the chosen blocks are the cheapest set for the measured costs; no
    # single width does better
    ws, cost = cheapest_blocks(ms, Y.shape[0])
    prog = B.progress()
    assert (prog["block_width"], prog["blocks"]) == (ws[0], len(ws))
    assert all(cost <= -(-Y.shape[0] // w) * ms[w] + 1e-12 for w in ms)
    B.run(10)
    m = p.matrix
    for f in range(Y.shape[0]):
        assert np.array_equal(B.image(f), orc.mlem(m.rowidx, m.elements, m.cols, Y[f], 3)[0])
    assert M.Batch(A, Y, block_width=32, mode="exact").tuning() == {1: 0.0, 4: 0.0, 8: 0.0,
                                                                     16: 0.0, 32: 0.0}


def test_batch_single_frame_blocks_after_launch_change(tiny_batch):
    # width-1 blocks run the matrix's single-image kernels: a launch change
    # between runs re-captures the batch's graph and stays bit-identical
    p, A, Y = tiny_batch
    m = p.matrix
    budgets = np.array([2, 7, 4, 1, 6, 3, 5, 7, 2, 6, 4], np.int32)
    B = M.Batch(A, Y, budgets, block_width=1, mode="exact")
    assert B.progress()["blocks"] == int(np.count_nonzero(budgets))
    B.run(2)
    A.set_launch(("dict", 16), ("tiles", 8))
    B.run(2)
    A.set_launch(("dict512", 16), ("dict512", 16))
    B.run(10)
    A.set_launch(("dict512", 32), ("dict512", 32))
    assert B.progress()["active"] == 0
    for f in range(Y.shape[0]):
        assert B.frame_info(f)[0] == budgets[f]
        xo = orc.mlem(m.rowidx, m.elements, m.cols, Y[f], int(budgets[f]))[0]
        assert np.array_equal(B.image(f), xo), f
    B.close()


# ---- FAST mode: the matrix's two-dimensional tiling --------------------------
TOL_FAST = 1e-4   # relative L-infinity per voxel vs FP64 MLEM (floor 1e-6 max)


def _vs_fp64(m, B, Y, frames=None):
    """Largest relative distance of the frames' images to FP64 MLEM after the
    same number of iterations, and of their log-likelihood histories."""
    worst, worst_ll = 0.0, 0.0
    for f in (range(Y.shape[0]) if frames is None else frames):
        it, ll = B.frame_info(f)
        x64, ll64, _ = orc.f64_mlem(m.rowidx, m.elements, m.cols, Y[f].astype(np.float64), it)
        worst = max(worst, rel_linf(B.image(f), x64))
        if it:
            worst_ll = max(worst_ll, rel_linf(ll, ll64))
    return worst, worst_ll


@pytest.mark.parametrize("width", [1, 4, 8, 16, 32, 0])
def test_fast_batch_budgets_vs_fp64(tiny_batch, width):
    p, A, Y = tiny_batch
    budgets = np.array([1, 2, 3, 5, 8, 13, 4, 0, 7, 9, 6], np.int32)
    B = M.Batch(A, Y, budgets, block_width=width, mode="fast")
    assert B.run(100) == budgets.max()
    assert B.progress()["frame_iterations"] == budgets.sum()
    for f in range(Y.shape[0]):
        assert B.frame_info(f)[0] == budgets[f]
    e, ell = _vs_fp64(p.matrix, B, Y)
    print(f"[fast batch] width {width}: rel Linf vs FP64 {e:.2e}, loglik {ell:.2e}")
    assert e < TOL_FAST and ell < 1e-6
    B.close()


def test_fast_batch_deterministic_graph_and_runs(tiny_batch):
    # the tiled sums are in a fixed order: the same bits run to run, with or
    # without graph replay, in one run or in several
    p, A, Y = tiny_batch
    budgets = np.array([3, 6, 2, 5, 4, 1, 6, 6, 3, 2, 5], np.int32)
    a = M.Batch(A, Y, budgets, block_width=8, use_graph=True)
    b = M.Batch(A, Y, budgets, block_width=8, use_graph=False)
    c = M.Batch(A, Y, budgets, block_width=8, use_graph=True)
    a.run(10)
    b.run(10)
    c.run(2)
    c.run(3)
    c.run(10)
    for f in range(Y.shape[0]):
        assert np.array_equal(a.image(f), b.image(f)), f
        assert np.array_equal(a.image(f), c.image(f)), f
        assert np.array_equal(a.frame_info(f)[1], c.frame_info(f)[1]), f
    for B in (a, b, c):
        B.close()


@pytest.mark.parametrize("width", [1, 8])
def test_fast_batch_early_stop(tiny_batch, width):
    p, A, Y = tiny_batch
    tol = 2e-4
    B = M.Batch(A, Y, np.full(Y.shape[0], 60, np.int32), rel_tol=tol, block_width=width)
    B.run(1000)
    for f in range(Y.shape[0]):
        it, ll = B.frame_info(f)
        assert 2 <= it <= 60
        rel = np.abs(np.diff(ll)) / np.abs(ll[1:])
        if it < 60:
            assert rel[-1] <= tol and np.all(rel[:-1] > tol)
    e, ell = _vs_fp64(p.matrix, B, Y)
    assert e < TOL_FAST and ell < 1e-6
    B.close()


@pytest.mark.parametrize("width", [4, 16, 0])
def test_fast_batch_small_workload(width):
    # 32^3 grid, 262,144 lines (two panels): 9 frames, budgets up to 12;
    # autotuned widths mix tiled blocks and lone frames on the single-image
    # FAST kernels, so frames move between the two column orders
    p = synth.make_problem("small")
    m = p.matrix
    A = M.Csr4Matrix.from_arrays(m.rows, m.cols, m.rowidx, m.elements, grid=m.grid)
    Y = make_frames(p, 9, 1e5, 1e7)
    budgets = np.array([12, 3, 7, 1, 10, 5, 12, 2, 9], np.int32)
    B = M.Batch(A, Y, budgets, block_width=width)
    B.run(50)
    n = M.calcColumnSums(A).astype(np.float64)
    for f in range(Y.shape[0]):
        ysum = float(Y[f].astype(np.float64).sum())
        assert abs(float(n @ B.image(f).astype(np.float64)) - ysum) <= 1e-5 * ysum
    e, ell = _vs_fp64(m, B, Y)
    print(f"[fast batch] small, width {width}: rel Linf vs FP64 {e:.2e}, loglik {ell:.2e}")
    assert e < TOL_FAST and ell < 1e-6
    B.close()
    A.close()


def test_fast_batch_file_order_matrix(tiny_problem):
    # a matrix opened without a grid: panels are runs of columns
    m = tiny_problem.matrix
    A = M.Csr4Matrix.from_arrays(m.rows, m.cols, m.rowidx, m.elements)
    Y = make_frames(tiny_problem, 5)
    B = M.Batch(A, Y, np.array([4, 2, 6, 1, 3], np.int32), block_width=4)
    B.run(10)
    e, ell = _vs_fp64(m, B, Y)
    assert e < TOL_FAST and ell < 1e-6
    B.close()
    A.close()
