"""Synthetic-input generator: Siddon lengths against geometry, determinism,
csr4 invariants, Poisson moments (CPU)."""
from __future__ import annotations

import numpy as np
import pytest

from mlem_amd import synth


def _clip_length(p0, p1, grid):
    """Length of the segment inside the centred box (slab method)."""
    p0, p1 = np.asarray(p0, float), np.asarray(p1, float)
    d = p1 - p0
    lo, hi = -0.5 * np.asarray(grid, float), 0.5 * np.asarray(grid, float)
    t0, t1 = 0.0, 1.0
    for a in range(3):
        if d[a] == 0:
            if not lo[a] < p0[a] < hi[a]:
                return 0.0
            continue
        u, v = sorted(((lo[a] - p0[a]) / d[a], (hi[a] - p0[a]) / d[a]))
        t0, t1 = max(t0, u), min(t1, v)
    return max(0.0, t1 - t0) * float(np.linalg.norm(d))


def _voxels_by_sampling(p0, p1, grid, n=200001):
    p0, p1 = np.asarray(p0, float), np.asarray(p1, float)
    t = np.linspace(0, 1, n)[:, None]
    pts = p0 + t * (p1 - p0) + 0.5 * np.asarray(grid, float)
    ijk = np.floor(pts).astype(np.int64)
    ok = np.all((ijk >= 0) & (ijk < np.asarray(grid)), axis=1)
    ijk = ijk[ok]
    return set((ijk[:, 0] + grid[0] * (ijk[:, 1] + grid[1] * ijk[:, 2])).tolist())


def test_axis_aligned_line():
    cols, lens = synth.siddon([-10, 0.5, 0.25], [10, 0.5, 0.25], (4, 4, 2))
    # y = 0.5 -> voxel row 2 (of 0..3), z = 0.25 -> layer 1
    assert list(cols) == [x + 4 * (2 + 4 * 1) for x in range(4)]
    assert np.allclose(lens, 1.0)


def test_diagonal_line():
    cols, lens = synth.siddon([-2, -2, 0.5], [2, 2, 0.5], (4, 4, 2))
    assert list(cols) == [0 + 4 * 0 + 16, 1 + 4 * 1 + 16, 2 + 4 * 2 + 16, 3 + 4 * 3 + 16]
    assert np.allclose(lens, np.sqrt(2.0))


def test_miss_returns_nothing():
    cols, _ = synth.siddon([-10, 5, 0], [10, 5, 0], (4, 4, 2))
    assert cols.size == 0


@pytest.mark.parametrize("seed", range(20))
def test_random_lines_against_geometry(seed):
    rng = np.random.default_rng(seed)
    grid = (7, 5, 6)
    p0 = rng.uniform(-6, 6, 3)
    p1 = rng.uniform(-6, 6, 3)
    cols, lens = synth.siddon(p0, p1, grid)
    L = _clip_length(p0, p1, grid)
    assert abs(float(lens.astype(np.float64).sum()) - L) < 1e-5 * max(L, 1)
    assert np.all(np.diff(cols.astype(np.int64)) > 0)
    assert np.all(lens > 0) and np.all(lens <= np.sqrt(3) + 1e-6)
    sampled = _voxels_by_sampling(p0, p1, grid)
    # every voxel crossed by more than a sliver is found by dense sampling
    big = set(cols[lens > 1e-3].tolist())
    assert big <= sampled and sampled <= set(cols.tolist())


def test_matrix_deterministic_and_thread_independent():
    a = synth.make_matrix((8, 8, 8), 3000, 77, threads=1)
    b = synth.make_matrix((8, 8, 8), 3000, 77, threads=4)
    assert a.nnz == b.nnz
    assert np.array_equal(a.rowidx, b.rowidx)
    assert np.array_equal(a.elements.view(np.uint64), b.elements.view(np.uint64))
    c = synth.make_matrix((8, 8, 8), 3000, 78)
    assert c.nnz != a.nnz or not np.array_equal(c.elements.view(np.uint64), a.elements.view(np.uint64))


def test_matrix_invariants(tiny_problem):
    m = tiny_problem.matrix
    rb = m.row_begin().astype(np.int64)
    re = m.rowidx.astype(np.int64)
    assert np.all(re - rb >= 1)              # every line meets the grid
    col = m.elements["column"].astype(np.int64)
    val = m.elements["value"]
    assert col.max() < m.cols
    assert np.all(val > 0) and np.all(val <= np.float32(np.sqrt(3)))
    inner = np.ones(col.size - 1, bool)
    inner[rb[1:] - 1] = False                 # boundaries between rows
    assert np.all(np.diff(col)[inner] > 0)    # ascending inside every row


def test_small_workload_size():
    # the 32^3 workload has ~10M elements (about 37 per line)
    m = synth.make_matrix((32, 32, 32), 16384, 1000)
    assert 30 < m.nnz / m.rows < 45


def test_phantom():
    x = synth.make_phantom((16, 16, 16), 4, 5, (2, 5))
    assert x.shape == (4096,)
    assert x.min() >= 0 and 0.5 <= x.max() <= 4 * 4
    s2 = synth.make_phantom((16, 16, 16), 4, 5, (2, 5), scale=np.full(4, 2.0))
    assert np.allclose(s2, 2 * x)


@pytest.mark.parametrize("lam", [0.3, 4.0, 12.0, 75.0, 2000.0])
def test_poisson_moments(lam):
    n = 200000
    y = synth.poisson(np.full(n, lam), 0.0, 123)
    m, v = y.mean(), y.var()
    se = np.sqrt(lam / n)
    assert abs(m - lam) < 6 * se
    assert abs(v - lam) < 0.05 * lam + 6 * se
    assert np.array_equal(y, synth.poisson(np.full(n, lam), 0.0, 123, threads=1))


def test_poisson_total_scaling():
    ybar = np.linspace(0.1, 3, 10000)
    y = synth.poisson(ybar, 1e6, 9)
    assert abs(y.sum() - 1e6) < 6 * np.sqrt(1e6)


def test_frame_parameters():
    c = synth.frame_counts(64)
    assert c.min() >= 1e5 and c.max() <= 1e8
    b = synth.frame_budgets(64)
    assert b.min() >= 10 and b.max() <= 100
    s = synth.frame_scales(3, 16)
    assert all(0.2 <= v.min() and v.max() < 2.0 for v in s)
