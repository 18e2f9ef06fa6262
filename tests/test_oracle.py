"""Pins the CPU oracle (oracle/mlem_oracle.c) before anything is compared
against it.

The reference has no tests, fixtures or golden outputs, and its sequential
driver cannot be compiled here (csr4matrix.hpp:16 needs Boost, absent from
this image). Each restated function is therefore pinned by checks that a
plausible mistake fails:
  * exact rational answers of the hand-built fixture (tests/golden);
  * summation order, no fused multiply-add, flush-to-zero, int->float
    conversion and rounding order of seqcsr4mlem.cpp:37-134, each by an input
    where the alternative gives a different fp32 result (found with numpy's
    IEEE float32 arithmetic, independent of the oracle);
  * dense brute force, closed forms (identity, single column, fixed point,
    block-diagonal decoupling) and invariants of MLEM (count preservation,
    monotone likelihood, non-negativity, homogeneity in x0 and y).
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rel_linf
from oracle import oracle as orc

EL = orc.ELEMENT_DTYPE


def csr(rows_list, ncols=None):
    """[(col, val), ...] per row -> (rowidx END offsets, elements)."""
    ends, acc = [], 0
    for r in rows_list:
        acc += len(r)
        ends.append(acc)
    el = np.zeros(acc, EL)
    k = 0
    for r in rows_list:
        for c, v in r:
            el[k] = (c, v)
            k += 1
    return np.array(ends, np.uint64), el


def dense(rowidx, el, ncols):
    A = np.zeros((len(rowidx), ncols))
    b = 0
    for i, e in enumerate(rowidx):
        for k in range(b, int(e)):
            A[i, el["column"][k]] += el["value"][k]
        b = int(e)
    return A


# ----------------------------------------------------- golden fixture --------
def test_golden_exact_values(golden_tiny):
    ref, rowidx, el = golden_tiny
    y = np.array(ref["sino"], np.int32)
    norm = orc.column_sums(rowidx, el, 4)
    assert np.array_equal(norm, np.float32(ref["norm"]))
    x0, v, sn, si = orc.init_image(norm, y)
    assert v == np.float32(ref["x0"]) and sn == 11.5 and si == 10.0
    x = x0
    for it in ref["iterates"]:
        f = orc.fwproj(rowidx, el, x)
        assert rel_linf(f, it["f"]) < 3e-7
        r = orc.correl(f, y)
        assert rel_linf(r, it["r"]) < 6e-7
        u = orc.bkproj(rowidx, el, r, 4)
        x = orc.update(u, norm, x)
        assert rel_linf(x, it["x"]) < 2e-6
    xf, s, z, sub, _ = orc.mlem(rowidx, el, 4, y, 3)
    assert np.array_equal(xf, x)
    x64, _, _ = orc.f64_mlem(rowidx, el, 4, y.astype(np.float64), 3)
    assert rel_linf(x64, ref["iterates"][-1]["x"]) < 1e-14


# ------------------------------------------------------- fp32 semantics ------
def test_forward_sums_in_element_order():
    # x = [-1e8, 1e8, 1, 1]: in order -> 2; reversed -> 0; pairwise -> 2.
    # x = [1, 1e8, -1e8, 1]: in order -> 1; pairwise -> 0.
    rowidx, el = csr([[(0, 1.0), (1, 1.0), (2, 1.0), (3, 1.0)]])
    assert orc.fwproj(rowidx, el, np.float32([-1e8, 1e8, 1, 1]))[0] == 2.0
    assert orc.fwproj(rowidx, el, np.float32([1, 1e8, -1e8, 1]))[0] == 1.0


def test_backward_sums_in_ascending_row_order():
    rowidx, el = csr([[(0, 1.0)], [(0, 1.0)], [(0, 1.0)], [(0, 1.0)]])
    assert orc.bkproj(rowidx, el, np.float32([-1e8, 1e8, 1, 1]), 1)[0] == 2.0
    assert orc.bkproj(rowidx, el, np.float32([1, 1e8, -1e8, 1]), 1)[0] == 1.0
    # column sums accumulate in the same order
    rowidx, el = csr([[(0, v)] for v in (-1e8, 1e8, 1.0, 1.0)])
    assert orc.column_sums(rowidx, el, 1)[0] == 2.0


def test_no_fused_multiply_add():
    a = np.float32(1 + 2.0 ** -12)
    c = -np.float32(1 + 2.0 ** -11)
    # round(a*a) == 1 + 2^-11 (tie to even) so c*1 + a*a == 0 without FMA;
    # an FMA gives 2^-24.
    assert np.float32(a * a) == -c
    rowidx, el = csr([[(0, float(c)), (1, float(a))]])
    assert orc.fwproj(rowidx, el, np.float32([1.0, a]))[0] == 0.0
    rowidx, el = csr([[(0, float(c))], [(0, float(a))]])
    assert orc.bkproj(rowidx, el, np.float32([1.0, a]), 1)[0] == 0.0


def test_flush_to_zero_without_denormals_are_zero():
    rowidx, el = csr([[(0, 1e-20)]])
    assert orc.fwproj(rowidx, el, np.float32([1e-20]))[0] == 0.0   # result flushed
    rowidx, el = csr([[(0, 1e10)]])
    f = orc.fwproj(rowidx, el, np.float32([1e-40]))[0]              # input kept
    assert f == np.float32(1e10) * np.float32(1e-40) and f > 0
    assert orc.correl(np.float32([3e38]), np.int32([1]))[0] == 0.0
    x = orc.update(np.float32([1e-15]), np.float32([1.0]), np.float32([1e-30]))
    assert x[0] == 0.0


def _search(pred, gen, n=200000, seed=1):
    rng = np.random.default_rng(seed)
    for _ in range(n):
        args = gen(rng)
        if pred(*args):
            return args
    raise AssertionError("no distinguishing input found")


def test_correl_converts_count_to_float_first():
    def gen(rng):
        return np.int32(rng.integers(2 ** 24 + 1, 2 ** 30) | 1), np.float32(rng.uniform(1, 100))
    y, f = _search(lambda y, f: np.float32(np.float32(y) / f) != np.float32(np.float64(y) / np.float64(f)), gen)
    r = orc.correl(np.float32([f, 0.0]), np.int32([y, 7]))
    assert r[0] == np.float32(np.float32(y) / f)
    assert r[1] == 0.0


def test_update_rounds_quotient_then_product():
    def gen(rng):
        return tuple(np.float32(v) for v in rng.uniform(0.1, 10, 3))
    x, u, n = _search(lambda x, u, n: np.float32(x * np.float32(u / n)) != np.float32(np.float32(x * u) / n), gen)
    out = orc.update(np.float32([u, u]), np.float32([n, 0.0]), np.float32([x, x]))
    assert out[0] == np.float32(x * np.float32(u / n))
    assert out[1] == np.float32(x * u)        # n == 0: multiply by u


def test_init_image_double_sums():
    norm = np.float32([1e8, 1.0, 1.0, 1.0])
    y = np.int32([2 ** 30, 2 ** 30, 3])
    x, v, sn, si = orc.init_image(norm, y)
    assert sn == 1e8 + 3 and si == 2.0 ** 31 + 3
    assert v == np.float32((2.0 ** 31 + 3) / (1e8 + 3))
    assert np.all(x == v)


# ---------------------------------------------------- brute force ------------
def test_dense_brute_force():
    rng = np.random.default_rng(5)
    A = rng.uniform(0.1, 2, (23, 17)) * (rng.uniform(size=(23, 17)) < 0.4)
    rows = [[(j, float(np.float32(A[i, j]))) for j in range(17) if A[i, j] != 0] for i in range(23)]
    rowidx, el = csr(rows)
    A = dense(rowidx, el, 17)
    x = rng.uniform(0.5, 2, 17)
    r = rng.uniform(0.5, 2, 23)
    assert rel_linf(orc.f64_fwproj(rowidx, el, x), A @ x) < 1e-14
    assert rel_linf(orc.f64_bkproj(rowidx, el, r, 17), A.T @ r) < 1e-14
    assert rel_linf(orc.f64_column_sums(rowidx, el, 17), A.sum(0)) < 1e-14
    assert rel_linf(orc.fwproj(rowidx, el, x.astype(np.float32)), A @ x.astype(np.float32)) < 1e-6
    assert rel_linf(orc.bkproj(rowidx, el, r.astype(np.float32), 17),
                    A.T @ r.astype(np.float32)) < 1e-6
    # one full MLEM iteration against dense numpy
    y = rng.integers(0, 50, 23)
    n = A.sum(0)
    x0 = np.full(17, y.sum() / n.sum())
    f = A @ x0
    rr = np.where(f != 0, y / np.where(f != 0, f, 1), 0)
    u = A.T @ rr
    x1 = x0 * np.where(n != 0, u / np.where(n != 0, n, 1), u)
    xo, _, _ = orc.f64_mlem(rowidx, el, 17, y.astype(np.float64), 1)
    assert rel_linf(xo, x1) < 1e-13


# ------------------------------------------------------- closed forms --------
def test_identity_matrix_gives_y_after_one_iteration():
    y = np.int32([3, 0, 7, 1, 12])
    rowidx, el = csr([[(i, 1.0)] for i in range(5)])
    x64, _, _ = orc.f64_mlem(rowidx, el, 5, y.astype(np.float64), 1)
    assert np.allclose(x64, y, rtol=1e-15, atol=0)
    x32 = orc.mlem(rowidx, el, 5, y, 1)[0]
    assert rel_linf(x32, y) < 3e-7


def test_single_column():
    a = np.float32([0.5, 1.25, 2.0, 0.75])
    y = np.int32([4, 9, 1, 6])
    rowidx, el = csr([[(0, float(v))] for v in a])
    for x0 in (None, np.float64([3.7])):
        x, _, _ = orc.f64_mlem(rowidx, el, 1, y.astype(np.float64), 1, x0)
        assert abs(x[0] - y.sum() / a.astype(np.float64).sum()) < 1e-14 * x[0]


def test_consistent_data_fixed_point_exact_in_fp32():
    rng = np.random.default_rng(11)
    A = rng.integers(0, 3, (40, 9)).astype(np.float32)
    assert np.linalg.matrix_rank(A) == 9
    xs = rng.integers(1, 5, 9).astype(np.float32)
    y = (A @ xs).astype(np.int32)
    rows = [[(j, float(A[i, j])) for j in range(9) if A[i, j]] for i in range(40)]
    rowidx, el = csr(rows)
    x = orc.mlem(rowidx, el, 9, y, 7, x0=xs)[0]
    assert np.array_equal(x, xs)


def test_block_diagonal_decouples_bitwise(tiny_problem):
    p = tiny_problem
    m = p.matrix
    rb = m.row_begin()
    half = m.rows // 2
    A1 = (m.rowidx[:half], m.elements[: int(m.rowidx[half - 1])])
    el2 = m.elements[int(m.rowidx[half - 1]):].copy()
    ri2 = m.rowidx[half:] - m.rowidx[half - 1]
    # block 2 uses columns shifted by cols: diag(A1, A2)
    el2s = el2.copy()
    el2s["column"] += m.cols
    rowidx = np.concatenate([A1[0], ri2 + A1[0][-1]])
    el = np.concatenate([A1[1], el2s])
    y1, y2 = p.y[:half], p.y[half:]
    x0 = np.full(2 * m.cols, 1.0, np.float32)
    xb = orc.mlem(rowidx, el, 2 * m.cols, np.concatenate([y1, y2]), 5, x0=x0)[0]
    xa1 = orc.mlem(A1[0], A1[1], m.cols, y1, 5, x0=x0[: m.cols])[0]
    xa2 = orc.mlem(ri2, el2, m.cols, y2, 5, x0=x0[m.cols:])[0]
    assert np.array_equal(xb, np.concatenate([xa1, xa2]))
    assert rb.size == m.rows


# --------------------------------------------------------- invariants --------
def test_invariants_fp64(tiny_problem):
    p = tiny_problem
    m = p.matrix
    y = p.y.astype(np.float64)
    K = 15
    x, ll, ft = orc.f64_mlem(m.rowidx, m.elements, m.cols, y, K + 1)
    f = orc.f64_fwproj(m.rowidx, m.elements, x)
    # count preservation: sum_i f_i(x_k) = sum_{i: f_i != 0} y_i for k >= 1
    ymask = y[orc.f64_fwproj(m.rowidx, m.elements, np.ones(m.cols)) != 0].sum()
    assert np.all(np.abs(ft[1:] - ymask) < 1e-11 * ymask)
    assert np.all(np.diff(ll) >= -1e-9 * np.abs(ll[:-1]))     # EM: L non-decreasing
    assert np.all(x >= 0)
    # homogeneity: degree 0 in x0, degree 1 in y
    x0 = np.full(m.cols, 0.37)
    xa, _, _ = orc.f64_mlem(m.rowidx, m.elements, m.cols, y, K, x0)
    xb, _, _ = orc.f64_mlem(m.rowidx, m.elements, m.cols, y, K, 5.0 * x0)
    assert rel_linf(xa, xb, 1e-9) < 1e-11
    xc, _, _ = orc.f64_mlem(m.rowidx, m.elements, m.cols, 3.0 * y, K, x0)
    assert rel_linf(xc, 3.0 * xa, 1e-9) < 1e-11
    assert f.shape == (m.rows,)


def test_fp32_restatement_close_to_fp64(tiny_problem):
    p = tiny_problem
    m = p.matrix
    x32 = orc.mlem(m.rowidx, m.elements, m.cols, p.y, 20)[0]
    x64, _, _ = orc.f64_mlem(m.rowidx, m.elements, m.cols, p.y.astype(np.float64), 20)
    assert rel_linf(x32, x64, 1e-4) < 1e-4


def test_openmp_restatement(tiny_problem):
    p = tiny_problem
    m = p.matrix
    xs = orc.mlem(m.rowidx, m.elements, m.cols, p.y, 10)[0]
    xo, t = orc.omp_mlem(m.rowidx, m.elements, m.cols, p.y, 10, threads=4)
    assert rel_linf(xo, xs, 1e-4) < 1e-5
    assert t.shape == (10,)
    # one thread merges nothing: bitwise equal to the sequential restatement
    # up to flush-to-zero, which the OpenMP driver does not set
    x1, _ = orc.omp_mlem(m.rowidx, m.elements, m.cols, p.y, 10, threads=1)
    assert rel_linf(x1, xs, 1e-4) < 1e-6


def test_per_iteration_stats(tiny_problem):
    p = tiny_problem
    m = p.matrix
    x, s, z, sub, t = orc.mlem(m.rowidx, m.elements, m.cols, p.y, 4)
    assert s[-1] == np.cumsum(x, dtype=np.float32)[-1]     # sequential fp32 sum
    assert z[-1] == int(np.sum(x == 0))
    assert sub[-1] == 0
