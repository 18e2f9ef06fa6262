"""Pins the oracle's FP64 reverse mode through K unrolled MLEM iterations
(loss 0.5 ||x_K - x_true||^2) by central finite differences along seeded
random directions and by the two homogeneity identities of MLEM:
x_K(x0) is degree-0 homogeneous, so <x0, dL/dx0> = 0; x_K(y) is degree-1
homogeneous, so <y, dL/dy> = <x_K - x_true, x_K>.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as orc

EL = orc.ELEMENT_DTYPE


def _problem(seed, rows=120, cols=40, empty_col=True):
    rng = np.random.default_rng(seed)
    lists = []
    for i in range(rows):
        L = int(rng.integers(1, 12))
        c = np.sort(rng.choice(cols - (1 if empty_col else 0), L, replace=False))
        v = rng.uniform(0.05, 1.5, L).astype(np.float32)
        lists.append(list(zip(c.tolist(), v.tolist())))
    lists[5] = []                                  # empty row: f = 0
    ends = np.cumsum([len(r) for r in lists]).astype(np.uint64)
    el = np.array([e for r in lists for e in r], dtype=EL)
    xt = rng.uniform(0.5, 3.0, cols)
    ybar = orc.f64_fwproj(ends, el, xt)
    y = rng.poisson(ybar).astype(np.float64) + rng.uniform(0, 0.5, rows)   # relaxed counts
    x0 = rng.uniform(0.5, 1.5, cols)
    return ends, el, cols, y, x0, xt, rng


def _loss(ends, el, cols, y, x0, K, xt):
    xK = orc.f64_unroll(ends, el, cols, y, x0, K)
    return 0.5 * float(np.sum((xK - xt) ** 2)), xK


@pytest.mark.parametrize("K", [1, 3, 8])
def test_adjoint_finite_differences(K):
    ends, el, cols, y, x0, xt, rng = _problem(K)
    L0, xK = _loss(ends, el, cols, y, x0, K, xt)
    _, gy, gx0 = orc.f64_unroll(ends, el, cols, y, x0, K, gK=xK - xt)
    for _ in range(16):
        dy = rng.standard_normal(y.size) * (y > 0)
        dx = rng.standard_normal(cols)
        for v, g, d in ((y, gy, dy), (x0, gx0, dx)):
            h = 1e-5 * np.linalg.norm(v) / max(np.linalg.norm(d), 1e-300)
            args_p = (y + h * d, x0) if v is y else (y, x0 + h * d)
            args_m = (y - h * d, x0) if v is y else (y, x0 - h * d)
            Lp = _loss(ends, el, cols, *args_p, K, xt)[0]
            Lm = _loss(ends, el, cols, *args_m, K, xt)[0]
            fd = (Lp - Lm) / (2 * h)
            an = float(np.dot(g, d))
            assert abs(fd - an) <= 1e-6 * max(abs(an), 1e-3 * np.linalg.norm(g) * np.linalg.norm(d))


@pytest.mark.parametrize("K", [2, 6])
def test_adjoint_homogeneity_identities(K):
    ends, el, cols, y, x0, xt, _ = _problem(10 + K)
    xK, gy, gx0 = orc.f64_unroll(ends, el, cols, y, x0, K, gK=orc.f64_unroll(ends, el, cols, y,
                                                                             x0, K) - xt)
    assert abs(np.dot(x0, gx0)) <= 1e-10 * np.linalg.norm(x0) * np.linalg.norm(gx0)
    want = float(np.dot(xK - xt, xK))
    assert abs(np.dot(y, gy) - want) <= 1e-10 * abs(want)
    assert gx0[cols - 1] == 0.0            # empty column: image pinned at 0 after one step


def test_unroll_matches_f64_mlem():
    ends, el, cols, y, x0, _, _ = _problem(3)
    xa = orc.f64_unroll(ends, el, cols, y, x0, 7)
    xb = orc.f64_mlem(ends, el, cols, y, 7, x0=x0)[0]
    assert np.array_equal(xa, xb)


def test_scaled_unroll_equals_dequantized():
    # power-of-two row scales: s * v exact, so scaled and plain runs agree
    ends, el, cols, y, x0, xt, rng = _problem(4)
    s = (2.0 ** rng.integers(-2, 3, ends.size)).astype(np.float32)
    row = np.repeat(np.arange(ends.size), np.diff(np.concatenate([[0], ends.astype(np.int64)])))
    el2 = el.copy()
    el2["value"] = el["value"] * s[row]
    g = rng.standard_normal(cols)
    a = orc.f64_unroll(ends, el, cols, y, x0, 4, gK=g, rscale=s)
    b = orc.f64_unroll(ends, el2, cols, y, x0, 4, gK=g)
    for u, v in zip(a, b):
        assert np.array_equal(u, v)
