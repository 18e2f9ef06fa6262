"""Pins the oracle's reduced-precision restatement (row scale + fp16/bf16
quotients, FP64 MLEM on a_ij = s_i q_ij) without the GPU.

fp16 rounding is checked against numpy's float16 conversion, bf16 rounding
against torch's bfloat16 conversion and a brute-force nearest search; the
scaled FP64 MLEM against the unscaled one through exact power-of-two scales.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import oracle as orc

EL = orc.ELEMENT_DTYPE


def _matrix(rng, rows=60, cols=90, with_special=True):
    lists = []
    for i in range(rows):
        L = int(rng.integers(0, 40))
        c = np.sort(rng.choice(cols, L, replace=False))
        v = rng.uniform(0.0, 1.8, L).astype(np.float32)
        if with_special and L > 3 and i % 7 == 0:
            v[1] = np.float32(1e-6) * v[0]       # fp16 subnormal quotient
            v[2] = np.float32(3e-9) * v[0]       # below half the fp16 quantum
        lists.append(list(zip(c.tolist(), v.tolist())))
    lists.append([(3, 0.0), (5, 0.0)])           # all-zero row: scale 1
    lists.append([])                              # empty row
    ends = np.cumsum([len(r) for r in lists]).astype(np.uint64)
    el = np.array([e for r in lists for e in r], dtype=EL)
    return ends, el, cols


def _row_of(ends, nnz):
    starts = np.concatenate([[0], ends[:-1]]).astype(np.int64)
    return np.repeat(np.arange(ends.size), np.diff(np.concatenate([[0], ends])).astype(np.int64)), starts


def test_row_scale_is_row_max():
    rng = np.random.default_rng(1)
    ends, el, _ = _matrix(rng)
    q, s = orc.quantize(ends, el, "fp16")
    b = 0
    for i, e in enumerate(ends):
        v = el["value"][b:int(e)]
        want = np.float32(np.max(np.abs(v))) if v.size and np.max(np.abs(v)) > 0 else np.float32(1)
        assert s[i] == want, i
        b = int(e)
    assert np.array_equal(q["column"], el["column"])


def test_fp16_matches_numpy_half():
    rng = np.random.default_rng(2)
    ends, el, _ = _matrix(rng)
    q, s = orc.quantize(ends, el, "fp16")
    row, _ = _row_of(ends, el.size)
    quot = el["value"] / s[row]                      # fp32 division, as the oracle
    want = quot.astype(np.float16).astype(np.float32)
    assert np.array_equal(q["value"], want)
    assert np.any((want > 0) & (want < 2.0 ** -14))  # subnormal fp16 quotients present
    assert np.all(np.abs(q["value"]) <= 1.0)


def test_fp16_ties_to_even():
    # quotients exactly halfway between fp16 neighbours round to even
    ulp = 2.0 ** -11                                  # fp16 spacing in [0.5, 1)
    vals = np.float32([1.0, 0.5 + 0.5 * ulp, 0.5 + 1.5 * ulp, 0.75 + 2.5 * ulp])
    ends = np.uint64([vals.size])
    el = np.zeros(vals.size, EL)
    el["column"] = np.arange(vals.size)
    el["value"] = vals
    q, s = orc.quantize(ends, el, "fp16")
    assert s[0] == 1.0
    assert list(q["value"]) == [1.0, 0.5, 0.5 + 2 * ulp, 0.75 + 2 * ulp]


def _nearest_bf16(x: np.ndarray) -> np.ndarray:
    """Brute force: the two bf16 neighbours of every value, nearest wins,
    ties to the even significand."""
    out = np.empty_like(x)
    for k, v in enumerate(x.astype(np.float64)):
        u = np.float32(v).view(np.uint32)
        lo = np.uint32(u & 0xFFFF0000)
        hi = np.uint32(lo + 0x10000)
        a = float(lo.view(np.float32))
        b = float(hi.view(np.float32))
        da, db = abs(v - a), abs(b - v)
        if da < db or (da == db and ((int(lo) >> 16) & 1) == 0):
            out[k] = np.float32(a)
        else:
            out[k] = np.float32(b)
    return out


def test_bf16_matches_torch_and_brute_force():
    rng = np.random.default_rng(3)
    ends, el, _ = _matrix(rng)
    q, s = orc.quantize(ends, el, "bf16")
    row, _ = _row_of(ends, el.size)
    quot = (el["value"] / s[row]).astype(np.float32)
    via_torch = torch.from_numpy(quot).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(q["value"], via_torch)
    assert np.array_equal(q["value"], _nearest_bf16(quot))


def test_scaled_f64_mlem_identity_and_pow2_scales():
    rng = np.random.default_rng(4)
    ends, el, cols = _matrix(rng, with_special=False)
    y = rng.integers(0, 30, ends.size).astype(np.float64)
    x1, ll1, ft1 = orc.f64_mlem(ends, el, cols, y, 7)
    xs, lls, fts = orc.f64_mlem_scaled(ends, el, np.ones(ends.size, np.float32), cols, y, 7)
    assert np.array_equal(x1, xs) and np.array_equal(ll1, lls) and np.array_equal(ft1, fts)
    # per-row powers of two: s_i * v_ij is exact in fp32, so the scaled run on
    # (v, s) and the plain run on s*v are the same arithmetic
    s = (2.0 ** rng.integers(-3, 4, ends.size)).astype(np.float32)
    row, _ = _row_of(ends, el.size)
    el2 = el.copy()
    el2["value"] = el["value"] * s[row]
    xa = orc.f64_mlem_scaled(ends, el, s, cols, y, 7)[0]
    xb = orc.f64_mlem(ends, el2, cols, y, 7)[0]
    assert np.array_equal(xa, xb)


def test_scaled_f64_mlem_uniform_scale_divides_image():
    # A -> cA maps x_k -> x_k / c (x0 = sum y / sum n scales the same way)
    rng = np.random.default_rng(5)
    ends, el, cols = _matrix(rng, with_special=False)
    y = rng.integers(0, 30, ends.size).astype(np.float64)
    x1 = orc.f64_mlem(ends, el, cols, y, 5)[0]
    x2 = orc.f64_mlem_scaled(ends, el, np.full(ends.size, 4.0, np.float32), cols, y, 5)[0]
    assert np.array_equal(x2, x1 / 4.0)


@pytest.mark.parametrize("kind", ["fp16", "bf16"])
def test_quantization_error_bound(kind):
    # half an ulp of a p-bit significand is at most 2^-p relative; fp16 adds
    # half its subnormal quantum (2^-25), the fp32 quotient one fp32 rounding
    rng = np.random.default_rng(6)
    ends, el, _ = _matrix(rng)
    q, s = orc.quantize(ends, el, kind)
    row, _ = _row_of(ends, el.size)
    sr = s[row].astype(np.float64)
    deq = q["value"].astype(np.float64) * sr
    a = el["value"].astype(np.float64)
    p = 11 if kind == "fp16" else 8
    sub = 2.0 ** -25 if kind == "fp16" else 0.0
    tol = sr * (2.0 ** -p * np.abs(a) / sr + sub) + np.abs(a) * 2.0 ** -23
    assert np.all(np.abs(deq - a) <= tol)
    # and the bound is not vacuous: some element uses more than half of it
    assert np.max(np.abs(deq - a) / np.maximum(tol, 1e-30)) > 0.5


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_parallel_f64_matches_sequential(threads):
    rng = np.random.default_rng(7)
    ends, el, cols = _matrix(rng, rows=300, cols=200, with_special=False)
    y = rng.integers(0, 30, ends.size).astype(np.float64)
    s = rng.uniform(0.5, 2, ends.size).astype(np.float32)
    for rs in (None, s):
        if rs is None:
            xa, la, fa = orc.f64_mlem(ends, el, cols, y, 9)
        else:
            xa, la, fa = orc.f64_mlem_scaled(ends, el, rs, cols, y, 9)
        xb, lb, fb = orc.f64_mlem_par(ends, el, cols, y, 9, rscale=rs, threads=threads)
        assert np.max(np.abs(xa - xb) / np.maximum(np.abs(xa), 1e-300)) < 1e-12
        assert np.max(np.abs(la - lb) / np.abs(la)) < 1e-13
        assert np.max(np.abs(fa - fb) / np.abs(fa)) < 1e-13
