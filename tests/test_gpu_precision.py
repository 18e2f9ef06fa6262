"""Reduced-precision matrix storage (fp16 / bf16 quotients with an fp32 row
scale) on the GPU against the oracle's restatement of the quantizer.

The device quantizes at open time; EXACT-mode projections must then equal,
bit for bit, the fp32 sequential projections of the oracle's quantized matrix
times the row scale, and reconstructions must stay within 1e-4 relative of
the FP64 oracle run on the dequantized matrix a_ij = s_i q_ij.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rel_linf
import mlem_amd as M
from mlem_amd import synth
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

KINDS = ["fp16", "bf16"]
# Relative L-infinity tolerance (floor 1e-6 of the maximum) of a GPU
# reconstruction with 16-bit values against FP64 MLEM on the dequantized matrix.
TOL_RECON_Q = 1e-4
TOL_PROJ = 2e-6


def _open(m, **kw):
    return M.Csr4Matrix.from_arrays(m.rows, m.cols, m.rowidx, m.elements, grid=m.grid, **kw)


def _row_index(rowidx):
    lens = np.diff(np.concatenate([[0], rowidx.astype(np.int64)]))
    return np.repeat(np.arange(rowidx.size), lens)


@pytest.fixture(scope="module", params=[(k, o) for k in KINDS for o in ("layout", "file")],
                ids=lambda p: f"{p[0]}-{p[1]}")
def qtiny(tiny_problem, request):
    kind, order = request.param
    m = tiny_problem.matrix
    A = _open(m, precision=kind, keep_order=(order == "file"))
    q, s = orc.quantize(m.rowidx, m.elements, kind)
    yield tiny_problem, A, q, s, kind
    A.close()


def test_column_sums_fp64_of_scaled_values(qtiny):
    # n_j = (float) sum_i s_i q_ij in FP64: at most one fp32 rounding away
    # from the FP64 sum taken in another order
    p, A, q, s, _ = qtiny
    m = p.matrix
    row = _row_index(m.rowidx)
    want = np.zeros(m.cols, np.float64)
    np.add.at(want, q["column"], q["value"].astype(np.float64) * s[row].astype(np.float64))
    got = M.calcColumnSums(A)
    assert rel_linf(got, want) <= 1.2e-7


def test_exact_projections_bitwise(qtiny):
    p, A, q, s, _ = qtiny
    m = p.matrix
    rng = np.random.default_rng(11)
    x = rng.uniform(0.1, 3, m.cols).astype(np.float32)
    f = M.calcFwProj(A, x, "exact")
    assert np.array_equal(f, orc.fwproj(m.rowidx, q, x) * s)
    ones = np.ones(m.cols, np.float32)
    assert np.array_equal(M.calcFwProj(A, ones, "exact"), orc.fwproj(m.rowidx, q, ones) * s)
    r = rng.uniform(0.1, 3, m.rows).astype(np.float32)
    u = M.calcBkProj(A, r, "exact")
    assert np.array_equal(u, orc.bkproj(m.rowidx, q, r * s, m.cols))


def test_fast_projections_vs_fp64(qtiny):
    p, A, q, s, _ = qtiny
    m = p.matrix
    rng = np.random.default_rng(12)
    x = rng.uniform(0.1, 3, m.cols).astype(np.float32)
    r = rng.uniform(0.1, 3, m.rows).astype(np.float32)
    s64 = s.astype(np.float64)
    f64 = orc.f64_fwproj(m.rowidx, q, x) * s64
    u64 = orc.f64_bkproj(m.rowidx, q, r.astype(np.float64) * s64, m.cols)
    assert rel_linf(M.calcFwProj(A, x, "fast"), f64) < TOL_PROJ
    assert rel_linf(M.calcBkProj(A, r, "fast"), u64) < TOL_PROJ
    assert rel_linf(M.calcBkProj(A, r, "atomic"), u64) < TOL_PROJ


@pytest.mark.parametrize("mode", ["fast", "exact", "atomic"])
def test_recon_vs_fp64_dequantized(qtiny, mode):
    p, A, q, s, _ = qtiny
    m = p.matrix
    K = 15
    x, ll, ft = M.mlem(A, p.y, None, K, mode=mode, stats=True)
    x64, ll64, ft64 = orc.f64_mlem_scaled(m.rowidx, q, s, m.cols, p.y.astype(np.float64), K)
    assert rel_linf(x, x64) < TOL_RECON_Q
    assert rel_linf(ll, ll64) < 1e-6
    assert rel_linf(ft, ft64) < 1e-6


# every dictionary tile size (16-bit tiles load 8 elements per lane and load,
# so each size has its own number of packs per lane) and the other families
FAMILIES_Q = [("dict", 8), ("dict", 16), ("dict", 24), ("dict", 32), ("dict", 48), ("dict", 64),
              ("dict512", 16), ("dict512", 24), ("dict512", 32), ("tiles", 8), ("groups", 16)]


@pytest.mark.parametrize("family", FAMILIES_Q, ids=lambda f: f"{f[0]}{f[1]}")
@pytest.mark.parametrize("kind", KINDS)
def test_small_families(kind, family):
    # 32^3 workload, 20 iterations, every kernel family on 16-bit values
    pr = synth.make_problem("small")
    m = pr.matrix
    A = _open(m, precision=kind)
    A.set_launch(family, family)
    q, s = orc.quantize(m.rowidx, m.elements, kind)
    x = M.mlem(A, pr.y, None, 20, mode="fast")
    x64 = orc.f64_mlem_scaled(m.rowidx, q, s, m.cols, pr.y.astype(np.float64), 20)[0]
    assert rel_linf(x, x64) < TOL_RECON_Q
    # reported, not a parity bar: distance to FP64 MLEM on the original values
    xo = orc.f64_mlem(m.rowidx, m.elements, m.cols, pr.y.astype(np.float64), 20)[0]
    err = rel_linf(x, xo)
    print(f"[precision] {kind} {family}: rel Linf vs fp64(original) {err:.3e}")
    assert err < 5e-2
    A.close()


@pytest.mark.parametrize("family", [f for f in FAMILIES_Q if f[0].startswith("dict")],
                         ids=lambda f: f"{f[0]}{f[1]}")
@pytest.mark.parametrize("kind", KINDS)
def test_exact_projections_every_tile_size(kind, family):
    # EXACT keeps its bitwise bar on every 16-bit dictionary tile size
    pr = synth.make_problem("small")
    m = pr.matrix
    A = _open(m, precision=kind)
    A.set_launch(family, family)
    q, s = orc.quantize(m.rowidx, m.elements, kind)
    rng = np.random.default_rng(5)
    x = rng.uniform(0.1, 3, m.cols).astype(np.float32)
    assert np.array_equal(M.calcFwProj(A, x, "exact"), orc.fwproj(m.rowidx, q, x) * s)
    r = rng.uniform(0.1, 3, m.rows).astype(np.float32)
    assert np.array_equal(M.calcBkProj(A, r, "exact"), orc.bkproj(m.rowidx, q, r * s, m.cols))
    A.close()


def test_open_flag_errors(tiny_problem):
    m = tiny_problem.matrix
    with pytest.raises(M.MlemError):
        M.Csr4Matrix.from_arrays(m.rows, m.cols, m.rowidx, m.elements,
                                 precision="fp16", no_csc=True)
    with pytest.raises(ValueError):
        M.Csr4Matrix.from_arrays(m.rows, m.cols, m.rowidx, m.elements, precision="fp8")
    A = _open(m, precision="bf16")
    with pytest.raises(M.MlemError):
        M.Batch(A, tiny_problem.y[None, :])
    A.close()


def test_fp16_subnormal_quotients_and_zero_rows():
    # rows whose small elements fall in the fp16 subnormal range (or below
    # half its quantum), an all-zero row and an empty row
    rows = [[(0, 1.5), (1, 1.5e-6), (2, 4e-9), (3, 0.75)],
            [(1, 0.0), (2, 0.0)],
            [],
            [(0, 2.0 ** -20), (3, 2.0 ** -6)],
            [(2, 0.3), (3, 0.30001)]]
    ends = np.cumsum([len(r) for r in rows]).astype(np.uint64)
    el = np.array([e for r in rows for e in r], dtype=M.ELEMENT_DTYPE)
    for kind in KINDS:
        A = M.Csr4Matrix.from_arrays(len(rows), 4, ends, el, precision=kind)
        q, s = orc.quantize(ends, el, kind)
        for c in range(4):
            x = np.zeros(4, np.float32)
            x[c] = 1.0
            assert np.array_equal(M.calcFwProj(A, x, "exact"), orc.fwproj(ends, q, x) * s), (kind, c)
        A.close()
