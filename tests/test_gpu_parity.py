"""HIP path against the oracle on the GPU (through the C ABI).

EXACT mode must reproduce the fp32 restatement of seqcsr4mlem bit for bit;
FAST and ATOMIC modes are held to stated relative tolerances against the fp32
restatement and the FP64 oracle. Sizes are those the oracle finishes in
seconds; the 200M-element workload is covered by size-independent
properties (adjointness, count preservation) in test_gpu_scale.py.
"""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, rel_linf
import mlem_amd as M
from mlem_amd import synth
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

# Relative L-infinity tolerances (floor: 1e-6 of the image maximum) of the
# FAST/ATOMIC GPU path: single projections vs FP64, and the 60-iteration
# reconstruction vs the fp32 sequential restatement.
TOL_PROJ = 2e-6
TOL_RECON = 1e-4       # ATOMIC reconstructions (unordered fp32 sums) vs fp32 / FP64
TOL_FAST_SEQ = 1e-5    # FAST reconstructions vs the fp32 restatement of seqcsr4mlem


def _open(m, **kw):
    return M.Csr4Matrix.from_arrays(m.rows, m.cols, m.rowidx, m.elements, grid=m.grid, **kw)


# "layout": rows and voxels renumbered on the device (the default);
# "file": file order kept (MLEM_OPEN_KEEP_ORDER).
@pytest.fixture(scope="module", params=["layout", "file"])
def tiny(tiny_problem, request):
    A = _open(tiny_problem.matrix, keep_order=(request.param == "file"))
    yield tiny_problem, A
    A.close()


@pytest.fixture(scope="module", params=["layout", "file"])
def small(request):
    p = synth.make_problem("small")
    A = _open(p.matrix, keep_order=(request.param == "file"))
    yield p, A
    A.close()


def test_golden_fixture_exact(golden_tiny):
    ref, rowidx, el = golden_tiny
    A = M.Csr4Matrix(str(GOLDEN / "tiny3x4.csr4"))
    assert (A.rows(), A.columns(), A.elements()) == (3, 4, 6)
    assert A.geometry() == ((2, 2, 1), 0)
    y = np.array(ref["sino"], np.int32)
    x = M.mlem(A, y, None, 3, mode="exact")
    xo = orc.mlem(rowidx, el, 4, y, 3)[0]
    assert np.array_equal(x, xo)
    assert rel_linf(x, ref["iterates"][-1]["x"]) < 2e-6


def test_column_sums_exact(tiny):
    p, A = tiny
    m = p.matrix
    assert np.array_equal(M.calcColumnSums(A), orc.column_sums(m.rowidx, m.elements, m.cols))


def test_steps_exact(tiny):
    p, A = tiny
    m = p.matrix
    rng = np.random.default_rng(3)
    x = rng.uniform(0.1, 3, m.cols).astype(np.float32)
    f = M.calcFwProj(A, x, "exact")
    assert np.array_equal(f, orc.fwproj(m.rowidx, m.elements, x))
    r = M.calcCorrel(A, f, p.y)
    assert np.array_equal(r, orc.correl(f, p.y))
    u = M.calcBkProj(A, r, "exact")
    assert np.array_equal(u, orc.bkproj(m.rowidx, m.elements, r, m.cols))
    norm = M.calcColumnSums(A)
    xg = M.calcUpdate(A, u, norm, x.copy())
    assert np.array_equal(xg, orc.update(u, norm, x))
    x0, v = M.initImage(A, norm, p.y)
    xo, vo, _, _ = orc.init_image(norm, p.y)
    assert v == vo and np.array_equal(x0, xo)


def test_steps_fast_and_atomic(tiny):
    p, A = tiny
    m = p.matrix
    rng = np.random.default_rng(4)
    x = rng.uniform(0.1, 3, m.cols).astype(np.float32)
    r = rng.uniform(0.1, 3, m.rows).astype(np.float32)
    f64 = orc.f64_fwproj(m.rowidx, m.elements, x)
    u64 = orc.f64_bkproj(m.rowidx, m.elements, r, m.cols)
    assert rel_linf(M.calcFwProj(A, x, "fast"), f64) < TOL_PROJ
    assert rel_linf(M.calcBkProj(A, r, "fast"), u64) < TOL_PROJ
    assert rel_linf(M.calcBkProj(A, r, "atomic"), u64) < TOL_PROJ


def test_recon_exact_bitwise_small(small):
    # the 32^3, 262,144-line workload: 60 iterations, reference init
    p, A = small
    m = p.matrix
    x = M.mlem(A, p.y, None, 60, mode="exact")
    xo = orc.mlem(m.rowidx, m.elements, m.cols, p.y, 60)[0]
    assert np.array_equal(x, xo)


@pytest.mark.parametrize("mode", ["fast", "atomic"])
def test_recon_fast_modes_small(small, mode):
    p, A = small
    m = p.matrix
    x = M.mlem(A, p.y, None, 60, mode=mode)
    xo = orc.mlem(m.rowidx, m.elements, m.cols, p.y, 60)[0]
    x64, _, _ = orc.f64_mlem(m.rowidx, m.elements, m.cols, p.y.astype(np.float64), 60)
    e32 = rel_linf(x, xo)
    e64 = rel_linf(x, x64)
    e_ref64 = rel_linf(xo, x64)
    print(f"{mode}: rel Linf vs fp32 seq {e32:.3e}, vs fp64 {e64:.3e}; seq vs fp64 {e_ref64:.3e}")
    # FAST is held to the plain-fp32 bar (1e-5 relative per voxel vs the
    # sequential driver); ATOMIC's unordered sums to 1e-4
    assert e32 < (TOL_FAST_SEQ if mode == "fast" else TOL_RECON)
    assert e64 < TOL_RECON


AUTOTUNE_CANDIDATES = ({("dict512", e) for e in (16, 24, 32)} |
                       {("dict", e) for e in (8, 16, 24, 32, 48, 64)} |
                       {("tiles", e) for e in (4, 8, 16)} |
                       {("groups", g) for g in (8, 16, 32, 64)})


def test_autotune_in_core_matrix():
    # the launch autotuner on the 32^3 / 262,144-line matrix held in HBM: it
    # keeps one of its candidates per projection, EXACT stays bit-identical
    # to the fp32 restatement and FAST within 1e-5 of it after 60 iterations
    p = synth.make_problem("small")
    m = p.matrix
    A = _open(m)
    try:
        chosen = A.autotune()
        print(f"[autotune] {chosen}")
        assert tuple(chosen["fwd"]) in AUTOTUNE_CANDIDATES
        assert tuple(chosen["bwd"]) in AUTOTUNE_CANDIDATES
        assert A.launch() == chosen
        xo = orc.mlem(m.rowidx, m.elements, m.cols, p.y, 60)[0]
        assert np.array_equal(M.mlem(A, p.y, None, 60, mode="exact"), xo)
        assert rel_linf(M.mlem(A, p.y, None, 60, mode="fast"), xo) < TOL_FAST_SEQ
    finally:
        A.close()


def test_fast_is_deterministic_and_graph_matches(tiny):
    p, A = tiny
    a = M.mlem(A, p.y, None, 10, mode="fast", use_graph=True)
    b = M.mlem(A, p.y, None, 10, mode="fast", use_graph=False)
    c = M.mlem(A, p.y, None, 10, mode="fast", use_graph=True)
    assert np.array_equal(a, b) and np.array_equal(a, c)


def test_recon_state_matches_one_shot(tiny):
    p, A = tiny
    full = M.mlem(A, p.y, None, 7, mode="fast")
    R = M.Recon(A, p.y, mode="fast")
    R.run(3)
    R.run(4)
    assert R.iterations == 7
    assert np.array_equal(R.image(), full)
    R2 = M.Recon(A, p.y, mode="fast")
    kf, kb, span = R2.run_timed(7)
    assert np.all(kf > 0) and np.all(kb > 0) and span >= kf.sum()
    assert np.array_equal(R2.image(), full)


def test_launch_change_between_runs_recaptures(tiny):
    # a reconstruction whose iteration is captured in a graph keeps working
    # after the matrix's kernels (and their tile buffers) are replaced
    p, A = tiny
    m = p.matrix
    A.set_launch(("dict512", 32), ("dict512", 32))
    R = M.Recon(A, p.y, mode="exact")
    R.run(3)
    A.set_launch(("dict", 16), ("tiles", 8))
    R.run(2)
    A.set_launch(("dict512", 16), ("dict512", 16))
    R.run(2)
    assert np.array_equal(R.image(), orc.mlem(m.rowidx, m.elements, m.cols, p.y, 7)[0])
    R.close()
    A.set_launch(("dict512", 32), ("dict512", 32))


def test_recon_statistics_grid_grows_with_launch_change(tiny):
    # a reconstruction created with statistics under a launch of few
    # workgroups keeps running after a change to 64 lanes per row (a grid of
    # 1024 workgroups on this matrix needs more statistics partials)
    import ctypes as C
    from mlem_amd._native import lib
    p, A = tiny
    m = p.matrix
    A.set_launch(("dict512", 32), ("dict512", 32))
    ll = np.zeros(1, np.float64)
    o = M._options("fast", False, use_graph=True, loglik=ll)
    y = np.ascontiguousarray(p.y, np.int32)
    h = C.c_void_p()
    M.check(lib().mlem_recon_create(A.handle, y.ctypes.data_as(C.c_void_p), y.size, None,
                                    A.columns(), C.byref(o), C.byref(h)))
    try:
        M.check(lib().mlem_recon_run(h, 2))
        A.set_launch(("groups", 64), ("groups", 64))
        M.check(lib().mlem_recon_run(h, 2))
        A.set_launch(("dict512", 32), ("dict512", 32))
        M.check(lib().mlem_recon_run(h, 2))
        x = np.empty(A.columns(), np.float32)
        M.check(lib().mlem_recon_image(h, x.ctypes.data_as(C.c_void_p), x.size))
    finally:
        lib().mlem_recon_destroy(h)
        A.set_launch(("dict512", 32), ("dict512", 32))
    xo = orc.mlem(m.rowidx, m.elements, m.cols, p.y, 6)[0]
    assert rel_linf(x, xo) < 1e-5


def test_stats_loglik_and_counts(tiny):
    p, A = tiny
    m = p.matrix
    K = 12
    x, ll, ft = M.mlem(A, p.y, None, K, mode="fast", stats=True)
    x64, ll64, ft64 = orc.f64_mlem(m.rowidx, m.elements, m.cols, p.y.astype(np.float64), K)
    assert rel_linf(ll, ll64) < 1e-6
    assert rel_linf(ft, ft64) < 1e-6
    assert np.all(np.diff(ll) > -1e-6 * np.abs(ll[:-1]))
    # stats do not change the image
    assert np.array_equal(x, M.mlem(A, p.y, None, K, mode="fast"))


def test_given_x0_and_homogeneity(tiny):
    p, A = tiny
    x0 = np.full(A.columns(), 0.25, np.float32)
    xa = M.mlem(A, p.y, x0.copy(), 10, mode="exact", x0_given=True)
    xo = orc.mlem(p.matrix.rowidx, p.matrix.elements, A.columns(), p.y, 10, x0=x0)[0]
    assert np.array_equal(xa, xo)
    xb = M.mlem(A, p.y, (4 * x0).copy(), 10, mode="fast", x0_given=True)
    assert rel_linf(xb, xa) < TOL_RECON


def test_edge_cases_empty_rows_and_columns():
    # rows with no elements, an empty column, one long row, zero counts
    rows = [[], [(0, 1.0), (2, 0.5)], [], [(c, 0.25) for c in range(0, 300, 2) if c != 4],
            [(1, 2.0)], []]
    ends = np.cumsum([len(r) for r in rows]).astype(np.uint64)
    el = np.array([e for r in rows for e in r], dtype=M.ELEMENT_DTYPE)
    y = np.int32([0, 5, 3, 40, 0, 9])
    A = M.Csr4Matrix.from_arrays(6, 301, ends, el)
    for mode in ("exact", "fast", "atomic"):
        x = M.mlem(A, y, None, 5, mode=mode)
        xo = orc.mlem(ends, el, 301, y, 5)[0]
        if mode == "exact":
            assert np.array_equal(x, xo)
        else:
            assert rel_linf(x, xo) < 1e-5
        assert x[299] == 0 and x[300] == 0 and x[4] == 0   # empty columns
    A.close()


def _matrix(rows, cols):
    ends = np.cumsum([len(r) for r in rows]).astype(np.uint64)
    el = np.array([e for r in rows for e in r], dtype=M.ELEMENT_DTYPE)
    return ends, el, M.Csr4Matrix.from_arrays(len(rows), cols, ends, el)


def test_exact_keeps_subnormal_inputs_and_flushes_subnormal_results():
    # The reference sets FTZ without DAZ (mlem_seq/seqcsr4mlem.cpp:194): a
    # subnormal matrix value, image value or forward projection is used as it
    # is; a subnormal product, sum or quotient becomes zero.
    ends, el, A = _matrix([[(0, 1e10)], [(0, 1e-40)], [(2, 1e-20)], [(1, 1e-39), (0, 3e-39)]], 3)
    x = np.float32([1e-40, 1e10, 1e-20])
    f = M.calcFwProj(A, x, "exact")
    assert np.array_equal(f, orc.fwproj(ends, el, x))
    assert f[0] == np.float32(1e10) * np.float32(1e-40) > 0      # subnormal image value kept
    assert f[1] == 0 and f[2] == 0                                # subnormal products flushed
    assert f[3] == np.float32(1e-39) * np.float32(1e10) > 0      # subnormal matrix value kept
    fr = np.float32([1e-40, 3e38, 2.0, 0.0])                      # subnormal f: 1/f overflows
    yr = np.int32([1, 1, 3, 4])
    r = M.calcCorrel(A, fr, yr)
    assert np.array_equal(r, orc.correl(fr, yr))
    assert np.isinf(r[0]) and r[1] == 0                           # tiny quotient flushed
    rb = np.float32([1e-40, 1e10, 5e37, 1.0])
    u = M.calcBkProj(A, rb, "exact")
    assert np.array_equal(u, orc.bkproj(ends, el, rb, 3))
    assert u[0] > 0 and u[1] == 0                                 # 1e-39 * 1.0 flushed
    norm = np.float32([1.0, 4.0, 2.0])
    x = np.float32([1e-40, 1e-40, 1.0])                           # subnormal image, after u/n
    for uu in (np.float32([1e10, 3.0, 1.0]), np.float32([1e-30, 1e-20, 1e-37])):
        xg = M.calcUpdate(A, uu, norm, x.copy())
        assert np.array_equal(xg, orc.update(uu, norm, x))
    assert M.calcUpdate(A, np.float32([1e10, 3.0, 1.0]), norm, x.copy())[0] > 0
    A.close()
    # column sums that stay subnormal are flushed (norm 0, the update then
    # multiplies by u); rows made only of subnormal values still project
    rows = [[(0, 1.0), (1, 0.5)], [(0, 1e-39), (2, 2e-39)], [(1, 0.25), (2, 1e-40), (3, 1.0)],
            [(3, 2.0)]]
    ends, el, A = _matrix(rows, 4)
    norm = M.calcColumnSums(A)
    assert np.array_equal(norm, orc.column_sums(ends, el, 4)) and norm[2] == 0
    y = np.int32([1000000, 3, 500000, 7])
    x = M.mlem(A, y, None, 6, mode="exact")
    xo = orc.mlem(ends, el, 4, y, 6)[0]
    assert np.array_equal(x, xo)
    # without the subnormal inputs (denormals-are-zero) the result differs
    el_daz = el.copy()
    for k in range(len(el_daz)):
        if 0 < abs(float(el_daz[k]["value"])) < 1.17549435e-38:
            el_daz[k]["value"] = 0.0
    assert not np.array_equal(xo, orc.mlem(ends, el_daz, 4, y, 6)[0])
    A.close()


FAMILIES = [("tiles", 4), ("tiles", 8), ("tiles", 16), ("dict", 8), ("dict", 16),
            ("dict", 24), ("dict", 32), ("dict", 48), ("dict", 64), ("dict512", 16),
            ("dict512", 24), ("dict512", 32), ("groups", 4), ("groups", 16), ("groups", 64)]


@pytest.fixture(scope="module")
def ragged():
    # row lengths 0..299 plus rows longer than every tile (5000 > 4096), and
    # columns from 1 to ~300 elements: every lane-group tail, partial tile and
    # long-segment path
    rng = np.random.default_rng(8)
    cols = 6000
    rows = [[]]
    for L in list(range(1, 300)) + [511, 512, 513, 1000, 2000, 2049, 4097, 5000]:
        c = np.sort(rng.choice(cols, L, replace=False))
        rows.append([(int(j), float(np.float32(rng.uniform(0.1, 1.7)))) for j in c])
    ends = np.cumsum([len(r) for r in rows]).astype(np.uint64)
    el = np.array([e for r in rows for e in r], dtype=M.ELEMENT_DTYPE)
    A = M.Csr4Matrix.from_arrays(len(rows), cols, ends, el)
    yield A, ends, el, cols, rng
    A.close()


def test_ragged_exact(ragged):
    A, ends, el, cols, rng = ragged
    x = rng.uniform(0.1, 3, cols).astype(np.float32)
    r = rng.uniform(0.1, 3, A.rows()).astype(np.float32)
    assert np.array_equal(M.calcFwProj(A, x, "exact"), orc.fwproj(ends, el, x))
    assert np.array_equal(M.calcBkProj(A, r, "exact"), orc.bkproj(ends, el, r, cols))


@pytest.mark.parametrize("family", FAMILIES, ids=lambda f: f"{f[0]}{f[1]}")
def test_ragged_fast_families(ragged, family):
    A, ends, el, cols, rng = ragged
    A.set_launch(family, family)
    assert A.launch() == {"fwd": family, "bwd": family}
    x = rng.uniform(0.1, 3, cols).astype(np.float32)
    r = rng.uniform(0.1, 3, A.rows()).astype(np.float32)
    assert rel_linf(M.calcFwProj(A, x, "fast"), orc.f64_fwproj(ends, el, x)) < TOL_PROJ
    assert rel_linf(M.calcBkProj(A, r, "fast"), orc.f64_bkproj(ends, el, r, cols)) < TOL_PROJ
    y = rng.integers(0, 50, A.rows()).astype(np.int32)
    xg, ll, _ = M.mlem(A, y, None, 6, mode="fast", stats=True)
    xo = orc.mlem(ends, el, cols, y, 6)[0]
    assert rel_linf(xg, xo) < 1e-5
    _, ll64, _ = orc.f64_mlem(ends, el, cols, y.astype(np.float64), 6)
    assert rel_linf(ll, ll64) < 1e-6


@pytest.mark.parametrize("family", FAMILIES, ids=lambda f: f"{f[0]}{f[1]}")
def test_small_fast_families(small, family):
    # every kernel family on the 32^3 workload, 20 iterations
    p, A = small
    m = p.matrix
    A.set_launch(family, family)
    x = M.mlem(A, p.y, None, 20, mode="fast")
    xo = orc.mlem(m.rowidx, m.elements, m.cols, p.y, 20)[0]
    assert rel_linf(x, xo) < TOL_RECON
    A.set_launch(("tiles", 8), ("tiles", 8))


@pytest.mark.parametrize("family", [("dict", 8), ("dict", 64), ("dict512", 16), ("dict512", 32)],
                         ids=lambda f: f"{f[0]}{f[1]}")
def test_short_rows_fill_tiles_by_segment_count(family):
    # 5000 rows of 0-6 elements and 3000 columns: dictionary tiles end at their
    # segment limit (256 or 512 rows), not at their element limit
    rng = np.random.default_rng(21)
    cols = 3000
    lens = rng.integers(0, 7, 5000)
    rows = [np.sort(rng.choice(cols, L, replace=False)) for L in lens]
    ends = np.cumsum(lens).astype(np.uint64)
    el = np.zeros(int(ends[-1]), dtype=M.ELEMENT_DTYPE)
    el["column"] = np.concatenate(rows).astype(np.uint32)
    el["value"] = rng.uniform(0.1, 1.7, el.size).astype(np.float32)
    A = M.Csr4Matrix.from_arrays(len(rows), cols, ends, el)
    A.set_launch(family, family)
    x = rng.uniform(0.1, 3, cols).astype(np.float32)
    r = rng.uniform(0.1, 3, A.rows()).astype(np.float32)
    assert rel_linf(M.calcFwProj(A, x, "fast"), orc.f64_fwproj(ends, el, x)) < TOL_PROJ
    assert rel_linf(M.calcBkProj(A, r, "fast"), orc.f64_bkproj(ends, el, r, cols)) < TOL_PROJ
    y = rng.integers(0, 50, A.rows()).astype(np.int32)
    xg = M.mlem(A, y, None, 6, mode="fast")
    assert rel_linf(xg, orc.mlem(ends, el, cols, y, 6)[0]) < 1e-5
    A.close()


@pytest.mark.parametrize("family", [("dict512", 16), ("dict512", 32), ("dict", 64)],
                         ids=lambda f: f"{f[0]}{f[1]}")
def test_exact_segments_longer_than_a_tile(family):
    # a 20000-element row and a column shared by 9000 rows exceed every tile
    # (8192 / 16384 elements): EXACT walks such a segment with one lane, the
    # other segments with ordered per-lane sums over the tile's products
    rng = np.random.default_rng(23)
    cols = 30000
    rows = [np.sort(rng.choice(np.arange(1, cols), 20000, replace=False))]
    for _ in range(9000):
        rows.append(np.concatenate([[0], np.sort(rng.choice(np.arange(1, cols),
                                                             int(rng.integers(0, 12)),
                                                             replace=False))]))
    ends = np.cumsum([len(r) for r in rows]).astype(np.uint64)
    el = np.zeros(int(ends[-1]), dtype=M.ELEMENT_DTYPE)
    el["column"] = np.concatenate(rows).astype(np.uint32)
    el["value"] = rng.uniform(0.1, 1.7, el.size).astype(np.float32)
    A = M.Csr4Matrix.from_arrays(len(rows), cols, ends, el)
    A.set_launch(family, family)
    x = rng.uniform(0.1, 3, cols).astype(np.float32)
    r = rng.uniform(0.1, 3, len(rows)).astype(np.float32)
    assert np.array_equal(M.calcFwProj(A, x, "exact"), orc.fwproj(ends, el, x))
    assert np.array_equal(M.calcBkProj(A, r, "exact"), orc.bkproj(ends, el, r, cols))
    y = rng.integers(0, 30, len(rows)).astype(np.int32)
    assert np.array_equal(M.mlem(A, y, None, 3, mode="exact"), orc.mlem(ends, el, cols, y, 3)[0])
    A.close()


def test_layout_with_partial_bricks():
    # 7x5x3 grid: bricks of 4x4x2 are cut at every face
    m = synth.make_matrix((7, 5, 3), 3000, 4242)
    rng = np.random.default_rng(9)
    y = rng.integers(0, 30, m.rows).astype(np.int32)
    A = _open(m)
    x = rng.uniform(0.1, 3, m.cols).astype(np.float32)
    r = rng.uniform(0.1, 3, m.rows).astype(np.float32)
    assert np.array_equal(M.calcFwProj(A, x, "exact"), orc.fwproj(m.rowidx, m.elements, x))
    assert np.array_equal(M.calcBkProj(A, r, "exact"), orc.bkproj(m.rowidx, m.elements, r, m.cols))
    assert np.array_equal(M.calcColumnSums(A), orc.column_sums(m.rowidx, m.elements, m.cols))
    assert np.array_equal(M.mlem(A, y, None, 8, mode="exact"),
                          orc.mlem(m.rowidx, m.elements, m.cols, y, 8)[0])
    x0 = rng.uniform(0.5, 1.5, m.cols).astype(np.float32)
    assert np.array_equal(M.mlem(A, y, x0.copy(), 4, mode="exact", x0_given=True),
                          orc.mlem(m.rowidx, m.elements, m.cols, y, 4, x0=x0)[0])
    A.close()


def test_errors_on_device(tiny):
    p, A = tiny
    with pytest.raises(M.MlemError) as e:
        M.calcFwProj(A, np.zeros(A.columns() + 1, np.float32))
    assert e.value.code == -1
    bad = p.matrix.elements.copy()
    bad["column"][5] = p.matrix.cols
    with pytest.raises(M.MlemError) as e:
        M.Csr4Matrix.from_arrays(p.matrix.rows, p.matrix.cols, p.matrix.rowidx, bad)
    assert e.value.code == -3 and "column index" in str(e.value)


def test_no_csc_matrix_atomic_only(tiny):
    p, _ = tiny
    A = _open(p.matrix, no_csc=True)
    x = M.mlem(A, p.y, None, 5, mode="atomic")
    xo = orc.mlem(p.matrix.rowidx, p.matrix.elements, A.columns(), p.y, 5)[0]
    assert rel_linf(x, xo) < 1e-5
    with pytest.raises(M.MlemError):
        M.mlem(A, p.y, None, 5, mode="fast")
    A.close()


# ------------------------------------------------------------------- CLI ----
CLI = ROOT / "mlem_amd" / "bin" / "mlem_cli"


def _cli(args, cwd, mode="exact"):
    env = dict(os.environ)
    env.pop("MLEM_MODE", None)
    if mode is not None:
        env["MLEM_MODE"] = mode
    return subprocess.run([str(CLI), *map(str, args)], cwd=cwd, env=env, capture_output=True,
                          text=True, timeout=120)


def test_cli_drop_in(tmp_path, tiny_problem):
    p = tiny_problem
    mtx, sino, out = tmp_path / "m.csr4", tmp_path / "s.sino", tmp_path / "o.img"
    p.matrix.write_csr4(str(mtx))
    p.y.tofile(sino)
    r = _cli([mtx, sino, out, 6], tmp_path)
    assert r.returncode == 0, r.stderr
    x = np.fromfile(out, np.float32)
    xo, s, z, sub, _ = orc.mlem(p.matrix.rowidx, p.matrix.elements, p.matrix.cols, p.y, 6)
    assert np.array_equal(x, xo)
    lines = r.stdout.splitlines()
    assert lines[0] == f"Matrix file: {mtx}" and lines[3] == "Iterations: 6"
    assert any(l.startswith("INIT: sumnorm=") for l in lines)
    fin = [l for l in lines if l.startswith("Finished iteration ")]
    assert len(fin) == 6 and fin[0].startswith("Finished iteration 1,  image sum: ")
    # the printed image sums are the sequential float sums of the reference
    printed = [float(l.split("image sum: ")[1].split(",")[0]) for l in fin]
    assert np.allclose(printed, s, rtol=1e-5)
    assert any(l.startswith("Zeros: ") for l in lines)


def test_cli_default_fast_mode(tmp_path, tiny_problem):
    # no MLEM_MODE in the environment: the drop-in runs FAST, within 1e-5 of
    # the fp32 restatement of seqcsr4mlem, and prints the same lines
    p = tiny_problem
    mtx, sino, out = tmp_path / "m.csr4", tmp_path / "s.sino", tmp_path / "o.img"
    p.matrix.write_csr4(str(mtx))
    p.y.tofile(sino)
    r = _cli([mtx, sino, out, 6], tmp_path, mode=None)
    assert r.returncode == 0, r.stderr
    x = np.fromfile(out, np.float32)
    xo = orc.mlem(p.matrix.rowidx, p.matrix.elements, p.matrix.cols, p.y, 6)[0]
    assert rel_linf(x, xo) < TOL_FAST_SEQ
    fin = [l for l in r.stdout.splitlines() if l.startswith("Finished iteration ")]
    assert len(fin) == 6
    # the same through the library in FAST mode: identical bits
    A = _open(p.matrix)
    assert np.array_equal(x, M.mlem(A, p.y, None, 6, mode="fast"))
    A.close()


def test_cli_size_mismatch(tmp_path, tiny_problem):
    p = tiny_problem
    mtx, sino = tmp_path / "m.csr4", tmp_path / "s.sino"
    p.matrix.write_csr4(str(mtx))
    p.y[:-1].tofile(sino)
    r = _cli([mtx, sino, tmp_path / "o.img", 2], tmp_path)
    assert r.returncode == 1 and "rows" in r.stderr


def test_cli_checkpoint_restart(tmp_path, tiny_problem):
    p = tiny_problem
    mtx, sino = tmp_path / "m.csr4", tmp_path / "s.sino"
    p.matrix.write_csr4(str(mtx))
    p.y.tofile(sino)
    full = tmp_path / "full.img"
    assert _cli([mtx, sino, full, 7], tmp_path).returncode == 0
    # a complete checkpoint after 3 iterations, then restart to 7
    xo3 = orc.mlem(p.matrix.rowidx, p.matrix.elements, p.matrix.cols, p.y, 3)[0]
    xo3.tofile(tmp_path / "img.chkpt")
    np.int32([3]).tofile(tmp_path / "iter.chkpt")
    out = tmp_path / "restart.img"
    r = _cli([mtx, sino, out, 7, 2], tmp_path)
    assert r.returncode == 0, r.stderr
    assert "I got Checkpoint at iteration 3" in r.stdout
    assert np.array_equal(np.fromfile(out, np.float32), np.fromfile(full, np.float32))
    assert not (tmp_path / "img.chkpt").exists()      # removed on clean exit
    # incomplete checkpoint is ignored
    xo3.tofile(tmp_path / "img.chkpt")
    r = _cli([mtx, sino, out, 7, 2], tmp_path)
    assert r.returncode == 0 and "omitted incomplete checkpoint" in r.stdout
    assert np.array_equal(np.fromfile(out, np.float32), np.fromfile(full, np.float32))
    # wrong-size checkpoint is an error
    np.zeros(5, np.float32).tofile(tmp_path / "img.chkpt")
    np.int32([3]).tofile(tmp_path / "iter.chkpt")
    r = _cli([mtx, sino, out, 7, 2], tmp_path)
    assert r.returncode == 1 and "Mismatch element number" in r.stderr
    # <checkpointing> = 0 still restarts from a complete checkpoint
    # (mpicsr4mlem.cpp:536-549) and neither writes nor removes the files
    xo3.tofile(tmp_path / "img.chkpt")
    np.int32([3]).tofile(tmp_path / "iter.chkpt")
    r = _cli([mtx, sino, out, 7, 0], tmp_path)
    assert r.returncode == 0, r.stderr
    assert "I got Checkpoint at iteration 3" in r.stdout
    assert np.array_equal(np.fromfile(out, np.float32), np.fromfile(full, np.float32))
    assert (tmp_path / "img.chkpt").exists() and (tmp_path / "iter.chkpt").exists()
    assert np.array_equal(np.fromfile(tmp_path / "img.chkpt", np.float32), xo3)
    # the 4-argument form is seqcsr4mlem: no checkpoints, the files are ignored
    r = _cli([mtx, sino, out, 7], tmp_path)
    assert r.returncode == 0 and "Checkpoint" not in r.stdout
    assert np.array_equal(np.fromfile(out, np.float32), np.fromfile(full, np.float32))


def test_cli_writes_checkpoints(tmp_path, tiny_problem):
    p = tiny_problem
    mtx, sino = tmp_path / "m.csr4", tmp_path / "s.sino"
    p.matrix.write_csr4(str(mtx))
    p.y.tofile(sino)
    # the checkpoint files exist while running and are deleted at the end;
    # check their content through the library, which writes them the same way
    chk = tmp_path / "chk"
    chk.mkdir()
    A = _open(p.matrix)
    x = M.mlem(A, p.y, None, 4, mode="exact", checkpoint_interval=2, checkpoint_dir=str(chk))
    assert not (chk / "img.chkpt").exists() and not (chk / "iter.chkpt").exists()
    assert np.array_equal(x, orc.mlem(p.matrix.rowidx, p.matrix.elements, p.matrix.cols, p.y, 4)[0])
    A.close()
