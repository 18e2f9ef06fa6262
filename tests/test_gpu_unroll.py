"""Differentiable unrolled MLEM on the GPU against the FP64 oracle.

Forward: with integer-valued counts the unrolled iterate is bit-identical to
the fp32 sequential restatement of mlem() from the same x0. Backward: the
hand-written adjoint (fp32 on the GPU) against the oracle's FP64 reverse
mode, central finite differences of the FP64 forward, the homogeneity
identities of MLEM, and checkpoint-interval independence (bitwise).
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rel_linf
import mlem_amd as M
from mlem_amd import synth
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

# Relative L2 error of an fp32 GPU gradient against the FP64 oracle's.
TOL_GRAD = 1e-4


def _open(m):
    return M.Csr4Matrix.from_arrays(m.rows, m.cols, m.rowidx, m.elements, grid=m.grid)


def _frames(p, n, seed):
    rng = np.random.default_rng(seed)
    ys = [p.y.astype(np.float64)]
    for f in range(1, n):
        ys.append(rng.poisson(p.ybar * rng.uniform(0.3, 2.0)).astype(np.float64)
                  + rng.uniform(0, 0.5, p.ybar.size) * (f % 2))
    return np.stack(ys)


@pytest.fixture(scope="module")
def tiny_u(tiny_problem):
    A = _open(tiny_problem.matrix)
    yield tiny_problem, A
    A.close()


def _rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def test_forward_bitwise_integer_counts(tiny_u):
    p, A = tiny_u
    m = p.matrix
    rng = np.random.default_rng(1)
    Y = np.stack([p.y, rng.poisson(p.ybar * 0.5)]).astype(np.int32)
    X0 = rng.uniform(0.5, 1.5, (2, m.cols)).astype(np.float32)
    u = M.Unrolled(A, Y.astype(np.float32), X0, K=7, keep_tape=False)
    for f in range(2):
        xo = orc.mlem(m.rowidx, m.elements, m.cols, Y[f], 7, x0=X0[f])[0]
        assert np.array_equal(u.xK[f], xo), f


def test_forward_default_x0_matches_mlem(tiny_u):
    p, A = tiny_u
    m = p.matrix
    u = M.Unrolled(A, p.y.astype(np.float32)[None, :], None, K=5, keep_tape=False)
    assert np.array_equal(u.xK[0], orc.mlem(m.rowidx, m.elements, m.cols, p.y, 5)[0])


@pytest.mark.parametrize("mode", ["exact", "fast"])
@pytest.mark.parametrize("K,C", [(6, 2), (6, 0), (5, 1), (5, 5), (7, 3)])
def test_gradient_vs_fp64_oracle(tiny_u, K, C, mode):
    p, A = tiny_u
    m = p.matrix
    Y = _frames(p, 5, 2)
    rng = np.random.default_rng(3)
    X0 = rng.uniform(0.5, 1.5, (5, m.cols))
    u = M.Unrolled(A, Y, X0, K=K, checkpoint_interval=C, mode=mode)
    info = u.info()
    Ce = C if C else int(round(np.sqrt(K)))
    assert info["checkpoint_interval"] == Ce and info["checkpoints"] == -(-K // Ce)
    xt = p.x_true
    g = u.xK - xt
    gy, gx0 = u.backward(g)
    assert u.info()["recomputed_iterations"] == K
    assert u.info()["spmm_launches"] == 6 * K
    for f in range(5):
        xK64, gy64, gx064 = orc.f64_unroll(m.rowidx, m.elements, m.cols, Y[f].astype(np.float32),
                                           X0[f].astype(np.float32), K,
                                           gK=g[f].astype(np.float64))
        assert rel_linf(u.xK[f], xK64) < 1e-5
        assert _rel_l2(gy[f], gy64) < TOL_GRAD, f
        assert _rel_l2(gx0[f], gx064) < TOL_GRAD, f
    u.close()


def test_gradient_independent_of_checkpointing_and_batch(tiny_u):
    p, A = tiny_u
    m = p.matrix
    Y = _frames(p, 3, 4)
    rng = np.random.default_rng(5)
    X0 = rng.uniform(0.5, 1.5, (3, m.cols)).astype(np.float32)
    G = rng.standard_normal((3, m.cols)).astype(np.float32)
    ref = None
    for C in (1, 2, 3, 8):
        u = M.Unrolled(A, Y, X0, K=8, checkpoint_interval=C)
        gy, gx0 = u.backward(G)
        if ref is None:
            ref = (u.xK.copy(), gy, gx0)
        else:
            assert np.array_equal(u.xK, ref[0])
            assert np.array_equal(gy, ref[1]) and np.array_equal(gx0, ref[2]), C
        u.close()
    # one frame alone (block of 4 lanes) gives the same bits as in the batch
    u = M.Unrolled(A, Y[1:2], X0[1:2], K=8, checkpoint_interval=2)
    gy1, gx01 = u.backward(G[1:2])
    assert np.array_equal(u.xK[0], ref[0][1])
    assert np.array_equal(gy1[0], ref[1][1]) and np.array_equal(gx01[0], ref[2][1])
    # a second backward on the same tape repeats the result
    gy2, gx02 = u.backward(G[1:2])
    assert np.array_equal(gy1, gy2) and np.array_equal(gx01, gx02)
    u.close()


def test_fast_gradient_deterministic_and_independent_of_checkpointing(tiny_u):
    # the tiled products sum in a fixed order: for one block width, x_K and
    # the gradients are the same bits whatever the checkpoint interval, and
    # a repeated backward pass gives them again
    p, A = tiny_u
    m = p.matrix
    Y = _frames(p, 3, 4)
    rng = np.random.default_rng(5)
    X0 = rng.uniform(0.5, 1.5, (3, m.cols)).astype(np.float32)
    G = rng.standard_normal((3, m.cols)).astype(np.float32)
    ref = None
    for C in (1, 2, 3, 8):
        u = M.Unrolled(A, Y, X0, K=8, checkpoint_interval=C, mode="fast")
        gy, gx0 = u.backward(G)
        if ref is None:
            ref = (u.xK.copy(), gy, gx0)
            gy2, gx02 = u.backward(G)
            assert np.array_equal(gy, gy2) and np.array_equal(gx0, gx02)
        else:
            assert np.array_equal(u.xK, ref[0])
            assert np.array_equal(gy, ref[1]) and np.array_equal(gx0, ref[2]), C
        u.close()
    ex = M.Unrolled(A, Y, X0, K=8, checkpoint_interval=2)
    assert rel_linf(ref[0], ex.xK) < 1e-5      # FAST next to the sequential sums
    ex.close()


def test_finite_differences_and_homogeneity(tiny_u):
    p, A = tiny_u
    m = p.matrix
    K = 6
    y = p.y.astype(np.float64) + 0.25
    rng = np.random.default_rng(6)
    x0 = rng.uniform(0.5, 1.5, m.cols)
    xt = p.x_true.astype(np.float64)
    u = M.Unrolled(A, y[None, :], x0[None, :], K=K, checkpoint_interval=2)
    gy, gx0 = u.backward((u.xK - xt).astype(np.float32))
    gy, gx0 = gy[0].astype(np.float64), gx0[0].astype(np.float64)
    u.close()
    y32, x032 = y.astype(np.float32).astype(np.float64), x0.astype(np.float32).astype(np.float64)

    def loss(yy, xx):
        xK = orc.f64_unroll(m.rowidx, m.elements, m.cols, yy, xx, K)
        return 0.5 * float(np.sum((xK - xt) ** 2))

    for _ in range(64):  # 64 seeded directions for each of y and x0
        for g, v, which in ((gy, y32, "y"), (gx0, x032, "x")):
            d = rng.standard_normal(v.size)
            h = 1e-5 * np.linalg.norm(v) / np.linalg.norm(d)
            if which == "y":
                fd = (loss(y32 + h * d, x032) - loss(y32 - h * d, x032)) / (2 * h)
            else:
                fd = (loss(y32, x032 + h * d) - loss(y32, x032 - h * d)) / (2 * h)
            assert abs(np.dot(g, d) - fd) <= 2e-4 * np.linalg.norm(g) * np.linalg.norm(d)
    xK = orc.f64_unroll(m.rowidx, m.elements, m.cols, y32, x032, K)
    assert abs(np.dot(x032, gx0)) <= 1e-4 * np.linalg.norm(x032) * np.linalg.norm(gx0)
    want = float(np.dot(xK - xt, xK))
    assert abs(np.dot(y32, gy) - want) <= 1e-4 * abs(want)


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_small_batch8_k20(mode):
    # 32^3 workload, batch of 8 relaxed-count sinograms, K = 20, checkpoints
    # every 5 iterations
    p = synth.make_problem("small")
    m = p.matrix
    A = _open(m)
    Y = _frames(p, 8, 7)
    u = M.Unrolled(A, Y, None, K=20, checkpoint_interval=5, mode=mode)
    assert u.info()["block_width"] == 8
    g = (u.xK - p.x_true).astype(np.float32)
    gy, gx0 = u.backward(g)
    for f in (0, 3, 7):
        y32 = Y[f].astype(np.float32)
        x0 = np.full(m.cols, np.float32(np.sum(y32.astype(np.float64)) /
                                        np.sum(M.calcColumnSums(A).astype(np.float64))))
        xK64, gy64, gx064 = orc.f64_unroll(m.rowidx, m.elements, m.cols, y32, x0, 20,
                                           gK=g[f].astype(np.float64))
        assert rel_linf(u.xK[f], xK64) < 1e-4
        assert _rel_l2(gy[f], gy64) < TOL_GRAD
        assert _rel_l2(gx0[f], gx064) < TOL_GRAD
    u.close()
    A.close()


def test_unrolled_argument_errors(tiny_u):
    p, A = tiny_u
    with pytest.raises(ValueError):
        M.Unrolled(A, p.y[None, :-1].astype(np.float32))
    with pytest.raises(M.MlemError):
        M.Unrolled(A, p.y[None, :].astype(np.float32), K=0)
    with pytest.raises(M.MlemError):
        M.Unrolled(A, p.y[None, :].astype(np.float32), K=2, mode="atomic")
    u = M.Unrolled(A, p.y[None, :].astype(np.float32), K=2, keep_tape=False)
    with pytest.raises(RuntimeError):
        u.backward(u.xK)
