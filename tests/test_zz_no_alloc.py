"""The allocation-free reconstruction entry point (mlem_reconstruct_noalloc)
against mlem_reconstruct: the same images and FP64 statistics bit for bit, in
FAST and EXACT mode, on two calls in a row (the second on a smaller matrix,
so it reuses scratch the first one left behind); EXACT also equals the fp32
restatement of seqcsr4mlem. Device buffers come from the HIP runtime the
library itself uses (ctypes), so no other allocator is involved."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import mlem_amd as M
from mlem_amd import synth
from mlem_amd._native import lib
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        lib()  # the library's HIP runtime is loaded first; the same soname resolves to it
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


H2D, D2H = 1, 2


class DevBuf:
    def __init__(self, host: np.ndarray):
        self.n = host.nbytes
        self.p = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.p), max(self.n, 16)) == 0
        assert hip().hipMemcpy(self.p, host.ctypes.data, self.n, H2D) == 0

    def get(self, dtype, count) -> np.ndarray:
        out = np.empty(count, dtype)
        assert hip().hipDeviceSynchronize() == 0
        assert hip().hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        hip().hipFree(self.p)


def noalloc(A, y, x0, iterations, mode, stats=True):
    """One call: device copies of y and x0, the iterations, the image back."""
    yd = DevBuf(np.ascontiguousarray(y, np.int32))
    xd = DevBuf(np.ascontiguousarray(x0, np.float32))
    ll = np.zeros(max(iterations, 1))
    ft = np.zeros(max(iterations, 1))
    rc = lib().mlem_reconstruct_noalloc(A.handle, yd.p, y.size, xd.p, x0.size, iterations,
                                        M.MODES[mode], ll.ctypes.data if stats else None,
                                        ft.ctypes.data if stats else None)
    try:
        if rc != 0:
            raise M.MlemError(rc, lib().mlem_last_error().decode())
        x = xd.get(np.float32, x0.size)
    finally:
        yd.free()
        xd.free()
    return x, ll[:iterations], ft[:iterations]


def _open(p):
    m = p.matrix
    return M.Csr4Matrix.from_arrays(m.rows, m.cols, m.rowidx, m.elements, grid=m.grid)


@pytest.mark.parametrize("mode", ["fast", "exact"])
def test_noalloc_matches_reconstruct_two_calls(mode):
    big, small = synth.make_problem("small"), synth.make_problem("tiny")
    for p, iters in ((big, 12), (small, 7)):        # second call: smaller input
        m = p.matrix
        A = _open(p)
        try:
            n = M.calcColumnSums(A)
            x0, _ = M.initImage(A, n, p.y)
            x, ll, ft = noalloc(A, p.y, x0, iters, mode)
            xr, llr, ftr = M.mlem(A, p.y, None, iters, mode=mode, stats=True)
            assert np.array_equal(x, xr), (p.name, mode)
            assert np.array_equal(ll, llr) and np.array_equal(ft, ftr), (p.name, mode)
            if mode == "exact":
                assert np.array_equal(x, orc.mlem(m.rowidx, m.elements, m.cols, p.y, iters)[0])
            # without statistics the call only enqueues; same image
            x2, _, _ = noalloc(A, p.y, x0, iters, mode, stats=False)
            assert np.array_equal(x2, xr)
        finally:
            A.close()


def test_noalloc_rejects_too_large_and_bad_arguments():
    # 10M empty rows need 80 MB of device-order sinogram and ratio scratch
    rows, cols = 10_000_000, 4
    rowidx = np.ones(rows, np.uint64)
    el = np.zeros(1, M.ELEMENT_DTYPE)
    el["value"] = 1.0
    A = M.Csr4Matrix.from_arrays(rows, cols, rowidx, el)
    try:
        with pytest.raises(M.MlemError, match="bytes of device scratch"):
            noalloc(A, np.zeros(rows, np.int32), np.ones(cols, np.float32), 1, "fast")
        with pytest.raises(M.MlemError, match="MLEM_MODE_FAST or MLEM_MODE_EXACT"):
            noalloc(A, np.zeros(rows, np.int32), np.ones(cols, np.float32), 1, "atomic")
    finally:
        A.close()
