"""ctypes wrapper of the CPU oracle (liboracle.so, built from mlem_oracle.c).

TEST INFRASTRUCTURE ONLY. Imported by tests/, __graft_entry__.smoke() and the
benchmark's CPU-baseline leg — never by the product package. See the header
of mlem_oracle.c for what each restatement follows in the reference.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path
from typing import Optional

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "lib" / "liboracle.so"
ELEMENT_DTYPE = np.dtype([("column", "<u4"), ("value", "<f4")])

_lib = None
_P = C.c_void_p
_u32 = C.c_uint32


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        sig = {
            "orc_column_sums": (None, [_u32, _u32, _P, _P, _P]),
            "orc_init_image": (C.c_float, [_u32, _P, _u32, _P, _P, _P, _P]),
            "orc_fwproj": (None, [_u32, _P, _P, _P, _P]),
            "orc_correl": (None, [_u32, _P, _P, _P]),
            "orc_bkproj": (None, [_u32, _u32, _P, _P, _P, _P]),
            "orc_update": (None, [_u32, _P, _P, _P]),
            "orc_mlem": (C.c_int, [_u32, _u32, _P, _P, _P, _P, C.c_int, C.c_int, _P, _P, _P, _P]),
            "orc_omp_mlem": (C.c_int, [_u32, _u32, _P, _P, _P, _P, C.c_int, C.c_int, C.c_int, _P]),
            "orc_f64_column_sums": (None, [_u32, _u32, _P, _P, _P]),
            "orc_f64_fwproj": (None, [_u32, _P, _P, _P, _P]),
            "orc_f64_bkproj": (None, [_u32, _u32, _P, _P, _P, _P]),
            "orc_f64_mlem": (C.c_int, [_u32, _u32, _P, _P, _P, _P, C.c_int, C.c_int, _P, _P]),
            "orc_f64_mlem_scaled": (C.c_int, [_u32, _u32, _P, _P, _P, _P, _P, C.c_int, C.c_int,
                                              _P, _P]),
            "orc_quantize": (None, [_u32, _P, _P, C.c_int, _P, _P]),
            "orc_f64_mlem_par": (C.c_int, [_u32, _u32, _P, _P, _P, _P, _P, C.c_int, C.c_int,
                                           C.c_int, _P, _P]),
            "orc_f64_unroll": (C.c_int, [_u32, _u32, _P, _P, _P, _P, _P, C.c_int, _P, _P, _P,
                                         _P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _mat(rowidx, elements):
    rowidx = np.ascontiguousarray(rowidx, np.uint64)
    el = np.ascontiguousarray(elements)
    if el.dtype != ELEMENT_DTYPE:
        el = el.view(ELEMENT_DTYPE)
    return rowidx, el


def column_sums(rowidx, elements, cols) -> np.ndarray:
    rowidx, el = _mat(rowidx, elements)
    n = np.empty(cols, np.float32)
    lib().orc_column_sums(rowidx.size, cols, _p(rowidx), _p(el), _p(n))
    return n


def init_image(norm, y):
    norm = np.ascontiguousarray(norm, np.float32)
    y = np.ascontiguousarray(y, np.int32)
    x = np.empty(norm.size, np.float32)
    sn, si = C.c_double(), C.c_double()
    v = lib().orc_init_image(norm.size, _p(norm), y.size, _p(y), _p(x), C.byref(sn), C.byref(si))
    return x, v, sn.value, si.value


def fwproj(rowidx, elements, x) -> np.ndarray:
    rowidx, el = _mat(rowidx, elements)
    x = np.ascontiguousarray(x, np.float32)
    f = np.empty(rowidx.size, np.float32)
    lib().orc_fwproj(rowidx.size, _p(rowidx), _p(el), _p(x), _p(f))
    return f


def correl(f, y) -> np.ndarray:
    f = np.ascontiguousarray(f, np.float32)
    y = np.ascontiguousarray(y, np.int32)
    r = np.empty(f.size, np.float32)
    lib().orc_correl(f.size, _p(f), _p(y), _p(r))
    return r


def bkproj(rowidx, elements, r, cols) -> np.ndarray:
    rowidx, el = _mat(rowidx, elements)
    r = np.ascontiguousarray(r, np.float32)
    u = np.empty(cols, np.float32)
    lib().orc_bkproj(rowidx.size, cols, _p(rowidx), _p(el), _p(r), _p(u))
    return u


def update(u, norm, x) -> np.ndarray:
    u = np.ascontiguousarray(u, np.float32)
    norm = np.ascontiguousarray(norm, np.float32)
    x = np.array(x, np.float32, copy=True)
    lib().orc_update(x.size, _p(u), _p(norm), _p(x))
    return x


def mlem(rowidx, elements, cols, y, iters, x0: Optional[np.ndarray] = None):
    """fp32 seqcsr4mlem restatement. Returns (x, image_sum, zeros, subnormals,
    iter_seconds)."""
    rowidx, el = _mat(rowidx, elements)
    y = np.ascontiguousarray(y, np.int32)
    x = np.zeros(cols, np.float32) if x0 is None else np.array(x0, np.float32, copy=True)
    n = max(iters, 1)
    s, z, sn, t = (np.zeros(n, np.float32), np.zeros(n, np.int32), np.zeros(n, np.int32),
                   np.zeros(n, np.float64))
    rc = lib().orc_mlem(rowidx.size, cols, _p(rowidx), _p(el), _p(y), _p(x), iters,
                        0 if x0 is None else 1, _p(s), _p(z), _p(sn), _p(t))
    if rc:
        raise MemoryError("oracle allocation failed")
    return x, s[:iters], z[:iters], sn[:iters], t[:iters]


def omp_mlem(rowidx, elements, cols, y, iters, x0=None, threads=0):
    rowidx, el = _mat(rowidx, elements)
    y = np.ascontiguousarray(y, np.int32)
    x = np.zeros(cols, np.float32) if x0 is None else np.array(x0, np.float32, copy=True)
    t = np.zeros(max(iters, 1), np.float64)
    rc = lib().orc_omp_mlem(rowidx.size, cols, _p(rowidx), _p(el), _p(y), _p(x), iters,
                            0 if x0 is None else 1, threads, _p(t))
    if rc:
        raise MemoryError("oracle allocation failed")
    return x, t[:iters]


def f64_column_sums(rowidx, elements, cols) -> np.ndarray:
    rowidx, el = _mat(rowidx, elements)
    n = np.empty(cols, np.float64)
    lib().orc_f64_column_sums(rowidx.size, cols, _p(rowidx), _p(el), _p(n))
    return n


def f64_fwproj(rowidx, elements, x) -> np.ndarray:
    rowidx, el = _mat(rowidx, elements)
    x = np.ascontiguousarray(x, np.float64)
    f = np.empty(rowidx.size, np.float64)
    lib().orc_f64_fwproj(rowidx.size, _p(rowidx), _p(el), _p(x), _p(f))
    return f


def f64_bkproj(rowidx, elements, r, cols) -> np.ndarray:
    rowidx, el = _mat(rowidx, elements)
    r = np.ascontiguousarray(r, np.float64)
    u = np.empty(cols, np.float64)
    lib().orc_f64_bkproj(rowidx.size, cols, _p(rowidx), _p(el), _p(r), _p(u))
    return u


def f64_mlem(rowidx, elements, cols, y, iters, x0=None):
    """FP64 MLEM. Returns (x, loglik[iters], fwd_total[iters]); loglik[k] is
    the Poisson log-likelihood of the iterate entering iteration k."""
    rowidx, el = _mat(rowidx, elements)
    y = np.ascontiguousarray(y, np.float64)
    x = np.zeros(cols, np.float64) if x0 is None else np.array(x0, np.float64, copy=True)
    n = max(iters, 1)
    ll, ft = np.zeros(n), np.zeros(n)
    rc = lib().orc_f64_mlem(rowidx.size, cols, _p(rowidx), _p(el), _p(y), _p(x), iters,
                            0 if x0 is None else 1, _p(ll), _p(ft))
    if rc:
        raise MemoryError("oracle allocation failed")
    return x, ll[:iters], ft[:iters]


QUANT_KINDS = {"fp16": 1, "bf16": 2}


def quantize(rowidx, elements, kind: str):
    """Reduced-precision matrix: (elements with q_ij values, rscale[rows])."""
    rowidx, el = _mat(rowidx, elements)
    out = np.empty_like(el)
    s = np.empty(rowidx.size, np.float32)
    lib().orc_quantize(rowidx.size, _p(rowidx), _p(el), QUANT_KINDS[kind], _p(out), _p(s))
    return out, s


def f64_mlem_scaled(rowidx, elements, rscale, cols, y, iters, x0=None):
    """FP64 MLEM on a_ij = rscale[i] * value_ij (see f64_mlem)."""
    rowidx, el = _mat(rowidx, elements)
    rscale = np.ascontiguousarray(rscale, np.float32)
    y = np.ascontiguousarray(y, np.float64)
    x = np.zeros(cols, np.float64) if x0 is None else np.array(x0, np.float64, copy=True)
    n = max(iters, 1)
    ll, ft = np.zeros(n), np.zeros(n)
    rc = lib().orc_f64_mlem_scaled(rowidx.size, cols, _p(rowidx), _p(el), _p(rscale), _p(y),
                                   _p(x), iters, 0 if x0 is None else 1, _p(ll), _p(ft))
    if rc:
        raise MemoryError("oracle allocation failed")
    return x, ll[:iters], ft[:iters]


def f64_mlem_par(rowidx, elements, cols, y, iters, rscale=None, x0=None, threads=0):
    """FP64 MLEM parallel over rows (large matrices); rscale as in
    f64_mlem_scaled (None: unscaled). Same results as f64_mlem up to the
    order of FP64 sums."""
    rowidx, el = _mat(rowidx, elements)
    rs = None if rscale is None else np.ascontiguousarray(rscale, np.float32)
    y = np.ascontiguousarray(y, np.float64)
    x = np.zeros(cols, np.float64) if x0 is None else np.array(x0, np.float64, copy=True)
    n = max(iters, 1)
    ll, ft = np.zeros(n), np.zeros(n)
    rc = lib().orc_f64_mlem_par(rowidx.size, cols, _p(rowidx), _p(el), _p(rs), _p(y), _p(x),
                                iters, 0 if x0 is None else 1, threads, _p(ll), _p(ft))
    if rc:
        raise MemoryError("oracle allocation failed")
    return x, ll[:iters], ft[:iters]


def f64_unroll(rowidx, elements, cols, y, x0, K, gK=None, rscale=None):
    """K unrolled MLEM iterations in FP64 from x0 with real counts y.
    Returns x_K, or (x_K, dL/dy, dL/dx0) when gK = dL/dx_K is given."""
    rowidx, el = _mat(rowidx, elements)
    rs = None if rscale is None else np.ascontiguousarray(rscale, np.float32)
    y = np.ascontiguousarray(y, np.float64)
    x0 = np.ascontiguousarray(x0, np.float64)
    xK = np.empty(cols, np.float64)
    g = None if gK is None else np.ascontiguousarray(gK, np.float64)
    gy = None if gK is None else np.empty(rowidx.size, np.float64)
    gx0 = None if gK is None else np.empty(cols, np.float64)
    rc = lib().orc_f64_unroll(rowidx.size, cols, _p(rowidx), _p(el), _p(rs), _p(y), _p(x0), K,
                              _p(xK), _p(g), _p(gy), _p(gx0))
    if rc:
        raise MemoryError("oracle allocation failed")
    return xK if gK is None else (xK, gy, gx0)
