"""MI355X-native MLEM reconstruction engine — Python face of libmlem.so.

The names mirror the reference's sequential driver and matrix class so that
tests read like that code:

  Csr4Matrix            helper_files_common/csr4matrix.hpp:30-40 (here: the
                        matrix lives in HBM, CSR + device-built CSC)
  calcColumnSums        mlem_seq/seqcsr4mlem.cpp:37-49
  initImage             seqcsr4mlem.cpp:52-66
  calcFwProj            seqcsr4mlem.cpp:69-91
  calcCorrel            seqcsr4mlem.cpp:94-105
  calcBkProj            seqcsr4mlem.cpp:108-120
  calcUpdate            seqcsr4mlem.cpp:123-134
  mlem                  seqcsr4mlem.cpp:137-189

Every call goes through the C ABI into HIP kernels; errors raise MlemError
(the reference throws std::runtime_error). There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

from ._native import (INIT_GIVEN, INIT_REFERENCE, MODE_ATOMIC, MODE_EXACT, MODE_FAST, MODES,
                      OPEN_BF16, OPEN_EXPAND_SYMMETRY, OPEN_FP16, OPEN_KEEP_ORDER, OPEN_NO_CSC,
                      MlemError, Options, check, lib)

__all__ = [
    "Csr4Matrix", "Recon", "MlemError", "ELEMENT_DTYPE", "probe", "calcColumnSums", "initImage",
    "calcFwProj", "calcCorrel", "calcBkProj", "calcUpdate", "mlem", "MODE_FAST", "MODE_EXACT",
    "MODE_ATOMIC", "calibration_copy", "device_sync",
]

# csr4 element record {u32 column, f32 value} (matrixelement.hpp:37-46)
ELEMENT_DTYPE = np.dtype([("column", "<u4"), ("value", "<f4")])


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _mode(mode) -> int:
    if isinstance(mode, str):
        return MODES[mode]
    return int(mode)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


PRECISIONS = {"fp32": 0, "fp16": OPEN_FP16, "bf16": OPEN_BF16}


def _flags(no_csc: bool, keep_order: bool, precision: str = "fp32") -> int:
    if precision not in PRECISIONS:
        raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
    return ((OPEN_NO_CSC if no_csc else 0) | (OPEN_KEEP_ORDER if keep_order else 0)
            | PRECISIONS[precision])


def probe(path: str) -> dict:
    """Header of a csr4 file (no GPU needed)."""
    rows, cols, grid, sym = C.c_uint32(), C.c_uint32(), (C.c_uint32 * 3)(), C.c_uint32()
    nnz = C.c_uint64()
    check(lib().mlem_probe(os.fsencode(path), C.byref(rows), C.byref(cols), C.byref(nnz), grid,
                           C.byref(sym)))
    return {"rows": rows.value, "cols": cols.value, "nnz": nnz.value, "grid": tuple(grid),
            "symflags": sym.value}


def expand_symmetry_file(src: str, dst: str) -> None:
    """Write the full matrix of a symmetry-reduced csr4 file (no GPU needed)."""
    check(lib().mlem_expand_symmetry_file(os.fsencode(src), os.fsencode(dst)))


class Csr4Matrix:
    """Device-resident csr4 system matrix (Csr4Matrix, csr4matrix.hpp:30-40).

    Csr4Matrix(path)                            load a csr4 file into HBM
    Csr4Matrix.from_arrays(rows, cols, rowidx, elements)   from host memory

    precision "fp16" / "bf16" stores every row as an fp32 scale (the row's
    largest value) times 16-bit quotients; "fp32" keeps the file's values.
    stream_block_bytes (not None) keeps the matrix in pinned host memory and
    streams it to the device in row blocks of at most that many bytes per
    part (0: sized from the free device memory), for matrices larger than HBM;
    the first blocks that fit stream_resident_bytes stay in device memory.
    """

    def __init__(self, path: Optional[str] = None, *, no_csc: bool = False,
                 keep_order: bool = False, precision: str = "fp32",
                 expand_symmetry: bool = False, stream_block_bytes: Optional[int] = None,
                 stream_resident_bytes: int = 0, _handle=None):
        self._h = None
        self.precision = precision
        if _handle is not None:
            self._h = _handle
        else:
            if path is None:
                raise ValueError("path required")
            h = C.c_void_p()
            flags = _flags(no_csc, keep_order, precision)
            if expand_symmetry:
                flags |= OPEN_EXPAND_SYMMETRY
            if stream_block_bytes is None:
                check(lib().mlem_open(os.fsencode(path), flags, C.byref(h)))
            else:
                check(lib().mlem_open_streamed(os.fsencode(path), flags,
                                               int(stream_block_bytes),
                                               int(stream_resident_bytes), C.byref(h)))
            self._h = h
        r, c, n = C.c_uint32(), C.c_uint32(), C.c_uint64()
        check(lib().mlem_dims(self._h, C.byref(r), C.byref(c), C.byref(n)))
        self._rows, self._cols, self._nnz = r.value, c.value, n.value

    @classmethod
    def from_arrays(cls, rows: int, cols: int, rowidx: np.ndarray, elements: np.ndarray,
                    grid=None, no_csc: bool = False, keep_order: bool = False,
                    precision: str = "fp32", stream_block_bytes: Optional[int] = None,
                    stream_resident_bytes: int = 0) -> "Csr4Matrix":
        """grid = (nx, ny, nz) enables the device locality layout."""
        flags = _flags(no_csc, keep_order, precision)
        rowidx = np.ascontiguousarray(rowidx, dtype=np.uint64)
        if elements.dtype != ELEMENT_DTYPE:
            elements = np.ascontiguousarray(elements).view(ELEMENT_DTYPE)
        elements = np.ascontiguousarray(elements)
        if rowidx.shape != (rows,):
            raise ValueError("rowidx must hold one END offset per row")
        nnz = int(elements.shape[0])
        g = None if grid is None else (C.c_uint32 * 3)(*grid)
        h = C.c_void_p()
        if stream_block_bytes is None:
            check(lib().mlem_open_host(rows, cols, nnz, _ptr(rowidx), _ptr(elements), g,
                                       flags, C.byref(h)))
        else:
            check(lib().mlem_open_host_streamed(rows, cols, nnz, _ptr(rowidx), _ptr(elements),
                                                g, flags, int(stream_block_bytes),
                                                int(stream_resident_bytes), C.byref(h)))
        return cls(_handle=h, precision=precision)

    def stream_blocks(self) -> int:
        """Row blocks of a streamed matrix (0 when the matrix is in HBM)."""
        n, r = C.c_uint32(), C.c_uint32()
        check(lib().mlem_stream_blocks(self._h, C.byref(n), C.byref(r)))
        return n.value

    def stream_passes(self) -> int:
        """Passes over the streamed blocks per FAST / EXACT iteration (1: each
        block's CSC built on the device; 2: CSC parts streamed too; 0: in HBM)."""
        n = C.c_int()
        check(lib().mlem_stream_passes(self._h, C.byref(n)))
        return n.value

    def resident_blocks(self) -> int:
        """Blocks of a streamed matrix kept in device memory."""
        n, r = C.c_uint32(), C.c_uint32()
        check(lib().mlem_stream_blocks(self._h, C.byref(n), C.byref(r)))
        return r.value

    def rows(self) -> int:
        return self._rows

    def columns(self) -> int:
        return self._cols

    def elements(self) -> int:
        return self._nnz

    def geometry(self) -> tuple:
        grid, sym = (C.c_uint32 * 3)(), C.c_uint32()
        check(lib().mlem_geometry(self._h, grid, C.byref(sym)))
        return tuple(grid), sym.value

    def launch(self) -> dict:
        """FAST-mode kernels: {"fwd": (family, param), "bwd": (family, param)},
        family "dict512", "dict" or "tiles" (param = elements per lane) or
        "groups" (lanes/row)."""
        v = [C.c_int() for _ in range(4)]
        check(lib().mlem_get_launch(self.handle, *[C.byref(a) for a in v]))
        fam = {0: "groups", 1: "tiles", 2: "dict", 3: "dict512"}
        return {"fwd": (fam[v[0].value], v[1].value), "bwd": (fam[v[2].value], v[3].value)}

    def set_launch(self, fwd, bwd) -> None:
        """fwd, bwd: (family, param) as returned by launch()."""
        fam = {"groups": 0, "tiles": 1, "dict": 2, "dict512": 3}
        check(lib().mlem_set_launch(self.handle, fam[fwd[0]], int(fwd[1]), fam[bwd[0]],
                                    int(bwd[1])))

    def autotune(self) -> dict:
        check(lib().mlem_autotune(self.handle))
        return self.launch()

    @property
    def handle(self):
        if self._h is None:
            raise MlemError(-6, "matrix is closed")
        return self._h

    def close(self) -> None:
        if self._h is not None:
            lib().mlem_close(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def calcColumnSums(matrix: Csr4Matrix) -> np.ndarray:
    norm = np.empty(matrix.columns(), np.float32)
    check(lib().mlem_column_sums(matrix.handle, _ptr(norm), norm.size))
    return norm


def initImage(matrix: Csr4Matrix, norm, lmsino) -> tuple[np.ndarray, float]:
    norm, y = _f32(norm), _i32(lmsino)
    x = np.empty(matrix.columns(), np.float32)
    init = C.c_float()
    check(lib().mlem_init_image(matrix.handle, _ptr(norm), norm.size, _ptr(y), y.size, _ptr(x),
                                C.byref(init)))
    return x, init.value


def calcFwProj(matrix: Csr4Matrix, image, mode="fast") -> np.ndarray:
    x = _f32(image)
    f = np.empty(matrix.rows(), np.float32)
    check(lib().mlem_fwproj(matrix.handle, _ptr(x), x.size, _ptr(f), f.size, _mode(mode)))
    return f


def calcCorrel(matrix: Csr4Matrix, fwproj, lmsino) -> np.ndarray:
    f, y = _f32(fwproj), _i32(lmsino)
    if f.size != y.size:
        raise MlemError(-1, "fwproj and sinogram sizes differ")
    r = np.empty(f.size, np.float32)
    check(lib().mlem_correl(matrix.handle, _ptr(f), _ptr(y), _ptr(r), r.size))
    return r


def calcBkProj(matrix: Csr4Matrix, correlation, mode="fast") -> np.ndarray:
    r = _f32(correlation)
    u = np.empty(matrix.columns(), np.float32)
    check(lib().mlem_bkproj(matrix.handle, _ptr(r), r.size, _ptr(u), u.size, _mode(mode)))
    return u


def calcUpdate(matrix: Csr4Matrix, update, norm, image: np.ndarray) -> np.ndarray:
    """x *= (n != 0) ? u/n : u, in place on `image` (float32, contiguous)."""
    u, n = _f32(update), _f32(norm)
    if image.dtype != np.float32 or not image.flags.c_contiguous:
        raise ValueError("image must be a contiguous float32 array")
    check(lib().mlem_update(matrix.handle, _ptr(u), _ptr(n), _ptr(image), image.size))
    return image


def _options(mode="fast", x0_given=False, checkpoint_interval=0, checkpoint_dir=None,
             verbose=False, use_graph=True, loglik=None, fwd_total=None) -> Options:
    o = Options()
    lib().mlem_options_default(C.byref(o))
    o.mode = _mode(mode)
    o.init = INIT_GIVEN if x0_given else INIT_REFERENCE
    o.checkpoint_interval = int(checkpoint_interval)
    o.checkpoint_dir = os.fsencode(checkpoint_dir) if checkpoint_dir else None
    o.verbose = 1 if verbose else 0
    o.use_graph = 1 if use_graph else 0
    if loglik is not None:
        o.loglik = loglik.ctypes.data_as(C.POINTER(C.c_double))
    if fwd_total is not None:
        o.fwd_total = fwd_total.ctypes.data_as(C.POINTER(C.c_double))
    return o


def mlem(matrix: Csr4Matrix, lmsino, image: Optional[np.ndarray], nIterations: int, *,
         mode="fast", x0_given=False, checkpoint_interval=0, checkpoint_dir=None,
         verbose=False, use_graph=True, stats=False):
    """Reconstruct; returns the image (and (loglik, fwd_total) when stats).

    image: float32[cols], receives the result; holds x0 when x0_given.
    """
    y = _i32(lmsino)
    if image is None:
        image = np.zeros(matrix.columns(), np.float32)
    if image.dtype != np.float32 or not image.flags.c_contiguous:
        raise ValueError("image must be a contiguous float32 array")
    ll = np.zeros(max(nIterations, 1), np.float64) if stats else None
    ft = np.zeros(max(nIterations, 1), np.float64) if stats else None
    o = _options(mode, x0_given, checkpoint_interval, checkpoint_dir, verbose, use_graph, ll, ft)
    check(lib().mlem_reconstruct(matrix.handle, _ptr(y), y.size, _ptr(image), image.size,
                                 int(nIterations), C.byref(o)))
    if stats:
        return image, ll[:nIterations], ft[:nIterations]
    return image


class Recon:
    """Device-resident reconstruction state (sinogram, image, ratio in HBM)."""

    def __init__(self, matrix: Csr4Matrix, lmsino, x0=None, *, mode="fast", use_graph=True):
        self.matrix = matrix
        y = _i32(lmsino)
        x = None if x0 is None else _f32(x0)
        o = _options(mode, x is not None, use_graph=use_graph)
        h = C.c_void_p()
        check(lib().mlem_recon_create(matrix.handle, _ptr(y), y.size, _ptr(x),
                                      matrix.columns(), C.byref(o), C.byref(h)))
        self._h = h

    def run(self, iterations: int) -> None:
        check(lib().mlem_recon_run(self._h, int(iterations)))

    def run_timed(self, iterations: int):
        """Returns (fwd_ms[iters], bwd_ms[iters], span_ms) from HIP events."""
        kf = np.zeros(max(iterations, 1), np.float32)
        kb = np.zeros(max(iterations, 1), np.float32)
        span = C.c_float()
        check(lib().mlem_recon_run_timed(self._h, int(iterations), _ptr(kf), _ptr(kb),
                                         C.byref(span)))
        return kf[:iterations], kb[:iterations], span.value

    def image(self) -> np.ndarray:
        x = np.empty(self.matrix.columns(), np.float32)
        check(lib().mlem_recon_image(self._h, _ptr(x), x.size))
        return x

    @property
    def iterations(self) -> int:
        return lib().mlem_recon_iterations(self._h)

    def close(self) -> None:
        if getattr(self, "_h", None) is not None:
            lib().mlem_recon_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many sinograms reconstructed against one matrix (RHS blocks).

    y: int32 [frames, rows]; budgets: int [frames] or None; rel_tol > 0 stops a
    frame on a small relative log-likelihood change. mode "fast": blocks run
    through the matrix's two-dimensional tiling (deterministic, within 1e-4
    of FP64 MLEM); "exact": every frame bit-identical to the sequential driver.
    """

    def __init__(self, matrix: Csr4Matrix, y, budgets=None, *, rel_tol=0.0, block_width=32,
                 use_graph=True, mode="fast"):
        from ._native import BatchOptions
        self.matrix = matrix
        y = np.ascontiguousarray(y, dtype=np.int32)
        if y.ndim != 2 or y.shape[1] != matrix.rows():
            raise ValueError("y must be [frames, rows]")
        self.frames = y.shape[0]
        b = None if budgets is None else np.ascontiguousarray(budgets, dtype=np.int32)
        if b is not None and b.shape != (self.frames,):
            raise ValueError(f"budgets must be [frames] = [{self.frames}], got {b.shape}")
        o = BatchOptions()
        lib().mlem_batch_options_default(C.byref(o))
        o.rel_tol = float(rel_tol)
        o.block_width = int(block_width)
        o.use_graph = 1 if use_graph else 0
        o.mode = _mode(mode)
        self.mode = mode
        h = C.c_void_p()
        check(lib().mlem_batch_create(matrix.handle, self.frames, _ptr(y), y.shape[1], _ptr(b),
                                      C.byref(o), C.byref(h)))
        self._h = h

    def run(self, max_iterations: int) -> int:
        n = C.c_int()
        check(lib().mlem_batch_run(self._h, int(max_iterations), C.byref(n)))
        return n.value

    def image(self, frame: int) -> np.ndarray:
        x = np.empty(self.matrix.columns(), np.float32)
        check(lib().mlem_batch_image(self._h, int(frame), _ptr(x), x.size))
        return x

    def frame_info(self, frame: int, capacity: int = 4096):
        it = C.c_int()
        ll = np.zeros(capacity, np.float64)
        check(lib().mlem_batch_frame_info(self._h, int(frame), C.byref(it), _ptr(ll), capacity))
        return it.value, ll[:min(it.value, capacity)]

    def progress(self) -> dict:
        a, fi, w, nb = C.c_uint32(), C.c_uint64(), C.c_int(), C.c_uint32()
        check(lib().mlem_batch_progress(self._h, C.byref(a), C.byref(fi), C.byref(w),
                                        C.byref(nb)))
        return {"active": a.value, "frame_iterations": fi.value, "block_width": w.value,
                "blocks": nb.value}

    def tuning(self) -> dict:
        """Measured ms per block-iteration by width (block_width=0 batches);
        width 1 is a lone frame on the single-image EXACT kernels."""
        ms = (C.c_double * 4)()
        check(lib().mlem_batch_tuning(self._h, ms))
        one = C.c_double()
        check(lib().mlem_batch_single_ms(self._h, C.byref(one)))
        return {1: one.value, **{w: ms[i] for i, w in enumerate((4, 8, 16, 32))}}

    def close(self) -> None:
        if getattr(self, "_h", None) is not None:
            lib().mlem_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Unrolled:
    """K unrolled MLEM iterations as a differentiable function of the counts
    y (real-valued, [frames, rows]) and the initial images x0 ([frames, cols];
    None: per-frame uniform sum(y)/sum(n), held constant).

    u = Unrolled(A, y, x0, K)       runs the forward pass; u.xK is x_K
    gy, gx0 = u.backward(dL/dx_K)   reverse mode (recomputes from checkpoints)

    mode="exact": every frame's products summed one by one in file element
    order (x_K bit-identical to mlem() for integer counts); mode="fast": the
    products through the matrix's panel tiling, fixed-order sums.
    """

    def __init__(self, matrix: Csr4Matrix, y, x0=None, K: int = 20, *,
                 checkpoint_interval: int = 0, keep_tape: bool = True, mode="exact"):
        y = np.ascontiguousarray(np.atleast_2d(y), np.float32)
        frames, rows = y.shape
        if rows != matrix.rows():
            raise ValueError("y must be [frames, rows]")
        cols = matrix.columns()
        if x0 is not None:
            x0 = np.ascontiguousarray(np.atleast_2d(x0), np.float32)
            if x0.shape != (frames, cols):
                raise ValueError("x0 must be [frames, cols]")
        self.matrix, self.frames, self.K = matrix, frames, int(K)
        self.xK = np.zeros((frames, cols), np.float32)
        t = C.c_void_p()
        check(lib().mlem_forward_k_mode(matrix.handle, frames, _ptr(y), rows, _ptr(x0), cols,
                                        self.K, int(checkpoint_interval), _mode(mode),
                                        _ptr(self.xK), C.byref(t) if keep_tape else None))
        self._t = t if keep_tape else None

    def backward(self, grad_xK):
        g = np.ascontiguousarray(np.atleast_2d(grad_xK), np.float32)
        if g.shape != self.xK.shape:
            raise ValueError("grad_xK must be [frames, cols]")
        if self._t is None:
            raise RuntimeError("forward pass ran without a tape")
        gy = np.zeros((self.frames, self.matrix.rows()), np.float32)
        gx0 = np.zeros_like(self.xK)
        check(lib().mlem_backward_k(self._t, _ptr(g), _ptr(gy), _ptr(gx0)))
        return gy, gx0

    def info(self) -> dict:
        v = [C.c_int() for _ in range(4)]
        n = C.c_uint64()
        check(lib().mlem_tape_info(self._t, *[C.byref(a) for a in v], C.byref(n)))
        return {"block_width": v[0].value, "checkpoint_interval": v[1].value,
                "checkpoints": v[2].value, "recomputed_iterations": v[3].value,
                "spmm_launches": n.value}

    def close(self) -> None:
        if getattr(self, "_t", None) is not None:
            lib().mlem_tape_destroy(self._t)
            self._t = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def calibration_copy(nbytes: int, repeats: int = 5) -> float:
    ms = C.c_float()
    check(lib().mlem_calibration_copy(int(nbytes), int(repeats), C.byref(ms)))
    return ms.value


def device_sync() -> None:
    check(lib().mlem_device_sync())
