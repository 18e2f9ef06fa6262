"""Synthetic MLEM workloads (libmlem_synth.so): csr4 matrices of random lines
of response through a unit-voxel grid (Siddon lengths), ellipsoid phantoms and
Poisson sinograms, all deterministic in their seeds.

Workloads (seeds: matrix 1000+k, phantom 2000+k, Poisson 3000+k, frame f
4000+f):
  small     32x32x32 grid, 262,144 lines, ~10M nnz, 8 ellipsoids, 1e7 counts
  pet       128x128x48 grid, 1,327,104 lines, ~200M nnz, 16 ellipsoids, 1e8
  large     192x192x64 grid, 4,194,304 lines, ~1B nnz, 16 ellipsoids, 1e9
  tiny      8x8x8 grid, 4,096 lines (tests)
"""
from __future__ import annotations

import ctypes as C
import weakref
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from ._native import synth_lib

ELEMENT_DTYPE = np.dtype([("column", "<u4"), ("value", "<f4")])

WORKLOADS = {
    "tiny": dict(key=9, grid=(8, 8, 8), lors=4096, ellipsoids=3, semi=(1.5, 3.0), counts=2e5),
    "small": dict(key=0, grid=(32, 32, 32), lors=262144, ellipsoids=8, semi=(2.0, 10.0),
                  counts=1e7),
    "pet": dict(key=1, grid=(128, 128, 48), lors=1327104, ellipsoids=16, semi=(4.0, 20.0),
                counts=1e8),
    "large": dict(key=3, grid=(192, 192, 64), lors=4194304, ellipsoids=16, semi=(6.0, 30.0),
                  counts=1e9),
}


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed with status {rc}")


@dataclass
class SynthMatrix:
    rows: int
    cols: int
    nnz: int
    grid: tuple
    rowidx: np.ndarray      # uint64[rows], exclusive END offsets (csr4)
    elements: np.ndarray    # ELEMENT_DTYPE[nnz]
    _keep: list = field(default_factory=list, repr=False)

    def row_begin(self) -> np.ndarray:
        b = np.zeros(self.rows, np.uint64)
        b[1:] = self.rowidx[:-1]
        return b

    def write_csr4(self, path: str) -> None:
        nx, ny, nz = self.grid
        _check(synth_lib().mlem_synth_write_csr4(str(path).encode(), nx, ny, nz, self.rows,
                                                 self.nnz, self.rowidx.ctypes.data,
                                                 self.elements.ctypes.data), "write_csr4")


def make_matrix(grid, lors: int, seed: int, threads: int = 0) -> SynthMatrix:
    nx, ny, nz = grid
    L = synth_lib()
    prow, pel, nnz = C.c_void_p(), C.c_void_p(), C.c_uint64()
    _check(L.mlem_synth_matrix(nx, ny, nz, lors, seed, threads, C.byref(prow), C.byref(pel),
                               C.byref(nnz)), "synth_matrix")
    n = nnz.value
    rowidx = np.ctypeslib.as_array(C.cast(prow, C.POINTER(C.c_uint64)), shape=(lors,))
    raw = np.ctypeslib.as_array(C.cast(pel, C.POINTER(C.c_uint64)), shape=(max(n, 1),))[:n]
    elements = raw.view(ELEMENT_DTYPE)
    m = SynthMatrix(lors, nx * ny * nz, n, tuple(grid), rowidx, elements)
    # free the malloc'd buffers when the arrays are gone
    weakref.finalize(rowidx, L.mlem_synth_free, prow)
    weakref.finalize(raw, L.mlem_synth_free, pel)
    m._keep = [raw]
    return m


def make_phantom(grid, ellipsoids: int, seed: int, semi=(2.0, 10.0),
                 scale: Optional[np.ndarray] = None) -> np.ndarray:
    nx, ny, nz = grid
    x = np.empty(nx * ny * nz, np.float32)
    sc = None
    if scale is not None:
        sc = np.ascontiguousarray(scale, np.float32)
        assert sc.size >= ellipsoids
    _check(synth_lib().mlem_synth_phantom(nx, ny, nz, ellipsoids, seed, float(semi[0]),
                                          float(semi[1]),
                                          None if sc is None else sc.ctypes.data,
                                          x.ctypes.data), "synth_phantom")
    return x


def project(m: SynthMatrix, x: np.ndarray, threads: int = 0) -> np.ndarray:
    """A x in double (the noiseless sinogram)."""
    x = np.ascontiguousarray(x, np.float32)
    ybar = np.empty(m.rows, np.float64)
    _check(synth_lib().mlem_synth_project(m.rows, m.rowidx.ctypes.data, m.elements.ctypes.data,
                                          x.ctypes.data, ybar.ctypes.data, threads), "project")
    return ybar


def poisson(ybar: np.ndarray, total: float, seed: int, threads: int = 0) -> np.ndarray:
    ybar = np.ascontiguousarray(ybar, np.float64)
    y = np.empty(ybar.size, np.int32)
    s = C.c_int64()
    _check(synth_lib().mlem_synth_poisson(ybar.size, ybar.ctypes.data, float(total), seed,
                                          threads, y.ctypes.data, C.byref(s)), "poisson")
    return y


def uniform(seed: int, first: int, n: int) -> np.ndarray:
    out = np.empty(n, np.float64)
    _check(synth_lib().mlem_synth_uniform(seed, first, n, out.ctypes.data), "uniform")
    return out


def siddon(p0, p1, grid):
    """Voxels and intersection lengths of the segment p0 -> p1."""
    a = np.ascontiguousarray(p0, np.float64)
    b = np.ascontiguousarray(p1, np.float64)
    nx, ny, nz = grid
    cap = nx + ny + nz + 3
    cols = np.zeros(cap, np.uint32)
    lens = np.zeros(cap, np.float32)
    k = synth_lib().mlem_synth_siddon(a.ctypes.data, b.ctypes.data, nx, ny, nz, cap,
                                      cols.ctypes.data, lens.ctypes.data)
    if k < 0:
        raise ValueError("bad arguments")
    return cols[:k], lens[:k]


@dataclass
class Problem:
    name: str
    matrix: SynthMatrix
    x_true: np.ndarray
    ybar: np.ndarray
    y: np.ndarray


def make_problem(name: str, threads: int = 0, lors: Optional[int] = None) -> Problem:
    """Matrix, phantom and Poisson sinogram of a named workload. `lors`
    overrides the number of lines (smaller test variants)."""
    w = WORKLOADS[name]
    k = w["key"]
    m = make_matrix(w["grid"], lors or w["lors"], 1000 + k, threads)
    x = make_phantom(w["grid"], w["ellipsoids"], 2000 + k, w["semi"])
    ybar = project(m, x, threads)
    y = poisson(ybar, w["counts"], 3000 + k, threads)
    return Problem(name, m, x, ybar, y)


def frame_scales(frames: int, ellipsoids: int) -> list[np.ndarray]:
    """Per-frame ellipsoid intensity multipliers, uniform in [0.2, 2)."""
    return [(0.2 + 1.8 * uniform(4000 + f, 0, ellipsoids)).astype(np.float32)
            for f in range(frames)]


def frame_counts(frames: int, lo: float = 1e5, hi: float = 1e8, seed: int = 4999) -> np.ndarray:
    """Total counts per frame, log-uniform in [lo, hi]."""
    u = uniform(seed, 0, frames)
    return np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))


def frame_budgets(frames: int, lo: int = 10, hi: int = 100, seed: int = 4998) -> np.ndarray:
    """Iteration budget per frame, uniform integers in [lo, hi]."""
    u = uniform(seed, 0, frames)
    return (lo + np.floor(u * (hi - lo + 1))).astype(np.int32)


def dynamic_frames(problem: Problem, frames: int, lo: float = 1e5, hi: float = 1e8) -> np.ndarray:
    """Sinograms of a dynamic series on a workload's matrix: frame f is the
    workload's phantom with per-ellipsoid intensity scales U(0.2, 2) (seed
    4000+f), projected and Poisson-sampled (seed 5000+f) at a total count
    log-uniform in [lo, hi]. int32 [frames, rows]."""
    w = WORKLOADS[problem.name]
    counts = frame_counts(frames, lo, hi)
    Y = np.empty((frames, problem.matrix.rows), np.int32)
    for f, sc in enumerate(frame_scales(frames, w["ellipsoids"])):
        x = make_phantom(w["grid"], w["ellipsoids"], 2000 + w["key"], w["semi"], scale=sc)
        Y[f] = poisson(project(problem.matrix, x), counts[f], 5000 + f)
    return Y
