"""csr4 symmetry-flag expansion (host code; no GPU).

The expansion regenerates the elements a symmetry-reduced csr4 file omits
(visualizematrix/symhandler.hpp:27-101, symhandler.cpp:33-115,
dispcsr4spy.cpp:126-140). Pinned by: hand-derived detector maps of small
scanners, involution/permutation/commutation of the maps, and the defining
properties of the expanded matrix — it contains every stored element, it is
closed under every generated symmetry, and every element is the image of a
stored one (the orbit of the stored set, nothing more).
"""
from __future__ import annotations

import itertools
import struct

import numpy as np
import pytest

import mlem_amd as M

EL = M.ELEMENT_DTYPE


def write_csr4(path, flags, rows, cols, rowidx, el, *, nblocks, blocksize, nrings, nlayers=1,
               grid):
    """csr4 v4 bytes from the documented layout (csr4matrix.cpp:32-147)."""
    scan = struct.pack("<I", nlayers)
    for _ in range(nlayers):
        scan += struct.pack("<5f", 2.0, 2.0, 10.0, 50.0, 0.08)
    scan += struct.pack("<II", nblocks, blocksize) + struct.pack(f"<{blocksize - 1}f",
                                                                 *([0.1] * (blocksize - 1)))
    scan += struct.pack("<I", nrings) + struct.pack(f"<{nrings - 1}f", *([0.2] * (nrings - 1)))
    scan += struct.pack("<3f", 1.0, 1.0, 1.0) + struct.pack("<3I", *grid)
    head = b"PCSR" + struct.pack("<BBxxIIQI", 4, flags, rows, cols, int(el.size), len(scan))
    with open(path, "wb") as f:
        f.write(head + scan + np.asarray(rowidx, "<u8").tobytes() + el.tobytes())


def read_csr4(path):
    info = M.probe(str(path))
    raw = open(path, "rb").read()
    scb = struct.unpack_from("<I", raw, 24)[0]
    off = 28 + scb
    rowidx = np.frombuffer(raw, "<u8", info["rows"], off)
    el = np.frombuffer(raw, EL, info["nnz"], off + 8 * info["rows"])
    return info, rowidx, el


def triples(rowidx, el):
    out = {}
    b = 0
    for r, e in enumerate(rowidx):
        for k in range(b, int(e)):
            out[(r, int(el["column"][k]))] = float(el["value"][k])
        b = int(e)
    return out


class Maps:
    """The scanner's mirror maps: detectors are numbered position-in-layer
    fastest, then layer, then ring; x and y mirror the position inside every
    layer ring about offsets blocksize-1 and blocksize*(nblocks/2+1)-1, z
    reverses the ring order; voxels are mirrored in the grid."""

    def __init__(self, nblocks, blocksize, nrings, nlayers, grid):
        self.L = nblocks * blocksize
        self.R = self.L * nlayers
        self.nc = self.R * nrings
        self.nrings = nrings
        self.off = {1: blocksize - 1, 2: blocksize * (nblocks // 2 + 1) - 1}
        self.grid = grid

    def det(self, d, s):
        ring, rest = divmod(d, self.R)
        layer, p = divmod(rest, self.L)
        for bit in (1, 2):
            if s & bit:
                p = (self.off[bit] - p) % self.L
        if s & 4:
            ring = self.nrings - 1 - ring
        return ring * self.R + layer * self.L + p

    def vox(self, v, s):
        nx, ny, nz = self.grid
        z, rem = divmod(v, nx * ny)
        y, x = divmod(rem, nx)
        if s & 1:
            x = nx - 1 - x
        if s & 2:
            y = ny - 1 - y
        if s & 4:
            z = nz - 1 - z
        return x + nx * (y + ny * z)

    def row(self, r, s):
        a, b = divmod(r, self.nc)
        a, b = self.det(a, s), self.det(b, s)
        return min(a, b) * self.nc + max(a, b)


def test_detector_maps_small_scanner_by_hand():
    # 2 blocks of 2 crystals, one layer, one ring: x maps 0<->1, 2<->3 and y
    # maps 0<->3, 1<->2 (offsets 1 and 3 on a ring of 4)
    m = Maps(nblocks=2, blocksize=2, nrings=1, nlayers=1, grid=(2, 2, 1))
    assert [m.det(d, 1) for d in range(4)] == [1, 0, 3, 2]
    assert [m.det(d, 2) for d in range(4)] == [3, 2, 1, 0]
    assert [m.det(d, 3) for d in range(4)] == [2, 3, 0, 1]
    m = Maps(nblocks=4, blocksize=3, nrings=2, nlayers=1, grid=(4, 4, 2))
    assert [m.det(d, 4) for d in (0, 5, 11, 12, 23)] == [12, 17, 23, 0, 11]


@pytest.mark.parametrize("cfg", [dict(nblocks=4, blocksize=3, nrings=2, nlayers=1),
                                 dict(nblocks=6, blocksize=2, nrings=3, nlayers=2),
                                 dict(nblocks=3, blocksize=4, nrings=1, nlayers=1)])
def test_maps_are_commuting_involutions(cfg):
    # with an odd number of blocks the y mirror is not a scanner symmetry
    # (symhandler.cpp:42-46): x and y then compose to a rotation
    m = Maps(**cfg, grid=(5, 4, 3))
    syms = [s for s in range(1, 8) if cfg["nblocks"] % 2 == 0 or not s & 2]
    for s in syms:
        perm = [m.det(d, s) for d in range(m.nc)]
        assert sorted(perm) == list(range(m.nc))
        assert all(m.det(m.det(d, s), s) == d for d in range(m.nc))
        assert all(m.vox(m.vox(v, s), s) == v for v in range(60))
    for a, b in itertools.product(syms, repeat=2):
        assert all(m.det(m.det(d, a), b) == m.det(d, a ^ b) for d in range(m.nc))
    if cfg["nblocks"] % 2:
        assert any(m.det(m.det(d, 3), 3) != d for d in range(m.nc))


def _reduced_problem(rng, flags, cfg, grid):
    m = Maps(**cfg, grid=grid)
    rows, cols = m.nc * m.nc, grid[0] * grid[1] * grid[2]
    stored = {}
    for _ in range(150):
        a, b = sorted(rng.choice(m.nc, 2, replace=False))
        r = int(a) * m.nc + int(b)
        for c in rng.choice(cols, int(rng.integers(1, 6)), replace=False):
            stored[(r, int(c))] = float(np.float32(rng.uniform(0.1, 2.0)))
    by_row = [[] for _ in range(rows)]
    for (r, c), v in sorted(stored.items()):
        by_row[r].append((c, v))
    rowidx = np.cumsum([len(x) for x in by_row]).astype(np.uint64)
    el = np.array([e for x in by_row for e in x], dtype=EL)
    return m, rows, cols, rowidx, el, stored


@pytest.mark.parametrize("flags", [1, 2, 4, 3, 5, 6, 7])
def test_expansion_is_the_orbit_of_the_stored_elements(tmp_path, flags):
    cfg = dict(nblocks=4, blocksize=3, nrings=2, nlayers=1)
    grid = (4, 4, 2)
    rng = np.random.default_rng(flags)
    m, rows, cols, rowidx, el, stored = _reduced_problem(rng, flags, cfg, grid)
    src, dst = tmp_path / "reduced.csr4", tmp_path / "full.csr4"
    write_csr4(src, flags, rows, cols, rowidx, el, grid=grid, **cfg)
    assert M.probe(str(src))["symflags"] == flags
    M.expand_symmetry_file(str(src), str(dst))
    info, xrow, xel = read_csr4(dst)
    assert info["symflags"] == 0 and info["rows"] == rows and info["cols"] == cols
    full = triples(xrow, xel)
    # columns ascending and unique inside every row
    b = 0
    for e in xrow:
        cs = xel["column"][b:int(e)]
        assert np.all(np.diff(cs.astype(np.int64)) > 0)
        b = int(e)
    syms = [s for s in range(1, 8) if s & ~flags == 0]
    # contains the stored elements with their values
    assert all(full[k] == v for k, v in stored.items())
    # closed under every generated symmetry
    for (r, c), v in full.items():
        for s in syms:
            assert (m.row(r, s), m.vox(c, s)) in full
    # nothing but images of stored elements
    images = {(m.row(r, s), m.vox(c, s)) for (r, c) in stored for s in [0] + syms}
    assert set(full) == images


def test_stored_value_wins_over_images(tmp_path):
    cfg = dict(nblocks=2, blocksize=2, nrings=1, nlayers=1)
    grid = (2, 2, 1)
    m = Maps(**cfg, grid=grid)
    rows, cols = 16, 4
    # row (0,1) voxel 0 = 1.0; its x image is row (1,0)->(0,1) voxel 1, which
    # is also stored with 5.0: the stored value stays
    by_row = [[] for _ in range(rows)]
    by_row[1] = [(0, 1.0), (1, 5.0)]
    rowidx = np.cumsum([len(x) for x in by_row]).astype(np.uint64)
    el = np.array([e for x in by_row for e in x], dtype=EL)
    src, dst = tmp_path / "r.csr4", tmp_path / "f.csr4"
    write_csr4(src, 1, rows, cols, rowidx, el, grid=grid, **cfg)
    M.expand_symmetry_file(str(src), str(dst))
    full = triples(*read_csr4(dst)[1:])
    assert m.row(1, 1) == 1 and m.vox(0, 1) == 1
    assert full == {(1, 0): 1.0, (1, 1): 5.0}


def test_unflagged_file_is_copied_and_bad_shapes_fail(tmp_path):
    cfg = dict(nblocks=2, blocksize=2, nrings=1, nlayers=1)
    rowidx = np.array([0] * 15 + [2], np.uint64)
    el = np.array([(0, 1.0), (3, 2.0)], dtype=EL)
    src, dst = tmp_path / "a.csr4", tmp_path / "b.csr4"
    write_csr4(src, 0, 16, 4, rowidx, el, grid=(2, 2, 1), **cfg)
    M.expand_symmetry_file(str(src), str(dst))
    assert open(src, "rb").read() == open(dst, "rb").read()
    # a flagged file whose rows are not detector pairs is refused
    bad = tmp_path / "bad.csr4"
    write_csr4(bad, 1, 15, 4, rowidx[1:], el, grid=(2, 2, 1), **cfg)
    with pytest.raises(M.MlemError, match="one row per detector pair"):
        M.expand_symmetry_file(str(bad), str(dst))
    # so is a y flag on a scanner with an odd number of blocks
    odd = tmp_path / "odd.csr4"
    write_csr4(odd, 2, 36, 4, np.array([0] * 35 + [2], np.uint64), el, grid=(2, 2, 1),
               nblocks=3, blocksize=2, nrings=1, nlayers=1)
    with pytest.raises(M.MlemError, match="odd number of blocks"):
        M.expand_symmetry_file(str(odd), str(dst))


@pytest.mark.gpu
def test_open_with_expansion_matches_expanded_file(tmp_path):
    import os
    import subprocess
    from conftest import ROOT
    from oracle import oracle as orc
    cfg = dict(nblocks=4, blocksize=3, nrings=2, nlayers=1)
    grid = (4, 4, 2)
    rng = np.random.default_rng(11)
    m, rows, cols, rowidx, el, stored = _reduced_problem(rng, 7, cfg, grid)
    src, dst = tmp_path / "reduced.csr4", tmp_path / "full.csr4"
    write_csr4(src, 7, rows, cols, rowidx, el, grid=grid, **cfg)
    M.expand_symmetry_file(str(src), str(dst))
    _, xrow, xel = read_csr4(dst)
    y = rng.integers(0, 40, rows).astype(np.int32)
    A = M.Csr4Matrix(str(src), expand_symmetry=True)
    B = M.Csr4Matrix(str(dst))
    assert A.elements() == B.elements() == xel.size
    assert A.geometry()[1] == 0
    for mode in ("exact", "fast"):
        xa = M.mlem(A, y, None, 8, mode=mode)
        assert np.array_equal(xa, M.mlem(B, y, None, 8, mode=mode))
    assert np.array_equal(M.mlem(A, y, None, 8, mode="exact"),
                          orc.mlem(xrow, xel, cols, y, 8)[0])
    # without the flag the stored elements are used as they are
    C_ = M.Csr4Matrix(str(src))
    assert C_.elements() == el.size and C_.geometry()[1] == 7
    for h in (A, B, C_):
        h.close()
    # command line: MLEM_SYMMETRY=expand
    sino = tmp_path / "y.sino"
    y.tofile(sino)
    cli = str(ROOT / "mlem_amd" / "bin" / "mlem_cli")
    outs = []
    for mat, env in ((src, "expand"), (dst, "stored")):
        out = tmp_path / f"x_{env}.img"
        subprocess.run([cli, str(mat), str(sino), str(out), "8"], check=True, capture_output=True,
                       env=dict(os.environ, MLEM_MODE="exact", MLEM_SYMMETRY=env))
        outs.append(np.fromfile(out, np.float32))
    assert np.array_equal(outs[0], outs[1])
