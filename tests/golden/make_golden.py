"""Writes the committed golden fixtures of tests/golden/.

The reference repository ships no tests, fixtures or known-answer vectors, so
these fixtures are built here from first principles:

  tiny3x4.csr4   a 3x4 csr4 v4 file assembled byte by byte with struct.pack
                 from the layout that Csr4Matrix::Csr4Matrix parses
                 (helper_files_common/csr4matrix.cpp:59-147): 28-byte header,
                 60-byte scanner block, u64 row END offsets, {u32,f32} records.
  tiny3x4.sino   raw int32 counts (helper_files_common/vector.hpp:33-53).
  tiny3x4.json   one MLEM iteration of seqcsr4mlem.cpp:137-189 evaluated in
                 exact rational arithmetic (fractions.Fraction): norm, x0, f,
                 r, u, x1, and x2 — the real-number answer that every fp32 and
                 FP64 implementation must approximate.

Run: python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import struct
from fractions import Fraction
from pathlib import Path

HERE = Path(__file__).resolve().parent

# 3 rows x 4 columns; column 2 of row 0 and column 3 of row 2 make column
# sums differ; every value is exactly representable in float32.
ROWS = [
    [(0, 1.0), (2, 2.0)],
    [(1, 0.5)],
    [(0, 3.0), (1, 1.0), (3, 4.0)],
]
NCOLS = 4
GRID = (2, 2, 1)
SINO = [2, 3, 5]


def csr4_bytes(rows, ncols, grid, flags=0) -> bytes:
    nnz = sum(len(r) for r in rows)
    scanner = b"".join([
        struct.pack("<I", 1),                          # nLayers
        struct.pack("<5f", 1.0, 1.0, 1.0, 1.0, 0.0),  # w h d r mu
        struct.pack("<I", 1),                          # nBlocks
        struct.pack("<I", 1),                          # blocksize (no gaps)
        struct.pack("<I", 1),                          # nRings (no gaps)
        struct.pack("<3f", 1.0, 1.0, 1.0),             # voxel size
        struct.pack("<3I", *grid),                     # grid dims
    ])
    assert len(scanner) == 60
    header = b"PCSR" + struct.pack("<BBxx", 4, flags) + struct.pack(
        "<IIQI", len(rows), ncols, nnz, len(scanner))
    assert len(header) == 28
    ends, acc = [], 0
    for r in rows:
        acc += len(r)
        ends.append(acc)
    rowidx = struct.pack(f"<{len(rows)}Q", *ends)
    elems = b"".join(struct.pack("<If", c, v) for r in rows for (c, v) in r)
    return header + scanner + rowidx + elems


def exact_mlem(rows, ncols, y, iters):
    F = Fraction
    norm = [F(0)] * ncols
    for r in rows:
        for c, v in r:
            norm[c] += F(v)
    x0 = F(sum(y)) / sum(norm)
    x = [x0] * ncols
    out = {"norm": [float(v) for v in norm], "x0": float(x0), "iterates": []}
    for it in range(iters):
        f = [sum(F(v) * x[c] for c, v in r) for r in rows]
        rr = [F(yi) / fi if fi != 0 else F(0) for yi, fi in zip(y, f)]
        u = [F(0)] * ncols
        for i, r in enumerate(rows):
            for c, v in r:
                u[c] += F(v) * rr[i]
        x = [xj * (uj / nj if nj != 0 else uj) for xj, uj, nj in zip(x, u, norm)]
        out["iterates"].append({"f": [float(v) for v in f], "r": [float(v) for v in rr],
                                "u": [float(v) for v in u], "x": [float(v) for v in x]})
    return out


def main() -> None:
    (HERE / "tiny3x4.csr4").write_bytes(csr4_bytes(ROWS, NCOLS, GRID))
    (HERE / "tiny3x4.sino").write_bytes(struct.pack(f"<{len(SINO)}i", *SINO))
    ref = exact_mlem(ROWS, NCOLS, SINO, 3)
    ref.update({"rows": len(ROWS), "cols": NCOLS, "nnz": sum(len(r) for r in ROWS),
                "grid": list(GRID), "sino": SINO,
                "elements": [[c, v] for r in ROWS for (c, v) in r],
                "rowidx": [sum(len(r) for r in ROWS[: i + 1]) for i in range(len(ROWS))]})
    (HERE / "tiny3x4.json").write_text(json.dumps(ref, indent=1) + "\n")
    print("wrote tiny3x4.csr4, tiny3x4.sino, tiny3x4.json")


if __name__ == "__main__":
    main()
