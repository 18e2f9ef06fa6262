// symmetry.cpp — expansion of symmetry-reduced csr4 matrices (see
// symmetry.hpp for the mapping and its reference).
#include "symmetry.hpp"

#include <algorithm>
#include <string>

#include "mlem.h"

namespace mlem {

namespace {

[[noreturn]] void fail(const std::string& msg) { throw Csr4Error{MLEM_EFORMAT, msg}; }

}  // namespace

SymmetryMaps::SymmetryMaps(const Csr4Scanner& sc) {
  nx = sc.grid[0];
  ny = sc.grid[1];
  nz = sc.grid[2];
  const uint32_t B = sc.blocksize(), NB = sc.nBlocks, NL = (uint32_t)sc.layers.size();
  const uint32_t NR = sc.nRings();
  const uint64_t per_layer = (uint64_t)B * NB, per_ring = per_layer * NL;
  const uint64_t n = per_ring * NR;
  if (n == 0 || n > 65535) fail("symmetry expansion: unsupported detector count " + std::to_string(n));
  crystals = (uint32_t)n;
  for (auto& l : lut) l.assign(crystals, 0);
  for (uint32_t i = 0; i < crystals; ++i) lut[0][i] = i;
  // Mirror inside each layer ring: detector v of a layer goes to
  // (L - v + off) mod L, with off = B - 1 for x and B*(NB/2 + 1) - 1 for y
  // (symhandler.cpp:64-86); z reverses the ring order (symhandler.cpp:88-94).
  const uint64_t offx = B - 1, offy = (uint64_t)B * (NB / 2 + 1) - 1;
  for (uint32_t r = 0; r < NR; ++r)
    for (uint32_t l = 0; l < NL; ++l)
      for (uint64_t v = 0; v < per_layer; ++v) {
        const uint64_t base = l * per_layer + r * per_ring;
        lut[1][v + base] = (uint32_t)((per_layer - v + offx) % per_layer + base);
        lut[2][v + base] = (uint32_t)((per_layer - v + offy) % per_layer + base);
      }
  for (uint32_t r = 0; r < NR; ++r)
    for (uint64_t v = 0; v < per_ring; ++v)
      lut[4][v + r * per_ring] = (uint32_t)(v + (uint64_t)(NR - 1 - r) * per_ring);
  // Combined symmetries: x first, then y, then z (symhandler.cpp:96-114).
  for (uint32_t i = 0; i < crystals; ++i) {
    lut[3][i] = lut[2][lut[1][i]];
    lut[5][i] = lut[4][lut[1][i]];
    lut[6][i] = lut[4][lut[2][i]];
    lut[7][i] = lut[4][lut[2][lut[1][i]]];
  }
}

uint32_t SymmetryMaps::voxel(uint32_t v, int s) const {
  const uint32_t area = nx * ny;
  uint32_t z = v / area, rem = v % area, y = rem / nx, x = rem % nx;
  if (s & 1) x = nx - 1 - x;
  if (s & 2) y = ny - 1 - y;
  if (s & 4) z = nz - 1 - z;
  return x + nx * (y + ny * z);
}

uint32_t SymmetryMaps::row(uint32_t r, int s) const {
  const uint32_t d1 = lut[s][r / crystals], d2 = lut[s][r % crystals];
  return d1 < d2 ? d1 * crystals + d2 : d2 * crystals + d1;
}

std::vector<int> generated_symmetries(unsigned flags) {
  std::vector<int> out;
  for (int x = 0; x <= (int)(flags & 1u); ++x)
    for (int y = 0; y <= (int)((flags >> 1) & 1u); ++y)
      for (int z = 0; z <= (int)((flags >> 2) & 1u); ++z)
        if (x || y || z) out.push_back(x + 2 * y + 4 * z);
  return out;
}

void expand_symmetry(const Csr4Scanner& sc, unsigned flags, uint32_t rows, uint32_t cols,
                     const std::vector<uint64_t>& rowidx, const Csr4Element* el,
                     std::vector<uint64_t>& out_rowidx, std::vector<Csr4Element>& out_el) {
  const std::vector<int> syms = generated_symmetries(flags & 7u);
  // With an odd number of blocks the xz-plane mirror is not a symmetry of the
  // scanner (symhandler.cpp:42-46).
  if ((flags & 2u) && (sc.nBlocks % 2u))
    fail("the y symmetry flag is set but the scanner has an odd number of blocks (" +
         std::to_string(sc.nBlocks) + ")");
  const SymmetryMaps maps(sc);
  if ((uint64_t)maps.crystals * maps.crystals != rows)
    fail("symmetry expansion needs one row per detector pair: " + std::to_string(rows) +
         " rows, " + std::to_string(maps.crystals) + " crystals");
  if ((uint64_t)maps.nx * maps.ny * maps.nz != cols)
    fail("symmetry expansion needs one column per voxel of the " + std::to_string(maps.nx) +
         "x" + std::to_string(maps.ny) + "x" + std::to_string(maps.nz) + " grid");
  auto begin = [&](uint32_t r) { return r ? rowidx[r - 1] : 0; };
  for (uint64_t k = 0; k < (rows ? rowidx[rows - 1] : 0); ++k)
    if (el[k].column >= cols) fail("matrix element column out of range");
  // Count, place (stored elements first, then every symmetry in order), then
  // sort every row by column keeping the placement order among equals.
  std::vector<uint64_t> start((size_t)rows + 1, 0);
  for (uint32_t r = 0; r < rows; ++r) {
    const uint64_t len = rowidx[r] - begin(r);
    start[(size_t)r + 1] += len;
    for (int s : syms) start[(size_t)maps.row(r, s) + 1] += len;
  }
  for (uint32_t r = 0; r < rows; ++r) start[(size_t)r + 1] += start[r];
  std::vector<Csr4Element> tmp(start[rows]);
  std::vector<uint64_t> fill(start.begin(), start.end() - 1);
  for (uint32_t r = 0; r < rows; ++r)
    for (uint64_t k = begin(r); k < rowidx[r]; ++k) tmp[fill[r]++] = el[k];
  for (int s : syms)
    for (uint32_t r = 0; r < rows; ++r) {
      const uint32_t rr = maps.row(r, s);
      for (uint64_t k = begin(r); k < rowidx[r]; ++k)
        tmp[fill[rr]++] = Csr4Element{maps.voxel(el[k].column, s), el[k].value};
    }
  out_el.clear();
  out_el.reserve(tmp.size());
  out_rowidx.assign(rows, 0);
  for (uint32_t r = 0; r < rows; ++r) {
    auto b = tmp.begin() + (ptrdiff_t)start[r], e = tmp.begin() + (ptrdiff_t)start[(size_t)r + 1];
    std::stable_sort(b, e, [](const Csr4Element& a, const Csr4Element& c) {
      return a.column < c.column;
    });
    for (auto it = b; it != e; ++it)
      if (it == b || it->column != (it - 1)->column) out_el.push_back(*it);
    out_rowidx[r] = out_el.size();
  }
}

}  // namespace mlem
