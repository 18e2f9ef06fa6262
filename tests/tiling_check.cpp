// Host-side checks of the box tiling (mlem_amd/csrc/panels.cpp) on small
// seeded matrices: structural invariants the GPU kernels rely on (bounds,
// partitions, chunk limits and alignment, launch order) and CPU emulations of
// the forward and backward products over the tiling, compared with the dense
// products in double precision. Built and run by tests/test_tiling.py.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "panels.hpp"

using namespace mlem;

#define CHECK(c)                                                         \
  do {                                                                   \
    if (!(c)) {                                                          \
      std::fprintf(stderr, "check failed (%s:%d): %s\n", __FILE__, __LINE__, #c); \
      std::exit(1);                                                      \
    }                                                                    \
  } while (0)

namespace {

struct Case {
  uint32_t nx, ny, nz;
  bool grid;
  uint32_t rows;
  int maxlen;
  unsigned seed;
};

// Rows are runs of neighbouring voxels (as a line of response crosses the
// grid), some of random voxels, some longer than a piece may be; columns
// ascending and distinct within a row; device columns a random permutation
// of the file columns.
void make_matrix(const Case& k, std::vector<uint32_t>* rowptr, std::vector<uint32_t>* csr2,
                 std::vector<uint32_t>* colmap) {
  const uint32_t cols = k.nx * k.ny * k.nz;
  std::mt19937 rng(k.seed);
  rowptr->assign(k.rows + 1, 0);
  csr2->clear();
  for (uint32_t r = 0; r < k.rows; ++r) {
    int len = (int)(rng() % (uint32_t)(k.maxlen + 1));
    if (r % 97 == 5) len = (int)std::min<uint32_t>(cols, 1500);
    std::vector<uint32_t> c;
    uint32_t x = rng() % k.nx, y = rng() % k.ny, z = rng() % k.nz;
    for (int i = 0; i < len; ++i) {
      c.push_back(x + k.nx * (y + k.ny * z));
      const uint32_t d = rng() % 3;
      if (d == 0) x = (x + 1) % k.nx;
      else if (d == 1) y = (y + 1) % k.ny;
      else z = (z + 1) % k.nz;
    }
    if (r % 13 == 0)
      for (auto& v : c) v = rng() % cols;
    std::sort(c.begin(), c.end());
    c.erase(std::unique(c.begin(), c.end()), c.end());
    for (uint32_t col : c) {
      const float v = (float)(rng() % 1000 + 1) / 1000.f;
      uint32_t bits;
      std::memcpy(&bits, &v, 4);
      csr2->push_back(col);
      csr2->push_back(bits);
    }
    (*rowptr)[r + 1] = (*rowptr)[r] + (uint32_t)c.size();
  }
  colmap->resize(cols);
  for (uint32_t i = 0; i < cols; ++i) (*colmap)[i] = i;
  std::shuffle(colmap->begin(), colmap->end(), rng);
  for (size_t e = 0; e < csr2->size(); e += 2) (*csr2)[e] = (*colmap)[(*csr2)[e]];
}

void run(const Case& k) {
  const uint32_t cols = k.nx * k.ny * k.nz;
  std::vector<uint32_t> rowptr, csr2, colmap;
  make_matrix(k, &rowptr, &csr2, &colmap);
  const uint64_t nnz = rowptr[k.rows];

  // ---- panel order: a permutation; boxes of at most kBoxVox voxels, each
  // one box of the grid (or one run of columns without a grid)
  std::vector<uint32_t> count_dev(cols, 0), count_file(cols);
  for (uint64_t e = 0; e < nnz; ++e) count_dev[csr2[2 * e]]++;
  for (uint32_t c = 0; c < cols; ++c) count_file[c] = count_dev[colmap[c]];
  const uint32_t g[3] = {k.nx, k.ny, k.nz};
  std::vector<uint32_t> pc_of_file, pstart;
  uint32_t boxgrid[3];
  panel_order(cols, k.grid ? g : nullptr, count_file.data(), &pc_of_file, &pstart, boxgrid);
  const uint32_t P = (uint32_t)pstart.size() - 1;
  CHECK(pstart.front() == 0 && pstart.back() == cols);
  CHECK((uint64_t)boxgrid[0] * boxgrid[1] * boxgrid[2] == P);
  std::vector<uint32_t> file_of_pc(cols, ~0u);
  for (uint32_t c = 0; c < cols; ++c) {
    CHECK(pc_of_file[c] < cols && file_of_pc[pc_of_file[c]] == ~0u);
    file_of_pc[pc_of_file[c]] = c;
  }
  for (uint32_t p = 0; p < P; ++p) {
    CHECK(pstart[p] < pstart[p + 1] && pstart[p + 1] - pstart[p] <= kBoxVox);
    // every voxel of the range lies in one box of the grid (or one run)
    for (uint32_t pc = pstart[p]; pc < pstart[p + 1]; ++pc) {
      const uint32_t c = file_of_pc[pc];
      uint32_t b;
      if (k.grid) {
        const uint32_t x = c % k.nx, y = (c / k.nx) % k.ny, z = c / (k.nx * k.ny);
        b = x / 16 + boxgrid[0] * (y / 16 + boxgrid[1] * (z / 4));
      } else {
        b = c / kBoxVox;
      }
      CHECK(b == p);
    }
  }
  std::vector<uint32_t> pc_of_dev(cols);
  for (uint32_t c = 0; c < cols; ++c) pc_of_dev[colmap[c]] = pc_of_file[c];

  HostPanels H;
  build_host_panels(k.rows, cols, nnz, rowptr.data(), csr2.data(), pc_of_dev, pstart, boxgrid, 4,
                    &H);
  CHECK(H.npanels == P);
  const uint64_t nseg = H.fseg.size() - 1;
  CHECK(H.pseg.size() == P + 1 && H.pseg[P] == nseg && H.rowslot[k.rows] == nseg);

  // ---- launch order: every box once, largest (most elements) first
  {
    std::vector<uint64_t> elems(P, 0);
    for (uint32_t p = 0; p < P; ++p)
      for (uint32_t pc = pstart[p]; pc < pstart[p + 1]; ++pc) elems[p] += count_file[file_of_pc[pc]];
    std::vector<char> seen(P, 0);
    uint64_t prev = ~0ull;
    for (uint32_t b : H.border) {
      CHECK(b < P && !seen[b]);
      seen[b] = 1;
      CHECK(elems[b] <= prev);
      prev = elems[b];
    }
  }

  // ---- operands and dense reference products (device rows, panel columns)
  const int W = 3;
  std::mt19937 rng(k.seed + 1);
  std::vector<double> X((size_t)cols * W), R((size_t)k.rows * W);
  for (auto& v : X) v = (double)(rng() % 1000) / 7.0;
  for (auto& v : R) v = (double)(rng() % 1000) / 11.0;
  std::vector<double> fref((size_t)k.rows * W, 0.0), uref((size_t)cols * W, 0.0);
  for (uint32_t r = 0; r < k.rows; ++r)
    for (uint32_t e = rowptr[r]; e < rowptr[r + 1]; ++e) {
      float a;
      std::memcpy(&a, &csr2[2ull * e + 1], 4);
      const uint32_t pc = pc_of_dev[csr2[2ull * e]];
      for (int f = 0; f < W; ++f) {
        fref[(size_t)r * W + f] += a * X[(size_t)pc * W + f];
        uref[(size_t)pc * W + f] += a * R[(size_t)r * W + f];
      }
    }

  // ---- forward: chunks partition each box's pieces, start aligned and fit
  // the staging limits; slots are a permutation; the emulated product
  // matches (the zero gaps between chunks add nothing)
  CHECK(H.pfchunk.size() == P + 1 && H.fval.size() == H.fcol.size());
  const uint64_t fend = H.fval.size() - kFChunkElems;  // the arrays end with a chunk of zeros
  std::vector<double> part(nseg * W, 0.0);
  std::vector<char> slot_seen(nseg, 0);
  for (uint32_t p = 0; p < P; ++p) {
    const uint32_t c0 = pstart[p], nv = pstart[p + 1] - c0;
    if (H.pseg[p] < H.pseg[p + 1]) CHECK(H.fchunk_piece[H.pfchunk[p]] == H.pseg[p]);
    for (uint32_t kk = H.pfchunk[p]; kk < H.pfchunk[p + 1]; ++kk) {
      const uint32_t s0 = H.fchunk_piece[kk], s1 = H.fchunk_piece[kk + 1];
      CHECK(s0 < s1 && s1 - s0 <= kFChunkPieces && s1 <= H.pseg[p + 1]);
      CHECK(H.fseg[s0] % kFAlign == 0 && H.fseg[s1] - H.fseg[s0] <= kFChunkElems);
    }
    for (uint32_t s = H.pseg[p]; s < H.pseg[p + 1]; ++s) {
      const uint32_t b = H.fseg[s], e = H.fseg[s + 1];
      CHECK(b < e && e <= fend && e - b <= kMaxSegment + kFAlign);
      CHECK(H.fslot[s] < nseg && !slot_seen[H.fslot[s]]);
      slot_seen[H.fslot[s]] = 1;
      for (uint32_t i = b; i < e; ++i) {
        CHECK(H.fcol[i] < nv);
        for (int f = 0; f < W; ++f)
          part[(size_t)H.fslot[s] * W + f] += H.fval[i] * X[(size_t)(c0 + H.fcol[i]) * W + f];
      }
    }
  }
  double ferr = 0.0, fmax = 0.0;
  for (uint32_t r = 0; r < k.rows; ++r)
    for (int f = 0; f < W; ++f) {
      double a = 0.0;
      for (uint32_t s = H.rowslot[r]; s < H.rowslot[r + 1]; ++s) a += part[(size_t)s * W + f];
      ferr = std::max(ferr, std::fabs(a - fref[(size_t)r * W + f]));
      fmax = std::max(fmax, std::fabs(fref[(size_t)r * W + f]));
    }
  CHECK(ferr <= 1e-9 * std::max(1.0, fmax));

  // ---- backward: chunks within limits and aligned, voxel offsets cover the
  // chunk's elements, pieces ascending (rows ascending) inside a voxel's list;
  // the emulated product matches
  const uint32_t nchunks = (uint32_t)H.chunk_piece.size() - 1;
  CHECK(H.boff.size() == (size_t)nchunks * (kBoxVox + 1) && H.bval.size() == H.bpc.size());
  std::vector<double> U((size_t)cols * W, 0.0);
  for (uint32_t p = 0; p < P; ++p) {
    const uint32_t c0 = pstart[p], nv = pstart[p + 1] - c0;
    for (uint32_t kk = H.pchunk[p]; kk < H.pchunk[p + 1]; ++kk) {
      const uint32_t np = H.chunk_piece[kk + 1] - H.chunk_piece[kk];
      const uint32_t ne = H.chunk_elem[kk + 1] - H.chunk_elem[kk];
      CHECK(np >= 1 && np <= kChunkPieces && ne <= kChunkElems);
      CHECK(H.chunk_elem[kk] % kBAlign == 0 && H.chunk_elem[kk] + kChunkElems <= H.bval.size());
      const uint16_t* off = &H.boff[(size_t)kk * (kBoxVox + 1)];
      CHECK(off[0] == 0 && off[kBoxVox] <= ne && ne - off[kBoxVox] < kBAlign);
      for (uint32_t v = 0; v < kBoxVox; ++v) {
        CHECK(off[v] <= off[v + 1]);
        if (v >= nv) CHECK(off[v] == off[v + 1]);
        uint32_t prev = 0;
        for (uint32_t i = off[v]; i < off[v + 1]; ++i) {
          const uint64_t e = H.chunk_elem[kk] + i;
          CHECK(H.bpc[e] < np && H.bpc[e] >= prev);
          prev = H.bpc[e];
          const uint32_t row = H.brow[H.chunk_piece[kk] + H.bpc[e]];
          CHECK(row < k.rows);
          for (int f = 0; f < W; ++f)
            U[(size_t)(c0 + v) * W + f] += H.bval[e] * R[(size_t)row * W + f];
        }
      }
    }
  }
  double uerr = 0.0, umax = 0.0;
  for (size_t i = 0; i < U.size(); ++i) {
    uerr = std::max(uerr, std::fabs(U[i] - uref[i]));
    umax = std::max(umax, std::fabs(uref[i]));
  }
  CHECK(uerr <= 1e-9 * std::max(1.0, umax));
  std::printf("ok %ux%ux%u grid=%d rows=%u nnz=%llu boxes=%u pieces=%llu fwd=%.1e bwd=%.1e\n",
              k.nx, k.ny, k.nz, (int)k.grid, k.rows, (unsigned long long)nnz, P,
              (unsigned long long)nseg, ferr, uerr);
}

}  // namespace

int main() {
  const Case cases[] = {
      {40, 36, 10, true, 3000, 60, 1},    // partial boxes on every side
      {16, 16, 4, true, 5000, 200, 2},    // exactly one box
      {33, 17, 9, false, 2000, 300, 3},   // no grid: runs of 1024 columns
      {64, 64, 8, true, 20000, 120, 4},   // many full boxes
      {8, 8, 8, true, 4096, 30, 5},       // boxes smaller than 1024 voxels
      {20, 20, 4, true, 0, 10, 6},        // no rows at all
  };
  for (const Case& c : cases) run(c);
  return 0;
}
