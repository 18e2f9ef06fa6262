// panels.cpp — host construction of the two-dimensional tiling (panels.hpp).
#include "panels.hpp"

#include <algorithm>
#include <cstring>
#include <numeric>
#include <stdexcept>
#include <thread>

namespace mlem {
namespace {

constexpr uint32_t kBoxX = 16, kBoxY = 16, kBoxZ = 4;
static_assert(kBoxX * kBoxY * kBoxZ == kBoxVox, "box shape");
static_assert(kMaxSegment <= kChunkElems && kMaxSegment <= kFChunkElems,
              "every piece fits a chunk");
static_assert(kChunkPieces <= 65536 && kChunkElems <= 65535, "16-bit chunk offsets");

int pick_threads(int threads) {
  if (threads > 0) return threads;
  const unsigned hc = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(16u, hc));
}

template <class F>
void parallel_for(int nthreads, F&& f) {
  if (nthreads <= 1) {
    f(0);
    return;
  }
  std::vector<std::thread> pool;
  pool.reserve((size_t)nthreads);
  for (int t = 0; t < nthreads; ++t) pool.emplace_back([&f, t] { f(t); });
  for (auto& th : pool) th.join();
}

}  // namespace

void panel_order(uint32_t cols, const uint32_t grid[3], const uint32_t* count,
                 std::vector<uint32_t>* pc_of_file, std::vector<uint32_t>* pstart,
                 uint32_t boxgrid[3]) {
  pc_of_file->assign(cols, 0);
  pstart->clear();
  const uint64_t nx = grid ? grid[0] : 0, ny = grid ? grid[1] : 0, nz = grid ? grid[2] : 0;
  // box of every file column, boxes numbered x fastest; file columns of every box
  std::vector<uint32_t> box_of(cols);
  uint32_t nboxes = 0;
  if (!nx || !ny || !nz || nx * ny * nz != cols) {
    for (uint32_t c = 0; c < cols; ++c) box_of[c] = c / kBoxVox;
    nboxes = (uint32_t)((cols + kBoxVox - 1) / kBoxVox);
    boxgrid[0] = nboxes;
    boxgrid[1] = boxgrid[2] = 1;
  } else {
    const uint64_t nbx = (nx + kBoxX - 1) / kBoxX, nby = (ny + kBoxY - 1) / kBoxY;
    const uint64_t nbz = (nz + kBoxZ - 1) / kBoxZ;
    for (uint64_t c = 0; c < cols; ++c) {
      const uint64_t x = c % nx, y = (c / nx) % ny, z = c / (nx * ny);
      box_of[c] = (uint32_t)(x / kBoxX + nbx * (y / kBoxY + nby * (z / kBoxZ)));
    }
    nboxes = (uint32_t)(nbx * nby * nbz);
    boxgrid[0] = (uint32_t)nbx;
    boxgrid[1] = (uint32_t)nby;
    boxgrid[2] = (uint32_t)nbz;
  }
  std::vector<uint32_t> first(nboxes + 1, 0), members(cols);
  for (uint32_t c = 0; c < cols; ++c) first[box_of[c] + 1]++;
  for (uint32_t p = 0; p < nboxes; ++p) first[p + 1] += first[p];
  {
    std::vector<uint32_t> cur(first.begin(), first.end() - 1);
    for (uint32_t c = 0; c < cols; ++c) members[cur[box_of[c]]++] = c;  // file order in a box
  }
  constexpr uint32_t kGroup = 8, kGroups = kBoxVox / kGroup;
  std::vector<uint32_t> pos;
  for (uint32_t p = 0; p < nboxes; ++p) {
    pstart->push_back(first[p]);
    const uint32_t nv = first[p + 1] - first[p];
    uint32_t* m = &members[first[p]];
    if (count)  // most elements first (ties: file order)
      std::stable_sort(m, m + nv, [&](uint32_t a, uint32_t b) { return count[a] > count[b]; });
    // positions in deal order: round j gives every group g its position 8g + j,
    // groups in snake order (ascending in even rounds, descending in odd ones)
    const uint32_t groups = std::min(kGroups, (nv + kGroup - 1) / kGroup);
    pos.clear();
    for (uint32_t j = 0; j < kGroup; ++j)
      for (uint32_t i = 0; i < groups; ++i) {
        const uint32_t g = (j % 2 == 0) ? i : groups - 1 - i;
        if (g * kGroup + j < nv) pos.push_back(g * kGroup + j);
      }
    for (uint32_t r = 0; r < nv; ++r) (*pc_of_file)[m[r]] = first[p] + pos[r];
  }
  pstart->push_back(cols);
}

void build_host_panels(uint32_t rows, uint32_t cols, uint64_t nnz, const uint32_t* rowptr,
                       const uint32_t* csr2, const std::vector<uint32_t>& pc_of_dev,
                       const std::vector<uint32_t>& pstart, const uint32_t boxgrid[3],
                       int threads, HostPanels* out) {
  HostPanels& H = *out;
  H = HostPanels();
  H.rows = rows;
  H.cols = cols;
  H.nnz = nnz;
  H.pc_of_dev = pc_of_dev;
  H.pstart = pstart;
  const uint32_t P = (uint32_t)pstart.size() - 1;
  H.npanels = P;
  if (pc_of_dev.size() != cols) throw std::invalid_argument("panel order size");
  if ((uint64_t)boxgrid[0] * boxgrid[1] * boxgrid[2] != P) throw std::invalid_argument("box grid");
  if (nnz > 0xFFFFFFFFull) throw std::invalid_argument("too many elements for 32-bit offsets");
  for (uint32_t p = 0; p < P; ++p)
    if (pstart[p + 1] - pstart[p] > kBoxVox) throw std::invalid_argument("box too large");

  // box and in-box offset of every device column
  std::vector<uint32_t> pid(cols);
  std::vector<uint16_t> lid(cols);
  {
    std::vector<uint32_t> panel_of_pc(cols);
    for (uint32_t p = 0; p < P; ++p)
      for (uint32_t c = pstart[p]; c < pstart[p + 1]; ++c) panel_of_pc[c] = p;
    for (uint32_t d = 0; d < cols; ++d) {
      const uint32_t pc = pc_of_dev[d];
      if (pc >= cols) throw std::invalid_argument("panel order out of range");
      pid[d] = panel_of_pc[pc];
      lid[d] = (uint16_t)(pc - pstart[pid[d]]);
    }
  }

  const int T = std::max(1, std::min<int>(pick_threads(threads), (int)std::max<uint32_t>(rows, 1)));
  // row ranges of the threads, balanced by elements
  std::vector<uint32_t> rr((size_t)T + 1, rows);
  rr[0] = 0;
  for (int t = 1; t < T; ++t) {
    const uint64_t target = nnz * (uint64_t)t / (uint64_t)T;
    rr[(size_t)t] = (uint32_t)(std::lower_bound(rowptr, rowptr + rows, (uint32_t)target) - rowptr);
  }
  for (int t = 1; t <= T; ++t) rr[(size_t)t] = std::max(rr[(size_t)t], rr[(size_t)t - 1]);

  // ---- pass 1: pieces per row, pieces and elements per (thread, box)
  std::vector<uint32_t> nsegrow(rows);
  std::vector<std::vector<uint64_t>> cseg((size_t)T, std::vector<uint64_t>(P, 0));
  std::vector<std::vector<uint64_t>> cel((size_t)T, std::vector<uint64_t>(P, 0));
  const uint32_t SP = kMaxSegment;
  parallel_for(T, [&](int t) {
    std::vector<uint32_t> stamp(P, 0xFFFFFFFFu), cnt(P, 0);
    std::vector<uint32_t> plist;
    for (uint32_t r = rr[(size_t)t]; r < rr[(size_t)t + 1]; ++r) {
      plist.clear();
      for (uint32_t k = rowptr[r]; k < rowptr[r + 1]; ++k) {
        const uint32_t p = pid[csr2[2ull * k]];
        cel[(size_t)t][p]++;
        if (stamp[p] != r) {
          stamp[p] = r;
          cnt[p] = 0;
          plist.push_back(p);
        }
        cnt[p]++;
      }
      uint32_t nd = 0;
      for (uint32_t p : plist) {  // a (row, box) piece longer than SP is split
        const uint32_t pieces = (cnt[p] + SP - 1) / SP;
        nd += pieces;
        cseg[(size_t)t][p] += pieces;
      }
      nsegrow[r] = nd;
    }
  });
  H.rowslot.assign((size_t)rows + 1, 0);
  for (uint32_t r = 0; r < rows; ++r) H.rowslot[r + 1] = H.rowslot[r] + nsegrow[r];
  const uint64_t nseg = H.rowslot[rows];
  if (nseg > 0xFFFFFFF0ull) throw std::invalid_argument("too many pieces");
  // box-major cursors: box p, then the threads in row order
  std::vector<std::vector<uint64_t>> scur((size_t)T, std::vector<uint64_t>(P)),
      ecur((size_t)T, std::vector<uint64_t>(P));
  H.pseg.assign((size_t)P + 1, 0);
  std::vector<uint64_t> pel((size_t)P + 1, 0);
  {
    uint64_t s = 0, e = 0;
    for (uint32_t p = 0; p < P; ++p) {
      H.pseg[p] = (uint32_t)s;
      pel[p] = e;
      for (int t = 0; t < T; ++t) {
        scur[(size_t)t][p] = s;
        ecur[(size_t)t][p] = e;
        s += cseg[(size_t)t][p];
        e += cel[(size_t)t][p];
      }
    }
    H.pseg[P] = (uint32_t)s;
    pel[P] = e;
  }
  // ---- pass 2: pieces in row order inside each box (the backward order):
  // first element, slot and row of every piece, elements piece-major
  std::vector<uint32_t> rseg(nseg + 1), rslot(nseg);
  std::vector<float> rval(nnz);
  std::vector<uint16_t> rcol(nnz);
  H.brow.assign(nseg, 0);
  parallel_for(T, [&](int t) {
    std::vector<uint32_t> stamp(P, 0xFFFFFFFFu), cnt(P, 0);
    std::vector<uint64_t> pos(P, 0);
    std::vector<uint32_t> plist;
    for (uint32_t r = rr[(size_t)t]; r < rr[(size_t)t + 1]; ++r) {
      plist.clear();
      for (uint32_t k = rowptr[r]; k < rowptr[r + 1]; ++k) {
        const uint32_t p = pid[csr2[2ull * k]];
        if (stamp[p] != r) {
          stamp[p] = r;
          cnt[p] = 0;
          plist.push_back(p);
        }
        cnt[p]++;
      }
      std::sort(plist.begin(), plist.end());
      uint32_t rank = 0;
      for (size_t j = 0; j < plist.size(); ++j) {
        const uint32_t p = plist[j];
        pos[p] = ecur[(size_t)t][p];
        for (uint32_t done = 0; done < cnt[p]; done += SP) {
          const uint64_t s = scur[(size_t)t][p]++;
          rseg[s] = (uint32_t)(ecur[(size_t)t][p] + done);
          rslot[s] = H.rowslot[r] + rank++;
          H.brow[s] = r;
        }
        ecur[(size_t)t][p] += cnt[p];
      }
      for (uint32_t k = rowptr[r]; k < rowptr[r + 1]; ++k) {
        const uint32_t d = csr2[2ull * k];
        const uint64_t q = pos[pid[d]]++;
        const uint32_t bits = csr2[2ull * k + 1];
        float v;
        static_assert(sizeof(v) == sizeof(bits), "fp32");
        std::memcpy(&v, &bits, 4);
        rval[q] = v;
        rcol[q] = lid[d];
      }
    }
  });
  rseg[nseg] = (uint32_t)nnz;

  const int TP = std::max(1, std::min<int>(T, (int)std::max<uint32_t>(P, 1)));
  // ---- forward: inside every box the pieces longest first (stable)
  // Chunks of both directions start at multiples of kFAlign / kBAlign
  // elements, so that their elements load in 16-byte pieces; the gaps hold
  // zero values (column / piece 0), and the arrays end with a chunk's worth
  // of zeros so that whole-chunk loads stay inside them.
  auto align_up = [](uint64_t x, uint64_t a) { return (x + a - 1) / a * a; };
  // ---- forward: inside every box the pieces longest first (stable), cut
  // greedily into chunks of at most kFChunkPieces pieces / kFChunkElems
  // elements
  std::vector<std::vector<uint32_t>> forder(P), fcut(P);
  std::vector<uint64_t> fsize(P, 0);
  parallel_for(TP, [&](int t) {
    for (uint32_t p = (uint32_t)t; p < P; p += (uint32_t)TP) {
      const uint32_t s0 = H.pseg[p], s1 = H.pseg[p + 1];
      std::vector<uint32_t>& idx = forder[p];
      idx.resize(s1 - s0);
      std::iota(idx.begin(), idx.end(), s0);
      std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) {
        return rseg[a + 1] - rseg[a] > rseg[b + 1] - rseg[b];
      });
      uint64_t size = 0;
      for (uint32_t i = 0; i < s1 - s0;) {
        fcut[p].push_back(i);
        uint32_t n = 0, j = i;
        while (j < s1 - s0 && j - i < kFChunkPieces &&
               n + (rseg[idx[j] + 1] - rseg[idx[j]]) <= kFChunkElems)
          n += rseg[idx[j] + 1] - rseg[idx[j]], ++j;
        size += align_up(n, kFAlign);
        i = j;
      }
      fsize[p] = size;
    }
  });
  std::vector<uint64_t> fbase((size_t)P + 1, 0);
  H.pfchunk.assign((size_t)P + 1, 0);
  for (uint32_t p = 0; p < P; ++p) {
    fbase[p + 1] = fbase[p] + fsize[p];
    H.pfchunk[p + 1] = H.pfchunk[p] + (uint32_t)fcut[p].size();
  }
  const uint64_t nfel = fbase[P];
  if (nfel + kFChunkElems > 0xFFFFFFFFull) throw std::invalid_argument("too many elements");
  H.fseg.assign(nseg + 1, 0);
  H.fslot.assign(nseg, 0);
  H.fval.assign(nfel + kFChunkElems, 0.f);
  H.fcol.assign(nfel + kFChunkElems, 0);
  H.fchunk_piece.assign((size_t)H.pfchunk[P] + 1, 0);
  parallel_for(TP, [&](int t) {
    for (uint32_t p = (uint32_t)t; p < P; p += (uint32_t)TP) {
      const uint32_t s0 = H.pseg[p], np = H.pseg[p + 1] - s0;
      const std::vector<uint32_t>& idx = forder[p];
      uint64_t e = fbase[p];
      uint32_t k = H.pfchunk[p];
      for (uint32_t i = 0, c = 0; i < np; ++i) {
        if (c < fcut[p].size() && fcut[p][c] == i) {  // a chunk starts: align
          e = fbase[p] + align_up(e - fbase[p], kFAlign);
          H.fchunk_piece[k++] = s0 + i;
          ++c;
        }
        const uint32_t s = idx[i], n = rseg[s + 1] - rseg[s];
        H.fseg[s0 + i] = (uint32_t)e;
        H.fslot[s0 + i] = rslot[s];
        std::memcpy(&H.fval[e], &rval[rseg[s]], 4ull * n);
        std::memcpy(&H.fcol[e], &rcol[rseg[s]], 2ull * n);
        e += n;
      }
    }
  });
  // a piece ends where the next one starts (the last of a chunk includes the
  // zero gap before the next chunk)
  H.fseg[nseg] = (uint32_t)nfel;
  H.fchunk_piece[H.pfchunk[P]] = (uint32_t)nseg;

  // ---- backward: chunks of the row-ordered pieces of every box
  std::vector<uint32_t> nck(P, 0);
  std::vector<uint64_t> bsize(P, 0);
  auto chunk_end = [&](uint32_t s, uint32_t s1) {  // greedy: pieces and elements bounded
    uint32_t e = s + 1;
    while (e < s1 && e - s < kChunkPieces && rseg[e + 1] - rseg[s] <= kChunkElems) ++e;
    return e;
  };
  parallel_for(TP, [&](int t) {
    for (uint32_t p = (uint32_t)t; p < P; p += (uint32_t)TP)
      for (uint32_t s = H.pseg[p], s1; s < H.pseg[p + 1]; s = s1) {
        s1 = chunk_end(s, H.pseg[p + 1]);
        nck[p]++;
        bsize[p] += align_up(rseg[s1] - rseg[s], kBAlign);
      }
  });
  H.pchunk.assign((size_t)P + 1, 0);
  std::vector<uint64_t> bbase((size_t)P + 1, 0);
  for (uint32_t p = 0; p < P; ++p) {
    H.pchunk[p + 1] = H.pchunk[p] + nck[p];
    bbase[p + 1] = bbase[p] + bsize[p];
  }
  const uint32_t nchunks = H.pchunk[P];
  const uint64_t nbel = bbase[P];
  if (nbel + kChunkElems > 0xFFFFFFFFull) throw std::invalid_argument("too many elements");
  H.chunk_piece.assign((size_t)nchunks + 1, 0);
  H.chunk_elem.assign((size_t)nchunks + 1, 0);
  H.boff.assign((size_t)nchunks * (kBoxVox + 1), 0);
  H.bpc.assign(nbel + kChunkElems, 0);
  H.bval.assign(nbel + kChunkElems, 0.f);
  parallel_for(TP, [&](int t) {
    std::vector<uint32_t> cur(kBoxVox + 1);
    for (uint32_t p = (uint32_t)t; p < P; p += (uint32_t)TP) {
      uint32_t k = H.pchunk[p];
      uint64_t at0 = bbase[p];
      for (uint32_t s = H.pseg[p]; s < H.pseg[p + 1]; ++k) {
        const uint32_t s1 = chunk_end(s, H.pseg[p + 1]);
        const uint32_t e0 = rseg[s], e1 = rseg[s1];
        H.chunk_piece[k] = s;
        H.chunk_elem[k] = (uint32_t)at0;
        // elements per voxel, then their offsets; elements placed voxel by
        // voxel, pieces (rows) ascending inside a voxel
        std::fill(cur.begin(), cur.end(), 0u);
        for (uint32_t e = e0; e < e1; ++e) cur[rcol[e] + 1]++;
        for (uint32_t v = 0; v < kBoxVox; ++v) cur[v + 1] += cur[v];
        uint16_t* off = &H.boff[(size_t)k * (kBoxVox + 1)];
        for (uint32_t v = 0; v <= kBoxVox; ++v) off[v] = (uint16_t)cur[v];
        for (uint32_t q = s; q < s1; ++q)
          for (uint32_t e = rseg[q]; e < rseg[q + 1]; ++e) {
            const uint64_t at = at0 + cur[rcol[e]]++;
            H.bpc[at] = (uint16_t)(q - s);
            H.bval[at] = rval[e];
          }
        at0 += align_up(e1 - e0, kBAlign);
        s = s1;
      }
    }
  });
  H.chunk_piece[nchunks] = (uint32_t)nseg;
  H.chunk_elem[nchunks] = (uint32_t)nbel;

  // launch order: boxes by decreasing element count (largest first)
  H.border.resize(P);
  std::iota(H.border.begin(), H.border.end(), 0u);
  std::stable_sort(H.border.begin(), H.border.end(), [&](uint32_t a, uint32_t b) {
    return pel[a + 1] - pel[a] > pel[b + 1] - pel[b];
  });
}

}  // namespace mlem
