// panels.cpp — host construction of the two-dimensional tiling (panels.hpp).
#include "panels.hpp"

#include <algorithm>
#include <cstring>
#include <numeric>
#include <stdexcept>
#include <thread>

namespace mlem {
namespace {

constexpr uint32_t kPanelX = 32, kPanelY = 32, kPanelZ = 16;  // 16384 voxels
constexpr uint32_t kFlatPanel = 16384;                          // without a grid

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

void panel_order(uint32_t cols, const uint32_t grid[3], std::vector<uint32_t>* pc_of_file,
                 std::vector<uint32_t>* pstart) {
  pc_of_file->assign(cols, 0);
  pstart->clear();
  const uint64_t nx = grid ? grid[0] : 0, ny = grid ? grid[1] : 0, nz = grid ? grid[2] : 0;
  if (!nx || !ny || !nz || nx * ny * nz != cols) {
    for (uint32_t c = 0; c < cols; ++c) (*pc_of_file)[c] = c;
    for (uint64_t p = 0; p < cols; p += kFlatPanel) pstart->push_back((uint32_t)p);
    pstart->push_back(cols);
    return;
  }
  const uint64_t npx = (nx + kPanelX - 1) / kPanelX, npy = (ny + kPanelY - 1) / kPanelY;
  const uint64_t npz = (nz + kPanelZ - 1) / kPanelZ;
  // sort key: panel, then 4x4x2 brick inside the panel (x fastest), then the
  // voxel inside the brick; ranks of the keys are the panel-order columns
  std::vector<uint64_t> key(cols);
  for (uint64_t c = 0; c < cols; ++c) {
    const uint64_t x = c % nx, y = (c / nx) % ny, z = c / (nx * ny);
    const uint64_t p = x / kPanelX + npx * (y / kPanelY + npy * (z / kPanelZ));
    const uint64_t lx = x % kPanelX, ly = y % kPanelY, lz = z % kPanelZ;
    const uint64_t b = lx / 4 + (kPanelX / 4) * (ly / 4 + (kPanelY / 4) * (lz / 2));
    const uint64_t in = lx % 4 + 4 * (ly % 4 + 4 * (lz % 2));
    key[c] = (p << 32) | (b * 32 + in);
  }
  std::vector<uint32_t> order(cols);
  std::iota(order.begin(), order.end(), 0u);
  std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return key[a] < key[b]; });
  uint64_t prev = ~0ull;
  for (uint32_t r = 0; r < cols; ++r) {
    const uint32_t c = order[r];
    (*pc_of_file)[c] = r;
    const uint64_t p = key[c] >> 32;
    if (p != prev) {
      pstart->push_back(r);
      prev = p;
    }
  }
  pstart->push_back(cols);
  (void)npz;
}

void build_host_panels(uint32_t rows, uint32_t cols, uint64_t nnz, const uint32_t* rowptr,
                       const uint32_t* csr2, const std::vector<uint32_t>& pc_of_dev,
                       const std::vector<uint32_t>& pstart, uint32_t chunk_rows, int threads,
                       HostPanels* out) {
  HostPanels& H = *out;
  H = HostPanels();
  H.rows = rows;
  H.cols = cols;
  H.nnz = nnz;
  H.chunk_rows = chunk_rows;
  H.pc_of_dev = pc_of_dev;
  H.pstart = pstart;
  const uint32_t P = (uint32_t)pstart.size() - 1;
  H.npanels = P;
  if (pc_of_dev.size() != cols) throw std::invalid_argument("panel order size");
  if (chunk_rows == 0) throw std::invalid_argument("chunk size");
  if (nnz > 0xFFFFFFFFull) throw std::invalid_argument("too many elements for 32-bit offsets");
  for (uint32_t p = 0; p < P; ++p)
    if (pstart[p + 1] - pstart[p] > 65536u) throw std::invalid_argument("panel too large");

  // panel and in-panel offset of every device column
  std::vector<uint32_t> pid(cols);
  std::vector<uint16_t> lid(cols);
  {
    std::vector<uint32_t> panel_of_pc(cols);
    for (uint32_t p = 0; p < P; ++p)
      for (uint32_t c = pstart[p]; c < pstart[p + 1]; ++c) panel_of_pc[c] = p;
    for (uint32_t d = 0; d < cols; ++d) {
      const uint32_t pc = pc_of_dev[d];
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

  // ---- forward, pass 1: segments per row, segments and elements per (thread, panel)
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
      for (uint32_t p : plist) {  // a (row, panel) piece longer than SP is split
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
  if (nseg > 0xFFFFFFF0ull) throw std::invalid_argument("too many segments");
  // panel-major cursors: panel p, then the threads in row order
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
  // ---- forward, pass 2: segment offsets and slots, elements
  H.fseg.assign(nseg + 1, 0);
  H.fslot.assign(nseg, 0);
  H.fval.assign(nnz, 0.f);
  H.fcol.assign(nnz, 0);
  std::vector<uint32_t> segrow(nseg);
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
          H.fseg[s] = (uint32_t)(ecur[(size_t)t][p] + done);
          H.fslot[s] = H.rowslot[r] + rank++;
          segrow[s] = r;
        }
        ecur[(size_t)t][p] += cnt[p];
      }
      for (uint32_t k = rowptr[r]; k < rowptr[r + 1]; ++k) {
        const uint32_t d = csr2[2ull * k];
        const uint64_t q = pos[pid[d]]++;
        uint32_t bits = csr2[2ull * k + 1];
        float v;
        static_assert(sizeof(v) == sizeof(bits), "fp32");
        std::memcpy(&v, &bits, 4);
        H.fval[q] = v;
        H.fcol[q] = lid[d];
      }
    }
  });
  H.fseg[nseg] = (uint32_t)nnz;

  // ---- backward: groups (panel, chunk, voxel)
  const uint32_t C = chunk_rows;
  H.pchunk.assign((size_t)P + 1, 0);
  for (uint32_t p = 0; p < P; ++p) {
    const uint32_t ns = H.pseg[p + 1] - H.pseg[p];
    H.pchunk[p + 1] = H.pchunk[p] + (ns + C - 1) / C;
  }
  const uint32_t nchunks = H.pchunk[P];
  // phase A (threads over panels): entries per group; non-empty chunks per voxel
  std::vector<std::vector<uint32_t>> gcount(P);
  std::vector<uint32_t> nchunk_pc(cols, 0);
  std::vector<uint64_t> ngroups_p(P, 0);
  const int TP = std::max(1, std::min<int>(T, (int)std::max<uint32_t>(P, 1)));
  parallel_for(TP, [&](int t) {
    for (uint32_t p = (uint32_t)t; p < P; p += (uint32_t)TP) {
      const uint32_t pv = pstart[p + 1] - pstart[p];
      const uint32_t nk = H.pchunk[p + 1] - H.pchunk[p];
      std::vector<uint32_t>& g = gcount[p];
      g.assign((size_t)nk * pv, 0);
      for (uint32_t s = H.pseg[p]; s < H.pseg[p + 1]; ++s) {
        const uint32_t k = (s - H.pseg[p]) / C;
        for (uint32_t e = H.fseg[s]; e < H.fseg[s + 1]; ++e) g[(size_t)k * pv + H.fcol[e]]++;
      }
      uint64_t ng = 0;
      for (uint32_t v = 0; v < pv; ++v) {
        uint32_t n = 0;  // groups of the voxel: non-empty chunks, long ones split
        for (uint32_t k = 0; k < nk; ++k) n += (g[(size_t)k * pv + v] + SP - 1) / SP;
        nchunk_pc[pstart[p] + v] = n;
        ng += n;
      }
      ngroups_p[p] = ng;
    }
  });
  H.colslot.assign((size_t)cols + 1, 0);
  for (uint32_t c = 0; c < cols; ++c) H.colslot[c + 1] = H.colslot[c] + nchunk_pc[c];
  std::vector<uint64_t> gbase((size_t)P + 1, 0);
  for (uint32_t p = 0; p < P; ++p) gbase[p + 1] = gbase[p] + ngroups_p[p];
  const uint64_t ngroups = gbase[P];
  if (ngroups > 0xFFFFFFF0ull) throw std::invalid_argument("too many groups");
  H.bseg.assign(ngroups + 1, 0);
  H.bslot.assign(ngroups, 0);
  H.brow.assign(nnz, 0);
  H.bval.assign(nnz, 0.f);
  H.cgroup.assign((size_t)nchunks + 1, 0);
  // phase B (threads over panels): group offsets and slots in (chunk, voxel)
  // order, then the entries of every segment in row order
  parallel_for(TP, [&](int t) {
    std::vector<uint64_t> gpos;
    for (uint32_t p = (uint32_t)t; p < P; p += (uint32_t)TP) {
      const uint32_t pv = pstart[p + 1] - pstart[p];
      const uint32_t nk = H.pchunk[p + 1] - H.pchunk[p];
      const std::vector<uint32_t>& g = gcount[p];
      gpos.assign((size_t)nk * pv, 0);
      std::vector<uint32_t> rank(pv, 0);
      uint64_t gi = gbase[p], e = pel[p];
      for (uint32_t k = 0; k < nk; ++k) {
        H.cgroup[H.pchunk[p] + k] = (uint32_t)gi;
        for (uint32_t v = 0; v < pv; ++v) {
          const uint32_t n = g[(size_t)k * pv + v];
          if (!n) continue;
          gpos[(size_t)k * pv + v] = e;
          for (uint32_t done = 0; done < n; done += SP) {
            H.bseg[gi] = (uint32_t)(e + done);
            H.bslot[gi] = H.colslot[pstart[p] + v] + rank[v]++;
            ++gi;
          }
          e += n;
        }
      }
      for (uint32_t s = H.pseg[p]; s < H.pseg[p + 1]; ++s) {
        const uint32_t k = (s - H.pseg[p]) / C;
        const uint32_t r = segrow[s];
        for (uint32_t q = H.fseg[s]; q < H.fseg[s + 1]; ++q) {
          const uint64_t at = gpos[(size_t)k * pv + H.fcol[q]]++;
          H.brow[at] = r;
          H.bval[at] = H.fval[q];
        }
      }
    }
  });
  H.cgroup[nchunks] = (uint32_t)ngroups;
  H.bseg[ngroups] = (uint32_t)nnz;
}

}  // namespace mlem
