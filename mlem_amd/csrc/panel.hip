// panel.hip — products of the system matrix with blocks of W interleaved
// right-hand sides through its two-dimensional tiling (panels.hpp), gfx950.
//
// Forward, f = A X: one 1024-lane workgroup per box loads the box's slice of
// X (at most 1024 voxels x W floats, 128 KB at W = 32) into LDS with
// coalesced 16-byte loads, then gives every piece of the box to 8 lanes: W/4
// quads of frames times 8/(W/4) shares of the elements (every 8/(W/4)-th
// element). Per element a lane reads its 16-byte quad of the voxel's run
// from LDS and accumulates with fused multiply-adds; a butterfly over the
// lanes adds the shares and the piece's sums go to its slot.
//
// Backward, u = A^T R: one 1024-lane workgroup per box; 8 lanes (quads times
// shares again) own 8 voxels of the box, whose sums stay in registers while
// the workgroup walks the box's pieces chunk by chunk: the chunk's runs of R
// (at most 256 rows) and its elements (at most 4096, voxel by voxel) are
// staged in LDS, then every lane group adds its voxels' elements. The
// complete sums are written once per voxel.
//
// Boxes are launched largest first. The slots of one row are contiguous; the
// epilogue kernels of the callers add them in box order, so every result is
// deterministic.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <numeric>
#include <vector>

#include "internal.hpp"
#include "mlem.h"
#include "panels.hpp"

namespace mlem {
namespace {

constexpr int kBoxThreads = 1024;

void check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) raise(MLEM_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

template <class T>
void upload(T** dst, const std::vector<T>& v, hipStream_t s) {
  MLEM_HIP(hipMalloc(dst, std::max<size_t>(16, sizeof(T) * v.size())));
  if (!v.empty())
    MLEM_HIP(hipMemcpyAsync(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, s));
}

template <class T>
void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

// Every piece (forward) and every voxel group (backward) gets 8 lanes
// whatever W is: W/4 lanes for the quads of frames (16 bytes each) times
// E = 32/W lanes that take every E-th element, their sums then added across
// the E lanes by a fixed butterfly.
template <int W>
struct Quad {
  static_assert(W % 4 == 0 && W <= 32, "right-hand sides come in quads, at most 32");
  static constexpr uint32_t L = W / 4;             // lanes per element (quads)
  static constexpr uint32_t E = 8 / L;             // lanes per piece and quad
  static constexpr uint32_t G = kBoxThreads / 8;   // pieces (voxel groups) per workgroup
};
// backward: voxels owned by one lane group (the groups cover a whole box)
constexpr uint32_t kGroupVox = kBoxVox / (kBoxThreads / 8);
static_assert(kGroupVox * (kBoxThreads / 8) == kBoxVox, "lane groups cover a box");

// Sum over the E lanes (lane distance L) of a piece and quad; every one of
// them ends with the same value (each step adds a pair in either order).
template <int W>
__device__ __forceinline__ float4 across_e(float4 v) {
#pragma unroll
  for (uint32_t m = Quad<W>::L; m < 8; m *= 2) {
    v.x += __shfl_xor(v.x, (int)m);
    v.y += __shfl_xor(v.y, (int)m);
    v.z += __shfl_xor(v.z, (int)m);
    v.w += __shfl_xor(v.w, (int)m);
  }
  return v;
}

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// 16-byte store that does not allocate in the caches (partial sums are read
// once, by the next kernel).
__device__ __forceinline__ void store_nt(float4* p, const float4& v) {
  const f32x4 w = {v.x, v.y, v.z, v.w};
  __builtin_nontemporal_store(w, reinterpret_cast<f32x4*>(p));
}

__device__ __forceinline__ void fma4(float a, const float4& g, float4& acc) {
  acc.x = fmaf(a, g.x, acc.x);
  acc.y = fmaf(a, g.y, acc.y);
  acc.z = fmaf(a, g.z, acc.z);
  acc.w = fmaf(a, g.w, acc.w);
}

// Value of lane m of this lane's run of L consecutive lanes (L = 1, 2, 4, 8;
// runs start at multiples of L), by DPP moves: quad permutes inside a quad,
// plus a row shift by 4 lanes into the other quad of a run of 8.
constexpr int qperm(int a, int b, int c, int d) { return a | (b << 2) | (c << 4) | (d << 6); }
template <uint32_t L, uint32_t m>
__device__ __forceinline__ uint32_t bcast(uint32_t x) {
  static_assert(m < L && (L == 1 || L == 2 || L == 4 || L == 8), "runs of 1, 2, 4 or 8 lanes");
  if constexpr (L == 1) {
    return x;
  } else if constexpr (L == 2) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, qperm(m, m, 2 + m, 2 + m), 0xF, 0xF, false);
  } else if constexpr (L == 4) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, qperm(m, m, m, m), 0xF, 0xF, false);
  } else {
    constexpr int k = m % 4;
    const int y = __builtin_amdgcn_mov_dpp((int)x, qperm(k, k, k, k), 0xF, 0xF, false);
    if constexpr (m < 4)  // upper quads (banks 1, 3) take the lower quad's value: row_shr:4
      return (uint32_t)__builtin_amdgcn_update_dpp(y, y, 0x114, 0xF, 0xA, false);
    else  // lower quads (banks 0, 2) take the upper quad's value: row_shl:4
      return (uint32_t)__builtin_amdgcn_update_dpp(y, y, 0x104, 0xF, 0x5, false);
  }
}

// Adds a batch of 8 staged elements [i, i + 8) of a list ending at e to the
// lane's sums: lane g8 = ei*L + q of its run of 8 reads element i + g8 from
// LDS, and share ei takes elements i + ei*L + m, m < L, from its run of L
// lanes; per element the lane reads its 16-byte quad of the operand row idx
// from LDS (past e: weight 0). The reads go out in groups of N before
// their multiply-adds.
template <uint32_t L, uint32_t m, uint32_t n, uint32_t N>
__device__ __forceinline__ void gather_rec(uint32_t idx, uint32_t vbits,
                                           const float4* __restrict__ op, uint32_t q, float4* g,
                                           float* v) {
  if constexpr (n < N) {
    g[n] = op[bcast<L, m + n>(idx) * L + q];
    v[n] = __uint_as_float(bcast<L, m + n>(vbits));
    gather_rec<L, m, n + 1, N>(idx, vbits, op, q, g, v);
  }
}
template <uint32_t L, uint32_t N, uint32_t m = 0>
__device__ __forceinline__ void batch8(uint32_t idx, float val, const float4* __restrict__ op,
                                       uint32_t q, float4& acc) {
  static_assert(L % N == 0, "reads in flight divide the batch");
  if constexpr (m < L) {
    float4 g[N];
    float v[N];
    gather_rec<L, m, 0, N>(idx, __float_as_uint(val), op, q, g, v);
#pragma unroll
    for (uint32_t n = 0; n < N; ++n) fma4(v[n], g[n], acc);
    batch8<L, N, m + N>(idx, val, op, q, acc);
  }
}

// Staging: every prefetch load is unconditional (indices clamped into the
// chunk) and every LDS store too (the arrays are padded to one entry per
// thread; entries past the chunk hold copies that are never read), so the
// loads are issued together, stay in flight while the current chunk is
// reduced and are waited for only at the next chunk's stores.

// Forward: part[fslot[s] * W + f] = sum over piece s of a * X[(c0 + col) * W + f]
// for the pieces of box border[blockIdx.x]; X in panel order. The box's
// pieces are staged in LDS chunk by chunk; the next chunk's elements and
// piece offsets are loaded into registers while the current one is reduced.
template <int W, bool kByBox>
__global__ __launch_bounds__(kBoxThreads) void k_box_fwd(
    const uint32_t* __restrict__ pstart, const uint32_t* __restrict__ border,
    const uint32_t* __restrict__ pfchunk, const uint2* __restrict__ fchunk,
    const uint32_t* __restrict__ fseg, const uint32_t* __restrict__ fslot,
    const float* __restrict__ fval, const uint16_t* __restrict__ fcol,
    const float* __restrict__ X, float* __restrict__ part) {
  constexpr uint32_t L = Quad<W>::L, E = Quad<W>::E, G = Quad<W>::G;
  constexpr uint32_t XJ = kBoxVox * L / kBoxThreads;
  static_assert(kFChunkElems == 4 * kBoxThreads && kFChunkPieces < kBoxThreads &&
                kFAlign % 4 == 0, "staging: four elements per thread, 16-byte aligned chunks");
  __shared__ float4 xs[kBoxVox * L];
  __shared__ float4 sval4[kFChunkElems / 4];
  __shared__ uint2 scol4[kFChunkElems / 4];
  const float* __restrict__ sval = reinterpret_cast<const float*>(sval4);
  const uint16_t* __restrict__ scol = reinterpret_cast<const uint16_t*>(scol4);
  __shared__ uint16_t soff[kBoxThreads + 1];  // soff[s + 1]: end of piece s in the chunk
  __shared__ uint32_t sslot[kBoxThreads];
  const uint32_t box = border[blockIdx.x];
  const uint32_t c0 = pstart[box], nv = pstart[box + 1] - c0;
  const uint32_t t = threadIdx.x, q = t % L, ei = (t / L) % E, grp = t / 8, g8 = t % 8;
  const uint32_t k0 = pfchunk[box], k1 = pfchunk[box + 1];
  f32x4 rv;
  u32x2 rc;
  uint32_t rend = 0, rslot = 0;
  auto fetch = [&](uint32_t k) __attribute__((always_inline)) {  // chunk k into registers
    const uint2 a = fchunk[k], b = fchunk[k + 1];
    const uint32_t np = b.x - a.x;
    // elements 4t .. 4t+3 of the chunk (the arrays end with a chunk of zeros);
    // streamed once: non-temporal
    rv = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(fval + a.y) + t);
    rc = __builtin_nontemporal_load(reinterpret_cast<const u32x2*>(fcol + a.y) + t);
    const uint32_t sp = a.x + min(t, np - 1);
    rend = fseg[sp + 1];
    if constexpr (!kByBox) rslot = fslot[sp];
  };
  if (k0 < k1) fetch(k0);
  {  // the box's slice of X
    const float4* __restrict__ X4 = reinterpret_cast<const float4*>(X) + (size_t)c0 * L;
    const uint32_t n = nv * L;
    float4 xr[XJ];
#pragma unroll
    for (uint32_t j = 0; j < XJ; ++j) xr[j] = X4[min(t + j * kBoxThreads, n - 1)];
#pragma unroll
    for (uint32_t j = 0; j < XJ; ++j) xs[t + j * kBoxThreads] = xr[j];
  }
  if (t == 0) soff[0] = 0;
  float4* __restrict__ part4 = reinterpret_cast<float4*>(part);
  for (uint32_t k = k0; k < k1; ++k) {
    const uint2 a = fchunk[k], b = fchunk[k + 1];
    const uint32_t np = b.x - a.x;
    __syncthreads();  // the previous chunk's readers are done with the LDS arrays
    sval4[t] = make_float4(rv.x, rv.y, rv.z, rv.w);
    scol4[t] = make_uint2(rc.x, rc.y);
    soff[t + 1] = (uint16_t)(rend - a.y);
    sslot[t] = kByBox ? a.x + t : rslot;
    __syncthreads();
    if (k + 1 < k1) fetch(k + 1);
    // pieces are longest first: rounds of G pieces go to the groups in snake
    // order, so that the waves get about the same work
    for (uint32_t r0 = 0; r0 < np; r0 += G) {
      const uint32_t sg = r0 + (((r0 / G) & 1u) ? G - 1 - grp : grp);
      if (sg >= np) continue;
      const uint32_t e = soff[sg + 1];
      float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
      for (uint32_t i = soff[sg]; i < e; i += 8) {
        const uint32_t my = i + g8, ic = min(my, e - 1);
        const uint32_t c = scol[ic];
        const float v = sval[ic];
        batch8<L, L>(c, my < e ? v : 0.f, xs, q, acc);
      }
      acc = across_e<W>(acc);
      if (ei == 0) store_nt(part4 + (size_t)sslot[sg] * L + q, acc);
    }
  }
}

// Backward: U[(c0 + v) * W + f] = sum over the box's elements of voxel v of
// a * R[row * W + f] (R in device row order), for box border[blockIdx.x].
// Lane group g owns voxels 8g .. 8g+7; per chunk the group adds those
// voxels' staged elements in order (E shares of every voxel's list). While a
// chunk is reduced, the next chunk's runs of R, elements and voxel offsets
// are loaded into registers, and the rows of the one after it.
template <int W>
__global__ __launch_bounds__(kBoxThreads) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_box_bwd(
    const uint32_t* __restrict__ pstart, const uint32_t* __restrict__ border,
    const uint32_t* __restrict__ pchunk, const uint2* __restrict__ chunk,
    const uint16_t* __restrict__ boff, const uint32_t* __restrict__ brow,
    const uint16_t* __restrict__ bpc, const float* __restrict__ bval,
    const float* __restrict__ R, float* __restrict__ U) {
  constexpr uint32_t L = Quad<W>::L, E = Quad<W>::E, V = kGroupVox;
  constexpr uint32_t kRq = (kChunkPieces * L + kBoxThreads - 1) / kBoxThreads;  // R quads
  static_assert(kChunkElems == 8 * kBoxThreads && kBoxVox == kBoxThreads && kBAlign % 8 == 0,
                "staging: eight elements per thread, 16-byte aligned chunks");
  __shared__ float4 rs[kRq * kBoxThreads];
  __shared__ float4 sval4[kChunkElems / 4];
  __shared__ uint4 spc4[kChunkElems / 8];
  const float* __restrict__ sval = reinterpret_cast<const float*>(sval4);
  const uint16_t* __restrict__ spc = reinterpret_cast<const uint16_t*>(spc4);
  __shared__ uint16_t soff[kBoxVox + 1];  // soff[v + 1]: end of voxel v's elements
  const uint32_t box = border[blockIdx.x];
  const uint32_t c0 = pstart[box], nv = pstart[box + 1] - c0;
  const uint32_t t = threadIdx.x, q = t % L, ei = (t / L) % E, v0 = (t / 8) * V, g8 = t % 8;
  const uint32_t k0 = pchunk[box], k1 = pchunk[box + 1];
  const float4* __restrict__ R4 = reinterpret_cast<const float4*>(R);
  // vector registers (not arrays), so they are never placed in memory
  typedef float rq_t __attribute__((ext_vector_type(4 * kRq)));
  typedef uint32_t rr_t __attribute__((ext_vector_type(kRq)));
  rq_t rR;
  rr_t rrow;
  f32x4 rva, rvb;
  u32x4 rp;
  uint32_t rend = 0;
  auto rows_of = [&](uint32_t k) __attribute__((always_inline)) {  // rows of chunk k's R quads
    const uint2 a = chunk[k], b = chunk[k + 1];
    const uint32_t nq = (b.x - a.x) * L;
#pragma unroll
    for (uint32_t j = 0; j < kRq; ++j) rrow[j] = brow[a.x + min(t + j * kBoxThreads, nq - 1) / L];

  };
  auto fetch = [&](uint32_t k) __attribute__((always_inline)) {  // chunk k (its rows in rrow)
    const uint2 a = chunk[k];
#pragma unroll
    for (uint32_t j = 0; j < kRq; ++j) {
      const float4 g = R4[(size_t)rrow[j] * L + (t + j * kBoxThreads) % L];
      rR[4 * j] = g.x;
      rR[4 * j + 1] = g.y;
      rR[4 * j + 2] = g.z;
      rR[4 * j + 3] = g.w;
    }
    // elements 8t .. 8t+7 of the chunk (the arrays end with a chunk of
    // zeros); streamed once: non-temporal
    rva = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(bval + a.y) + 2 * t);
    rvb = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(bval + a.y) + 2 * t + 1);
    rp = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(bpc + a.y) + t);
    rend = boff[(size_t)k * (kBoxVox + 1) + t + 1];
  };
  if (k0 < k1) {
    rows_of(k0);
    fetch(k0);
    if (k0 + 1 < k1) rows_of(k0 + 1);
  }
  if (t == 0) soff[0] = 0;
  float4 acc[V];
#pragma unroll
  for (uint32_t j = 0; j < V; ++j) acc[j] = make_float4(0.f, 0.f, 0.f, 0.f);
  for (uint32_t k = k0; k < k1; ++k) {
    __syncthreads();  // the previous chunk's readers are done with the LDS arrays
#pragma unroll
    for (uint32_t j = 0; j < kRq; ++j)
      rs[t + j * kBoxThreads] = make_float4(rR[4 * j], rR[4 * j + 1], rR[4 * j + 2], rR[4 * j + 3]);
    sval4[2 * t] = make_float4(rva.x, rva.y, rva.z, rva.w);
    sval4[2 * t + 1] = make_float4(rvb.x, rvb.y, rvb.z, rvb.w);
    spc4[t] = make_uint4(rp.x, rp.y, rp.z, rp.w);
    soff[t + 1] = (uint16_t)rend;
    __syncthreads();
    if (k + 1 < k1) {
      fetch(k + 1);
      if (k + 2 < k1) rows_of(k + 2);
    }
#pragma unroll
    for (uint32_t j = 0; j < V; ++j) {
      const uint32_t e = soff[v0 + j + 1];
      for (uint32_t i = soff[v0 + j]; i < e; i += 8) {
        const uint32_t my = i + g8, ic = min(my, e - 1);
        const uint32_t pc = spc[ic];
        const float v = sval[ic];
        batch8<L, (L < 4 ? L : 4)>(pc, my < e ? v : 0.f, rs, q, acc[j]);  // registers: 4
      }
    }
  }
  float4* __restrict__ U4 = reinterpret_cast<float4*>(U);
#pragma unroll
  for (uint32_t j = 0; j < V; ++j) {
    const float4 u = across_e<W>(acc[j]);
    if (ei == 0 && v0 + j < nv) U4[(size_t)(c0 + v0 + j) * L + q] = u;
  }
}

}  // namespace

PanelFormat::~PanelFormat() {
  dfree(pc_of_file);
  dfree(pc_of_dev);
  dfree(dev_of_pc);
  dfree(norm);
  dfree(pstart);
  dfree(border);
  dfree(pseg);
  dfree(fseg);
  dfree(fslot);
  dfree(fval);
  dfree(fcol);
  dfree(rowslot);
  dfree(rowpiece);
  dfree(pfchunk);
  dfree(fchunk);
  dfree(brow);
  dfree(pchunk);
  dfree(chunk);
  dfree(boff);
  dfree(bpc);
  dfree(bval);
  dfree(colslot);
}

const PanelFormat& ensure_panels(DevMatrix& m) {
  if (m.panels) return *m.panels;
  if (m.streamed || !m.csr) raise(MLEM_EINVAL, "the tiled product needs a matrix held in HBM");
  if (m.vk != kF32) raise(MLEM_EINVAL, "the tiled product needs fp32 matrix values");
  auto t0 = std::chrono::steady_clock::now();
  hipStream_t s = m.stream;
  // device-order CSR and row pointers, device column of each file column
  std::vector<uint32_t> csr2(2 * m.nnz), rowptr((size_t)m.rows + 1), colmap(m.cols);
  std::vector<float> norm(m.cols);
  if (m.nnz)
    MLEM_HIP(hipMemcpyAsync(csr2.data(), m.csr, 8ull * m.nnz, hipMemcpyDeviceToHost, s));
  MLEM_HIP(hipMemcpyAsync(rowptr.data(), m.rowptr, 4ull * (m.rows + 1), hipMemcpyDeviceToHost, s));
  if (m.colmap && m.cols)
    MLEM_HIP(hipMemcpyAsync(colmap.data(), m.colmap, 4ull * m.cols, hipMemcpyDeviceToHost, s));
  if (m.cols)
    MLEM_HIP(hipMemcpyAsync(norm.data(), m.norm, 4ull * m.cols, hipMemcpyDeviceToHost, s));
  MLEM_HIP(hipStreamSynchronize(s));
  if (!m.colmap) std::iota(colmap.begin(), colmap.end(), 0u);

  std::vector<uint32_t> pc_of_file, pstart;
  uint32_t boxgrid[3] = {0, 0, 0};
  const bool has_grid = (uint64_t)m.grid[0] * m.grid[1] * m.grid[2] == m.cols && m.cols;
  std::vector<uint32_t> count_dev(m.cols, 0), count_file(m.cols);
  for (uint64_t k = 0; k < m.nnz; ++k) count_dev[csr2[2 * k]]++;
  for (uint32_t c = 0; c < m.cols; ++c) count_file[c] = count_dev[colmap[c]];
  panel_order(m.cols, has_grid ? m.grid : nullptr, count_file.data(), &pc_of_file, &pstart,
              boxgrid);
  std::vector<uint32_t> pc_of_dev(m.cols), dev_of_pc(m.cols);
  for (uint32_t c = 0; c < m.cols; ++c) pc_of_dev[colmap[c]] = pc_of_file[c];
  for (uint32_t d = 0; d < m.cols; ++d) dev_of_pc[pc_of_dev[d]] = d;
  std::vector<float> pnorm(m.cols);
  for (uint32_t c = 0; c < m.cols; ++c) pnorm[c] = norm[dev_of_pc[c]];

  HostPanels H;
  try {
    build_host_panels(m.rows, m.cols, m.nnz, rowptr.data(), csr2.data(), pc_of_dev, pstart,
                      boxgrid, 0, &H);
  } catch (const std::invalid_argument& e) {
    raise(MLEM_EINVAL, std::string("matrix tiling: ") + e.what());
  }
  std::vector<uint32_t>().swap(csr2);
  const uint32_t nchunks = (uint32_t)H.chunk_piece.size() - 1;
  const uint32_t nfchunks = (uint32_t)H.fchunk_piece.size() - 1;
  std::vector<uint2> fchunk((size_t)nfchunks + 1);
  for (uint32_t k = 0; k <= nfchunks; ++k)
    fchunk[k] = make_uint2(H.fchunk_piece[k], H.fseg[H.fchunk_piece[k]]);
  std::vector<uint2> chunk((size_t)nchunks + 1);
  for (uint32_t k = 0; k <= nchunks; ++k) chunk[k] = make_uint2(H.chunk_piece[k], H.chunk_elem[k]);
  std::vector<uint32_t> colslot((size_t)m.cols + 1);
  std::iota(colslot.begin(), colslot.end(), 0u);

  std::unique_ptr<PanelFormat> P(new PanelFormat());
  P->rows = m.rows;
  P->cols = m.cols;
  P->npanels = H.npanels;
  P->nchunks = nchunks;
  P->nfchunks = nfchunks;
  P->nseg = H.fseg.size() - 1;
  P->ngroups = m.cols;
  upload(&P->pc_of_file, pc_of_file, s);
  upload(&P->pc_of_dev, pc_of_dev, s);
  upload(&P->dev_of_pc, dev_of_pc, s);
  upload(&P->norm, pnorm, s);
  upload(&P->pstart, H.pstart, s);
  upload(&P->border, H.border, s);
  upload(&P->pseg, H.pseg, s);
  upload(&P->fseg, H.fseg, s);
  upload(&P->fslot, H.fslot, s);
  upload(&P->fval, H.fval, s);
  upload(&P->fcol, H.fcol, s);
  upload(&P->rowslot, H.rowslot, s);
  std::vector<uint32_t> rowpiece(H.fslot.size());
  for (size_t k = 0; k < H.fslot.size(); ++k) rowpiece[H.fslot[k]] = (uint32_t)k;
  upload(&P->rowpiece, rowpiece, s);
  upload(&P->pfchunk, H.pfchunk, s);
  upload(&P->fchunk, fchunk, s);
  upload(&P->brow, H.brow, s);
  upload(&P->pchunk, H.pchunk, s);
  upload(&P->chunk, chunk, s);
  upload(&P->boff, H.boff, s);
  upload(&P->bpc, H.bpc, s);
  upload(&P->bval, H.bval, s);
  upload(&P->colslot, colslot, s);
  MLEM_HIP(hipStreamSynchronize(s));  // host vectors go out of scope
  P->build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  m.panels = P.release();
  return *m.panels;
}

void free_panels(DevMatrix& m) {
  delete m.panels;
  m.panels = nullptr;
}

template <int W>
void launch_panel_fwd(const PanelFormat& P, const float* X, float* part, uint32_t /*grid*/,
                      hipStream_t s) {
  if (!P.npanels) return;
  hipLaunchKernelGGL((k_box_fwd<W, panel_slots_by_box(W)>), dim3(P.npanels), dim3(kBoxThreads), 0, s, P.pstart,
                     P.border, P.pfchunk, P.fchunk, P.fseg, P.fslot, P.fval, P.fcol, X, part);
  check_launch("tiled forward launch");
}

template <int W>
void launch_panel_bwd(const PanelFormat& P, const float* R, float* part, uint32_t /*grid*/,
                      hipStream_t s) {
  if (!P.npanels) return;
  hipLaunchKernelGGL((k_box_bwd<W>), dim3(P.npanels), dim3(kBoxThreads), 0, s, P.pstart,
                     P.border, P.pchunk, P.chunk, P.boff, P.brow, P.bpc, P.bval, R, part);
  check_launch("tiled backward launch");
}

#define MLEM_PANEL_WIDTH(W)                                                                  \
  template void launch_panel_fwd<W>(const PanelFormat&, const float*, float*, uint32_t,      \
                                    hipStream_t);                                            \
  template void launch_panel_bwd<W>(const PanelFormat&, const float*, float*, uint32_t, hipStream_t);
MLEM_PANEL_WIDTH(4)
MLEM_PANEL_WIDTH(8)
MLEM_PANEL_WIDTH(16)
MLEM_PANEL_WIDTH(32)
#undef MLEM_PANEL_WIDTH

}  // namespace mlem
