// panel.hip — products of the system matrix with blocks of W interleaved
// right-hand sides through its two-dimensional tiling (panels.hpp), gfx950.
//
// Forward, f = A X: a workgroup takes a tile of segments of one panel
// ((row, panel) pieces of rows, at most kMaxSegment elements each), stages
// their values and 16-bit in-panel columns in LDS with coalesced loads and
// gives every segment to 8 lanes: W/4 quads of frames times 8/(W/4) shares of
// the elements (every 8/(W/4)-th element). Per element a lane gathers one
// 16-byte quad of X[(panel start + column) * W + ...] (the quads of the W/4
// lanes make the element's W*4-byte run, from the panel's 2 MB slice of X at
// W = 32) and accumulates with fused multiply-adds; a butterfly over the
// lanes adds the shares and the segment's sums go to its slot. Backward,
// u = A^T R: the same over the (panel, row chunk, voxel) groups, gathering
// R[row * W + ...] from the chunk's rows. Workgroup b runs on XCD b % 8 and XCD x walks the
// x-th contiguous eighth of the tiles, which are stored panel after panel
// (chunk after chunk): the 256 workgroups of one XCD work through the same
// panel (chunk) together, so its gathered slice stays in that XCD's L2.
//
// The slots of one row (voxel) are contiguous; the epilogue kernels of the
// callers add them in panel (chunk) order, so every result is deterministic.
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

constexpr int kBlock = 256;
constexpr uint32_t kPTile = 2048;  // elements staged per tile
constexpr uint32_t kChunkRows = 16384;
constexpr int kUnroll = 8;  // 16-byte gathers in flight per lane

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

static_assert(kMaxSegment <= kPTile, "every segment fits a staged tile");

// Greedy tiles over segments [s, e) with element offsets off[]: at most kBlock
// segments and kPTile elements per tile (no segment is longer than that).
void add_tiles(const std::vector<uint32_t>& off, uint32_t s, uint32_t e, uint32_t base,
               std::vector<uint4>* tiles) {
  while (s < e) {
    uint32_t t = s + 1;
    while (t < e && t - s < (uint32_t)kBlock && off[t + 1] - off[s] <= kPTile) ++t;
    tiles->push_back(make_uint4(s, off[s], base, 0));
    s = t;
  }
}

// A lane carries four right-hand sides (one 16-byte gather per element): the
// address unit handles a fixed number of lanes per clock, so 16-byte lanes
// move four times the bytes of 4-byte ones. Every segment gets 8 lanes
// whatever W is: W/4 lanes for the quads times E = 32/W lanes that take
// every E-th element, their sums then added across the E lanes by a fixed
// butterfly, so narrow blocks keep as many gathers in flight as wide ones.
template <int W>
struct Quad {
  static_assert(W % 4 == 0 && W <= 32, "right-hand sides come in quads, at most 32");
  static constexpr uint32_t L = W / 4;        // lanes per element (quads)
  static constexpr uint32_t E = 8 / L;        // lanes per segment and quad
  static constexpr uint32_t G = kBlock / 8;   // segments in flight per workgroup
};

// Sum over the E lanes (lane distance L) of a segment and quad; every one of
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

// 16-byte store that does not allocate in the caches (partial sums are read
// once, by the next kernel).
__device__ __forceinline__ void store_nt(float4* p, const float4& v) {
  const f32x4 w = {v.x, v.y, v.z, v.w};
  __builtin_nontemporal_store(w, reinterpret_cast<f32x4*>(p));
}

__device__ __forceinline__ uint2 load_nt(const uint2* p) {
  const u32x2 w = __builtin_nontemporal_load(reinterpret_cast<const u32x2*>(p));
  return make_uint2(w.x, w.y);
}

// Gathers through a buffer resource: one 32-bit offset register per load
// instead of a 64-bit address (reads past `bytes` return zeros).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes,
                                           0x00020000);
}
__device__ __forceinline__ float4 gather4(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  const f32x4 v = __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
  return make_float4(v.x, v.y, v.z, v.w);
}

__device__ __forceinline__ void fma4(float a, const float4& g, float4& acc) {
  acc.x = fmaf(a, g.x, acc.x);
  acc.y = fmaf(a, g.y, acc.y);
  acc.z = fmaf(a, g.z, acc.z);
  acc.w = fmaf(a, g.w, acc.w);
}

// This lane's share of one staged segment's dot products (elements b + ei,
// b + ei + E, ... of [b, e)) with its four vectors: U gathers issued
// together, then fused multiply-adds in element order. off0 is the byte
// offset in X of the lane's quad of the panel's first column; each further
// column is W*4 bytes on.
template <int W, int U>
__device__ __forceinline__ float4 pdot(const float* vals, const uint16_t* cols, uint32_t b,
                                       uint32_t e, uint32_t ei, __amdgpu_buffer_rsrc_t rx,
                                       uint32_t off0) {
  constexpr uint32_t E = Quad<W>::E;
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  for (uint32_t k = b + ei; k < e; k += U * E) {
    float4 g[U];
#pragma unroll
    for (int j = 0; j < U; ++j)
      g[j] = gather4(rx, off0 + (uint32_t)cols[min(k + j * E, e - 1)] * (W * 4u));
    // the values are read once the gathers are in flight (fewer live registers)
#pragma unroll
    for (int j = 0; j < U; ++j)
      if (k + j * E < e) fma4(vals[k + j * E], g[j], acc);
  }
  return across_e<W>(acc);
}

template <int W, int U>
__device__ __forceinline__ float4 rdot(const uint2* ent, uint32_t b, uint32_t e, uint32_t ei,
                                       __amdgpu_buffer_rsrc_t rr, uint32_t off0) {
  constexpr uint32_t E = Quad<W>::E;
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  for (uint32_t k = b + ei; k < e; k += U * E) {
    float4 g[U];
#pragma unroll
    for (int j = 0; j < U; ++j) g[j] = gather4(rr, off0 + ent[min(k + j * E, e - 1)].x * (W * 4u));
#pragma unroll
    for (int j = 0; j < U; ++j)
      if (k + j * E < e) fma4(__uint_as_float(ent[k + j * E].y), g[j], acc);
  }
  return across_e<W>(acc);
}

// Forward tiles: part[fslot[s] * W + f] = sum over segment s of a * X[.].
// Both tile kernels are held to 64 VGPRs (8 waves per SIMD): the launch is
// a persistent grid of 8 workgroups per CU, all resident at once.
template <int W>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8))) void k_panel_fwd(
    const uint32_t* __restrict__ fseg, const uint32_t* __restrict__ fslot,
    const float* __restrict__ fval, const uint16_t* __restrict__ fcol,
    const uint4* __restrict__ tiles, uint32_t ntiles, const float* __restrict__ X,
    uint32_t xbytes, float* __restrict__ part) {
  constexpr uint32_t L = Quad<W>::L, E = Quad<W>::E, G = Quad<W>::G;
  __shared__ float sval[kPTile];
  __shared__ uint16_t scol[kPTile];
  __shared__ uint32_t S[kBlock + 1];
  __shared__ uint32_t SL[kBlock];
  const uint32_t t = threadIdx.x, q = t % L, ei = (t / L) % E, grp = t / 8;
  const uint32_t xcd = blockIdx.x & 7u, per_xcd = gridDim.x >> 3;
  const uint32_t t0 = (uint32_t)((uint64_t)ntiles * xcd / 8);
  const uint32_t t1 = (uint32_t)((uint64_t)ntiles * (xcd + 1) / 8);
  float4* __restrict__ part4 = reinterpret_cast<float4*>(part);
  for (uint32_t tile = t0 + (blockIdx.x >> 3); tile < t1; tile += per_xcd) {
    const uint4 d0 = tiles[tile], d1 = tiles[tile + 1];
    const uint32_t s0 = d0.x, e0 = d0.y, nseg = d1.x - s0, count = d1.y - e0;
    __syncthreads();  // the previous tile's readers are done with the LDS arrays
    for (uint32_t i = t; i < count; i += kBlock) {  // streamed once: non-temporal
      sval[i] = __builtin_nontemporal_load(fval + e0 + i);
      scol[i] = __builtin_nontemporal_load(fcol + e0 + i);
    }
    for (uint32_t i = t; i <= nseg; i += kBlock) S[i] = fseg[s0 + i] - e0;
    for (uint32_t i = t; i < nseg; i += kBlock) SL[i] = fslot[s0 + i];
    __syncthreads();
    const auto rx = rsrc(X, xbytes);
    const uint32_t off0 = d0.z * (W * 4u) + q * 16u;
    for (uint32_t sg = grp; sg < nseg; sg += G) {
      const float4 d = pdot<W, kUnroll>(sval, scol, S[sg], S[sg + 1], ei, rx, off0);
      if (ei == 0) store_nt(part4 + (size_t)SL[sg] * L + q, d);
    }
  }
}

// Backward tiles: part[bslot[g] * W + f] = sum over group g of a * R[row * W + f].
template <int W>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8))) void k_panel_bwd(
    const uint32_t* __restrict__ bseg, const uint32_t* __restrict__ bslot,
    const uint2* __restrict__ bent, const uint4* __restrict__ tiles, uint32_t ntiles,
    const float* __restrict__ R, uint32_t rbytes, float* __restrict__ part) {
  constexpr uint32_t L = Quad<W>::L, E = Quad<W>::E, G = Quad<W>::G;
  __shared__ uint2 sent[kPTile];
  __shared__ uint32_t S[kBlock + 1];
  __shared__ uint32_t SL[kBlock];
  const uint32_t t = threadIdx.x, q = t % L, ei = (t / L) % E, grp = t / 8;
  const uint32_t xcd = blockIdx.x & 7u, per_xcd = gridDim.x >> 3;
  const uint32_t t0 = (uint32_t)((uint64_t)ntiles * xcd / 8);
  const uint32_t t1 = (uint32_t)((uint64_t)ntiles * (xcd + 1) / 8);
  const auto rr = rsrc(R, rbytes);
  const uint32_t off0 = q * 16u;
  float4* __restrict__ part4 = reinterpret_cast<float4*>(part);
  for (uint32_t tile = t0 + (blockIdx.x >> 3); tile < t1; tile += per_xcd) {
    const uint4 d0 = tiles[tile], d1 = tiles[tile + 1];
    const uint32_t s0 = d0.x, e0 = d0.y, nseg = d1.x - s0, count = d1.y - e0;
    __syncthreads();
    for (uint32_t i = t; i < count; i += kBlock) sent[i] = load_nt(bent + e0 + i);
    for (uint32_t i = t; i <= nseg; i += kBlock) S[i] = bseg[s0 + i] - e0;
    for (uint32_t i = t; i < nseg; i += kBlock) SL[i] = bslot[s0 + i];
    __syncthreads();
    for (uint32_t sg = grp; sg < nseg; sg += G) {
      const float4 d = rdot<W, kUnroll>(sent, S[sg], S[sg + 1], ei, rr, off0);
      if (ei == 0) store_nt(part4 + (size_t)SL[sg] * L + q, d);
    }
  }
}

}  // namespace

PanelFormat::~PanelFormat() {
  dfree(pc_of_file);
  dfree(pc_of_dev);
  dfree(dev_of_pc);
  dfree(norm);
  dfree(fseg);
  dfree(fslot);
  dfree(fval);
  dfree(fcol);
  dfree(rowslot);
  dfree(ftiles);
  dfree(bseg);
  dfree(bslot);
  dfree(bent);
  dfree(colslot);
  dfree(btiles);
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
  const bool has_grid = (uint64_t)m.grid[0] * m.grid[1] * m.grid[2] == m.cols && m.cols;
  panel_order(m.cols, has_grid ? m.grid : nullptr, &pc_of_file, &pstart);
  std::vector<uint32_t> pc_of_dev(m.cols), dev_of_pc(m.cols);
  for (uint32_t c = 0; c < m.cols; ++c) pc_of_dev[colmap[c]] = pc_of_file[c];
  for (uint32_t d = 0; d < m.cols; ++d) dev_of_pc[pc_of_dev[d]] = d;
  std::vector<float> pnorm(m.cols);
  for (uint32_t c = 0; c < m.cols; ++c) pnorm[c] = norm[dev_of_pc[c]];

  HostPanels H;
  try {
    build_host_panels(m.rows, m.cols, m.nnz, rowptr.data(), csr2.data(), pc_of_dev, pstart,
                      kChunkRows, 0, &H);
  } catch (const std::invalid_argument& e) {
    raise(MLEM_EINVAL, std::string("matrix tiling: ") + e.what());
  }
  std::vector<uint32_t>().swap(csr2);
  std::vector<uint4> ft, bt;
  for (uint32_t p = 0; p < H.npanels; ++p) add_tiles(H.fseg, H.pseg[p], H.pseg[p + 1], H.pstart[p], &ft);
  ft.push_back(make_uint4((uint32_t)(H.fseg.size() - 1), (uint32_t)m.nnz, 0, 0));
  const uint32_t nchunks = (uint32_t)H.cgroup.size() - 1;
  for (uint32_t c = 0; c < nchunks; ++c) add_tiles(H.bseg, H.cgroup[c], H.cgroup[c + 1], 0, &bt);
  bt.push_back(make_uint4((uint32_t)(H.bseg.size() - 1), (uint32_t)m.nnz, 0, 0));
  std::vector<uint2> bent(m.nnz);
  for (uint64_t k = 0; k < m.nnz; ++k) {
    uint32_t bits;
    std::memcpy(&bits, &H.bval[k], 4);
    bent[k] = make_uint2(H.brow[k], bits);
  }

  std::unique_ptr<PanelFormat> P(new PanelFormat());
  P->rows = m.rows;
  P->cols = m.cols;
  P->npanels = H.npanels;
  P->nchunks = nchunks;
  P->nseg = H.fseg.size() - 1;
  P->ngroups = H.bseg.size() - 1;
  P->nftiles = (uint32_t)ft.size() - 1;
  P->nbtiles = (uint32_t)bt.size() - 1;
  upload(&P->pc_of_file, pc_of_file, s);
  upload(&P->pc_of_dev, pc_of_dev, s);
  upload(&P->dev_of_pc, dev_of_pc, s);
  upload(&P->norm, pnorm, s);
  upload(&P->fseg, H.fseg, s);
  upload(&P->fslot, H.fslot, s);
  upload(&P->fval, H.fval, s);
  upload(&P->fcol, H.fcol, s);
  upload(&P->rowslot, H.rowslot, s);
  upload(&P->ftiles, ft, s);
  upload(&P->bseg, H.bseg, s);
  upload(&P->bslot, H.bslot, s);
  upload(&P->bent, bent, s);
  upload(&P->colslot, H.colslot, s);
  upload(&P->btiles, bt, s);
  MLEM_HIP(hipStreamSynchronize(s));  // host vectors go out of scope
  P->build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  m.panels = P.release();
  return *m.panels;
}

void free_panels(DevMatrix& m) {
  delete m.panels;
  m.panels = nullptr;
}

// Byte size of a [n][W] fp32 block for a buffer resource (32-bit sizes).
uint32_t block_bytes(uint64_t n, int W) {
  const uint64_t b = 4ull * n * (uint64_t)W;
  if (b > 0xFFFFFFFFull) raise(MLEM_EINVAL, "right-hand-side block larger than 4 GiB");
  return (uint32_t)b;
}

template <int W>
void launch_panel_fwd(const PanelFormat& P, const float* X, float* part, uint32_t grid,
                      hipStream_t s) {
  hipLaunchKernelGGL((k_panel_fwd<W>), dim3(grid), dim3(kBlock), 0, s, P.fseg, P.fslot, P.fval,
                     P.fcol, P.ftiles, P.nftiles, X, block_bytes(P.cols, W), part);
  check_launch("tiled forward launch");
}

template <int W>
void launch_panel_bwd(const PanelFormat& P, const float* R, float* part, uint32_t grid,
                      hipStream_t s) {
  hipLaunchKernelGGL((k_panel_bwd<W>), dim3(grid), dim3(kBlock), 0, s, P.bseg, P.bslot, P.bent,
                     P.btiles, P.nbtiles, R, block_bytes(P.rows, W), part);
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
