// kernels.hip — MLEM kernels for gfx950 (CDNA4, wave64).
//
// Hot path, per iteration:
//   k_fwd  : CSR forward projection f = A x; each row is owned by a group of
//            G lanes (G in 4..64) that stream the row's 8-byte {column,value}
//            records with non-temporal loads (the matrix is read once per pass
//            and must not evict x from L2), gather x, and reduce with
//            cross-lane shuffles. The epilogue fuses the ratio
//            r = (f != 0) ? y/f : 0 (reference calcCorrel,
//            mlem_seq/seqcsr4mlem.cpp:94-105) and, optionally, the FP64
//            Poisson log-likelihood and sum of f with a deterministic
//            last-block reduction.
//   k_bwd  : CSC gather backprojection u = A^T r with the multiplicative
//            update fused in the epilogue, x *= (n != 0) ? u/n : u (reference
//            calcBkProj + calcUpdate, seqcsr4mlem.cpp:108-134). No atomics,
//            bit-for-bit deterministic.
//   k_scatter + k_update : the atomic-scatter alternative over the CSR.
//
// G = 1 with EXACT = true gives the reference's rounding sequence: one lane
// walks a row (column) in ascending order, multiply and add rounded
// separately (this file is compiled with -ffp-contract=off), results flushed
// to zero (-fgpu-flush-denormals-to-zero), IEEE-correct division.
//
// Launch mapping: workgroup b runs on XCD b % 8; logical_block() renumbers
// so that each XCD owns one contiguous eighth of the rows (columns), which
// keeps the CSC pass's r gathers inside one XCD's L2.
#include <rocprim/device/device_radix_sort.hpp>

#include "internal.hpp"
#include "mlem.h"

namespace mlem {
namespace {

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int kBlock = 256;
constexpr int kWave = 64;

__device__ __forceinline__ u32x2 ld_stream(const uint2* p) {
  return __builtin_nontemporal_load(reinterpret_cast<const u32x2*>(p));
}

__device__ __forceinline__ uint32_t logical_block() {
  const uint32_t b = blockIdx.x, g = gridDim.x;  // g is a multiple of 8
  return (b & 7u) * (g >> 3) + (b >> 3);
}

template <int G>
__device__ __forceinline__ float group_sum(float v) {
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, G);
  return v;
}

// Dot product of one sparse row (or CSC column) with a dense vector.
template <int G, bool EXACT>
__device__ __forceinline__ float seg_dot(const uint2* __restrict__ a, const float* __restrict__ v,
                                         uint32_t beg, uint32_t end, uint32_t lane) {
  if constexpr (EXACT) {
    float acc = 0.f;
    for (uint32_t k = beg; k < end; ++k) {
      const u32x2 e = ld_stream(a + k);
      acc = acc + __uint_as_float(e.y) * v[e.x];
    }
    return acc;
  } else {
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
    uint32_t k = beg + lane;
    for (; k + 3 * G < end; k += 4 * G) {
      const u32x2 e0 = ld_stream(a + k);
      const u32x2 e1 = ld_stream(a + k + G);
      const u32x2 e2 = ld_stream(a + k + 2 * G);
      const u32x2 e3 = ld_stream(a + k + 3 * G);
      a0 = __builtin_fmaf(__uint_as_float(e0.y), v[e0.x], a0);
      a1 = __builtin_fmaf(__uint_as_float(e1.y), v[e1.x], a1);
      a2 = __builtin_fmaf(__uint_as_float(e2.y), v[e2.x], a2);
      a3 = __builtin_fmaf(__uint_as_float(e3.y), v[e3.x], a3);
    }
    for (; k < end; k += G) {
      const u32x2 e = ld_stream(a + k);
      a0 = __builtin_fmaf(__uint_as_float(e.y), v[e.x], a0);
    }
    return (a0 + a1) + (a2 + a3);
  }
}

// Deterministic FP64 block + grid reduction of (ll, tot). Every block writes
// its partial; the block that finishes last sums all partials in a fixed
// order and appends the result to the ring.
__device__ void stats_finish(double ll, double tot, double2* __restrict__ partials,
                             unsigned* __restrict__ counter, double2* __restrict__ results,
                             unsigned ring) {
  __shared__ double s_ll[kBlock / kWave], s_tot[kBlock / kWave];
  __shared__ int s_last;
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    ll += __shfl_xor(ll, off);
    tot += __shfl_xor(tot, off);
  }
  const int wave = threadIdx.x / kWave, wl = threadIdx.x % kWave;
  if (wl == 0) {
    s_ll[wave] = ll;
    s_tot[wave] = tot;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double bl = 0.0, bt = 0.0;
    for (int w = 0; w < kBlock / kWave; ++w) {
      bl += s_ll[w];
      bt += s_tot[w];
    }
    __hip_atomic_store(&partials[blockIdx.x].x, bl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&partials[blockIdx.x].y, bt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    const unsigned done = atomicAdd(&counter[0], 1u);
    s_last = (done == gridDim.x - 1);
  }
  __syncthreads();
  if (!s_last) return;
  __threadfence();
  double pl = 0.0, pt = 0.0;
  for (unsigned i = threadIdx.x; i < gridDim.x; i += kBlock) {
    pl += __hip_atomic_load(&partials[i].x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    pt += __hip_atomic_load(&partials[i].y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    pl += __shfl_xor(pl, off);
    pt += __shfl_xor(pt, off);
  }
  __syncthreads();
  if (wl == 0) {
    s_ll[wave] = pl;
    s_tot[wave] = pt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double bl = 0.0, bt = 0.0;
    for (int w = 0; w < kBlock / kWave; ++w) {
      bl += s_ll[w];
      bt += s_tot[w];
    }
    const unsigned slot = atomicAdd(&counter[1], 1u);
    results[slot % ring] = make_double2(bl, bt);
    atomicExch(&counter[0], 0u);
  }
}

// ---------------------------------------------------------------- forward --
template <int G, bool EXACT, bool RATIO, bool STATS>
__global__ __launch_bounds__(kBlock) void k_fwd(
    const uint32_t* __restrict__ rowptr, const uint2* __restrict__ csr,
    const float* __restrict__ x, const int32_t* __restrict__ y, float* __restrict__ out,
    uint32_t rows, uint32_t chunk, double2* __restrict__ partials, unsigned* __restrict__ counter,
    double2* __restrict__ results, unsigned ring) {
  constexpr uint32_t GPB = kBlock / G;
  const uint32_t lane = threadIdx.x % G;
  const uint32_t grp = threadIdx.x / G;
  const uint64_t r0 = (uint64_t)logical_block() * chunk;
  const uint64_t r1 = r0 + chunk < rows ? r0 + chunk : rows;
  double ll = 0.0, tot = 0.0;
  for (uint64_t base = r0; base < r1; base += GPB) {
    const uint32_t row = (uint32_t)(base + grp);
    const bool active = base + grp < r1;
    float acc = 0.f;
    if (active) acc = seg_dot<G, EXACT>(csr, x, rowptr[row], rowptr[row + 1], lane);
    acc = group_sum<G>(acc);
    if (active && lane == 0) {
      if constexpr (RATIO) {
        const int32_t yi = y[row];
        out[row] = (acc != 0.f) ? (float)yi / acc : 0.f;
        if constexpr (STATS) {
          tot += (double)acc;
          if (acc > 0.f) ll += (yi != 0 ? (double)yi * log((double)acc) : 0.0) - (double)acc;
        }
      } else {
        out[row] = acc;
      }
    }
  }
  if constexpr (STATS) stats_finish(ll, tot, partials, counter, results, ring);
}

// --------------------------------------------------------------- backward --
template <int G, bool EXACT, bool UPDATE>
__global__ __launch_bounds__(kBlock) void k_bwd(const uint32_t* __restrict__ colptr,
                                                const uint2* __restrict__ csc,
                                                const float* __restrict__ r,
                                                const float* __restrict__ norm,
                                                float* __restrict__ out, uint32_t cols,
                                                uint32_t chunk) {
  constexpr uint32_t GPB = kBlock / G;
  const uint32_t lane = threadIdx.x % G;
  const uint32_t grp = threadIdx.x / G;
  const uint64_t c0 = (uint64_t)logical_block() * chunk;
  const uint64_t c1 = c0 + chunk < cols ? c0 + chunk : cols;
  for (uint64_t base = c0; base < c1; base += GPB) {
    const uint32_t col = (uint32_t)(base + grp);
    const bool active = base + grp < c1;
    float acc = 0.f;
    if (active) acc = seg_dot<G, EXACT>(csc, r, colptr[col], colptr[col + 1], lane);
    acc = group_sum<G>(acc);
    if (active && lane == 0) {
      if constexpr (UPDATE) {
        const float n = norm[col];
        out[col] = out[col] * ((n != 0.f) ? acc / n : acc);
      } else {
        out[col] = acc;
      }
    }
  }
}

// ------------------------------------------------------ atomic alternative --
template <int G>
__global__ __launch_bounds__(kBlock) void k_scatter(const uint32_t* __restrict__ rowptr,
                                                    const uint2* __restrict__ csr,
                                                    const float* __restrict__ r,
                                                    float* __restrict__ u, uint32_t rows,
                                                    uint32_t chunk) {
  constexpr uint32_t GPB = kBlock / G;
  const uint32_t lane = threadIdx.x % G;
  const uint32_t grp = threadIdx.x / G;
  const uint64_t r0 = (uint64_t)logical_block() * chunk;
  const uint64_t r1 = r0 + chunk < rows ? r0 + chunk : rows;
  for (uint64_t base = r0; base < r1; base += GPB) {
    const uint32_t row = (uint32_t)(base + grp);
    if (base + grp >= r1) continue;
    const float rr = r[row];
    if (rr == 0.f) continue;
    const uint32_t beg = rowptr[row], end = rowptr[row + 1];
    for (uint32_t k = beg + lane; k < end; k += G) {
      const u32x2 e = ld_stream(csr + k);
      unsafeAtomicAdd(&u[e.x], __uint_as_float(e.y) * rr);
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_update(uint32_t cols, const float* __restrict__ u,
                                                   const float* __restrict__ norm,
                                                   float* __restrict__ x, float* zero_u) {
  const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= cols) return;
  const float n = norm[j], uj = u[j];
  x[j] = x[j] * ((n != 0.f) ? uj / n : uj);
  if (zero_u) zero_u[j] = 0.f;
}

__global__ __launch_bounds__(kBlock) void k_correl(uint32_t rows, const float* __restrict__ f,
                                                   const int32_t* __restrict__ y,
                                                   float* __restrict__ r) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= rows) return;
  const float fi = f[i];
  r[i] = (fi != 0.f) ? (float)y[i] / fi : 0.f;
}

// Column sums from the CSC: one lane per column, ascending rows, fp32 — the
// accumulation order of calcColumnSums (seqcsr4mlem.cpp:37-49).
__global__ __launch_bounds__(kBlock) void k_colsum_csc(const uint32_t* __restrict__ colptr,
                                                       const uint2* __restrict__ csc,
                                                       float* __restrict__ norm, uint32_t cols) {
  const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= cols) return;
  float s = 0.f;
  for (uint32_t k = colptr[j]; k < colptr[j + 1]; ++k) s = s + __uint_as_float(csc[k].y);
  norm[j] = s;
}

__global__ __launch_bounds__(kBlock) void k_colsum_atomic(const uint2* __restrict__ csr,
                                                          uint64_t nnz, float* __restrict__ norm) {
  for (uint64_t k = blockIdx.x * (uint64_t)kBlock + threadIdx.x; k < nnz;
       k += (uint64_t)gridDim.x * kBlock) {
    const u32x2 e = ld_stream(csr + k);
    unsafeAtomicAdd(&norm[e.x], __uint_as_float(e.y));
  }
}

__global__ __launch_bounds__(kBlock) void k_check_columns(const uint2* __restrict__ csr,
                                                          uint64_t nnz, uint32_t cols,
                                                          unsigned* __restrict__ bad) {
  unsigned cnt = 0;
  for (uint64_t k = blockIdx.x * (uint64_t)kBlock + threadIdx.x; k < nnz;
       k += (uint64_t)gridDim.x * kBlock)
    cnt += (csr[k].x >= cols) ? 1u : 0u;
  if (cnt) atomicAdd(bad, cnt);
}

__global__ __launch_bounds__(kBlock) void k_rowptr(const uint64_t* __restrict__ rowidx,
                                                   uint32_t rows, uint32_t* __restrict__ rowptr) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i == 0) rowptr[0] = 0;
  if (i < rows) rowptr[i + 1] = (uint32_t)rowidx[i];
}

// CSC keys/values: key = column, value = {row, value bits} as one u64 (the
// CSC record layout), one G=16 lane group per row.
__global__ __launch_bounds__(kBlock) void k_csc_keys(const uint32_t* __restrict__ rowptr,
                                                     const uint2* __restrict__ csr, uint32_t rows,
                                                     uint32_t* __restrict__ keys,
                                                     unsigned long long* __restrict__ vals) {
  constexpr int G = 16;
  const uint64_t t = blockIdx.x * (uint64_t)kBlock + threadIdx.x;
  const uint64_t row = t / G;
  const uint32_t lane = (uint32_t)(t % G);
  if (row >= rows) return;
  for (uint32_t k = rowptr[row] + lane; k < rowptr[row + 1]; k += G) {
    const uint2 e = csr[k];
    keys[k] = e.x;
    vals[k] = ((unsigned long long)e.y << 32) | (unsigned long long)row;
  }
}

// colptr from the column-sorted keys: colptr[c] = first position with key >= c.
__global__ __launch_bounds__(kBlock) void k_colptr(const uint32_t* __restrict__ skeys, uint64_t nnz,
                                                   uint32_t cols, uint32_t* __restrict__ colptr) {
  const uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x;
  if (i >= nnz) return;
  const uint32_t c = skeys[i];
  if (i == 0) {
    for (uint32_t cc = 0; cc <= c; ++cc) colptr[cc] = 0u;
  } else {
    const uint32_t p = skeys[i - 1];
    if (p != c)
      for (uint32_t cc = p + 1; cc <= c; ++cc) colptr[cc] = (uint32_t)i;
  }
  if (i == nnz - 1)
    for (uint64_t cc = (uint64_t)c + 1; cc <= cols; ++cc) colptr[cc] = (uint32_t)nnz;
}

__global__ __launch_bounds__(kBlock) void k_fill_u32(uint32_t* p, uint32_t v, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kBlock)
    p[i] = v;
}

__global__ __launch_bounds__(kBlock) void k_fill(float* p, float v, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kBlock)
    p[i] = v;
}

__global__ __launch_bounds__(kBlock) void k_copy(const uint4* __restrict__ src,
                                                 uint4* __restrict__ dst, uint64_t n16) {
  for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < n16;
       i += (uint64_t)gridDim.x * kBlock) {
    const u32x4 v = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(src) + i);
    __builtin_nontemporal_store(v, reinterpret_cast<u32x4*>(dst) + i);
  }
}

// ------------------------------------------------------------- launch glue --
struct Geometry {
  uint32_t grid;
  uint32_t chunk;
};

// Blocks of 256 lanes, 256/G groups each; `n` rows or columns. Grid rounded
// up to a multiple of 8 (one share per XCD); every block owns `chunk` items.
Geometry geometry(uint32_t n, int G, int blocks_per_cu, int cus, bool cap_for_stats) {
  const uint64_t gpb = (uint64_t)(kBlock / G);
  uint64_t need = (n + gpb - 1) / gpb;
  if (need == 0) need = 1;
  uint64_t grid = need;
  int bpc = blocks_per_cu;
  if (cap_for_stats && bpc <= 0) bpc = 8;
  if (bpc > 0) grid = std::min<uint64_t>(need, (uint64_t)bpc * (uint64_t)cus);
  grid = (grid + 7) / 8 * 8;
  uint64_t chunk = (n + grid - 1) / grid;
  chunk = (chunk + gpb - 1) / gpb * gpb;
  if (chunk == 0) chunk = gpb;
  return Geometry{(uint32_t)grid, (uint32_t)chunk};
}

inline uint32_t blocks_for(uint64_t n) {
  uint64_t b = (n + kBlock - 1) / kBlock;
  return (uint32_t)std::max<uint64_t>(1, b);  // n < 2^32 + 1: at most 2^24 blocks
}

template <int G, bool EXACT>
void fwd_g(const DevMatrix& m, const float* x, const int32_t* y, float* out, bool ratio,
           const StatsState* st, hipStream_t s) {
  const bool stats = ratio && st != nullptr;
  const Geometry g = geometry(m.rows, G, m.fwd.blocks_per_cu, m.cus, stats);
  if (stats && g.grid > st->max_blocks) raise(MLEM_ESTATE, "statistics buffer too small for launch");
  const dim3 grid(g.grid), block(kBlock);
  if (stats)
    hipLaunchKernelGGL((k_fwd<G, EXACT, true, true>), grid, block, 0, s, m.rowptr, m.csr, x, y, out,
                       m.rows, g.chunk, st->partials, st->counter, st->results, st->ring);
  else if (ratio)
    hipLaunchKernelGGL((k_fwd<G, EXACT, true, false>), grid, block, 0, s, m.rowptr, m.csr, x, y,
                       out, m.rows, g.chunk, nullptr, nullptr, nullptr, 0u);
  else
    hipLaunchKernelGGL((k_fwd<G, EXACT, false, false>), grid, block, 0, s, m.rowptr, m.csr, x, y,
                       out, m.rows, g.chunk, nullptr, nullptr, nullptr, 0u);
}

template <int G, bool EXACT>
void bwd_g(const DevMatrix& m, const float* r, const float* norm, float* out, bool update,
           hipStream_t s) {
  const Geometry g = geometry(m.cols, G, m.bwd.blocks_per_cu, m.cus, false);
  const dim3 grid(g.grid), block(kBlock);
  if (update)
    hipLaunchKernelGGL((k_bwd<G, EXACT, true>), grid, block, 0, s, m.colptr, m.csc, r, norm, out,
                       m.cols, g.chunk);
  else
    hipLaunchKernelGGL((k_bwd<G, EXACT, false>), grid, block, 0, s, m.colptr, m.csc, r, norm, out,
                       m.cols, g.chunk);
}

template <int G>
void scatter_g(const DevMatrix& m, const float* r, float* u, hipStream_t s) {
  const Geometry g = geometry(m.rows, G, m.scatter.blocks_per_cu, m.cus, false);
  hipLaunchKernelGGL((k_scatter<G>), dim3(g.grid), dim3(kBlock), 0, s, m.rowptr, m.csr, r, u,
                     m.rows, g.chunk);
}

void check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) raise(MLEM_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace

unsigned stats_max_blocks(const DevMatrix& m) {
  const Geometry g = geometry(m.rows, m.fwd.lanes, m.fwd.blocks_per_cu, m.cus, true);
  return g.grid;
}

void launch_fwd(const DevMatrix& m, int mode, const float* x, const int32_t* y, float* out,
                bool ratio, const StatsState* stats, hipStream_t s) {
  if (mode == MLEM_MODE_EXACT) {
    fwd_g<1, true>(m, x, y, out, ratio, stats, s);
  } else {
    switch (m.fwd.lanes) {
      case 4: fwd_g<4, false>(m, x, y, out, ratio, stats, s); break;
      case 8: fwd_g<8, false>(m, x, y, out, ratio, stats, s); break;
      case 16: fwd_g<16, false>(m, x, y, out, ratio, stats, s); break;
      case 32: fwd_g<32, false>(m, x, y, out, ratio, stats, s); break;
      case 64: fwd_g<64, false>(m, x, y, out, ratio, stats, s); break;
      default: raise(MLEM_EINVAL, "unsupported forward lane count");
    }
  }
  check_launch("forward projection launch");
}

void launch_bwd(const DevMatrix& m, int mode, const float* r, const float* norm, float* out,
                bool update, hipStream_t s) {
  if (!m.csc) raise(MLEM_ESTATE, "matrix was opened without a CSC transpose");
  if (mode == MLEM_MODE_EXACT) {
    bwd_g<1, true>(m, r, norm, out, update, s);
  } else {
    switch (m.bwd.lanes) {
      case 4: bwd_g<4, false>(m, r, norm, out, update, s); break;
      case 8: bwd_g<8, false>(m, r, norm, out, update, s); break;
      case 16: bwd_g<16, false>(m, r, norm, out, update, s); break;
      case 32: bwd_g<32, false>(m, r, norm, out, update, s); break;
      case 64: bwd_g<64, false>(m, r, norm, out, update, s); break;
      default: raise(MLEM_EINVAL, "unsupported backward lane count");
    }
  }
  check_launch("backprojection launch");
}

void launch_scatter(const DevMatrix& m, const float* r, float* u, hipStream_t s) {
  switch (m.scatter.lanes) {
    case 4: scatter_g<4>(m, r, u, s); break;
    case 8: scatter_g<8>(m, r, u, s); break;
    case 16: scatter_g<16>(m, r, u, s); break;
    case 32: scatter_g<32>(m, r, u, s); break;
    case 64: scatter_g<64>(m, r, u, s); break;
    default: raise(MLEM_EINVAL, "unsupported scatter lane count");
  }
  check_launch("scatter launch");
}

void launch_update(uint32_t cols, const float* u, const float* norm, float* x, bool zero_u,
                   hipStream_t s) {
  hipLaunchKernelGGL(k_update, dim3(blocks_for(cols)), dim3(kBlock), 0, s, cols, u, norm, x,
                     zero_u ? const_cast<float*>(u) : nullptr);
  check_launch("update launch");
}

void launch_correl(uint32_t rows, const float* f, const int32_t* y, float* r, hipStream_t s) {
  hipLaunchKernelGGL(k_correl, dim3(blocks_for(rows)), dim3(kBlock), 0, s, rows, f, y, r);
  check_launch("correl launch");
}

void launch_colsums_csc(const DevMatrix& m, float* norm, hipStream_t s) {
  hipLaunchKernelGGL(k_colsum_csc, dim3(blocks_for(m.cols)), dim3(kBlock), 0, s, m.colptr, m.csc,
                     norm, m.cols);
  check_launch("column sums launch");
}

void launch_colsums_atomic(const DevMatrix& m, float* norm, hipStream_t s) {
  hipLaunchKernelGGL(k_fill, dim3(blocks_for(m.cols)), dim3(kBlock), 0, s, norm, 0.f,
                     (uint64_t)m.cols);
  hipLaunchKernelGGL(k_colsum_atomic, dim3(std::min<uint32_t>(blocks_for(m.nnz), 8 * m.cus)),
                     dim3(kBlock), 0, s, m.csr, m.nnz, norm);
  check_launch("column sums launch");
}

void launch_check_columns(const DevMatrix& m, unsigned* bad, hipStream_t s) {
  hipLaunchKernelGGL(k_check_columns, dim3(std::min<uint32_t>(blocks_for(m.nnz), 8 * m.cus)),
                     dim3(kBlock), 0, s, m.csr, m.nnz, m.cols, bad);
  check_launch("column check launch");
}

void launch_rowptr_from_rowidx(const uint64_t* rowidx, uint32_t rows, uint32_t* rowptr,
                               hipStream_t s) {
  hipLaunchKernelGGL(k_rowptr, dim3(blocks_for((uint64_t)rows + 1)), dim3(kBlock), 0, s, rowidx,
                     rows, rowptr);
  check_launch("row pointer launch");
}

void launch_fill(float* p, float v, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_fill, dim3(blocks_for(n)), dim3(kBlock), 0, s, p, v, (uint64_t)n);
  check_launch("fill launch");
}

void launch_copy(const uint4* src, uint4* dst, size_t n16, hipStream_t s) {
  hipLaunchKernelGGL(k_copy, dim3(blocks_for(n16)), dim3(kBlock), 0, s, src, dst, (uint64_t)n16);
  check_launch("copy launch");
}

void build_csc(DevMatrix& m, hipStream_t s) {
  const uint64_t nnz = m.nnz;
  MLEM_HIP(hipMalloc(&m.colptr, sizeof(uint32_t) * ((size_t)m.cols + 1)));
  MLEM_HIP(hipMalloc(&m.csc, sizeof(uint2) * std::max<uint64_t>(nnz, 1)));
  if (nnz == 0) {
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for((uint64_t)m.cols + 1)), dim3(kBlock), 0, s,
                       m.colptr, 0u, (uint64_t)m.cols + 1);
    check_launch("colptr fill");
    return;
  }
  uint32_t* keys = nullptr;
  uint32_t* skeys = nullptr;
  unsigned long long* vals = nullptr;
  void* temp = nullptr;
  struct Guard {
    void** ptrs[4];
    ~Guard() {
      for (void** p : ptrs)
        if (*p) (void)hipFree(*p);
    }
  } guard{{(void**)&keys, (void**)&skeys, (void**)&vals, &temp}};
  MLEM_HIP(hipMalloc(&keys, sizeof(uint32_t) * nnz));
  MLEM_HIP(hipMalloc(&skeys, sizeof(uint32_t) * nnz));
  MLEM_HIP(hipMalloc(&vals, sizeof(unsigned long long) * nnz));
  const uint64_t threads = (uint64_t)m.rows * 16;
  hipLaunchKernelGGL(k_csc_keys, dim3((uint32_t)((threads + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                     s, m.rowptr, m.csr, m.rows, keys, vals);
  check_launch("csc key launch");
  const uint32_t maxcol = m.cols > 0 ? m.cols - 1 : 0;
  const int end_bit = maxcol == 0 ? 1 : 32 - __builtin_clz(maxcol);
  size_t temp_bytes = 0;
  MLEM_HIP(rocprim::radix_sort_pairs(nullptr, temp_bytes, keys, skeys, vals,
                                     reinterpret_cast<unsigned long long*>(m.csc), (size_t)nnz, 0,
                                     end_bit, s));
  MLEM_HIP(hipMalloc(&temp, std::max<size_t>(temp_bytes, 16)));
  MLEM_HIP(rocprim::radix_sort_pairs(temp, temp_bytes, keys, skeys, vals,
                                     reinterpret_cast<unsigned long long*>(m.csc), (size_t)nnz, 0,
                                     end_bit, s));
  hipLaunchKernelGGL(k_colptr, dim3((uint32_t)((nnz + kBlock - 1) / kBlock)), dim3(kBlock), 0, s,
                     skeys, nnz, m.cols, m.colptr);
  check_launch("colptr launch");
  MLEM_HIP(hipStreamSynchronize(s));
}

}  // namespace mlem
