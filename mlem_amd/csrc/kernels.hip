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
// so that each XCD works through runs of 256 consecutive tiles (rows,
// columns), the 8 XCDs on 8 neighbouring runs at a time: a run's gathers stay
// in one XCD's L2, and the XCDs together sweep one region of the vectors, which
// the last-level cache holds for all of them. (One contiguous eighth per XCD
// was 2-4% slower on the 192x192x64 matrix; runs of 32 or 1024 tiles were
// between the two.)
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

// Block i of XCD x (b = 8i + x) takes entry i % C of run x in row i / C of
// runs; the last, partial row of runs is split evenly over the XCDs. A
// bijection of [0, g) for every g that is a multiple of 8.
__device__ __forceinline__ uint32_t logical_block() {
  constexpr uint32_t C = 256;  // tiles per run
  const uint32_t b = blockIdx.x, g = gridDim.x;  // g is a multiple of 8
  const uint32_t q = g >> 3, x = b & 7u, i = b >> 3;
  const uint32_t c = i / C, o = i % C, full = q / C;
  const uint32_t w = c < full ? C : q - full * C;
  return c * 8u * C + x * w + o;
}

// v from the lane that DPP control CTRL selects within the 16-lane row.
template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// Sum over each group of G consecutive lanes; every lane of the group gets
// the (fixed-order) sum. G = 16 and G = 8 use row DPP moves (no LDS traffic):
// swap within pairs, within quads, mirror the half-rows (G = 8 stops here),
// mirror the row.
template <int G>
__device__ __forceinline__ float group_sum(float v) {
  if constexpr (G == 16) {
    v = v + dpp<0xB1>(v);   // quad_perm [1,0,3,2]
    v = v + dpp<0x4E>(v);   // quad_perm [2,3,0,1]
    v = v + dpp<0x141>(v);  // row_half_mirror
    v = v + dpp<0x140>(v);  // row_mirror
    return v;
  } else if constexpr (G == 8) {
    v = v + dpp<0xB1>(v);   // quad_perm [1,0,3,2]
    v = v + dpp<0x4E>(v);   // quad_perm [2,3,0,1]
    v = v + dpp<0x141>(v);  // row_half_mirror
    return v;
  } else {
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, G);
    return v;
  }
}

// Dot product of one sparse row (or CSC column) with a dense vector. EXACT
// continues a running sum that starts at `init` (0 for a whole segment).
template <int G, bool EXACT>
__device__ __forceinline__ float seg_dot(const uint2* __restrict__ a, const float* __restrict__ v,
                                         uint32_t beg, uint32_t end, uint32_t lane,
                                         float init = 0.f) {
  if constexpr (EXACT) {
    // one lane, ascending element order; the records and gathers of 8
    // elements are issued together (clamped at the segment end), the products
    // then added one by one
    constexpr int U = 8;
    float acc = init;
    for (uint32_t k = beg; k < end; k += U) {
      u32x2 e[U];
      float g[U];
#pragma unroll
      for (int j = 0; j < U; ++j) e[j] = ld_stream(a + min(k + j, end - 1));
#pragma unroll
      for (int j = 0; j < U; ++j) g[j] = v[e[j].x];
#pragma unroll
      for (int j = 0; j < U; ++j)
        if (k + j < end) acc = acc + __uint_as_float(e[j].y) * g[j];
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

// Deterministic FP64 reduction of (ll, tot), first stage: every workgroup
// reduces its lanes in a fixed order and stores its partial. The second
// stage, k_stats_reduce, runs as the next kernel in the stream; the kernel
// boundary orders the partials before it, so the projection kernels need no
// device-scope fence or atomic of their own (with a fence and a counter
// atomic in every workgroup, the EXACT forward with statistics took twice
// as long as without statistics).
template <int NB = kBlock>
__device__ void stats_partial(double ll, double tot, double2* __restrict__ partials) {
  __shared__ double s_ll[NB / kWave], s_tot[NB / kWave];
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
    for (int w = 0; w < NB / kWave; ++w) {
      bl += s_ll[w];
      bt += s_tot[w];
    }
    partials[blockIdx.x] = make_double2(bl, bt);
  }
}

// Second stage, one 1024-lane workgroup: sum the n partials in a fixed order
// (lane t takes partials t, t + 1024, ... into two alternating accumulators,
// so two loads are in flight) and append (ll, tot) to the ring at slot
// counter[1] (then advance it).
constexpr int kReduceLanes = 1024;
__global__ __launch_bounds__(kReduceLanes) void k_stats_reduce(
    const double2* __restrict__ partials, unsigned n, unsigned* __restrict__ counter,
    double2* __restrict__ results, unsigned ring) {
  __shared__ double s_ll[kReduceLanes / kWave], s_tot[kReduceLanes / kWave];
  double l0 = 0.0, t0 = 0.0, l1 = 0.0, t1 = 0.0;
  unsigned i = threadIdx.x;
  for (; i + kReduceLanes < n; i += 2 * kReduceLanes) {
    const double2 p = partials[i], q = partials[i + kReduceLanes];
    l0 += p.x;
    t0 += p.y;
    l1 += q.x;
    t1 += q.y;
  }
  if (i < n) {
    const double2 p = partials[i];
    l0 += p.x;
    t0 += p.y;
  }
  double pl = l0 + l1, pt = t0 + t1;
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    pl += __shfl_xor(pl, off);
    pt += __shfl_xor(pt, off);
  }
  const int wave = threadIdx.x / kWave, wl = threadIdx.x % kWave;
  if (wl == 0) {
    s_ll[wave] = pl;
    s_tot[wave] = pt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double bl = 0.0, bt = 0.0;
    for (int w = 0; w < kReduceLanes / kWave; ++w) {
      bl += s_ll[w];
      bt += s_tot[w];
    }
    const unsigned slot = counter[1];
    results[slot % ring] = make_double2(bl, bt);
    counter[1] = slot + 1;
  }
}

// Forward epilogue: r = (f != 0) ? y/f : 0 (or f itself), optional FP64
// log-likelihood and sum of f. With reduced-precision values the row sum is
// q_i.x and the projection is s_i * (q_i.x); the ratio y_i / (q_i.x) already
// carries the factor s_i that the backprojection of q needs, so s_i only
// enters the statistics and the plain projection.
template <bool RATIO, bool STATS>
struct FwdEpilogue {
  const int32_t* __restrict__ y;
  float* __restrict__ out;
  const float* __restrict__ rscale = nullptr;
  double ll = 0.0, tot = 0.0;
  int32_t yi = 0;
  float s = 1.f;
  __device__ void prefetch(uint32_t row) {
    if constexpr (RATIO) yi = y[row];
    if (rscale && (STATS || !RATIO)) s = rscale[row];
  }
  __device__ void finish(uint32_t row, float f) {
    if constexpr (RATIO) {
      out[row] = (f != 0.f) ? div_rn((float)yi, f) : 0.f;
      if constexpr (STATS) {
        const double fd = (double)f * (double)s;
        tot += fd;
        if (fd > 0.0) ll += (yi != 0 ? (double)yi * log(fd) : 0.0) - fd;
      }
    } else {
      out[row] = rscale ? f * s : f;
    }
  }
};

// Backward epilogue: x *= (n != 0) ? u/n : u (or u itself).
template <bool UPDATE>
struct BwdEpilogue {
  const float* __restrict__ norm;
  float* __restrict__ out;
  float n = 0.f, x = 0.f;
  __device__ void prefetch(uint32_t col) {
    if constexpr (UPDATE) {
      n = norm[col];
      x = out[col];
    }
  }
  __device__ void finish(uint32_t col, float u) {
    if constexpr (UPDATE)
      out[col] = x * ((n != 0.f) ? div_rn(u, n) : u);
    else
      out[col] = u;
  }
};

// ---------------------------------------------------------------- forward --
template <int G, bool EXACT, bool RATIO, bool STATS>
__global__ __launch_bounds__(kBlock) void k_fwd(
    const uint32_t* __restrict__ rowptr, const uint2* __restrict__ csr,
    const float* __restrict__ x, const int32_t* __restrict__ y, float* __restrict__ out,
    const float* __restrict__ rscale, uint32_t rows, uint32_t chunk,
    double2* __restrict__ partials) {
  constexpr uint32_t GPB = kBlock / G;
  const uint32_t lane = threadIdx.x % G;
  const uint32_t grp = threadIdx.x / G;
  const uint64_t r0 = (uint64_t)logical_block() * chunk;
  const uint64_t r1 = r0 + chunk < rows ? r0 + chunk : rows;
  FwdEpilogue<RATIO, STATS> epi{y, out, rscale};
  for (uint64_t base = r0; base < r1; base += GPB) {
    const uint32_t row = (uint32_t)(base + grp);
    const bool active = base + grp < r1;
    float acc = 0.f;
    if (active) acc = seg_dot<G, EXACT>(csr, x, rowptr[row], rowptr[row + 1], lane);
    acc = group_sum<G>(acc);
    if (active && lane == 0) {
      epi.prefetch(row);
      epi.finish(row, acc);
    }
  }
  if constexpr (STATS) stats_partial(epi.ll, epi.tot, partials);
}

// --------------------------------------------------------------- backward --
// OUT: what a column's sum u becomes. kBwdSet: out = u; kBwdUpdate:
// out *= (n != 0) ? u/n : u; kBwdAdd: out += u, the sum of one row block's
// part of the column (EXACT continues out's running sum element by element,
// so adding the blocks in ascending row order reproduces the whole column's
// sequential sum).
constexpr int kBwdSet = 0, kBwdUpdate = 1, kBwdAdd = 2;
template <int G, bool EXACT, int OUT>
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
    if (active) {
      const float init = (EXACT && OUT == kBwdAdd) ? out[col] : 0.f;
      acc = seg_dot<G, EXACT>(csc, r, colptr[col], colptr[col + 1], lane, init);
    }
    acc = group_sum<G>(acc);
    if (active && lane == 0) {
      if constexpr (OUT == kBwdUpdate) {
        const float n = norm[col];
        out[col] = out[col] * ((n != 0.f) ? div_rn(acc, n) : acc);
      } else if constexpr (OUT == kBwdAdd && !EXACT) {
        out[col] = out[col] + acc;
      } else {
        out[col] = acc;
      }
    }
  }
}

// ------------------------------------------------------------ tiled SpMV --
// Sums of the products P over the segments [S[s], S[s+1]), s < nseg, of an
// NB-lane block: G-lane group g takes segments g, g + NB/G, ... and writes
// Sum[s]. Lanes read aligned 16-byte quads, the first and last masked to the
// segment, and the group adds its lanes' partial sums in a fixed tree. The
// next segment's bounds are read before this segment's quads. (Two or four
// quads in flight per lane measured 2-6% slower: the LDS is shared with the
// other block's product phase.)
template <int NB, int G>
__device__ __forceinline__ void segment_sums(const float* __restrict__ P,
                                             const uint32_t* __restrict__ S, uint32_t nseg,
                                             float* __restrict__ Sum, uint32_t t) {
  const float4* __restrict__ P4 = reinterpret_cast<const float4*>(P);
  const uint32_t lane = t % G;
  uint32_t sg = t / G;
  uint32_t b = S[min(sg, (uint32_t)NB)], end = S[min(sg + 1, (uint32_t)NB)];
  while (sg < nseg) {
    const uint32_t next = sg + NB / G;
    const uint32_t nb = S[min(next, (uint32_t)NB)], ne = S[min(next + 1, (uint32_t)NB)];
    float acc = 0.f;
    for (uint32_t qd = (b >> 2) + lane; 4 * qd < end; qd += G) {
      const float4 p = P4[qd];
      const uint32_t k = 4 * qd;
      acc += (k >= b) ? p.x : 0.f;
      acc += (k + 1 >= b && k + 1 < end) ? p.y : 0.f;
      acc += (k + 2 >= b && k + 2 < end) ? p.z : 0.f;
      acc += (k + 3 >= b && k + 3 < end) ? p.w : 0.f;
    }
    acc = group_sum<G>(acc);
    if (lane == 0) Sum[sg] = acc;
    sg = next;
    b = nb;
    end = ne;
  }
}

// A tile is a run of whole segments (rows, or CSC columns) holding at most
// T = 256*EPT elements and at most T segments, or one longer segment alone
// (partitioned once at load time). A 256-lane block streams the tile's
// element range with EPT independent coalesced loads per lane, issues all
// its gathers, writes the products to LDS and sums every segment there with
// 16-lane groups; the epilogue (ratio or multiplicative update) runs on the
// segment sum. All of a tile's loads are in flight at once, which is what
// keeps HBM busy: short segments no longer serialise load -> gather ->
// reduce per wave.
// Loads and gathers are unconditional (indices clamped to the tile's last
// element; the duplicates are masked or never read), so the compiler keeps
// all EPT element loads, then all EPT gathers, in flight together instead of
// wrapping each in a branch that waits for the previous one. A tile holds at
// most 256 segments: lane t owns segment t, fetches its end offset and its
// epilogue operands (Epi::prefetch) together with the element loads, and
// finishes it (Epi::finish) from the LDS sum, so no load sits on the tail.
template <int EPT, class Epi>
__device__ __forceinline__ void tile_spmv(const uint32_t* __restrict__ segptr,
                                          const uint2* __restrict__ el,
                                          const uint2* __restrict__ tiles, uint32_t tile,
                                          const float* __restrict__ v, Epi& epi) {
  constexpr uint32_t T = kBlock * EPT;
  constexpr int G = 16;
  __shared__ __align__(16) float P[T];
  __shared__ uint32_t S[kBlock + 1];
  __shared__ float Sum[kBlock];
  __shared__ float red[kBlock / kWave];
  const uint32_t t = threadIdx.x;
  const uint2 d0 = tiles[tile], d1 = tiles[tile + 1];  // {first segment, first element}
  const uint32_t s0 = d0.x, e0 = d0.y, s1 = d1.x, e1 = d1.y;
  const uint32_t nseg = s1 - s0;                        // <= kBlock
  if (t < nseg) epi.prefetch(s0 + t);
  if (e1 - e0 > T) {
    // a single segment longer than a tile: strided partial sums, block sum
    float acc = 0.f;
    for (uint32_t base = e0; base < e1; base += T) {
      u32x2 e[EPT];
      float g[EPT];
#pragma unroll
      for (int j = 0; j < EPT; ++j) e[j] = ld_stream(el + min(base + t + j * kBlock, e1 - 1));
#pragma unroll
      for (int j = 0; j < EPT; ++j) g[j] = v[e[j].x];
#pragma unroll
      for (int j = 0; j < EPT; ++j) {
        const float p = __uint_as_float(e[j].y) * g[j];
        acc += (base + t + j * kBlock < e1) ? p : 0.f;
      }
    }
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if (t % kWave == 0) red[t / kWave] = acc;
    __syncthreads();
    if (t == 0) {
      float sum = 0.f;
      for (int w = 0; w < kBlock / kWave; ++w) sum += red[w];
      epi.finish(s0, sum);
    }
    return;
  }
  uint32_t send = 0;
  if (t < nseg) send = segptr[s0 + t + 1] - e0;
  if (e1 > e0) {  // block-uniform
    u32x2 e[EPT];
    float g[EPT];
    const uint32_t last = e1 - 1;
#pragma unroll
    for (int j = 0; j < EPT; ++j) e[j] = ld_stream(el + min(e0 + t + j * kBlock, last));
#pragma unroll
    for (int j = 0; j < EPT; ++j) g[j] = v[e[j].x];
#pragma unroll
    for (int j = 0; j < EPT; ++j) P[t + j * kBlock] = __uint_as_float(e[j].y) * g[j];
  }
  if (t < nseg) S[t + 1] = send;
  if (t == 0) S[0] = 0;
  __syncthreads();
  segment_sums<kBlock, G>(P, S, nseg, Sum, t);
  __syncthreads();
  if (t < nseg) epi.finish(s0 + t, Sum[t]);
}

template <int EPT, bool RATIO, bool STATS>
__global__ __launch_bounds__(kBlock) void k_fwd_tiles(
    const uint32_t* __restrict__ rowptr, const uint2* __restrict__ csr,
    const uint2* __restrict__ tiles, uint32_t ntiles, const float* __restrict__ x,
    const int32_t* __restrict__ y, float* __restrict__ out, const float* __restrict__ rscale,
    double2* __restrict__ partials) {
  const uint32_t tile = logical_block();
  FwdEpilogue<RATIO, STATS> epi{y, out, rscale};
  if (tile < ntiles) tile_spmv<EPT>(rowptr, csr, tiles, tile, x, epi);
  if constexpr (STATS) stats_partial(epi.ll, epi.tot, partials);
}

template <int EPT, bool UPDATE>
__global__ __launch_bounds__(kBlock) void k_bwd_tiles(const uint32_t* __restrict__ colptr,
                                                      const uint2* __restrict__ csc,
                                                      const uint2* __restrict__ tiles,
                                                      uint32_t ntiles, const float* __restrict__ r,
                                                      const float* __restrict__ norm,
                                                      float* __restrict__ out) {
  const uint32_t tile = logical_block();
  if (tile >= ntiles) return;
  BwdEpilogue<UPDATE> epi{norm, out};
  tile_spmv<EPT>(colptr, csc, tiles, tile, r, epi);
}

// ------------------------------------------------------- dictionary tiles --
// The same tiles, stored for gathers that cost less than one cache-line
// request per element. At load time every tile of at most T elements gets a
// sorted dictionary of the distinct indices it gathers (voxels for the
// forward pass, rows for the backward pass) and each element is stored as
// {fp32 value, u16 position in that dictionary} (SoA, every tile starting on
// a 64-element boundary). At run time a block gathers each distinct entry
// once, with sorted (coalescing) addresses, into LDS; elements read their
// operand from LDS. Tiles of one long segment use the 8-byte records and
// direct gathers. Blocks have NB lanes (256 or 512); a tile holds at most NB
// whole segments and NB*EPT elements.

template <int NB, int EPT>
__device__ __forceinline__ float long_segment_sum(const uint2* __restrict__ el,
                                                  const float* __restrict__ v, uint32_t e0,
                                                  uint32_t e1, float* red) {
  constexpr uint32_t T = NB * EPT;
  const uint32_t t = threadIdx.x;
  float acc = 0.f;
  for (uint32_t base = e0; base < e1; base += T) {
    u32x2 e[EPT];
    float g[EPT];
#pragma unroll
    for (int j = 0; j < EPT; ++j) e[j] = ld_stream(el + min(base + t + j * NB, e1 - 1));
#pragma unroll
    for (int j = 0; j < EPT; ++j) g[j] = v[e[j].x];
#pragma unroll
    for (int j = 0; j < EPT; ++j) {
      const float p = __uint_as_float(e[j].y) * g[j];
      acc += (base + t + j * NB < e1) ? p : 0.f;
    }
  }
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  if (t % kWave == 0) red[t / kWave] = acc;
  __syncthreads();
  float sum = 0.f;
  for (int w = 0; w < NB / kWave; ++w) sum += red[w];
  return sum;
}

// Elements of a dictionary tile are loaded NP at a time (NP = 4 or 8): NP
// fp32 words, or NP 16-bit values (fp16 / bf16) packed two to a word, and NP
// 16-bit dictionary positions. The 16-bit value forms widen exactly to fp32.
template <int NW>
struct Words;
template <>
struct Words<1> {
  using T = uint32_t;
  __device__ static uint32_t at(T w, int) { return w; }
};
template <>
struct Words<2> {
  using T = u32x2;
  __device__ static uint32_t at(T w, int i) { return w[i]; }
};
template <>
struct Words<4> {
  using T = u32x4;
  __device__ static uint32_t at(T w, int i) { return w[i]; }
};

template <int VK, int NP>
struct ValPack {
  using WS = Words<VK == kF32 ? NP : NP / 2>;
  using W = typename WS::T;
  __device__ static float get(W w, int q) {
    if constexpr (VK == kF32) {
      return __uint_as_float(WS::at(w, q));
    } else {
      const uint32_t h = (WS::at(w, q >> 1) >> (16 * (q & 1))) & 0xFFFFu;
      if constexpr (VK == kF16)
        return (float)__builtin_bit_cast(_Float16, (uint16_t)h);
      else
        return __uint_as_float(h << 16);
    }
  }
};

template <int NP>
struct PosPack {
  using WS = Words<NP / 2>;
  using W = typename WS::T;
  __device__ static uint32_t get(W w, int q) { return (WS::at(w, q >> 1) >> (16 * (q & 1))) & 0xFFFFu; }
};

// Elements per stream load: fp32 values take 16-byte value loads and 8-byte
// position loads (1-3% faster than pairs); 16-bit values take 8 elements per
// load, 16 bytes of values and 16 of positions (on the 192x192x64 matrix the
// fp16 forward 3% and the backward 8% faster than 4 per load).
template <int VK>
constexpr int dict_pack() { return VK == kF32 ? 4 : 8; }

// Raw buffer resource over [base, base + bytes): 32-bit byte offsets held in
// one VGPR (no 64-bit address per load), loads at or past `bytes` return 0.
// Word 3 is the gfx9-family raw-buffer configuration.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes,
                                           0x00020000);
}
constexpr int kCacheNT = 2;  // cache-policy bit: non-temporal (streamed once)
// Dictionary tiles of at least this many elements store the products over the
// gathered dictionary values in LDS (half the LDS, one more barrier).
constexpr uint32_t kDictAliasMin = 6144;

template <class W, int AUX>
struct BufferLoad;
template <int AUX>
struct BufferLoad<uint32_t, AUX> {
  __device__ static uint32_t load(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(uint32_t, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, AUX));
  }
};
template <int AUX>
struct BufferLoad<u32x2, AUX> {
  __device__ static u32x2 load(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(u32x2, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, AUX));
  }
};
template <int AUX>
struct BufferLoad<u32x4, AUX> {
  __device__ static u32x4 load(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(u32x4, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, AUX));
  }
};

// The dictionary of one tile as stored: round j (one entry per lane of the
// block) is one 64-word slot per wave; slot s of the tile holds 64 32-bit indices, or,
// when meta[s] has bit 31 clear, 64 16-bit offsets from meta[s] in its first
// 32 words. Entries past the dictionary resolve to a valid index.
struct DictSlots {
  __amdgpu_buffer_rsrc_t words;  // the tile's slot words
  const uint32_t* meta;          // the tile's slot descriptors
  uint32_t wave, lane;           // wave index (uniform) and lane within it
};

// Loads of one dictionary tile: first the H element packs (values and
// dictionary positions) that stay in registers, then R rounds of NB
// dictionary indices, then the R gathers they address; the gathered values
// go to V last, so the element stream is in flight while the gathers
// complete. (Issuing the indices before the stream, so that the gathers need
// not wait for it, measured 4-15% slower: more registers, fewer waves.)
template <int NB, int R, int H, class W, class L>
__device__ __forceinline__ void load_dict_tile(const DictSlots& ds, __amdgpu_buffer_rsrc_t rv,
                                               __amdgpu_buffer_rsrc_t rval,
                                               __amdgpu_buffer_rsrc_t rlix, uint32_t t, float* V,
                                               W (&val)[H], L (&li)[H]) {
#pragma unroll
  for (int j = 0; j < H; ++j) {
    val[j] = BufferLoad<W, kCacheNT>::load(rval, (t + j * NB) * (uint32_t)sizeof(W));
    li[j] = BufferLoad<L, kCacheNT>::load(rlix, (t + j * NB) * (uint32_t)sizeof(L));
  }
  __builtin_amdgcn_sched_barrier(0);
  uint32_t ix[R];
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const uint32_t m = ds.meta[(NB / kWave) * j + ds.wave];
    const uint32_t w = NB * j + kWave * ds.wave + ((m >> 31) ? ds.lane : ds.lane >> 1);
    ix[j] = BufferLoad<uint32_t, kCacheNT>::load(ds.words, w * 4u);  // read once per pass
  }
  __builtin_amdgcn_sched_barrier(0);  // one wait between the index and the gather loads
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const uint32_t m = ds.meta[(NB / kWave) * j + ds.wave];
    if (!(m >> 31)) ix[j] = m + ((ix[j] >> (16 * (ds.lane & 1))) & 0xFFFFu);
  }
  float g[R];
#pragma unroll
  for (int j = 0; j < R; ++j) g[j] = __uint_as_float(BufferLoad<uint32_t, 0>::load(rv, ix[j] * 4u));
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int j = 0; j < R; ++j) V[t + j * NB] = g[j];
}

// load_dict_tile with the smallest multiple of 4 rounds >= `rounds` (at most
// EPT): a chain of block-uniform branches over fully unrolled variants.
template <int NB, int R, int EPT, int H, class W, class L>
__device__ __forceinline__ void load_dict_rounds(uint32_t rounds, const DictSlots& ds,
                                                 __amdgpu_buffer_rsrc_t rv,
                                                 __amdgpu_buffer_rsrc_t rval,
                                                 __amdgpu_buffer_rsrc_t rlix, uint32_t t, float* V,
                                                 W (&val)[H], L (&li)[H]) {
  if constexpr (R >= EPT) {
    load_dict_tile<NB, EPT>(ds, rv, rval, rlix, t, V, val, li);
  } else {
    if (rounds <= (uint32_t)R)
      load_dict_tile<NB, R>(ds, rv, rval, rlix, t, V, val, li);
    else
      load_dict_rounds<NB, R + 4, EPT>(rounds, ds, rv, rval, rlix, t, V, val, li);
  }
}

// EXACT: lane t adds segment t's products (each one fp32-rounded) one by one
// in element order — the reference's sequential sum — instead of the 8-lane
// tree; a segment longer than a tile is walked by one lane.
template <int NB, int EPT, int VK, bool EXACT, class Epi>
__device__ __forceinline__ void dict_spmv(const uint32_t* __restrict__ segptr,
                                          const uint2* __restrict__ el,
                                          const uint4* __restrict__ tiles, uint32_t tile,
                                          const void* __restrict__ dval,
                                          const uint16_t* __restrict__ dlix,
                                          const uint32_t* __restrict__ dict,
                                          const uint32_t* __restrict__ dmeta,
                                          const float* __restrict__ v, uint32_t nv, Epi& epi) {
  constexpr int NP = dict_pack<VK>();
  static_assert(EPT % NP == 0, "elements are loaded in packs of NP");
  using VP = ValPack<VK, NP>;
  using LP = PosPack<NP>;
  using W = typename VP::W;
  using L = typename LP::W;
  using F = float __attribute__((ext_vector_type(NP)));
  constexpr uint32_t T = NB * EPT;
  constexpr int H = EPT / NP;
  constexpr int G = 8;  // lanes per segment sum (pet rows of ~127 elements: 2-5% faster than 16)
  // Large tiles store the products over the dictionary values; smaller tiles
  // keep two buffers.
  constexpr bool kAlias = T >= kDictAliasMin;
  __shared__ __align__(32) float V[T];                  // gathered dictionary values
  __shared__ __align__(32) float Pbuf[kAlias ? 2 : T];  // element products, tile order
  float* P = kAlias ? V : Pbuf;
  __shared__ uint32_t S[NB + 1];
  __shared__ float Sum[NB];
  __shared__ float red[NB / kWave];
  const uint32_t t = threadIdx.x;
  const uint4 d0 = tiles[tile], d1 = tiles[tile + 1];
  const uint32_t s0 = d0.x, e0 = d0.y, a0 = d0.z, g0 = d0.w;
  const uint32_t nseg = d1.x - s0, count = d1.y - e0, nwords = d1.w - g0;
  if (t < nseg) epi.prefetch(s0 + t);
  if (count > T) {
    if constexpr (EXACT) {
      if (t == 0) epi.finish(s0, seg_dot<1, true>(el, v, e0, d1.y, 0));
    } else {
      const float sum = long_segment_sum<NB, EPT>(el, v, e0, d1.y, red);
      if (t == 0) epi.finish(s0, sum);
    }
    return;
  }
  uint32_t send = 0;
  if (t < nseg) send = segptr[s0 + t + 1] - e0;
  W val[H];
  L li[H];
  if (count > 0) {  // block-uniform
    constexpr uint32_t vb = VK == kF32 ? 4 : 2;
    const auto rval = buffer_rsrc(static_cast<const char*>(dval) + (size_t)a0 * vb, T * vb);
    const auto rlix = buffer_rsrc(dlix + a0, T * 2);
    // Rounds past the tile's own read words past its buffer (zeros) under
    // the next tile's slot descriptors (the descriptor array is padded by
    // EPT rounds of zeros), which resolve to valid indices as well.
    const DictSlots ds{buffer_rsrc(dict + g0, nwords * 4), dmeta + (g0 >> 6),
                       (uint32_t)__builtin_amdgcn_readfirstlane(t / kWave), t % kWave};
    const auto rv = buffer_rsrc(v, nv * 4);
    // Gather only as many NB-entry rounds as the dictionary needs, rounded
    // up to a multiple of 4 (block-uniform branch; each variant issues all
    // its loads before the first wait).
    const uint32_t rounds = __builtin_amdgcn_readfirstlane(nwords / NB);
    load_dict_rounds<NB, 4, EPT>(rounds, ds, rv, rval, rlix, t, V, val, li);
  }
  if (t < nseg) S[t + 1] = send;
  if (t == 0) S[0] = 0;
  __syncthreads();
  // Products of pack j of this lane are elements NP*(t + j*NB) + q, stored as
  // one NP*4-byte run per pack: no bank conflicts for fp32; two lanes per bank
  // group for 16-bit packs (a layout that swaps pack halves in every other
  // 8 lanes removes the conflicts but measured 2% slower: its reads need
  // the swapped indices).
  if constexpr (kAlias) {
    F pr[H];
    if (count > 0) {
#pragma unroll
      for (int j = 0; j < H; ++j)
#pragma unroll
        for (int q = 0; q < NP; ++q) pr[j][q] = VP::get(val[j], q) * V[LP::get(li[j], q)];
    }
    __syncthreads();  // every read of V is done before products overwrite it
    if (count > 0) {
#pragma unroll
      for (int j = 0; j < H; ++j)  // one NP*4-byte store per pack (see below)
        reinterpret_cast<F*>(P)[t + j * NB] = pr[j];
    }
  } else if (count > 0) {
#pragma unroll
    for (int j = 0; j < H; ++j) {
      F pr;
#pragma unroll
      for (int q = 0; q < NP; ++q) pr[q] = VP::get(val[j], q) * V[LP::get(li[j], q)];
      reinterpret_cast<F*>(P)[t + j * NB] = pr;
    }
  }
  __syncthreads();
  if constexpr (EXACT) {
    if (t < nseg) {
      const uint32_t b = S[t], end = S[t + 1];
      float acc = 0.f;
      uint32_t k = b;
      for (; k + 8 <= end; k += 8) {  // 8 reads issued, then 8 ordered adds
        float p[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) p[j] = P[k + j];
#pragma unroll
        for (int j = 0; j < 8; ++j) acc = acc + p[j];
      }
      for (; k < end; ++k) acc = acc + P[k];
      epi.finish(s0 + t, acc);
    }
    return;
  }
  segment_sums<NB, G>(P, S, nseg, Sum, t);
  __syncthreads();
  if (t < nseg) epi.finish(s0 + t, Sum[t]);
}

template <int NB, int EPT, int VK, bool EXACT, bool RATIO, bool STATS>
__global__ __launch_bounds__(NB) void k_fwd_dict(
    const uint32_t* __restrict__ rowptr, const uint2* __restrict__ csr,
    const uint4* __restrict__ tiles, uint32_t ntiles, const void* __restrict__ dval,
    const uint16_t* __restrict__ dlix, const uint32_t* __restrict__ dict,
    const uint32_t* __restrict__ dmeta, const float* __restrict__ x, uint32_t cols,
    const int32_t* __restrict__ y,
    float* __restrict__ out, const float* __restrict__ rscale, double2* __restrict__ partials) {
  const uint32_t tile = logical_block();
  FwdEpilogue<RATIO, STATS> epi{y, out, rscale};
  if (tile < ntiles)
    dict_spmv<NB, EPT, VK, EXACT>(rowptr, csr, tiles, tile, dval, dlix, dict, dmeta, x, cols, epi);
  if constexpr (STATS) stats_partial<NB>(epi.ll, epi.tot, partials);
}

template <int NB, int EPT, int VK, bool EXACT, bool UPDATE>
__global__ __launch_bounds__(NB) void k_bwd_dict(
    const uint32_t* __restrict__ colptr, const uint2* __restrict__ csc,
    const uint4* __restrict__ tiles, uint32_t ntiles, const void* __restrict__ dval,
    const uint16_t* __restrict__ dlix, const uint32_t* __restrict__ dict,
    const uint32_t* __restrict__ dmeta, const float* __restrict__ r, uint32_t rows,
    const float* __restrict__ norm,
    float* __restrict__ out) {
  const uint32_t tile = logical_block();
  if (tile >= ntiles) return;
  BwdEpilogue<UPDATE> epi{norm, out};
  dict_spmv<NB, EPT, VK, EXACT>(colptr, csc, tiles, tile, dval, dlix, dict, dmeta, r, rows, epi);
}

// Packs every 64-entry dictionary slot in place, one wave per slot: when the
// slot's ascending indices span less than 2^16 they become 16-bit offsets
// from the first (two per word, in the slot's first 32 words); meta[slot] =
// first index, with bit 31 set when the slot keeps its 32-bit indices.
// Unused entries at a tile's end hold 0 (below the first index) and become
// offset 0.
__global__ __launch_bounds__(kBlock) void k_dict_pack(uint32_t* __restrict__ dict,
                                                      uint32_t* __restrict__ meta,
                                                      uint64_t nslots) {
  const uint64_t s = (uint64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
  if (s >= nslots) return;  // whole waves
  const uint32_t l = threadIdx.x % kWave;
  const uint32_t v = dict[s * kWave + l];
  const uint32_t base = __shfl(v, 0);
  uint32_t hi = v;
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) hi = max(hi, (uint32_t)__shfl_xor(hi, off));
  const bool wide = hi - base > 0xFFFFu;
  if (!wide) {
    const uint32_t o = v >= base ? v - base : 0u;
    const uint32_t o1 = __shfl_down(o, 1);
    if ((l & 1) == 0) dict[s * kWave + (l >> 1)] = o | (o1 << 16);
  }
  if (l == 0) meta[s] = base | (wide ? 0x80000000u : 0u);
}

constexpr uint32_t pow2_ceil(uint32_t n) {
  uint32_t p = 1;
  while (p < n) p <<= 1;
  return p;
}

// Dictionary construction, one block per tile. Pass 1 (tiles4 == nullptr)
// counts the distinct indices of every tile; pass 2 writes the dictionary
// and the {value, position} element arrays. Keys (index << 16 | element
// position) are bitonic-sorted in LDS; a block scan of "new index" flags
// gives every element its dictionary position.
template <int EPT, int VK>
__global__ __launch_bounds__(kBlock) void k_dict_build(const uint2* __restrict__ el,
                                                       const uint2* __restrict__ tiles2,
                                                       const uint4* __restrict__ tiles4,
                                                       uint32_t* __restrict__ nd_out,
                                                       uint32_t* __restrict__ dict,
                                                       void* __restrict__ dval,
                                                       uint16_t* __restrict__ dlix) {
  constexpr uint32_t T = kBlock * EPT;
  constexpr uint32_t TP = pow2_ceil(T);  // bitonic sort length
  __shared__ unsigned long long K[TP];
  __shared__ uint32_t cnt[kBlock + 1];
  const uint32_t tile = blockIdx.x, t = threadIdx.x;
  uint32_t e0, e1, a0 = 0, g0 = 0;
  if (tiles4) {
    const uint4 a = tiles4[tile], b = tiles4[tile + 1];
    e0 = a.y;
    e1 = b.y;
    a0 = a.z;
    g0 = a.w;
  } else {
    e0 = tiles2[tile].y;
    e1 = tiles2[tile + 1].y;
  }
  const uint32_t count = e1 - e0;
  if (count > T) {  // one long segment: no dictionary
    if (!tiles4 && t == 0) nd_out[tile] = 0;
    return;
  }
  for (uint32_t i = t; i < TP; i += kBlock)
    K[i] = i < count ? (((unsigned long long)el[e0 + i].x << 16) | i) : ~0ull;
  __syncthreads();
  for (uint32_t k = 2; k <= TP; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t i = t; i < TP; i += kBlock) {
        const uint32_t ixj = i ^ j;
        if (ixj > i) {
          const unsigned long long a = K[i], b = K[ixj];
          const bool up = (i & k) == 0;
          if ((a > b) == up) {
            K[i] = b;
            K[ixj] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  // thread t owns sorted entries [t*EPT, t*EPT + EPT)
  const uint32_t q0 = t * EPT;
  uint32_t mine = 0;
  for (uint32_t q = q0; q < q0 + EPT && q < count; ++q)
    mine += (q == 0 || (K[q] >> 16) != (K[q - 1] >> 16)) ? 1u : 0u;
  cnt[t] = mine;
  __syncthreads();
  if (t == 0) {
    uint32_t acc = 0;
    for (int i = 0; i < kBlock; ++i) {
      const uint32_t c = cnt[i];
      cnt[i] = acc;
      acc += c;
    }
    cnt[kBlock] = acc;
  }
  __syncthreads();
  if (!tiles4) {
    if (t == 0) nd_out[tile] = cnt[kBlock];
    return;
  }
  uint32_t rank = cnt[t];
  for (uint32_t q = q0; q < q0 + EPT && q < count; ++q) {
    const uint32_t idx = (uint32_t)(K[q] >> 16), pos = (uint32_t)(K[q] & 0xFFFFu);
    if (q == 0 || (K[q] >> 16) != (K[q - 1] >> 16)) dict[g0 + rank++] = idx;
    dlix[a0 + pos] = (uint16_t)(rank - 1);
    const uint32_t bits = el[e0 + pos].y;  // fp32 bits; 16-bit kinds hold exact values
    if constexpr (VK == kF32)
      static_cast<uint32_t*>(dval)[a0 + pos] = bits;
    else if constexpr (VK == kF16)
      static_cast<uint16_t*>(dval)[a0 + pos] =
          __builtin_bit_cast(uint16_t, (_Float16)__uint_as_float(bits));
    else
      static_cast<uint16_t*>(dval)[a0 + pos] = (uint16_t)(bits >> 16);
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
  x[j] = x[j] * ((n != 0.f) ? div_rn(uj, n) : uj);
  if (zero_u) zero_u[j] = 0.f;
}

__global__ __launch_bounds__(kBlock) void k_correl(uint32_t rows, const float* __restrict__ f,
                                                   const int32_t* __restrict__ y,
                                                   float* __restrict__ r) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= rows) return;
  const float fi = f[i];
  r[i] = (fi != 0.f) ? div_rn((float)y[i], fi) : 0.f;
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

// ------------------------------------------------- reduced-precision values --
// Per-row scale s_i = max_k |a_ik| (1 for an all-zero row) and in-place
// quantization of the row's values to q_ik = round(a_ik / s_i) in fp16
// (round to nearest even, subnormals kept) or bf16 (round to nearest even on
// the fp32 bits). The fp32 record keeps q_ik, which the 16-bit type holds
// exactly. One 16-lane group per row, file row order.
template <int VK>
__device__ __forceinline__ float quantize_value(float v) {
  if constexpr (VK == kF16) {
    return (float)(_Float16)v;
  } else {
    const uint32_t u = __float_as_uint(v);
    return __uint_as_float((u + 0x7FFFu + ((u >> 16) & 1u)) & 0xFFFF0000u);
  }
}

template <int VK>
__global__ __launch_bounds__(kBlock) void k_quantize_rows(const uint32_t* __restrict__ rowptr,
                                                          uint2* __restrict__ csr, uint32_t rows,
                                                          float* __restrict__ rscale) {
  constexpr int G = 16;
  const uint64_t t = blockIdx.x * (uint64_t)kBlock + threadIdx.x;
  const uint64_t row = t / G;
  const uint32_t lane = (uint32_t)(t % G);
  const bool live = row < rows;  // the whole 16-lane group agrees
  const uint32_t b = live ? rowptr[row] : 0u, e = live ? rowptr[row + 1] : 0u;
  float mx = 0.f;
  for (uint32_t k = b + lane; k < e; k += G) mx = fmaxf(mx, fabsf(__uint_as_float(csr[k].y)));
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, G));
  if (!live) return;
  const float s = (mx > 0.f) ? mx : 1.f;
  for (uint32_t k = b + lane; k < e; k += G)
    csr[k].y = __float_as_uint(quantize_value<VK>(div_rn(__uint_as_float(csr[k].y), s)));
  if (lane == 0) rscale[row] = s;
}

// norm_j = (float) sum_i (double) s_i * (double) q_ij, ascending file row
// order (the CSC order); row indices are device rows, as is rscale.
__global__ __launch_bounds__(kBlock) void k_colsum_scaled(const uint32_t* __restrict__ colptr,
                                                          const uint2* __restrict__ csc,
                                                          const float* __restrict__ rscale,
                                                          float* __restrict__ norm, uint32_t cols) {
  const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= cols) return;
  double s = 0.0;
  for (uint32_t k = colptr[j]; k < colptr[j + 1]; ++k) {
    const uint2 e = csc[k];
    s += (double)rscale[e.x] * (double)__uint_as_float(e.y);
  }
  norm[j] = (float)s;
}

// r_i *= s_i (a backprojection of q needs s_i r_i).
__global__ __launch_bounds__(kBlock) void k_scale_rows(float* __restrict__ r,
                                                       const float* __restrict__ rscale,
                                                       uint32_t rows) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i < rows) r[i] = r[i] * rscale[i];
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

// Second stage of a statistics-carrying forward launch of `blocks`
// workgroups: one workgroup appends the launch's (ll, tot) to the ring.
void stats_reduce(const StatsState* st, unsigned blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_stats_reduce, dim3(1), dim3(kReduceLanes), 0, s, st->partials, blocks,
                     st->counter, st->results, st->ring);
}

template <int G, bool EXACT>
void fwd_g(const DevMatrix& m, const float* x, const int32_t* y, float* out, bool ratio,
           const StatsState* st, hipStream_t s) {
  const bool stats = ratio && st != nullptr;
  const Geometry g = geometry(m.rows, G, m.fwd.blocks_per_cu, m.cus, stats);
  if (stats && g.grid > st->max_blocks) raise(MLEM_ESTATE, "statistics buffer too small for launch");
  const dim3 grid(g.grid), block(kBlock);
  if (stats) {
    hipLaunchKernelGGL((k_fwd<G, EXACT, true, true>), grid, block, 0, s, m.rowptr, m.csr, x, y, out,
                       m.rscale, m.rows, g.chunk, st->partials);
    stats_reduce(st, grid.x, s);
  } else if (ratio)
    hipLaunchKernelGGL((k_fwd<G, EXACT, true, false>), grid, block, 0, s, m.rowptr, m.csr, x, y,
                       out, m.rscale, m.rows, g.chunk, nullptr);
  else
    hipLaunchKernelGGL((k_fwd<G, EXACT, false, false>), grid, block, 0, s, m.rowptr, m.csr, x, y,
                       out, m.rscale, m.rows, g.chunk, nullptr);
}

template <int G, bool EXACT>
void bwd_g(const DevMatrix& m, const float* r, const float* norm, float* out, bool update,
           hipStream_t s) {
  const Geometry g = geometry(m.cols, G, m.bwd.blocks_per_cu, m.cus, false);
  const dim3 grid(g.grid), block(kBlock);
  if (update)
    hipLaunchKernelGGL((k_bwd<G, EXACT, kBwdUpdate>), grid, block, 0, s, m.colptr, m.csc, r, norm,
                       out, m.cols, g.chunk);
  else
    hipLaunchKernelGGL((k_bwd<G, EXACT, kBwdSet>), grid, block, 0, s, m.colptr, m.csc, r, norm,
                       out, m.cols, g.chunk);
}

// One row block of a streamed matrix: its CSR (block-local row pointers,
// global columns; y and out start at the block's first row) and its CSC
// (global rows), one lane group per row or column.
template <int G, bool EXACT>
void fwd_block_g(const uint32_t* rowptr, const uint2* csr, uint32_t rows, const float* x,
                 const int32_t* y, float* out, bool ratio, const StatsState* st, int cus,
                 hipStream_t s) {
  const bool stats = ratio && st != nullptr;
  const Geometry g = geometry(rows, G, 0, cus, stats);
  if (stats && g.grid > st->max_blocks) raise(MLEM_ESTATE, "statistics buffer too small for launch");
  const dim3 grid(g.grid), block(kBlock);
  if (stats) {
    hipLaunchKernelGGL((k_fwd<G, EXACT, true, true>), grid, block, 0, s, rowptr, csr, x, y, out,
                       (const float*)nullptr, rows, g.chunk, st->partials);
    stats_reduce(st, grid.x, s);
  } else if (ratio)
    hipLaunchKernelGGL((k_fwd<G, EXACT, true, false>), grid, block, 0, s, rowptr, csr, x, y, out,
                       (const float*)nullptr, rows, g.chunk, nullptr);
  else
    hipLaunchKernelGGL((k_fwd<G, EXACT, false, false>), grid, block, 0, s, rowptr, csr, x, y, out,
                       (const float*)nullptr, rows, g.chunk, nullptr);
}

template <int G, bool EXACT>
void bwd_block_add_g(const uint32_t* colptr, const uint2* csc, uint32_t cols, const float* r,
                     float* u, int cus, hipStream_t s) {
  const Geometry g = geometry(cols, G, 0, cus, false);
  hipLaunchKernelGGL((k_bwd<G, EXACT, kBwdAdd>), dim3(g.grid), dim3(kBlock), 0, s, colptr, csc, r,
                     (const float*)nullptr, u, cols, g.chunk);
}

template <int G>
void scatter_block_g(const uint32_t* rowptr, const uint2* csr, uint32_t rows, const float* r,
                     float* u, int cus, hipStream_t s) {
  const Geometry g = geometry(rows, G, 0, cus, false);
  hipLaunchKernelGGL((k_scatter<G>), dim3(g.grid), dim3(kBlock), 0, s, rowptr, csr, r, u, rows,
                     g.chunk);
}

template <int G>
void scatter_g(const DevMatrix& m, const float* r, float* u, hipStream_t s) {
  const Geometry g = geometry(m.rows, G, m.scatter.blocks_per_cu, m.cus, false);
  hipLaunchKernelGGL((k_scatter<G>), dim3(g.grid), dim3(kBlock), 0, s, m.rowptr, m.csr, r, u,
                     m.rows, g.chunk);
}

uint32_t tile_grid(uint32_t ntiles) { return (std::max<uint32_t>(ntiles, 1) + 7) / 8 * 8; }

template <int EPT>
void fwd_t(const DevMatrix& m, const float* x, const int32_t* y, float* out, bool ratio,
           const StatsState* st, hipStream_t s) {
  if (m.tiles_fwd_ept != EPT || !m.tiles_fwd) raise(MLEM_ESTATE, "forward tiles not prepared");
  const bool stats = ratio && st != nullptr;
  const dim3 grid(tile_grid(m.ntiles_fwd)), block(kBlock);
  if (stats && grid.x > st->max_blocks) raise(MLEM_ESTATE, "statistics buffer too small for launch");
  if (stats) {
    hipLaunchKernelGGL((k_fwd_tiles<EPT, true, true>), grid, block, 0, s, m.rowptr, m.csr,
                       m.tiles_fwd, m.ntiles_fwd, x, y, out, m.rscale, st->partials);
    stats_reduce(st, grid.x, s);
  } else if (ratio)
    hipLaunchKernelGGL((k_fwd_tiles<EPT, true, false>), grid, block, 0, s, m.rowptr, m.csr,
                       m.tiles_fwd, m.ntiles_fwd, x, y, out, m.rscale, nullptr);
  else
    hipLaunchKernelGGL((k_fwd_tiles<EPT, false, false>), grid, block, 0, s, m.rowptr, m.csr,
                       m.tiles_fwd, m.ntiles_fwd, x, y, out, m.rscale, nullptr);
}

template <int EPT>
void bwd_t(const DevMatrix& m, const float* r, const float* norm, float* out, bool update,
           hipStream_t s) {
  if (m.tiles_bwd_ept != EPT || !m.tiles_bwd) raise(MLEM_ESTATE, "backward tiles not prepared");
  const dim3 grid(tile_grid(m.ntiles_bwd)), block(kBlock);
  if (update)
    hipLaunchKernelGGL((k_bwd_tiles<EPT, true>), grid, block, 0, s, m.colptr, m.csc, m.tiles_bwd,
                       m.ntiles_bwd, r, norm, out);
  else
    hipLaunchKernelGGL((k_bwd_tiles<EPT, false>), grid, block, 0, s, m.colptr, m.csc, m.tiles_bwd,
                       m.ntiles_bwd, r, norm, out);
}

void check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) raise(MLEM_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

template <int NB, int EPT, int VK, bool EXACT>
void fwd_dvx(const DevMatrix& m, const float* x, const int32_t* y, float* out, bool ratio,
             const StatsState* st, hipStream_t s) {
  const DictFormat& D = m.dfwd;
  const bool stats = ratio && st != nullptr;
  const dim3 grid(tile_grid(D.ntiles)), block(NB);
  if (stats && grid.x > st->max_blocks) raise(MLEM_ESTATE, "statistics buffer too small for launch");
  if (stats) {
    hipLaunchKernelGGL((k_fwd_dict<NB, EPT, VK, EXACT, true, true>), grid, block, 0, s, m.rowptr,
                       m.csr, D.tiles, D.ntiles, D.dval, D.dlix, D.dict, D.dmeta, x, m.cols, y, out,
                       m.rscale, st->partials);
    stats_reduce(st, grid.x, s);
  } else if (ratio)
    hipLaunchKernelGGL((k_fwd_dict<NB, EPT, VK, EXACT, true, false>), grid, block, 0, s, m.rowptr,
                       m.csr, D.tiles, D.ntiles, D.dval, D.dlix, D.dict, D.dmeta, x, m.cols, y, out,
                       m.rscale, nullptr);
  else
    hipLaunchKernelGGL((k_fwd_dict<NB, EPT, VK, EXACT, false, false>), grid, block, 0, s, m.rowptr,
                       m.csr, D.tiles, D.ntiles, D.dval, D.dlix, D.dict, D.dmeta, x, m.cols, y, out,
                       m.rscale, nullptr);
}

template <int NB, int EPT, int VK>
void fwd_dv(const DevMatrix& m, const float* x, const int32_t* y, float* out, bool ratio,
            bool exact, const StatsState* st, hipStream_t s) {
  if (exact) return fwd_dvx<NB, EPT, VK, true>(m, x, y, out, ratio, st, s);
  fwd_dvx<NB, EPT, VK, false>(m, x, y, out, ratio, st, s);
}

template <int NB, int EPT>
void fwd_d(const DevMatrix& m, const float* x, const int32_t* y, float* out, bool ratio,
           bool exact, const StatsState* st, hipStream_t s) {
  const DictFormat& D = m.dfwd;
  if (D.nb != NB || D.ept != EPT || !D.tiles)
    raise(MLEM_ESTATE, "forward dictionary not prepared");
  if (m.cols >= (1u << 30)) raise(MLEM_EINVAL, "dictionary kernels need cols < 2^30");
  switch (D.vk) {
    case kF32: fwd_dv<NB, EPT, kF32>(m, x, y, out, ratio, exact, st, s); break;
    case kF16: fwd_dv<NB, EPT, kF16>(m, x, y, out, ratio, exact, st, s); break;
    default: fwd_dv<NB, EPT, kBF16>(m, x, y, out, ratio, exact, st, s); break;
  }
}

template <int NB, int EPT, int VK, bool EXACT>
void bwd_dvx(const DevMatrix& m, const float* r, const float* norm, float* out, bool update,
             hipStream_t s) {
  const DictFormat& D = m.dbwd;
  const dim3 grid(tile_grid(D.ntiles)), block(NB);
  if (update)
    hipLaunchKernelGGL((k_bwd_dict<NB, EPT, VK, EXACT, true>), grid, block, 0, s, m.colptr, m.csc,
                       D.tiles, D.ntiles, D.dval, D.dlix, D.dict, D.dmeta, r, m.rows, norm, out);
  else
    hipLaunchKernelGGL((k_bwd_dict<NB, EPT, VK, EXACT, false>), grid, block, 0, s, m.colptr, m.csc,
                       D.tiles, D.ntiles, D.dval, D.dlix, D.dict, D.dmeta, r, m.rows, norm, out);
}

template <int NB, int EPT, int VK>
void bwd_dv(const DevMatrix& m, const float* r, const float* norm, float* out, bool update,
            bool exact, hipStream_t s) {
  if (exact) return bwd_dvx<NB, EPT, VK, true>(m, r, norm, out, update, s);
  bwd_dvx<NB, EPT, VK, false>(m, r, norm, out, update, s);
}

template <int NB, int EPT>
void bwd_d(const DevMatrix& m, const float* r, const float* norm, float* out, bool update,
           bool exact, hipStream_t s) {
  const DictFormat& D = m.dbwd;
  if (D.nb != NB || D.ept != EPT || !D.tiles)
    raise(MLEM_ESTATE, "backward dictionary not prepared");
  if (m.rows >= (1u << 30)) raise(MLEM_EINVAL, "dictionary kernels need rows < 2^30");
  switch (D.vk) {
    case kF32: bwd_dv<NB, EPT, kF32>(m, r, norm, out, update, exact, s); break;
    case kF16: bwd_dv<NB, EPT, kF16>(m, r, norm, out, update, exact, s); break;
    default: bwd_dv<NB, EPT, kBF16>(m, r, norm, out, update, exact, s); break;
  }
}

// Greedy partition of segments [0, n) into tiles of at most T elements and
// `maxseg` segments (one per lane); a segment longer than T forms a tile of
// its own.
std::vector<uint32_t> partition_tiles(const std::vector<uint32_t>& ptr, uint32_t T,
                                      uint32_t maxseg = kBlock) {
  const uint32_t n = (uint32_t)ptr.size() - 1;
  std::vector<uint32_t> tiles{0};
  uint32_t s = 0;
  while (s < n) {
    uint32_t e = s;
    if (ptr[s + 1] - ptr[s] > T) {
      e = s + 1;
    } else {
      uint64_t cnt = 0;
      while (e < n && e - s < maxseg && cnt + (ptr[e + 1] - ptr[e]) <= T) {
        cnt += ptr[e + 1] - ptr[e];
        ++e;
      }
    }
    tiles.push_back(e);
    s = e;
  }
  return tiles;
}

// Tile descriptors {first segment, first element}, ntiles + 1 entries.
void upload_tiles(const std::vector<uint32_t>& ptr, int ept, uint2*& dev, uint32_t& ntiles,
                  int& cur, hipStream_t s) {
  if (cur == ept && dev) return;
  const std::vector<uint32_t> t = partition_tiles(ptr, (uint32_t)(kBlock * ept));
  std::vector<uint2> d(t.size());
  for (size_t i = 0; i < t.size(); ++i) d[i] = make_uint2(t[i], ptr[t[i]]);
  if (dev) (void)hipFree(dev);
  dev = nullptr;
  cur = 0;
  MLEM_HIP(hipMalloc(&dev, sizeof(uint2) * d.size()));
  MLEM_HIP(hipMemcpyAsync(dev, d.data(), sizeof(uint2) * d.size(), hipMemcpyHostToDevice, s));
  MLEM_HIP(hipStreamSynchronize(s));
  ntiles = (uint32_t)t.size() - 1;
  cur = ept;
}

// Build the dictionary format of one projection from its device-order
// segments (host copy of the pointers, device 8-byte records).
// The format depends on the tile size T = NB*EPT and the segment limit NB;
// the build kernel runs 256-lane blocks over the same tiles.
template <int NB, int EPT, int VK>
void build_dict(const std::vector<uint32_t>& ptr, const uint2* el, DictFormat& D, hipStream_t s) {
  constexpr uint32_t T = NB * EPT;
  static_assert(T % kBlock == 0, "build blocks cover whole tiles");
  constexpr int EB = T / kBlock;  // elements per build lane
  const std::vector<uint32_t> tl = partition_tiles(ptr, T, NB);
  const uint32_t nt = (uint32_t)tl.size() - 1;
  std::vector<uint2> d2(tl.size());
  for (size_t i = 0; i < tl.size(); ++i) d2[i] = make_uint2(tl[i], ptr[tl[i]]);
  struct Buf {
    void* p = nullptr;
    ~Buf() {
      if (p) (void)hipFree(p);
    }
  } t2, ndb;
  MLEM_HIP(hipMalloc(&t2.p, sizeof(uint2) * d2.size()));
  MLEM_HIP(hipMalloc(&ndb.p, 4ull * std::max<uint32_t>(nt, 1)));
  MLEM_HIP(hipMemcpyAsync(t2.p, d2.data(), sizeof(uint2) * d2.size(), hipMemcpyHostToDevice, s));
  std::vector<uint32_t> nd(nt);
  if (nt) {
    hipLaunchKernelGGL((k_dict_build<EB, VK>), dim3(nt), dim3(kBlock), 0, s, (const uint2*)el,
                       (const uint2*)t2.p, (const uint4*)nullptr, (uint32_t*)ndb.p,
                       (uint32_t*)nullptr, (void*)nullptr, (uint16_t*)nullptr);
    check_launch("dictionary count launch");
    MLEM_HIP(hipMemcpyAsync(nd.data(), ndb.p, 4ull * nt, hipMemcpyDeviceToHost, s));
    MLEM_HIP(hipStreamSynchronize(s));
  }
  std::vector<uint4> d4(tl.size());
  uint64_t acc_a = 0, acc_g = 0;
  for (uint32_t i = 0; i <= nt; ++i) {
    d4[i] = make_uint4(tl[i], ptr[tl[i]], (uint32_t)acc_a, (uint32_t)acc_g);
    if (i == nt) break;
    const uint32_t count = ptr[tl[i + 1]] - ptr[tl[i]];
    if (count <= T) acc_a += (count + 63) / 64 * 64;
    acc_g += (nd[i] + NB - 1) / NB * NB;  // whole rounds of one slot per wave
  }
  if (acc_a + T > 0xFFFFFFFFull || acc_g > 0xFFFFFFFFull)
    raise(MLEM_EINVAL, "matrix too large for the dictionary format");
  free_dict(D);
  D.ntiles = nt;
  D.vk = VK;
  D.ndict = std::max<uint64_t>(acc_g, NB);
  D.nval = acc_a + T;  // tail: full-tile loads of the last tile stay in bounds
  const uint64_t vbytes = (VK == kF32 ? 4ull : 2ull) * D.nval;
  const uint64_t nslots = D.ndict / kWave;
  const uint64_t nmeta = nslots + (uint64_t)(NB / kWave) * EPT;  // tail: rounded-up rounds
  MLEM_HIP(hipMalloc(&D.tiles, sizeof(uint4) * d4.size()));
  MLEM_HIP(hipMalloc(&D.dict, 4ull * D.ndict));
  MLEM_HIP(hipMalloc(&D.dmeta, 4ull * nmeta));
  MLEM_HIP(hipMalloc(&D.dval, vbytes));
  MLEM_HIP(hipMalloc(&D.dlix, 2ull * D.nval));
  MLEM_HIP(hipMemsetAsync(D.dict, 0, 4ull * D.ndict, s));
  MLEM_HIP(hipMemsetAsync(D.dmeta, 0, 4ull * nmeta, s));
  MLEM_HIP(hipMemsetAsync(D.dval, 0, vbytes, s));
  MLEM_HIP(hipMemsetAsync(D.dlix, 0, 2ull * D.nval, s));
  MLEM_HIP(hipMemcpyAsync(D.tiles, d4.data(), sizeof(uint4) * d4.size(), hipMemcpyHostToDevice, s));
  if (nt) {
    hipLaunchKernelGGL((k_dict_build<EB, VK>), dim3(nt), dim3(kBlock), 0, s, (const uint2*)el,
                       (const uint2*)nullptr, (const uint4*)D.tiles, (uint32_t*)nullptr, D.dict,
                       D.dval, D.dlix);
    check_launch("dictionary fill launch");
    const uint64_t per = kBlock / kWave;
    hipLaunchKernelGGL(k_dict_pack, dim3((unsigned)((nslots + per - 1) / per)), dim3(kBlock), 0, s,
                       D.dict, D.dmeta, nslots);
    check_launch("dictionary pack launch");
  }
  MLEM_HIP(hipStreamSynchronize(s));
  D.nb = NB;
  D.ept = EPT;
}

template <int VK>
void build_dict_vk(const std::vector<uint32_t>& ptr, const uint2* el, int nb, int ept,
                   DictFormat& D, hipStream_t s) {
  switch (nb * 1000 + ept) {
    case 256008: build_dict<256, 8, VK>(ptr, el, D, s); break;
    case 256016: build_dict<256, 16, VK>(ptr, el, D, s); break;
    case 256024: build_dict<256, 24, VK>(ptr, el, D, s); break;
    case 256032: build_dict<256, 32, VK>(ptr, el, D, s); break;
    case 256048: build_dict<256, 48, VK>(ptr, el, D, s); break;
    case 256064: build_dict<256, 64, VK>(ptr, el, D, s); break;
    case 512016: build_dict<512, 16, VK>(ptr, el, D, s); break;
    case 512024: build_dict<512, 24, VK>(ptr, el, D, s); break;
    case 512032: build_dict<512, 32, VK>(ptr, el, D, s); break;
    default: raise(MLEM_EINVAL, "unsupported dictionary tile size");
  }
}

void ensure_dict(const std::vector<uint32_t>& ptr, const uint2* el, int nb, int ept, int vk,
                 DictFormat& D, hipStream_t s) {
  if (D.nb == nb && D.ept == ept && D.vk == vk && D.tiles) return;
  switch (vk) {
    case kF32: build_dict_vk<kF32>(ptr, el, nb, ept, D, s); break;
    case kF16: build_dict_vk<kF16>(ptr, el, nb, ept, D, s); break;
    case kBF16: build_dict_vk<kBF16>(ptr, el, nb, ept, D, s); break;
    default: raise(MLEM_EINVAL, "unsupported value storage");
  }
}

}  // namespace

void free_dict(DictFormat& d) {
  if (d.tiles) (void)hipFree(d.tiles);
  if (d.dict) (void)hipFree(d.dict);
  if (d.dmeta) (void)hipFree(d.dmeta);
  if (d.dval) (void)hipFree(d.dval);
  if (d.dlix) (void)hipFree(d.dlix);
  d = DictFormat();
}

void prepare_tiles(DevMatrix& m, hipStream_t s) {
  if (m.h_rowptr.size() != (size_t)m.rows + 1) {
    m.h_rowptr.resize((size_t)m.rows + 1);
    MLEM_HIP(hipMemcpyAsync(m.h_rowptr.data(), m.rowptr, 4ull * (m.rows + 1ull),
                            hipMemcpyDeviceToHost, s));
    MLEM_HIP(hipStreamSynchronize(s));
  }
  if (m.csc && m.h_colptr.size() != (size_t)m.cols + 1) {
    m.h_colptr.resize((size_t)m.cols + 1);
    MLEM_HIP(hipMemcpyAsync(m.h_colptr.data(), m.colptr, 4ull * (m.cols + 1ull),
                            hipMemcpyDeviceToHost, s));
    MLEM_HIP(hipStreamSynchronize(s));
  }
  if (m.fwd.kind == kTiles)
    upload_tiles(m.h_rowptr, m.fwd.ept, m.tiles_fwd, m.ntiles_fwd, m.tiles_fwd_ept, s);
  if (m.bwd.kind == kTiles && m.csc)
    upload_tiles(m.h_colptr, m.bwd.ept, m.tiles_bwd, m.ntiles_bwd, m.tiles_bwd_ept, s);
  if (is_dict(m.fwd.kind))
    ensure_dict(m.h_rowptr, m.csr, dict_lanes(m.fwd.kind), m.fwd.ept, m.vk, m.dfwd, s);
  if (is_dict(m.bwd.kind) && m.csc)
    ensure_dict(m.h_colptr, m.csc, dict_lanes(m.bwd.kind), m.bwd.ept, m.vk, m.dbwd, s);
}

// Blocks of the statistics-producing forward launch: the larger of the FAST
// launch configured now and the EXACT one (one lane per row).
unsigned stats_max_blocks(const DevMatrix& m) {
  unsigned fast;
  if (is_dict(m.fwd.kind))
    fast = tile_grid(m.dfwd.ntiles);
  else if (m.fwd.kind == kTiles)
    fast = tile_grid(m.ntiles_fwd);
  else
    fast = geometry(m.rows, m.fwd.lanes, m.fwd.blocks_per_cu, m.cus, true).grid;
  const unsigned exact = geometry(m.rows, 1, m.fwd.blocks_per_cu, m.cus, true).grid;
  return std::max(fast, exact);
}

void launch_fwd(const DevMatrix& m, int mode, const float* x, const int32_t* y, float* out,
                bool ratio, const StatsState* stats, hipStream_t s) {
  if (m.streamed) {
    stream_fwd(m, mode, x, y, out, ratio, stats, s);
    return;
  }
  // EXACT: the dictionary kernels (either block width) with per-lane ordered
  // sums when they are configured, else one lane per row
  const bool exact = mode == MLEM_MODE_EXACT;
  if (exact && !is_dict(m.fwd.kind)) {
    fwd_g<1, true>(m, x, y, out, ratio, stats, s);
  } else if (is_dict(m.fwd.kind)) {
    switch (dict_lanes(m.fwd.kind) * 1000 + m.fwd.ept) {
      case 256008: fwd_d<256, 8>(m, x, y, out, ratio, exact, stats, s); break;
      case 256016: fwd_d<256, 16>(m, x, y, out, ratio, exact, stats, s); break;
      case 256024: fwd_d<256, 24>(m, x, y, out, ratio, exact, stats, s); break;
      case 256032: fwd_d<256, 32>(m, x, y, out, ratio, exact, stats, s); break;
      case 256048: fwd_d<256, 48>(m, x, y, out, ratio, exact, stats, s); break;
      case 256064: fwd_d<256, 64>(m, x, y, out, ratio, exact, stats, s); break;
      case 512016: fwd_d<512, 16>(m, x, y, out, ratio, exact, stats, s); break;
      case 512024: fwd_d<512, 24>(m, x, y, out, ratio, exact, stats, s); break;
      case 512032: fwd_d<512, 32>(m, x, y, out, ratio, exact, stats, s); break;
      default: raise(MLEM_EINVAL, "unsupported forward dictionary tile size");
    }
  } else if (m.fwd.kind == kTiles) {
    switch (m.fwd.ept) {
      case 4: fwd_t<4>(m, x, y, out, ratio, stats, s); break;
      case 8: fwd_t<8>(m, x, y, out, ratio, stats, s); break;
      case 16: fwd_t<16>(m, x, y, out, ratio, stats, s); break;
      default: raise(MLEM_EINVAL, "unsupported forward tile size");
    }
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
  if (m.streamed) {
    stream_bwd(m, mode, r, norm, out, update, s);
    return;
  }
  if (!m.csc) raise(MLEM_ESTATE, "matrix was opened without a CSC transpose");
  const bool exact = mode == MLEM_MODE_EXACT;
  if (exact && !is_dict(m.bwd.kind)) {
    bwd_g<1, true>(m, r, norm, out, update, s);
  } else if (is_dict(m.bwd.kind)) {
    switch (dict_lanes(m.bwd.kind) * 1000 + m.bwd.ept) {
      case 256008: bwd_d<256, 8>(m, r, norm, out, update, exact, s); break;
      case 256016: bwd_d<256, 16>(m, r, norm, out, update, exact, s); break;
      case 256024: bwd_d<256, 24>(m, r, norm, out, update, exact, s); break;
      case 256032: bwd_d<256, 32>(m, r, norm, out, update, exact, s); break;
      case 256048: bwd_d<256, 48>(m, r, norm, out, update, exact, s); break;
      case 256064: bwd_d<256, 64>(m, r, norm, out, update, exact, s); break;
      case 512016: bwd_d<512, 16>(m, r, norm, out, update, exact, s); break;
      case 512024: bwd_d<512, 24>(m, r, norm, out, update, exact, s); break;
      case 512032: bwd_d<512, 32>(m, r, norm, out, update, exact, s); break;
      default: raise(MLEM_EINVAL, "unsupported backward dictionary tile size");
    }
  } else if (m.bwd.kind == kTiles) {
    switch (m.bwd.ept) {
      case 4: bwd_t<4>(m, r, norm, out, update, s); break;
      case 8: bwd_t<8>(m, r, norm, out, update, s); break;
      case 16: bwd_t<16>(m, r, norm, out, update, s); break;
      default: raise(MLEM_EINVAL, "unsupported backward tile size");
    }
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

void launch_fwd_block(const uint32_t* rowptr, const uint2* csr, uint32_t rows, int lanes,
                      bool exact, const float* x, const int32_t* y, float* out, bool ratio,
                      const StatsState* st, int cus, hipStream_t s) {
  if (exact) {
    fwd_block_g<1, true>(rowptr, csr, rows, x, y, out, ratio, st, cus, s);
  } else {
    switch (lanes) {
      case 4: fwd_block_g<4, false>(rowptr, csr, rows, x, y, out, ratio, st, cus, s); break;
      case 8: fwd_block_g<8, false>(rowptr, csr, rows, x, y, out, ratio, st, cus, s); break;
      case 16: fwd_block_g<16, false>(rowptr, csr, rows, x, y, out, ratio, st, cus, s); break;
      case 32: fwd_block_g<32, false>(rowptr, csr, rows, x, y, out, ratio, st, cus, s); break;
      case 64: fwd_block_g<64, false>(rowptr, csr, rows, x, y, out, ratio, st, cus, s); break;
      default: raise(MLEM_EINVAL, "unsupported forward lane count");
    }
  }
  check_launch("streamed forward projection launch");
}

void launch_bwd_block_add(const uint32_t* colptr, const uint2* csc, uint32_t cols, int lanes,
                          bool exact, const float* r, float* u, int cus, hipStream_t s) {
  if (exact) {
    bwd_block_add_g<1, true>(colptr, csc, cols, r, u, cus, s);
  } else {
    switch (lanes) {
      case 4: bwd_block_add_g<4, false>(colptr, csc, cols, r, u, cus, s); break;
      case 8: bwd_block_add_g<8, false>(colptr, csc, cols, r, u, cus, s); break;
      case 16: bwd_block_add_g<16, false>(colptr, csc, cols, r, u, cus, s); break;
      case 32: bwd_block_add_g<32, false>(colptr, csc, cols, r, u, cus, s); break;
      case 64: bwd_block_add_g<64, false>(colptr, csc, cols, r, u, cus, s); break;
      default: raise(MLEM_EINVAL, "unsupported backward lane count");
    }
  }
  check_launch("streamed backprojection launch");
}

void launch_scatter(const DevMatrix& m, const float* r, float* u, hipStream_t s) {
  if (m.streamed) {  // u += A^T r, one block after another (no atomics needed)
    stream_bwd_add(m, false, r, u, s);
    return;
  }
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

void launch_scatter_block(const uint32_t* rowptr, const uint2* csr, uint32_t rows, int lanes,
                          const float* r, float* u, int cus, hipStream_t s) {
  switch (lanes) {
    case 4: scatter_block_g<4>(rowptr, csr, rows, r, u, cus, s); break;
    case 8: scatter_block_g<8>(rowptr, csr, rows, r, u, cus, s); break;
    case 16: scatter_block_g<16>(rowptr, csr, rows, r, u, cus, s); break;
    case 32: scatter_block_g<32>(rowptr, csr, rows, r, u, cus, s); break;
    case 64: scatter_block_g<64>(rowptr, csr, rows, r, u, cus, s); break;
    default: raise(MLEM_EINVAL, "unsupported scatter lane count");
  }
  check_launch("streamed scatter launch");
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

void launch_quantize(DevMatrix& m, int vk, float* rscale, hipStream_t s) {
  if (m.rows == 0) return;
  const dim3 grid(blocks_for((uint64_t)m.rows * 16)), block(kBlock);
  if (vk == kF16)
    hipLaunchKernelGGL(k_quantize_rows<kF16>, grid, block, 0, s, m.rowptr, m.csr, m.rows, rscale);
  else if (vk == kBF16)
    hipLaunchKernelGGL(k_quantize_rows<kBF16>, grid, block, 0, s, m.rowptr, m.csr, m.rows, rscale);
  else
    raise(MLEM_EINVAL, "unsupported value storage");
  check_launch("quantization launch");
}

void launch_colsums_scaled(const DevMatrix& m, float* norm, hipStream_t s) {
  if (!m.csc || !m.rscale) raise(MLEM_ESTATE, "scaled column sums need the CSC and row scales");
  hipLaunchKernelGGL(k_colsum_scaled, dim3(blocks_for(m.cols)), dim3(kBlock), 0, s, m.colptr,
                     m.csc, m.rscale, norm, m.cols);
  check_launch("column sums launch");
}

void launch_scale_rows(const DevMatrix& m, float* r, hipStream_t s) {
  if (!m.rscale || m.rows == 0) return;
  hipLaunchKernelGGL(k_scale_rows, dim3(blocks_for(m.rows)), dim3(kBlock), 0, s, r, m.rscale,
                     m.rows);
  check_launch("row scale launch");
}

void launch_fill(float* p, float v, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_fill, dim3(blocks_for(n)), dim3(kBlock), 0, s, p, v, (uint64_t)n);
  check_launch("fill launch");
}

void launch_copy(const uint4* src, uint4* dst, size_t n16, hipStream_t s) {
  hipLaunchKernelGGL(k_copy, dim3(blocks_for(n16)), dim3(kBlock), 0, s, src, dst, (uint64_t)n16);
  check_launch("copy launch");
}

}  // namespace mlem
