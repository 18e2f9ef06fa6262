// spmm.hpp — sparse matrix times a block of W interleaved dense vectors
// (device code shared by the batched and the unrolled/adjoint paths).
//
// The W right-hand sides of one block are stored interleaved per element
// (V[j*W + f]), so the W lanes that own one matrix segment gather one
// contiguous W*4-byte run per record and every matrix record streamed from
// HBM serves W vectors. A 256-lane workgroup walks whole tiles of segments
// (rows of the CSR or columns of the CSC) in a persistent loop, stages each
// tile's 8-byte records in LDS with coalesced loads, and gives every segment
// to a group of W lanes; lane f accumulates vector f sequentially in element
// order with separate multiply and add (this code is compiled with
// -ffp-contract=off), so every vector's result is the reference driver's
// sequential fp32 sum (mlem_seq/seqcsr4mlem.cpp:69-91 for the CSR,
// :108-120 for the CSC in ascending row order).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <vector>

namespace mlem {
namespace spmm {

constexpr int kBlock = 256;
constexpr uint32_t kTile = 2048;  // records staged per tile (16 KB of LDS)
#ifndef MLEM_SPMM_UNROLL
#define MLEM_SPMM_UNROLL 16
#endif
constexpr int kUnroll = MLEM_SPMM_UNROLL;  // gathers in flight per lane

// Sequential fp32 dot product of one segment's records with vector f of an
// interleaved block. The gathers of U consecutive records are issued
// together (the segment's last group with clamped indices), the products are
// then added one by one in element order.
template <int W, int U = kUnroll>
__device__ __forceinline__ float seg_dot(const uint2* rec, uint32_t b, uint32_t e,
                                         const float* __restrict__ v, uint32_t f) {
  float acc = 0.f;
  for (uint32_t k = b; k < e; k += U) {
    float a[U], g[U];
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const uint2 r = rec[min(k + j, e - 1)];
      a[j] = __uint_as_float(r.y);
      g[j] = v[(size_t)r.x * W + f];
    }
#pragma unroll
    for (int j = 0; j < U; ++j)
      if (k + j < e) acc = acc + a[j] * g[j];
  }
  return acc;
}

// Walk the tiles [0, ntiles) of a segment partition (tiles[t] = {first
// segment, first element}) and call epi(segment, f, dot) for every segment
// and vector. Tiles hold at most 256 segments and kTile records, or one
// longer segment (read from HBM, 4 records at a time: its 64-bit record
// addresses would otherwise set the kernel's register count). Workgroup b runs on XCD b % 8 (the grid is
// a multiple of 8): XCD x walks the x-th contiguous eighth of the tiles, so
// neighbouring segments (rows of one sinogram region, voxels of one brick
// range) share that XCD's L2.
template <int W, class Epi>
__device__ __forceinline__ void walk(const uint32_t* __restrict__ ptr,
                                     const uint2* __restrict__ el,
                                     const uint2* __restrict__ tiles, uint32_t ntiles,
                                     const float* __restrict__ v, Epi& epi) {
  constexpr int G = kBlock / W;
  __shared__ uint2 rec[kTile];
  __shared__ uint32_t S[kBlock + 1];
  const uint32_t t = threadIdx.x, f = t % W, grp = t / W;
  const uint32_t xcd = blockIdx.x & 7u, per_xcd = gridDim.x >> 3;
  const uint32_t t0 = (uint32_t)((uint64_t)ntiles * xcd / 8);
  const uint32_t t1 = (uint32_t)((uint64_t)ntiles * (xcd + 1) / 8);
  for (uint32_t tile = t0 + (blockIdx.x >> 3); tile < t1; tile += per_xcd) {
    const uint2 d0 = tiles[tile], d1 = tiles[tile + 1];
    const uint32_t s0 = d0.x, e0 = d0.y, nseg = d1.x - s0, count = d1.y - e0;
    __syncthreads();  // the previous tile's readers are done with rec / S
    const bool staged = count <= kTile;
    if (staged)
      for (uint32_t i = t; i < count; i += kBlock) rec[i] = el[e0 + i];
    for (uint32_t i = t; i <= nseg; i += kBlock) S[i] = ptr[s0 + i] - e0;
    __syncthreads();
    for (uint32_t sg = grp; sg < nseg; sg += G) {
      const float d = staged ? seg_dot<W>(rec, S[sg], S[sg + 1], v, f)
                             : seg_dot<W, 4>(el + e0, S[sg], S[sg + 1], v, f);
      epi(s0 + sg, f, d);
    }
  }
}

// Greedy partition of segments [0, n) into tiles of at most T records and
// 256 segments; a segment longer than T forms a tile of its own. Returns the
// first segment of every tile plus n.
inline std::vector<uint32_t> partition(const std::vector<uint32_t>& ptr, uint32_t T) {
  const uint32_t n = (uint32_t)ptr.size() - 1;
  std::vector<uint32_t> tiles{0};
  uint32_t s = 0;
  while (s < n) {
    uint32_t e = s;
    if (ptr[s + 1] - ptr[s] > T) {
      e = s + 1;
    } else {
      uint64_t cnt = 0;
      while (e < n && e - s < (uint32_t)kBlock && cnt + (ptr[e + 1] - ptr[e]) <= T) {
        cnt += ptr[e + 1] - ptr[e];
        ++e;
      }
    }
    tiles.push_back(e);
    s = e;
  }
  return tiles;
}

}  // namespace spmm
}  // namespace mlem
