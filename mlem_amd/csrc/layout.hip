// layout.hip — load-time work on the device: validation, the CSC transpose,
// and the locality layout of rows and voxels (gfx950).
//
// Locality layout. The projections are bound by the cache lines their
// gathers touch: a line of response visits voxels whose linear numbers
// (x + y*nx + z*nx*ny) are mostly far apart, and rows in file order are
// unrelated lines. At load time, when the voxel grid is known, the device
// copy is renumbered:
//   voxels  -> 4x4x2 bricks (32 floats = one 128-byte line per brick), so a
//              line of response finds several of its voxels in one cache line,
//              the bricks grouped 8x8x2 so that consecutive tiles stay close;
//   rows    -> sinogram order (projected direction, then radial offset and
//              axial centre in 4 x 4 blocks), so the rows one CU works on
//              at a time are near-parallel neighbours that gather the same
//              lines of x, and the rows a column gathers in the
//              backprojection are clustered in r.
// Only indices change: every row keeps its elements in file order and every
// CSC column keeps its rows in ascending FILE order (the CSC is sorted from
// the file-order CSR before rows are renumbered), so each row sum and column
// sum adds the same terms in the same order as before and EXACT mode stays
// bit-identical to the reference's sequential sums. Vectors cross the API in
// file order; the gathers/scatters below translate at the boundary.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <vector>

#include "internal.hpp"
#include "mlem.h"

namespace mlem {
namespace {

constexpr int kBlock = 256;

inline uint32_t blocks_for(uint64_t n) { return (uint32_t)std::max<uint64_t>(1, (n + kBlock - 1) / kBlock); }

void check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) raise(MLEM_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
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
  const uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x;
  if (i == 0) rowptr[0] = 0;
  if (i < rows) rowptr[i + 1] = (uint32_t)rowidx[i];
}

// Sinogram sort key of every row from its first and last voxel (projected
// direction) and its voxel centroid (radial offset, axial centre): 64
// direction bins, 256 offset bins, 64 axial bins. Within a direction the
// (offset, axial) plane is ordered in blocks of 4 x 4 bins, so the rows through
// a compact group of voxels (a backprojection tile) fill fewer lines of r than
// in offset-major order (pet: backprojection 3-4% faster, forward 2%):
// 6 bits direction | 6 bits offset / 4 | 4 bits axial / 4 | 2 bits offset % 4 |
// 2 bits axial % 4. Empty rows sort last.
__global__ __launch_bounds__(kBlock) void k_row_keys(const uint32_t* __restrict__ rowptr,
                                                     const uint2* __restrict__ csr, uint32_t rows,
                                                     uint32_t nx, uint32_t ny, uint32_t nz,
                                                     uint32_t* __restrict__ keys) {
  constexpr int G = 16;
  const uint64_t t = blockIdx.x * (uint64_t)kBlock + threadIdx.x;
  const uint32_t lane = (uint32_t)(t % G);
  const uint64_t row = t / G;
  const bool active = row < rows;
  uint32_t beg = 0, end = 0;
  if (active) {
    beg = rowptr[row];
    end = rowptr[row + 1];
  }
  const uint32_t nxy = nx * ny;
  float sx = 0.f, sy = 0.f, sz = 0.f;
  for (uint32_t k = beg + lane; k < end; k += G) {
    const uint32_t c = csr[k].x;
    sx += (float)(c % nx);
    sy += (float)((c / nx) % ny);
    sz += (float)(c / nxy);
  }
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) {
    sx += __shfl_xor(sx, off, G);
    sy += __shfl_xor(sy, off, G);
    sz += __shfl_xor(sz, off, G);
  }
  if (!active || lane != 0) return;
  if (end == beg) {
    keys[row] = 0xFFFFFFFFu;
    return;
  }
  const float n = (float)(end - beg);
  const uint32_t c0 = csr[beg].x, c1 = csr[end - 1].x;
  const float x0 = (float)(c0 % nx), y0 = (float)((c0 / nx) % ny);
  const float x1 = (float)(c1 % nx), y1 = (float)((c1 / nx) % ny);
  const float pi = 3.14159265358979f;
  float phi = atan2f(y1 - y0, x1 - x0);
  if (phi < 0.f) phi += pi;
  if (phi >= pi) phi -= pi;
  const float mx = div_rn(sx, n) - 0.5f * nx, my = div_rn(sy, n) - 0.5f * ny, mz = div_rn(sz, n);
  const float s = -sinf(phi) * mx + cosf(phi) * my;
  const float smax = 0.5f * sqrtf((float)nx * nx + (float)ny * ny) + 1.f;
  const uint32_t pb = min((uint32_t)(div_rn(phi, pi) * 64.f), 63u);
  const uint32_t sb = (uint32_t)min(max(div_rn(s + smax, 2.f * smax) * 256.f, 0.f), 255.f);
  const uint32_t zb = (uint32_t)min(max(div_rn(mz, (float)nz) * 64.f, 0.f), 63.f);
  keys[row] = (pb << 14) | ((sb >> 2) << 8) | ((zb >> 2) << 4) | ((sb & 3u) << 2) | (zb & 3u);
}

__global__ __launch_bounds__(kBlock) void k_iota(uint32_t* __restrict__ p, uint32_t n) {
  const uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x;
  if (i < n) p[i] = (uint32_t)i;
}

__global__ __launch_bounds__(kBlock) void k_invert(const uint32_t* __restrict__ perm, uint32_t n,
                                                   uint32_t* __restrict__ inv) {
  const uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x;
  if (i < n) inv[perm[i]] = (uint32_t)i;
}

// Length of device row i (file row rowold[i]); entry n is 0 for the scan.
__global__ __launch_bounds__(kBlock) void k_perm_lengths(const uint32_t* __restrict__ rowptr,
                                                         const uint32_t* __restrict__ rowold,
                                                         uint32_t rows, uint32_t* __restrict__ len) {
  const uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x;
  if (i < rows) {
    const uint32_t o = rowold[i];
    len[i] = rowptr[o + 1] - rowptr[o];
  } else if (i == rows) {
    len[i] = 0;
  }
}

// Device-order CSR: row i copies file row rowold[i] in element order, with
// column numbers translated through colmap.
__global__ __launch_bounds__(kBlock) void k_copy_rows(const uint32_t* __restrict__ rowptr,
                                                      const uint2* __restrict__ csr,
                                                      const uint32_t* __restrict__ rowold,
                                                      const uint32_t* __restrict__ colmap,
                                                      const uint32_t* __restrict__ nrowptr,
                                                      uint2* __restrict__ ncsr, uint32_t rows) {
  constexpr int G = 16;
  const uint64_t t = blockIdx.x * (uint64_t)kBlock + threadIdx.x;
  const uint64_t row = t / G;
  const uint32_t lane = (uint32_t)(t % G);
  if (row >= rows) return;
  const uint32_t o = rowold ? rowold[row] : (uint32_t)row;
  const uint32_t b = rowptr[o], e = rowptr[o + 1], d = nrowptr[row];
  for (uint32_t k = lane; b + k < e; k += G) {
    uint2 v = csr[b + k];
    if (colmap) v.x = colmap[v.x];
    ncsr[d + k] = v;
  }
}

// CSC keys/values from the FILE-order CSR: key = device column, value =
// {device row, value bits} as one u64 (the CSC record layout).
__global__ __launch_bounds__(kBlock) void k_csc_keys(const uint32_t* __restrict__ rowptr,
                                                     const uint2* __restrict__ csr, uint32_t rows,
                                                     const uint32_t* __restrict__ colmap,
                                                     const uint32_t* __restrict__ rowpos,
                                                     uint32_t* __restrict__ keys,
                                                     unsigned long long* __restrict__ vals) {
  constexpr int G = 16;
  const uint64_t t = blockIdx.x * (uint64_t)kBlock + threadIdx.x;
  const uint64_t row = t / G;
  const uint32_t lane = (uint32_t)(t % G);
  if (row >= rows) return;
  const uint32_t drow = rowpos ? rowpos[row] : (uint32_t)row;
  for (uint32_t k = rowptr[row] + lane; k < rowptr[row + 1]; k += G) {
    const uint2 e = csr[k];
    keys[k] = colmap ? colmap[e.x] : e.x;
    vals[k] = ((unsigned long long)e.y << 32) | (unsigned long long)drow;
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

// colptr[c] = first position with key >= c, by binary search (one lane per
// column, c = 0..cols): no lane walks a long run of empty columns.
__global__ __launch_bounds__(kBlock) void k_colptr_search(const uint32_t* __restrict__ skeys,
                                                          uint64_t nnz, uint32_t cols,
                                                          uint32_t* __restrict__ colptr) {
  const uint64_t c = blockIdx.x * (uint64_t)kBlock + threadIdx.x;
  if (c > cols) return;
  uint64_t lo = 0, hi = nnz;  // answer in [lo, hi]
  while (lo < hi) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if (skeys[mid] < c)
      lo = mid + 1;
    else
      hi = mid;
  }
  colptr[c] = (uint32_t)lo;
}

__global__ __launch_bounds__(kBlock) void k_fill_u32(uint32_t* p, uint32_t v, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kBlock)
    p[i] = v;
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_gather(T* __restrict__ out, const T* __restrict__ in,
                                                   const uint32_t* __restrict__ idx, uint32_t n) {
  const uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x;
  if (i < n) out[i] = in[idx[i]];
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_scatter_perm(T* __restrict__ out,
                                                         const T* __restrict__ in,
                                                         const uint32_t* __restrict__ idx,
                                                         uint32_t n) {
  const uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x;
  if (i < n) out[idx[i]] = in[i];
}

struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  template <class T>
  T* as() {
    return static_cast<T*>(p);
  }
  void alloc(size_t bytes) { MLEM_HIP(hipMalloc(&p, std::max<size_t>(bytes, 16))); }
};

int bits_for(uint32_t maxval) { return maxval == 0 ? 1 : 32 - __builtin_clz(maxval); }

}  // namespace

// -------------------------------------------------------------- launchers --
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

std::vector<uint32_t> brick_colmap(uint32_t nx, uint32_t ny, uint32_t nz) {
  constexpr uint32_t BX = 4, BY = 4, BZ = 2, BV = BX * BY * BZ;
  // bricks in raster order within groups of 8 x 8 x 2 bricks (32 x 32 x 4
  // voxels), the groups in raster order: the tiles one XCD works through in
  // turn stay in one compact region, whose rows share lines of r in L2
  // (backprojection of the 192x192x64 matrix 5% faster than plain raster order)
  constexpr uint64_t SX = 8, SY = 8, SZ = 2;
  const uint64_t nsx = (nx + BX * SX - 1) / (BX * SX), nsy = (ny + BY * SY - 1) / (BY * SY),
                 nsz = (nz + BZ * SZ - 1) / (BZ * SZ);
  const uint64_t nbx = nsx * SX, nby = nsy * SY, nbz = nsz * SZ;
  const uint64_t cols = (uint64_t)nx * ny * nz;
  const uint64_t space = nbx * nby * nbz * BV;
  std::vector<uint32_t> key(cols);
  std::vector<uint8_t> used(space, 0);
  for (uint64_t c = 0; c < cols; ++c) {
    const uint64_t x = c % nx, y = (c / nx) % ny, z = c / ((uint64_t)nx * ny);
    const uint64_t bx = x / BX, by = y / BY, bz = z / BZ;
    const uint64_t b = ((bx / SX) + nsx * ((by / SY) + nsy * (bz / SZ))) * (SX * SY * SZ) +
                       (bx % SX) + SX * ((by % SY) + SY * (bz % SZ));
    const uint64_t in = (x % BX) + BX * ((y % BY) + BY * (z % BZ));
    key[c] = (uint32_t)(b * BV + in);
    used[b * BV + in] = 1;
  }
  // rank the (sparse when the grid is not a multiple of the brick) keys
  std::vector<uint32_t> rank(space);
  uint32_t acc = 0;
  for (uint64_t k = 0; k < space; ++k) {
    rank[k] = acc;
    acc += used[k];
  }
  for (uint64_t c = 0; c < cols; ++c) key[c] = rank[key[c]];
  return key;
}

namespace {
size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }
size_t block_csc_sort_bytes(uint64_t nnz, uint32_t cols) {
  size_t tb = 0;
  MLEM_HIP(rocprim::radix_sort_pairs(nullptr, tb, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                     (unsigned long long*)nullptr, (unsigned long long*)nullptr,
                                     (size_t)nnz, 0, bits_for(cols ? cols - 1 : 0),
                                     hipStream_t(nullptr)));
  return tb;
}
}  // namespace

size_t block_csc_scratch(uint64_t nnz, uint32_t cols) {
  return up256(4 * nnz) * 2 + up256(8 * nnz) + up256(block_csc_sort_bytes(nnz, cols));
}

void build_block_csc(const uint32_t* rowptr, const uint2* csr, uint32_t rows, uint32_t cols,
                     uint64_t nnz, void* scratch, size_t scratch_bytes, uint32_t* colptr,
                     uint2* csc, hipStream_t s) {
  if (nnz == 0) {
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for((uint64_t)cols + 1)), dim3(kBlock), 0, s, colptr,
                       0u, (uint64_t)cols + 1);
    check_launch("block colptr fill");
    return;
  }
  char* p = static_cast<char*>(scratch);
  uint32_t* keys = reinterpret_cast<uint32_t*>(p);
  uint32_t* skeys = reinterpret_cast<uint32_t*>(p + up256(4 * nnz));
  auto* vals = reinterpret_cast<unsigned long long*>(p + 2 * up256(4 * nnz));
  void* temp = p + 2 * up256(4 * nnz) + up256(8 * nnz);
  size_t tb = block_csc_sort_bytes(nnz, cols);
  if (2 * up256(4 * nnz) + up256(8 * nnz) + tb > scratch_bytes)
    raise(MLEM_ESTATE, "block CSC scratch too small");
  hipLaunchKernelGGL(k_csc_keys, dim3(blocks_for((uint64_t)rows * 16)), dim3(kBlock), 0, s, rowptr,
                     csr, rows, (const uint32_t*)nullptr, (const uint32_t*)nullptr, keys, vals);
  check_launch("block csc key launch");
  MLEM_HIP(rocprim::radix_sort_pairs(temp, tb, keys, skeys, vals,
                                     reinterpret_cast<unsigned long long*>(csc), (size_t)nnz, 0,
                                     bits_for(cols ? cols - 1 : 0), s));
  hipLaunchKernelGGL(k_colptr_search, dim3(blocks_for((uint64_t)cols + 1)), dim3(kBlock), 0, s,
                     skeys, nnz, cols, colptr);
  check_launch("block colptr launch");
}

void build_layout(DevMatrix& m, bool reorder_rows, bool reorder_cols, bool want_csc,
                  hipStream_t s) {
  const bool geom = m.grid[0] && m.grid[1] && m.grid[2] &&
                    (uint64_t)m.grid[0] * m.grid[1] * m.grid[2] == m.cols;
  reorder_rows = reorder_rows && geom && m.rows > 1 && m.nnz > 0;
  reorder_cols = reorder_cols && geom && m.cols > 1;

  if (reorder_cols) {
    std::vector<uint32_t> cm = brick_colmap(m.grid[0], m.grid[1], m.grid[2]);
    MLEM_HIP(hipMalloc(&m.colmap, 4ull * m.cols));
    MLEM_HIP(hipMemcpyAsync(m.colmap, cm.data(), 4ull * m.cols, hipMemcpyHostToDevice, s));
    MLEM_HIP(hipStreamSynchronize(s));
  }
  if (reorder_rows) {
    DevBuf keys, skeys, iota;
    keys.alloc(4ull * m.rows);
    skeys.alloc(4ull * m.rows);
    iota.alloc(4ull * m.rows);
    MLEM_HIP(hipMalloc(&m.rowold, 4ull * m.rows));
    MLEM_HIP(hipMalloc(&m.rowpos, 4ull * m.rows));
    hipLaunchKernelGGL(k_row_keys, dim3(blocks_for((uint64_t)m.rows * 16)), dim3(kBlock), 0, s,
                       m.rowptr, m.csr, m.rows, m.grid[0], m.grid[1], m.grid[2],
                       keys.as<uint32_t>());
    hipLaunchKernelGGL(k_iota, dim3(blocks_for(m.rows)), dim3(kBlock), 0, s, iota.as<uint32_t>(),
                       m.rows);
    check_launch("row key launch");
    size_t tb = 0;
    MLEM_HIP(rocprim::radix_sort_pairs(nullptr, tb, keys.as<uint32_t>(), skeys.as<uint32_t>(),
                                       iota.as<uint32_t>(), m.rowold, (size_t)m.rows, 0, 32, s));
    DevBuf temp;
    temp.alloc(tb);
    MLEM_HIP(rocprim::radix_sort_pairs(temp.p, tb, keys.as<uint32_t>(), skeys.as<uint32_t>(),
                                       iota.as<uint32_t>(), m.rowold, (size_t)m.rows, 0, 32, s));
    hipLaunchKernelGGL(k_invert, dim3(blocks_for(m.rows)), dim3(kBlock), 0, s, m.rowold, m.rows,
                       m.rowpos);
    check_launch("row permutation launch");
    MLEM_HIP(hipStreamSynchronize(s));
  }

  // CSC from the FILE-order CSR (stable: file row order inside each column).
  if (want_csc) {
    MLEM_HIP(hipMalloc(&m.colptr, sizeof(uint32_t) * ((size_t)m.cols + 1)));
    MLEM_HIP(hipMalloc(&m.csc, sizeof(uint2) * std::max<uint64_t>(m.nnz, 1)));
    if (m.nnz == 0) {
      hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for((uint64_t)m.cols + 1)), dim3(kBlock), 0, s,
                         m.colptr, 0u, (uint64_t)m.cols + 1);
      check_launch("colptr fill");
    } else {
      DevBuf keys, skeys, vals, temp;
      keys.alloc(4ull * m.nnz);
      skeys.alloc(4ull * m.nnz);
      vals.alloc(8ull * m.nnz);
      hipLaunchKernelGGL(k_csc_keys, dim3(blocks_for((uint64_t)m.rows * 16)), dim3(kBlock), 0, s,
                         m.rowptr, m.csr, m.rows, m.colmap, m.rowpos, keys.as<uint32_t>(),
                         vals.as<unsigned long long>());
      check_launch("csc key launch");
      const int end_bit = bits_for(m.cols - 1);
      size_t tb = 0;
      auto* out = reinterpret_cast<unsigned long long*>(m.csc);
      MLEM_HIP(rocprim::radix_sort_pairs(nullptr, tb, keys.as<uint32_t>(), skeys.as<uint32_t>(),
                                         vals.as<unsigned long long>(), out, (size_t)m.nnz, 0,
                                         end_bit, s));
      temp.alloc(tb);
      MLEM_HIP(rocprim::radix_sort_pairs(temp.p, tb, keys.as<uint32_t>(), skeys.as<uint32_t>(),
                                         vals.as<unsigned long long>(), out, (size_t)m.nnz, 0,
                                         end_bit, s));
      hipLaunchKernelGGL(k_colptr, dim3(blocks_for(m.nnz)), dim3(kBlock), 0, s,
                         skeys.as<uint32_t>(), m.nnz, m.cols, m.colptr);
      check_launch("colptr launch");
      MLEM_HIP(hipStreamSynchronize(s));
    }
  }

  // Device-order CSR.
  if (reorder_rows || reorder_cols) {
    DevBuf len;
    len.alloc(4ull * ((uint64_t)m.rows + 1));
    uint32_t* nrowptr = nullptr;
    uint2* ncsr = nullptr;
    MLEM_HIP(hipMalloc(&nrowptr, 4ull * ((uint64_t)m.rows + 1)));
    DevBuf guard_rp;
    guard_rp.p = nrowptr;
    MLEM_HIP(hipMalloc(&ncsr, 8ull * std::max<uint64_t>(m.nnz, 1)));
    DevBuf guard_csr;
    guard_csr.p = ncsr;
    if (reorder_rows) {
      hipLaunchKernelGGL(k_perm_lengths, dim3(blocks_for((uint64_t)m.rows + 1)), dim3(kBlock), 0,
                         s, m.rowptr, m.rowold, m.rows, len.as<uint32_t>());
      check_launch("row length launch");
      size_t tb = 0;
      MLEM_HIP(rocprim::exclusive_scan(nullptr, tb, len.as<uint32_t>(), nrowptr, 0u,
                                       (size_t)m.rows + 1, rocprim::plus<uint32_t>(), s));
      DevBuf temp;
      temp.alloc(tb);
      MLEM_HIP(rocprim::exclusive_scan(temp.p, tb, len.as<uint32_t>(), nrowptr, 0u,
                                       (size_t)m.rows + 1, rocprim::plus<uint32_t>(), s));
    } else {
      MLEM_HIP(hipMemcpyAsync(nrowptr, m.rowptr, 4ull * ((uint64_t)m.rows + 1),
                              hipMemcpyDeviceToDevice, s));
    }
    hipLaunchKernelGGL(k_copy_rows, dim3(blocks_for((uint64_t)m.rows * 16)), dim3(kBlock), 0, s,
                       m.rowptr, m.csr, m.rowold, m.colmap, nrowptr, ncsr, m.rows);
    check_launch("row copy launch");
    MLEM_HIP(hipStreamSynchronize(s));
    (void)hipFree(m.rowptr);
    (void)hipFree(m.csr);
    m.rowptr = nrowptr;
    m.csr = ncsr;
    guard_rp.p = nullptr;
    guard_csr.p = nullptr;
  }
}

// ------------------------------------------------- boundary translations --
void to_device_cols(const DevMatrix& m, const float* file_order, float* dev, hipStream_t s) {
  if (!m.colmap) {
    if (file_order != dev)
      MLEM_HIP(hipMemcpyAsync(dev, file_order, 4ull * m.cols, hipMemcpyDeviceToDevice, s));
    return;
  }
  hipLaunchKernelGGL(k_scatter_perm<float>, dim3(blocks_for(m.cols)), dim3(kBlock), 0, s, dev,
                     file_order, m.colmap, m.cols);
  check_launch("column scatter launch");
}

void to_file_cols(const DevMatrix& m, const float* dev, float* file_order, hipStream_t s) {
  if (!m.colmap) {
    if (file_order != dev)
      MLEM_HIP(hipMemcpyAsync(file_order, dev, 4ull * m.cols, hipMemcpyDeviceToDevice, s));
    return;
  }
  hipLaunchKernelGGL(k_gather<float>, dim3(blocks_for(m.cols)), dim3(kBlock), 0, s, file_order,
                     dev, m.colmap, m.cols);
  check_launch("column gather launch");
}

template <class T>
void rows_to_device(const DevMatrix& m, const T* file_order, T* dev, hipStream_t s) {
  if (!m.rowold) {
    if (file_order != dev)
      MLEM_HIP(hipMemcpyAsync(dev, file_order, sizeof(T) * m.rows, hipMemcpyDeviceToDevice, s));
    return;
  }
  hipLaunchKernelGGL(k_gather<T>, dim3(blocks_for(m.rows)), dim3(kBlock), 0, s, dev, file_order,
                     m.rowold, m.rows);
  check_launch("row gather launch");
}

template <class T>
void rows_to_file(const DevMatrix& m, const T* dev, T* file_order, hipStream_t s) {
  if (!m.rowold) {
    if (file_order != dev)
      MLEM_HIP(hipMemcpyAsync(file_order, dev, sizeof(T) * m.rows, hipMemcpyDeviceToDevice, s));
    return;
  }
  hipLaunchKernelGGL(k_scatter_perm<T>, dim3(blocks_for(m.rows)), dim3(kBlock), 0, s, file_order,
                     dev, m.rowold, m.rows);
  check_launch("row scatter launch");
}

template void rows_to_device<float>(const DevMatrix&, const float*, float*, hipStream_t);
template void rows_to_device<int32_t>(const DevMatrix&, const int32_t*, int32_t*, hipStream_t);
template void rows_to_file<float>(const DevMatrix&, const float*, float*, hipStream_t);

}  // namespace mlem
