// batch.hip — many sinograms (frames) reconstructed against one system
// matrix: the RHS-block (SpMM) form of MLEM on gfx950.
//
// Frames are grouped into blocks of W (4, 8, 16 or 32) right-hand sides. A
// block's image, counts and ratios are stored interleaved per voxel / row
// (X[col*W + f]), so the lanes that own one voxel's frames gather a whole
// contiguous W*4-byte run (one 128-byte line at W = 32) and every matrix
// record streamed from HBM serves W frames.
//
// MLEM_MODE_FAST (the default): a block's products run through the matrix's
// box tiling (panel.hip): sums per (row, box) piece, added in a fixed order
// by the epilogue kernel k_tiled_ratio (ratio and statistics), and complete
// per-voxel backprojections, applied by k_tiled_update (masked update);
// deterministic, not the reference's summation order.
// MLEM_MODE_EXACT: each 256-lane block walks whole tiles of rows (CSC
// columns), stages the tile's 8-byte records in LDS with coalesced loads, and
// gives every row (column) to a group of W lanes, lane f accumulating frame f
// sequentially in file element order with separate multiply and add: per
// frame, the result is bit-identical to the reference's sequential driver
// (mlem_seq/seqcsr4mlem.cpp:69-134).
//
// Per frame and iteration, kernels after the forward pass produce the FP64
// Poisson log-likelihood sum_{f_i != 0} (y_i log f_i - f_i) and sum_i f_i with
// a deterministic two-stage reduction; the host stops a frame when its budget
// is spent or the relative log-likelihood change drops below the tolerance,
// freezes it, and repacks the remaining frames into fewer / narrower blocks.
// An autotuned batch may also give a frame a block of its own (width 1): it
// then runs on the single-image kernels in the batch's mode (in EXACT mode
// the same sequential fp32 arithmetic per frame), so the few frames left at
// the end of a dynamic series do not pay for a whole block each.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "internal.hpp"
#include "mlem.h"
#include "spmm.hpp"

using namespace mlem;

namespace {

using spmm::kBlock;
using spmm::kTile;
constexpr unsigned kRing = 1024;  // per-iteration statistics ring

inline uint32_t blocks_for(uint64_t n) { return (uint32_t)std::max<uint64_t>(1, (n + kBlock - 1) / kBlock); }

void check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) raise(MLEM_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

// Deterministic per-frame FP64 reduction of (ll, tot), first stage: every
// workgroup sums its lane groups in a fixed order and stores one partial per
// frame. The second stage (k_batch_stats_reduce) is the next kernel in the
// stream, so no workgroup needs a device-scope fence or atomic.
template <int W>
__device__ void batch_stats_partial(double ll, double tot, double2* __restrict__ partials) {
  constexpr int G = kBlock / W;
  __shared__ double s_ll[kBlock], s_tot[kBlock];
  const uint32_t t = threadIdx.x, f = t % W;
  s_ll[t] = ll;
  s_tot[t] = tot;
  __syncthreads();
  if (t < (uint32_t)W) {
    double a = 0.0, b = 0.0;
    for (int k = 0; k < G; ++k) {
      a += s_ll[k * W + f];
      b += s_tot[k * W + f];
    }
    partials[(size_t)blockIdx.x * W + f] = make_double2(a, b);
  }
}

// Second stage, one 1024-lane workgroup: per frame, the sum of the n
// workgroup partials in a fixed order (lane group g takes workgroups g,
// g + G, ... into two alternating accumulators), appended to the ring at
// slot counter[1].
constexpr int kReduceLanes = 1024;
template <int W>
__global__ __launch_bounds__(kReduceLanes) void k_batch_stats_reduce(
    const double2* __restrict__ partials, unsigned n, unsigned* __restrict__ counter,
    double2* __restrict__ results, unsigned ring) {
  constexpr int G = kReduceLanes / W;
  __shared__ double s_ll[kReduceLanes], s_tot[kReduceLanes];
  const uint32_t t = threadIdx.x, f = t % W, g = t / W;
  double a0 = 0.0, b0 = 0.0, a1 = 0.0, b1 = 0.0;
  uint32_t blk = g;
  for (; blk + G < n; blk += 2 * G) {
    const double2 p = partials[(size_t)blk * W + f], q = partials[(size_t)(blk + G) * W + f];
    a0 += p.x;
    b0 += p.y;
    a1 += q.x;
    b1 += q.y;
  }
  if (blk < n) {
    const double2 p = partials[(size_t)blk * W + f];
    a0 += p.x;
    b0 += p.y;
  }
  s_ll[t] = a0 + a1;
  s_tot[t] = b0 + b1;
  __syncthreads();
  const unsigned slot = counter[1];
  if (t < (uint32_t)W) {
    double sa = 0.0, sb = 0.0;
    for (int k = 0; k < G; ++k) {
      sa += s_ll[k * W + f];
      sb += s_tot[k * W + f];
    }
    results[(size_t)(slot % ring) * W + f] = make_double2(sa, sb);
  }
  __syncthreads();  // every lane has read the slot before it advances
  if (t == 0) counter[1] = slot + 1;
}

// Forward epilogue: ratio r = (f != 0) ? y/f : 0, and f kept for the
// statistics kernel. The FP64 log-likelihood is not computed here: its
// registers would cap the projection at 5 waves per SIMD instead of 7.
template <int W>
struct FwdRatio {
  const int32_t* __restrict__ Y;
  float* __restrict__ R;
  float* __restrict__ Fp;
  __device__ void operator()(uint32_t row, uint32_t f, float F) {
    const size_t i = (size_t)row * W + f;
    const int32_t yi = Y[i];
    R[i] = (F != 0.f) ? div_rn((float)yi, F) : 0.f;
    Fp[i] = F;
  }
};

// Backward epilogue: x *= (n != 0) ? u/n : u for live lanes (bit set in
// the block's mask); finished frames keep their image.
template <int W>
struct BwdUpdate {
  const float* __restrict__ norm;
  float* __restrict__ X;
  uint32_t mask;
  __device__ void operator()(uint32_t col, uint32_t f, float u) {
    if (!((mask >> f) & 1u)) return;
    const float n = norm[col];
    const size_t i = (size_t)col * W + f;
    X[i] = X[i] * ((n != 0.f) ? div_rn(u, n) : u);
  }
};

// Forward projection + ratio for one RHS block (f kept in Fp).
template <int W>
__global__ __launch_bounds__(kBlock) void k_fwd_batch(
    const uint32_t* __restrict__ rowptr, const uint2* __restrict__ csr,
    const uint2* __restrict__ tiles, uint32_t ntiles,
    const float* __restrict__ X, const int32_t* __restrict__ Y, float* __restrict__ R,
    float* __restrict__ Fp) {
  FwdRatio<W> epi{Y, R, Fp};
  spmm::walk<W>(rowptr, csr, tiles, ntiles, X, epi);
}

// Per-frame FP64 statistics of a block's forward projection: lane f of row
// group g sums, over rows g, g + groups, ..., tot += f_i and, where f_i > 0,
// ll += y_i log f_i - f_i; one partial per workgroup and frame.
template <int W>
__global__ __launch_bounds__(kBlock) void k_batch_loglik(const float* __restrict__ Fp,
                                                         const int32_t* __restrict__ Y,
                                                         uint32_t rows,
                                                         double2* __restrict__ partials) {
  constexpr uint32_t G = kBlock / W;
  const uint32_t f = threadIdx.x % W, g = threadIdx.x / W;
  const uint32_t stride = gridDim.x * G;
  double ll = 0.0, tot = 0.0;
  for (uint32_t row = blockIdx.x * G + g; row < rows; row += stride) {
    const size_t i = (size_t)row * W + f;
    const float F = Fp[i];
    const int32_t yi = Y[i];
    tot += (double)F;
    if (F > 0.f) ll += (yi != 0 ? (double)yi * log((double)F) : 0.0) - (double)F;
  }
  batch_stats_partial<W>(ll, tot, partials);
}

// Backprojection + update for one RHS block over the CSC column tiles.
template <int W>
__global__ __launch_bounds__(kBlock) void k_bwd_batch(
    const uint32_t* __restrict__ colptr, const uint2* __restrict__ csc,
    const uint2* __restrict__ tiles, uint32_t ntiles,
    const float* __restrict__ R, const float* __restrict__ norm, float* __restrict__ X,
    const uint32_t* __restrict__ mask) {
  BwdUpdate<W> epi{norm, X, *mask};
  spmm::walk<W>(colptr, csc, tiles, ntiles, R, epi);
}

// The tiled (FAST) epilogues give each lane four right-hand sides (16-byte
// loads and stores): W/4 lanes per row or voxel.
template <int W>
struct Quads {
  static_assert(W % 4 == 0, "right-hand sides come in quads");
  static constexpr uint32_t L = W / 4;
  static constexpr uint32_t G = kBlock / L;
};

// batch_stats_partial for lanes that carry four frames each (lane q of a
// group holds frames 4q .. 4q+3): one partial per workgroup and frame, the
// groups summed in a fixed order.
template <int W>
__device__ void batch_stats_partial4(const double (&ll)[4], const double (&tot)[4],
                                     double2* __restrict__ partials) {
  constexpr uint32_t L = Quads<W>::L, G = Quads<W>::G;
  __shared__ double s_ll[4 * kBlock], s_tot[4 * kBlock];
  const uint32_t t = threadIdx.x;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    s_ll[4 * t + j] = ll[j];
    s_tot[4 * t + j] = tot[j];
  }
  __syncthreads();
  if (t < (uint32_t)W) {
    double a = 0.0, b = 0.0;
    for (uint32_t g = 0; g < G; ++g) {
      const uint32_t i = 4 * (g * L + t / 4) + t % 4;
      a += s_ll[i];
      b += s_tot[i];
    }
    partials[(size_t)blockIdx.x * W + t] = make_double2(a, b);
  }
}

__device__ __forceinline__ void add4(float4& a, const float4& b) {
  a.x += b.x;
  a.y += b.y;
  a.z += b.z;
  a.w += b.w;
}

__device__ __forceinline__ float ratio1(int32_t y, float f) { return (f != 0.f) ? div_rn((float)y, f) : 0.f; }

__device__ __forceinline__ void stat1(int32_t y, float f, double& ll, double& tot) {
  tot += (double)f;
  if (f > 0.f) ll += (y != 0 ? (double)y * log((double)f) : 0.0) - (double)f;
}

// Epilogue of the tiled forward product (FAST mode): a row's piece sums
// added in box order (read through rowpiece where the forward writes them
// box by box), then ratio r = (f != 0) ? y/f : 0 and the frames'
// FP64 statistics (row groups g, g + groups, ...); one partial per workgroup
// and frame.
// (64 VGPRs: the grid of 8 workgroups per CU is resident at once)
template <int W>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8))) void k_tiled_ratio(const uint32_t* __restrict__ rowslot,
                                                        const uint32_t* __restrict__ rowpiece,
                                                        const float* __restrict__ part,
                                                        const int32_t* __restrict__ Y,
                                                        float* __restrict__ R, uint32_t rows,
                                                        double2* __restrict__ partials) {
  constexpr uint32_t L = Quads<W>::L, G = Quads<W>::G;
  constexpr bool kByBox = panel_slots_by_box(W);  // slots gathered through rowpiece
  const uint32_t q = threadIdx.x % L, g = threadIdx.x / L;
  const uint32_t stride = gridDim.x * G;
  const float4* __restrict__ p4 = reinterpret_cast<const float4*>(part);
  const int4* __restrict__ y4 = reinterpret_cast<const int4*>(Y);
  float4* __restrict__ r4 = reinterpret_cast<float4*>(R);
  auto slot = [&](uint32_t k) { return kByBox ? rowpiece[k] : k; };
  double ll[4] = {0.0, 0.0, 0.0, 0.0}, tot[4] = {0.0, 0.0, 0.0, 0.0};
  for (uint32_t row = blockIdx.x * G + g; row < rows; row += stride) {
    uint32_t s = rowslot[row];
    const uint32_t e = rowslot[row + 1];
    float4 F = make_float4(0.f, 0.f, 0.f, 0.f);
    for (; s + 2 <= e; s += 2) {
      const float4 a0 = p4[(size_t)slot(s) * L + q], a1 = p4[(size_t)slot(s + 1) * L + q];
      add4(F, a0);
      add4(F, a1);
    }
    if (s < e) add4(F, p4[(size_t)slot(s) * L + q]);
    const size_t i = (size_t)row * L + q;
    const int4 y = y4[i];
    r4[i] = make_float4(ratio1(y.x, F.x), ratio1(y.y, F.y), ratio1(y.z, F.z), ratio1(y.w, F.w));
    stat1(y.x, F.x, ll[0], tot[0]);
    stat1(y.y, F.y, ll[1], tot[1]);
    stat1(y.z, F.z, ll[2], tot[2]);
    stat1(y.w, F.w, ll[3], tot[3]);
  }
  batch_stats_partial4<W>(ll, tot, partials);
}

__device__ __forceinline__ float update1(float x, float u, float n) { return x * ((n != 0.f) ? div_rn(u, n) : u); }

// Epilogue of the tiled backprojection (FAST mode): u = the voxel's group
// slots added in chunk order, then x *= (n != 0) ? u/n : u for live lanes.
// Voxels in panel order.
template <int W>
__global__ __launch_bounds__(kBlock) void k_tiled_update(const uint32_t* __restrict__ colslot,
                                                         const float* __restrict__ part,
                                                         const float* __restrict__ norm,
                                                         float* __restrict__ X,
                                                         const uint32_t* __restrict__ mask,
                                                         uint32_t cols) {
  constexpr uint32_t L = Quads<W>::L, G = Quads<W>::G;
  const uint32_t q = threadIdx.x % L, g = threadIdx.x / L;
  const uint32_t live = (*mask >> (4 * q)) & 0xFu;
  if (!live) return;
  const uint32_t stride = gridDim.x * G;
  const float4* __restrict__ p4 = reinterpret_cast<const float4*>(part);
  float4* __restrict__ x4 = reinterpret_cast<float4*>(X);
  for (uint32_t c = blockIdx.x * G + g; c < cols; c += stride) {
    uint32_t s = colslot[c];
    const uint32_t e = colslot[c + 1];
    float4 u = make_float4(0.f, 0.f, 0.f, 0.f);
    for (; s + 4 <= e; s += 4) {
      const float4 a0 = p4[(size_t)s * L + q], a1 = p4[(size_t)(s + 1) * L + q];
      const float4 a2 = p4[(size_t)(s + 2) * L + q], a3 = p4[(size_t)(s + 3) * L + q];
      add4(u, a0);
      add4(u, a1);
      add4(u, a2);
      add4(u, a3);
    }
    for (; s < e; ++s) add4(u, p4[(size_t)s * L + q]);
    const float n = norm[c];
    const size_t i = (size_t)c * L + q;
    float4 x = x4[i];
    if (live & 1u) x.x = update1(x.x, u.x, n);
    if (live & 2u) x.y = update1(x.y, u.y, n);
    if (live & 4u) x.z = update1(x.z, u.z, n);
    if (live & 8u) x.w = update1(x.w, u.w, n);
    x4[i] = x;
  }
}

// out[c*W + lane] = in[src(c)] for one frame: file-order image <-> block lane.
__global__ __launch_bounds__(kBlock) void k_lane_in(float* __restrict__ X, int W, int lane,
                                                    const float* __restrict__ file_img,
                                                    const uint32_t* __restrict__ colmap,
                                                    uint32_t cols) {
  const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= cols) return;
  const uint32_t d = colmap ? colmap[c] : c;
  X[(size_t)d * W + lane] = file_img[c];
}

__global__ __launch_bounds__(kBlock) void k_lane_out(const float* __restrict__ X, int W, int lane,
                                                     float* __restrict__ file_img,
                                                     const uint32_t* __restrict__ colmap,
                                                     uint32_t cols) {
  const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= cols) return;
  const uint32_t d = colmap ? colmap[c] : c;
  file_img[c] = X[(size_t)d * W + lane];
}

__global__ __launch_bounds__(kBlock) void k_lane_fill(float* __restrict__ X, int W, int lane,
                                                      float v, uint32_t cols) {
  const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
  if (c < cols) X[(size_t)c * W + lane] = v;
}

__global__ __launch_bounds__(kBlock) void k_lane_y(int32_t* __restrict__ Y, int W, int lane,
                                                   const int32_t* __restrict__ file_y,
                                                   const uint32_t* __restrict__ rowold,
                                                   uint32_t rows) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= rows) return;
  Y[(size_t)i * W + lane] = file_y[rowold ? rowold[i] : i];
}

__global__ __launch_bounds__(kBlock) void k_lane_move(const float* __restrict__ src, int Ws,
                                                      int ls, float* __restrict__ dst, int Wd,
                                                      int ld, uint32_t n) {
  const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
  if (c < n) dst[(size_t)c * Wd + ld] = src[(size_t)c * Ws + ls];
}

// A frame moved between blocks whose columns are in different orders:
// dst[map[c]*Wd + ld] = src[c*Ws + ls].
__global__ __launch_bounds__(kBlock) void k_lane_move_map(const float* __restrict__ src, int Ws,
                                                          int ls, float* __restrict__ dst, int Wd,
                                                          int ld, const uint32_t* __restrict__ map,
                                                          uint32_t n) {
  const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
  if (c < n) dst[(size_t)map[c] * Wd + ld] = src[(size_t)c * Ws + ls];
}

__global__ __launch_bounds__(kBlock) void k_lane_move_i32(const int32_t* __restrict__ src, int Ws,
                                                          int ls, int32_t* __restrict__ dst,
                                                          int Wd, int ld, uint32_t n) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) dst[(size_t)i * Wd + ld] = src[(size_t)i * Ws + ls];
}

template <class T>
void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

struct RhsBlock {
  int W = 0;
  std::vector<int> frames;  // frame in each lane; -1 = idle lane
  float* X = nullptr;       // [cols][W], device voxel order
  int32_t* Y = nullptr;     // [rows][W], device row order
  float* R = nullptr;       // [rows][W]
  float* F = nullptr;       // [rows][W] forward projections (statistics; W > 1)
  uint32_t* mask = nullptr; // bit f set: lane f still iterating
  double2* partials = nullptr;
  unsigned pmax = 0;           // partials held, in workgroups (width 1)
  unsigned* counter = nullptr;
  double2* results = nullptr;  // [kRing][W]
  unsigned iters = 0;          // iterations run by this block (ring slots used)
  ~RhsBlock() {
    dfree(X);
    dfree(Y);
    dfree(R);
    dfree(F);
    dfree(mask);
    dfree(partials);
    dfree(counter);
    dfree(results);
  }
};

}  // namespace

struct mlem_batch {
  mlem_matrix* mat = nullptr;
  uint32_t F = 0;
  int wmax = 32;
  // MLEM_MODE_FAST: blocks of width > 1 run through the matrix's tiling P,
  // their columns in panel order, partial sums in fpart / upart (sized for
  // wmax); MLEM_MODE_EXACT: per-frame sequential sums over the CSR / CSC
  int mode = MLEM_MODE_FAST;
  const PanelFormat* P = nullptr;
  float* fpart = nullptr;
  float* upart = nullptr;
  // Tiled blocks are independent within an iteration: they alternate between
  // the matrix's stream (partials fpart / upart) and a second stream `side`
  // (partials fpart2 / upart2), forked and joined by events, so that one
  // block's bandwidth-bound epilogues overlap another block's LDS-bound
  // products (a captured iteration holds the two branches). Lone frames run
  // on the matrix's stream.
  float* fpart2 = nullptr;
  float* upart2 = nullptr;
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  // measured cost (ms) of one block-iteration at widths 4, 8, 16, 32
  // (autotuned batches; 0 = not measured)
  double wcost[4] = {0, 0, 0, 0};
  // measured cost (ms) of one iteration of a lone frame on the single-image
  // EXACT kernels (autotuned batches; 0 = never used)
  double wcost1 = 0.0;
  uint64_t graph_gen = 0;  // matrix launch generation the graph was captured at
  double tol = 0.0;
  bool use_graph = false;
  std::vector<int> budget, done;
  std::vector<std::vector<double>> ll;   // per frame: log-likelihood of x_{k-1}, k = 1..done
  std::vector<char> active;
  std::vector<std::unique_ptr<RhsBlock>> blocks;
  std::vector<std::unique_ptr<RhsBlock>> spare;  // blocks kept for reuse by repacking
  float* final_x = nullptr;    // [F][cols] file order: images of finished frames
  int32_t* y_file = nullptr;   // [F][rows] file order (kept for repacking)
  uint2* tiles_f = nullptr;
  uint2* tiles_b = nullptr;
  uint32_t nt_f = 0, nt_b = 0;
  uint32_t grid = 0;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  uint64_t frame_iterations = 0;
  ~mlem_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    blocks.clear();
    dfree(final_x);
    dfree(y_file);
    dfree(tiles_f);
    dfree(tiles_b);
    dfree(fpart);
    dfree(upart);
    dfree(fpart2);
    dfree(upart2);
    if (side) {
      (void)hipStreamSynchronize(side);
      (void)hipStreamDestroy(side);
    }
    if (fork) (void)hipEventDestroy(fork);
    if (join) (void)hipEventDestroy(join);
  }
};

namespace {

void drop_graph(mlem_batch* B) {
  if (B->exec) (void)hipGraphExecDestroy(B->exec);
  if (B->graph) (void)hipGraphDestroy(B->graph);
  B->exec = nullptr;
  B->graph = nullptr;
}

std::unique_ptr<RhsBlock> new_block(mlem_batch* B, int W) {
  const DevMatrix& m = B->mat->m;
  std::unique_ptr<RhsBlock> b(new RhsBlock());
  b->W = W;
  b->frames.assign(W, -1);
  MLEM_HIP(hipMalloc(&b->X, 4ull * m.cols * W));
  MLEM_HIP(hipMalloc(&b->Y, 4ull * m.rows * W));
  MLEM_HIP(hipMalloc(&b->R, 4ull * m.rows * W));
  if (W > 1 && B->mode == MLEM_MODE_EXACT) MLEM_HIP(hipMalloc(&b->F, 4ull * m.rows * W));
  MLEM_HIP(hipMalloc(&b->mask, 4));
  // width 1 runs the single-image forward kernel, whose grid follows the
  // matrix's launch configuration
  b->pmax = W == 1 ? std::max(B->grid, stats_max_blocks(m)) : B->grid * W;
  MLEM_HIP(hipMalloc(&b->partials, sizeof(double2) * (size_t)b->pmax));
  MLEM_HIP(hipMalloc(&b->counter, 2 * sizeof(unsigned)));
  MLEM_HIP(hipMalloc(&b->results, sizeof(double2) * (size_t)kRing * W));
  MLEM_HIP(hipMemsetAsync(b->counter, 0, 2 * sizeof(unsigned), m.stream));
  MLEM_HIP(hipMemsetAsync(b->X, 0, 4ull * m.cols * W, m.stream));
  MLEM_HIP(hipMemsetAsync(b->Y, 0, 4ull * m.rows * W, m.stream));
  return b;
}

// A block of width W for repacking: a spare one when the batch holds one
// (idle lanes may keep an earlier frame's values; the lane mask keeps the
// update off them and their statistics are never read), else a new one.
// Repacking then costs no device allocation or free per finished frame.
std::unique_ptr<RhsBlock> take_block(mlem_batch* B, int W) {
  const DevMatrix& m = B->mat->m;
  for (size_t i = 0; i < B->spare.size(); ++i) {
    if (B->spare[i]->W != W) continue;
    std::unique_ptr<RhsBlock> b = std::move(B->spare[i]);
    B->spare.erase(B->spare.begin() + (std::ptrdiff_t)i);
    b->frames.assign(W, -1);
    b->iters = 0;
    MLEM_HIP(hipMemsetAsync(b->counter, 0, 2 * sizeof(unsigned), m.stream));
    const unsigned need = stats_max_blocks(m);
    if (W == 1 && need > b->pmax) {
      MLEM_HIP(hipStreamSynchronize(m.stream));
      dfree(b->partials);
      MLEM_HIP(hipMalloc(&b->partials, sizeof(double2) * (size_t)need));
      b->pmax = need;
    }
    return b;
  }
  return new_block(B, W);
}

void upload_mask(mlem_batch* B, RhsBlock& b) {
  uint32_t mask = 0;
  for (int l = 0; l < b.W; ++l)
    if (b.frames[l] >= 0 && B->active[b.frames[l]]) mask |= 1u << l;
  MLEM_HIP(hipMemcpyAsync(b.mask, &mask, 4, hipMemcpyHostToDevice, B->mat->m.stream));
  MLEM_HIP(hipStreamSynchronize(B->mat->m.stream));
}

// Columns of a block's image: panel order for the tiled (FAST) blocks,
// device order otherwise; the map takes file columns there.
inline bool tiled_block(const mlem_batch* B, int W) { return B->mode == MLEM_MODE_FAST && W > 1; }
inline const uint32_t* file_to_block(const mlem_batch* B, int W) {
  return tiled_block(B, W) ? B->P->pc_of_file : B->mat->m.colmap;
}

// set 0: partials fpart / upart; set 1: fpart2 / upart2 (the side stream's)
template <int W>
void launch_tiled_block(mlem_batch* B, RhsBlock& b, hipStream_t s, int set) {
  const DevMatrix& m = B->mat->m;
  const PanelFormat& P = *B->P;
  float* fpart = set ? B->fpart2 : B->fpart;
  float* upart = set ? B->upart2 : B->upart;
  launch_panel_fwd<W>(P, b.X, fpart, B->grid, s);
  hipLaunchKernelGGL((k_tiled_ratio<W>), dim3(B->grid), dim3(kBlock), 0, s, P.rowslot,
                     P.rowpiece, fpart,
                     b.Y, b.R, m.rows, b.partials);
  hipLaunchKernelGGL((k_batch_stats_reduce<W>), dim3(1), dim3(kReduceLanes), 0, s, b.partials,
                     B->grid, b.counter, b.results, kRing);
  launch_panel_bwd<W>(P, b.R, upart, B->grid, s);
  hipLaunchKernelGGL((k_tiled_update<W>), dim3(B->grid), dim3(kBlock), 0, s, P.colslot, upart,
                     P.norm, b.X, b.mask, m.cols);
}

template <int W>
void launch_block(mlem_batch* B, RhsBlock& b, hipStream_t s, int set = 0) {
  if (tiled_block(B, W)) {
    launch_tiled_block<W>(B, b, s, set);
    return;
  }
  const DevMatrix& m = B->mat->m;
  hipLaunchKernelGGL((k_fwd_batch<W>), dim3(B->grid), dim3(kBlock), 0, s, m.rowptr, m.csr,
                     B->tiles_f, B->nt_f, b.X, b.Y, b.R, b.F);
  hipLaunchKernelGGL((k_batch_loglik<W>), dim3(B->grid), dim3(kBlock), 0, s, b.F, b.Y, m.rows,
                     b.partials);
  hipLaunchKernelGGL((k_batch_stats_reduce<W>), dim3(1), dim3(kReduceLanes), 0, s, b.partials,
                     B->grid, b.counter, b.results, kRing);
  hipLaunchKernelGGL((k_bwd_batch<W>), dim3(B->grid), dim3(kBlock), 0, s, m.colptr, m.csc,
                     B->tiles_b, B->nt_b, b.R, m.norm, b.X, b.mask);
}

// A block of one frame: the single-image EXACT projections (forward with the
// fused ratio and the frame's FP64 statistics into the block's ring,
// backprojection with the fused update). X, Y and R of a width-1 block are
// the plain device-order vectors these kernels take.
void launch_single(mlem_batch* B, RhsBlock& b, hipStream_t s) {
  const DevMatrix& m = B->mat->m;
  StatsState st;
  st.partials = b.partials;
  st.counter = b.counter;
  st.results = b.results;
  st.ring = kRing;
  st.max_blocks = b.pmax;
  launch_fwd(m, B->mode, b.X, b.Y, b.R, true, &st, s);
  launch_bwd(m, B->mode, b.R, m.norm, b.X, true, s);
}

void launch_any(mlem_batch* B, RhsBlock& b, hipStream_t s, int set) {
  switch (b.W) {
    case 1: launch_single(B, b, s); break;
    case 4: launch_block<4>(B, b, s, set); break;
    case 8: launch_block<8>(B, b, s, set); break;
    case 16: launch_block<16>(B, b, s, set); break;
    case 32: launch_block<32>(B, b, s, set); break;
    default: raise(MLEM_EINVAL, "unsupported block width");
  }
}

void enqueue_iteration(mlem_batch* B, hipStream_t s) {
  size_t tiled = 0;
  for (auto& bp : B->blocks) tiled += tiled_block(B, bp->W) ? 1 : 0;
  if (!B->side || tiled < 2) {
    for (auto& bp : B->blocks) launch_any(B, *bp, s, 0);
    check_launch("batched iteration launch");
    return;
  }
  // fork: the side stream starts after everything before this iteration
  MLEM_HIP(hipEventRecord(B->fork, s));
  MLEM_HIP(hipStreamWaitEvent(B->side, B->fork, 0));
  size_t k = 0;
  for (auto& bp : B->blocks) {
    if (!tiled_block(B, bp->W)) {
      launch_any(B, *bp, s, 0);
      continue;
    }
    const int set = (int)(k++ % 2);
    launch_any(B, *bp, set ? B->side : s, set);
  }
  // join: the iteration ends when both streams have finished
  MLEM_HIP(hipEventRecord(B->join, B->side));
  MLEM_HIP(hipStreamWaitEvent(s, B->join, 0));
  check_launch("batched iteration launch");
}

// Measure one block-iteration (forward + backward) at every width on a
// scratch block (images 1, counts 0, all lanes live); median of 3.
void tune_widths(mlem_batch* B) {
  const DevMatrix& m = B->mat->m;
  hipEvent_t a, b;
  MLEM_HIP(hipEventCreate(&a));
  MLEM_HIP(hipEventCreate(&b));
  struct Ev {
    hipEvent_t a, b;
    ~Ev() {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
  } ev{a, b};
  static const int widths[5] = {1, 4, 8, 16, 32};
  for (int k = 0; k < 5; ++k) {
    const int W = widths[k];
    std::unique_ptr<RhsBlock> blk = new_block(B, W);
    for (int l = 0; l < W; ++l)
      hipLaunchKernelGGL(k_lane_fill, dim3(blocks_for(m.cols)), dim3(kBlock), 0, m.stream, blk->X,
                         W, l, 1.f, m.cols);
    const uint32_t all = W == 32 ? 0xFFFFFFFFu : ((1u << W) - 1u);
    MLEM_HIP(hipMemcpyAsync(blk->mask, &all, 4, hipMemcpyHostToDevice, m.stream));
    std::vector<float> t;
    for (int r = 0; r < 4; ++r) {
      MLEM_HIP(hipEventRecord(a, m.stream));
      switch (W) {
        case 1: launch_single(B, *blk, m.stream); break;
        case 4: launch_block<4>(B, *blk, m.stream); break;
        case 8: launch_block<8>(B, *blk, m.stream); break;
        case 16: launch_block<16>(B, *blk, m.stream); break;
        default: launch_block<32>(B, *blk, m.stream); break;
      }
      MLEM_HIP(hipEventRecord(b, m.stream));
      MLEM_HIP(hipEventSynchronize(b));
      float ms = 0.f;
      MLEM_HIP(hipEventElapsedTime(&ms, a, b));
      if (r) t.push_back(ms);  // first launch warms up
    }
    check_launch("batch width tuning");
    std::sort(t.begin(), t.end());
    (W == 1 ? B->wcost1 : B->wcost[k - 1]) = t[t.size() / 2];
    B->spare.push_back(std::move(blk));
  }
}

// Block widths for A live frames, widest first. With measured costs: the
// blocks (each of width <= wmax) that hold A frames at the least summed
// block-iteration cost, e.g. one 32-wide and one 8-wide block for 40 frames
// rather than two 32-wide ones, or a few single-frame blocks once fewer
// frames are left than make a block pay (dynamic programming over the frames
// covered). Without: blocks of the narrowest power of two that holds them
// (at most wmax), or one single-frame block per frame when wmax is 1.
std::vector<int> choose_widths(const mlem_batch* B, int A) {
  std::vector<int> ws;
  if (A <= 0) return ws;
  if (B->wcost[0] > 0.0) {
    std::vector<double> best((size_t)A + 1, -1.0);
    std::vector<int> pick((size_t)A + 1, 0);
    best[0] = 0.0;
    for (int a = 1; a <= A; ++a) {
      for (int i = 0, W = 4; i < 4 && W <= B->wmax; ++i, W *= 2) {
        const double c = best[(size_t)std::max(0, a - W)] + B->wcost[i];
        if (best[(size_t)a] < 0.0 || c < best[(size_t)a]) {
          best[(size_t)a] = c;
          pick[(size_t)a] = W;
        }
      }
      if (B->wcost1 > 0.0 && best[(size_t)a - 1] + B->wcost1 < best[(size_t)a]) {
        best[(size_t)a] = best[(size_t)a - 1] + B->wcost1;
        pick[(size_t)a] = 1;
      }
    }
    for (int a = A; a > 0; a = std::max(0, a - pick[(size_t)a])) ws.push_back(pick[(size_t)a]);
    std::sort(ws.rbegin(), ws.rend());
    return ws;
  }
  if (B->wmax == 1) {
    ws.assign((size_t)A, 1);
    return ws;
  }
  int W = 4;
  while (W < B->wmax && W < A) W *= 2;
  ws.assign((size_t)((A + W - 1) / W), W);
  return ws;
}

// Bring the blocks to the widths chosen for the still-active frames;
// finished frames' images move to final_x. Blocks of a width the new layout
// keeps stay where they are with their frames in their lanes (those holding
// the most live frames first); only the frames of dropped blocks move, into
// free lanes of the kept or new blocks.
void repack(mlem_batch* B) {
  DevMatrix& m = B->mat->m;
  int A = 0;
  for (uint32_t f = 0; f < B->F; ++f) A += B->active[f] ? 1 : 0;
  const std::vector<int> ws = choose_widths(B, A);
  bool same = ws.size() == B->blocks.size();
  for (size_t i = 0; same && i < ws.size(); ++i) same = B->blocks[i]->W == ws[i];
  // the single-image kernels have no lane mask: a finished frame's block of
  // width 1 must go
  for (auto& b : B->blocks)
    if (b->W == 1 && b->frames[0] >= 0 && !B->active[b->frames[0]]) same = false;
  // save finished frames still living in a block; their lanes become idle
  for (auto& b : B->blocks)
    for (int l = 0; l < b->W; ++l) {
      const int f = b->frames[l];
      if (f >= 0 && !B->active[f]) {
        hipLaunchKernelGGL(k_lane_out, dim3(blocks_for(m.cols)), dim3(kBlock), 0, m.stream, b->X,
                           b->W, l, B->final_x + (size_t)f * m.cols, file_to_block(B, b->W),
                           m.cols);
        b->frames[l] = -1;
      }
    }
  check_launch("lane copy");
  if (same) {
    for (auto& b : B->blocks) upload_mask(B, *b);
    return;
  }
  drop_graph(B);
  auto live = [](const RhsBlock& b) {
    int n = 0;
    for (int f : b.frames) n += f >= 0 ? 1 : 0;
    return n;
  };
  std::vector<std::unique_ptr<RhsBlock>> old = std::move(B->blocks), nbk;
  B->blocks.clear();
  for (int W : ws) {
    int pick = -1;
    for (size_t j = 0; j < old.size(); ++j)
      if (old[j] && old[j]->W == W && (pick < 0 || live(*old[j]) > live(*old[(size_t)pick])))
        pick = (int)j;
    if (pick >= 0)
      nbk.push_back(std::move(old[(size_t)pick]));
    else
      nbk.push_back(take_block(B, W));
  }
  // frames of the dropped blocks into free lanes (blocks in width order)
  size_t bi = 0;
  int l = 0;
  for (auto& src : old) {
    if (!src) continue;
    for (int ls = 0; ls < src->W; ++ls) {
      const int f = src->frames[ls];
      if (f < 0) continue;
      while (bi < nbk.size() && (l == nbk[bi]->W || nbk[bi]->frames[l] >= 0)) {
        if (++l >= nbk[bi]->W) {
          ++bi;
          l = 0;
        }
      }
      if (bi == nbk.size()) raise(MLEM_ESTATE, "batched frame lost during repacking");
      RhsBlock& d = *nbk[bi];
      if (tiled_block(B, src->W) == tiled_block(B, d.W))
        hipLaunchKernelGGL(k_lane_move, dim3(blocks_for(m.cols)), dim3(kBlock), 0, m.stream,
                           src->X, src->W, ls, d.X, d.W, l, m.cols);
      else  // panel order <-> device order (a tiled block and a lone frame)
        hipLaunchKernelGGL(k_lane_move_map, dim3(blocks_for(m.cols)), dim3(kBlock), 0, m.stream,
                           src->X, src->W, ls, d.X, d.W, l,
                           tiled_block(B, src->W) ? B->P->dev_of_pc : B->P->pc_of_dev, m.cols);
      hipLaunchKernelGGL(k_lane_move_i32, dim3(blocks_for(m.rows)), dim3(kBlock), 0, m.stream,
                         src->Y, src->W, ls, d.Y, d.W, l, m.rows);
      d.frames[l] = f;
    }
  }
  check_launch("repack");
  MLEM_HIP(hipStreamSynchronize(m.stream));
  for (auto& b : old)
    if (b) B->spare.push_back(std::move(b));
  // a single-frame block left without a frame would still run its kernels
  for (auto& b : nbk)
    if (b->W == 1 && b->frames[0] < 0) B->spare.push_back(std::move(b));
  nbk.erase(std::remove(nbk.begin(), nbk.end(), nullptr), nbk.end());
  std::stable_sort(nbk.begin(), nbk.end(),
                   [](const std::unique_ptr<RhsBlock>& x, const std::unique_ptr<RhsBlock>& y) {
                     return x->W > y->W;
                   });
  B->blocks = std::move(nbk);
  for (auto& b : B->blocks) upload_mask(B, *b);
}

}  // namespace

extern "C" {

void mlem_batch_options_default(mlem_batch_options* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->rel_tol = 0.0;
  o->block_width = 32;
  o->use_graph = 1;
  o->mode = MLEM_MODE_FAST;
}

int mlem_batch_create(mlem_matrix* h, uint32_t frames, const int32_t* y, uint32_t rows,
                      const int32_t* budgets, const mlem_batch_options* opts, mlem_batch** out) {
  return guarded([&] {
    if (!h || !out || !y || frames == 0) raise(MLEM_EINVAL, "bad argument");
    *out = nullptr;
    DevMatrix& m = h->m;
    if (rows != m.rows) raise(MLEM_EINVAL, "sinogram length does not match the matrix rows");
    if (m.streamed) raise(MLEM_EINVAL, "batched reconstruction needs a matrix held in HBM");
    if (!m.csc) raise(MLEM_ESTATE, "batched reconstruction needs the CSC transpose");
    if (m.vk != kF32) raise(MLEM_EINVAL, "batched reconstruction needs fp32 matrix values");
    mlem_batch_options o;
    if (opts)
      o = *opts;
    else
      mlem_batch_options_default(&o);
    if (o.block_width != 0 && o.block_width != 1 && o.block_width != 4 && o.block_width != 8 &&
        o.block_width != 16 && o.block_width != 32)
      raise(MLEM_EINVAL, "block_width must be 0 (autotune), 1, 4, 8, 16 or 32");
    if (o.mode != MLEM_MODE_FAST && o.mode != MLEM_MODE_EXACT)
      raise(MLEM_EINVAL, "batch mode must be MLEM_MODE_FAST or MLEM_MODE_EXACT");
    const bool autotune = o.block_width == 0;
    std::unique_ptr<mlem_batch> B(new mlem_batch());
    B->mat = h;
    B->F = frames;
    B->wmax = autotune ? 32 : o.block_width;
    B->mode = o.mode;
    B->tol = o.rel_tol;
    B->use_graph = o.use_graph != 0;
    B->budget.assign(frames, 0x7FFFFFFF);
    if (budgets)
      for (uint32_t f = 0; f < frames; ++f) {
        if (budgets[f] < 0) raise(MLEM_EINVAL, "negative iteration budget");
        B->budget[f] = budgets[f];
      }
    B->done.assign(frames, 0);
    B->ll.assign(frames, {});
    B->active.assign(frames, 1);
    for (uint32_t f = 0; f < frames; ++f)
      if (B->budget[f] == 0) B->active[f] = 0;
    B->grid = (uint32_t)std::max(8, 8 * m.cus);
    B->graph_gen = m.launch_gen;
    if (B->mode == MLEM_MODE_FAST && B->wmax > 1) {
      // the tiling (built once per matrix) and one block's partial sums
      B->P = &ensure_panels(m);
      MLEM_HIP(hipMalloc(&B->fpart, 4ull * std::max<uint64_t>(B->P->nseg, 1) * B->wmax));
      MLEM_HIP(hipMalloc(&B->upart, 4ull * std::max<uint64_t>(B->P->ngroups, 1) * B->wmax));
      if (frames >= 8) {  // room for two tiled blocks: a second stream
        MLEM_HIP(hipMalloc(&B->fpart2, 4ull * std::max<uint64_t>(B->P->nseg, 1) * B->wmax));
        MLEM_HIP(hipMalloc(&B->upart2, 4ull * std::max<uint64_t>(B->P->ngroups, 1) * B->wmax));
        MLEM_HIP(hipStreamCreateWithFlags(&B->side, hipStreamNonBlocking));
        MLEM_HIP(hipEventCreateWithFlags(&B->fork, hipEventDisableTiming));
        MLEM_HIP(hipEventCreateWithFlags(&B->join, hipEventDisableTiming));
      }
    }
    // tiles (device-order pointers are on the host after open)
    prepare_tiles(m, m.stream);
    const std::vector<uint32_t> tf = spmm::partition(m.h_rowptr, kTile),
                                tb = spmm::partition(m.h_colptr, kTile);
    std::vector<uint2> df(tf.size()), db(tb.size());
    for (size_t i = 0; i < tf.size(); ++i) df[i] = make_uint2(tf[i], m.h_rowptr[tf[i]]);
    for (size_t i = 0; i < tb.size(); ++i) db[i] = make_uint2(tb[i], m.h_colptr[tb[i]]);
    B->nt_f = (uint32_t)tf.size() - 1;
    B->nt_b = (uint32_t)tb.size() - 1;
    MLEM_HIP(hipMalloc(&B->tiles_f, sizeof(uint2) * df.size()));
    MLEM_HIP(hipMalloc(&B->tiles_b, sizeof(uint2) * db.size()));
    MLEM_HIP(hipMemcpyAsync(B->tiles_f, df.data(), sizeof(uint2) * df.size(),
                            hipMemcpyHostToDevice, m.stream));
    MLEM_HIP(hipMemcpyAsync(B->tiles_b, db.data(), sizeof(uint2) * db.size(),
                            hipMemcpyHostToDevice, m.stream));
    MLEM_HIP(hipMalloc(&B->final_x, 4ull * m.cols * frames));
    MLEM_HIP(hipMalloc(&B->y_file, 4ull * m.rows * frames));
    MLEM_HIP(hipMemcpyAsync(B->y_file, y, 4ull * m.rows * frames, hipMemcpyHostToDevice,
                            m.stream));
    // x0 per frame: (float)(sum y / sum norm), initImage (seqcsr4mlem.cpp:52-66)
    std::vector<float> x0(frames);
    for (uint32_t f = 0; f < frames; ++f) {
      double s = 0.0;
      const int32_t* yf = y + (size_t)f * rows;
      for (uint32_t i = 0; i < rows; ++i) s += yf[i];
      x0[f] = (float)(s / m.sumnorm);
      hipLaunchKernelGGL(k_lane_fill, dim3(blocks_for(m.cols)), dim3(kBlock), 0, m.stream,
                         B->final_x + (size_t)f * m.cols, 1, 0, x0[f], m.cols);
    }
    if (autotune) tune_widths(B.get());
    // initial blocks: every frame with a budget
    int A = 0;
    for (uint32_t f = 0; f < frames; ++f) A += B->active[f];
    const std::vector<int> ws = choose_widths(B.get(), A);
    for (int W : ws) B->blocks.push_back(take_block(B.get(), W));
    size_t bi = 0;
    int l = 0;
    for (uint32_t f = 0; f < frames; ++f) {
      if (!B->active[f]) continue;
      if (l == B->blocks[bi]->W) {
        ++bi;
        l = 0;
      }
      RhsBlock& b = *B->blocks[bi];
      b.frames[l] = (int)f;
      hipLaunchKernelGGL(k_lane_fill, dim3(blocks_for(m.cols)), dim3(kBlock), 0, m.stream, b.X,
                         b.W, l, x0[f], m.cols);
      hipLaunchKernelGGL(k_lane_y, dim3(blocks_for(m.rows)), dim3(kBlock), 0, m.stream, b.Y, b.W,
                         l, B->y_file + (size_t)f * m.rows, m.rowold, m.rows);
      ++l;
    }
    check_launch("batch setup");
    for (auto& b : B->blocks) upload_mask(B.get(), *b);
    MLEM_HIP(hipStreamSynchronize(m.stream));
    *out = B.release();
  });
}

int mlem_batch_run(mlem_batch* B, int max_iterations, int* iterations_run) {
  return guarded([&] {
    if (!B || max_iterations < 0) raise(MLEM_EINVAL, "bad argument");
    DevMatrix& m = B->mat->m;
    if (B->graph_gen != m.launch_gen) {
      // the matrix's launch configuration changed since the last run: the
      // single-image kernels of width-1 blocks now use other tile buffers
      // and perhaps a larger statistics grid
      drop_graph(B);
      MLEM_HIP(hipStreamSynchronize(m.stream));
      const unsigned need = stats_max_blocks(m);
      for (auto& b : B->blocks)
        if (b->W == 1 && need > b->pmax) {
          dfree(b->partials);
          MLEM_HIP(hipMalloc(&b->partials, sizeof(double2) * (size_t)need));
          b->pmax = need;
        }
      B->graph_gen = m.launch_gen;
    }
    int it = 0;
    for (; it < max_iterations && !B->blocks.empty(); ++it) {
      if (B->use_graph) {
        if (!B->exec) {
          MLEM_HIP(hipStreamBeginCapture(m.stream, hipStreamCaptureModeThreadLocal));
          try {
            enqueue_iteration(B, m.stream);
          } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(m.stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
          }
          MLEM_HIP(hipStreamEndCapture(m.stream, &B->graph));
          MLEM_HIP(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0));
        }
        MLEM_HIP(hipGraphLaunch(B->exec, m.stream));
      } else {
        enqueue_iteration(B, m.stream);
      }
      // per-frame statistics of this iteration
      bool changed = false;
      std::vector<std::vector<double2>> res(B->blocks.size());
      for (size_t i = 0; i < B->blocks.size(); ++i) {
        const RhsBlock& b = *B->blocks[i];
        res[i].resize((size_t)b.W);
        MLEM_HIP(hipMemcpyAsync(res[i].data(), b.results + (size_t)(b.iters % kRing) * b.W,
                                sizeof(double2) * b.W, hipMemcpyDeviceToHost, m.stream));
      }
      MLEM_HIP(hipStreamSynchronize(m.stream));
      for (size_t i = 0; i < B->blocks.size(); ++i) {
        auto& b = B->blocks[i];
        const std::vector<double2>& r = res[i];
        b->iters++;
        for (int l = 0; l < b->W; ++l) {
          const int f = b->frames[l];
          if (f < 0 || !B->active[f]) continue;
          B->done[f]++;
          B->frame_iterations++;
          B->ll[f].push_back(r[l].x);
          const int k = B->done[f];
          bool stop = k >= B->budget[f];
          if (!stop && B->tol > 0.0 && k >= 2) {
            const double a = B->ll[f][k - 1], p = B->ll[f][k - 2];
            stop = std::fabs(a - p) <= B->tol * std::fabs(a);
          }
          if (stop) {
            B->active[f] = 0;
            changed = true;
          }
        }
      }
      if (changed) repack(B);
    }
    if (iterations_run) *iterations_run = it;
  });
}

int mlem_batch_image(mlem_batch* B, uint32_t frame, float* x, uint32_t cols) {
  return guarded([&] {
    if (!B || !x || frame >= B->F) raise(MLEM_EINVAL, "bad argument");
    DevMatrix& m = B->mat->m;
    if (cols != m.cols) raise(MLEM_EINVAL, "image size does not match the matrix columns");
    float* src = B->final_x + (size_t)frame * m.cols;
    for (auto& b : B->blocks)
      for (int l = 0; l < b->W; ++l)
        if (b->frames[l] == (int)frame)
          hipLaunchKernelGGL(k_lane_out, dim3(blocks_for(m.cols)), dim3(kBlock), 0, m.stream,
                             b->X, b->W, l, src, file_to_block(B, b->W), m.cols);
    check_launch("image copy");
    MLEM_HIP(hipMemcpyAsync(x, src, 4ull * cols, hipMemcpyDeviceToHost, m.stream));
    MLEM_HIP(hipStreamSynchronize(m.stream));
  });
}

int mlem_batch_frame_info(mlem_batch* B, uint32_t frame, int* iterations, double* loglik,
                          int capacity) {
  return guarded([&] {
    if (!B || frame >= B->F) raise(MLEM_EINVAL, "bad argument");
    if (iterations) *iterations = B->done[frame];
    if (loglik)
      for (int k = 0; k < capacity && k < (int)B->ll[frame].size(); ++k) loglik[k] = B->ll[frame][k];
  });
}

int mlem_batch_progress(const mlem_batch* B, uint32_t* active_frames, uint64_t* frame_iterations,
                        int* block_width, uint32_t* blocks) {
  return guarded([&] {
    if (!B) raise(MLEM_EINVAL, "null handle");
    uint32_t a = 0;
    for (char c : B->active) a += c ? 1 : 0;
    if (active_frames) *active_frames = a;
    if (frame_iterations) *frame_iterations = B->frame_iterations;
    if (block_width) *block_width = B->blocks.empty() ? 0 : B->blocks[0]->W;
    if (blocks) *blocks = (uint32_t)B->blocks.size();
  });
}

int mlem_batch_tuning(const mlem_batch* B, double block_ms[4]) {
  return guarded([&] {
    if (!B || !block_ms) raise(MLEM_EINVAL, "null argument");
    for (int i = 0; i < 4; ++i) block_ms[i] = B->wcost[i];
  });
}

int mlem_batch_single_ms(const mlem_batch* B, double* ms) {
  return guarded([&] {
    if (!B || !ms) raise(MLEM_EINVAL, "null argument");
    *ms = B->wcost1;
  });
}

void mlem_batch_destroy(mlem_batch* B) {
  if (!B) return;
  (void)hipStreamSynchronize(B->mat->m.stream);
  delete B;
}

}  // extern "C"
