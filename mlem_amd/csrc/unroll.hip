// unroll.hip — differentiable unrolled MLEM on gfx950.
//
// Forward: K MLEM iterations over a batch of sinograms with real-valued
// counts y (the reference's iteration, mlem_seq/seqcsr4mlem.cpp:137-189,
// with y relaxed to float), from given initial images x0:
//   f = A x_k,  r = (f != 0) ? y/f : 0,  u = A^T r,  x_{k+1} = x_k (u/n)
// (u instead of u/n where n = 0). The iterate is kept every C iterations
// (the tape); nothing else is stored.
//
// Backward (reverse mode) from g = dL/dx_K, one checkpoint segment at a time
// from the last: the segment's iterations are recomputed from its
// checkpoint (storing x_k, f_k, r_k, u_k for those C iterations only), then
// for k from the segment's end down to its start
//   gu = g x_k/n,  g <- g u_k/n               (elementwise, n = 0: drop /n)
//   gr = A gu:      dL/dy += gr/f_k,  gf = -gr r_k/f_k   (0 where f_k = 0)
//   g += A^T gf
// and dL/dx0 = g. The batch's frames are lanes of one block (W = 4..32), so
// each matrix element read from HBM serves every frame. MLEM_MODE_EXACT:
// every product is the RHS-block SpMM of spmm.hpp, each frame's sums
// sequential fp32 (deterministic; the same for any batch composition; for
// integer counts the forward pass is bit-identical to mlem()).
// MLEM_MODE_FAST: every product runs through the matrix's box tiling
// (panel.hip): the forward's sums per (row, box) piece are added in a fixed
// order by k_slot_sums before the epilogue, the backward's sums are complete
// per voxel (one slot each); column vectors are then held in panel order.
// Deterministic for a given block width.
//
// Memory per block of W frames: ceil(K/C) checkpoints of cols*W floats plus
// C iterations of (2 rows + 2 cols)*W floats; C = round(sqrt(K)) by default.
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

inline uint32_t blocks_for(uint64_t n) {
  return (uint32_t)std::max<uint64_t>(1, (n + kBlock - 1) / kBlock);
}

void check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) raise(MLEM_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

// ---- SpMM epilogues ------------------------------------------------------
// Forward step: F = f, R = (f != 0) ? y/f : 0.
template <int W>
struct StepFwd {
  const float* __restrict__ Y;
  float* __restrict__ F;
  float* __restrict__ R;
  __device__ void operator()(uint32_t row, uint32_t l, float f) {
    const size_t i = (size_t)row * W + l;
    F[i] = f;
    R[i] = (f != 0.f) ? div_rn(Y[i], f) : 0.f;
  }
};

// Backward step: U = u (if kept), Xn = X (n != 0 ? u/n : u).
template <int W>
struct StepBwd {
  const float* __restrict__ norm;
  const float* __restrict__ X;
  float* __restrict__ Xn;
  float* __restrict__ U;
  __device__ void operator()(uint32_t col, uint32_t l, float u) {
    const size_t i = (size_t)col * W + l;
    const float n = norm[col];
    if (U) U[i] = u;
    Xn[i] = X[i] * ((n != 0.f) ? div_rn(u, n) : u);
  }
};

// Adjoint of the ratio: gr = A gu; GY += gr/f, GF = -gr r/f (0 where f = 0).
template <int W>
struct AdjFwd {
  const float* __restrict__ F;
  const float* __restrict__ R;
  float* __restrict__ GY;
  float* __restrict__ GF;
  __device__ void operator()(uint32_t row, uint32_t l, float gr) {
    const size_t i = (size_t)row * W + l;
    const float f = F[i];
    if (f != 0.f) {
      GY[i] = GY[i] + div_rn(gr, f);
      GF[i] = div_rn(-(gr * R[i]), f);
    } else {
      GF[i] = 0.f;
    }
  }
};

// Adjoint of the backprojection: G += A^T gf.
template <int W>
struct AdjBwd {
  float* __restrict__ G;
  __device__ void operator()(uint32_t col, uint32_t l, float h) {
    const size_t i = (size_t)col * W + l;
    G[i] = G[i] + h;
  }
};

template <int W, class Epi>
__global__ __launch_bounds__(kBlock) void k_spmm(const uint32_t* __restrict__ ptr,
                                                 const uint2* __restrict__ el,
                                                 const uint2* __restrict__ tiles, uint32_t ntiles,
                                                 const float* __restrict__ v, Epi epi) {
  spmm::walk<W>(ptr, el, tiles, ntiles, v, epi);
}

// Fixed-order sums of a tiled product's partials: lane (i, l) adds
// part[s*W + l] over the slots s in [first[i], first[i+1]) (the slots of row
// i in panel order, or of panel-order column i in chunk order) and hands the
// sum to the epilogue.
template <int W, class Epi>
__global__ __launch_bounds__(kBlock) void k_slot_sums(const uint32_t* __restrict__ first,
                                                      const uint32_t* __restrict__ index,
                                                      uint32_t n, const float* __restrict__ part,
                                                      Epi epi) {
  const uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x;
  if (i >= (uint64_t)n * W) return;
  const uint32_t r = (uint32_t)(i / W), l = (uint32_t)(i % W);
  const uint32_t s0 = first[r], s1 = first[r + 1];
  float acc = 0.f;
  for (uint32_t k = s0; k < s1; ++k) acc = acc + part[(size_t)(index ? index[k] : k) * W + l];
  epi(r, l, acc);
}

// gu = g x/n, g <- g u/n (n = 0: without /n), over cols*W.
__global__ __launch_bounds__(kBlock) void k_adj_update(float* __restrict__ G,
                                                       float* __restrict__ GU,
                                                       const float* __restrict__ X,
                                                       const float* __restrict__ U,
                                                       const float* __restrict__ norm, int W,
                                                       uint64_t n) {
  const uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x;
  if (i >= n) return;
  const float nj = norm[i / (uint64_t)W];
  const float g = G[i];
  if (nj != 0.f) {
    GU[i] = g * div_rn(X[i], nj);
    G[i] = g * div_rn(U[i], nj);
  } else {
    GU[i] = g * X[i];
    G[i] = g * U[i];
  }
}

// File-order vector <-> lane of an interleaved block. Columns: device
// position map[c] of file column c (nullptr = identity). Rows: file row
// old[i] of device row i (nullptr = identity).
__global__ __launch_bounds__(kBlock) void k_cols_in(float* __restrict__ X, int W, int lane,
                                                    const float* __restrict__ file,
                                                    const uint32_t* __restrict__ map,
                                                    uint32_t n) {
  const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
  if (c < n) X[(size_t)(map ? map[c] : c) * W + lane] = file[c];
}

__global__ __launch_bounds__(kBlock) void k_cols_out(const float* __restrict__ X, int W, int lane,
                                                     float* __restrict__ file,
                                                     const uint32_t* __restrict__ map,
                                                     uint32_t n) {
  const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
  if (c < n) file[c] = X[(size_t)(map ? map[c] : c) * W + lane];
}

__global__ __launch_bounds__(kBlock) void k_rows_in(float* __restrict__ Y, int W, int lane,
                                                    const float* __restrict__ file,
                                                    const uint32_t* __restrict__ old,
                                                    uint32_t n) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) Y[(size_t)i * W + lane] = file[old ? old[i] : i];
}

__global__ __launch_bounds__(kBlock) void k_rows_out(const float* __restrict__ Y, int W, int lane,
                                                     float* __restrict__ file,
                                                     const uint32_t* __restrict__ old,
                                                     uint32_t n) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) file[old ? old[i] : i] = Y[(size_t)i * W + lane];
}

__global__ __launch_bounds__(kBlock) void k_zero(float* __restrict__ p, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kBlock)
    p[i] = 0.f;
}

struct DevBuf {
  float* p = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  void alloc(uint64_t floats) { MLEM_HIP(hipMalloc(&p, 4ull * std::max<uint64_t>(floats, 1))); }
};

struct Block {
  int frames[32];   // frame of every lane, -1 = idle
  DevBuf Y;         // [rows][W]
  DevBuf ckpt;      // [nseg][cols][W]: x_{s*C}
  DevBuf xa, xb;    // [cols][W] forward ping-pong / x_K
  DevBuf sx;        // [C-1][cols][W]: recomputed iterates inside a segment
  DevBuf sf, sr;    // [C][rows][W]
  DevBuf su;        // [C][cols][W]
  DevBuf G, GU;     // [cols][W]
  DevBuf GF, GY;    // [rows][W]
  float* xK = nullptr;  // xa or xb after the forward pass
};

}  // namespace

struct mlem_tape {
  mlem_matrix* mat = nullptr;
  uint32_t F = 0;
  int W = 4, K = 0, C = 1, nseg = 0;
  int mode = MLEM_MODE_EXACT;
  // MLEM_MODE_FAST: the matrix's tiling, partial-sum buffers (slots x W),
  // and the column order (panel order) and column sums the vectors use
  const PanelFormat* P = nullptr;
  float* fpart = nullptr;
  float* upart = nullptr;
  const uint32_t* colmap = nullptr;
  const float* norm = nullptr;
  uint32_t grid = 0;
  uint2* tiles_f = nullptr;
  uint2* tiles_b = nullptr;
  uint32_t nt_f = 0, nt_b = 0;
  std::vector<std::unique_ptr<Block>> blocks;
  bool forward_done = false;
  uint64_t spmm_launches = 0;
  int recomputed = 0;
  ~mlem_tape() {
    blocks.clear();
    if (tiles_f) (void)hipFree(tiles_f);
    if (tiles_b) (void)hipFree(tiles_b);
    if (fpart) (void)hipFree(fpart);
    if (upart) (void)hipFree(upart);
  }
};

namespace {

// One product with the block's W vectors v: A v (csc = false, epilogue per
// row) or A^T v (csc = true, epilogue per column).
template <int W, class Epi>
void spmm_launch(mlem_tape* T, bool csc, const float* v, const Epi& epi) {
  const DevMatrix& m = T->mat->m;
  if (T->mode == MLEM_MODE_FAST) {
    const PanelFormat& P = *T->P;
    const uint32_t n = csc ? m.cols : m.rows;
    float* part = csc ? T->upart : T->fpart;
    if (csc)
      launch_panel_bwd<W>(P, v, part, T->grid, m.stream);
    else
      launch_panel_fwd<W>(P, v, part, T->grid, m.stream);
    hipLaunchKernelGGL((k_slot_sums<W, Epi>), dim3(blocks_for((uint64_t)n * W)), dim3(kBlock), 0,
                       m.stream, csc ? P.colslot : P.rowslot,
                       (!csc && panel_slots_by_box(W)) ? P.rowpiece : nullptr, n, part, epi);
  } else if (csc) {
    hipLaunchKernelGGL((k_spmm<W, Epi>), dim3(T->grid), dim3(kBlock), 0, m.stream, m.colptr, m.csc,
                       T->tiles_b, T->nt_b, v, epi);
  } else {
    hipLaunchKernelGGL((k_spmm<W, Epi>), dim3(T->grid), dim3(kBlock), 0, m.stream, m.rowptr, m.csr,
                       T->tiles_f, T->nt_f, v, epi);
  }
  ++T->spmm_launches;
}

// One MLEM iteration of a block: x -> xn, keeping f, r (and u if U).
template <int W>
void step(mlem_tape* T, Block& b, const float* x, float* xn, float* F, float* R, float* U) {
  spmm_launch<W>(T, false, x, StepFwd<W>{b.Y.p, F, R});
  spmm_launch<W>(T, true, R, StepBwd<W>{T->norm, x, xn, U});
}

template <int W>
void forward_block(mlem_tape* T, Block& b) {
  const DevMatrix& m = T->mat->m;
  const uint64_t cw = (uint64_t)m.cols * W;
  const float* x = b.ckpt.p;  // x_0
  for (int k = 0; k < T->K; ++k) {
    float* xn;
    if ((k + 1) % T->C == 0 && k + 1 < T->K)
      xn = b.ckpt.p + (uint64_t)((k + 1) / T->C) * cw;
    else
      xn = (x == b.xa.p) ? b.xb.p : b.xa.p;
    // f and r of the forward pass are not kept: the segment buffers serve as
    // scratch
    step<W>(T, b, x, xn, b.sf.p, b.sr.p, nullptr);
    x = xn;
  }
  b.xK = const_cast<float*>(x);
  check_launch("unrolled forward launch");
}

template <int W>
void backward_block(mlem_tape* T, Block& b) {
  const DevMatrix& m = T->mat->m;
  const uint64_t cw = (uint64_t)m.cols * W, rw = (uint64_t)m.rows * W;
  hipLaunchKernelGGL(k_zero, dim3(std::min<uint32_t>(blocks_for(rw), 4096)), dim3(kBlock), 0,
                     m.stream, b.GY.p, rw);
  for (int seg = T->nseg - 1; seg >= 0; --seg) {
    const int s = seg * T->C, e = std::min(s + T->C, T->K);
    // recompute the segment, keeping x_k, f_k, r_k, u_k
    std::vector<const float*> xs(e - s);
    xs[0] = b.ckpt.p + (uint64_t)seg * cw;
    for (int k = s; k < e; ++k) {
      const int i = k - s;
      float* xn = (k + 1 < e) ? b.sx.p + (uint64_t)i * cw : b.xa.p;  // x_e is not needed
      step<W>(T, b, xs[i], xn, b.sf.p + (uint64_t)i * rw, b.sr.p + (uint64_t)i * rw,
              b.su.p + (uint64_t)i * cw);
      if (k + 1 < e) xs[i + 1] = xn;
    }
    T->recomputed += e - s;
    for (int k = e - 1; k >= s; --k) {
      const int i = k - s;
      hipLaunchKernelGGL(k_adj_update, dim3(blocks_for(cw)), dim3(kBlock), 0, m.stream, b.G.p,
                         b.GU.p, xs[i], b.su.p + (uint64_t)i * cw, T->norm, W, cw);
      spmm_launch<W>(T, false, b.GU.p,
                     AdjFwd<W>{b.sf.p + (uint64_t)i * rw, b.sr.p + (uint64_t)i * rw, b.GY.p,
                               b.GF.p});
      spmm_launch<W>(T, true, b.GF.p, AdjBwd<W>{b.G.p});
    }
  }
  check_launch("unrolled backward launch");
}

template <class Fn>
void dispatch(int W, Fn&& fn) {
  switch (W) {
    case 4: fn(std::integral_constant<int, 4>()); break;
    case 8: fn(std::integral_constant<int, 8>()); break;
    case 16: fn(std::integral_constant<int, 16>()); break;
    case 32: fn(std::integral_constant<int, 32>()); break;
    default: raise(MLEM_EINVAL, "unsupported block width");
  }
}

}  // namespace

extern "C" {

int mlem_forward_k(mlem_matrix* h, uint32_t frames, const float* y, uint32_t rows,
                   const float* x0, uint32_t cols, int K, int checkpoint_interval, float* xK,
                   mlem_tape** tape) {
  return mlem_forward_k_mode(h, frames, y, rows, x0, cols, K, checkpoint_interval,
                             MLEM_MODE_EXACT, xK, tape);
}

int mlem_forward_k_mode(mlem_matrix* h, uint32_t frames, const float* y, uint32_t rows,
                        const float* x0, uint32_t cols, int K, int checkpoint_interval, int mode,
                        float* xK, mlem_tape** tape) {
  return guarded([&] {
    if (tape) *tape = nullptr;
    if (!h || !y || frames == 0) raise(MLEM_EINVAL, "bad argument");
    DevMatrix& m = h->m;
    if (rows != m.rows || cols != m.cols)
      raise(MLEM_EINVAL, "vector lengths do not match the matrix dimensions");
    if (K < 1) raise(MLEM_EINVAL, "K must be at least 1");
    if (checkpoint_interval < 0) raise(MLEM_EINVAL, "negative checkpoint interval");
    if (m.streamed) raise(MLEM_EINVAL, "unrolled MLEM needs a matrix held in HBM");
    if (!m.csc) raise(MLEM_ESTATE, "unrolled MLEM needs the CSC transpose");
    if (m.vk != kF32) raise(MLEM_EINVAL, "unrolled MLEM needs fp32 matrix values");
    if (mode != MLEM_MODE_EXACT && mode != MLEM_MODE_FAST)
      raise(MLEM_EINVAL, "unrolled MLEM runs MLEM_MODE_EXACT or MLEM_MODE_FAST");
    std::unique_ptr<mlem_tape> T(new mlem_tape());
    T->mat = h;
    T->F = frames;
    T->mode = mode;
    T->K = K;
    T->C = checkpoint_interval > 0 ? std::min(checkpoint_interval, K)
                                   : std::max(1, (int)std::lround(std::sqrt((double)K)));
    T->nseg = (K + T->C - 1) / T->C;
    int W = 4;
    while (W < 32 && W < (int)frames) W *= 2;
    T->W = W;
    T->grid = (uint32_t)std::max(8, 8 * m.cus);
    T->colmap = m.colmap;
    T->norm = m.norm;
    if (mode == MLEM_MODE_FAST) {
      T->P = &ensure_panels(m);
      T->colmap = T->P->pc_of_file;
      T->norm = T->P->norm;
      MLEM_HIP(hipMalloc(&T->fpart, 4ull * std::max<uint64_t>(T->P->nseg, 1) * W));
      MLEM_HIP(hipMalloc(&T->upart, 4ull * std::max<uint64_t>(T->P->ngroups, 1) * W));
    }
    prepare_tiles(m, m.stream);
    if (mode == MLEM_MODE_EXACT) {  // record tiles of the sequential SpMM walker
      const std::vector<uint32_t> tf = spmm::partition(m.h_rowptr, kTile),
                                  tb = spmm::partition(m.h_colptr, kTile);
      std::vector<uint2> df(tf.size()), db(tb.size());
      for (size_t i = 0; i < tf.size(); ++i) df[i] = make_uint2(tf[i], m.h_rowptr[tf[i]]);
      for (size_t i = 0; i < tb.size(); ++i) db[i] = make_uint2(tb[i], m.h_colptr[tb[i]]);
      T->nt_f = (uint32_t)tf.size() - 1;
      T->nt_b = (uint32_t)tb.size() - 1;
      MLEM_HIP(hipMalloc(&T->tiles_f, sizeof(uint2) * df.size()));
      MLEM_HIP(hipMalloc(&T->tiles_b, sizeof(uint2) * db.size()));
      MLEM_HIP(hipMemcpyAsync(T->tiles_f, df.data(), sizeof(uint2) * df.size(),
                              hipMemcpyHostToDevice, m.stream));
      MLEM_HIP(hipMemcpyAsync(T->tiles_b, db.data(), sizeof(uint2) * db.size(),
                              hipMemcpyHostToDevice, m.stream));
    }

    const uint64_t cw = (uint64_t)m.cols * W, rw = (uint64_t)m.rows * W;
    DevBuf stage;  // one frame's file-order vector
    stage.alloc(std::max<uint64_t>(m.rows, m.cols));
    const size_t nb = (frames + W - 1) / W;
    for (size_t bi = 0; bi < nb; ++bi) {
      std::unique_ptr<Block> b(new Block());
      b->Y.alloc(rw);
      b->ckpt.alloc(cw * T->nseg);
      b->xa.alloc(cw);
      b->xb.alloc(cw);
      b->sf.alloc(rw * T->C);
      b->sr.alloc(rw * T->C);
      if (tape) {
        b->sx.alloc(cw * std::max(1, T->C - 1));
        b->su.alloc(cw * T->C);
        b->G.alloc(cw);
        b->GU.alloc(cw);
        b->GF.alloc(rw);
        b->GY.alloc(rw);
      }
      hipLaunchKernelGGL(k_zero, dim3(std::min<uint32_t>(blocks_for(rw), 4096)), dim3(kBlock), 0,
                         m.stream, b->Y.p, rw);
      hipLaunchKernelGGL(k_zero, dim3(std::min<uint32_t>(blocks_for(cw), 4096)), dim3(kBlock), 0,
                         m.stream, b->ckpt.p, cw);
      for (int l = 0; l < W; ++l) {
        const uint64_t f = bi * W + l;
        b->frames[l] = f < frames ? (int)f : -1;
        if (f >= frames) continue;
        const float* yf = y + f * rows;
        MLEM_HIP(hipMemcpyAsync(stage.p, yf, 4ull * rows, hipMemcpyHostToDevice, m.stream));
        hipLaunchKernelGGL(k_rows_in, dim3(blocks_for(rows)), dim3(kBlock), 0, m.stream, b->Y.p, W,
                           l, stage.p, m.rowold, rows);
        if (x0) {
          MLEM_HIP(hipMemcpyAsync(stage.p, x0 + f * cols, 4ull * cols, hipMemcpyHostToDevice,
                                  m.stream));
        } else {
          // x0 = (float)(sum y / sum n) with double sums (seqcsr4mlem.cpp:52-66),
          // y taken in eight interleaved partial sums (exact for integer
          // counts, as the reference's single running sum is)
          double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
          uint32_t i = 0;
          for (; i + 8 <= rows; i += 8)
            for (int j = 0; j < 8; ++j) acc[j] += yf[i + j];
          for (; i < rows; ++i) acc[0] += yf[i];
          const double sy = ((acc[0] + acc[1]) + (acc[2] + acc[3])) +
                            ((acc[4] + acc[5]) + (acc[6] + acc[7]));
          launch_fill(stage.p, (float)(sy / m.sumnorm), cols, m.stream);
        }
        // the staging buffer is reused by the next frame in stream order
        hipLaunchKernelGGL(k_cols_in, dim3(blocks_for(cols)), dim3(kBlock), 0, m.stream,
                           b->ckpt.p, W, l, stage.p, T->colmap, cols);
      }
      check_launch("unrolled input launch");
      T->blocks.push_back(std::move(b));
    }
    for (auto& b : T->blocks) dispatch(W, [&](auto w) { forward_block<decltype(w)::value>(T.get(), *b); });
    if (xK) {
      for (auto& b : T->blocks)
        for (int l = 0; l < W; ++l) {
          if (b->frames[l] < 0) continue;
          hipLaunchKernelGGL(k_cols_out, dim3(blocks_for(cols)), dim3(kBlock), 0, m.stream, b->xK,
                             W, l, stage.p, T->colmap, cols);
          MLEM_HIP(hipMemcpyAsync(xK + (size_t)b->frames[l] * cols, stage.p, 4ull * cols,
                                  hipMemcpyDeviceToHost, m.stream));
        }
    }
    MLEM_HIP(hipStreamSynchronize(m.stream));
    T->forward_done = true;
    if (tape) *tape = T.release();
  });
}

int mlem_backward_k(mlem_tape* T, const float* grad_xK, float* grad_y, float* grad_x0) {
  return guarded([&] {
    if (!T || !grad_xK) raise(MLEM_EINVAL, "bad argument");
    if (!T->forward_done) raise(MLEM_ESTATE, "forward pass missing");
    const DevMatrix& m = T->mat->m;
    const uint32_t rows = m.rows, cols = m.cols;
    const int W = T->W;
    const uint64_t cw = (uint64_t)cols * W;
    DevBuf stage;  // one frame's file-order vector, reused in stream order
    stage.alloc(std::max<uint64_t>(rows, cols));
    for (auto& b : T->blocks) {
      hipLaunchKernelGGL(k_zero, dim3(std::min<uint32_t>(blocks_for(cw), 4096)), dim3(kBlock), 0,
                         m.stream, b->G.p, cw);
      for (int l = 0; l < W; ++l) {
        if (b->frames[l] < 0) continue;
        MLEM_HIP(hipMemcpyAsync(stage.p, grad_xK + (size_t)b->frames[l] * cols, 4ull * cols,
                                hipMemcpyHostToDevice, m.stream));
        hipLaunchKernelGGL(k_cols_in, dim3(blocks_for(cols)), dim3(kBlock), 0, m.stream, b->G.p,
                           W, l, stage.p, T->colmap, cols);
      }
      dispatch(W, [&](auto w) { backward_block<decltype(w)::value>(T, *b); });
      for (int l = 0; l < W; ++l) {
        const int f = b->frames[l];
        if (f < 0) continue;
        if (grad_y) {
          hipLaunchKernelGGL(k_rows_out, dim3(blocks_for(rows)), dim3(kBlock), 0, m.stream,
                             b->GY.p, W, l, stage.p, m.rowold, rows);
          MLEM_HIP(hipMemcpyAsync(grad_y + (size_t)f * rows, stage.p, 4ull * rows,
                                  hipMemcpyDeviceToHost, m.stream));
        }
        if (grad_x0) {
          hipLaunchKernelGGL(k_cols_out, dim3(blocks_for(cols)), dim3(kBlock), 0, m.stream,
                             b->G.p, W, l, stage.p, T->colmap, cols);
          MLEM_HIP(hipMemcpyAsync(grad_x0 + (size_t)f * cols, stage.p, 4ull * cols,
                                  hipMemcpyDeviceToHost, m.stream));
        }
      }
    }
    check_launch("unrolled output launch");
    MLEM_HIP(hipStreamSynchronize(m.stream));
  });
}

int mlem_tape_info(const mlem_tape* T, int* block_width, int* checkpoint_interval,
                   int* checkpoints, int* recomputed_iterations, uint64_t* spmm_launches) {
  return guarded([&] {
    if (!T) raise(MLEM_EINVAL, "null tape");
    if (block_width) *block_width = T->W;
    if (checkpoint_interval) *checkpoint_interval = T->C;
    if (checkpoints) *checkpoints = T->nseg;
    if (recomputed_iterations) *recomputed_iterations = T->recomputed;
    if (spmm_launches) *spmm_launches = T->spmm_launches;
  });
}

void mlem_tape_destroy(mlem_tape* T) {
  if (!T) return;
  (void)hipStreamSynchronize(T->mat->m.stream);
  delete T;
}

}  // extern "C"
