// internal.hpp — device-side matrix/state structures and kernel launchers
// shared by the C ABI (api.cpp) and the kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "csr4.hpp"

namespace mlem {

// Error raised inside the library; the C ABI converts it to a status code.
struct ApiError {
  int code;
  std::string msg;
};

// Every kernel runs in the fp32 mode of the reference driver, which sets
// flush-to-zero but not denormals-are-zero (mlem_seq/seqcsr4mlem.cpp:194):
// subnormal inputs are used as they are, subnormal results become zero. The
// code is built with -fdenormal-fp-math-f32=preserve-sign,ieee (MODE register
// FP32 denormal field 1, "flush outputs").
//
// fp32 division goes through div_rn: the compiler's own fp32 division would
// switch to full denormals for its fused steps and then write field 0
// (inputs flushed as well) for the rest of the kernel. div_rn divides in
// fp64 instead (fp64 keeps denormals, no mode switch): the quotient of two
// fp32 values rounded to fp64 and then to fp32 is the correctly rounded fp32
// quotient, since 53 >= 2*24 + 2. As x86 under FTZ, a quotient is flushed to
// zero when it is tiny after rounding, i.e. when its rounding to 24 bits with
// unbounded exponent is below 2^-126, which is |a/b| < 2^-126 - 2^-151.
__device__ __forceinline__ float div_rn(float a, float b) {
  const double q = (double)a / (double)b;
  constexpr double kTiny = 0x1.ffffffp-127;  // 2^-126 - 2^-151
  return (float)(__builtin_fabs(q) < kTiny ? __builtin_copysign(0.0, q) : q);
}

[[noreturn]] void raise(int code, const std::string& msg);
void hip_check(hipError_t e, const char* what);
#define MLEM_HIP(call) ::mlem::hip_check((call), #call)

// Thread-local message of the last failing C-ABI call (api.cpp).
void set_error(const std::string& msg);
void clear_error();

// Run f and translate exceptions into a C-ABI status code.
template <class F>
int guarded(F&& f) {
  try {
    f();
    clear_error();
    return 0;
  } catch (const ApiError& e) {
    set_error(e.msg);
    return e.code;
  } catch (const Csr4Error& e) {
    set_error(e.msg);
    return e.code;
  } catch (const std::bad_alloc&) {
    set_error("host allocation failed");
    return -4;
  } catch (const std::exception& e) {
    set_error(e.what());
    return -1;
  } catch (...) {
    set_error("unknown error");
    return -1;
  }
}

// FAST-mode kernel family of a projection. kDict and kDictWide are the same
// dictionary tiles run by 256-lane and 512-lane blocks.
enum KernelKind { kGroups = 0, kTiles = 1, kDict = 2, kDictWide = 3 };
inline bool is_dict(int kind) { return kind == kDict || kind == kDictWide; }
inline int dict_lanes(int kind) { return kind == kDictWide ? 512 : 256; }

// Dictionary-tile storage of one projection (built on the device at load
// time for `nb` lanes per block and `ept` elements per lane: tiles of at most
// nb*ept elements and nb segments; 0 = not built).
struct DictFormat {
  int nb = 0;
  int ept = 0;
  int vk = 0;                 // value storage (ValueKind): fp32, fp16, bf16
  uint32_t ntiles = 0;
  uint4* tiles = nullptr;     // [ntiles+1] {first segment, first element,
                              //  aligned element start, first dict word}
  uint32_t* dict = nullptr;   // distinct gather indices, sorted per tile, in
                              //  64-word slots of 64 entries, nb/64 per round
  uint32_t* dmeta = nullptr;  // per slot: first index | 0x80000000 if 32-bit
  void* dval = nullptr;       // element values, tile-aligned (4 or 2 bytes)
  uint16_t* dlix = nullptr;   // element positions in the tile dictionary
  uint64_t ndict = 0, nval = 0;  // dict words, element slots
};

// Storage precision of the matrix values.
enum ValueKind { kF32 = 0, kF16 = 1, kBF16 = 2 };

// Launch configuration of one projection (forward or backward).
struct GroupLaunch {
  int kind = kTiles;       // kGroups: lane group per row/column; kTiles: tiles
  int lanes = 16;          // kGroups: lanes per row (or column)
  int ept = 8;             // kTiles: elements per lane, tile = 256*ept elements
  int blocks_per_cu = 0;   // kGroups: 0 = one block per 256/lanes rows
};

// A matrix kept in pinned host memory and streamed to the device in row
// blocks (out-of-core). Each block holds its rows as CSR (block-local row
// pointers, global columns) and their transpose as CSC (global rows); a
// projection copies its part of every block in turn through two device
// buffers, the copy of the next block overlapping the kernels of this one.
struct StreamBlock {
  uint32_t row0 = 0, rows = 0;  // first row, rows
  uint64_t nnz = 0;
  uint64_t csr_off = 0, csr_bytes = 0;  // parts within the host buffer
  uint64_t csc_off = 0, csc_bytes = 0;
  char* resident = nullptr;  // device copy of both parts (CSR, then CSC), or none
};
struct StreamedMatrix {
  char* host = nullptr;  // pinned: every block's CSR part, then its CSC part
  uint64_t host_bytes = 0;
  std::vector<StreamBlock> blocks;
  uint64_t part_bytes = 0;  // largest part: the size of each device buffer
  char* dev[2] = {nullptr, nullptr};
  hipStream_t copy = nullptr;
  hipEvent_t copied[2] = {nullptr, nullptr};  // part in dev[i]
  hipEvent_t freed[2] = {nullptr, nullptr};   // kernels reading dev[i] done
  float* u = nullptr;                          // [cols] backprojection sums
  int lanes_fwd = 16, lanes_bwd = 4;           // FAST lane groups per row / column
  // A streamed block's CSC built on the device from its CSR part (colptr,
  // elements, sort scratch for the largest streamed block), so that a
  // reconstruction iteration streams each block once; null when it did not
  // fit in device memory next to the two part buffers (two passes then).
  char* csc_dev = nullptr;
  uint64_t csc_scratch_bytes = 0;
  uint64_t max_stream_nnz = 0;
  StreamedMatrix() = default;
  StreamedMatrix(const StreamedMatrix&) = delete;
  StreamedMatrix& operator=(const StreamedMatrix&) = delete;
  ~StreamedMatrix();
};

// Two-dimensional tiling of the matrix for products with W interleaved
// right-hand sides (panels.hpp; built on first use by ensure_panels). Vectors
// of those products index columns in panel order.
struct PanelFormat {
  uint32_t rows = 0, cols = 0;
  uint32_t npanels = 0, nchunks = 0;   // boxes, backward chunks
  uint32_t nfchunks = 0;               // forward chunks
  uint64_t nseg = 0, ngroups = 0;      // forward slots (pieces), backward slots (voxels)
  uint32_t* pc_of_file = nullptr;      // [cols] panel-order column of each file column
  uint32_t* pc_of_dev = nullptr;       // [cols] ... of each device column
  uint32_t* dev_of_pc = nullptr;       // [cols] device column of each panel-order column
  float* norm = nullptr;               // [cols] column sums in panel order
  uint32_t* pstart = nullptr;          // [npanels+1] first panel-order column of each box
  uint32_t* border = nullptr;          // [npanels] launch order (largest box first)
  uint32_t* pseg = nullptr;            // [npanels+1] first piece of each box
  uint32_t* fseg = nullptr;            // [nseg+1] first element of each piece (forward order)
  uint32_t* fslot = nullptr;           // [nseg] row-major slot of each piece
  float* fval = nullptr;               // [nnz]
  uint16_t* fcol = nullptr;            // [nnz] column inside the box
  uint32_t* rowslot = nullptr;         // [rows+1] first slot of each device row
  uint32_t* rowpiece = nullptr;        // [nseg] forward-order piece of each row-major slot
  uint32_t* pfchunk = nullptr;         // [npanels+1] first forward chunk of each box
  uint2* fchunk = nullptr;             // [nfchunks+1] {first piece, first element}
  uint32_t* brow = nullptr;            // [nseg] device row of each piece (backward order)
  uint32_t* pchunk = nullptr;          // [npanels+1] first chunk of each box
  uint2* chunk = nullptr;              // [nchunks+1] {first piece, first element}
  uint16_t* boff = nullptr;            // [nchunks*(kBoxVox+1)] voxel element offsets
  uint16_t* bpc = nullptr;             // [nnz] piece inside the chunk
  float* bval = nullptr;               // [nnz]
  uint32_t* colslot = nullptr;         // [cols+1] slot of each panel-order column (identity)
  double build_seconds = 0.0;
  PanelFormat() = default;
  PanelFormat(const PanelFormat&) = delete;
  PanelFormat& operator=(const PanelFormat&) = delete;
  ~PanelFormat();
};

struct DevMatrix {
  uint32_t rows = 0, cols = 0;
  uint64_t nnz = 0;
  uint32_t grid[3] = {0, 0, 0};
  uint32_t symflags = 0;
  int device = 0;
  int cus = 256;
  hipStream_t stream = nullptr;

  uint32_t* rowptr = nullptr;  // [rows+1]
  uint2* csr = nullptr;        // [nnz] {column, value bits}
  uint32_t* colptr = nullptr;  // [cols+1]
  uint2* csc = nullptr;        // [nnz] {row, value bits}, rows ascending
  float* norm = nullptr;       // [cols] fp32 column sums, ascending row order

  // Locality layout (nullptr = identity): device column of each file column,
  // file row of each device row, device row of each file row.
  uint32_t* colmap = nullptr;  // [cols]
  uint32_t* rowold = nullptr;  // [rows]
  uint32_t* rowpos = nullptr;  // [rows]
  double sumnorm = 0.0;        // sequential double sum of norm (initImage)

  // Reduced-precision values: every record holds q_ij (exactly representable
  // in fp16 / bf16) and a_ij = rscale[i] * q_ij. The projections run on q;
  // the scales enter the FP64 column sums and the FP64 statistics only.
  int vk = kF32;
  float* rscale = nullptr;     // [rows] device row order, nullptr for fp32

  GroupLaunch fwd, bwd, scatter;
  // Bumped whenever the launch configuration (and with it the tile and
  // dictionary buffers) may change; captured iterations are re-captured.
  uint64_t launch_gen = 0;

  // Tiled kernels: device-order pointers kept on the host, and the tile
  // partition (first segment of every tile, plus the end) for the current
  // elements-per-lane setting.
  std::vector<uint32_t> h_rowptr, h_colptr;
  uint2* tiles_fwd = nullptr;
  uint32_t ntiles_fwd = 0;
  int tiles_fwd_ept = 0;
  uint2* tiles_bwd = nullptr;
  uint32_t ntiles_bwd = 0;
  int tiles_bwd_ept = 0;
  DictFormat dfwd, dbwd;

  // per-handle scratch for the host-buffer step functions
  float* sx = nullptr;   // [cols]
  float* su = nullptr;   // [cols]
  float* sf = nullptr;   // [rows]
  float* sr = nullptr;   // [rows]
  int32_t* sy = nullptr; // [rows]
  float* tx = nullptr;   // [cols] file-order staging
  float* tf = nullptr;   // [rows] file-order staging

  StreamedMatrix* streamed = nullptr;  // out-of-core matrix (no device CSR/CSC)
  PanelFormat* panels = nullptr;       // tiling for many right-hand sides (or none)
};

// Deterministic FP64 reduction state for the forward pass: per-block partial
// (sum of y log f - f, sum of f) pairs, a completion counter, and a ring of
// per-iteration results indexed by a device-side iteration counter.
struct StatsState {
  double2* partials = nullptr;  // [max blocks]
  unsigned* counter = nullptr;  // [2]: [1] = next result slot ([0] unused)
  double2* results = nullptr;   // [ring]
  unsigned ring = 0;
  unsigned max_blocks = 0;
};

// ---- launchers (kernels.hip) -------------------------------------------
// Grid size of a statistics-carrying forward launch (partials needed).
unsigned stats_max_blocks(const DevMatrix& m);
// Copy the device-order row/column pointers to the host (once, after the
// layout is built) and (re)build the tile partitions and dictionary formats
// the launch configuration needs.
void prepare_tiles(DevMatrix& m, hipStream_t s);
// Release a dictionary format's device memory.
void free_dict(DictFormat& d);
// Forward projection. ratio: out = (f != 0) ? y/f : 0, else out = f.
void launch_fwd(const DevMatrix& m, int mode, const float* x, const int32_t* y,
                float* out, bool ratio, const StatsState* stats, hipStream_t s);
// Backprojection through the CSC. update: x *= (n != 0) ? u/n : u written to
// `out` (which holds x), else out = u.
void launch_bwd(const DevMatrix& m, int mode, const float* r, const float* norm,
                float* out, bool update, hipStream_t s);
// Backprojection by atomic scatter over the CSR: u += A^T r (u pre-zeroed).
void launch_scatter(const DevMatrix& m, const float* r, float* u, hipStream_t s);
// One row block of a streamed matrix (kernels.hip).
void launch_fwd_block(const uint32_t* rowptr, const uint2* csr, uint32_t rows, int lanes,
                      bool exact, const float* x, const int32_t* y, float* out, bool ratio,
                      const StatsState* st, int cus, hipStream_t s);
void launch_bwd_block_add(const uint32_t* colptr, const uint2* csc, uint32_t cols, int lanes,
                          bool exact, const float* r, float* u, int cus, hipStream_t s);
// u += (block rows)^T r by atomic scatter over one row block's CSR.
void launch_scatter_block(const uint32_t* rowptr, const uint2* csr, uint32_t rows, int lanes,
                          const float* r, float* u, int cus, hipStream_t s);
// Streamed matrices (stream.cpp): build the blocks from a row index (END
// offsets) and an element source; the projections over every block.
using ElementFetch = std::function<void(void* dst, uint64_t first, uint64_t count)>;
StreamedMatrix* build_streamed(uint32_t rows, uint32_t cols, uint64_t nnz,
                               const uint64_t* rowidx, const ElementFetch& fetch,
                               uint64_t block_bytes, uint64_t resident_bytes, hipStream_t s);
void stream_fwd(const DevMatrix& m, int mode, const float* x, const int32_t* y, float* out,
                bool ratio, const StatsState* st, hipStream_t s);
void stream_bwd(const DevMatrix& m, int mode, const float* r, const float* norm, float* out,
                bool update, hipStream_t s);
void stream_bwd_add(const DevMatrix& m, bool exact, const float* r, float* u, hipStream_t s);
// One pass over the CSR blocks: per block the forward projection with the
// fused ratio into r, then u += (block rows)^T r by atomic scatter while the
// block is on the device (ATOMIC mode: each block crosses PCIe once per
// iteration instead of twice).
void stream_fwd_scatter(const DevMatrix& m, const float* x, const int32_t* y, float* r, float* u,
                        const StatsState* st, hipStream_t s);
// One FAST / EXACT reconstruction iteration over a streamed matrix: per block
// the forward projection with the fused ratio into r, then (with csc_dev) the
// block's CSC built on the device and its backprojection added to u, so each
// CSR part crosses PCIe once; without csc_dev the two passes of stream_fwd and
// stream_bwd. Then x *= (n != 0) ? u/n : u.
void stream_iteration(const DevMatrix& m, int mode, float* x, const int32_t* y, float* r,
                      const StatsState* st, hipStream_t s);
// CSC of one row block (block-local row pointers, global columns) built on the
// device: a stable sort by column keeps the block-local rows ascending inside
// every column. block_csc_scratch gives the scratch bytes it needs.
size_t block_csc_scratch(uint64_t nnz, uint32_t cols);
void build_block_csc(const uint32_t* rowptr, const uint2* csr, uint32_t rows, uint32_t cols,
                     uint64_t nnz, void* scratch, size_t scratch_bytes, uint32_t* colptr,
                     uint2* csc, hipStream_t s);
// x *= (n != 0) ? u/n : u; optionally u = 0 afterwards.
void launch_update(uint32_t cols, const float* u, const float* norm, float* x,
                   bool zero_u, hipStream_t s);
// r = (f != 0) ? y/f : 0.
void launch_correl(uint32_t rows, const float* f, const int32_t* y, float* r,
                   hipStream_t s);
// norm from the CSC, sequential per column in ascending row order.
void launch_colsums_csc(const DevMatrix& m, float* norm, hipStream_t s);
// norm by atomic scatter (matrices opened without a CSC).
void launch_colsums_atomic(const DevMatrix& m, float* norm, hipStream_t s);
// Reduced precision: row scales (file row order) and in-place quantization
// of the file-order CSR values to fp16 / bf16 (kept as exact fp32 words).
void launch_quantize(DevMatrix& m, int vk, float* rscale, hipStream_t s);
// norm_j = (float) sum_i (double) s_i q_ij from the CSC, ascending row order.
void launch_colsums_scaled(const DevMatrix& m, float* norm, hipStream_t s);
// r_i *= s_i in device row order (no-op for fp32 values).
void launch_scale_rows(const DevMatrix& m, float* r, hipStream_t s);
// ---- products with W right-hand sides through the tiling (panel.hip) ------
// Build (once per matrix, on the host from the device CSR) and return the
// tiling; free_panels releases it.
const PanelFormat& ensure_panels(DevMatrix& m);
void free_panels(DevMatrix& m);
// part[fslot * W + f] = piece dot products with X (panel order, [cols][W]);
// part[c * W + f] = (A^T R)_c for every panel-order column c, R [rows][W] in
// device row order (one slot per column: colslot is the identity).
// One workgroup per box; grid is not used. For W <= 8 (panel_slots_by_box)
// the forward writes piece s (forward order) to part[s * W + f] instead, so
// that a box's sums are written contiguously; the slot sums of row r then
// read part[rowpiece[k] * W + f] for k in [rowslot[r], rowslot[r+1]).
constexpr bool panel_slots_by_box(int W) { return W <= 8; }
template <int W>
void launch_panel_fwd(const PanelFormat& P, const float* X, float* part, uint32_t grid,
                      hipStream_t s);
template <int W>
void launch_panel_bwd(const PanelFormat& P, const float* R, float* part, uint32_t grid,
                      hipStream_t s);
// ---- load-time layout (layout.hip) ----------------------------------------
// Renumber rows (sinogram order) and voxels (4x4x2 bricks) when the grid is
// known, build the CSC (file row order inside every column) and the
// device-order CSR. Requires validated column indices.
void build_layout(DevMatrix& m, bool reorder_rows, bool reorder_cols, bool want_csc,
                  hipStream_t s);
// File-order <-> device-order vectors (device buffers, distinct pointers).
void to_device_cols(const DevMatrix& m, const float* file_order, float* dev, hipStream_t s);
void to_file_cols(const DevMatrix& m, const float* dev, float* file_order, hipStream_t s);
template <class T>
void rows_to_device(const DevMatrix& m, const T* file_order, T* dev, hipStream_t s);
template <class T>
void rows_to_file(const DevMatrix& m, const T* dev, T* file_order, hipStream_t s);
// Count column indices >= cols in the CSR (validation).
void launch_check_columns(const DevMatrix& m, unsigned* bad, hipStream_t s);
// Build rowptr (u32) from the csr4 END-offset row index already on device.
void launch_rowptr_from_rowidx(const uint64_t* rowidx, uint32_t rows,
                               uint32_t* rowptr, hipStream_t s);
// Fill a float buffer.
void launch_fill(float* p, float v, size_t n, hipStream_t s);
// Streaming copy kernel (byte-counter calibration).
void launch_copy(const uint4* src, uint4* dst, size_t n16, hipStream_t s);

}  // namespace mlem

// The C-ABI matrix handle (include/mlem.h): the device matrix plus load
// timings. Defined here so every source of the library can reach it.
struct mlem_matrix {
  mlem::DevMatrix m;
  double norm_seconds = 0.0;  // layout, CSC build and column sums at open
  double load_seconds = 0.0;  // file/host -> HBM
  ~mlem_matrix();
};
