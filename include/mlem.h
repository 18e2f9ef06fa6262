/*
 * mlem.h — C ABI of the MI355X MLEM reconstruction engine (libmlem.so).
 *
 * The entry points below replace the plain-MLEM path of the reference's
 * sequential driver (mlem_seq/seqcsr4mlem.cpp) and the csr4 matrix class it
 * calls (helper_files_common/csr4matrix.{hpp,cpp}). Each declaration names the
 * reference function or method whose role it takes over.
 *
 * Conventions
 *   - C linkage, no exceptions cross the boundary.
 *   - Every function returns an int status: MLEM_OK (0) or a negative code;
 *     mlem_last_error() then returns a thread-local message.
 *   - Host pointers are caller-allocated; sizes are passed explicitly and are
 *     checked against the matrix dimensions.
 *   - A matrix handle owns one HIP stream and its device buffers. A handle is
 *     not thread-safe; use one handle per host thread.
 *   - Matrix storage on the device: CSR {u32 column, f32 value} (the csr4
 *     element layout, 8 B) plus a CSC transpose {u32 row, f32 value} built on
 *     the device with rows kept in ascending order inside every column, and
 *     per-tile dictionary formats of both built from them for the kernels.
 *
 * Compute modes (mlem_mode):
 *   MLEM_MODE_FAST   fp32 products summed in fixed trees per row (column);
 *                    deterministic run to run.
 *   MLEM_MODE_EXACT  each row's (column's) products summed one by one in
 *                    ascending element order, separate multiply and add,
 *                    flush-to-zero: reproduces the reference's fp32 rounding
 *                    sequence bit for bit.
 *   MLEM_MODE_ATOMIC forward as FAST; backprojection by fp32 atomic scatter
 *                    over the CSR (non-deterministic summation order).
 */
#ifndef MLEM_H
#define MLEM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Only the entry points below are exported from libmlem.so. */
#define MLEM_API __attribute__((visibility("default")))

/* ---- status codes ---------------------------------------------------- */
#define MLEM_OK 0
#define MLEM_EINVAL (-1)  /* bad argument or size mismatch */
#define MLEM_EIO (-2)     /* file could not be opened / read / written */
#define MLEM_EFORMAT (-3) /* not a csr4 v4 file, truncated, inconsistent */
#define MLEM_ENOMEM (-4)  /* host or device allocation failed */
#define MLEM_EDEVICE (-5) /* no usable GPU, or a HIP call failed */
#define MLEM_ESTATE (-6)  /* call not valid in the handle's current state */

/* ---- compute modes --------------------------------------------------- */
#define MLEM_MODE_FAST 0
#define MLEM_MODE_EXACT 1
#define MLEM_MODE_ATOMIC 2

/* ---- open flags ------------------------------------------------------ */
#define MLEM_OPEN_NO_CSC 0x1u /* skip the CSC build (ATOMIC mode only)   */
#define MLEM_OPEN_KEEP_ORDER 0x2u /* keep file row/voxel order on the device
                                     (no locality layout)                   */
/* Reduced-precision storage of the matrix values: every row is scaled by its
 * largest value (an fp32 per-row scale s_i) and the quotients q_ij are stored
 * in fp16 or bf16; column indices stay exact. Projections accumulate in fp32;
 * the column sums n_j = sum_i s_i q_ij and the log-likelihood run in FP64.
 * MLEM only sees a_ij = s_i q_ij. */
#define MLEM_OPEN_FP16 0x4u
#define MLEM_OPEN_BF16 0x8u
/* Files whose header sets symmetry flags (x 1, y 2, z 4; csr4matrix.hpp:
 * 147-149) store a matrix reduced by the scanner's mirror symmetries. With
 * this flag mlem_open regenerates the omitted elements on the host before the
 * upload (see mlem_expand_symmetry_file); without it the stored elements are
 * used as they are, as the reference's MLEM drivers do. */
#define MLEM_OPEN_EXPAND_SYMMETRY 0x10u

typedef struct mlem_matrix mlem_matrix; /* device-resident system matrix */
typedef struct mlem_recon mlem_recon;   /* one reconstruction's state    */

/* Thread-local message describing the last failing call ("" if none). */
MLEM_API const char* mlem_last_error(void);

/* Library version string, e.g. "mlem-amd 0.1 gfx950". */
MLEM_API const char* mlem_version(void);

/* ---- matrix ------------------------------------------------------------ */

/* Open a csr4 v4 file, validate it, stream the element array into HBM once
 * and build the CSC transpose on the device.
 * Replaces Csr4Matrix::Csr4Matrix(const std::string&)
 * (helper_files_common/csr4matrix.cpp:32-157) together with
 * Csr4Matrix::mapRows(0, rows()) (csr4matrix.cpp:166-224). Errors carry the
 * reference's messages ("Cannot open file ...", "... is not a CSR matrix",
 * "... has wrong CSR version (need v4)", "... is too short ...",
 * "Cannot find scanner geometry configuration",
 * "Error reading scanner geometry configuration"). Unlike the reference the
 * size check is 64-bit and row offsets and column indices are validated. */
MLEM_API int mlem_open(const char* path, unsigned flags, mlem_matrix** out);

/* Parse and validate a csr4 file's header, scanner block and row index
 * without touching a GPU: the header half of Csr4Matrix::Csr4Matrix
 * (csr4matrix.cpp:32-147), same error messages as mlem_open. grid and
 * symflags may be NULL. */
MLEM_API int mlem_probe(const char* path, uint32_t* rows, uint32_t* cols,
                        uint64_t* nnz, uint32_t grid[3], uint32_t* symflags);

/* Write the full matrix of a symmetry-reduced csr4 file (no GPU needed):
 * every stored element plus its images under the symmetries its flags
 * generate, as the reference's matrix viewer regenerates them
 * (visualizematrix/symhandler.hpp:27-101, symhandler.cpp:33-115,
 * dispcsr4spy.cpp:126-140): voxels mirrored in the grid, detectors mapped
 * through the scanner's mirror tables, rows (detector pairs) re-sorted. One
 * value per (row, column), the stored one first; columns ascending; the
 * output's symmetry flags are cleared. A file without flags is copied. */
MLEM_API int mlem_expand_symmetry_file(const char* in_path, const char* out_path);

/* Same as mlem_open for a matrix already in host memory: rowidx holds the
 * csr4 row index (nRows u64 exclusive END offsets, csr4matrix.cpp:141-147,
 * csr4matrix.hpp:138-149) and elements holds nnz 8-byte {u32 column,
 * f32 value} records (helper_files_common/matrixelement.hpp:37-46).
 * grid (may be NULL) gives the voxel grid nx, ny, nz with voxel number
 * x + y*nx + z*nx*ny (helper_files_common/scannerconfig.cpp:30-34); when
 * nx*ny*nz == cols the device copy uses the locality layout.
 * Counterpart of constructing a Csr4Matrix and installing the element array
 * with Csr4Matrix::setData (csr4matrix.hpp:46). */
MLEM_API int mlem_open_host(uint32_t rows, uint32_t cols, uint64_t nnz,
                   const uint64_t* rowidx, const void* elements,
                   const uint32_t grid[3], unsigned flags, mlem_matrix** out);

/* Out-of-core matrices (larger than HBM). The elements stay in pinned host
 * memory in blocks of whole rows; a block holds its rows (CSR) and their
 * transpose (CSC), each part at most block_bytes bytes (0: an eighth of the
 * device memory free at open). The first blocks whose parts fit in
 * resident_bytes are also kept in device memory (0: none). Every forward
 * projection streams the CSR parts of the other blocks and every
 * backprojection their CSC parts over PCIe through two device buffers, the
 * copy of one block overlapping the kernels of the block before. A
 * reconstruction iteration streams each CSR part once and builds the block's
 * CSC on the device right after its forward projection (mlem_stream_passes).
 * A column's backprojection continues its running sum from block to block in
 * ascending row order, so EXACT results equal the in-core ones bit for bit.
 * In a reconstruction in MLEM_MODE_ATOMIC each block's CSR part crosses PCIe
 * once per iteration: its forward projection and ratio, then the atomic
 * scatter of its rows, run while it is on the device (the reference's
 * out-of-core driver copies the parts for its forward loop and again for
 * its backprojection loop, mlem_cuda/versions/version_8.cu:214-252).
 * Every step function, mode and mlem_reconstruct / mlem_recon_* work on such a
 * handle (iterations run without hipGraph capture; statistics need at most 4096
 * blocks); the launch configuration is fixed (lane groups), and the batched
 * and unrolled paths, 16-bit values and the locality layout do not apply.
 * Flags: MLEM_OPEN_EXPAND_SYMMETRY (file only); the others are ignored or
 * rejected. Replaces the split-and-copy loop of
 * mlem_cuda/versions/version_7.cu:205-260 with the row splitting of
 * mlem_cuda/helper_files/helper_v_7_8.cu:4-72 (parts sized by free device
 * memory). */
MLEM_API int mlem_open_streamed(const char* path, unsigned flags, uint64_t block_bytes,
                                uint64_t resident_bytes, mlem_matrix** out);
MLEM_API int mlem_open_host_streamed(uint32_t rows, uint32_t cols, uint64_t nnz,
                                     const uint64_t* rowidx, const void* elements,
                                     const uint32_t grid[3], unsigned flags,
                                     uint64_t block_bytes, uint64_t resident_bytes,
                                     mlem_matrix** out);
/* Row blocks of a streamed matrix and how many of them are resident in
 * device memory (resident may be NULL); 0 for a matrix held in HBM. No
 * reference counterpart (version_7.cu prints its part count). */
MLEM_API int mlem_stream_blocks(const mlem_matrix* m, uint32_t* blocks, uint32_t* resident);
/* Passes over the streamed blocks per FAST / EXACT reconstruction iteration:
 * 1 when each block's CSC is built on the device from its CSR part (the
 * usual case), 2 when that did not fit in device memory at open and the
 * host-built CSC parts are streamed too; 0 for a matrix held in HBM. ATOMIC
 * iterations always make one pass. No reference counterpart. */
MLEM_API int mlem_stream_passes(const mlem_matrix* m, int* passes);

/* Release device memory, stream and handle. Replaces Csr4Matrix::~Csr4Matrix
 * (csr4matrix.cpp:159-163). Accepts NULL. */
MLEM_API void mlem_close(mlem_matrix* m);

/* rows(), columns(), elements() (csr4matrix.hpp:35-37). */
MLEM_API int mlem_dims(const mlem_matrix* m, uint32_t* rows, uint32_t* cols,
              uint64_t* nnz);

/* Grid dimensions and symmetry flags from the csr4 scanner block
 * (csr4matrix.cpp:62-65,126-129; Csr4Matrix::scannerConfig/symconfig,
 * csr4matrix.hpp:39-40). Zeros for matrices opened with mlem_open_host. */
MLEM_API int mlem_geometry(const mlem_matrix* m, uint32_t grid[3], uint32_t* symflags);

/* The handle's HIP stream (a hipStream_t), for callers that order their own
 * work against the library's. */
MLEM_API void* mlem_stream(mlem_matrix* m);

/* ---- step functions on host buffers (kernel-level parity) --------------- */

/* norm_j = sum_i a_ij in ascending row order, fp32.
 * Replaces calcColumnSums (mlem_seq/seqcsr4mlem.cpp:37-49). */
MLEM_API int mlem_column_sums(mlem_matrix* m, float* norm, uint32_t cols);

/* x_j = (float)(sum(y) / sum(norm)) with both sums in double, index order.
 * Replaces initImage (seqcsr4mlem.cpp:52-66). initial may be NULL. */
MLEM_API int mlem_init_image(mlem_matrix* m, const float* norm, uint32_t cols,
                    const int32_t* y, uint32_t rows, float* x, float* initial);

/* f = A x. Replaces calcFwProj (seqcsr4mlem.cpp:69-91). */
MLEM_API int mlem_fwproj(mlem_matrix* m, const float* x, uint32_t cols, float* f,
                uint32_t rows, int mode);

/* r_i = (f_i != 0) ? (float)y_i / f_i : 0.
 * Replaces calcCorrel (seqcsr4mlem.cpp:94-105). */
MLEM_API int mlem_correl(mlem_matrix* m, const float* f, const int32_t* y, float* r,
                uint32_t rows);

/* u = A^T r. Replaces calcBkProj (seqcsr4mlem.cpp:108-120). */
MLEM_API int mlem_bkproj(mlem_matrix* m, const float* r, uint32_t rows, float* u,
                uint32_t cols, int mode);

/* x_j *= (n_j != 0) ? u_j / n_j : u_j.
 * Replaces calcUpdate (seqcsr4mlem.cpp:123-134). */
MLEM_API int mlem_update(mlem_matrix* m, const float* u, const float* norm, float* x,
                uint32_t cols);

/* ---- whole reconstruction ------------------------------------------------ */

#define MLEM_INIT_REFERENCE 0 /* x0 = (float)(sum y / sum norm), initImage */
#define MLEM_INIT_GIVEN 1     /* x0 = caller's x_inout                     */
#define MLEM_CHECKPOINT_RESTORE_ONLY (-1) /* see mlem_options.checkpoint_interval */

typedef struct mlem_options {
  int mode;               /* MLEM_MODE_*                                  */
  int init;               /* MLEM_INIT_*                                  */
  int checkpoint_interval;/* 0: no checkpoint files (seqcsr4mlem);
                             >0: restart from img.chkpt/iter.chkpt when both
                             exist, write them every N iterations, remove
                             them at the end;
                             MLEM_CHECKPOINT_RESTORE_ONLY: restart from them
                             when both exist, write and remove none (the
                             reference's <checkpointing> = 0,
                             mpicsr4mlem.cpp:536-549,589,657)            */
  const char* checkpoint_dir; /* directory of the checkpoint files (NULL = CWD) */
  int verbose;            /* 1: print the reference driver's stdout lines */
  int use_graph;          /* 1: run iterations from a captured hipGraph   */
  double* loglik;         /* optional [iterations]: FP64 Poisson log-likelihood
                             of each iterate's forward projection          */
  double* fwd_total;      /* optional [iterations]: FP64 sum_i f_i         */
} mlem_options;

/* Fill *o with the defaults (FAST, reference init, no checkpoints, quiet,
 * graph capture on). */
MLEM_API void mlem_options_default(mlem_options* o);

/* Run `iterations` MLEM iterations for one int32 sinogram. x_inout receives
 * the image (and supplies x0 under MLEM_INIT_GIVEN). With verbose=1 the
 * reference's stdout lines are printed (INIT line, per-iteration image sum,
 * zero and subnormal counts). Replaces mlem() (seqcsr4mlem.cpp:137-189); the
 * checkpoint files and restart follow mpicsr4mlem.cpp:95-154,536-553 with the
 * interval meaning of hybCalcFunc.cu:852-853. */
MLEM_API int mlem_reconstruct(mlem_matrix* m, const int32_t* y, uint32_t rows,
                     float* x_inout, uint32_t cols, int iterations,
                     const mlem_options* opts);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the iterations of mlem_reconstruct on device-resident
 * vectors, allocating nothing. y (rows int32 counts) and x (cols fp32 image:
 * x0 in, x_K out) are DEVICE pointers in file row / voxel order; mode is
 * MLEM_MODE_FAST or MLEM_MODE_EXACT; loglik and fwd_total (host,
 * [iterations], may be NULL) receive the FP64 statistics of every iteration.
 * The work runs on the matrix's stream (mlem_stream): make y and x ready
 * before the call. With both statistics arrays NULL the call only enqueues
 * (no synchronization; it can be captured into a HIP graph), otherwise it
 * returns when done. Its scratch (device-order copies of y and x, the ratio
 * vector, FP64 statistics partials) is one static 64 MiB device array per
 * GPU, shared by every call on that GPU: calls that overlap on one GPU race.
 * Matrices that need more scratch (4 * (2 rows + cols) bytes plus the
 * statistics) get MLEM_EINVAL; streamed matrices are not supported. */
MLEM_API int mlem_reconstruct_noalloc(mlem_matrix* m, const int32_t* y, uint32_t rows,
                                      float* x, uint32_t cols, int iterations, int mode,
                                      double* loglik, double* fwd_total);

/* ---- device-resident reconstruction state --------------------------------
 * The same iteration as mlem_reconstruct, split so that the sinogram and the
 * image stay in HBM between calls (used by the benchmark and the batched
 * driver). No reference counterpart beyond the loop body of mlem()
 * (seqcsr4mlem.cpp:163-188). */

/* Upload y, compute x0 (opts->init), allocate r. */
MLEM_API int mlem_recon_create(mlem_matrix* m, const int32_t* y, uint32_t rows,
                      const float* x0, uint32_t cols, const mlem_options* opts,
                      mlem_recon** out);
MLEM_API void mlem_recon_destroy(mlem_recon* r);
/* Enqueue `iterations` iterations on the matrix stream and wait for them. */
MLEM_API int mlem_recon_run(mlem_recon* r, int iterations);
/* As mlem_recon_run, with HIP events around every kernel launch; returns the
 * per-kernel durations in milliseconds (arrays of length `iterations`, any
 * may be NULL) and the whole span. kernel_fwd: forward+ratio launches;
 * kernel_bwd: backprojection+update launches (ATOMIC mode: scatter and
 * update summed; a streamed matrix in ATOMIC mode scatters each row block
 * right after its forward projection, so kernel_fwd holds both and
 * kernel_bwd the update; in FAST / EXACT mode a streamed matrix is timed
 * with its two passes, the forward over the CSR parts and the
 * backprojection over the CSC parts, same results as the one-pass
 * mlem_recon_run). */
MLEM_API int mlem_recon_run_timed(mlem_recon* r, int iterations, float* kernel_fwd_ms,
                         float* kernel_bwd_ms, float* span_ms);
/* Copy the current image to the host. */
MLEM_API int mlem_recon_image(mlem_recon* r, float* x, uint32_t cols);
/* Iterations completed so far. */
MLEM_API int mlem_recon_iterations(const mlem_recon* r);

/* ---- batched reconstruction: many sinograms, one matrix ------------------
 * Frames (e.g. dynamic PET frames) are reconstructed together in blocks of up
 * to block_width right-hand sides; each matrix record streamed from HBM
 * serves a whole block. Per frame the iteration is that of mlem()
 * (seqcsr4mlem.cpp:137-189); in MLEM_MODE_EXACT the image after k iterations
 * is bit-identical to k iterations of the sequential driver. A frame stops after budgets[f]
 * iterations or, with rel_tol > 0, after iteration k >= 2 once
 * |L(x_{k-1}) - L(x_{k-2})| <= rel_tol * |L(x_{k-1})|, L being the FP64
 * Poisson log-likelihood sum_{f_i != 0} (y_i log f_i - f_i) that iteration k
 * computes for the image entering it. Finished frames are frozen and the rest
 * repacked into fewer / narrower blocks. No reference counterpart (the
 * reference reconstructs one sinogram per run). */
typedef struct mlem_batch mlem_batch;
typedef struct mlem_batch_options {
  double rel_tol;   /* early-stop tolerance on the log-likelihood (0 = off) */
  int block_width;  /* widest RHS block: 4, 8, 16 or 32 frames;
                       1 = every frame alone on the single-image
                       kernels of the batch's mode;
                       0 = autotune: one iteration is timed at every width
                       (and alone) at creation and live frames are packed
                       into the blocks that minimise the summed measured
                       time, single frames included                       */
  int use_graph;    /* 1: replay each iteration from a captured hipGraph  */
  int mode;         /* MLEM_MODE_FAST (default): blocks run through the
                       matrix's box tiling (16x16x4-voxel boxes whose slice
                       of a block, or whose voxel sums, stay in one
                       workgroup's LDS or registers), fp32 fused
                       multiply-adds summed in a fixed order: deterministic,
                       at the fp32 floor of FP64 MLEM; lone frames on the
                       single-image FAST kernels.
                       MLEM_MODE_EXACT: every frame's image is bit-identical
                       to the sequential driver's (per-frame sequential sums
                       over the CSR / CSC, lone frames on the EXACT kernels) */
} mlem_batch_options;
MLEM_API void mlem_batch_options_default(mlem_batch_options* o);
/* y: frames*rows int32 counts, frame-major; budgets: frames iteration limits
 * (NULL: unlimited). x0 of each frame follows initImage. */
MLEM_API int mlem_batch_create(mlem_matrix* m, uint32_t frames, const int32_t* y,
                               uint32_t rows, const int32_t* budgets,
                               const mlem_batch_options* o, mlem_batch** out);
/* Run up to max_iterations batched iterations (fewer if every frame stops). */
MLEM_API int mlem_batch_run(mlem_batch* b, int max_iterations, int* iterations_run);
MLEM_API int mlem_batch_image(mlem_batch* b, uint32_t frame, float* x, uint32_t cols);
/* Iterations done by a frame and its log-likelihood history (capacity entries). */
MLEM_API int mlem_batch_frame_info(mlem_batch* b, uint32_t frame, int* iterations,
                                   double* loglik, int capacity);
MLEM_API int mlem_batch_progress(const mlem_batch* b, uint32_t* active_frames,
                                 uint64_t* frame_iterations, int* block_width,
                                 uint32_t* blocks);
/* Measured milliseconds of one block-iteration at widths 4, 8, 16, 32
 * (zeros unless the batch was created with block_width 0). */
MLEM_API int mlem_batch_tuning(const mlem_batch* b, double block_ms[4]);
/* Measured milliseconds of one iteration of a lone frame on the single-image
 * EXACT kernels, the cost of a width-1 block (0 unless block_width was 0).
 * No reference counterpart. */
MLEM_API int mlem_batch_single_ms(const mlem_batch* b, double* ms);
MLEM_API void mlem_batch_destroy(mlem_batch* b);

/* ---- differentiable unrolled MLEM ----------------------------------------
 * x_K = K iterations of mlem() (seqcsr4mlem.cpp:137-189) from given initial
 * images, with real-valued counts y, as a function of (y, x0); reverse mode
 * through the multiplicative update gives dL/dy and dL/dx0 from dL/dx_K.
 * The frames of a batch are the right-hand sides of one SpMM block. The tape
 * keeps the iterate every checkpoint_interval iterations (0: round(sqrt K));
 * mlem_backward_k recomputes each segment once from its checkpoint. Where
 * f_i = 0 the ratio is 0 and its derivative is taken as 0. No reference
 * counterpart (the reference has no gradients). */
typedef struct mlem_tape mlem_tape;
/* y: frames*rows float counts; x0: frames*cols initial images (NULL: per
 * frame (float)(sum y / sum n), then treated as a constant); xK: frames*cols
 * output (may be NULL); tape: receives the tape for mlem_backward_k (NULL:
 * forward only). All arrays frame-major, file row / voxel order. */
MLEM_API int mlem_forward_k(mlem_matrix* m, uint32_t frames, const float* y, uint32_t rows,
                            const float* x0, uint32_t cols, int K, int checkpoint_interval,
                            float* xK, mlem_tape** tape);
/* mlem_forward_k with the products' arithmetic chosen (mlem_forward_k is
 * MLEM_MODE_EXACT): MLEM_MODE_EXACT sums every frame's products one by one in
 * file element order (for integer counts x_K is bit-identical to mlem());
 * MLEM_MODE_FAST runs the products through the matrix's box tiling (the
 * batched FAST kernels), fixed-order sums, deterministic for a given block
 * width. mlem_backward_k runs in the tape's mode. No reference counterpart. */
MLEM_API int mlem_forward_k_mode(mlem_matrix* m, uint32_t frames, const float* y,
                                 uint32_t rows, const float* x0, uint32_t cols, int K,
                                 int checkpoint_interval, int mode, float* xK, mlem_tape** tape);
/* grad_xK: frames*cols dL/dx_K; grad_y (frames*rows) and grad_x0
 * (frames*cols) may be NULL. May be called more than once per tape. */
MLEM_API int mlem_backward_k(mlem_tape* t, const float* grad_xK, float* grad_y,
                             float* grad_x0);
MLEM_API int mlem_tape_info(const mlem_tape* t, int* block_width, int* checkpoint_interval,
                            int* checkpoints, int* recomputed_iterations,
                            uint64_t* spmm_launches);
/* Destroy a tape (before closing its matrix). */
MLEM_API void mlem_tape_destroy(mlem_tape* t);

/* ---- launch configuration (autotuner) --------------------------------------
 * FAST-mode kernel family and its parameter, per projection:
 *   MLEM_KERNEL_TILES  param = elements per lane (4, 8 or 16): 256-lane
 *                      blocks stream tiles of 256*param elements made of
 *                      whole rows (columns) and reduce them in LDS;
 *   MLEM_KERNEL_GROUPS param = lanes per row (column): 4, 8, 16, 32 or 64.
 *   MLEM_KERNEL_DICT   param = elements per lane (8, 16, 24, 32, 48 or 64):
 *                      256-lane blocks over tiles of at most 256*param
 *                      elements and 256 whole rows (columns) whose distinct
 *                      gather indices are stored as a sorted per-tile
 *                      dictionary (built on the device on first use); each
 *                      distinct x (r) value is gathered once into LDS.
 *   MLEM_KERNEL_DICT_WIDE param = elements per lane (16, 24 or 32): the same
 *                      with 512-lane blocks, tiles of at most 512*param
 *                      elements and 512 rows (columns) (default, param 32).
 * No reference counterpart (the reference hard-codes 1024-thread blocks,
 * mlem_cuda/versions/version_5.cu:121-126). */
#define MLEM_KERNEL_GROUPS 0
#define MLEM_KERNEL_TILES 1
#define MLEM_KERNEL_DICT 2
#define MLEM_KERNEL_DICT_WIDE 3
MLEM_API int mlem_get_launch(const mlem_matrix* m, int* fwd_kernel,
                             int* fwd_param, int* bwd_kernel, int* bwd_param);
MLEM_API int mlem_set_launch(mlem_matrix* m, int fwd_kernel, int fwd_param,
                             int bwd_kernel, int bwd_param);
/* Time every forward and backward configuration with HIP events on this
 * matrix and keep the fastest of each. */
MLEM_API int mlem_autotune(mlem_matrix* m);

/* ---- measurement helper ---------------------------------------------------
 * Copy `bytes` (multiple of 16) from one device buffer to another with a plain
 * streaming kernel, `repeats` times; average kernel time in *ms. Used to
 * calibrate hardware byte counters. No reference counterpart. */
MLEM_API int mlem_calibration_copy(size_t bytes, int repeats, float* ms);

/* Synchronize the device (hipDeviceSynchronize). */
MLEM_API int mlem_device_sync(void);

#ifdef __cplusplus
}
#endif
#endif /* MLEM_H */
