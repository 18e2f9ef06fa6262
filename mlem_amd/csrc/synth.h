/*
 * synth.h — deterministic synthetic inputs for the MLEM engine
 * (libmlem_synth.so, host code, OpenMP).
 *
 * The reference's matrix and sinogram generators are external to it
 * (README.md:35-36), so this library has no reference counterpart. It writes
 * the same csr4 matrices and raw int32 sinograms the reference reads.
 *
 * Geometry: unit voxels, grid centred at the origin, voxel number
 * x + y*nx + z*nx*ny (helper_files_common/scannerconfig.cpp:30-34). Every
 * matrix row is a line of response between two points on the lateral surface
 * of a cylinder of radius 1.2*max(nx,ny)/2 and height nz (axis along z);
 * lines that miss the grid are redrawn. Values are Siddon ray/voxel
 * intersection lengths (R. L. Siddon, Med. Phys. 12(2), 1985) in float32,
 * columns ascending within each row.
 *
 * Randomness: SplitMix64 (Steele, Lea, Flood, OOPSLA 2014). Item k of a
 * stream with seed S starts from state mix(S ^ (k+1)*0xD1B54A32D192ED03), so
 * results do not depend on the thread count. Poisson variates: inversion for
 * mean < 10, PTRS transformed rejection (W. Hoermann, Insurance: Mathematics
 * and Economics 12, 1993) otherwise.
 *
 * All functions return 0 on success, -1 on bad arguments, -4 on allocation
 * failure. `threads` <= 0 means the OpenMP default.
 */
#ifndef MLEM_SYNTH_H
#define MLEM_SYNTH_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* System matrix with `nlors` rows and nx*ny*nz columns. *rowidx (nlors u64
 * END offsets, csr4 convention) and *elements (nnz {u32 col, f32 val}) are
 * malloc'd; release them with mlem_synth_free. */
int mlem_synth_matrix(uint32_t nx, uint32_t ny, uint32_t nz, uint32_t nlors,
                      uint64_t seed, int threads, uint64_t** rowidx,
                      void** elements, uint64_t* nnz);

void mlem_synth_free(void* p);

/* Sum of `nell` axis-aligned ellipsoids: centres uniform in the inner 70% of
 * the grid, semi-axes uniform in [semi_lo, semi_hi] voxels, intensities
 * uniform in [0.5, 4), each multiplied by scale[e] when scale != NULL.
 * x has nx*ny*nz entries. */
int mlem_synth_phantom(uint32_t nx, uint32_t ny, uint32_t nz, uint32_t nell,
                       uint64_t seed, float semi_lo, float semi_hi,
                       const float* scale, float* x);

/* ybar = A x in double (row-parallel). */
int mlem_synth_project(uint32_t rows, const uint64_t* rowidx,
                       const void* elements, const float* x, double* ybar,
                       int threads);

/* y_i ~ Poisson(ybar_i * total / sum(ybar)), int32. total <= 0 keeps ybar
 * unscaled. Returns the realised sum of y in *sum_y when non-NULL. */
int mlem_synth_poisson(uint32_t rows, const double* ybar, double total,
                       uint64_t seed, int threads, int32_t* y,
                       int64_t* sum_y);

/* Uniform doubles in [0,1) from the SplitMix64 stream `seed`, items
 * first..first+n-1 (one draw per item). */
int mlem_synth_uniform(uint64_t seed, uint64_t first, uint64_t n, double* out);

/* Trace one segment p0 -> p1 through the centred nx*ny*nz grid. Writes at
 * most `cap` (column, length) pairs, columns ascending; returns the number of
 * pairs, 0 when the segment misses the grid, -1 on bad arguments. */
int mlem_synth_siddon(const double p0[3], const double p1[3], uint32_t nx,
                      uint32_t ny, uint32_t nz, uint32_t cap, uint32_t* cols,
                      float* lens);

/* Write a csr4 v4 file with the minimal 60-byte unit-voxel scanner block. */
int mlem_synth_write_csr4(const char* path, uint32_t nx, uint32_t ny,
                          uint32_t nz, uint32_t rows, uint64_t nnz,
                          const uint64_t* rowidx, const void* elements);

#ifdef __cplusplus
}
#endif
#endif
