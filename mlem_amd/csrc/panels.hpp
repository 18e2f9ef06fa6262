// panels.hpp — the system matrix as a two-dimensional tiling for products
// with many right-hand sides (host-side construction, no GPU needed).
//
// A product with W interleaved right-hand sides gathers one W*4-byte run per
// matrix element (128 bytes at W = 32). Over the whole matrix those runs are
// tens of GB, far more than the element stream, and they only stay cheap if
// the vector runs that the workgroups of one XCD gather at a time fit in its
// L2. Two tilings keep them there:
//
//   forward (f = A X): the voxels are grouped into panels, boxes of the grid
//   (32x32x16 voxels: 2 MB of X at W = 32). Every row is cut into segments,
//   one per panel it crosses (about 4.4 per row on the 128x128x48 grid); the
//   segments of one panel are stored together, rows ascending, with their
//   columns as 16-bit offsets inside the panel. A panel's segments gather only
//   that panel's slice of X. Each segment's dot product is written to its own
//   slot; the slots of one row are contiguous (row-major), so a second pass
//   adds them in panel order.
//
//   backward (u = A^T R): inside each panel the segments (rows ascending) are
//   split into chunks of `chunk_rows` segments; the chunk's rows' runs of R
//   are the gathered working set (2 MB at W = 32, 16384 rows). For every
//   (panel, chunk, voxel) group the elements are stored as {row, value},
//   rows ascending; each group's dot product goes to a slot, the slots of one
//   voxel contiguous, added in chunk order by a second pass.
//
// Sums are deterministic (fixed order), not the sequential order of the
// reference driver (mlem_seq/seqcsr4mlem.cpp:69-120).
#pragma once

#include <cstdint>
#include <vector>

namespace mlem {

// Longest segment (forward) or group (backward): longer pieces are split into
// several, each with its own slot, so that every segment fits a staged tile.
constexpr uint32_t kMaxSegment = 512;

struct HostPanels {
  uint32_t rows = 0, cols = 0;
  uint64_t nnz = 0;
  uint32_t npanels = 0;
  uint32_t chunk_rows = 0;
  // column order of the tiling ("panel order"): panels one after the other,
  // the voxels of a panel in 4x4x2 bricks
  std::vector<uint32_t> pc_of_dev;    // [cols] panel-order column of each device column
  std::vector<uint32_t> pstart;       // [npanels+1] first panel-order column of each panel
  // forward: segments (row, panel), panel-major, rows ascending
  std::vector<uint32_t> fseg;         // [nseg+1] first element of each segment
  std::vector<uint32_t> fslot;        // [nseg] row-major slot of each segment
  std::vector<float> fval;            // [nnz] values, segment-major
  std::vector<uint16_t> fcol;         // [nnz] column inside the segment's panel
  std::vector<uint32_t> rowslot;      // [rows+1] first slot of each device row
  std::vector<uint32_t> pseg;         // [npanels+1] first segment of each panel
  // backward: groups (panel, chunk, voxel), rows ascending inside a group
  std::vector<uint32_t> bseg;         // [ngroups+1] first entry of each group
  std::vector<uint32_t> bslot;        // [ngroups] voxel-major slot of each group
  std::vector<uint32_t> brow;         // [nnz] device row of each entry
  std::vector<float> bval;            // [nnz] value of each entry
  std::vector<uint32_t> colslot;      // [cols+1] first slot of each panel-order column
  std::vector<uint32_t> pchunk;       // [npanels+1] first chunk (global numbering) of each panel
  std::vector<uint32_t> cgroup;       // [nchunks+1] first group of each chunk
};

// Panel-order column of every file column (voxel x + nx*y + nx*ny*z,
// helper_files_common/scannerconfig.cpp:30-34) for a grid of nx*ny*nz ==
// cols voxels, with the first column of every panel; without a grid (any
// zero) columns keep their order and panels are runs of 16384 columns.
void panel_order(uint32_t cols, const uint32_t grid[3], std::vector<uint32_t>* pc_of_file,
                 std::vector<uint32_t>* pstart);

// Build the tiling from the device-order CSR (row pointers, {column, value
// bits} records) and the panel-order column of every device column.
// threads = 0: hardware concurrency (at most 16).
void build_host_panels(uint32_t rows, uint32_t cols, uint64_t nnz, const uint32_t* rowptr,
                       const uint32_t* csr2 /* [2*nnz] */, const std::vector<uint32_t>& pc_of_dev,
                       const std::vector<uint32_t>& pstart, uint32_t chunk_rows, int threads,
                       HostPanels* out);

}  // namespace mlem
