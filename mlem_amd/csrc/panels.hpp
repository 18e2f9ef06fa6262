// panels.hpp — the system matrix as a two-dimensional tiling for products
// with many right-hand sides (host-side construction, no GPU needed).
//
// A product with W interleaved right-hand sides needs one W*4-byte run of
// the dense operand per matrix element (128 bytes at W = 32). Gathered from
// the caches element by element, those runs are tens of GB per product. The
// tiling cuts the voxel grid into boxes small enough that one box's slice of
// a W-wide block fits in a workgroup's LDS (16x16x4 = 1024 voxels, 128 KB at
// W = 32), so that the runs are read from LDS and reused by every row that
// crosses the box:
//
//   pieces: every row is cut into pieces, one per box it crosses (about 10
//   per row on the 128x128x48 grid, 12 elements each); a piece's elements
//   keep their file order and carry their column as a 16-bit offset inside
//   the box.
//
//   forward (f = A X): a workgroup loads its box's slice of X into LDS and
//   writes one W-wide dot product per piece into the piece's slot. The slots
//   of one row are contiguous, in box order, so a second pass adds them in a
//   fixed order. Inside a box the pieces are stored longest first, so the
//   lanes of a wave get pieces of about the same length, and staged in LDS
//   in chunks (at most 256 pieces and 4096 elements).
//
//   backward (u = A^T R): a workgroup holds its box's voxel sums in
//   registers and walks the box's pieces (rows ascending) in chunks: the
//   chunk's runs of R are staged in LDS, then every voxel adds its elements
//   of the chunk (chunk order, then rows ascending). Every voxel's sum is
//   complete at the end, one slot per voxel.
//
// Sums are deterministic (fixed order), not the sequential order of the
// reference driver (mlem_seq/seqcsr4mlem.cpp:69-120).
#pragma once

#include <cstdint>
#include <vector>

namespace mlem {

constexpr uint32_t kBoxVox = 1024;      // voxels per box (at most)
constexpr uint32_t kMaxSegment = 512;   // longer pieces are split, each with its own slot
constexpr uint32_t kChunkPieces = 512;   // backward: pieces staged per chunk (at most)
constexpr uint32_t kChunkElems = 8192;   // backward: elements staged per chunk (at most)
constexpr uint32_t kFChunkPieces = 256;  // forward: pieces staged per chunk (at most)
constexpr uint32_t kFChunkElems = 4096;  // forward: elements staged per chunk (at most)
constexpr uint32_t kFAlign = 4;          // forward chunks start at multiples of 4 elements
constexpr uint32_t kBAlign = 8;          // backward chunks start at multiples of 8 elements

struct HostPanels {
  uint32_t rows = 0, cols = 0;
  uint64_t nnz = 0;
  uint32_t npanels = 0;  // boxes
  // column order of the tiling ("panel order"): boxes one after the other
  std::vector<uint32_t> pc_of_dev;    // [cols] panel-order column of each device column
  std::vector<uint32_t> pstart;       // [npanels+1] first panel-order column of each box
  // forward: pieces box-major, longest first inside a box
  std::vector<uint32_t> pseg;         // [npanels+1] first piece of each box
  std::vector<uint32_t> fseg;         // [nseg+1] first element of each piece
  std::vector<uint32_t> fslot;        // [nseg] row-major slot of each piece
  std::vector<float> fval;            // values, piece-major, chunks aligned (zero gaps)
  std::vector<uint16_t> fcol;         // column inside the piece's box (0 in the gaps)
  std::vector<uint32_t> rowslot;      // [rows+1] first slot of each device row
  std::vector<uint32_t> pfchunk;      // [npanels+1] first forward chunk of each box
  std::vector<uint32_t> fchunk_piece; // [nfchunks+1] first piece of each forward chunk
  // backward: pieces box-major, rows ascending inside a box, in chunks
  std::vector<uint32_t> brow;         // [nseg] device row of each piece
  std::vector<uint32_t> pchunk;       // [npanels+1] first chunk of each box
  std::vector<uint32_t> chunk_piece;  // [nchunks+1] first piece of each chunk
  std::vector<uint32_t> chunk_elem;   // [nchunks+1] first element of each chunk
  std::vector<uint16_t> boff;         // [nchunks*(kBoxVox+1)] element offset of each
                                      //  in-box voxel, relative to the chunk's first
  std::vector<uint16_t> bpc;          // piece of each element inside its chunk, chunks
                                      //  aligned (zero gaps)
  std::vector<float> bval;            // value of each element
  std::vector<uint32_t> border;       // [npanels] launch order: decreasing element count
};

// Panel-order column of every file column (voxel x + nx*y + nx*ny*z,
// helper_files_common/scannerconfig.cpp:30-34) for a grid of nx*ny*nz ==
// cols voxels, with the first column of every box; without a grid (any
// zero) boxes are runs of kBoxVox columns. Inside a box the voxels are
// numbered so that the backward's lane groups (8 consecutive positions each,
// 8 groups per wave) get balanced work: with count[c] = elements of file
// column c (may be null), the voxels sorted by count are dealt to the
// groups in snake order, one position per group per round. boxgrid receives
// the number of boxes along x, y, z, or {P, 1, 1} without a grid.
void panel_order(uint32_t cols, const uint32_t grid[3], const uint32_t* count,
                 std::vector<uint32_t>* pc_of_file, std::vector<uint32_t>* pstart,
                 uint32_t boxgrid[3]);

// Build the tiling from the device-order CSR (row pointers, {column, value
// bits} records) and the panel-order column of every device column.
// boxgrid: as returned by panel_order. threads = 0: hardware concurrency
// (at most 16).
void build_host_panels(uint32_t rows, uint32_t cols, uint64_t nnz, const uint32_t* rowptr,
                       const uint32_t* csr2 /* [2*nnz] */, const std::vector<uint32_t>& pc_of_dev,
                       const std::vector<uint32_t>& pstart, const uint32_t boxgrid[3],
                       int threads, HostPanels* out);

}  // namespace mlem
