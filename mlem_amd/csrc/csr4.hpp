// csr4.hpp — host-side reader/writer of the csr4 v4 system-matrix format.
//
// Layout (little endian), as parsed by Csr4Matrix::Csr4Matrix
// (helper_files_common/csr4matrix.cpp:32-147):
//   0  char[4] "PCSR"            4 u8 version (=4)   5 u8 flags (bit0 x, bit1 y,
//   bit2 z symmetry, csr4matrix.hpp:180-182)   6-7 pad   8 u32 nRows
//   12 u32 nColumns   16 u64 nnz   24 u32 scanConfigBytes
//   28 scanner block (csr4matrix.cpp:100-129): u32 nLayers, 5 f32 per layer,
//      u32 nBlocks, u32 blocksize, (blocksize-1) f32 gaps, u32 nRings,
//      (nRings-1) f32 gaps, 3 f32 voxel size, 3 u32 grid dims
//   28+scb  u64 rowidx[nRows]: exclusive END offset of every row
//   then    {u32 column, f32 value}[nnz] (matrixelement.hpp:37-46)
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace mlem {

struct Csr4Element {
  uint32_t column;
  float value;
};
static_assert(sizeof(Csr4Element) == 8, "csr4 element is 8 bytes");

struct Csr4Layer {
  float w, h, d, r, mu;
};

struct Csr4Scanner {
  std::vector<Csr4Layer> layers;
  uint32_t nBlocks = 1;
  std::vector<float> blockGaps;  // blocksize-1 entries
  std::vector<float> ringGaps;   // nRings-1 entries
  float voxel[3] = {1.f, 1.f, 1.f};
  uint32_t grid[3] = {0, 0, 0};
  uint32_t blocksize() const { return (uint32_t)blockGaps.size() + 1; }
  uint32_t nRings() const { return (uint32_t)ringGaps.size() + 1; }
};

struct Csr4Header {
  uint8_t version = 4;
  uint8_t flags = 0;
  uint32_t nRows = 0;
  uint32_t nColumns = 0;
  uint64_t nnz = 0;
  uint32_t scanConfigBytes = 0;
  Csr4Scanner scanner;
  uint64_t rowidxOffset = 0;   // file offset of rowidx[]
  uint64_t dataOffset = 0;     // file offset of the element array
  uint64_t fileSize = 0;
};

// Thrown by the reader; the C ABI turns it into a status code + message.
struct Csr4Error {
  int code;  // MLEM_EIO or MLEM_EFORMAT
  std::string msg;
};

// Open `path`, parse and validate the header and scanner block, read the row
// index. The file descriptor stays open in *fd for streaming the elements.
// Throws Csr4Error.
void csr4_read_header(const std::string& path, Csr4Header& h,
                      std::vector<uint64_t>& rowidx, int* fd);

// Validate a row index: non-decreasing, last entry == nnz. Throws Csr4Error.
void csr4_check_rowidx(const uint64_t* rowidx, uint32_t nRows, uint64_t nnz);

// Serialized size of a scanner block.
uint32_t csr4_scanner_bytes(const Csr4Scanner& s);

// Write a complete csr4 file. Throws Csr4Error.
void csr4_write(const std::string& path, uint8_t flags, uint32_t nRows,
                uint32_t nColumns, uint64_t nnz, const Csr4Scanner& scanner,
                const uint64_t* rowidx, const Csr4Element* elements);

// Minimal scanner block for a unit-voxel grid: 1 layer, blocksize 1, 1 ring
// (60 bytes, the smallest block the reader accepts).
Csr4Scanner csr4_unit_scanner(uint32_t nx, uint32_t ny, uint32_t nz);

// Raw headerless binary vectors (helper_files_common/vector.hpp:33-87).
// Reads `path` as T[]; throws Csr4Error("Cannot open file <path>") when the
// file is missing or its size is not a multiple of sizeof(T).
std::vector<int32_t> read_raw_i32(const std::string& path);
std::vector<float> read_raw_f32(const std::string& path);
void write_raw(const std::string& path, const void* data, size_t bytes,
               bool sync);

}  // namespace mlem
