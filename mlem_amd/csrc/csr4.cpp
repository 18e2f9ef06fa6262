// csr4.cpp — csr4 v4 reader/writer and raw vector I/O (host code).
//
// Parsing order and error messages follow Csr4Matrix::Csr4Matrix
// (helper_files_common/csr4matrix.cpp:32-147); raw vector I/O follows
// Vector<T> (helper_files_common/vector.hpp:33-53,77-87,102-114). Deviations,
// all of them stricter: the "too short" check is done in 64 bits (the
// reference computes it in uint32 and wraps for files over 4 GiB,
// csr4matrix.cpp:85-86), the scanner block is parsed with bounds checks, and
// the row index is checked to be non-decreasing and to end at nnz.
#include "csr4.hpp"

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cerrno>
#include <cstring>
#include <sstream>

#include "mlem.h"

namespace mlem {

namespace {
constexpr uint32_t kMinHeader = 28;       // csr4matrix.hpp:178
constexpr uint32_t kMinScanConfig = 60;   // csr4matrix.hpp:179

[[noreturn]] void fail(int code, const std::string& m) { throw Csr4Error{code, m}; }

bool read_full(int fd, void* dst, size_t n, uint64_t off) {
  char* p = static_cast<char*>(dst);
  while (n > 0) {
    ssize_t got = pread(fd, p, n, (off_t)off);
    if (got < 0 && errno == EINTR) continue;
    if (got <= 0) return false;
    p += got;
    n -= (size_t)got;
    off += (uint64_t)got;
  }
  return true;
}

bool write_full(int fd, const void* src, size_t n) {
  const char* p = static_cast<const char*>(src);
  while (n > 0) {
    ssize_t put = ::write(fd, p, n);
    if (put < 0 && errno == EINTR) continue;
    if (put <= 0) return false;
    p += put;
    n -= (size_t)put;
  }
  return true;
}

// Bounded little-endian cursor over the scanner block.
struct Cursor {
  const char* p;
  const char* end;
  bool ok = true;
  template <class T>
  T get() {
    T v{};
    if (end - p < (ptrdiff_t)sizeof(T)) {
      ok = false;
      return v;
    }
    std::memcpy(&v, p, sizeof(T));
    p += sizeof(T);
    return v;
  }
};
}  // namespace

uint32_t csr4_scanner_bytes(const Csr4Scanner& s) {
  return (uint32_t)(4 + 20 * s.layers.size() + 4 + 4 + 4 * s.blockGaps.size() +
                    4 + 4 * s.ringGaps.size() + 12 + 12);
}

Csr4Scanner csr4_unit_scanner(uint32_t nx, uint32_t ny, uint32_t nz) {
  Csr4Scanner s;
  s.layers.push_back(Csr4Layer{1.f, 1.f, 1.f, 1.f, 0.f});
  s.nBlocks = 1;
  s.grid[0] = nx;
  s.grid[1] = ny;
  s.grid[2] = nz;
  return s;
}

void csr4_check_rowidx(const uint64_t* rowidx, uint32_t nRows, uint64_t nnz) {
  uint64_t prev = 0;
  for (uint32_t r = 0; r < nRows; ++r) {
    if (rowidx[r] < prev || rowidx[r] > nnz) {
      std::stringstream ss;
      ss << "Row index is not a non-decreasing sequence of end offsets <= nnz"
         << " (row " << r << ")";
      fail(MLEM_EFORMAT, ss.str());
    }
    prev = rowidx[r];
  }
  if (prev != nnz) {
    std::stringstream ss;
    ss << "Row index ends at " << prev << " but the header declares nnz=" << nnz;
    fail(MLEM_EFORMAT, ss.str());
  }
}

void csr4_read_header(const std::string& fn, Csr4Header& h,
                      std::vector<uint64_t>& rowidx, int* fdOut) {
  int fd = open(fn.c_str(), O_RDONLY | O_CLOEXEC);
  if (fd < 0) fail(MLEM_EIO, "Cannot open file " + fn);
  struct FdGuard {
    int fd;
    bool keep = false;
    ~FdGuard() {
      if (!keep) close(fd);
    }
  } guard{fd};

  struct stat st;
  if (fstat(fd, &st) != 0) fail(MLEM_EIO, "Cannot open file " + fn);
  h.fileSize = (uint64_t)st.st_size;
  if (h.fileSize < kMinHeader) fail(MLEM_EFORMAT, "File too short " + fn);

  char hdr[kMinHeader];
  if (!read_full(fd, hdr, kMinHeader, 0)) fail(MLEM_EIO, "Cannot read header " + fn);
  std::string magic(hdr, 4);
  h.version = (uint8_t)hdr[4];
  h.flags = (uint8_t)hdr[5];
  std::memcpy(&h.nRows, hdr + 8, 4);
  std::memcpy(&h.nColumns, hdr + 12, 4);
  std::memcpy(&h.nnz, hdr + 16, 8);
  std::memcpy(&h.scanConfigBytes, hdr + 24, 4);

  if (magic != "PCSR")
    fail(MLEM_EFORMAT, "Input file '" + fn + "' is not a CSR matrix");
  if (h.version != 4)
    fail(MLEM_EFORMAT, "Input file '" + fn + "' has wrong CSR version (need v4)");

  // 64-bit size check (no wrap for large files), guarded against overflow.
  const uint64_t maxU64 = ~0ull;
  if (h.nnz > (maxU64 - kMinHeader - h.scanConfigBytes - 8ull * h.nRows) / 8ull)
    fail(MLEM_EFORMAT, "Input file '" + fn + "' declares an impossible nnz");
  uint64_t need = kMinHeader + (uint64_t)h.scanConfigBytes + 8ull * h.nRows + 8ull * h.nnz;
  if (h.fileSize < need) {
    std::stringstream ss;
    ss << "Input file '" << fn << "' is too short: " << h.fileSize << " B"
       << " instead of at least " << need << " B";
    fail(MLEM_EFORMAT, ss.str());
  }
  if (h.scanConfigBytes < kMinScanConfig)
    fail(MLEM_EFORMAT, "Cannot find scanner geometry configuration");

  std::vector<char> scb(h.scanConfigBytes);
  if (!read_full(fd, scb.data(), scb.size(), kMinHeader))
    fail(MLEM_EIO, "Cannot read header " + fn);
  Cursor c{scb.data(), scb.data() + scb.size()};
  Csr4Scanner& s = h.scanner;
  s = Csr4Scanner();
  uint32_t nLayers = c.get<uint32_t>();
  for (uint32_t i = 0; c.ok && i < nLayers; ++i) {
    Csr4Layer L;
    L.w = c.get<float>();
    L.h = c.get<float>();
    L.d = c.get<float>();
    L.r = c.get<float>();
    L.mu = c.get<float>();
    s.layers.push_back(L);
  }
  s.nBlocks = c.get<uint32_t>();
  uint32_t blocksize = c.get<uint32_t>();
  for (uint32_t i = 0; c.ok && blocksize > 0 && i < blocksize - 1; ++i)
    s.blockGaps.push_back(c.get<float>());
  uint32_t nRings = c.get<uint32_t>();
  for (uint32_t i = 0; c.ok && nRings > 0 && i < nRings - 1; ++i)
    s.ringGaps.push_back(c.get<float>());
  for (int k = 0; k < 3; ++k) s.voxel[k] = c.get<float>();
  for (int k = 0; k < 3; ++k) s.grid[k] = c.get<uint32_t>();
  if (!c.ok || blocksize == 0 || nRings == 0 ||
      (size_t)(c.p - scb.data()) != h.scanConfigBytes)
    fail(MLEM_EFORMAT, "Error reading scanner geometry configuration");

  h.rowidxOffset = kMinHeader + (uint64_t)h.scanConfigBytes;
  h.dataOffset = h.rowidxOffset + 8ull * h.nRows;
  rowidx.resize(h.nRows);
  if (h.nRows > 0 && !read_full(fd, rowidx.data(), 8ull * h.nRows, h.rowidxOffset))
    fail(MLEM_EIO, "Cannot read row indexes " + fn);
  csr4_check_rowidx(rowidx.data(), h.nRows, h.nnz);

  if (fdOut) {
    guard.keep = true;
    *fdOut = fd;
  }
}

void csr4_write(const std::string& path, uint8_t flags, uint32_t nRows,
                uint32_t nColumns, uint64_t nnz, const Csr4Scanner& s,
                const uint64_t* rowidx, const Csr4Element* elements) {
  int fd = open(path.c_str(), O_WRONLY | O_CREAT | O_TRUNC | O_CLOEXEC, 0644);
  if (fd < 0)
    fail(MLEM_EIO, "Cannot open file " + path + "; reason: " + std::strerror(errno));
  std::vector<char> head;
  auto put = [&](const void* p, size_t n) {
    const char* c = static_cast<const char*>(p);
    head.insert(head.end(), c, c + n);
  };
  const uint8_t version = 4, pad = 0;
  const uint32_t scb = csr4_scanner_bytes(s);
  put("PCSR", 4);
  put(&version, 1);
  put(&flags, 1);
  put(&pad, 1);
  put(&pad, 1);
  put(&nRows, 4);
  put(&nColumns, 4);
  put(&nnz, 8);
  put(&scb, 4);
  uint32_t nLayers = (uint32_t)s.layers.size();
  put(&nLayers, 4);
  for (const auto& L : s.layers) {
    put(&L.w, 4);
    put(&L.h, 4);
    put(&L.d, 4);
    put(&L.r, 4);
    put(&L.mu, 4);
  }
  put(&s.nBlocks, 4);
  uint32_t bs = s.blocksize();
  put(&bs, 4);
  for (float g : s.blockGaps) put(&g, 4);
  uint32_t nr = s.nRings();
  put(&nr, 4);
  for (float g : s.ringGaps) put(&g, 4);
  put(s.voxel, 12);
  put(s.grid, 12);
  bool ok = write_full(fd, head.data(), head.size()) &&
            (nRows == 0 || write_full(fd, rowidx, 8ull * nRows)) &&
            (nnz == 0 || write_full(fd, elements, 8ull * nnz));
  ok = (close(fd) == 0) && ok;
  if (!ok) fail(MLEM_EIO, "Cannot write file " + path);
}

namespace {
template <class T>
std::vector<T> read_raw(const std::string& fn) {
  int fd = open(fn.c_str(), O_RDONLY | O_CLOEXEC);
  if (fd < 0) fail(MLEM_EIO, "Cannot open file " + fn);
  struct stat st;
  if (fstat(fd, &st) != 0 || (uint64_t)st.st_size % sizeof(T) != 0) {
    close(fd);
    fail(MLEM_EIO, "Cannot open file " + fn);
  }
  std::vector<T> v((size_t)st.st_size / sizeof(T));
  bool ok = v.empty() || read_full(fd, v.data(), (size_t)st.st_size, 0);
  close(fd);
  if (!ok) fail(MLEM_EIO, "Cannot read file " + fn);
  return v;
}
}  // namespace

std::vector<int32_t> read_raw_i32(const std::string& p) { return read_raw<int32_t>(p); }
std::vector<float> read_raw_f32(const std::string& p) { return read_raw<float>(p); }

void write_raw(const std::string& path, const void* data, size_t bytes, bool sync) {
  int fd = open(path.c_str(), O_WRONLY | O_CREAT | O_TRUNC | O_CLOEXEC, 0644);
  if (fd < 0)
    fail(MLEM_EIO, "Cannot open file " + path + "; reason: " + std::strerror(errno));
  bool ok = bytes == 0 || write_full(fd, data, bytes);
  if (ok && sync) ok = fsync(fd) == 0;
  ok = (close(fd) == 0) && ok;
  if (!ok) fail(MLEM_EIO, "Cannot write file " + path);
}

}  // namespace mlem
