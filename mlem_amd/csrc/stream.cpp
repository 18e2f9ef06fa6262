// stream.cpp — out-of-core matrices: row blocks in pinned host memory, each
// projection streaming its part of every block over PCIe.
//
// Reference counterpart: mlem_cuda/versions/version_7.cu:205-260 (the matrix
// split into row parts that are copied to the device in turn before each
// projection) with the splitting of mlem_cuda/helper_files/helper_v_7_8.cu:4-72
// (parts sized by the free device memory, balanced by elements). Here every
// block also carries the CSC of its rows, so the backprojection gathers per
// column like the in-core path instead of scattering; a column's sum continues
// from block to block in ascending row order, which keeps EXACT mode equal to
// the sequential sums of seqcsr4mlem.cpp:37-49,108-120.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "internal.hpp"
#include "mlem.h"

namespace mlem {

namespace {

uint64_t align16(uint64_t n) { return (n + 15) & ~15ull; }

// Largest element count of one block: the kernels address a block's elements
// with 32-bit offsets.
constexpr uint64_t kMaxBlockNnz = (1ull << 32) - 1024;

}  // namespace

StreamedMatrix::~StreamedMatrix() {
  for (StreamBlock& b : blocks)
    if (b.resident) (void)hipFree(b.resident);
  for (int i = 0; i < 2; ++i) {
    if (copied[i]) (void)hipEventDestroy(copied[i]);
    if (freed[i]) (void)hipEventDestroy(freed[i]);
    if (dev[i]) (void)hipFree(dev[i]);
  }
  if (u) (void)hipFree(u);
  if (csc_dev) (void)hipFree(csc_dev);
  if (copy) (void)hipStreamDestroy(copy);
  if (host) (void)hipHostFree(host);
}

StreamedMatrix* build_streamed(uint32_t rows, uint32_t cols, uint64_t nnz,
                               const uint64_t* rowidx, const ElementFetch& fetch,
                               uint64_t block_bytes, uint64_t resident_bytes, hipStream_t s) {
  const uint64_t colpart = align16(4ull * ((uint64_t)cols + 1));
  if (block_bytes == 0) {
    // an eighth of the free device memory per part, as the reference sizes
    // its parts by cudaMemGetInfo (helper_v_7_8.cu:6-30): the two part
    // buffers and one block's device CSC with its sort scratch (about four
    // parts) then fit
    size_t freeb = 0, total = 0;
    MLEM_HIP(hipMemGetInfo(&freeb, &total));
    block_bytes = freeb / 8;
  }
  auto row_len = [&](uint32_t r) { return rowidx[r] - (r ? rowidx[r - 1] : 0); };

  // Greedy row blocks: both parts of a block fit `block_bytes`.
  std::unique_ptr<StreamedMatrix> S(new StreamedMatrix());
  uint64_t off = 0;
  for (uint32_t r0 = 0; r0 < rows || (rows == 0 && S->blocks.empty());) {
    StreamBlock B;
    B.row0 = r0;
    uint64_t n = 0;
    uint32_t r = r0;
    while (r < rows) {
      const uint64_t n2 = n + row_len(r);
      const uint64_t csr = align16(4ull * ((uint64_t)(r - r0) + 2)) + 8 * n2;
      const uint64_t csc = colpart + 8 * n2;
      if (r > r0 && (csr > block_bytes || csc > block_bytes || n2 > kMaxBlockNnz)) break;
      if (r == r0 && (csr > block_bytes || csc > block_bytes || n2 > kMaxBlockNnz))
        raise(MLEM_EINVAL, "block_bytes (" + std::to_string(block_bytes) + ") cannot hold row " +
                               std::to_string(r) + " and the column pointers of a block");
      n = n2;
      ++r;
    }
    B.rows = r - r0;
    B.nnz = n;
    B.csr_off = off;
    B.csr_bytes = align16(4ull * ((uint64_t)B.rows + 1)) + 8 * n;
    off += align16(B.csr_bytes);
    B.csc_off = off;
    B.csc_bytes = colpart + 8 * n;
    off += align16(B.csc_bytes);
    S->part_bytes = std::max(S->part_bytes, std::max(B.csr_bytes, B.csc_bytes));
    S->blocks.push_back(B);
    if (rows == 0) break;
    r0 = r;
  }
  S->host_bytes = std::max<uint64_t>(off, 16);
  MLEM_HIP(hipHostMalloc((void**)&S->host, S->host_bytes, hipHostMallocDefault));

  // Fill the blocks in parallel: CSR part from the source, CSC part by a
  // counting sort over the block's columns (rows ascending in every column).
  const unsigned nthreads = (unsigned)std::max<size_t>(
      1, std::min<size_t>({S->blocks.size(), 16, std::max(1u, std::thread::hardware_concurrency())}));
  std::atomic<uint64_t> bad{0};
  std::vector<std::string> errors(nthreads);
  auto work = [&](unsigned tid) {
    try {
      std::vector<uint32_t> next(cols);
      for (size_t b = tid; b < S->blocks.size(); b += nthreads) {
        const StreamBlock& B = S->blocks[b];
        char* csr = S->host + B.csr_off;
        char* csc = S->host + B.csc_off;
        uint32_t* rp = reinterpret_cast<uint32_t*>(csr);
        const uint64_t base = B.row0 ? rowidx[B.row0 - 1] : 0;
        rp[0] = 0;
        for (uint32_t k = 0; k < B.rows; ++k) rp[k + 1] = (uint32_t)(rowidx[B.row0 + k] - base);
        uint2* el = reinterpret_cast<uint2*>(csr + align16(4ull * ((uint64_t)B.rows + 1)));
        if (B.nnz) fetch(el, base, B.nnz);
        uint32_t* cp = reinterpret_cast<uint32_t*>(csc);
        std::fill(cp, cp + (uint64_t)cols + 1, 0u);
        uint64_t nbad = 0;
        for (uint64_t k = 0; k < B.nnz; ++k) {
          if (el[k].x >= cols)
            ++nbad;
          else
            ++cp[el[k].x + 1];
        }
        if (nbad) {
          bad += nbad;
          continue;
        }
        for (uint32_t j = 0; j < cols; ++j) cp[j + 1] += cp[j];
        std::copy(cp, cp + cols, next.begin());
        uint2* ce = reinterpret_cast<uint2*>(csc + colpart);
        for (uint32_t k = 0; k < B.rows; ++k)
          for (uint32_t e = rp[k]; e < rp[k + 1]; ++e)
            ce[next[el[e].x]++] = make_uint2(B.row0 + k, el[e].y);
      }
    } catch (const ApiError& e) {
      errors[tid] = e.msg;
    } catch (const std::exception& e) {
      errors[tid] = e.what();
    }
  };
  std::vector<std::thread> pool;
  for (unsigned t = 1; t < nthreads; ++t) pool.emplace_back(work, t);
  work(0);
  for (auto& t : pool) t.join();
  for (const std::string& e : errors)
    if (!e.empty()) raise(MLEM_EIO, e);
  if (bad)
    raise(MLEM_EFORMAT, std::to_string(bad.load()) + " matrix elements have a column index >= " +
                            std::to_string(cols));

  // Blocks kept in device memory: the first ones whose two parts fit the
  // resident budget; only the others are streamed.
  uint64_t kept = 0, streamed_part = 0;
  for (StreamBlock& B : S->blocks) {
    const uint64_t both = align16(B.csr_bytes) + B.csc_bytes;
    if (kept + both <= resident_bytes && streamed_part == 0) {
      MLEM_HIP(hipMalloc((void**)&B.resident, both));
      MLEM_HIP(hipMemcpyAsync(B.resident, S->host + B.csr_off, B.csr_bytes,
                              hipMemcpyHostToDevice, s));
      MLEM_HIP(hipMemcpyAsync(B.resident + align16(B.csr_bytes), S->host + B.csc_off,
                              B.csc_bytes, hipMemcpyHostToDevice, s));
      kept += both;
    } else {
      streamed_part = std::max(streamed_part, std::max(B.csr_bytes, B.csc_bytes));
      S->max_stream_nnz = std::max(S->max_stream_nnz, B.nnz);
    }
  }
  S->part_bytes = streamed_part;
  for (int i = 0; i < 2; ++i) {
    MLEM_HIP(hipMalloc((void**)&S->dev[i], std::max<uint64_t>(S->part_bytes, 16)));
    MLEM_HIP(hipEventCreateWithFlags(&S->copied[i], hipEventDisableTiming));
    MLEM_HIP(hipEventCreateWithFlags(&S->freed[i], hipEventDisableTiming));
    MLEM_HIP(hipEventRecord(S->freed[i], s));
  }
  MLEM_HIP(hipStreamCreateWithFlags(&S->copy, hipStreamNonBlocking));
  MLEM_HIP(hipMalloc((void**)&S->u, 4ull * std::max<uint32_t>(cols, 1)));
  MLEM_HIP(hipMemsetAsync(S->u, 0, 4ull * std::max<uint32_t>(cols, 1), s));
  if (S->max_stream_nnz > 0) {
    // device CSC of one streamed block; if it does not fit, iterations keep
    // streaming the host-built CSC parts
    S->csc_scratch_bytes = block_csc_scratch(S->max_stream_nnz, cols);
    const uint64_t bytes = colpart + align16(8 * S->max_stream_nnz) + S->csc_scratch_bytes;
    if (hipMalloc((void**)&S->csc_dev, bytes) != hipSuccess) {
      (void)hipGetLastError();
      S->csc_dev = nullptr;
    }
  }
  // lane groups of about four elements per lane (4..64), as the in-core
  // lane-group kernels choose them
  auto lanes_for = [](double avg) {
    int g = 4;
    while (g < 64 && g * 4 < avg) g *= 2;
    return g;
  };
  const uint64_t nb = S->blocks.size();
  S->lanes_fwd = lanes_for(rows ? (double)nnz / rows : 1.0);
  S->lanes_bwd = lanes_for(cols && nb ? (double)nnz / ((double)cols * (double)nb) : 1.0);
  MLEM_HIP(hipStreamSynchronize(s));
  return S.release();
}

namespace {

// Run `kernel(block, device part)` for every block on stream s, the copy of
// each block's part (CSR or CSC) on the copy stream: block b waits for its
// copy, the copy into a buffer waits for the kernels of the block before.
template <class K>
void stream_pass(const DevMatrix& m, bool csc, K&& kernel, hipStream_t s) {
  StreamedMatrix& S = *m.streamed;
  size_t n_streamed = 0;
  for (size_t b = 0; b < S.blocks.size(); ++b) {
    const StreamBlock& B = S.blocks[b];
    if (B.resident) {  // in device memory: no copy
      kernel(B, csc ? B.resident + align16(B.csr_bytes) : B.resident);
      continue;
    }
    const int i = (int)(n_streamed++ & 1);
    const uint64_t off = csc ? B.csc_off : B.csr_off;
    const uint64_t n = csc ? B.csc_bytes : B.csr_bytes;
    MLEM_HIP(hipStreamWaitEvent(S.copy, S.freed[i], 0));
    MLEM_HIP(hipMemcpyAsync(S.dev[i], S.host + off, n, hipMemcpyHostToDevice, S.copy));
    MLEM_HIP(hipEventRecord(S.copied[i], S.copy));
    MLEM_HIP(hipStreamWaitEvent(s, S.copied[i], 0));
    kernel(B, S.dev[i]);
    MLEM_HIP(hipEventRecord(S.freed[i], s));
  }
}

}  // namespace

void stream_fwd(const DevMatrix& m, int mode, const float* x, const int32_t* y, float* out,
                bool ratio, const StatsState* st, hipStream_t s) {
  const StreamedMatrix& S = *m.streamed;
  const bool exact = mode == MLEM_MODE_EXACT;
  stream_pass(
      m, false,
      [&](const StreamBlock& B, char* d) {
        // every block launches (also an empty one): statistics expect one
        // result per block
        const uint32_t* rp = reinterpret_cast<const uint32_t*>(d);
        const uint2* el = reinterpret_cast<const uint2*>(d + align16(4ull * ((uint64_t)B.rows + 1)));
        launch_fwd_block(rp, el, B.rows, S.lanes_fwd, exact, x, y ? y + B.row0 : nullptr,
                         out + B.row0, ratio, st, m.cus, s);
      },
      s);
}

void stream_bwd_add(const DevMatrix& m, bool exact, const float* r, float* u, hipStream_t s) {
  const StreamedMatrix& S = *m.streamed;
  if (m.cols == 0) return;
  const uint64_t colpart = align16(4ull * ((uint64_t)m.cols + 1));
  stream_pass(
      m, true,
      [&](const StreamBlock& B, char* d) {
        (void)B;
        launch_bwd_block_add(reinterpret_cast<const uint32_t*>(d),
                             reinterpret_cast<const uint2*>(d + colpart), m.cols, S.lanes_bwd, exact,
                             r, u, m.cus, s);
      },
      s);
}

void stream_fwd_scatter(const DevMatrix& m, const float* x, const int32_t* y, float* r, float* u,
                        const StatsState* st, hipStream_t s) {
  const StreamedMatrix& S = *m.streamed;
  stream_pass(
      m, false,
      [&](const StreamBlock& B, char* d) {
        const uint32_t* rp = reinterpret_cast<const uint32_t*>(d);
        const uint2* el = reinterpret_cast<const uint2*>(d + align16(4ull * ((uint64_t)B.rows + 1)));
        launch_fwd_block(rp, el, B.rows, S.lanes_fwd, false, x, y + B.row0, r + B.row0, true, st,
                         m.cus, s);
        if (B.rows) launch_scatter_block(rp, el, B.rows, S.lanes_fwd, r + B.row0, u, m.cus, s);
      },
      s);
}

void stream_iteration(const DevMatrix& m, int mode, float* x, const int32_t* y, float* r,
                      const StatsState* st, hipStream_t s) {
  const StreamedMatrix& S = *m.streamed;
  if (!S.csc_dev) {
    stream_fwd(m, mode, x, y, r, true, st, s);
    stream_bwd(m, mode, r, m.norm, x, true, s);
    return;
  }
  const bool exact = mode == MLEM_MODE_EXACT;
  const uint64_t colpart = align16(4ull * ((uint64_t)m.cols + 1));
  uint32_t* colptr = reinterpret_cast<uint32_t*>(S.csc_dev);
  uint2* csc = reinterpret_cast<uint2*>(S.csc_dev + colpart);
  void* scratch = S.csc_dev + colpart + align16(8 * S.max_stream_nnz);
  launch_fill(S.u, 0.f, m.cols, s);
  stream_pass(
      m, false,
      [&](const StreamBlock& B, char* d) {
        const uint32_t* rp = reinterpret_cast<const uint32_t*>(d);
        const uint2* el = reinterpret_cast<const uint2*>(d + align16(4ull * ((uint64_t)B.rows + 1)));
        launch_fwd_block(rp, el, B.rows, S.lanes_fwd, exact, x, y + B.row0, r + B.row0, true, st,
                         m.cus, s);
        if (B.resident) {  // its CSC part is in device memory already (global rows)
          const char* c = B.resident + align16(B.csr_bytes);
          launch_bwd_block_add(reinterpret_cast<const uint32_t*>(c),
                               reinterpret_cast<const uint2*>(c + colpart), m.cols, S.lanes_bwd,
                               exact, r, S.u, m.cus, s);
        } else {  // block-local rows
          build_block_csc(rp, el, B.rows, m.cols, B.nnz, scratch, S.csc_scratch_bytes, colptr, csc,
                          s);
          launch_bwd_block_add(colptr, csc, m.cols, S.lanes_bwd, exact, r + B.row0, S.u, m.cus, s);
        }
      },
      s);
  launch_update(m.cols, S.u, m.norm, x, false, s);
}

void stream_bwd(const DevMatrix& m, int mode, const float* r, const float* norm, float* out,
                bool update, hipStream_t s) {
  float* acc = update ? m.streamed->u : out;
  launch_fill(acc, 0.f, m.cols, s);
  stream_bwd_add(m, mode == MLEM_MODE_EXACT, r, acc, s);
  if (update) launch_update(m.cols, acc, norm, out, false, s);
}

}  // namespace mlem
