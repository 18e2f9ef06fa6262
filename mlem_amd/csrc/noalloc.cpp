// noalloc.cpp — MLEM iterations on device-resident vectors that allocate no
// device memory (for callers that cannot call hipMalloc: code under HIP graph
// capture, tight serving loops, harnesses that forbid allocation).
//
// Everything the iterations need beyond the caller's sinogram and image (the
// device-order copies of both, the ratio vector, the per-workgroup FP64
// statistics partials, their completion counter and results) lives in one
// file-scope device array of kScratchBytes. Each GPU holds its own copy of it
// for the life of the process; its address on a device is looked up once,
// with that device current, and kept. Every call on one GPU shares that GPU's
// array: calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "internal.hpp"
#include "mlem.h"

using namespace mlem;

namespace {

constexpr size_t kScratchBytes = 64ull << 20;  // 64 MiB per GPU
constexpr unsigned kRing = 1024;               // statistics results per read-back

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
std::vector<void*> g_scratch;  // device address of `scratch` on each GPU

// Address of this device's copy of the scratch array (looked up on the first
// call on the device, with the device current).
char* scratch_on(int device) {
  std::lock_guard<std::mutex> lock(g_mu);
  if ((int)g_scratch.size() <= device) g_scratch.resize((size_t)device + 1, nullptr);
  if (!g_scratch[(size_t)device]) {
    int cur = 0;
    MLEM_HIP(hipGetDevice(&cur));
    if (cur != device) MLEM_HIP(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    MLEM_HIP(e);
    if (!p) raise(MLEM_EDEVICE, "scratch array has no device address");
    g_scratch[(size_t)device] = p;
  }
  return static_cast<char*>(g_scratch[(size_t)device]);
}

size_t take(size_t* off, size_t bytes) {
  const size_t at = *off;
  *off = (at + bytes + 255) & ~(size_t)255;
  return at;
}

}  // namespace

extern "C" {

int mlem_reconstruct_noalloc(mlem_matrix* h, const int32_t* y, uint32_t rows, float* x,
                             uint32_t cols, int iterations, int mode, double* loglik,
                             double* fwd_total) {
  return guarded([&] {
    if (!h) raise(MLEM_EINVAL, "null matrix handle");
    DevMatrix& m = h->m;
    if (rows != m.rows || cols != m.cols)
      raise(MLEM_EINVAL, "vector sizes (" + std::to_string(rows) + " rows, " +
                             std::to_string(cols) + " columns) do not match the matrix (" +
                             std::to_string(m.rows) + " x " + std::to_string(m.cols) + ")");
    if (m.streamed) raise(MLEM_EINVAL, "mlem_reconstruct_noalloc needs a matrix held in HBM");
    if (!m.csc) raise(MLEM_ESTATE, "matrix was opened without a CSC transpose");
    if (mode != MLEM_MODE_FAST && mode != MLEM_MODE_EXACT)
      raise(MLEM_EINVAL, "mlem_reconstruct_noalloc runs MLEM_MODE_FAST or MLEM_MODE_EXACT");
    if (iterations < 0) raise(MLEM_EINVAL, "negative iteration count");
    if ((rows && !y) || (cols && !x)) raise(MLEM_EINVAL, "null vector");
    const bool stats = loglik || fwd_total;
    // scratch layout: y, x (device order), r, statistics partials, counter, results
    const unsigned maxb = stats_max_blocks(m);
    size_t off = 0;
    const size_t oy = take(&off, 4ull * rows), ox = take(&off, 4ull * cols);
    const size_t orr = take(&off, 4ull * rows), op = take(&off, sizeof(double2) * maxb);
    const size_t oc = take(&off, 2 * sizeof(unsigned)), ores = take(&off, sizeof(double2) * kRing);
    if (off > kScratchBytes)
      raise(MLEM_EINVAL, "mlem_reconstruct_noalloc needs " + std::to_string(off) +
                             " bytes of device scratch for this matrix, more than its " +
                             std::to_string(kScratchBytes) + " (use mlem_reconstruct)");
    char* base = scratch_on(m.device);
    int32_t* ys = reinterpret_cast<int32_t*>(base + oy);
    float* xs = reinterpret_cast<float*>(base + ox);
    float* rs = reinterpret_cast<float*>(base + orr);
    StatsState st;
    st.partials = reinterpret_cast<double2*>(base + op);
    st.counter = reinterpret_cast<unsigned*>(base + oc);
    st.results = reinterpret_cast<double2*>(base + ores);
    st.ring = kRing;
    st.max_blocks = maxb;
    hipStream_t s = m.stream;
    rows_to_device(m, y, ys, s);
    to_device_cols(m, x, xs, s);
    // the results slot counter starts at 0 (a fill kernel, so that the call
    // can be captured into a graph: no memset nodes)
    launch_fill(reinterpret_cast<float*>(st.counter), 0.f, 2, s);
    double2 res[kRing];
    for (int k = 0; k < iterations; ++k) {
      launch_fwd(m, mode, xs, ys, rs, true, stats ? &st : nullptr, s);
      launch_bwd(m, mode, rs, m.norm, xs, true, s);
      if (stats && ((k + 1) % (int)kRing == 0 || k + 1 == iterations)) {
        const int first = k - k % (int)kRing, n = k + 1 - first;
        MLEM_HIP(hipMemcpyAsync(res, st.results, sizeof(double2) * n, hipMemcpyDeviceToHost, s));
        MLEM_HIP(hipStreamSynchronize(s));
        for (int j = 0; j < n; ++j) {
          if (loglik) loglik[first + j] = res[j].x;
          if (fwd_total) fwd_total[first + j] = res[j].y;
        }
      }
    }
    to_file_cols(m, xs, x, s);
    if (stats) MLEM_HIP(hipStreamSynchronize(s));
  });
}

}  // extern "C"
