// api.cpp — C ABI of libmlem.so: matrix loading into HBM, step functions,
// reconstruction driver, checkpoint/restart, timing.
//
// Reference counterparts: Csr4Matrix (helper_files_common/csr4matrix.cpp),
// mlem() and its step functions (mlem_seq/seqcsr4mlem.cpp:37-189), and the
// checkpoint/restart logic of mlem_mpi/mpicsr4mlem.cpp:95-154,536-553 with
// the interval meaning of mlem_cuda/helper_files/hybCalcFunc.cu:852-853.
#include <hip/hip_runtime.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "csr4.hpp"
#include "internal.hpp"
#include "mlem.h"
#include "symmetry.hpp"

using namespace mlem;

namespace mlem {

[[noreturn]] void raise(int code, const std::string& msg) { throw ApiError{code, msg}; }

void hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return;
  const int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? MLEM_ENOMEM
                                                                                : MLEM_EDEVICE;
  raise(code, std::string(what) + " failed: " + hipGetErrorString(e));
}

}  // namespace mlem

namespace {
thread_local std::string g_error;
}  // namespace

namespace mlem {
void set_error(const std::string& msg) { g_error = msg; }
void clear_error() { g_error.clear(); }
}  // namespace mlem

namespace {

int env_int(const char* name, int dflt) {
  const char* v = std::getenv(name);
  if (!v || !*v) return dflt;
  return std::atoi(v);
}

double seconds_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Pick the lanes-per-row group from the average segment length: about four
// elements per lane per pass, clamped to 4..64.
int lanes_for(double avg) {
  int g = 4;
  while (g < 64 && g * 4 < avg) g *= 2;
  return g;
}

void select_device(int* dev, int* cus) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= 0)
    raise(MLEM_EDEVICE, "no GPU is visible to the HIP runtime (libmlem needs an MI355X, gfx950)");
  MLEM_HIP(hipGetDevice(dev));
  hipDeviceProp_t p;
  MLEM_HIP(hipGetDeviceProperties(&p, *dev));
  if (std::strncmp(p.gcnArchName, "gfx950", 6) != 0)
    raise(MLEM_EDEVICE, std::string("libmlem is built for gfx950; device is ") + p.gcnArchName);
  *cus = p.multiProcessorCount > 0 ? p.multiProcessorCount : 256;
}

template <class T>
void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

// Pinned double-buffered host->device streaming of a large byte range. The
// producer fills a pinned chunk (file read or memcpy); the copy engine drains
// it while the next chunk is produced.
template <class Produce>
void stream_to_device(void* dst, uint64_t bytes, hipStream_t s, Produce&& produce) {
  if (bytes == 0) return;
  const uint64_t kChunk = 64ull << 20;
  void* pin[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  struct Cleanup {
    void** pin;
    hipEvent_t* ev;
    ~Cleanup() {
      for (int i = 0; i < 2; ++i) {
        if (ev[i]) (void)hipEventDestroy(ev[i]);
        if (pin[i]) (void)hipHostFree(pin[i]);
      }
    }
  } cleanup{pin, ev};
  for (int i = 0; i < 2; ++i) {
    MLEM_HIP(hipHostMalloc(&pin[i], kChunk, hipHostMallocDefault));
    MLEM_HIP(hipEventCreateWithFlags(&ev[i], hipEventDisableTiming));
  }
  bool used[2] = {false, false};
  char* d = static_cast<char*>(dst);
  for (uint64_t off = 0, i = 0; off < bytes; off += kChunk, ++i) {
    const int b = (int)(i & 1);
    const uint64_t n = std::min(kChunk, bytes - off);
    if (used[b]) MLEM_HIP(hipEventSynchronize(ev[b]));
    produce(pin[b], off, n);
    MLEM_HIP(hipMemcpyAsync(d + off, pin[b], n, hipMemcpyHostToDevice, s));
    MLEM_HIP(hipEventRecord(ev[b], s));
    used[b] = true;
  }
  MLEM_HIP(hipStreamSynchronize(s));
}

}  // namespace

// --------------------------------------------------------------- handles --
mlem_matrix::~mlem_matrix() {
  {
    dfree(m.rowptr);
    dfree(m.csr);
    dfree(m.colptr);
    dfree(m.csc);
    dfree(m.norm);
    dfree(m.sx);
    dfree(m.su);
    dfree(m.sf);
    dfree(m.sr);
    dfree(m.sy);
    dfree(m.tx);
    dfree(m.tf);
    dfree(m.colmap);
    dfree(m.rowold);
    dfree(m.rowpos);
    dfree(m.rscale);
    dfree(m.tiles_fwd);
    dfree(m.tiles_bwd);
    free_dict(m.dfwd);
    free_dict(m.dbwd);
    delete m.streamed;
    m.streamed = nullptr;
    free_panels(m);
    if (m.stream) (void)hipStreamDestroy(m.stream);
  }
}

struct mlem_recon {
  mlem_matrix* mat = nullptr;
  int mode = MLEM_MODE_FAST;
  int done = 0;
  bool use_graph = false;
  bool want_stats = false;
  int32_t* y = nullptr;
  float* x = nullptr;
  float* r = nullptr;
  float* u = nullptr;   // ATOMIC mode accumulator
  float* xs = nullptr;  // file-order image staging
  StatsState st;
  std::vector<double> loglik, fwdtot;  // per completed iteration
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint64_t graph_gen = 0;  // matrix launch generation the graph was captured at
  std::vector<hipEvent_t> events;
  ~mlem_recon() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (hipEvent_t e : events) (void)hipEventDestroy(e);
    dfree(y);
    dfree(x);
    dfree(r);
    dfree(u);
    dfree(xs);
    dfree(st.partials);
    dfree(st.counter);
    dfree(st.results);
  }
};

namespace {

void check_dims(const mlem_matrix* h, uint32_t rows, uint32_t cols) {
  if (!h) raise(MLEM_EINVAL, "null matrix handle");
  if (rows != h->m.rows || cols != h->m.cols)
    raise(MLEM_EINVAL, "vector sizes (" + std::to_string(rows) + " rows, " + std::to_string(cols) +
                           " columns) do not match the matrix (" + std::to_string(h->m.rows) +
                           " x " + std::to_string(h->m.cols) + ")");
}

// Time one kernel variant (median of a few launches) with HIP events.
template <class Launch>
float time_variant(hipStream_t s, Launch&& launch, int reps) {
  hipEvent_t a, b;
  MLEM_HIP(hipEventCreate(&a));
  MLEM_HIP(hipEventCreate(&b));
  std::vector<float> t;
  launch();  // warm
  for (int i = 0; i < reps; ++i) {
    MLEM_HIP(hipEventRecord(a, s));
    launch();
    MLEM_HIP(hipEventRecord(b, s));
    MLEM_HIP(hipEventSynchronize(b));
    float ms = 0.f;
    MLEM_HIP(hipEventElapsedTime(&ms, a, b));
    t.push_back(ms);
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  std::sort(t.begin(), t.end());
  return t[t.size() / 2];
}

// Launch configuration of a kernel family and its parameter (validated).
GroupLaunch make_launch(int kind, int param) {
  GroupLaunch g;
  g.kind = kind;
  if (kind == kTiles) {
    if (param != 4 && param != 8 && param != 16)
      raise(MLEM_EINVAL, "tile elements per lane must be 4, 8 or 16");
    g.ept = param;
  } else if (kind == kDict) {
    if (param != 8 && param != 16 && param != 24 && param != 32 && param != 48 && param != 64)
      raise(MLEM_EINVAL, "dictionary tile elements per lane must be 8, 16, 24, 32, 48 or 64");
    g.ept = param;
  } else if (kind == kDictWide) {
    if (param != 16 && param != 24 && param != 32)
      raise(MLEM_EINVAL, "wide dictionary tile elements per lane must be 16, 24 or 32");
    g.ept = param;
  } else if (kind == kGroups) {
    if (param != 4 && param != 8 && param != 16 && param != 32 && param != 64)
      raise(MLEM_EINVAL, "lanes per row must be 4, 8, 16, 32 or 64");
    g.lanes = param;
  } else {
    raise(MLEM_EINVAL, "unknown kernel family");
  }
  return g;
}

// Candidate FAST-mode configurations of one projection.
std::vector<GroupLaunch> launch_candidates() {
  std::vector<GroupLaunch> c;
  for (int e : {16, 24, 32}) c.push_back(make_launch(kDictWide, e));
  for (int e : {8, 16, 24, 32, 48, 64}) c.push_back(make_launch(kDict, e));
  for (int e : {4, 8, 16}) c.push_back(make_launch(kTiles, e));
  for (int g : {8, 16, 32, 64}) c.push_back(make_launch(kGroups, g));
  return c;
}

// Time every forward and backward configuration on this matrix with HIP
// events (median of 3 launches on real-sized vectors) and keep the fastest.
void autotune(DevMatrix& m) {
  if (m.nnz == 0) return;
  float *x = nullptr, *r = nullptr, *xx = nullptr;
  int32_t* y = nullptr;
  struct F {
    void** p[4];
    ~F() {
      for (void** q : p)
        if (*q) (void)hipFree(*q);
    }
  } f{{(void**)&x, (void**)&r, (void**)&y, (void**)&xx}};
  MLEM_HIP(hipMalloc(&x, 4ull * std::max<uint32_t>(m.cols, 1)));
  MLEM_HIP(hipMalloc(&xx, 4ull * std::max<uint32_t>(m.cols, 1)));
  MLEM_HIP(hipMalloc(&r, 4ull * std::max<uint32_t>(m.rows, 1)));
  MLEM_HIP(hipMalloc(&y, 4ull * std::max<uint32_t>(m.rows, 1)));
  launch_fill(x, 1.f, m.cols, m.stream);
  launch_fill(r, 1.f, m.rows, m.stream);
  launch_fill(reinterpret_cast<float*>(y), 0.f, m.rows, m.stream);
  float best = 1e30f;
  GroupLaunch bestF = m.fwd, bestB = m.bwd;
  for (const GroupLaunch& g : launch_candidates()) {
    m.fwd = g;
    prepare_tiles(m, m.stream);
    const float t = time_variant(
        m.stream, [&] { launch_fwd(m, MLEM_MODE_FAST, x, y, r, true, nullptr, m.stream); }, 3);
    if (t < best) {
      best = t;
      bestF = g;
    }
  }
  m.fwd = bestF;
  if (m.csc) {
    best = 1e30f;
    for (const GroupLaunch& g : launch_candidates()) {
      m.bwd = g;
      prepare_tiles(m, m.stream);
      const float t = time_variant(
          m.stream, [&] { launch_bwd(m, MLEM_MODE_FAST, r, m.norm, xx, false, m.stream); }, 3);
      if (t < best) {
        best = t;
        bestB = g;
      }
    }
    m.bwd = bestB;
  }
  prepare_tiles(m, m.stream);
  MLEM_HIP(hipStreamSynchronize(m.stream));
}

void ensure_scratch(DevMatrix& m);

// Shared tail of mlem_open / mlem_open_host once rowptr and csr are on the
// device: validate columns, build the CSC, column sums, launch tuning.
void finish_open(mlem_matrix* h, unsigned flags) {
  DevMatrix& m = h->m;
  auto t0 = std::chrono::steady_clock::now();
  if ((flags & MLEM_OPEN_FP16) && (flags & MLEM_OPEN_BF16))
    raise(MLEM_EINVAL, "MLEM_OPEN_FP16 and MLEM_OPEN_BF16 are exclusive");
  if ((flags & (MLEM_OPEN_FP16 | MLEM_OPEN_BF16)) && (flags & MLEM_OPEN_NO_CSC))
    raise(MLEM_EINVAL, "16-bit values need the CSC (MLEM_OPEN_NO_CSC not allowed)");
  if (m.nnz > 0) {
    unsigned* bad = nullptr;
    MLEM_HIP(hipMalloc(&bad, sizeof(unsigned)));
    unsigned hbad = 0;
    hipError_t e = hipMemsetAsync(bad, 0, sizeof(unsigned), m.stream);
    if (e == hipSuccess) {
      launch_check_columns(m, bad, m.stream);
      e = hipMemcpyAsync(&hbad, bad, sizeof(unsigned), hipMemcpyDeviceToHost, m.stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(m.stream);
    (void)hipFree(bad);
    MLEM_HIP(e);
    if (hbad)
      raise(MLEM_EFORMAT, std::to_string(hbad) + " matrix elements have a column index >= " +
                              std::to_string(m.cols));
  }
  const int vk = (flags & MLEM_OPEN_FP16) ? kF16 : (flags & MLEM_OPEN_BF16) ? kBF16 : kF32;
  struct Scratch {
    float* p = nullptr;
    ~Scratch() {
      if (p) (void)hipFree(p);
    }
  } rscale_file;
  if (vk != kF32) {
    // Row scales and 16-bit values are fixed before the layout is built, so
    // the CSC and both dictionary formats inherit the quantized values.
    MLEM_HIP(hipMalloc(&m.rscale, 4ull * std::max<uint32_t>(m.rows, 1)));
    MLEM_HIP(hipMalloc(&rscale_file.p, 4ull * std::max<uint32_t>(m.rows, 1)));
    launch_quantize(m, vk, rscale_file.p, m.stream);
  }
  const bool reorder = !(flags & MLEM_OPEN_KEEP_ORDER);
  build_layout(m, reorder, reorder, !(flags & MLEM_OPEN_NO_CSC), m.stream);
  MLEM_HIP(hipMalloc(&m.norm, 4ull * std::max<uint32_t>(m.cols, 1)));
  if (vk != kF32) {
    if (m.rows) rows_to_device(m, (const float*)rscale_file.p, m.rscale, m.stream);
    MLEM_HIP(hipStreamSynchronize(m.stream));
    m.vk = vk;
    launch_colsums_scaled(m, m.norm, m.stream);
  } else if (m.csc) {
    launch_colsums_csc(m, m.norm, m.stream);
  } else {
    launch_colsums_atomic(m, m.norm, m.stream);
  }
  ensure_scratch(m);
  to_file_cols(m, m.norm, m.tx, m.stream);
  std::vector<float> norm(m.cols);
  if (m.cols)
    MLEM_HIP(hipMemcpyAsync(norm.data(), m.tx, 4ull * m.cols, hipMemcpyDeviceToHost, m.stream));
  MLEM_HIP(hipStreamSynchronize(m.stream));
  // initImage's sum of the norm: double, file index order (seqcsr4mlem.cpp:55-56).
  double sn = 0.0;
  for (uint32_t j = 0; j < m.cols; ++j) sn += norm[j];
  m.sumnorm = sn;
  h->norm_seconds = seconds_since(t0);

  const double avgRow = m.rows ? (double)m.nnz / m.rows : 1.0;
  // 16384-element dictionary tiles run by 512-lane blocks, two per CU (LDS-
  // and register-bound); mlem_autotune measures the alternatives on the real
  // matrix.
  m.fwd = make_launch(kDictWide, 32);
  m.bwd = make_launch(kDictWide, 32);
  m.scatter = make_launch(kGroups, lanes_for(avgRow));
  prepare_tiles(m, m.stream);
  if (env_int("MLEM_AUTOTUNE", 0) > 0) autotune(m);
}

mlem_matrix* new_matrix(uint32_t rows, uint32_t cols, uint64_t nnz) {
  // u32 element offsets on the device, with headroom for the kernels'
  // unrolled index arithmetic (k + 4*64).
  if (nnz > (1ull << 32) - 1024)
    raise(MLEM_EINVAL, "matrices with 2^32-1024 or more elements are not supported (nnz=" +
                           std::to_string(nnz) + ")");
  if (rows == 0xFFFFFFFFu || cols == 0xFFFFFFFFu) raise(MLEM_EINVAL, "dimension too large");
  std::unique_ptr<mlem_matrix> h(new mlem_matrix());
  DevMatrix& m = h->m;
  m.rows = rows;
  m.cols = cols;
  m.nnz = nnz;
  select_device(&m.device, &m.cus);
  MLEM_HIP(hipStreamCreateWithFlags(&m.stream, hipStreamNonBlocking));
  MLEM_HIP(hipMalloc(&m.rowptr, 4ull * ((uint64_t)rows + 1)));
  MLEM_HIP(hipMalloc(&m.csr, 8ull * std::max<uint64_t>(nnz, 1)));
  return h.release();
}

void upload_rowidx(DevMatrix& m, const uint64_t* rowidx) {
  if (m.rows == 0) {
    MLEM_HIP(hipMemsetAsync(m.rowptr, 0, 4, m.stream));
    return;
  }
  uint64_t* d = nullptr;
  MLEM_HIP(hipMalloc(&d, 8ull * m.rows));
  hipError_t e = hipMemcpyAsync(d, rowidx, 8ull * m.rows, hipMemcpyHostToDevice, m.stream);
  if (e == hipSuccess) {
    launch_rowptr_from_rowidx(d, m.rows, m.rowptr, m.stream);
    e = hipStreamSynchronize(m.stream);
  }
  (void)hipFree(d);
  MLEM_HIP(e);
}

void ensure_scratch(DevMatrix& m) {
  if (!m.sx) MLEM_HIP(hipMalloc(&m.sx, 4ull * std::max<uint32_t>(m.cols, 1)));
  if (!m.su) MLEM_HIP(hipMalloc(&m.su, 4ull * std::max<uint32_t>(m.cols, 1)));
  if (!m.sf) MLEM_HIP(hipMalloc(&m.sf, 4ull * std::max<uint32_t>(m.rows, 1)));
  if (!m.sr) MLEM_HIP(hipMalloc(&m.sr, 4ull * std::max<uint32_t>(m.rows, 1)));
  if (!m.sy) MLEM_HIP(hipMalloc(&m.sy, 4ull * std::max<uint32_t>(m.rows, 1)));
  if (!m.tx) MLEM_HIP(hipMalloc(&m.tx, 4ull * std::max<uint32_t>(m.cols, 1)));
  if (!m.tf) MLEM_HIP(hipMalloc(&m.tf, 4ull * std::max<uint32_t>(m.rows, 1)));
}

void h2d(void* d, const void* h, size_t n, hipStream_t s) {
  if (n) MLEM_HIP(hipMemcpyAsync(d, h, n, hipMemcpyHostToDevice, s));
}
void d2h(void* h, const void* d, size_t n, hipStream_t s) {
  if (n) MLEM_HIP(hipMemcpyAsync(h, d, n, hipMemcpyDeviceToHost, s));
}

constexpr unsigned kStatsRing = 4096;

// Forward launches per projection: one, or one per block of a streamed matrix.
unsigned fwd_launches(const DevMatrix& m) {
  return m.streamed ? (unsigned)std::max<size_t>(m.streamed->blocks.size(), 1) : 1u;
}

void enqueue_iteration(mlem_recon* R) {
  DevMatrix& m = R->mat->m;
  const StatsState* st = R->want_stats ? &R->st : nullptr;
  if (R->mode == MLEM_MODE_ATOMIC) {
    if (m.streamed) {
      stream_fwd_scatter(m, R->x, R->y, R->r, R->u, st, m.stream);
    } else {
      launch_fwd(m, MLEM_MODE_FAST, R->x, R->y, R->r, true, st, m.stream);
      launch_scatter(m, R->r, R->u, m.stream);
    }
    launch_update(m.cols, R->u, m.norm, R->x, true, m.stream);
  } else if (m.streamed) {
    stream_iteration(m, R->mode, R->x, R->y, R->r, st, m.stream);
  } else {
    launch_fwd(m, R->mode, R->x, R->y, R->r, true, st, m.stream);
    launch_bwd(m, R->mode, R->r, m.norm, R->x, true, m.stream);
  }
}

// Iteration k's statistics: the per-launch results of its forward launches
// (one, or one per streamed block), summed in launch order.
void collect_stats(mlem_recon* R, int first, int count) {
  if (!R->want_stats || count <= 0) return;
  const unsigned per = fwd_launches(R->mat->m);
  std::vector<double2> res(kStatsRing);
  d2h(res.data(), R->st.results, sizeof(double2) * kStatsRing, R->mat->m.stream);
  MLEM_HIP(hipStreamSynchronize(R->mat->m.stream));
  for (int k = 0; k < count; ++k) {
    double ll = 0.0, tot = 0.0;
    for (unsigned b = 0; b < per; ++b) {
      const double2 v = res[((uint64_t)(first + k) * per + b) % kStatsRing];
      ll += v.x;
      tot += v.y;
    }
    R->loglik.push_back(ll);
    R->fwdtot.push_back(tot);
  }
}

void run_iterations(mlem_recon* R, int iterations) {
  DevMatrix& m = R->mat->m;
  if (R->want_stats && stats_max_blocks(m) > R->st.max_blocks) {
    // a launch change since creation needs more workgroup partials; the
    // captured graph holds the old buffer
    if (R->exec) {
      (void)hipGraphExecDestroy(R->exec);
      (void)hipGraphDestroy(R->graph);
      R->exec = nullptr;
      R->graph = nullptr;
    }
    MLEM_HIP(hipStreamSynchronize(m.stream));
    dfree(R->st.partials);
    R->st.max_blocks = stats_max_blocks(m);
    MLEM_HIP(hipMalloc(&R->st.partials, sizeof(double2) * R->st.max_blocks));
  }
  int left = iterations;
  while (left > 0) {
    const int n =
        R->want_stats ? std::min<int>(left, (int)(kStatsRing / fwd_launches(m))) : left;
    if (R->use_graph) {
      if (R->exec && R->graph_gen != m.launch_gen) {  // kernels or buffers changed
        (void)hipGraphExecDestroy(R->exec);
        (void)hipGraphDestroy(R->graph);
        R->exec = nullptr;
        R->graph = nullptr;
      }
      if (!R->exec) {
        R->graph_gen = m.launch_gen;
        MLEM_HIP(hipStreamBeginCapture(m.stream, hipStreamCaptureModeThreadLocal));
        try {
          enqueue_iteration(R);
        } catch (...) {
          hipGraph_t g = nullptr;
          (void)hipStreamEndCapture(m.stream, &g);
          if (g) (void)hipGraphDestroy(g);
          throw;
        }
        MLEM_HIP(hipStreamEndCapture(m.stream, &R->graph));
        MLEM_HIP(hipGraphInstantiate(&R->exec, R->graph, nullptr, nullptr, 0));
      }
      for (int i = 0; i < n; ++i) MLEM_HIP(hipGraphLaunch(R->exec, m.stream));
    } else {
      for (int i = 0; i < n; ++i) enqueue_iteration(R);
    }
    MLEM_HIP(hipStreamSynchronize(m.stream));
    collect_stats(R, R->done, n);
    R->done += n;
    left -= n;
  }
}

mlem_recon* create_recon(mlem_matrix* h, const int32_t* y, uint32_t rows, const float* x0,
                         uint32_t cols, const mlem_options* o, float* initial_out,
                         double* sumin_out) {
  check_dims(h, rows, cols);
  if (!y) raise(MLEM_EINVAL, "null sinogram");
  mlem_options dflt;
  mlem_options_default(&dflt);
  if (!o) o = &dflt;
  if (o->mode < MLEM_MODE_FAST || o->mode > MLEM_MODE_ATOMIC) raise(MLEM_EINVAL, "unknown mode");
  if (o->init == MLEM_INIT_GIVEN && !x0) raise(MLEM_EINVAL, "MLEM_INIT_GIVEN needs an image");
  if (o->mode != MLEM_MODE_ATOMIC && !h->m.csc && !h->m.streamed)
    raise(MLEM_ESTATE, "matrix was opened without a CSC; only MLEM_MODE_ATOMIC is available");
  DevMatrix& m = h->m;
  std::unique_ptr<mlem_recon> R(new mlem_recon());
  R->mat = h;
  R->mode = o->mode;
  // a streamed iteration spans two streams and host copies: run it directly
  R->use_graph = o->use_graph != 0 && !m.streamed;
  R->want_stats = (o->loglik != nullptr) || (o->fwd_total != nullptr);
  if (R->want_stats && fwd_launches(m) > kStatsRing)
    raise(MLEM_EINVAL, "statistics need at most " + std::to_string(kStatsRing) +
                           " streamed blocks");
  MLEM_HIP(hipMalloc(&R->y, 4ull * std::max<uint32_t>(rows, 1)));
  MLEM_HIP(hipMalloc(&R->x, 4ull * std::max<uint32_t>(cols, 1)));
  MLEM_HIP(hipMalloc(&R->r, 4ull * std::max<uint32_t>(rows, 1)));
  MLEM_HIP(hipMalloc(&R->xs, 4ull * std::max<uint32_t>(cols, 1)));
  if (R->mode == MLEM_MODE_ATOMIC) {
    MLEM_HIP(hipMalloc(&R->u, 4ull * std::max<uint32_t>(cols, 1)));
    launch_fill(R->u, 0.f, cols, m.stream);
  }
  if (R->want_stats) {
    R->st.ring = kStatsRing;
    R->st.max_blocks = stats_max_blocks(m);
    MLEM_HIP(hipMalloc(&R->st.partials, sizeof(double2) * R->st.max_blocks));
    MLEM_HIP(hipMalloc(&R->st.counter, 2 * sizeof(unsigned)));
    MLEM_HIP(hipMalloc(&R->st.results, sizeof(double2) * kStatsRing));
    MLEM_HIP(hipMemsetAsync(R->st.counter, 0, 2 * sizeof(unsigned), m.stream));
  }
  // sinogram into device row order (r is free until the first iteration)
  int32_t* ystage = reinterpret_cast<int32_t*>(R->r);
  h2d(ystage, y, 4ull * rows, m.stream);
  rows_to_device(m, (const int32_t*)ystage, R->y, m.stream);
  // initImage (seqcsr4mlem.cpp:52-66): sum of int32 counts in double is exact
  // in any order; the norm's double sum was taken in index order at open.
  double sumin = 0.0;
  for (uint32_t i = 0; i < rows; ++i) sumin += y[i];
  if (sumin_out) *sumin_out = sumin;
  if (o->init == MLEM_INIT_GIVEN) {
    h2d(R->xs, x0, 4ull * cols, m.stream);
    to_device_cols(m, R->xs, R->x, m.stream);
  } else {
    const float initial = (float)(sumin / m.sumnorm);
    if (initial_out) *initial_out = initial;
    launch_fill(R->x, initial, cols, m.stream);
  }
  MLEM_HIP(hipStreamSynchronize(m.stream));
  return R.release();
}

// Current image in file voxel order into host memory (enqueued, not synced).
void read_image(mlem_recon* R, float* host) {
  DevMatrix& m = R->mat->m;
  to_file_cols(m, R->x, R->xs, m.stream);
  d2h(host, R->xs, 4ull * m.cols, m.stream);
}

bool file_exists(const std::string& p) {
  struct stat st;
  return stat(p.c_str(), &st) == 0;
}

}  // namespace

// ----------------------------------------------------------------- C ABI --
extern "C" {

const char* mlem_last_error(void) { return g_error.c_str(); }

const char* mlem_version(void) { return "mlem-amd 0.1 gfx950"; }

int mlem_open(const char* path, unsigned flags, mlem_matrix** out) {
  return guarded([&] {
    if (!path || !out) raise(MLEM_EINVAL, "null argument");
    *out = nullptr;
    Csr4Header hd;
    std::vector<uint64_t> rowidx;
    int fd = -1;
    csr4_read_header(path, hd, rowidx, &fd);
    struct FdClose {
      int fd;
      ~FdClose() {
        if (fd >= 0) close(fd);
      }
    } fdc{fd};
    auto t0 = std::chrono::steady_clock::now();
    const std::string fn(path);
    auto read_at = [&](void* dst, uint64_t off, uint64_t n) {
      char* p = static_cast<char*>(dst);
      uint64_t done = 0;
      while (done < n) {
        ssize_t got = pread(fd, p + done, n - done, (off_t)(hd.dataOffset + off + done));
        if (got <= 0) raise(MLEM_EIO, "Cannot read matrix elements from " + fn);
        done += (uint64_t)got;
      }
    };
    std::unique_ptr<mlem_matrix> h;
    if ((flags & MLEM_OPEN_EXPAND_SYMMETRY) && (hd.flags & 7u)) {
      // regenerate the omitted elements on the host, then upload the full matrix
      std::vector<Csr4Element> el(hd.nnz);
      if (hd.nnz) read_at(el.data(), 0, 8ull * hd.nnz);
      std::vector<uint64_t> xrow;
      std::vector<Csr4Element> xel;
      expand_symmetry(hd.scanner, hd.flags & 7u, hd.nRows, hd.nColumns, rowidx, el.data(), xrow,
                      xel);
      el.clear();
      el.shrink_to_fit();
      h.reset(new_matrix(hd.nRows, hd.nColumns, xel.size()));
      upload_rowidx(h->m, xrow.data());
      const char* src = reinterpret_cast<const char*>(xel.data());
      stream_to_device(h->m.csr, 8ull * xel.size(), h->m.stream,
                       [&](void* dst, uint64_t off, uint64_t n) { std::memcpy(dst, src + off, n); });
      h->m.symflags = 0;  // the matrix in HBM is complete
    } else {
      h.reset(new_matrix(hd.nRows, hd.nColumns, hd.nnz));
      upload_rowidx(h->m, rowidx.data());
      stream_to_device(h->m.csr, 8ull * hd.nnz, h->m.stream, read_at);
      h->m.symflags = hd.flags & 7u;
    }
    h->m.grid[0] = hd.scanner.grid[0];
    h->m.grid[1] = hd.scanner.grid[1];
    h->m.grid[2] = hd.scanner.grid[2];
    h->load_seconds = seconds_since(t0);
    finish_open(h.get(), flags);
    *out = h.release();
  });
}

int mlem_probe(const char* path, uint32_t* rows, uint32_t* cols, uint64_t* nnz, uint32_t grid[3],
               uint32_t* symflags) {
  return guarded([&] {
    if (!path) raise(MLEM_EINVAL, "null argument");
    Csr4Header hd;
    std::vector<uint64_t> rowidx;
    csr4_read_header(path, hd, rowidx, nullptr);
    if (rows) *rows = hd.nRows;
    if (cols) *cols = hd.nColumns;
    if (nnz) *nnz = hd.nnz;
    if (grid)
      for (int k = 0; k < 3; ++k) grid[k] = hd.scanner.grid[k];
    if (symflags) *symflags = hd.flags & 7u;
  });
}

int mlem_expand_symmetry_file(const char* in_path, const char* out_path) {
  return guarded([&] {
    if (!in_path || !out_path) raise(MLEM_EINVAL, "null argument");
    Csr4Header hd;
    std::vector<uint64_t> rowidx;
    int fd = -1;
    csr4_read_header(in_path, hd, rowidx, &fd);
    struct FdClose {
      int fd;
      ~FdClose() {
        if (fd >= 0) close(fd);
      }
    } fdc{fd};
    std::vector<Csr4Element> el(hd.nnz);
    const uint64_t bytes = 8ull * hd.nnz;
    char* p = reinterpret_cast<char*>(el.data());
    for (uint64_t done = 0; done < bytes;) {
      ssize_t got = pread(fd, p + done, bytes - done, (off_t)(hd.dataOffset + done));
      if (got <= 0) raise(MLEM_EIO, std::string("Cannot read matrix elements from ") + in_path);
      done += (uint64_t)got;
    }
    if (!(hd.flags & 7u)) {
      csr4_write(out_path, hd.flags, hd.nRows, hd.nColumns, hd.nnz, hd.scanner, rowidx.data(),
                 el.data());
      return;
    }
    std::vector<uint64_t> xrow;
    std::vector<Csr4Element> xel;
    expand_symmetry(hd.scanner, hd.flags & 7u, hd.nRows, hd.nColumns, rowidx, el.data(), xrow,
                    xel);
    csr4_write(out_path, (uint8_t)(hd.flags & ~7u), hd.nRows, hd.nColumns, xel.size(),
               hd.scanner, xrow.data(), xel.data());
  });
}

int mlem_open_host(uint32_t rows, uint32_t cols, uint64_t nnz, const uint64_t* rowidx,
                   const void* elements, const uint32_t grid[3], unsigned flags,
                   mlem_matrix** out) {
  return guarded([&] {
    if (!out || (rows && !rowidx) || (nnz && !elements)) raise(MLEM_EINVAL, "null argument");
    *out = nullptr;
    csr4_check_rowidx(rowidx, rows, nnz);
    auto t0 = std::chrono::steady_clock::now();
    std::unique_ptr<mlem_matrix> h(new_matrix(rows, cols, nnz));
    if (grid)
      for (int k = 0; k < 3; ++k) h->m.grid[k] = grid[k];
    upload_rowidx(h->m, rowidx);
    const char* src = static_cast<const char*>(elements);
    stream_to_device(h->m.csr, 8ull * nnz, h->m.stream,
                     [&](void* dst, uint64_t off, uint64_t n) { std::memcpy(dst, src + off, n); });
    h->load_seconds = seconds_since(t0);
    finish_open(h.get(), flags);
    *out = h.release();
  });
}

namespace {

// Handle of a streamed matrix: no device CSR/CSC, the row blocks in pinned
// host memory; column sums by a streamed EXACT backprojection of ones, which
// adds every column's values in ascending row order (calcColumnSums,
// seqcsr4mlem.cpp:37-49).
mlem_matrix* open_streamed(uint32_t rows, uint32_t cols, uint64_t nnz, const uint64_t* rowidx,
                           const ElementFetch& fetch, const uint32_t grid[3], unsigned flags,
                           uint64_t block_bytes, uint64_t resident_bytes) {
  if (flags & (MLEM_OPEN_FP16 | MLEM_OPEN_BF16))
    raise(MLEM_EINVAL, "streamed matrices hold fp32 values");
  if (rows == 0xFFFFFFFFu || cols == 0xFFFFFFFFu) raise(MLEM_EINVAL, "dimension too large");
  std::unique_ptr<mlem_matrix> h(new mlem_matrix());
  DevMatrix& m = h->m;
  m.rows = rows;
  m.cols = cols;
  m.nnz = nnz;
  if (grid)
    for (int k = 0; k < 3; ++k) m.grid[k] = grid[k];
  select_device(&m.device, &m.cus);
  MLEM_HIP(hipStreamCreateWithFlags(&m.stream, hipStreamNonBlocking));
  auto t0 = std::chrono::steady_clock::now();
  m.streamed =
      build_streamed(rows, cols, nnz, rowidx, fetch, block_bytes, resident_bytes, m.stream);
  h->load_seconds = seconds_since(t0);
  t0 = std::chrono::steady_clock::now();
  MLEM_HIP(hipMalloc(&m.norm, 4ull * std::max<uint32_t>(cols, 1)));
  ensure_scratch(m);
  launch_fill(m.norm, 0.f, cols, m.stream);
  launch_fill(m.sr, 1.f, rows, m.stream);
  stream_bwd_add(m, true, m.sr, m.norm, m.stream);
  std::vector<float> norm(cols);
  d2h(norm.data(), m.norm, 4ull * cols, m.stream);
  MLEM_HIP(hipStreamSynchronize(m.stream));
  double sn = 0.0;  // initImage's double sum, index order (seqcsr4mlem.cpp:55-56)
  for (uint32_t j = 0; j < cols; ++j) sn += norm[j];
  m.sumnorm = sn;
  h->norm_seconds = seconds_since(t0);
  m.fwd = make_launch(kGroups, m.streamed->lanes_fwd);
  m.bwd = make_launch(kGroups, m.streamed->lanes_bwd);
  m.scatter = m.bwd;
  return h.release();
}

}  // namespace

int mlem_open_streamed(const char* path, unsigned flags, uint64_t block_bytes,
                       uint64_t resident_bytes, mlem_matrix** out) {
  return guarded([&] {
    if (!path || !out) raise(MLEM_EINVAL, "null argument");
    *out = nullptr;
    Csr4Header hd;
    std::vector<uint64_t> rowidx;
    int fd = -1;
    csr4_read_header(path, hd, rowidx, &fd);
    struct FdClose {
      int fd;
      ~FdClose() {
        if (fd >= 0) close(fd);
      }
    } fdc{fd};
    const std::string fn(path);
    std::unique_ptr<mlem_matrix> h;
    if ((flags & MLEM_OPEN_EXPAND_SYMMETRY) && (hd.flags & 7u)) {
      std::vector<Csr4Element> el(hd.nnz);
      char* p = reinterpret_cast<char*>(el.data());
      for (uint64_t done = 0, n = 8ull * hd.nnz; done < n;) {
        ssize_t got = pread(fd, p + done, n - done, (off_t)(hd.dataOffset + done));
        if (got <= 0) raise(MLEM_EIO, "Cannot read matrix elements from " + fn);
        done += (uint64_t)got;
      }
      std::vector<uint64_t> xrow;
      std::vector<Csr4Element> xel;
      expand_symmetry(hd.scanner, hd.flags & 7u, hd.nRows, hd.nColumns, rowidx, el.data(), xrow,
                      xel);
      el.clear();
      el.shrink_to_fit();
      const char* src = reinterpret_cast<const char*>(xel.data());
      h.reset(open_streamed(
          hd.nRows, hd.nColumns, xel.size(), xrow.data(),
          [&](void* dst, uint64_t first, uint64_t count) {
            std::memcpy(dst, src + 8 * first, 8 * count);
          },
          hd.scanner.grid, flags, block_bytes, resident_bytes));
    } else {
      h.reset(open_streamed(
          hd.nRows, hd.nColumns, hd.nnz, rowidx.data(),
          [&](void* dst, uint64_t first, uint64_t count) {
            char* p = static_cast<char*>(dst);
            const uint64_t n = 8 * count;
            for (uint64_t done = 0; done < n;) {
              ssize_t got = pread(fd, p + done, n - done,
                                  (off_t)(hd.dataOffset + 8 * first + done));
              if (got <= 0) raise(MLEM_EIO, "Cannot read matrix elements from " + fn);
              done += (uint64_t)got;
            }
          },
          hd.scanner.grid, flags, block_bytes, resident_bytes));
      h->m.symflags = hd.flags & 7u;
    }
    *out = h.release();
  });
}

int mlem_open_host_streamed(uint32_t rows, uint32_t cols, uint64_t nnz, const uint64_t* rowidx,
                            const void* elements, const uint32_t grid[3], unsigned flags,
                            uint64_t block_bytes, uint64_t resident_bytes, mlem_matrix** out) {
  return guarded([&] {
    if (!out || (rows && !rowidx) || (nnz && !elements)) raise(MLEM_EINVAL, "null argument");
    *out = nullptr;
    csr4_check_rowidx(rowidx, rows, nnz);
    const char* src = static_cast<const char*>(elements);
    *out = open_streamed(
        rows, cols, nnz, rowidx,
        [&](void* dst, uint64_t first, uint64_t count) {
          std::memcpy(dst, src + 8 * first, 8 * count);
        },
        grid, flags, block_bytes, resident_bytes);
  });
}

int mlem_stream_passes(const mlem_matrix* h, int* passes) {
  return guarded([&] {
    if (!h || !passes) raise(MLEM_EINVAL, "null argument");
    const StreamedMatrix* S = h->m.streamed;
    *passes = !S ? 0 : S->csc_dev ? 1 : 2;
  });
}

int mlem_stream_blocks(const mlem_matrix* h, uint32_t* blocks, uint32_t* resident) {
  return guarded([&] {
    if (!h || !blocks) raise(MLEM_EINVAL, "null argument");
    const StreamedMatrix* S = h->m.streamed;
    *blocks = S ? (uint32_t)S->blocks.size() : 0u;
    if (resident) {
      *resident = 0;
      if (S)
        for (const StreamBlock& B : S->blocks) *resident += B.resident ? 1u : 0u;
    }
  });
}

void mlem_close(mlem_matrix* m) {
  if (!m) return;
  if (m->m.stream) (void)hipStreamSynchronize(m->m.stream);
  delete m;
}

int mlem_dims(const mlem_matrix* h, uint32_t* rows, uint32_t* cols, uint64_t* nnz) {
  return guarded([&] {
    if (!h) raise(MLEM_EINVAL, "null matrix handle");
    if (rows) *rows = h->m.rows;
    if (cols) *cols = h->m.cols;
    if (nnz) *nnz = h->m.nnz;
  });
}

int mlem_geometry(const mlem_matrix* h, uint32_t grid[3], uint32_t* symflags) {
  return guarded([&] {
    if (!h) raise(MLEM_EINVAL, "null matrix handle");
    if (grid)
      for (int k = 0; k < 3; ++k) grid[k] = h->m.grid[k];
    if (symflags) *symflags = h->m.symflags;
  });
}

void* mlem_stream(mlem_matrix* h) { return h ? (void*)h->m.stream : nullptr; }

int mlem_column_sums(mlem_matrix* h, float* norm, uint32_t cols) {
  return guarded([&] {
    check_dims(h, h ? h->m.rows : 0, cols);
    if (!norm && cols) raise(MLEM_EINVAL, "null output");
    ensure_scratch(h->m);
    to_file_cols(h->m, h->m.norm, h->m.tx, h->m.stream);
    d2h(norm, h->m.tx, 4ull * cols, h->m.stream);
    MLEM_HIP(hipStreamSynchronize(h->m.stream));
  });
}

int mlem_init_image(mlem_matrix* h, const float* norm, uint32_t cols, const int32_t* y,
                    uint32_t rows, float* x, float* initial) {
  return guarded([&] {
    check_dims(h, rows, cols);
    if ((cols && (!norm || !x)) || (rows && !y)) raise(MLEM_EINVAL, "null argument");
    // initImage is two sequential double sums and one division on the host,
    // exactly as seqcsr4mlem.cpp:52-66; the broadcast runs on the device.
    double sumnorm = 0.0, sumin = 0.0;
    for (uint32_t j = 0; j < cols; ++j) sumnorm += norm[j];
    for (uint32_t i = 0; i < rows; ++i) sumin += y[i];
    const float v = (float)(sumin / sumnorm);
    if (initial) *initial = v;
    DevMatrix& m = h->m;
    ensure_scratch(m);
    launch_fill(m.sx, v, cols, m.stream);
    d2h(x, m.sx, 4ull * cols, m.stream);
    MLEM_HIP(hipStreamSynchronize(m.stream));
  });
}

int mlem_fwproj(mlem_matrix* h, const float* x, uint32_t cols, float* f, uint32_t rows, int mode) {
  return guarded([&] {
    check_dims(h, rows, cols);
    if ((cols && !x) || (rows && !f)) raise(MLEM_EINVAL, "null argument");
    DevMatrix& m = h->m;
    ensure_scratch(m);
    h2d(m.tx, x, 4ull * cols, m.stream);
    to_device_cols(m, m.tx, m.sx, m.stream);
    launch_fwd(m, mode == MLEM_MODE_EXACT ? MLEM_MODE_EXACT : MLEM_MODE_FAST, m.sx, nullptr, m.sf,
               false, nullptr, m.stream);
    rows_to_file(m, (const float*)m.sf, m.tf, m.stream);
    d2h(f, m.tf, 4ull * rows, m.stream);
    MLEM_HIP(hipStreamSynchronize(m.stream));
  });
}

int mlem_correl(mlem_matrix* h, const float* f, const int32_t* y, float* r, uint32_t rows) {
  return guarded([&] {
    check_dims(h, rows, h ? h->m.cols : 0);
    if (rows && (!f || !y || !r)) raise(MLEM_EINVAL, "null argument");
    DevMatrix& m = h->m;
    ensure_scratch(m);
    h2d(m.sf, f, 4ull * rows, m.stream);
    h2d(m.sy, y, 4ull * rows, m.stream);
    launch_correl(rows, m.sf, m.sy, m.sr, m.stream);
    d2h(r, m.sr, 4ull * rows, m.stream);
    MLEM_HIP(hipStreamSynchronize(m.stream));
  });
}

int mlem_bkproj(mlem_matrix* h, const float* r, uint32_t rows, float* u, uint32_t cols, int mode) {
  return guarded([&] {
    check_dims(h, rows, cols);
    if ((rows && !r) || (cols && !u)) raise(MLEM_EINVAL, "null argument");
    DevMatrix& m = h->m;
    ensure_scratch(m);
    h2d(m.tf, r, 4ull * rows, m.stream);
    rows_to_device(m, (const float*)m.tf, m.sr, m.stream);
    launch_scale_rows(m, m.sr, m.stream);  // 16-bit values: A^T r = Q^T (s r)
    if (mode == MLEM_MODE_ATOMIC) {
      launch_fill(m.su, 0.f, cols, m.stream);
      launch_scatter(m, m.sr, m.su, m.stream);
    } else {
      launch_bwd(m, mode, m.sr, m.norm, m.su, false, m.stream);
    }
    to_file_cols(m, m.su, m.tx, m.stream);
    d2h(u, m.tx, 4ull * cols, m.stream);
    MLEM_HIP(hipStreamSynchronize(m.stream));
  });
}

int mlem_update(mlem_matrix* h, const float* u, const float* norm, float* x, uint32_t cols) {
  return guarded([&] {
    check_dims(h, h ? h->m.rows : 0, cols);
    if (cols && (!u || !norm || !x)) raise(MLEM_EINVAL, "null argument");
    DevMatrix& m = h->m;
    ensure_scratch(m);
    float* dn = m.tx;  // elementwise: file order throughout
    h2d(m.su, u, 4ull * cols, m.stream);
    h2d(dn, norm, 4ull * cols, m.stream);
    h2d(m.sx, x, 4ull * cols, m.stream);
    launch_update(cols, m.su, dn, m.sx, false, m.stream);
    d2h(x, m.sx, 4ull * cols, m.stream);
    MLEM_HIP(hipStreamSynchronize(m.stream));
  });
}

void mlem_options_default(mlem_options* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->mode = MLEM_MODE_FAST;
  o->init = MLEM_INIT_REFERENCE;
  o->use_graph = 1;
}

int mlem_recon_create(mlem_matrix* h, const int32_t* y, uint32_t rows, const float* x0,
                      uint32_t cols, const mlem_options* opts, mlem_recon** out) {
  return guarded([&] {
    if (!out) raise(MLEM_EINVAL, "null argument");
    *out = create_recon(h, y, rows, x0, cols, opts, nullptr, nullptr);
  });
}

void mlem_recon_destroy(mlem_recon* r) {
  if (!r) return;
  if (r->mat && r->mat->m.stream) (void)hipStreamSynchronize(r->mat->m.stream);
  delete r;
}

int mlem_recon_run(mlem_recon* R, int iterations) {
  return guarded([&] {
    if (!R || iterations < 0) raise(MLEM_EINVAL, "bad argument");
    run_iterations(R, iterations);
  });
}

int mlem_recon_run_timed(mlem_recon* R, int iterations, float* kf, float* kb, float* span) {
  return guarded([&] {
    if (!R || iterations < 0) raise(MLEM_EINVAL, "bad argument");
    if (R->want_stats) raise(MLEM_ESTATE, "timed runs do not record statistics");
    DevMatrix& m = R->mat->m;
    const size_t need = 2 * (size_t)iterations + 1;
    while (R->events.size() < need) {
      hipEvent_t e;
      MLEM_HIP(hipEventCreate(&e));
      R->events.push_back(e);
    }
    hipEvent_t* ev = R->events.data();
    MLEM_HIP(hipEventRecord(ev[0], m.stream));
    for (int i = 0; i < iterations; ++i) {
      if (R->mode == MLEM_MODE_ATOMIC) {
        if (m.streamed) {  // one fused pass: its time counts as the forward's
          stream_fwd_scatter(m, R->x, R->y, R->r, R->u, nullptr, m.stream);
          MLEM_HIP(hipEventRecord(ev[2 * i + 1], m.stream));
        } else {
          launch_fwd(m, MLEM_MODE_FAST, R->x, R->y, R->r, true, nullptr, m.stream);
          MLEM_HIP(hipEventRecord(ev[2 * i + 1], m.stream));
          launch_scatter(m, R->r, R->u, m.stream);
        }
        launch_update(m.cols, R->u, m.norm, R->x, true, m.stream);
      } else {
        launch_fwd(m, R->mode, R->x, R->y, R->r, true, nullptr, m.stream);
        MLEM_HIP(hipEventRecord(ev[2 * i + 1], m.stream));
        launch_bwd(m, R->mode, R->r, m.norm, R->x, true, m.stream);
      }
      MLEM_HIP(hipEventRecord(ev[2 * i + 2], m.stream));
    }
    MLEM_HIP(hipEventSynchronize(ev[2 * iterations]));
    for (int i = 0; i < iterations; ++i) {
      float a = 0.f, b = 0.f;
      MLEM_HIP(hipEventElapsedTime(&a, ev[2 * i], ev[2 * i + 1]));
      MLEM_HIP(hipEventElapsedTime(&b, ev[2 * i + 1], ev[2 * i + 2]));
      if (kf) kf[i] = a;
      if (kb) kb[i] = b;
    }
    if (span) MLEM_HIP(hipEventElapsedTime(span, ev[0], ev[2 * iterations]));
    R->done += iterations;
  });
}

int mlem_recon_image(mlem_recon* R, float* x, uint32_t cols) {
  return guarded([&] {
    if (!R) raise(MLEM_EINVAL, "null handle");
    check_dims(R->mat, R->mat->m.rows, cols);
    if (cols && !x) raise(MLEM_EINVAL, "null output");
    read_image(R, x);
    MLEM_HIP(hipStreamSynchronize(R->mat->m.stream));
  });
}

int mlem_recon_iterations(const mlem_recon* R) { return R ? R->done : -1; }

int mlem_reconstruct(mlem_matrix* h, const int32_t* y, uint32_t rows, float* x, uint32_t cols,
                     int iterations, const mlem_options* opts) {
  return guarded([&] {
    check_dims(h, rows, cols);
    if (iterations < 0) raise(MLEM_EINVAL, "negative iteration count");
    if (cols && !x) raise(MLEM_EINVAL, "null image");
    mlem_options o;
    if (opts)
      o = *opts;
    else
      mlem_options_default(&o);
    const bool verbose = o.verbose != 0;
    const std::string dir = o.checkpoint_dir && *o.checkpoint_dir ? o.checkpoint_dir : ".";
    const std::string imgChk = dir + "/img.chkpt", iterChk = dir + "/iter.chkpt";
    if (verbose)
      std::cout << "Calculated norm, elapsed time: " << h->norm_seconds << "s\n";

    if (o.checkpoint_interval < MLEM_CHECKPOINT_RESTORE_ONLY)
      raise(MLEM_EINVAL, "invalid checkpoint interval " + std::to_string(o.checkpoint_interval));
    // Restart from a complete checkpoint whenever both files exist, whatever
    // the interval (mpicsr4mlem.cpp:536-553); an interval of 0 is the
    // sequential driver, which has no checkpoints.
    int start = 0;
    std::vector<float> restored;
    if (o.checkpoint_interval != 0) {
      const bool hi = file_exists(imgChk), ht = file_exists(iterChk);
      if (hi && ht) {
        restored = read_raw_f32(imgChk);
        if (restored.size() != cols) raise(MLEM_EIO, "Mismatch element number " + imgChk);
        std::vector<int32_t> it = read_raw_i32(iterChk);
        if (it.size() < 1) raise(MLEM_EIO, "Cannot open file " + iterChk);
        start = it[0];
        if (start < 0) raise(MLEM_EIO, "invalid iteration in " + iterChk);
        if (verbose) std::cout << "I got Checkpoint at iteration " << start << "\n";
      } else if (hi || ht) {
        if (verbose) std::cout << "omitted incomplete checkpoint\n";
      }
    }
    mlem_options co = o;
    const float* x0 = x;
    if (!restored.empty()) {
      co.init = MLEM_INIT_GIVEN;
      x0 = restored.data();
    }
    float initial = 0.f;
    double sumin = 0.0;
    std::unique_ptr<mlem_recon> R(create_recon(h, y, rows, x0, cols, &co, &initial, &sumin));
    if (verbose && co.init == MLEM_INIT_REFERENCE)
      std::cout << "INIT: sumnorm=" << h->m.sumnorm << ", sumin=" << sumin
                << ", initial value=" << initial << std::endl;
    if (verbose) std::cout << "Starting " << iterations - std::min(start, iterations) << " MLEM iterations" << std::endl;

    const bool per_iter = verbose || o.checkpoint_interval > 0;
    std::vector<float> img(per_iter ? cols : 0);
    auto t0 = std::chrono::steady_clock::now();
    if (!per_iter) {
      run_iterations(R.get(), std::max(0, iterations - start));
    } else {
      for (int iter = start; iter < iterations; ++iter) {
        run_iterations(R.get(), 1);
        const double el = seconds_since(t0);
        read_image(R.get(), img.data());
        MLEM_HIP(hipStreamSynchronize(h->m.stream));
        if (verbose) {
          // per-iteration statistics of seqcsr4mlem.cpp:172-185
          int cz = 0, cs = 0;
          float s = 0.0f;
          for (uint32_t j = 0; j < cols; ++j) {
            s += img[j];
            const int c = std::fpclassify(img[j]);
            if (c == FP_ZERO) ++cz;
            if (c == FP_SUBNORMAL) ++cs;
          }
          std::cout << "Finished iteration " << iter + 1 << ", ";
          std::cout << " image sum: " << s << ", ";
          std::cout << "elapsed time: " << el << "s\n";
          std::cout << "Zeros: " << cz << " (" << (float)cz / cols * 100.0 << " %), "
                    << " subnormals: " << cs << " (" << (float)cs / cols * 100.0 << " %)\n";
        }
        if (o.checkpoint_interval > 0 && (iter + 1) % o.checkpoint_interval == 0) {
          write_raw(imgChk, img.data(), 4ull * cols, true);
          const int32_t next = iter + 1;
          write_raw(iterChk, &next, sizeof(next), true);
        }
        t0 = std::chrono::steady_clock::now();
      }
    }
    read_image(R.get(), x);
    MLEM_HIP(hipStreamSynchronize(h->m.stream));
    const int ran = (int)R->loglik.size();
    for (int k = 0; k < ran; ++k) {
      if (o.loglik) o.loglik[start + k] = R->loglik[k];
      if (o.fwd_total) o.fwd_total[start + k] = R->fwdtot[k];
    }
    if (o.checkpoint_interval > 0) {  // clean exit (mpicsr4mlem.cpp:656-660)
      std::remove(imgChk.c_str());
      std::remove(iterChk.c_str());
    }
  });
}

int mlem_get_launch(const mlem_matrix* h, int* fk, int* fp, int* bk, int* bp) {
  return guarded([&] {
    if (!h) raise(MLEM_EINVAL, "null matrix handle");
    const GroupLaunch &f = h->m.fwd, &b = h->m.bwd;
    if (fk) *fk = f.kind;
    if (fp) *fp = f.kind == kGroups ? f.lanes : f.ept;
    if (bk) *bk = b.kind;
    if (bp) *bp = b.kind == kGroups ? b.lanes : b.ept;
  });
}

int mlem_set_launch(mlem_matrix* h, int fk, int fp, int bk, int bp) {
  return guarded([&] {
    if (!h) raise(MLEM_EINVAL, "null matrix handle");
    if (h->m.streamed) raise(MLEM_EINVAL, "streamed matrices have a fixed launch configuration");
    const GroupLaunch f = make_launch(fk, fp), b = make_launch(bk, bp);
    ++h->m.launch_gen;
    h->m.fwd = f;
    h->m.bwd = b;
    prepare_tiles(h->m, h->m.stream);
  });
}

int mlem_autotune(mlem_matrix* h) {
  return guarded([&] {
    if (!h) raise(MLEM_EINVAL, "null matrix handle");
    if (h->m.streamed) raise(MLEM_EINVAL, "streamed matrices have a fixed launch configuration");
    ++h->m.launch_gen;
    autotune(h->m);
  });
}

int mlem_calibration_copy(size_t bytes, int repeats, float* ms) {
  return guarded([&] {
    if (bytes == 0 || bytes % 16 != 0 || repeats <= 0) raise(MLEM_EINVAL, "bad size");
    int dev = 0, cus = 0;
    select_device(&dev, &cus);
    void *a = nullptr, *b = nullptr;
    hipStream_t s = nullptr;
    struct C {
      void** a;
      void** b;
      hipStream_t* s;
      ~C() {
        if (*a) (void)hipFree(*a);
        if (*b) (void)hipFree(*b);
        if (*s) (void)hipStreamDestroy(*s);
      }
    } c{&a, &b, &s};
    MLEM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    MLEM_HIP(hipMalloc(&a, bytes));
    MLEM_HIP(hipMalloc(&b, bytes));
    MLEM_HIP(hipMemsetAsync(a, 1, bytes, s));
    float t = time_variant(s, [&] { launch_copy((const uint4*)a, (uint4*)b, bytes / 16, s); }, repeats);
    if (ms) *ms = t;
  });
}

int mlem_device_sync(void) {
  return guarded([&] { MLEM_HIP(hipDeviceSynchronize()); });
}

}  // extern "C"
