/*
 * mlem_oracle.c — CPU oracle for the MLEM hot path. TEST INFRASTRUCTURE ONLY:
 * it is linked by the tests, the smoke check and the benchmark's CPU-baseline
 * leg, never by the product library.
 *
 * Three restatements of the reference's plain MLEM:
 *   orc_*        fp32, the exact arithmetic of mlem_seq/seqcsr4mlem.cpp:
 *                sequential sums in element order, multiply and add rounded
 *                separately (built with -ffp-contract=off, no FMA),
 *                flush-to-zero on (seqcsr4mlem.cpp:194), IEEE division.
 *   orc_omp_*    the OpenMP driver mlem_openmp/openmpcsr4mlem.cpp:69-128:
 *                dynamic row schedule forward, per-thread private update
 *                arrays merged under a critical section (merge order is not
 *                deterministic); no flush-to-zero (the driver sets none).
 *   orc_f64_*    the same algorithm entirely in double precision: the
 *                accuracy reference for every fp32 result.
 *
 * Matrix arguments use the csr4 body layout: rowidx[r] = exclusive END offset
 * of row r (helper_files_common/csr4matrix.hpp:138-149), elements are
 * {u32 column, f32 value} (helper_files_common/matrixelement.hpp:37-46).
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <xmmintrin.h>

typedef struct {
  uint32_t column;
  float value;
} orc_elem;

#define ROW_BEGIN(rowidx, r) ((r) == 0 ? 0 : (rowidx)[(r)-1])

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* Flush-to-zero for SSE results, as _MM_SET_FLUSH_ZERO_MODE(ON)
 * (seqcsr4mlem.cpp:194). No denormals-are-zero. */
static unsigned ftz_on(void) {
  unsigned old = _mm_getcsr();
  _mm_setcsr(old | _MM_FLUSH_ZERO_ON);
  return old;
}
static void ftz_restore(unsigned old) { _mm_setcsr(old); }

/* ---------------------------------------------------------------- fp32 -- */

/* calcColumnSums, seqcsr4mlem.cpp:37-49 */
static void column_sums(uint32_t rows, uint32_t cols, const uint64_t* rowidx,
                        const orc_elem* el, float* norm) {
  for (uint32_t j = 0; j < cols; ++j) norm[j] = 0.0f;
  for (uint32_t r = 0; r < rows; ++r)
    for (uint64_t k = ROW_BEGIN(rowidx, r); k < rowidx[r]; ++k)
      norm[el[k].column] += el[k].value;
}

/* initImage, seqcsr4mlem.cpp:52-66 */
static float init_image(uint32_t cols, const float* norm, uint32_t rows,
                        const int32_t* y, float* x, double* sumnorm_out,
                        double* sumin_out) {
  double sumnorm = 0.0, sumin = 0.0;
  for (uint32_t j = 0; j < cols; ++j) sumnorm += norm[j];
  for (uint32_t i = 0; i < rows; ++i) sumin += y[i];
  float initial = (float)(sumin / sumnorm);
  for (uint32_t j = 0; j < cols; ++j) x[j] = initial;
  if (sumnorm_out) *sumnorm_out = sumnorm;
  if (sumin_out) *sumin_out = sumin;
  return initial;
}

/* calcFwProj, seqcsr4mlem.cpp:69-91 */
static void fwproj(uint32_t rows, const uint64_t* rowidx, const orc_elem* el,
                   const float* x, float* f) {
  for (uint32_t r = 0; r < rows; ++r) {
    float res = 0.0f;
    for (uint64_t k = ROW_BEGIN(rowidx, r); k < rowidx[r]; ++k)
      res += el[k].value * x[el[k].column];
    f[r] = res;
  }
}

/* calcCorrel, seqcsr4mlem.cpp:94-105 */
static void correl(uint32_t rows, const float* f, const int32_t* y, float* r) {
  for (uint32_t i = 0; i < rows; ++i)
    r[i] = (f[i] != 0.0f) ? ((float)y[i] / f[i]) : 0.0f;
}

/* calcBkProj, seqcsr4mlem.cpp:108-120 */
static void bkproj(uint32_t rows, uint32_t cols, const uint64_t* rowidx,
                   const orc_elem* el, const float* r, float* u) {
  for (uint32_t j = 0; j < cols; ++j) u[j] = 0.0f;
  for (uint32_t i = 0; i < rows; ++i)
    for (uint64_t k = ROW_BEGIN(rowidx, i); k < rowidx[i]; ++k)
      u[el[k].column] += el[k].value * r[i];
}

/* calcUpdate, seqcsr4mlem.cpp:123-134 */
static void update(uint32_t cols, const float* u, const float* norm, float* x) {
  for (uint32_t j = 0; j < cols; ++j)
    x[j] *= (norm[j] != 0.0f) ? (u[j] / norm[j]) : u[j];
}

void orc_column_sums(uint32_t rows, uint32_t cols, const uint64_t* rowidx,
                     const orc_elem* el, float* norm) {
  unsigned m = ftz_on();
  column_sums(rows, cols, rowidx, el, norm);
  ftz_restore(m);
}

float orc_init_image(uint32_t cols, const float* norm, uint32_t rows,
                     const int32_t* y, float* x, double* sumnorm,
                     double* sumin) {
  unsigned m = ftz_on();
  float v = init_image(cols, norm, rows, y, x, sumnorm, sumin);
  ftz_restore(m);
  return v;
}

void orc_fwproj(uint32_t rows, const uint64_t* rowidx, const orc_elem* el,
                const float* x, float* f) {
  unsigned m = ftz_on();
  fwproj(rows, rowidx, el, x, f);
  ftz_restore(m);
}

void orc_correl(uint32_t rows, const float* f, const int32_t* y, float* r) {
  unsigned m = ftz_on();
  correl(rows, f, y, r);
  ftz_restore(m);
}

void orc_bkproj(uint32_t rows, uint32_t cols, const uint64_t* rowidx,
                const orc_elem* el, const float* r, float* u) {
  unsigned m = ftz_on();
  bkproj(rows, cols, rowidx, el, r, u);
  ftz_restore(m);
}

void orc_update(uint32_t cols, const float* u, const float* norm, float* x) {
  unsigned m = ftz_on();
  update(cols, u, norm, x);
  ftz_restore(m);
}

/* mlem(), seqcsr4mlem.cpp:137-189. init_given = 0: x0 from initImage;
 * 1: x_inout holds x0. Per-iteration outputs (any may be NULL): the float
 * image sum, FP_ZERO and FP_SUBNORMAL counts (seqcsr4mlem.cpp:172-179) and
 * the wall time of the four steps. Returns 0, or -4 on allocation failure. */
int orc_mlem(uint32_t rows, uint32_t cols, const uint64_t* rowidx,
             const orc_elem* el, const int32_t* y, float* x, int iters,
             int init_given, float* image_sum, int32_t* zeros,
             int32_t* subnormals, double* iter_seconds) {
  float* norm = (float*)malloc(sizeof(float) * (cols ? cols : 1));
  float* f = (float*)malloc(sizeof(float) * (rows ? rows : 1));
  float* r = (float*)malloc(sizeof(float) * (rows ? rows : 1));
  float* u = (float*)malloc(sizeof(float) * (cols ? cols : 1));
  if (!norm || !f || !r || !u) {
    free(norm); free(f); free(r); free(u);
    return -4;
  }
  unsigned m = ftz_on();
  column_sums(rows, cols, rowidx, el, norm);
  if (!init_given) init_image(cols, norm, rows, y, x, NULL, NULL);
  for (int it = 0; it < iters; ++it) {
    double t0 = now_s();
    fwproj(rows, rowidx, el, x, f);
    correl(rows, f, y, r);
    bkproj(rows, cols, rowidx, el, r, u);
    update(cols, u, norm, x);
    double t1 = now_s();
    if (iter_seconds) iter_seconds[it] = t1 - t0;
    float s = 0.0f;
    int32_t cz = 0, cs = 0;
    for (uint32_t j = 0; j < cols; ++j) {
      s += x[j];
      int c = fpclassify(x[j]);
      if (c == FP_ZERO) ++cz;
      if (c == FP_SUBNORMAL) ++cs;
    }
    if (image_sum) image_sum[it] = s;
    if (zeros) zeros[it] = cz;
    if (subnormals) subnormals[it] = cs;
  }
  ftz_restore(m);
  free(norm); free(f); free(r); free(u);
  return 0;
}

/* -------------------------------------------------------------- OpenMP -- */

/* openmpcsr4mlem.cpp:69-91 (forward) and :107-128 (backward, private
 * arrays + critical merge); norm, init, correl and update sequential as in
 * that driver. threads <= 0: OpenMP default. */
int orc_omp_mlem(uint32_t rows, uint32_t cols, const uint64_t* rowidx,
                 const orc_elem* el, const int32_t* y, float* x, int iters,
                 int init_given, int threads, double* iter_seconds) {
  float* norm = (float*)malloc(sizeof(float) * (cols ? cols : 1));
  float* f = (float*)malloc(sizeof(float) * (rows ? rows : 1));
  float* r = (float*)malloc(sizeof(float) * (rows ? rows : 1));
  float* u = (float*)malloc(sizeof(float) * (cols ? cols : 1));
  if (!norm || !f || !r || !u) {
    free(norm); free(f); free(r); free(u);
    return -4;
  }
  int T = threads > 0 ? threads : omp_get_max_threads();
  int failed = 0;
  column_sums(rows, cols, rowidx, el, norm);
  if (!init_given) init_image(cols, norm, rows, y, x, NULL, NULL);
  for (int it = 0; it < iters && !failed; ++it) {
    double t0 = now_s();
#pragma omp parallel for schedule(dynamic) num_threads(T)
    for (int64_t i = 0; i < (int64_t)rows; ++i) {
      float res = 0.0f;
      for (uint64_t k = ROW_BEGIN(rowidx, i); k < rowidx[i]; ++k)
        res += el[k].value * x[el[k].column];
      f[i] = res;
    }
    correl(rows, f, y, r);
    for (uint32_t j = 0; j < cols; ++j) u[j] = 0.0f;
#pragma omp parallel num_threads(T)
    {
      float* priv = (float*)calloc(cols ? cols : 1, sizeof(float));
      if (!priv) {
#pragma omp atomic write
        failed = 1;
      }
#pragma omp for schedule(dynamic)
      for (int64_t i = 0; i < (int64_t)rows; ++i) {
        if (!priv) continue;
        for (uint64_t k = ROW_BEGIN(rowidx, i); k < rowidx[i]; ++k)
          priv[el[k].column] += el[k].value * r[i];
      }
#pragma omp critical
      {
        if (priv)
          for (uint32_t j = 0; j < cols; ++j) u[j] += priv[j];
      }
      free(priv);
    }
    update(cols, u, norm, x);
    if (iter_seconds) iter_seconds[it] = now_s() - t0;
  }
  free(norm); free(f); free(r); free(u);
  return failed ? -4 : 0;
}

/* ---------------------------------------------------------------- FP64 -- */
/* a_ij = value, or (double) rscale[i] * value when a row scale is given
 * (reduced-precision storage: the value is the 16-bit quotient q_ij). The
 * product of an fp32 scale and a 16-bit quotient is exact in double. */
#define A64(rscale, i, e) ((rscale) ? (double)(rscale)[i] * (double)(e).value : (double)(e).value)

static void f64_column_sums(uint32_t rows, uint32_t cols, const uint64_t* rowidx,
                            const orc_elem* el, const float* rscale, double* norm) {
  for (uint32_t j = 0; j < cols; ++j) norm[j] = 0.0;
  for (uint32_t r = 0; r < rows; ++r)
    for (uint64_t k = ROW_BEGIN(rowidx, r); k < rowidx[r]; ++k)
      norm[el[k].column] += A64(rscale, r, el[k]);
}

static void f64_fwproj(uint32_t rows, const uint64_t* rowidx, const orc_elem* el,
                       const float* rscale, const double* x, double* f) {
  for (uint32_t r = 0; r < rows; ++r) {
    double res = 0.0;
    for (uint64_t k = ROW_BEGIN(rowidx, r); k < rowidx[r]; ++k)
      res += A64(rscale, r, el[k]) * x[el[k].column];
    f[r] = res;
  }
}

static void f64_bkproj(uint32_t rows, uint32_t cols, const uint64_t* rowidx,
                       const orc_elem* el, const float* rscale, const double* r, double* u) {
  for (uint32_t j = 0; j < cols; ++j) u[j] = 0.0;
  for (uint32_t i = 0; i < rows; ++i)
    for (uint64_t k = ROW_BEGIN(rowidx, i); k < rowidx[i]; ++k)
      u[el[k].column] += A64(rscale, i, el[k]) * r[i];
}

void orc_f64_column_sums(uint32_t rows, uint32_t cols, const uint64_t* rowidx,
                         const orc_elem* el, double* norm) {
  f64_column_sums(rows, cols, rowidx, el, NULL, norm);
}

void orc_f64_fwproj(uint32_t rows, const uint64_t* rowidx, const orc_elem* el,
                    const double* x, double* f) {
  f64_fwproj(rows, rowidx, el, NULL, x, f);
}

void orc_f64_bkproj(uint32_t rows, uint32_t cols, const uint64_t* rowidx,
                    const orc_elem* el, const double* r, double* u) {
  f64_bkproj(rows, cols, rowidx, el, NULL, r, u);
}

/* The MLEM iteration in double on a_ij = rscale[i] * value (rscale may be
 * NULL: a_ij = value). y holds the counts as doubles (int32 counts, or
 * relaxed real counts). init_given = 0: x0 = sum(y)/sum(norm).
 * Per-iteration outputs (may be NULL): Poisson log-likelihood of the iterate
 * entering iteration k, L = sum_{i: f_i != 0} (y_i log f_i - f_i), and
 * sum_i f_i. Returns 0 or -4. */
int orc_f64_mlem_scaled(uint32_t rows, uint32_t cols, const uint64_t* rowidx,
                        const orc_elem* el, const float* rscale, const double* y, double* x,
                        int iters, int init_given, double* loglik, double* fwd_total) {
  double* norm = (double*)malloc(sizeof(double) * (cols ? cols : 1));
  double* f = (double*)malloc(sizeof(double) * (rows ? rows : 1));
  double* u = (double*)malloc(sizeof(double) * (cols ? cols : 1));
  if (!norm || !f || !u) {
    free(norm); free(f); free(u);
    return -4;
  }
  f64_column_sums(rows, cols, rowidx, el, rscale, norm);
  if (!init_given) {
    double sn = 0.0, sy = 0.0;
    for (uint32_t j = 0; j < cols; ++j) sn += norm[j];
    for (uint32_t i = 0; i < rows; ++i) sy += y[i];
    for (uint32_t j = 0; j < cols; ++j) x[j] = sy / sn;
  }
  for (int it = 0; it < iters; ++it) {
    f64_fwproj(rows, rowidx, el, rscale, x, f);
    double ll = 0.0, tot = 0.0;
    for (uint32_t i = 0; i < rows; ++i) {
      tot += f[i];
      if (f[i] != 0.0) {
        ll += (y[i] != 0.0 ? y[i] * log(f[i]) : 0.0) - f[i];
        f[i] = y[i] / f[i];
      }
    }
    if (loglik) loglik[it] = ll;
    if (fwd_total) fwd_total[it] = tot;
    f64_bkproj(rows, cols, rowidx, el, rscale, f, u);
    for (uint32_t j = 0; j < cols; ++j)
      x[j] *= (norm[j] != 0.0) ? (u[j] / norm[j]) : u[j];
  }
  free(norm); free(f); free(u);
  return 0;
}

int orc_f64_mlem(uint32_t rows, uint32_t cols, const uint64_t* rowidx,
                 const orc_elem* el, const double* y, double* x, int iters,
                 int init_given, double* loglik, double* fwd_total) {
  return orc_f64_mlem_scaled(rows, cols, rowidx, el, NULL, y, x, iters, init_given, loglik,
                             fwd_total);
}

/* ------------------------------------------------- reduced precision -- */
/* Round an fp32 value to the nearest fp16 value (ties to even, subnormals
 * kept, no overflow handling: |v| <= 1 here), returned as fp32. Works on the
 * binary representation: keep 11 significant bits (fewer in the subnormal
 * range, where the fp16 quantum is 2^-24). */
static float round_fp16(float v) {
  if (v == 0.0f) return v;
  int e;
  const double m = frexp((double)v, &e);          /* v = m 2^e, 0.5 <= |m| < 1 */
  int bits = 11;                                  /* fp16: 10 stored + hidden  */
  if (e - 1 < -14) bits = 11 - (-14 - (e - 1));   /* subnormal: quantum 2^-24  */
  if (bits < 0) return 0.0f * v;
  const double scaled = ldexp(m, bits);           /* integer part = kept bits  */
  double r = nearbyint(scaled);                   /* default mode: ties even   */
  return (float)ldexp(r, e - bits);
}

/* Round to the nearest bf16 value (ties to even): keep 8 significant bits
 * of the fp32 significand. */
static float round_bf16(float v) {
  if (v == 0.0f) return v;
  int e;
  const double m = frexp((double)v, &e);
  return (float)ldexp(nearbyint(ldexp(m, 8)), e - 8);
}

/* The reduced-precision matrix: s_i = max_k |a_ik| (1 for an all-zero row),
 * q_ik = round(a_ik / s_i) with the quotient rounded to fp32 first (fp32
 * division, flush-to-zero), then to fp16 (kind 1) or bf16 (kind 2). Writes
 * out[k] = {column, q_ik} and rscale[i] = s_i. */
void orc_quantize(uint32_t rows, const uint64_t* rowidx, const orc_elem* el, int kind,
                  orc_elem* out, float* rscale) {
  const unsigned old = ftz_on();
  for (uint32_t r = 0; r < rows; ++r) {
    float mx = 0.0f;
    for (uint64_t k = ROW_BEGIN(rowidx, r); k < rowidx[r]; ++k)
      mx = fabsf(el[k].value) > mx ? fabsf(el[k].value) : mx;
    const float s = mx > 0.0f ? mx : 1.0f;
    rscale[r] = s;
    for (uint64_t k = ROW_BEGIN(rowidx, r); k < rowidx[r]; ++k) {
      volatile float qf = el[k].value / s;
      out[k].column = el[k].column;
      out[k].value = kind == 1 ? round_fp16(qf) : round_bf16(qf);
    }
  }
  ftz_restore(old);
}

/* FP64 MLEM as orc_f64_mlem_scaled, parallel over rows for large matrices:
 * static row partition, per-thread private backprojection arrays summed in
 * thread order (deterministic for a given thread count; the order of the
 * FP64 sums differs from the sequential version only at the 1e-16 level).
 * threads <= 0: OpenMP default. */
int orc_f64_mlem_par(uint32_t rows, uint32_t cols, const uint64_t* rowidx,
                     const orc_elem* el, const float* rscale, const double* y, double* x,
                     int iters, int init_given, int threads, double* loglik,
                     double* fwd_total) {
  const int T = threads > 0 ? threads : omp_get_max_threads();
  double* norm = (double*)malloc(sizeof(double) * (cols ? cols : 1));
  double* f = (double*)malloc(sizeof(double) * (rows ? rows : 1));
  double* priv = (double*)malloc(sizeof(double) * (size_t)T * (cols ? cols : 1));
  double* tl = (double*)malloc(sizeof(double) * 2 * (size_t)T);
  if (!norm || !f || !priv || !tl) {
    free(norm); free(f); free(priv); free(tl);
    return -4;
  }
  /* column sums through the same private arrays */
#pragma omp parallel num_threads(T)
  {
    const int t = omp_get_thread_num(), nt = omp_get_num_threads();
    double* p = priv + (size_t)t * cols;
    for (uint32_t j = 0; j < cols; ++j) p[j] = 0.0;
    const uint32_t r0 = (uint32_t)((uint64_t)rows * t / nt), r1 = (uint32_t)((uint64_t)rows * (t + 1) / nt);
    for (uint32_t r = r0; r < r1; ++r)
      for (uint64_t k = ROW_BEGIN(rowidx, r); k < rowidx[r]; ++k)
        p[el[k].column] += A64(rscale, r, el[k]);
#pragma omp barrier
#pragma omp for schedule(static)
    for (int64_t j = 0; j < (int64_t)cols; ++j) {
      double s = 0.0;
      for (int q = 0; q < nt; ++q) s += priv[(size_t)q * cols + j];
      norm[j] = s;
    }
  }
  if (!init_given) {
    double sn = 0.0, sy = 0.0;
    for (uint32_t j = 0; j < cols; ++j) sn += norm[j];
    for (uint32_t i = 0; i < rows; ++i) sy += y[i];
    for (uint32_t j = 0; j < cols; ++j) x[j] = sy / sn;
  }
  for (int it = 0; it < iters; ++it) {
#pragma omp parallel num_threads(T)
    {
      const int t = omp_get_thread_num(), nt = omp_get_num_threads();
      const uint32_t r0 = (uint32_t)((uint64_t)rows * t / nt), r1 = (uint32_t)((uint64_t)rows * (t + 1) / nt);
      double ll = 0.0, tot = 0.0;
      for (uint32_t r = r0; r < r1; ++r) {
        double res = 0.0;
        for (uint64_t k = ROW_BEGIN(rowidx, r); k < rowidx[r]; ++k)
          res += A64(rscale, r, el[k]) * x[el[k].column];
        tot += res;
        if (res != 0.0) {
          ll += (y[r] != 0.0 ? y[r] * log(res) : 0.0) - res;
          res = y[r] / res;
        }
        f[r] = res;
      }
      tl[2 * t] = ll;
      tl[2 * t + 1] = tot;
      double* p = priv + (size_t)t * cols;
      for (uint32_t j = 0; j < cols; ++j) p[j] = 0.0;
      for (uint32_t r = r0; r < r1; ++r)
        for (uint64_t k = ROW_BEGIN(rowidx, r); k < rowidx[r]; ++k)
          p[el[k].column] += A64(rscale, r, el[k]) * f[r];
#pragma omp barrier
#pragma omp for schedule(static)
      for (int64_t j = 0; j < (int64_t)cols; ++j) {
        double u = 0.0;
        for (int q = 0; q < nt; ++q) u += priv[(size_t)q * cols + j];
        x[j] *= (norm[j] != 0.0) ? (u / norm[j]) : u;
      }
#pragma omp single
      {
        double a = 0.0, b = 0.0;
        for (int q = 0; q < nt; ++q) {
          a += tl[2 * q];
          b += tl[2 * q + 1];
        }
        if (loglik) loglik[it] = a;
        if (fwd_total) fwd_total[it] = b;
      }
    }
  }
  free(norm); free(f); free(priv); free(tl);
  return 0;
}

/* ------------------------------------------------ unrolled MLEM adjoint -- */
/* K MLEM iterations in FP64 from a given x0 with real-valued counts y,
 *   f = A x_k,  r = (f != 0) ? y/f : 0,  u = A^T r,
 *   x_{k+1} = x_k * (n != 0 ? u/n : u),  n = A^T 1,
 * and, when gK (= dL/dx_K) is given, reverse mode through the stored
 * iterates:
 *   gu = g * (n != 0 ? x_k/n : x_k),   g_direct = g * (n != 0 ? u/n : u),
 *   gr = A gu,  dL/dy += gr/f,  gf = -gr * y / f^2  (0 where f == 0),
 *   g_k = g_direct + A^T gf,          dL/dx0 = g_0.
 * a_ij = rscale[i] * value (rscale may be NULL). xK, gy, gx0 may be NULL.
 * Returns 0 or -4. */
int orc_f64_unroll(uint32_t rows, uint32_t cols, const uint64_t* rowidx, const orc_elem* el,
                   const float* rscale, const double* y, const double* x0, int K, double* xK,
                   const double* gK, double* gy, double* gx0) {
  const size_t C = cols ? cols : 1, R = rows ? rows : 1, KK = K > 0 ? (size_t)K : 1;
  double* norm = (double*)malloc(sizeof(double) * C);
  double* xs = (double*)malloc(sizeof(double) * C * (KK + 1));
  double* fs = (double*)malloc(sizeof(double) * R * KK);
  double* us = (double*)malloc(sizeof(double) * C * KK);
  double* r = (double*)malloc(sizeof(double) * R);
  double* g = (double*)malloc(sizeof(double) * C);
  double* gu = (double*)malloc(sizeof(double) * C);
  double* h = (double*)malloc(sizeof(double) * C);
  if (!norm || !xs || !fs || !us || !r || !g || !gu || !h) {
    free(norm); free(xs); free(fs); free(us); free(r); free(g); free(gu); free(h);
    return -4;
  }
  f64_column_sums(rows, cols, rowidx, el, rscale, norm);
  memcpy(xs, x0, sizeof(double) * cols);
  for (int k = 0; k < K; ++k) {
    const double* x = xs + (size_t)k * C;
    double* f = fs + (size_t)k * R;
    double* u = us + (size_t)k * C;
    double* xn = xs + (size_t)(k + 1) * C;
    f64_fwproj(rows, rowidx, el, rscale, x, f);
    for (uint32_t i = 0; i < rows; ++i) r[i] = (f[i] != 0.0) ? y[i] / f[i] : 0.0;
    f64_bkproj(rows, cols, rowidx, el, rscale, r, u);
    for (uint32_t j = 0; j < cols; ++j)
      xn[j] = x[j] * ((norm[j] != 0.0) ? u[j] / norm[j] : u[j]);
  }
  if (xK) memcpy(xK, xs + (size_t)K * C, sizeof(double) * cols);
  if (gK) {
    memcpy(g, gK, sizeof(double) * cols);
    if (gy)
      for (uint32_t i = 0; i < rows; ++i) gy[i] = 0.0;
    for (int k = K - 1; k >= 0; --k) {
      const double* x = xs + (size_t)k * C;
      const double* f = fs + (size_t)k * R;
      const double* u = us + (size_t)k * C;
      for (uint32_t j = 0; j < cols; ++j) {
        const double inv = (norm[j] != 0.0) ? 1.0 / norm[j] : 1.0;
        gu[j] = g[j] * x[j] * inv;
        g[j] = g[j] * u[j] * inv;  /* direct part */
      }
      f64_fwproj(rows, rowidx, el, rscale, gu, r);  /* r <- gr */
      for (uint32_t i = 0; i < rows; ++i) {
        if (f[i] != 0.0) {
          if (gy) gy[i] += r[i] / f[i];
          r[i] = -r[i] * y[i] / (f[i] * f[i]);
        } else {
          r[i] = 0.0;
        }
      }
      f64_bkproj(rows, cols, rowidx, el, rscale, r, h);
      for (uint32_t j = 0; j < cols; ++j) g[j] += h[j];
    }
    if (gx0) memcpy(gx0, g, sizeof(double) * cols);
  }
  free(norm); free(xs); free(fs); free(us); free(r); free(g); free(gu); free(h);
  return 0;
}
