// synth.cpp — deterministic csr4 system matrices, phantoms and Poisson
// sinograms for the MLEM engine's benchmarks and tests (host code).
#include "synth.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "csr4.hpp"

namespace {

inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct SplitMix64 {
  uint64_t s;
  SplitMix64(uint64_t seed, uint64_t item) : s(mix64(seed ^ ((item + 1) * 0xD1B54A32D192ED03ull))) {}
  uint64_t next() {
    s += 0x9E3779B97F4A7C15ull;
    return mix64(s);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

struct Seg {
  uint32_t col;
  float len;
};

// Siddon traversal of the segment p0->p1 through the centred unit-voxel grid.
// Appends (voxel, length) pairs, columns ascending and unique. Returns false
// when the line misses the grid.
bool siddon(const double p0[3], const double p1[3], const uint32_t n[3],
            std::vector<double>& alphas, std::vector<Seg>& out) {
  out.clear();
  double d[3], lo[3];
  double amin = 0.0, amax = 1.0;
  for (int a = 0; a < 3; ++a) {
    d[a] = p1[a] - p0[a];
    lo[a] = -0.5 * n[a];
    const double hi = 0.5 * n[a];
    if (d[a] == 0.0) {
      if (p0[a] <= lo[a] || p0[a] >= hi) return false;
      continue;
    }
    double t0 = (lo[a] - p0[a]) / d[a], t1 = (hi - p0[a]) / d[a];
    if (t0 > t1) std::swap(t0, t1);
    amin = std::max(amin, t0);
    amax = std::min(amax, t1);
  }
  if (!(amax > amin)) return false;
  const double L = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  alphas.clear();
  alphas.push_back(amin);
  alphas.push_back(amax);
  for (int a = 0; a < 3; ++a) {
    if (d[a] == 0.0) continue;
    // plane coordinates lo + i, i = 0..n, crossed strictly inside (amin, amax)
    const double ca = p0[a] + amin * d[a] - lo[a];
    const double cb = p0[a] + amax * d[a] - lo[a];
    const double cmin = std::min(ca, cb), cmax = std::max(ca, cb);
    long i0 = (long)std::floor(cmin) + 1, i1 = (long)std::ceil(cmax) - 1;
    i0 = std::max(i0, 0L);
    i1 = std::min(i1, (long)n[a]);
    for (long i = i0; i <= i1; ++i) {
      double t = (lo[a] + (double)i - p0[a]) / d[a];
      if (t > amin && t < amax) alphas.push_back(t);
    }
  }
  std::sort(alphas.begin(), alphas.end());
  const double minLen = 1e-6;
  for (size_t k = 0; k + 1 < alphas.size(); ++k) {
    const double len = (alphas[k + 1] - alphas[k]) * L;
    if (len < minLen) continue;
    const double mid = 0.5 * (alphas[k] + alphas[k + 1]);
    long v[3];
    for (int a = 0; a < 3; ++a) {
      v[a] = (long)std::floor(p0[a] + mid * d[a] - lo[a]);
      v[a] = std::min(std::max(v[a], 0L), (long)n[a] - 1);
    }
    uint32_t col = (uint32_t)(v[0] + (long)n[0] * (v[1] + (long)n[1] * v[2]));
    out.push_back(Seg{col, (float)std::min(len, std::sqrt(3.0))});
  }
  if (out.empty()) return false;
  std::sort(out.begin(), out.end(), [](const Seg& a, const Seg& b) { return a.col < b.col; });
  size_t w = 0;  // merge duplicates (numerically split segments of one voxel)
  for (size_t k = 0; k < out.size(); ++k) {
    if (w > 0 && out[w - 1].col == out[k].col) {
      out[w - 1].len = std::min(out[w - 1].len + out[k].len, (float)std::sqrt(3.0));
    } else {
      out[w++] = out[k];
    }
  }
  out.resize(w);
  return true;
}

// One line of response: two uniform points on the cylinder surface, redrawn
// until the line meets the grid.
void trace_lor(uint64_t seed, uint64_t lor, const uint32_t n[3], std::vector<double>& alphas,
               std::vector<Seg>& segs) {
  SplitMix64 rng(seed, lor);
  const double R = 1.2 * 0.5 * (double)std::max(n[0], n[1]);
  const double H = (double)n[2];
  const double twoPi = 6.283185307179586;
  for (int attempt = 0; attempt < 100000; ++attempt) {
    double th0 = twoPi * rng.uniform(), z0 = (rng.uniform() - 0.5) * H;
    double th1 = twoPi * rng.uniform(), z1 = (rng.uniform() - 0.5) * H;
    double p0[3] = {R * std::cos(th0), R * std::sin(th0), z0};
    double p1[3] = {R * std::cos(th1), R * std::sin(th1), z1};
    if (siddon(p0, p1, n, alphas, segs)) return;
  }
  segs.clear();
}

// Poisson variate, inversion for small means, PTRS otherwise.
int32_t poisson(double lam, SplitMix64& rng) {
  if (!(lam > 0.0)) return 0;
  if (lam < 10.0) {
    const double u = rng.uniform();
    double p = std::exp(-lam), F = p;
    int32_t k = 0;
    while (u > F && k < 10000) {
      ++k;
      p *= lam / k;
      F += p;
    }
    return k;
  }
  const double slam = std::sqrt(lam), loglam = std::log(lam);
  const double b = 0.931 + 2.53 * slam;
  const double a = -0.059 + 0.02483 * b;
  const double invalpha = 1.1239 + 1.1328 / (b - 3.4);
  const double vr = 0.9277 - 3.6224 / (b - 2.0);
  for (;;) {
    const double U = rng.uniform() - 0.5;
    const double V = rng.uniform();
    const double us = 0.5 - std::fabs(U);
    const double kd = std::floor((2.0 * a / us + b) * U + lam + 0.43);
    if (us >= 0.07 && V <= vr) return (int32_t)kd;
    if (kd < 0.0 || (us < 0.013 && V > us)) continue;
    if (std::log(V) + std::log(invalpha) - std::log(a / (us * us) + b) <=
        -lam + kd * loglam - std::lgamma(kd + 1.0))
      return (int32_t)kd;
  }
}

int nthreads(int t) { return t > 0 ? t : omp_get_max_threads(); }

}  // namespace

extern "C" {

int mlem_synth_matrix(uint32_t nx, uint32_t ny, uint32_t nz, uint32_t nlors, uint64_t seed,
                      int threads, uint64_t** rowidxOut, void** elementsOut, uint64_t* nnzOut) {
  if (!rowidxOut || !elementsOut || !nnzOut || nx == 0 || ny == 0 || nz == 0 || nlors == 0)
    return -1;
  if ((uint64_t)nx * ny * nz > 0xFFFFFFFFull) return -1;
  const uint32_t n[3] = {nx, ny, nz};
  uint64_t* rowidx = (uint64_t*)std::malloc(8ull * nlors);
  if (!rowidx) return -4;
  const int T = nthreads(threads);
  // pass 1: row lengths
#pragma omp parallel num_threads(T)
  {
    std::vector<double> al;
    std::vector<Seg> segs;
#pragma omp for schedule(dynamic, 4096)
    for (int64_t r = 0; r < (int64_t)nlors; ++r) {
      trace_lor(seed, (uint64_t)r, n, al, segs);
      rowidx[r] = segs.size();
    }
  }
  uint64_t acc = 0;
  for (uint32_t r = 0; r < nlors; ++r) {
    acc += rowidx[r];
    rowidx[r] = acc;
  }
  mlem::Csr4Element* el = (mlem::Csr4Element*)std::malloc(std::max<uint64_t>(acc, 1) * 8ull);
  if (!el) {
    std::free(rowidx);
    return -4;
  }
  // pass 2: elements (same streams, same lines)
#pragma omp parallel num_threads(T)
  {
    std::vector<double> al;
    std::vector<Seg> segs;
#pragma omp for schedule(dynamic, 4096)
    for (int64_t r = 0; r < (int64_t)nlors; ++r) {
      trace_lor(seed, (uint64_t)r, n, al, segs);
      uint64_t beg = r ? rowidx[r - 1] : 0;
      for (size_t k = 0; k < segs.size(); ++k) el[beg + k] = mlem::Csr4Element{segs[k].col, segs[k].len};
    }
  }
  *rowidxOut = rowidx;
  *elementsOut = el;
  *nnzOut = acc;
  return 0;
}

void mlem_synth_free(void* p) { std::free(p); }

int mlem_synth_phantom(uint32_t nx, uint32_t ny, uint32_t nz, uint32_t nell, uint64_t seed,
                       float semi_lo, float semi_hi, const float* scale, float* x) {
  if (!x || nx == 0 || ny == 0 || nz == 0 || semi_lo <= 0.f || semi_hi < semi_lo) return -1;
  const uint32_t n[3] = {nx, ny, nz};
  struct Ell {
    double c[3], s[3], v;
  };
  std::vector<Ell> ells(nell);
  for (uint32_t e = 0; e < nell; ++e) {
    SplitMix64 rng(seed, e);
    for (int a = 0; a < 3; ++a) ells[e].c[a] = (rng.uniform() - 0.5) * 0.7 * n[a];
    for (int a = 0; a < 3; ++a) ells[e].s[a] = semi_lo + (semi_hi - semi_lo) * rng.uniform();
    ells[e].v = 0.5 + 3.5 * rng.uniform();
    if (scale) ells[e].v *= scale[e];
  }
  const int64_t nvox = (int64_t)nx * ny * nz;
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < nvox; ++j) {
    const int64_t ix = j % nx, iy = (j / nx) % ny, iz = j / ((int64_t)nx * ny);
    const double p[3] = {ix + 0.5 - 0.5 * nx, iy + 0.5 - 0.5 * ny, iz + 0.5 - 0.5 * nz};
    double v = 0.0;
    for (const Ell& E : ells) {
      double q = 0.0;
      for (int a = 0; a < 3; ++a) {
        const double t = (p[a] - E.c[a]) / E.s[a];
        q += t * t;
      }
      if (q <= 1.0) v += E.v;
    }
    x[j] = (float)v;
  }
  return 0;
}

int mlem_synth_project(uint32_t rows, const uint64_t* rowidx, const void* elements, const float* x,
                       double* ybar, int threads) {
  if (!rowidx || !elements || !x || !ybar) return -1;
  const mlem::Csr4Element* el = (const mlem::Csr4Element*)elements;
#pragma omp parallel for schedule(dynamic, 4096) num_threads(nthreads(threads))
  for (int64_t r = 0; r < (int64_t)rows; ++r) {
    const uint64_t b = r ? rowidx[r - 1] : 0, e = rowidx[r];
    double s = 0.0;
    for (uint64_t k = b; k < e; ++k) s += (double)el[k].value * (double)x[el[k].column];
    ybar[r] = s;
  }
  return 0;
}

int mlem_synth_poisson(uint32_t rows, const double* ybar, double total, uint64_t seed, int threads,
                       int32_t* y, int64_t* sum_y) {
  if (!ybar || !y) return -1;
  double s = 0.0;
  for (uint32_t r = 0; r < rows; ++r) s += ybar[r];
  const double scale = (total > 0.0 && s > 0.0) ? total / s : 1.0;
  int64_t acc = 0;
#pragma omp parallel for schedule(dynamic, 4096) num_threads(nthreads(threads)) reduction(+ : acc)
  for (int64_t r = 0; r < (int64_t)rows; ++r) {
    SplitMix64 rng(seed, (uint64_t)r);
    y[r] = poisson(ybar[r] * scale, rng);
    acc += y[r];
  }
  if (sum_y) *sum_y = acc;
  return 0;
}

int mlem_synth_uniform(uint64_t seed, uint64_t first, uint64_t n, double* out) {
  if (!out) return -1;
  for (uint64_t k = 0; k < n; ++k) {
    SplitMix64 rng(seed, first + k);
    out[k] = rng.uniform();
  }
  return 0;
}

int mlem_synth_siddon(const double p0[3], const double p1[3], uint32_t nx, uint32_t ny,
                      uint32_t nz, uint32_t cap, uint32_t* cols, float* lens) {
  if (!p0 || !p1 || nx == 0 || ny == 0 || nz == 0 || (cap && (!cols || !lens))) return -1;
  const uint32_t n[3] = {nx, ny, nz};
  std::vector<double> al;
  std::vector<Seg> segs;
  if (!siddon(p0, p1, n, al, segs)) return 0;
  const size_t k = std::min<size_t>(cap, segs.size());
  for (size_t i = 0; i < k; ++i) {
    cols[i] = segs[i].col;
    lens[i] = segs[i].len;
  }
  return (int)segs.size();
}

int mlem_synth_write_csr4(const char* path, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t rows,
                          uint64_t nnz, const uint64_t* rowidx, const void* elements) {
  if (!path || (rows && !rowidx) || (nnz && !elements)) return -1;
  try {
    mlem::csr4_write(path, 0, rows, nx * ny * nz, nnz, mlem::csr4_unit_scanner(nx, ny, nz), rowidx,
                     (const mlem::Csr4Element*)elements);
  } catch (const mlem::Csr4Error&) {
    return -2;
  }
  return 0;
}

}  // extern "C"
