// Interpolating B-spline coefficients (spline prefilter) for gfx950.
//
// Replaces `spline_coeff` / `spline_coeff_nd` (imported by
// nitorch/_C/grid.py:4 from the third-party torch-interpol and restated in
// the reference's nitorch/_C/_ts/coeff.py): along each axis, every line of
// samples is multiplied by the filter gain and then, per pole z of the
// order's B-spline, run through a causal recursion c[i] = x[i] + z c[i-1]
// and an anticausal one c[i] = z (c[i+1] - c[i]), whose first values come
// from the boundary condition (coeff.py:68-248):
//   zero / dct1      -> mirror rules (dct1_initial / dct1_final),
//   replicate / dct2 -> half-sample rules (dct2_initial / dct2_final),
//   dft              -> periodic rules (dft_initial / dft_final).
//
// MI355X layout: one lane owns one line; a 64-lane work-group stages its 64
// lines in LDS ([k][64 + 1] with a pad column, so the column walks of the
// recursions and the row-wise staging are both free of bank conflicts), so
// a line is read from HBM once and written once per axis whatever the
// number of poles.  Lines along the contiguous axis are staged from one
// contiguous chunk of 64 rows; lines along the other axes are staged one
// position at a time across 64 neighbouring lines (coalesced rows).  Lines
// longer than the LDS holds fall back to the same recursion in global
// memory.  Initial-value sums run in double over the horizon
// ceil(-30 / ln|z|) the reference uses.
#include <hip/hip_runtime.h>
#include <cmath>

#include "common.h"
#include "nitorch_amd.h"

namespace ni_amd {

#define CF_LINES 64
#define CF_PAD (CF_LINES + 1)
#define CF_GBLOCK 256

enum CoeffRule : int { CF_MIRROR = 0, CF_HALF = 1, CF_WRAP = 2 };

struct CoeffPlan {
  int np;           // number of poles (1..3)
  int rule;         // CoeffRule
  int m[3];         // horizon ceil(-30 / ln|z|) per pole
  double z[3];
  double gain;
};

// Poles of the order's interpolation filter (coeff.py:34-57).
static int coeff_poles(int order, double z[3]) {
  switch (order) {
    case 2: z[0] = std::sqrt(8.) - 3.; return 1;
    case 3: z[0] = std::sqrt(3.) - 2.; return 1;
    case 4:
      z[0] = std::sqrt(664. - std::sqrt(438976.)) + std::sqrt(304.) - 19.;
      z[1] = std::sqrt(664. + std::sqrt(438976.)) - std::sqrt(304.) - 19.;
      return 2;
    case 5:
      z[0] = std::sqrt(67.5 - std::sqrt(4436.25)) + std::sqrt(26.25) - 6.5;
      z[1] = std::sqrt(67.5 + std::sqrt(4436.25)) - std::sqrt(26.25) - 6.5;
      return 2;
    case 6:
      z[0] = -0.488294589303044755130118038883789062112279161239377608394;
      z[1] = -0.081679271076237512597937765737059080653379610398148178525368;
      z[2] = -0.00141415180832581775108724397655859252786416905534669851652709;
      return 3;
    case 7:
      z[0] = -0.5352804307964381655424037816816460718339231523426924148812;
      z[1] = -0.122554615192326690515272264359357343605486549427295558490763;
      z[2] = -0.0091486948096082769285930216516478534156925639545994482648003;
      return 3;
    default: return 0;
  }
}

template <typename T>
struct LdsLine {
  T *p;
  __device__ __forceinline__ T get(int k) const { return p[k * CF_PAD]; }
  __device__ __forceinline__ void set(int k, T v) const { p[k * CF_PAD] = v; }
};

template <typename T>
struct GlobalLine {
  T *p;
  long long s;
  __device__ __forceinline__ T get(int k) const { return p[(long long)k * s]; }
  __device__ __forceinline__ void set(int k, T v) const { p[(long long)k * s] = v; }
};

// z^e for an integer e >= 0 (z may be negative).
__device__ __forceinline__ double ipow(double z, int e) {
  double r = 1.;
  for (; e > 0; e >>= 1, z *= z)
    if (e & 1) r *= z;
  return r;
}

// Causal initial value c[0] of pole z on the (gained) line (coeff.py:68-176).
template <typename T, class LN>
__device__ double coeff_initial(const LN &a, int n, double z, int m, int rule) {
  if (rule == CF_WRAP) {
    const int mm = m < n ? m : n;
    double acc = (double)a.get(0), zk = z;
#pragma unroll 4
    for (int k = 1; k < mm; ++k) { acc += zk * (double)a.get(n - k); zk *= z; }
    return acc / (1. - zk);   // zk = z^mm
  }
  if (rule == CF_MIRROR) {
    if (m < n) {
      double acc = (double)a.get(0), zk = z;
#pragma unroll 4
      for (int k = 1; k < m; ++k) { acc += zk * (double)a.get(k); zk *= z; }
      return acc;
    }
    const double zn = ipow(z, n - 1);
    double acc = (double)a.get(0) + zn * (double)a.get(n - 1), zk = z;
    for (int k = 1; k < n - 1; ++k) { acc += (zk + zn * zn / zk) * (double)a.get(k); zk *= z; }
    return acc / (1. - zn * zn);
  }
  // half-sample symmetric: weights z^k + z^(2n-1-k) for 0 < k < n-1; on lines
  // longer than 2m+2 every z^(2n-1-k) term is below z^m and is dropped
  const double x0 = (double)a.get(0);
  const double zn = ipow(z, n);
  const double last = zn * (1. + 1. / (z + zn * zn));
  double acc = x0 + last * (double)a.get(n - 1);
  const bool full = n <= 2 * m + 2;
  const int kend = full ? n - 1 : (m + 1 < n - 1 ? m + 1 : n - 1);
  double zk = z;
#pragma unroll 4
  for (int k = 1; k < kend; ++k) { acc += zk * (double)a.get(k); zk *= z; }
  if (full) {
    double zr = zn * z;   // z^(n+1) at k = n-2
    for (int k = n - 2; k >= 1; --k) { acc += zr * (double)a.get(k); zr *= z; }
  }
  return acc * (z / (1. - zn * zn)) + x0;
}

// Anticausal initial value c[n-1] of pole z on the causal output
// (coeff.py:179-222).
template <typename T, class LN>
__device__ double coeff_final(const LN &a, int n, double z, int m, int rule) {
  if (rule == CF_WRAP) {
    const int mm = m < n ? m : n;
    double acc = z * (double)a.get(n - 1), zk = z * z;
#pragma unroll 4
    for (int j = 0; j < mm - 1; ++j) { acc += zk * (double)a.get(j); zk *= z; }
    return acc / (ipow(z, mm) - 1.);
  }
  if (rule == CF_MIRROR)
    return (z * (double)a.get(n - 2) + (double)a.get(n - 1)) * (z / (z * z - 1.));
  return (double)a.get(n - 1) * (z / (z - 1.));
}

// All poles of one line (coeff.py:251-275); the gain is already applied.
template <typename T, class LN>
__device__ void coeff_line(const LN &a, int n, const CoeffPlan &P) {
  for (int q = 0; q < P.np; ++q) {
    const double z = P.z[q];
    const T zt = (T)z;
    T prev = (T)coeff_initial<T>(a, n, z, P.m[q], P.rule);
    a.set(0, prev);
    // unrolled so that the reads of the next samples are in flight while the
    // recursion runs (they do not depend on it)
#pragma unroll 8
    for (int k = 1; k < n; ++k) {
      prev = a.get(k) + zt * prev;
      a.set(k, prev);
    }
    prev = (T)coeff_final<T>(a, n, z, P.m[q], P.rule);
    a.set(n - 1, prev);
#pragma unroll 8
    for (int k = n - 2; k >= 0; --k) {
      prev = zt * (prev - a.get(k));
      a.set(k, prev);
    }
  }
}

// (line, k) = (e / n, e % n) of a lane's chunk position e = t + 64 j,
// advanced by 64 per step without a division.
struct CfWalk {
  int line, k, q, r;
  __device__ CfWalk(int t, int n) : line(t / n), k(t % n), q(CF_LINES / n), r(CF_LINES % n) {}
  __device__ __forceinline__ void step(int n) {
    k += r;
    line += q;
    const bool c = k >= n;
    k -= c ? n : 0;
    line += c;
  }
};

// L lines of length n; line l starts at (l / I) * n * I + l % I with stride I.
template <typename T>
__global__ __launch_bounds__(CF_LINES) void coeff_lds(T *__restrict__ data, long long L, int n, long long I,
                                                      CoeffPlan P) {
  extern __shared__ char cf_smem[];
  T *buf = (T *)cf_smem;
  const long long l0 = (long long)blockIdx.x * CF_LINES;
  const int t = threadIdx.x;
  const int nl = (int)(L - l0 < CF_LINES ? L - l0 : CF_LINES);
  const T g = (T)P.gain;
  if (I == 1) {
    // rows l0 .. l0 + nl - 1 are one contiguous chunk of nl * n values
    const T *src = data + l0 * n;
    const int cnt = (nl * n - t + CF_LINES - 1) / CF_LINES;
    CfWalk w(t, n);
#pragma unroll 16
    for (int j = 0; j < cnt; ++j) {
      buf[w.k * CF_PAD + w.line] = src[t + j * CF_LINES] * g;
      w.step(n);
    }
  } else if (t < nl) {
    const long long l = l0 + t;
    const T *src = data + (l / I) * n * I + l % I;
#pragma unroll 32
    for (int k = 0; k < n; ++k) buf[k * CF_PAD + t] = src[(long long)k * I] * g;
  }
  __syncthreads();
  if (t < nl) coeff_line<T>(LdsLine<T>{buf + t}, n, P);
  __syncthreads();
  if (I == 1) {
    T *dst = data + l0 * n;
    const int cnt = (nl * n - t + CF_LINES - 1) / CF_LINES;
    CfWalk w(t, n);
#pragma unroll 16
    for (int j = 0; j < cnt; ++j) {
      dst[t + j * CF_LINES] = buf[w.k * CF_PAD + w.line];
      w.step(n);
    }
  } else if (t < nl) {
    const long long l = l0 + t;
    T *dst = data + (l / I) * n * I + l % I;
#pragma unroll 32
    for (int k = 0; k < n; ++k) dst[(long long)k * I] = buf[k * CF_PAD + t];
  }
}

// Lines too long for LDS: the same recursion in place in global memory.
template <typename T>
__global__ __launch_bounds__(CF_GBLOCK) void coeff_global(T *__restrict__ data, long long L, int n, long long I,
                                                          CoeffPlan P) {
  const long long l = (long long)blockIdx.x * CF_GBLOCK + threadIdx.x;
  if (l >= L) return;
  const GlobalLine<T> a{data + (l / I) * n * I + l % I, I};
  const T g = (T)P.gain;
  for (int k = 0; k < n; ++k) a.set(k, a.get(k) * g);
  coeff_line<T>(a, n, P);
}

// Dynamic LDS a coeff_lds work-group may use on the current device, cached
// per device (64 KB if the query fails).
template <typename T>
static size_t coeff_lds_limit() {
  static size_t cache[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 65536;
  if (cache[dev]) return cache[dev];
  int v = 0;
  size_t lim = 65536;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess &&
      (size_t)v > lim &&
      hipFuncSetAttribute((const void *)coeff_lds<T>, hipFuncAttributeMaxDynamicSharedMemorySize, v) ==
          hipSuccess)
    lim = (size_t)v;
  (void)hipGetLastError();
  cache[dev] = lim;
  return lim;
}

template <typename T>
static void coeff_axis(T *data, long long L, int n, long long I, const CoeffPlan &P, hipStream_t st) {
  const size_t bytes = (size_t)n * CF_PAD * sizeof(T);
  if (bytes <= coeff_lds_limit<T>()) {
    const long long nb = (L + CF_LINES - 1) / CF_LINES;   // < 2^31: checked by the caller
    hipLaunchKernelGGL(coeff_lds<T>, dim3((unsigned)nb), dim3(CF_LINES), bytes, st, data, L, n, I, P);
  } else {
    const long long nb = (L + CF_GBLOCK - 1) / CF_GBLOCK;
    hipLaunchKernelGGL(coeff_global<T>, dim3((unsigned)nb), dim3(CF_GBLOCK), 0, st, data, L, n, I, P);
  }
}

static int coeff_rule(int bound) {
  switch (bound) {
    case B_ZERO: case B_DCT1: return CF_MIRROR;
    case B_REPLICATE: case B_DCT2: return CF_HALF;
    case B_DFT: return CF_WRAP;
    default: return -1;
  }
}

}  // namespace ni_amd

using namespace ni_amd;

extern "C" {

int ni_spline_coeff(int32_t dtype, int64_t nbatch, int32_t ndim, const int64_t *shape, const int32_t *bound,
                    const int32_t *order, void *data, void *stream) {
  set_error("");
  NI_CHECK(shape && bound && order, "spline_coeff: missing argument");
  NI_CHECK(ndim >= 1 && ndim <= 8, "spline_coeff: expected 1 to 8 filtered dimensions, got %d", ndim);
  NI_CHECK(nbatch >= 0, "spline_coeff: negative batch size");
  long long total = nbatch;
  for (int d = 0; d < ndim; ++d) {
    NI_CHECK(shape[d] >= 0, "spline_coeff: negative size along dimension %d", d);
    NI_CHECK(order[d] >= 0 && order[d] <= 7, "spline_coeff: interpolation order %d not in 0..7", order[d]);
    total *= shape[d];
  }
  NI_CHECK(dtype == NI_FLOAT32 || dtype == NI_FLOAT64, "spline_coeff: unsupported dtype %d", dtype);
  if (total == 0) return 0;
  NI_CHECK(data, "spline_coeff: missing data");
  hipStream_t st = (hipStream_t)stream;
  for (int d = 0; d < ndim; ++d) {
    CoeffPlan P;
    P.np = coeff_poles(order[d], P.z);
    if (P.np == 0 || shape[d] == 1) continue;
    P.rule = coeff_rule(bound[d]);
    NI_CHECK(P.rule >= 0, "spline_coeff: bound %d is not supported (zero, dct1, replicate, dct2 and dft are)",
             bound[d]);
    NI_CHECK(shape[d] <= (1LL << 30), "spline_coeff: line of %lld samples is too long", (long long)shape[d]);
    P.gain = 1.0;
    for (int q = 0; q < P.np; ++q) {
      P.gain *= (1. - P.z[q]) * (1. - 1. / P.z[q]);
      P.m[q] = (int)std::ceil(-30. / std::log(std::fabs(P.z[q])));
    }
    long long I = 1;
    for (int e = d + 1; e < ndim; ++e) I *= shape[e];
    const long long L = total / shape[d];
    NI_CHECK(L / CF_LINES < (1LL << 31) - 1, "spline_coeff: %lld lines exceed the launch limit", L);
    if (dtype == NI_FLOAT32) coeff_axis<float>((float *)data, L, (int)shape[d], I, P, st);
    else coeff_axis<double>((double *)data, L, (int)shape[d], I, P, st);
    NI_HIP_CHECK(hipGetLastError());
  }
  return 0;
}

}  // extern "C"
