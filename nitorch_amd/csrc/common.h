// Shared device/host helpers for the nitorch_amd HIP library (gfx950).
//
//  * boundary folding (index + sign) for the reference's BoundType codes
//    (nitorch/_C/src/bounds.h:7-17, semantics of bounds_common.h:28-219);
//  * B-spline weight / derivative evaluation from the generated piece tables
//    (semantics of interpolation_common.h:57-969);
//  * error reporting for the C ABI (thread-local message, integer status);
//  * stream-ordered scratch from a library-private memory pool.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cmath>
#include <mutex>

#include "spline_tables.h"

namespace ni_amd {

// --------------------------------------------------------------------------
// Stream-ordered scratch.  Allocations come from a memory pool per device
// owned by the library, which keeps up to SCRATCH_KEEP bytes reserved when
// they are freed (the device's default pool releases freed memory at every
// synchronisation, so each call would map it again: ~70 us a call, which
// dominates small pulls and pushes).  Falls back to hipMallocAsync on the
// default pool when a private pool cannot be made.
// --------------------------------------------------------------------------
#define SCRATCH_KEEP (uint64_t(256) << 20)

inline hipMemPool_t scratch_pool() {
  static std::mutex mu;
  static hipMemPool_t pools[64] = {};
  static bool tried[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
    (void)hipGetLastError();
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(mu);
  if (!tried[dev]) {
    tried[dev] = true;
    hipMemPoolProps props = {};
    props.allocType = hipMemAllocationTypePinned;
    props.handleTypes = hipMemHandleTypeNone;
    props.location.type = hipMemLocationTypeDevice;
    props.location.id = dev;
    hipMemPool_t p = nullptr;
    if (hipMemPoolCreate(&p, &props) == hipSuccess) {
      uint64_t keep = SCRATCH_KEEP;
      if (hipMemPoolSetAttribute(p, hipMemPoolAttrReleaseThreshold, &keep) == hipSuccess) pools[dev] = p;
      else (void)hipMemPoolDestroy(p);
    }
    (void)hipGetLastError();
  }
  return pools[dev];
}

inline hipError_t scratch_alloc(void **ptr, size_t bytes, hipStream_t st) {
  if (hipMemPool_t p = scratch_pool()) {
    if (hipMallocFromPoolAsync(ptr, bytes, p, st) == hipSuccess) return hipSuccess;
    (void)hipGetLastError();
  }
  return hipMallocAsync(ptr, bytes, st);
}

enum BoundCode : int {
  B_REPLICATE = 0, B_DCT1 = 1, B_DCT2 = 2, B_DST1 = 3, B_DST2 = 4,
  B_DFT = 5, B_SLIDING = 6, B_ZERO = 7, B_NOCHECK = 8
};

// --------------------------------------------------------------------------
// Boundary folding.  Returns the in-range index of `i` on an axis of length
// `n` and sets `s` to the multiplicative sign (+1, -1, or 0 = "outside, reads
// as zero").  In-range indices short-circuit (every bound maps them to
// themselves with sign +1), so the integer modulo only runs on border taps.
// --------------------------------------------------------------------------
template <typename I>
__host__ __device__ inline I pos_mod(I a, I m) {
  I r = a % m;
  return r < 0 ? r + m : r;
}

template <typename I>
__host__ __device__ inline I fold(int b, I i, I n, int &s) {
  s = 1;
  if (i >= 0 && i < n) return i;
  switch (b) {
    case B_REPLICATE:
      return i < 0 ? I(0) : n - 1;
    case B_DCT1: {
      if (n == 1) return 0;
      I p = 2 * (n - 1);
      I k = pos_mod<I>(i < 0 ? -i : i, p);
      return k >= n ? p - k : k;
    }
    case B_DCT2: {
      I p = 2 * n;
      I k = pos_mod<I>(i, p);
      return k >= n ? p - 1 - k : k;
    }
    case B_DST2: {
      I p = 2 * n;
      I k = pos_mod<I>(i, p);
      I j = i < 0 ? n - i - 1 : i;
      s = ((j / n) & 1) ? -1 : 1;
      return k >= n ? p - 1 - k : k;
    }
    case B_DST1: {
      if (n == 1) return 0;
      I p = 2 * (n + 1);
      I j = i < 0 ? n - i - 1 : i;
      j = j % p;
      if (j % (n + 1) == n) s = 0;
      else s = ((j / (n + 1)) & 1) ? -1 : 1;
      I k = (i == -1) ? I(0) : (i < 0 ? -i - 2 : i);
      k = pos_mod<I>(k, p);
      if (k == n) return n - 1;
      return k > n ? p - k - 2 : k;
    }
    case B_DFT:
      return pos_mod<I>(i, n);
    default:  // zero / nocheck / sliding: outside reads as zero
      s = 0;
      return 0;
  }
}

// --------------------------------------------------------------------------
// Spline evaluation.  For order O, the taps of a sample at coordinate x are
// lo, lo+1, ..., lo+O with lo = floor(x - (O-1)/2) (round(x) for O = 0);
// tap k always falls on the same polynomial piece, so the piece index is a
// compile-time constant when O is.
// --------------------------------------------------------------------------
template <int O>
__host__ __device__ constexpr int tap_piece(int k) {
  return (O & 1) ? (k <= (O - 1) / 2 ? (O - 1) / 2 - k : k - (O + 1) / 2)
                 : (k <= O / 2 ? O / 2 - k : k - O / 2);
}

template <int O>
__host__ __device__ constexpr double piece_centre(int p) {
  return (O & 1) ? p + 0.5 : double(p);
}

template <typename T, int ND>
__host__ __device__ inline T horner(const double (&c)[ND], T u) {
  T r = T(c[ND - 1]);
#pragma unroll
  for (int i = ND - 2; i >= 0; --i) r = r * u + T(c[i]);
  return r;
}

// weight / derivative / second derivative of tap k at signed distance d
template <typename T, int O>
__host__ __device__ inline T spl_w(int k, T d) {
  if constexpr (O == 0) {
    return T(1);
  } else {
    const int p = tap_piece<O>(k);
    T u = (d < T(0) ? -d : d) - T(piece_centre<O>(p));
    return horner<T, SplTab<O>::NDW>(SplTab<O>::W[p], u);
  }
}

template <typename T, int O>
__host__ __device__ inline T spl_g(int k, T d) {
  if constexpr (O == 0) {
    return T(0);
  } else {
    const int p = tap_piece<O>(k);
    T a = d < T(0) ? -d : d;
    T g = horner<T, SplTab<O>::NDG>(SplTab<O>::G[p], a - T(piece_centre<O>(p)));
    return d < T(0) ? -g : g;
  }
}

template <typename T, int O>
__host__ __device__ inline T spl_h(int k, T d) {
  if constexpr (O <= 1) {
    return T(0);
  } else if constexpr (O == 2) {
    T a = d < T(0) ? -d : d;
    return a < T(0.5) ? T(-2) : T(1);  // discontinuous: outer piece at |d| = 1/2
  } else {
    const int p = tap_piece<O>(k);
    T a = d < T(0) ? -d : d;
    return horner<T, SplTab<O>::NDH>(SplTab<O>::H[p], a - T(piece_centre<O>(p)));
  }
}

// Runtime-order variants (anisotropic orders, generic kernel).
template <typename T, int O>
__host__ __device__ inline void spl_all_rt(int k, T d, T &w, T &g, T &h) {
  w = spl_w<T, O>(k, d);
  g = spl_g<T, O>(k, d);
  h = spl_h<T, O>(k, d);
}

template <typename T>
__host__ __device__ inline void spl_eval(int order, int k, T d, T &w, T &g, T &h) {
  switch (order) {
    case 0: spl_all_rt<T, 0>(k, d, w, g, h); break;
    case 1: spl_all_rt<T, 1>(k, d, w, g, h); break;
    case 2: spl_all_rt<T, 2>(k, d, w, g, h); break;
    case 3: spl_all_rt<T, 3>(k, d, w, g, h); break;
    case 4: spl_all_rt<T, 4>(k, d, w, g, h); break;
    case 5: spl_all_rt<T, 5>(k, d, w, g, h); break;
    case 6: spl_all_rt<T, 6>(k, d, w, g, h); break;
    default: spl_all_rt<T, 7>(k, d, w, g, h); break;
  }
}

template <typename T>
__host__ __device__ inline T first_tap_offset(int order) {
  return T(order - 1) * T(0.5);
}

// The outermost taps (k = 0 and k = O) lie on the spline's last piece,
// (h - |d|)^O / O! with h = (O + 1) / 2.  Evaluated from the coordinate x
// and the first tap lo: s = (lo + h) - x for k = 0, s = x - (lo + O - h)
// for k = O.  Both differences are exact whenever s is small (Sterbenz), so
// a tiny outer weight keeps the relative accuracy of T; computed as
// h - |x - tap|, it would carry the rounding of x - tap (for |x| < 1 an ulp
// of the tap distance, relative error 1e-3 on a weight of 1e-4 in float).
__host__ __device__ constexpr double inv_fact(int n) {
  return n <= 1 ? 1.0 : inv_fact(n - 1) / n;
}

template <typename T, typename I>
__host__ __device__ inline T spl_outer_s(int O, int k, T x, I lo) {
  return k == 0 ? (T(lo) + T(0.5 * (O + 1))) - x : x - (T(lo) + T(0.5 * (O - 1)));
}

// weight, first and second derivative (w.r.t. x) on the last piece
template <typename T, int O>
__host__ __device__ inline void spl_outer(bool left, T s, T &w, T &g, T &h) {
  T p[O + 1];
  p[0] = T(1);
#pragma unroll
  for (int i = 1; i <= O; ++i) p[i] = p[i - 1] * s;
  // constexpr variables: the factorials fold at compile time (a call of
  // the recursive inv_fact left in device code needs a dynamic stack)
  constexpr double f0 = inv_fact(O), f1 = inv_fact(O - 1), f2 = inv_fact(O - 2);
  w = p[O] * T(f0);
  const T g0 = O >= 1 ? p[O >= 1 ? O - 1 : 0] * T(f1) : T(0);
  g = left ? -g0 : g0;
  h = O >= 2 ? p[O >= 2 ? O - 2 : 0] * T(f2) : T(0);
}

template <typename T>
__host__ __device__ inline void spl_outer_rt(int order, bool left, T s, T &w, T &g, T &h) {
  switch (order) {
    case 1: spl_outer<T, 1>(left, s, w, g, h); break;
    case 2: spl_outer<T, 2>(left, s, w, g, h); break;
    case 3: spl_outer<T, 3>(left, s, w, g, h); break;
    case 4: spl_outer<T, 4>(left, s, w, g, h); break;
    case 5: spl_outer<T, 5>(left, s, w, g, h); break;
    case 6: spl_outer<T, 6>(left, s, w, g, h); break;
    default: spl_outer<T, 7>(left, s, w, g, h); break;
  }
}

// --------------------------------------------------------------------------
// Error reporting for the C ABI.
// --------------------------------------------------------------------------
void set_error(const char *fmt, ...);
const char *get_error();

}  // namespace ni_amd

#define NI_CHECK(cond, ...)                     \
  do {                                          \
    if (!(cond)) {                              \
      ::ni_amd::set_error(__VA_ARGS__);         \
      return 1;                                 \
    }                                           \
  } while (0)

#define NI_HIP_CHECK(expr)                                               \
  do {                                                                   \
    hipError_t _e = (expr);                                              \
    if (_e != hipSuccess) {                                              \
      ::ni_amd::set_error("%s: HIP error %s", __func__,                  \
                          hipGetErrorString(_e));                        \
      return 2;                                                          \
    }                                                                    \
  } while (0)
