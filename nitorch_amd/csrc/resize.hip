// Generic resize of [N, C, *spatial] tensors (1-3 spatial dims) for gfx950.
//
// Replaces ni::resize (nitorch/_C/src/resize.cpp:25-59, implementation
// nitorch/_C/src/impl/resize_common.cpp): every target voxel w samples the
// source at the separable affine coordinate x = scale * w + shift (per axis,
// computed in double and rounded to the data type as the reference does)
// with the order-`order` B-spline and the boundary folding of the pull --
// with no out-of-bound masking (the reference's resize always interpolates).
// `do_adjoint` swaps the roles (source = output, target = input) and splats
// instead (the transpose), into a zeroed source.  shift / scale per axis
// follow prepare_affine (resize_common.cpp:1548-1567):
//   edge:   shift = (Ds/Dt - 1) / 2,            scale = Ds / Dt
//   center: shift = 0,                          scale = (Ds - 1) / (Dt - 1)
//   first:  shift = 0,                          scale = 1 / f
//   last:   shift = f (Ds - 1) / (Dt - 1),      scale = 1 / f
// with f the factor (inverted when do_adjoint), and `normalize` multiplies
// the result by the scales with the reference's fall-through switch
// (resize_common.cpp:1640-1646): 1-D by s0 * (s0 s1) * (s0 s1 s2), 2-D by
// (s0 s1) * (s0 s1 s2), 3-D by s0 s1 s2.
//
// MI355X layout: one lane per target voxel, z fastest, so the lanes of a
// wave read neighbouring source voxels and write (or, adjoint, scatter
// into) one contiguous run; per-axis taps, weights, signs and folded
// offsets are computed once per lane and reused over the channels.
#include <hip/hip_runtime.h>
#include <cmath>

#include "common.h"
#include "nitorch_amd.h"

namespace ni_amd {

#define RS_BLOCK 256
#define RS_MAXT 8

struct RsArgs {
  int N, C;
  int ns[3], nt[3];   // source / target sizes (1 for unused axes)
  int bound[3], order[3];
  double shift[3], scale[3];
  int accumulate;     // pull: add into the target instead of overwriting it
  int nearest_dbl;    // isotropic nearest: round scale * w + shift in double
};

template <typename T> __device__ inline T rs_floor(T x);
template <> __device__ inline float rs_floor<float>(float x) { return floorf(x); }
template <> __device__ inline double rs_floor<double>(double x) { return floor(x); }
template <typename T> __device__ inline T rs_round(T x);
template <> __device__ inline float rs_round<float>(float x) { return roundf(x); }
template <> __device__ inline double rs_round<double>(double x) { return round(x); }

// Taps of target index w along one axis: count, weights, signed source offsets.
template <typename T>
__device__ inline int rs_taps(const RsArgs &a, int d, int w, long long stride, T (&wt)[RS_MAXT],
                              long long (&off)[RS_MAXT], int (&sg)[RS_MAXT]) {
#pragma clang fp contract(off)
  const int o = a.order[d];
  // product and sum rounded separately, as the reference computes them (a
  // fused multiply-add moves coordinates such as 4.499999999999999 onto 4.5)
  const double xd = a.scale[d] * (double)w + a.shift[d];
  const T x = (T)xd;
  const int lo = o == 0 ? (a.nearest_dbl ? (int)round(xd) : (int)rs_round<T>(x))
                        : (int)rs_floor<T>(x - first_tap_offset<T>(o));
#pragma unroll
  for (int t = 0; t < RS_MAXT; ++t) {
    if (t > o) break;
    const int b = lo + t;
    T g, h;
    spl_eval<T>(o, t, x - T(b), wt[t], g, h);
    int s;
    const int j = fold<int>(a.bound[d], b, a.ns[d], s);
    sg[t] = s;
    off[t] = (long long)j * stride;
  }
  return o + 1;
}

// ADJ == false: out (target) [+]= sum_taps w * s * in(source)
// ADJ == true : out (source) += w * s * in(target), by atomics
template <typename T, bool ADJ>
__global__ __launch_bounds__(RS_BLOCK) void resize_kernel(const T *__restrict__ in, T *__restrict__ out,
                                                          RsArgs a) {
  const long long Vt = (long long)a.nt[0] * a.nt[1] * a.nt[2];
  const long long Vs = (long long)a.ns[0] * a.ns[1] * a.ns[2];
  const long long gid = (long long)blockIdx.x * RS_BLOCK + threadIdx.x;
  if (gid >= (long long)a.N * Vt) return;
  const int n = (int)(gid / Vt);
  const long long v = gid - (long long)n * Vt;
  const int wz = (int)(v % a.nt[2]);
  const int wy = (int)((v / a.nt[2]) % a.nt[1]);
  const int wx = (int)(v / ((long long)a.nt[2] * a.nt[1]));
  T w0[RS_MAXT], w1[RS_MAXT], w2[RS_MAXT];
  long long o0[RS_MAXT], o1[RS_MAXT], o2[RS_MAXT];
  int s0[RS_MAXT], s1[RS_MAXT], s2[RS_MAXT];
  const int t0 = rs_taps<T>(a, 0, wx, (long long)a.ns[1] * a.ns[2], w0, o0, s0);
  const int t1 = rs_taps<T>(a, 1, wy, (long long)a.ns[2], w1, o1, s1);
  const int t2 = rs_taps<T>(a, 2, wz, 1, w2, o2, s2);
  const long long C = a.C;
  if constexpr (!ADJ) {
    const T *src = in + (long long)n * C * Vs;
    T *tgt = out + (long long)n * C * Vt + v;
    for (long long c = 0; c < C; ++c) {
      const T *sc = src + c * Vs;
      T acc = a.accumulate ? tgt[c * Vt] : T(0);
      // same order as the reference: z taps outer, x taps inner,
      // weight (wz * wy) * wx
      for (int k = 0; k < t2; ++k)
        for (int j = 0; j < t1; ++j) {
          const int syz = s2[k] * s1[j];
          const T wyz = w2[k] * w1[j];
          const long long oyz = o2[k] + o1[j];
          for (int i = 0; i < t0; ++i) {
            const int s = syz * s0[i];
            if (s) {
              const T val = sc[oyz + o0[i]];
              acc += (s < 0 ? -val : val) * (wyz * w0[i]);
            }
          }
        }
      tgt[c * Vt] = acc;
    }
  } else {
    const T *tg = in + (long long)n * C * Vt + v;
    T *src = out + (long long)n * C * Vs;
    for (int k = 0; k < t2; ++k)
      for (int j = 0; j < t1; ++j) {
        const int syz = s2[k] * s1[j];
        const T wyz = w2[k] * w1[j];
        const long long oyz = o2[k] + o1[j];
        for (int i = 0; i < t0; ++i) {
          const int s = syz * s0[i];
          if (!s) continue;
          const T wxyz = wyz * w0[i];
          for (long long c = 0; c < C; ++c) {
            const T val = tg[c * Vt] * wxyz;
            unsafeAtomicAdd(src + c * Vs + oyz + o0[i], s < 0 ? -val : val);
          }
        }
      }
  }
}

template <typename T>
__global__ __launch_bounds__(RS_BLOCK) void rs_fill(T *__restrict__ p, long long n, T v) {
  const long long i = (long long)blockIdx.x * RS_BLOCK + threadIdx.x;
  if (i < n) p[i] = v;
}

template <typename T>
__global__ __launch_bounds__(RS_BLOCK) void rs_scale(T *__restrict__ p, long long n, T f) {
  const long long i = (long long)blockIdx.x * RS_BLOCK + threadIdx.x;
  if (i < n) p[i] *= f;
}

static unsigned rs_blocks(long long n) { return (unsigned)((n + RS_BLOCK - 1) / RS_BLOCK); }

// shift and scale of one axis (resize_common.cpp:1548-1567); GridAlignType
// edge = 0, center = 1, first = 2, last = 3
static void rs_affine(double f, double Ds, double Dt, int mode, double &shift, double &scale) {
  shift = 0.;
  scale = 1. / f;
  switch (mode) {
    case 0: shift = 0.5 * (Ds / Dt - 1); scale = Ds / Dt; break;
    case 1: shift = 0.; scale = (Ds - 1) / (Dt - 1); break;
    case 3: shift = f * (Ds - 1) / (Dt - 1); break;
    default: break;
  }
}

template <typename T>
static void rs_launch(const T *input, T *output, const RsArgs &a, bool adjoint, hipStream_t st) {
  const long long Vt = (long long)a.nt[0] * a.nt[1] * a.nt[2];
  const long long nthreads = (long long)a.N * Vt;
  if (adjoint) {
    const long long ns = (long long)a.N * a.C * a.ns[0] * a.ns[1] * a.ns[2];
    hipLaunchKernelGGL(rs_fill<T>, dim3(rs_blocks(ns)), dim3(RS_BLOCK), 0, st, output, ns, T(0));
    hipLaunchKernelGGL((resize_kernel<T, true>), dim3(rs_blocks(nthreads)), dim3(RS_BLOCK), 0, st, input, output, a);
  } else {
    hipLaunchKernelGGL((resize_kernel<T, false>), dim3(rs_blocks(nthreads)), dim3(RS_BLOCK), 0, st, input, output, a);
  }
}

}  // namespace ni_amd

using namespace ni_amd;

extern "C" {

int ni_resize(int32_t dtype, int32_t dim, int64_t N, int64_t C, const int64_t *in_shape, const int64_t *out_shape,
              const double *factor, const int32_t *bound, const int32_t *order, const int32_t *mode,
              int32_t do_adjoint, int32_t normalize, int32_t accumulate, const void *input, void *output,
              void *stream) {
  set_error("");
  NI_CHECK(in_shape && out_shape && factor && bound && order && mode && input && output, "resize: missing argument");
  NI_CHECK(dim >= 1 && dim <= 3, "resize: expected a 1D, 2D or 3D tensor (got %d spatial dims)", dim);
  NI_CHECK(dtype == NI_FLOAT32 || dtype == NI_FLOAT64, "resize: unsupported dtype %d", dtype);
  NI_CHECK(N > 0 && C > 0, "resize: expected a non empty tensor");
  RsArgs a;
  a.N = (int)N;
  a.C = (int)C;
  // the isotropic nearest / linear kernels (and quadratic in 3-D) overwrite
  // a given output, the others add into it (resize_common.cpp:440-530)
  const bool iso = order[0] == order[1] && order[0] == order[2];
  const bool overwrite = iso && (order[0] <= 1 || (order[0] == 2 && dim == 3));
  a.accumulate = (do_adjoint || overwrite) ? 0 : (accumulate != 0);
  a.nearest_dbl = iso && order[0] == 0;
  // source / target (resize.cpp:52-53)
  const int64_t *ss = do_adjoint ? out_shape : in_shape;
  const int64_t *ts = do_adjoint ? in_shape : out_shape;
  double f[3], shift[3], scale[3];
  f[0] = factor[0]; f[1] = factor[1]; f[2] = factor[2];
  for (int d = 0; d < 3; ++d) {
    if (do_adjoint) f[d] = 1. / f[d];
    const double Ds = d < dim ? (double)ss[d] : 1., Dt = d < dim ? (double)ts[d] : 1.;
    NI_CHECK(Ds > 0 && Dt > 0, "resize: expected a non empty tensor");
    rs_affine(f[d], Ds, Dt, mode[d], shift[d], scale[d]);
    NI_CHECK(order[d] >= 0 && order[d] <= 7, "resize: interpolation order %d not in 0..7", order[d]);
    if (d < dim) {
      a.ns[d] = (int)ss[d];
      a.nt[d] = (int)ts[d];
      a.bound[d] = bound[d];
      a.order[d] = order[d];
      a.shift[d] = shift[d];
      a.scale[d] = scale[d];
    } else {
      // axes the reference's 1-D / 2-D kernels do not touch: identity
      a.ns[d] = a.nt[d] = 1;
      a.bound[d] = B_DCT2;
      a.order[d] = 0;
      a.shift[d] = 0.;
      a.scale[d] = 1.;
    }
  }
  hipStream_t st = (hipStream_t)stream;
  if (dtype == NI_FLOAT32) rs_launch<float>((const float *)input, (float *)output, a, do_adjoint != 0, st);
  else rs_launch<double>((const double *)input, (double *)output, a, do_adjoint != 0, st);
  NI_HIP_CHECK(hipGetLastError());
  if (normalize) {
    double m = scale[0] * scale[1] * scale[2];   // 3-D (the switch's default)
    if (dim == 2) m *= scale[0] * scale[1];
    if (dim == 1) m *= scale[0] * (scale[0] * scale[1]);
    const int64_t *os = do_adjoint ? ss : ts;
    long long n = N * C;
    for (int d = 0; d < dim; ++d) n *= os[d];
    if (dtype == NI_FLOAT32)
      hipLaunchKernelGGL(rs_scale<float>, dim3(rs_blocks(n)), dim3(RS_BLOCK), 0, st, (float *)output, n, (float)m);
    else
      hipLaunchKernelGGL(rs_scale<double>, dim3(rs_blocks(n)), dim3(RS_BLOCK), 0, st, (double *)output, n, m);
    NI_HIP_CHECK(hipGetLastError());
  }
  return 0;
}

}  // extern "C"
