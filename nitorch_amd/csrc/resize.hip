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
#include <algorithm>
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

// One axis of the separable forward resize: in [A, nin, B] -> out [A, nout, B],
// out[a, j, b] = (accumulate ? out : 0) + sum_t w_t(j) s_t in[a, fold(lo_j + t), b].
// Lanes run along b (contiguous), so every tap is a coalesced row read.
template <typename T>
__global__ __launch_bounds__(RS_BLOCK) void resize_axis(const T *__restrict__ in, T *__restrict__ out,
                                                        long long A, long long B, RsArgs a, int d,
                                                        int accumulate) {
  const int nout = a.nt[d], nin = a.ns[d];
  const long long gid = (long long)blockIdx.x * RS_BLOCK + threadIdx.x;
  if (gid >= A * nout * B) return;
  const long long b = gid % B;
  const int j = (int)((gid / B) % nout);
  const long long o = gid / (B * nout);
  T w[RS_MAXT];
  long long off[RS_MAXT];
  int sg[RS_MAXT];
  const int nt = rs_taps<T>(a, d, j, B, w, off, sg);
  const T *src = in + o * nin * B + b;
  T acc = accumulate ? out[gid] : T(0);
  for (int t = 0; t < nt; ++t) {
    if (sg[t]) {
      const T v = src[off[t]];
      acc += (sg[t] < 0 ? -v : v) * w[t];
    }
  }
  out[gid] = acc;
}

// Logical block id with the blocks of each XCD contiguous: the hardware
// deals consecutive block ids round-robin over the 8 XCDs, so output rows
// that share input rows would otherwise fetch them through different L2s.
__device__ __forceinline__ unsigned rs_xcd_block() {
  const unsigned b = blockIdx.x, n = gridDim.x;
  const unsigned xcd = b % 8u, local = b / 8u, q = n / 8u, r = n % 8u;
  return xcd * q + (xcd < r ? xcd : r) + local;
}

// Tap table of one axis: for each target index j, RS_MAXT entries of folded
// source index (-1: no tap or sign 0, which the reference never reads) and
// weight times sign.
template <typename T>
__global__ __launch_bounds__(RS_BLOCK) void rs_table(RsArgs a, int d, T *__restrict__ tw, int *__restrict__ ti) {
  const int j = blockIdx.x * RS_BLOCK + threadIdx.x;
  if (j >= a.nt[d]) return;
  T w[RS_MAXT];
  long long off[RS_MAXT];
  int sg[RS_MAXT];
  const int nt = rs_taps<T>(a, d, j, 1, w, off, sg);
#pragma unroll
  for (int t = 0; t < RS_MAXT; ++t) {
    const bool on = t < nt && sg[t] != 0;
    tw[j * RS_MAXT + t] = on ? (sg[t] < 0 ? -w[t] : w[t]) : T(0);
    ti[j * RS_MAXT + t] = on ? (int)off[t] : -1;
  }
}

// The same pass for long rows (B >= 64): work-group row r = (a, j) is uniform,
// so its taps are read once from the table (scalar loads); lanes cover b.
template <typename T>
__global__ __launch_bounds__(RS_BLOCK) void resize_rows(const T *__restrict__ in, T *__restrict__ out, int nbx,
                                                        long long B, int nin, int nout, int ntap,
                                                        const T *__restrict__ tw, const int *__restrict__ ti,
                                                        int accumulate) {
  const unsigned blk = rs_xcd_block();
  const int row = (int)(blk / (unsigned)nbx);
  const long long b = (long long)(blk - (unsigned)row * (unsigned)nbx) * RS_BLOCK + threadIdx.x;
  if (b >= B) return;
  const int j = row % nout;
  const long long o = row / nout;
  const T *src = in + o * nin * B + b;
  T *dst = out + (long long)row * B + b;
  T acc = accumulate ? *dst : T(0);
  for (int t = 0; t < ntap; ++t) {
    const int i = ti[j * RS_MAXT + t];
    if (i >= 0) acc += src[(long long)i * B] * tw[j * RS_MAXT + t];
  }
  *dst = acc;
}

// Long rows with B % 4 == 0: four consecutive values per lane (16-byte
// loads and stores for float).
template <typename T>
__global__ __launch_bounds__(RS_BLOCK) void resize_rows4(const T *__restrict__ in, T *__restrict__ out, int nbx,
                                                         long long B, int nin, int nout, int ntap,
                                                         const T *__restrict__ tw, const int *__restrict__ ti,
                                                         int accumulate) {
  using V = HIP_vector_type<T, 4>;
  const unsigned blk = rs_xcd_block();
  const int row = (int)(blk / (unsigned)nbx);
  const long long b = ((long long)(blk - (unsigned)row * (unsigned)nbx) * RS_BLOCK + threadIdx.x) * 4;
  if (b >= B) return;
  const int j = row % nout;
  const long long o = row / nout;
  const T *src = in + o * nin * B + b;
  V *dst = (V *)(out + (long long)row * B + b);
  V acc = accumulate ? *dst : V(T(0));
  for (int t = 0; t < ntap; ++t) {
    const int i = ti[j * RS_MAXT + t];
    if (i >= 0) {
      const V v = *(const V *)(src + (long long)i * B);
      const T w = tw[j * RS_MAXT + t];
      acc.x += v.x * w; acc.y += v.y * w; acc.z += v.z * w; acc.w += v.w * w;
    }
  }
  *dst = acc;
}

// The contiguous axis (B == 1): a work-group stages `rows` whole input rows
// (one contiguous chunk) and the axis's tap table in LDS, then writes its
// rows x nout outputs as one contiguous chunk.
template <typename T>
__global__ __launch_bounds__(RS_BLOCK) void resize_contig(const T *__restrict__ in, T *__restrict__ out,
                                                          long long A, int rows, int nin, int nout, int ntap,
                                                          const T *__restrict__ tw, const int *__restrict__ ti,
                                                          int accumulate) {
  extern __shared__ char rs_smem[];
  T *lw = (T *)rs_smem;                                  // nout * RS_MAXT weights
  int *li = (int *)(lw + (size_t)nout * RS_MAXT);        // nout * RS_MAXT indices
  T *lx = (T *)(li + (size_t)nout * RS_MAXT + ((sizeof(T) == 8) ? ((nout * RS_MAXT) & 1) : 0));
  const long long r0 = (long long)blockIdx.x * rows;
  const int nr = (int)(A - r0 < rows ? A - r0 : rows);
  for (int k = threadIdx.x; k < nout * RS_MAXT; k += RS_BLOCK) { lw[k] = tw[k]; li[k] = ti[k]; }
  const T *src = in + r0 * nin;
  for (int k = threadIdx.x; k < nr * nin; k += RS_BLOCK) lx[k] = src[k];
  __syncthreads();
  T *dst = out + r0 * nout;
  for (int k = threadIdx.x; k < nr * nout; k += RS_BLOCK) {
    const int r = k / nout, j = k - r * nout;
    const T *x = lx + r * nin;
    T acc = accumulate ? dst[k] : T(0);
    for (int t = 0; t < ntap; ++t) {
      const int i = li[j * RS_MAXT + t];
      if (i >= 0) acc += x[i] * lw[j * RS_MAXT + t];
    }
    dst[k] = acc;
  }
}

// Short rows (B < 64) with fewer than 2^31 outputs: 32-bit index math.
template <typename T>
__global__ __launch_bounds__(RS_BLOCK) void resize_short(const T *__restrict__ in, T *__restrict__ out, int total,
                                                         int B, int nin, int nout, int ntap,
                                                         const T *__restrict__ tw, const int *__restrict__ ti,
                                                         int accumulate) {
  const int gid = blockIdx.x * RS_BLOCK + threadIdx.x;
  if (gid >= total) return;
  int b = 0, r = gid;
  if (B > 1) { b = gid % B; r = gid / B; }
  const int j = r % nout, o = r / nout;
  const T *src = in + ((long long)o * nin) * B + b;
  T acc = accumulate ? out[gid] : T(0);
  for (int t = 0; t < ntap; ++t) {
    const int i = ti[j * RS_MAXT + t];
    if (i >= 0) acc += src[(long long)i * B] * tw[j * RS_MAXT + t];
  }
  out[gid] = acc;
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

// Forward resize as separable passes x -> y -> z (axes of size 1 on both
// sides are skipped): each pass streams its input once with coalesced tap
// rows, instead of one lane gathering (O+1)^3 taps per channel from three
// strided axes.  Intermediates come from a stream-ordered allocation; false
// if it fails (the caller then uses the direct kernel).
template <typename T>
static bool rs_separable(const T *input, T *output, const RsArgs &a, hipStream_t st) {
  int axes[3], np = 0;
  for (int d = 0; d < 3; ++d)
    if (!(a.ns[d] == 1 && a.nt[d] == 1)) axes[np++] = d;
  if (np == 0) return false;
  const long long NC = (long long)a.N * a.C;
  long long cur[3] = {a.ns[0], a.ns[1], a.ns[2]};
  // sizes of the intermediates (after each pass but the last)
  long long tmp_elems[2] = {0, 0};
  {
    long long c[3] = {a.ns[0], a.ns[1], a.ns[2]};
    for (int p = 0; p + 1 < np; ++p) {
      c[axes[p]] = a.nt[axes[p]];
      tmp_elems[p % 2] = std::max(tmp_elems[p % 2], NC * c[0] * c[1] * c[2]);
    }
  }
  // buffer: the two intermediates, then per pass a tap table of nt[d] x
  // RS_MAXT weights and as many int indices
  long long tab_w[3], tab_i[3], tot = tmp_elems[0] + tmp_elems[1];
  for (int p = 0; p < np; ++p) {
    tab_w[p] = tot;
    tot += (long long)a.nt[axes[p]] * RS_MAXT;
    tab_i[p] = tot;
    tot += ((long long)a.nt[axes[p]] * RS_MAXT * sizeof(int) + sizeof(T) - 1) / sizeof(T);
  }
  T *buf = nullptr;
  if (scratch_alloc((void **)&buf, (size_t)tot * sizeof(T), st) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  T *tmp[2] = {buf, buf + tmp_elems[0]};
  for (int p = 0; p < np; ++p)
    hipLaunchKernelGGL(rs_table<T>, dim3(rs_blocks(a.nt[axes[p]])), dim3(RS_BLOCK), 0, st, a, axes[p],
                       buf + tab_w[p], (int *)(buf + tab_i[p]));
  const T *src = input;
  for (int p = 0; p < np; ++p) {
    const int d = axes[p];
    long long A = NC, B = 1;
    for (int e = 0; e < d; ++e) A *= cur[e];
    for (int e = d + 1; e < 3; ++e) B *= cur[e];
    const bool last = p + 1 == np;
    T *dst = last ? output : tmp[p % 2];
    const long long n = A * a.nt[d] * B;
    const int acc = last ? a.accumulate : 0;
    const long long nbx = (B + RS_BLOCK - 1) / RS_BLOCK;
    const T *tw = buf + tab_w[p];
    const int *ti = (const int *)(buf + tab_i[p]);
    const int ntap = a.order[d] + 1;
    const long long nbx4 = (B / 4 + RS_BLOCK - 1) / RS_BLOCK;
    // contiguous axis: rows of nin inputs staged in LDS (about 16 KB of them)
    const int crows = (int)std::max<long long>(1, std::min<long long>(A, 16384 / ((long long)a.ns[d] * sizeof(T))));
    const size_t clds = (size_t)a.nt[d] * RS_MAXT * (sizeof(T) + sizeof(int)) + sizeof(T) +
                        (size_t)crows * a.ns[d] * sizeof(T);
    // 4-wide accesses need both base pointers aligned to the vector (a
    // caller's view may start at any element)
    const bool al4 = ((uintptr_t)src % (4 * sizeof(T))) == 0 && ((uintptr_t)dst % (4 * sizeof(T))) == 0;
    if (B >= 64 && B % 4 == 0 && al4 && A * a.nt[d] * nbx4 < (1LL << 31))
      hipLaunchKernelGGL(resize_rows4<T>, dim3((unsigned)(A * a.nt[d] * nbx4)), dim3(RS_BLOCK), 0, st, src, dst,
                         (int)nbx4, B, a.ns[d], a.nt[d], ntap, tw, ti, acc);
    else if (B >= 64 && A * a.nt[d] * nbx < (1LL << 31))
      hipLaunchKernelGGL(resize_rows<T>, dim3((unsigned)(A * a.nt[d] * nbx)), dim3(RS_BLOCK), 0, st, src, dst,
                         (int)nbx, B, a.ns[d], a.nt[d], ntap, tw, ti, acc);
    else if (B == 1 && clds <= 32768 && (A + crows - 1) / crows < (1LL << 31) &&
             (long long)crows * std::max(a.ns[d], a.nt[d]) < (1LL << 30))
      hipLaunchKernelGGL(resize_contig<T>, dim3((unsigned)((A + crows - 1) / crows)), dim3(RS_BLOCK), clds, st,
                         src, dst, A, crows, a.ns[d], a.nt[d], ntap, tw, ti, acc);
    else if (n < (1LL << 31) - RS_BLOCK)
      hipLaunchKernelGGL(resize_short<T>, dim3(rs_blocks(n)), dim3(RS_BLOCK), 0, st, src, dst, (int)n, (int)B,
                         a.ns[d], a.nt[d], ntap, tw, ti, acc);
    else
      hipLaunchKernelGGL(resize_axis<T>, dim3(rs_blocks(n)), dim3(RS_BLOCK), 0, st, src, dst, A, B, a, d, acc);
    cur[d] = a.nt[d];
    src = dst;
  }
  (void)hipFreeAsync(buf, st);
  return true;
}

template <typename T>
static void rs_launch(const T *input, T *output, const RsArgs &a, bool adjoint, hipStream_t st) {
  const long long Vt = (long long)a.nt[0] * a.nt[1] * a.nt[2];
  const long long nthreads = (long long)a.N * Vt;
  if (adjoint) {
    const long long ns = (long long)a.N * a.C * a.ns[0] * a.ns[1] * a.ns[2];
    hipLaunchKernelGGL(rs_fill<T>, dim3(rs_blocks(ns)), dim3(RS_BLOCK), 0, st, output, ns, T(0));
    hipLaunchKernelGGL((resize_kernel<T, true>), dim3(rs_blocks(nthreads)), dim3(RS_BLOCK), 0, st, input, output, a);
  } else if (!rs_separable<T>(input, output, a, st)) {
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
