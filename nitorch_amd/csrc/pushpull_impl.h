// Push/pull kernels shared by the pushpull translation units (see pushpull.hip
// for the semantics and design notes).
#pragma once
#include <cstdio>
#include <cstring>

#include "common.h"

namespace ni_amd {

enum Flag : int {
  F_PULL = 1, F_PUSH = 2, F_COUNT = 4, F_GRAD = 8, F_SGRAD = 16,
  F_TRGT = 32,   // a target tensor is given (push values / grad weighting)
  F_KTRGT = 64,  // target carries a trailing spatial-gradient dim (trgt_K = dim)
};

template <typename T>
struct PPArgs {
  const T *src;
  const T *trgt;
  const T *grid;
  T *out;
  T *gout;
  int64_t N, C;
  int64_t sX, sY, sZ;      // source spatial size (missing dims = 1)
  int64_t tX, tY, tZ;      // target spatial size
  int64_t src_s[5];        // N C X Y Z
  int64_t trgt_s[6];       // N C X Y Z K
  int64_t grid_s[5];       // N X Y Z D
  int64_t out_s[6];        // N C X Y Z K
  int64_t gout_s[5];       // N X Y Z D
  int bound[3];
  int order[3];
  int dim;
  int extrap;
  int flags;
  // caller-provided scratch for the tile path's list of scattered tiles
  // (ints, capacity `scratch_cap`); null: the tile path allocates it on
  // the stream
  int *scratch;
  int64_t scratch_cap;
};

#define NI_BLOCK 256
#define NI_CB 8           // channel register block
#define NI_NXCD 8         // XCDs per MI355X

// Map the hardware block id (dispatched round-robin over the XCDs) to a
// logical block id such that every XCD owns one contiguous range.
__device__ inline int64_t xcd_block(int64_t nblocks) {
  int64_t b = blockIdx.x;
  int64_t xcd = b % NI_NXCD, local = b / NI_NXCD;
  int64_t q = nblocks / NI_NXCD, r = nblocks % NI_NXCD;
  return xcd * q + (xcd < r ? xcd : r) + local;
}

template <typename T> __device__ inline T dev_floor(T x);
template <> __device__ inline float dev_floor<float>(float x) { return floorf(x); }
template <> __device__ inline double dev_floor<double>(double x) { return floor(x); }
template <typename T> __device__ inline T dev_round(T x);
template <> __device__ inline float dev_round<float>(float x) { return roundf(x); }
template <> __device__ inline double dev_round<double>(double x) { return round(x); }

__device__ inline void atomic_add(float *p, float v) { unsafeAtomicAdd(p, v); }
__device__ inline void atomic_add(double *p, double v) { unsafeAtomicAdd(p, v); }

// --------------------------------------------------------------------------
// Per-sample routine.
//   T    : float / double
//   I    : int32_t / int64_t offsets
//   ORD  : isotropic order 0..7 known at compile time, or -1 (runtime, any)
//   FL   : compile-time flags, or -1 (read a.flags)
//   DIM3 : true when the call is known to be 3-D
// --------------------------------------------------------------------------
template <typename T, typename I, int ORD, int FL, bool DIM3>
__device__ __forceinline__ void pp_sample(const PPArgs<T> &a, I n, I w, I h, I d) {
  constexpr int MAXT = ORD >= 0 ? ORD + 1 : 8;
  const int flags = FL >= 0 ? FL : a.flags;
#define HAS(f) ((FL >= 0) ? ((FL & (f)) != 0) : ((flags & (f)) != 0))
  const int dim = DIM3 ? 3 : a.dim;
  const I C = (I)a.C;

  // ---- coordinates ----
  const T *gp = a.grid + n * (I)a.grid_s[0] + w * (I)a.grid_s[1] + h * (I)a.grid_s[2] + d * (I)a.grid_s[3];
  const I gD = (I)a.grid_s[4];
  T xs[3];
  xs[0] = gp[0];
  xs[1] = dim > 1 ? gp[gD] : T(0);
  xs[2] = dim > 2 ? gp[2 * gD] : T(0);
  const I ssz[3] = {(I)a.sX, (I)a.sY, (I)a.sZ};

  // output pointers for this sample (pull / sgrad: target-sized)
  const I obase = n * (I)a.out_s[0] + w * (I)a.out_s[2] + h * (I)a.out_s[3] + d * (I)a.out_s[4];
  T *gop = HAS(F_GRAD) ? a.gout + n * (I)a.gout_s[0] + w * (I)a.gout_s[1] + h * (I)a.gout_s[2] + d * (I)a.gout_s[3] : nullptr;
  const I goD = (I)a.gout_s[4];

  // ---- out-of-bound rule (check1d/2d/3d) ----
  if (!(a.extrap & 1)) {
    const T tol = (a.extrap & 2) ? T(0.55f) : T(0.05f);
    bool inb = true;
#pragma unroll
    for (int k = 0; k < 3; ++k)
      if (k < dim) inb = inb && (xs[k] >= -tol && xs[k] < T(ssz[k] - 1) + tol);
    if (!inb) {
      if (HAS(F_PULL) || HAS(F_SGRAD)) {
        const int nk = HAS(F_SGRAD) ? dim : 1;
        for (I c = 0; c < C; ++c)
          for (int k = 0; k < nk; ++k) a.out[obase + c * (I)a.out_s[1] + k * (I)a.out_s[5]] = T(0);
      }
      if (HAS(F_GRAD))
        for (int k = 0; k < dim; ++k) gop[k * goD] = T(0);
      return;
    }
  }

  const bool iso = (a.order[0] == a.order[1]) && (a.order[0] == a.order[2]);
  const int ord0 = ORD >= 0 ? ORD : a.order[0];

  // strides of source (gather) and output (scatter) along the 3 axes
  const I src_sd[3] = {(I)a.src_s[2], (I)a.src_s[3], (I)a.src_s[4]};
  const I out_sd[3] = {(I)a.out_s[2], (I)a.out_s[3], (I)a.out_s[4]};

  // ================== nearest neighbour (isotropic order 0) ==================
  if ((ORD == 0) || (ORD < 0 && iso && ord0 == 0)) {
    I oi = 0, oo = 0;
    int s = 1;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (k >= dim) continue;
      int sk;
      I j = fold<I>(a.bound[k], (I)dev_round<T>(xs[k]), ssz[k], sk);
      s *= sk;
      oi += j * src_sd[k];
      oo += j * out_sd[k];
    }
    if (HAS(F_PULL)) {
      const T *sp = a.src + n * (I)a.src_s[0] + oi;
      for (I c = 0; c < C; ++c) {
        T v = s ? sp[c * (I)a.src_s[1]] : T(0);
        a.out[obase + c * (I)a.out_s[1]] = s < 0 ? -v : v;
      }
    } else if (HAS(F_SGRAD)) {
      for (I c = 0; c < C; ++c)
        for (int k = 0; k < dim; ++k) a.out[obase + c * (I)a.out_s[1] + k * (I)a.out_s[5]] = T(0);
    } else if (HAS(F_PUSH) && !HAS(F_KTRGT)) {
      if (s) {
        const T *tp = a.trgt + n * (I)a.trgt_s[0] + w * (I)a.trgt_s[2] + h * (I)a.trgt_s[3] + d * (I)a.trgt_s[4];
        T *op = a.out + n * (I)a.out_s[0] + oo;
        for (I c = 0; c < C; ++c) {
          T v = tp[c * (I)a.trgt_s[1]];
          atomic_add(op + c * (I)a.out_s[1], s < 0 ? -v : v);
        }
      }
    } else if (HAS(F_COUNT)) {
      if (s) atomic_add(a.out + n * (I)a.out_s[0] + oo, T(s));
    }
    if (HAS(F_GRAD))
      for (int k = 0; k < dim; ++k) gop[k * goD] = T(0);
    return;
  }

  // ================== tap tables ==================
  const bool need_g = HAS(F_GRAD) || HAS(F_SGRAD) || (HAS(F_PUSH) && HAS(F_KTRGT));
  const bool need_h = HAS(F_GRAD) && HAS(F_KTRGT);
  const bool gather = HAS(F_PULL) || HAS(F_SGRAD) || HAS(F_GRAD);
  const bool scatter = HAS(F_PUSH) || HAS(F_COUNT);
  const bool lin = (ORD == 1) || (ORD < 0 && iso && ord0 == 1);

  T tw[3][MAXT], tg[3][MAXT], th[3][MAXT];
  I ti[3][MAXT];       // gather offsets (source strides)
  I to[3][MAXT];       // scatter offsets (output strides)
  int ts[3][MAXT];     // signs
  int nt[3];

#pragma unroll
  for (int k = 0; k < 3; ++k) {
    if (k >= dim) {
      nt[k] = 1;
      tw[k][0] = T(1); tg[k][0] = T(0); th[k][0] = T(0);
      ti[k][0] = 0; to[k][0] = 0; ts[k][0] = 1;
      continue;
    }
    const T x = xs[k];
    // reference quirk (pushpull_common.cpp:591) deliberately not reproduced:
    // each axis uses its own order.
    const int ok = ORD >= 0 ? ORD : a.order[k];
    nt[k] = ok + 1;
    if (lin) {
      // linear fast path weights (pushpull_common.cpp:1658-1667)
      // (weights as distances to the far tap: exact when small)
      const I j0 = (I)dev_floor<T>(x);
      tw[k][0] = T(j0 + 1) - x; tw[k][1] = x - T(j0);
      tg[k][0] = T(-1);     tg[k][1] = T(1);
      th[k][0] = T(0);      th[k][1] = T(0);
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        int s;
        I j = fold<I>(a.bound[k], j0 + t, ssz[k], s);
        ts[k][t] = s;
        ti[k][t] = j * src_sd[k];
        to[k][t] = j * out_sd[k];
      }
    } else {
      const I lo = (ok == 0) ? (I)dev_round<T>(x) : (I)dev_floor<T>(x - first_tap_offset<T>(ok));
#pragma unroll
      for (int t = 0; t < MAXT; ++t) {
        if (ORD < 0 && t > ok) break;
        const I b = lo + t;
        const T dd = x - T(b);
        if (ok >= 1 && (t == 0 || t == ok)) {   // last piece, from x itself
          const T s = spl_outer_s<T, I>(ok, t, x, lo);
          T gw, hw;
          if constexpr (ORD >= 1) spl_outer<T, (ORD >= 1 ? ORD : 1)>(t == 0, s, tw[k][t], gw, hw);
          else spl_outer_rt<T>(ok, t == 0, s, tw[k][t], gw, hw);
          tg[k][t] = gw;
          th[k][t] = hw;
        } else if constexpr (ORD >= 0) {
          tw[k][t] = spl_w<T, (ORD >= 0 ? ORD : 0)>(t, dd);
          if (need_g) tg[k][t] = spl_g<T, (ORD >= 0 ? ORD : 0)>(t, dd);
          if (need_h) th[k][t] = spl_h<T, (ORD >= 0 ? ORD : 0)>(t, dd);
        } else {
          spl_eval<T>(ok, t, dd, tw[k][t], tg[k][t], th[k][t]);
        }
        int s;
        I j = fold<I>(a.bound[k], b, ssz[k], s);
        ts[k][t] = s;
        ti[k][t] = j * src_sd[k];
        to[k][t] = j * out_sd[k];
      }
    }
  }
  (void)gather; (void)scatter;

  const T *srcn = a.src ? a.src + n * (I)a.src_s[0] : nullptr;
  const I src_c = (I)a.src_s[1];
  T *outn = a.out ? a.out + n * (I)a.out_s[0] : nullptr;
  const I out_c = (I)a.out_s[1], out_k = (I)a.out_s[5];
  const T *tp = (HAS(F_TRGT) && a.trgt) ? a.trgt + n * (I)a.trgt_s[0] + w * (I)a.trgt_s[2] + h * (I)a.trgt_s[3] + d * (I)a.trgt_s[4] : nullptr;
  const I trgt_c = (I)a.trgt_s[1], trgt_k = (I)a.trgt_s[5];

  T og0 = 0, og1 = 0, og2 = 0;   // grid gradient accumulators

  // ---------------- count: no channels ----------------
  if (HAS(F_COUNT) && !HAS(F_PUSH)) {
    for (int kz = 0; kz < nt[2]; ++kz)
      for (int ky = 0; ky < nt[1]; ++ky) {
        const T wyz = tw[1][ky];
#pragma unroll
        for (int kx = 0; kx < MAXT; ++kx) {
          if (kx >= nt[0]) break;
          const int s = ts[0][kx] * ts[1][ky] * ts[2][kz];
          if (s) {
            const T ww = (tw[0][kx] * wyz) * tw[2][kz];
            atomic_add(outn + to[0][kx] + to[1][ky] + to[2][kz], s < 0 ? -ww : ww);
          }
        }
      }
  }

  // ---------------- channel-blocked main loop ----------------
  const bool do_main = HAS(F_PULL) || HAS(F_SGRAD) || HAS(F_PUSH) || HAS(F_GRAD);
  if (do_main) {
    for (I c0 = 0; c0 < C; c0 += NI_CB) {
      const int ncb = (C - c0) < NI_CB ? (int)(C - c0) : NI_CB;
      T acc[NI_CB], acy[NI_CB], acz[NI_CB];
      T tv0[NI_CB], tv1[NI_CB], tv2[NI_CB];
#pragma unroll
      for (int c = 0; c < NI_CB; ++c) {
        acc[c] = acy[c] = acz[c] = T(0);
        tv0[c] = tv1[c] = tv2[c] = T(0);
        if (c < ncb && tp) {
          const T *tpc = tp + (c0 + c) * trgt_c;
          tv0[c] = tpc[0];
          if (HAS(F_KTRGT)) {
            if (dim > 1) tv1[c] = tpc[trgt_k];
            if (dim > 2) tv2[c] = tpc[2 * trgt_k];
          }
        }
      }
      const T *srcc = srcn ? srcn + c0 * src_c : nullptr;
      T *outc = outn ? outn + c0 * out_c : nullptr;

      for (int kz = 0; kz < nt[2]; ++kz) {
        for (int ky = 0; ky < nt[1]; ++ky) {
#pragma unroll
          for (int kx = 0; kx < MAXT; ++kx) {
            if (kx >= nt[0]) break;
            const int s = ts[0][kx] * ts[1][ky] * ts[2][kz];
            if (s == 0) continue;
            const T sg = s < 0 ? T(-1) : T(1);
            const T wx = tw[0][kx], wy = tw[1][ky], wz = tw[2][kz];
            const T w3 = ((wx * wy) * wz) * sg;
            T gx = 0, gy = 0, gz = 0;
            if (need_g) {
              gx = ((tg[0][kx] * wy) * wz) * sg;
              gy = ((wx * tg[1][ky]) * wz) * sg;
              gz = ((wx * wy) * tg[2][kz]) * sg;
            }
            const I oi = ti[0][kx] + ti[1][ky] + ti[2][kz];
            const I oo = to[0][kx] + to[1][ky] + to[2][kz];

            if (HAS(F_PULL) || HAS(F_SGRAD) || HAS(F_GRAD)) {
              T dot0 = 0, dot1 = 0, dot2 = 0;
#pragma unroll
              for (int c = 0; c < NI_CB; ++c) {
                if (c >= ncb) break;
                const T v = srcc[c * src_c + oi];
                if (HAS(F_PULL)) acc[c] += v * w3;
                if (HAS(F_SGRAD)) {
                  acc[c] += v * gx;
                  acy[c] += v * gy;
                  acz[c] += v * gz;
                }
                if (HAS(F_GRAD)) {
                  if (HAS(F_KTRGT)) {
                    dot0 += v * tv0[c]; dot1 += v * tv1[c]; dot2 += v * tv2[c];
                  } else if (HAS(F_TRGT)) {
                    dot0 += v * tv0[c];
                  } else {
                    dot0 += v;
                  }
                }
              }
              if (HAS(F_GRAD)) {
                if (HAS(F_KTRGT)) {
                  const T hx = th[0][kx], hy = th[1][ky], hz = th[2][kz];
                  const T gxx = tg[0][kx], gyy = tg[1][ky], gzz = tg[2][kz];
                  og0 += (((hx * wy) * wz) * dot0 + ((gxx * gyy) * wz) * dot1 + ((gxx * wy) * gzz) * dot2) * sg;
                  og1 += (((gxx * gyy) * wz) * dot0 + ((wx * hy) * wz) * dot1 + ((wx * gyy) * gzz) * dot2) * sg;
                  og2 += (((gxx * wy) * gzz) * dot0 + ((wx * gyy) * gzz) * dot1 + ((wx * wy) * hz) * dot2) * sg;
                } else {
                  og0 += gx * dot0;
                  og1 += gy * dot0;
                  og2 += gz * dot0;
                }
              }
            }
            if (HAS(F_PUSH)) {
#pragma unroll
              for (int c = 0; c < NI_CB; ++c) {
                if (c >= ncb) break;
                T v;
                if (HAS(F_KTRGT)) v = gx * tv0[c] + gy * tv1[c] + gz * tv2[c];
                else v = w3 * tv0[c];
                atomic_add(outc + c * out_c + oo, v);
              }
            }
          }
        }
      }
      // ---- write gathered outputs for this channel block ----
      if (HAS(F_PULL)) {
#pragma unroll
        for (int c = 0; c < NI_CB; ++c)
          if (c < ncb) a.out[obase + (c0 + c) * out_c] = acc[c];
      } else if (HAS(F_SGRAD)) {
#pragma unroll
        for (int c = 0; c < NI_CB; ++c)
          if (c < ncb) {
            T *o = a.out + obase + (c0 + c) * out_c;
            o[0] = acc[c];
            if (dim > 1) o[out_k] = acy[c];
            if (dim > 2) o[2 * out_k] = acz[c];
          }
      }
    }
  }
  if (HAS(F_GRAD)) {
    gop[0] = og0;
    if (dim > 1) gop[goD] = og1;
    if (dim > 2) gop[2 * goD] = og2;
  }
#undef HAS
}

template <typename T, typename I, int ORD, int FL, bool DIM3>
__global__ __launch_bounds__(NI_BLOCK) void pp_kernel(PPArgs<T> a, int64_t total, int64_t nblocks) {
  const int64_t lb = xcd_block(nblocks);
  const I gid = (I)(lb * NI_BLOCK + threadIdx.x);
  if ((int64_t)gid >= total) return;
  const I tZ = (I)a.tZ, tY = (I)a.tY, tX = (I)a.tX;
  I r = gid;
  const I d = r % tZ; r /= tZ;
  const I h = r % tY; r /= tY;
  const I w = r % tX;
  const I n = r / tX;
  pp_sample<T, I, ORD, FL, DIM3>(a, n, w, h, d);
}

// zero-fill of a contiguous output (push / count)
template <typename T>
__global__ __launch_bounds__(NI_BLOCK) void fill_zero(T *p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * NI_BLOCK + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * NI_BLOCK;
  for (; i < n; i += stride) p[i] = T(0);
}

template <typename T>
static int zero_fill(T *p, int64_t n, hipStream_t st) {
  if (n <= 0) return 0;
  int64_t blocks = (n + NI_BLOCK - 1) / NI_BLOCK;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(fill_zero<T>, dim3((unsigned)blocks), dim3(NI_BLOCK), 0, st, p, n);
  NI_HIP_CHECK(hipGetLastError());
  return 0;
}


// ---------------- per-order dispatch (pushpull_o*.hip) ----------------
template <typename T, typename I, int ORD, int FL>
static void launch_fast(const PPArgs<T> &a, int64_t total, hipStream_t st) {
  int64_t nb = (total + NI_BLOCK - 1) / NI_BLOCK;
  hipLaunchKernelGGL((pp_kernel<T, I, ORD, FL, true>), dim3((unsigned)nb), dim3(NI_BLOCK), 0, st, a, total, nb);
}

// Launch the order-ORD 3-D float kernel for a.flags; false if that mode has
// no specialisation (the caller then uses the generic kernel).
template <int ORD> bool dispatch_order(const PPArgs<float> &a, int64_t total, hipStream_t st);
template <> bool dispatch_order<0>(const PPArgs<float> &, int64_t, hipStream_t);
template <> bool dispatch_order<1>(const PPArgs<float> &, int64_t, hipStream_t);
template <> bool dispatch_order<2>(const PPArgs<float> &, int64_t, hipStream_t);
template <> bool dispatch_order<3>(const PPArgs<float> &, int64_t, hipStream_t);
template <> bool dispatch_order<4>(const PPArgs<float> &, int64_t, hipStream_t);
template <> bool dispatch_order<5>(const PPArgs<float> &, int64_t, hipStream_t);
template <> bool dispatch_order<6>(const PPArgs<float> &, int64_t, hipStream_t);
template <> bool dispatch_order<7>(const PPArgs<float> &, int64_t, hipStream_t);

// Gathers, scatters and their combinations run on the LDS-tiled kernels
// (pushpull_tile.h); modes with a spatial-gradient target (trgt_K = dim)
// keep the direct per-sample kernel.  Plain push and count of order 0 (one
// tap a sample) also run the direct kernel: C atomics a sample straight to
// L2 cost less than staging a box in LDS, adding into it and flushing it
// (batch 4, 8 channels, 192^3: push 0.88 vs 1.29 ms, count 0.16 vs 0.26 ms).  At
// order 1 the 8 taps a sample make the direct atomics 5x slower than the
// tiles.
template <int ORD> constexpr bool direct_scatter() { return ORD == 0; }

#define NI_DEFINE_DISPATCH_ORDER(ORD)                                            \
  template <> bool dispatch_order<ORD>(const PPArgs<float> &a, int64_t total, hipStream_t st) { \
    switch (a.flags) {                                                          \
      case F_PULL: launch_tile<ORD, TM_PULL>(a, st); return true;               \
      case F_SGRAD: launch_tile<ORD, TM_SGRAD>(a, st); return true;             \
      case F_PUSH | F_TRGT:                                                     \
        if constexpr (direct_scatter<ORD>())                                    \
          launch_fast<float, int32_t, ORD, F_PUSH | F_TRGT>(a, total, st);      \
        else                                                                    \
          launch_tile<ORD, TM_PUSH>(a, st);                                     \
        return true;                                                            \
      case F_COUNT:                                                             \
        if constexpr (direct_scatter<ORD>())                                    \
          launch_fast<float, int32_t, ORD, F_COUNT>(a, total, st);              \
        else                                                                    \
          launch_tile<ORD, TM_COUNT>(a, st);                                    \
        return true;                                                            \
      case F_PUSH | F_GRAD | F_TRGT:                                            \
        launch_tile<ORD, TM_PUSH>(a, st); launch_tile<ORD, TM_GGRAD>(a, st); return true; \
      case F_GRAD | F_TRGT: launch_tile<ORD, TM_GGRAD>(a, st); return true;     \
      case F_PULL | F_GRAD | F_TRGT:                                            \
        launch_tile<ORD, TM_PULL>(a, st); launch_tile<ORD, TM_GGRAD>(a, st); return true; \
      case F_GRAD: launch_tile<ORD, TM_GGRAD1>(a, st); return true;             \
      default: return false;                                                    \
    }                                                                           \
  }

}  // namespace ni_amd
