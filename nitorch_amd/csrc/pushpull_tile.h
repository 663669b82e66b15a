// LDS-tiled pull / push / count for 3-D float fields (isotropic order ORD).
//
// One work-group owns a TL_X x TL_Y x TL_Z tile of samples (grid points).
// It reduces the bounding box of the first taps of its in-bound samples,
// and when the box (extended by the ORD + 1 taps) fits in LDS:
//  * pull  : stages the box of the source -- folded by the boundary
//            condition and signed, so the taps need no folding -- in LDS
//            with coalesced loads, then gathers every tap from LDS;
//  * push / count : accumulates the tile's splats into an LDS box with
//            integer LDS atomics on a fixed-point scale chosen per tile and
//            channel (2^k with  n_in * max|value| * 2^k < 2^30, so no sum can
//            overflow; LDS float atomics are far slower than integer ones on
//            gfx950), then adds the box to the output once, folded and
//            signed, with one FP32 global atomic per non-zero box element.
//            Rounding: <= 2^-(21 - log2 n_in) * max|value| per splat, and the
//            in-tile sum is exact and order-independent.
// The box has fixed dimensions per order (every tap offset is then an
// immediate of the LDS instruction); channels are staged NCC at a time.
// A tile whose box does not fit (scattered or far out-of-range
// coordinates with extrapolation) is left to a second kernel that runs the
// direct per-sample routine pp_sample for its samples, so the result never
// depends on the data's smoothness, only the speed does.
//
// Weights, tap order (z outer, x inner) and the out-of-bound rule are those
// of pp_sample (nitorch/_C/src/impl/pushpull_common.cpp: OOB :926-974,
// interpolate3d / splat3d :1176-1290, linear :1653-1858, nearest
// :2314-2341); folding and signs follow bounds_common.h:28-219.
#pragma once
#include <climits>

#include "pushpull_impl.h"

namespace ni_amd {

#define TL_X 4
#define TL_Y 8
#define TL_Z 16
#define TL_SAMPLES (TL_X * TL_Y * TL_Z)
#define TL_PER_THREAD (TL_SAMPLES / NI_BLOCK)
#define TL_SLACK 2     // extra box extent per axis for the spread of a tile's coordinates

// Box geometry of order ORD, fixed at compile time so that every tap offset
// is an immediate of the LDS instruction: the box covers the tile plus the
// ORD + 1 taps plus TL_SLACK voxels per axis.
template <int ORD>
struct TileBox {
  static constexpr int DX = TL_X + ORD + TL_SLACK;
  static constexpr int DY = TL_Y + ORD + TL_SLACK;
  static constexpr int DZ = TL_Z + ORD + TL_SLACK;
  static constexpr int DYZ = DY * DZ;
  static constexpr int VOL = DX * DYZ;
  static constexpr int MAX_NCC = ORD <= 3 ? 4 : 2;   // channels staged at once (LDS budget)
};

// Gather modes stage the source box; scatter modes accumulate one.
//   TM_PULL  : out[c]      = sum_t w(t) src(t, c)                   (F_PULL)
//   TM_SGRAD : out[c][d]   = sum_t dw_d(t) src(t, c)                (F_SGRAD)
//   TM_GGRAD : gout[d]     = sum_t dw_d(t) sum_c src(t, c) trgt(c)  (F_GRAD | F_TRGT,
//              or F_GRAD with a single channel and no target: trgt = 1)
//   TM_PUSH  : out(t, c)  += w(t) trgt(c)                           (F_PUSH | F_TRGT)
//   TM_COUNT : out(t)     += w(t)                                   (F_COUNT)
enum TileMode { TM_PULL = 0, TM_PUSH = 1, TM_COUNT = 2, TM_SGRAD = 3, TM_GGRAD = 4, TM_GGRAD1 = 5 };

template <int MODE>
struct TileMeta {
  static constexpr bool gather = MODE == TM_PULL || MODE == TM_SGRAD || MODE == TM_GGRAD || MODE == TM_GGRAD1;
  static constexpr bool deriv = MODE == TM_SGRAD || MODE == TM_GGRAD || MODE == TM_GGRAD1;
};

// First tap index and the ORD + 1 weights (and, when DERIV, first
// derivatives) of coordinate x, as pp_sample computes them.
template <int ORD, bool DERIV = false>
__device__ __forceinline__ int tl_taps(float x, float (&w)[ORD + 1], float (&g)[ORD + 1]) {
  if constexpr (ORD == 0) {
    w[0] = 1.f;
    g[0] = 0.f;
    return (int)roundf(x);
  } else if constexpr (ORD == 1) {
    const int j0 = (int)floorf(x);
    const float d1 = x - (float)j0;
    w[0] = 1.f - d1;
    w[1] = d1;
    g[0] = -1.f;
    g[1] = 1.f;
    return j0;
  } else {
    const int lo = (int)floorf(x - first_tap_offset<float>(ORD));
#pragma unroll
    for (int t = 0; t <= ORD; ++t) {
      w[t] = spl_w<float, ORD>(t, x - (float)(lo + t));
      if (DERIV) g[t] = spl_g<float, ORD>(t, x - (float)(lo + t));
    }
    return lo;
  }
}

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}

template <int MODE>
struct TileFlags {
  static constexpr int value = MODE == TM_PULL ? F_PULL : MODE == TM_PUSH ? (F_PUSH | F_TRGT)
                               : MODE == TM_COUNT ? F_COUNT : MODE == TM_SGRAD ? F_SGRAD
                               : MODE == TM_GGRAD ? (F_GRAD | F_TRGT) : F_GRAD;
};

// NCC: channels per chunk (compile time, so that the channel loop of a tap
// is branch-free; lanes of a partial last chunk compute unused channels).
template <int ORD, int MODE, int NCC>
__global__ __launch_bounds__(NI_BLOCK) void tile_kernel(PPArgs<float> a, int ntx, int nty, int ntz,
                                                        int64_t nblocks) {
  constexpr int NT = ORD + 1;
  using B = TileBox<ORD>;
  constexpr int DY = B::DY, DZ = B::DZ, DYZ = B::DYZ, VOL = B::VOL;
  static_assert(NCC >= 1 && NCC <= B::MAX_NCC, "channel chunk exceeds the LDS budget");
  __shared__ float box[NCC * VOL];
  __shared__ int red[NI_BLOCK / 64][6];

  int64_t r = xcd_block(nblocks);
  const int tz = (int)(r % ntz); r /= ntz;
  const int ty = (int)(r % nty); r /= nty;
  const int tx = (int)(r % ntx);
  const int n = (int)(r / ntx);
  const int X = (int)a.tX, Y = (int)a.tY, Z = (int)a.tZ;
  const int sz[3] = {(int)a.sX, (int)a.sY, (int)a.sZ};

  // ---- samples of this thread: coordinates, in-bound flag, taps ----
  int px[TL_PER_THREAD], py[TL_PER_THREAD], pz[TL_PER_THREAD];
  bool valid[TL_PER_THREAD], inb[TL_PER_THREAD];
  constexpr bool GATHER = TileMeta<MODE>::gather, DERIV = TileMeta<MODE>::deriv;
  constexpr int NG = DERIV ? NT : 1;
  int lo[TL_PER_THREAD][3];
  float w[TL_PER_THREAD][3][NT];
  float gw[TL_PER_THREAD][3][NG];   // derivative weights (gradient modes)
  int mn[3] = {INT_MAX, INT_MAX, INT_MAX}, mx[3] = {INT_MIN, INT_MIN, INT_MIN};
  const float tol = (a.extrap & 2) ? 0.55f : 0.05f;
#pragma unroll
  for (int k = 0; k < TL_PER_THREAD; ++k) {
    const int s = threadIdx.x + k * NI_BLOCK;
    px[k] = tx * TL_X + s / (TL_Y * TL_Z);
    py[k] = ty * TL_Y + (s / TL_Z) % TL_Y;
    pz[k] = tz * TL_Z + s % TL_Z;
    valid[k] = px[k] < X && py[k] < Y && pz[k] < Z;
    inb[k] = false;
    if (valid[k]) {
      const float *gp = a.grid + (int)(n * a.grid_s[0] + px[k] * a.grid_s[1] + py[k] * a.grid_s[2] + pz[k] * a.grid_s[3]);
      const int gD = (int)a.grid_s[4];
      const float xs[3] = {gp[0], gp[gD], gp[2 * gD]};
      bool ok = true;
      if (!(a.extrap & 1))
#pragma unroll
        for (int d = 0; d < 3; ++d) ok = ok && (xs[d] >= -tol && xs[d] < (float)(sz[d] - 1) + tol);
      inb[k] = ok;
      if (ok) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          float gtmp[NT];
          lo[k][d] = tl_taps<ORD, DERIV>(xs[d], w[k][d], gtmp);
#pragma unroll
          for (int t = 0; t < NG; ++t) gw[k][d][t] = gtmp[t];
          mn[d] = min(mn[d], lo[k][d]);
          mx[d] = max(mx[d], lo[k][d]);
        }
      }
    }
  }

  // ---- bounding box of the work-group's first taps ----
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    mn[d] = wave_min(mn[d]);
    mx[d] = wave_max(mx[d]);
  }
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      red[threadIdx.x >> 6][d] = mn[d];
      red[threadIdx.x >> 6][3 + d] = mx[d];
    }
  }
  __syncthreads();
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    mn[d] = red[0][d];
    mx[d] = red[0][3 + d];
#pragma unroll
    for (int wv = 1; wv < NI_BLOCK / 64; ++wv) {
      mn[d] = min(mn[d], red[wv][d]);
      mx[d] = max(mx[d], red[wv][3 + d]);
    }
  }

  const int oC = (int)a.out_s[1], oK = (int)a.out_s[5];
  if (mn[0] > mx[0]) {   // no in-bound sample in this tile
    if (MODE == TM_PULL || MODE == TM_SGRAD) {
      const int C = (int)a.C, nk = MODE == TM_SGRAD ? 3 : 1;
#pragma unroll
      for (int k = 0; k < TL_PER_THREAD; ++k)
        if (valid[k]) {
          float *op = a.out + (int)(n * a.out_s[0] + px[k] * a.out_s[2] + py[k] * a.out_s[3] + pz[k] * a.out_s[4]);
          for (int c = 0; c < C; ++c)
            for (int d = 0; d < nk; ++d) op[c * oC + d * oK] = 0.f;
        }
    } else if (MODE == TM_GGRAD || MODE == TM_GGRAD1) {
#pragma unroll
      for (int k = 0; k < TL_PER_THREAD; ++k)
        if (valid[k]) {
          float *gp = a.gout + (int)(n * a.gout_s[0] + px[k] * a.gout_s[1] + py[k] * a.gout_s[2] + pz[k] * a.gout_s[3]);
          for (int d = 0; d < 3; ++d) gp[d * a.gout_s[4]] = 0.f;
        }
    }
    return;
  }
  int ext[3];
  bool fits = true;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int64_t e = (int64_t)mx[d] - (int64_t)mn[d] + NT;
    fits = fits && e <= (d == 0 ? B::DX : d == 1 ? DY : DZ);
    ext[d] = (int)(e < 0x7fffffff ? e : 0x7fffffff);
  }
  if (!fits) return;   // scattered tile: left to tile_fallback
  int off[TL_PER_THREAD];
#pragma unroll
  for (int k = 0; k < TL_PER_THREAD; ++k)
    off[k] = inb[k] ? ((lo[k][0] - mn[0]) * DY + (lo[k][1] - mn[1])) * DZ + (lo[k][2] - mn[2]) : 0;

  const int C = MODE == TM_COUNT ? 1 : (int)a.C;
  // scatter: number of in-bound samples (bounds every fixed-point box sum)
  int n_in = 0;
  if (!GATHER) {
#pragma unroll
    for (int k = 0; k < TL_PER_THREAD; ++k) n_in += __syncthreads_count(inb[k]);
  }
  __shared__ float s_red[NI_BLOCK / 64][NCC];
  __shared__ float s_scale[NCC], s_inv[NCC];

  // folded offsets (in elements of the staged / scattered tensor) and signs
  // of the box positions along each axis: fold() runs once per position
  constexpr int NF = B::DX + DY + DZ;
  __shared__ int fofs[NF];
  __shared__ float fsg[NF];
  {
    const int64_t *st = GATHER ? a.src_s : a.out_s;
    const int st0 = (int)st[2], st1 = (int)st[3], st2 = (int)st[4];
    for (int i = threadIdx.x; i < NF; i += NI_BLOCK) {
      const int d = i < B::DX ? 0 : i < B::DX + DY ? 1 : 2;
      const int b = i - (d == 0 ? 0 : d == 1 ? B::DX : B::DX + DY);
      const int e = d == 0 ? ext[0] : d == 1 ? ext[1] : ext[2];
      const int m = d == 0 ? mn[0] : d == 1 ? mn[1] : mn[2];
      const int nn = d == 0 ? sz[0] : d == 1 ? sz[1] : sz[2];
      const int bd = d == 0 ? a.bound[0] : d == 1 ? a.bound[1] : a.bound[2];
      int sg = 0, j = 0;
      if (b < e) j = fold<int>(bd, m + b, nn, sg);
      fofs[i] = j * (d == 0 ? st0 : d == 1 ? st1 : st2);
      fsg[i] = (float)sg;
    }
  }
  __syncthreads();

  float og[TL_PER_THREAD][3];   // grid gradient (TM_GGRAD*), summed over chunks
#pragma unroll
  for (int k = 0; k < TL_PER_THREAD; ++k) og[k][0] = og[k][1] = og[k][2] = 0.f;

  for (int c0 = 0; c0 < C; c0 += NCC) {
    const int nc = (C - c0) < NCC ? (C - c0) : NCC;
    if (c0) __syncthreads();   // previous chunk fully consumed
    // keep the per-tap weight products inside the chunk loop (hoisting
    // them would hold (ORD + 1)^3 products per sample in registers)
#pragma unroll
    for (int k = 0; k < TL_PER_THREAD; ++k)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
#pragma unroll
        for (int t = 0; t < NT; ++t) asm volatile("" : "+v"(w[k][d][t]));
#pragma unroll
        for (int t = 0; t < NG; ++t) asm volatile("" : "+v"(gw[k][d][t]));
      }
    const int nbox = nc * VOL;

    if (GATHER) {
      // ---- stage the folded, signed source box ----
      const float *src = a.src + (int)(n * a.src_s[0] + c0 * a.src_s[1]);
      const int sC = (int)a.src_s[1];
      // only the ext[0] x ext[1] x ext[2] corner of the box is ever read:
      // enumerate it directly (floor divisions by float reciprocals, exact
      // after the +-1 fix-ups for indices below 2^22)
      const int E2 = ext[2], E12 = ext[1] * ext[2], E = ext[0] * E12;
      const float rE = 1.f / (float)E, r12 = 1.f / (float)E12, r2 = 1.f / (float)E2;
      const int nfill = nc * E;
#pragma unroll 4
      for (int i = threadIdx.x; i < nfill; i += NI_BLOCK) {
        int c = (int)((float)i * rE);
        c -= c * E > i;
        c += (c + 1) * E <= i;
        const int rr = i - c * E;
        int bx = (int)((float)rr * r12);
        bx -= bx * E12 > rr;
        bx += (bx + 1) * E12 <= rr;
        const int rem = rr - bx * E12;
        int by = (int)((float)rem * r2);
        by -= by * E2 > rem;
        by += (by + 1) * E2 <= rem;
        const int bz = rem - by * E2;
        const float sg = fsg[bx] * fsg[B::DX + by] * fsg[B::DX + DY + bz];
        float v = 0.f;
        if (sg != 0.f) v = src[c * sC + fofs[bx] + fofs[B::DX + by] + fofs[B::DX + DY + bz]] * sg;
        box[c * VOL + (bx * DY + by) * DZ + bz] = v;
      }
    } else {
      for (int i = threadIdx.x; i < NCC * VOL; i += NI_BLOCK) box[i] = 0.f;
    }
    // scatter: values of this chunk and the per-channel fixed-point scales
    float tv[TL_PER_THREAD][NCC];
    if (!GATHER) {
      float mloc[NCC];
#pragma unroll
      for (int c = 0; c < NCC; ++c) mloc[c] = 0.f;
#pragma unroll
      for (int k = 0; k < TL_PER_THREAD; ++k) {
#pragma unroll
        for (int c = 0; c < NCC; ++c) tv[k][c] = 0.f;
        if (!inb[k]) continue;
        if (MODE == TM_PUSH) {
          const float *tp = a.trgt + (int)(n * a.trgt_s[0] + (int64_t)c0 * a.trgt_s[1] + px[k] * a.trgt_s[2] +
                                           py[k] * a.trgt_s[3] + pz[k] * a.trgt_s[4]);
          const int tC = (int)a.trgt_s[1];
#pragma unroll
          for (int c = 0; c < NCC; ++c) tv[k][c] = c < nc ? tp[c * tC] : 0.f;
        } else {
          tv[k][0] = 1.f;
        }
#pragma unroll
        for (int c = 0; c < NCC; ++c) mloc[c] = fmaxf(mloc[c], fabsf(tv[k][c]));
      }
#pragma unroll
      for (int c = 0; c < NCC; ++c) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) mloc[c] = fmaxf(mloc[c], __shfl_xor(mloc[c], o, 64));
        if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6][c] = mloc[c];
      }
      __syncthreads();
      if (threadIdx.x < NCC) {
        float m = s_red[0][threadIdx.x];
#pragma unroll
        for (int wv = 1; wv < NI_BLOCK / 64; ++wv) m = fmaxf(m, s_red[wv][threadIdx.x]);
        int e = 0;
        if (m > 0.f && n_in > 0) {
          int q, kb = 0;
          frexpf(m, &q);                       // m < 2^q
          while ((1 << kb) < n_in) ++kb;       // n_in <= 2^kb
          e = 30 - q - kb;
          e = e > 126 ? 126 : (e < -126 ? -126 : e);
        }
        s_scale[threadIdx.x] = ldexpf(1.f, e);
        s_inv[threadIdx.x] = ldexpf(1.f, -e);
      }
    }
    __syncthreads();

#pragma unroll
    for (int k = 0; k < TL_PER_THREAD; ++k) {
      if (MODE == TM_PULL) {
        if (!valid[k]) continue;
        float acc[NCC];
#pragma unroll
        for (int c = 0; c < NCC; ++c) acc[c] = 0.f;
        if (inb[k]) {
          // z innermost: consecutive taps are adjacent LDS words (paired
          // reads); the sum order differs from pp_sample's (z outer) only
          // by rounding
          const float *bp = box + off[k];
          // kx rolled: keeps the LDS reads of one (kx, ky) row in flight
          // without the unrolled nest's register demand
#pragma unroll 1
          for (int kx = 0; kx < NT; ++kx)
#pragma unroll
            for (int ky = 0; ky < NT; ++ky) {
              const float wxy = w[k][0][kx] * w[k][1][ky];
#pragma unroll
              for (int kz = 0; kz < NT; ++kz) {
                const float ww = wxy * w[k][2][kz];
                const int o = kx * DYZ + ky * DZ + kz;
#pragma unroll
                for (int c = 0; c < NCC; ++c) acc[c] += bp[c * VOL + o] * ww;
              }
              // bound the LDS reads in flight to one row of taps
              __builtin_amdgcn_sched_barrier(0);
            }
        }
        float *op = a.out + (int)(n * a.out_s[0] + (int64_t)c0 * oC + px[k] * a.out_s[2] + py[k] * a.out_s[3] +
                                  pz[k] * a.out_s[4]);
#pragma unroll
        for (int c = 0; c < NCC; ++c)
          if (c < nc) op[c * oC] = acc[c];
      } else if (MODE == TM_SGRAD) {
        if (!valid[k]) continue;
        float ax[NCC], ay[NCC], az[NCC];
#pragma unroll
        for (int c = 0; c < NCC; ++c) ax[c] = ay[c] = az[c] = 0.f;
        if (inb[k]) {
          const float *bp = box + off[k];
#pragma unroll 1
          for (int kz = 0; kz < NT; ++kz)
#pragma unroll
            for (int ky = 0; ky < NT; ++ky)
#pragma unroll
              for (int kx = 0; kx < NT; ++kx) {
                const float wx = w[k][0][kx], wy = w[k][1][ky], wz = w[k][2][kz];
                const float gx = (gw[k][0][kx] * wy) * wz, gy = (wx * gw[k][1][ky]) * wz,
                            gz = (wx * wy) * gw[k][2][kz];
                const int o = kx * DYZ + ky * DZ + kz;
#pragma unroll
                for (int c = 0; c < NCC; ++c) {
                  const float v = bp[c * VOL + o];
                  ax[c] += v * gx;
                  ay[c] += v * gy;
                  az[c] += v * gz;
                }
                if (kx == NT - 1) __builtin_amdgcn_sched_barrier(0);
              }
        }
        float *op = a.out + (int)(n * a.out_s[0] + (int64_t)c0 * oC + px[k] * a.out_s[2] + py[k] * a.out_s[3] +
                                  pz[k] * a.out_s[4]);
#pragma unroll
        for (int c = 0; c < NCC; ++c)
          if (c < nc) {
            op[c * oC] = ax[c];
            op[c * oC + oK] = ay[c];
            op[c * oC + 2 * oK] = az[c];
          }
      } else if (MODE == TM_GGRAD || MODE == TM_GGRAD1) {
        if (!inb[k]) continue;
        float tvv[NCC];
        if (MODE == TM_GGRAD) {
          const float *tp = a.trgt + (int)(n * a.trgt_s[0] + (int64_t)c0 * a.trgt_s[1] + px[k] * a.trgt_s[2] +
                                           py[k] * a.trgt_s[3] + pz[k] * a.trgt_s[4]);
          const int tC = (int)a.trgt_s[1];
#pragma unroll
          for (int c = 0; c < NCC; ++c) tvv[c] = c < nc ? tp[c * tC] : 0.f;
        } else {
#pragma unroll
          for (int c = 0; c < NCC; ++c) tvv[c] = c < nc ? 1.f : 0.f;
        }
        const float *bp = box + off[k];
#pragma unroll 1
        for (int kz = 0; kz < NT; ++kz)
#pragma unroll
          for (int ky = 0; ky < NT; ++ky)
#pragma unroll
            for (int kx = 0; kx < NT; ++kx) {
              const float wx = w[k][0][kx], wy = w[k][1][ky], wz = w[k][2][kz];
              const float gx = (gw[k][0][kx] * wy) * wz, gy = (wx * gw[k][1][ky]) * wz,
                          gz = (wx * wy) * gw[k][2][kz];
              const int o = kx * DYZ + ky * DZ + kz;
              float dot = 0.f;
              // channels >= nc of a partial chunk hold stale LDS: select, not multiply by 0
#pragma unroll
              for (int c = 0; c < NCC; ++c) dot += c < nc ? bp[c * VOL + o] * tvv[c] : 0.f;
              og[k][0] += gx * dot;
              og[k][1] += gy * dot;
              og[k][2] += gz * dot;
              if (kx == NT - 1) __builtin_amdgcn_sched_barrier(0);
            }
      } else {
        if (!inb[k]) continue;
        float sv[NCC];
#pragma unroll
        for (int c = 0; c < NCC; ++c) sv[c] = tv[k][c] * s_scale[c];   // exact (power of two)
        int *bp = (int *)box + off[k];
#pragma unroll
        for (int kz = 0; kz < NT; ++kz)
#pragma unroll
          for (int ky = 0; ky < NT; ++ky)
#pragma unroll
            for (int kx = 0; kx < NT; ++kx) {
              const float ww = (w[k][0][kx] * w[k][1][ky]) * w[k][2][kz];
              const int o = kx * DYZ + ky * DZ + kz;
#pragma unroll
              for (int c = 0; c < NCC; ++c) atomicAdd(bp + c * VOL + o, (int)rintf(ww * sv[c]));
            }
      }
    }

    if (!GATHER) {
      // ---- add the box to the output, folded and signed ----
      __syncthreads();
      float *out = a.out + (int)(n * a.out_s[0] + (int64_t)c0 * oC);
      const int *ibox = (const int *)box;
      const int E2 = ext[2], E12 = ext[1] * ext[2], E = ext[0] * E12;
      const float rE = 1.f / (float)E, r12 = 1.f / (float)E12, r2 = 1.f / (float)E2;
      const int nflush = nc * E;
      for (int i = threadIdx.x; i < nflush; i += NI_BLOCK) {
        int c = (int)((float)i * rE);
        c -= c * E > i;
        c += (c + 1) * E <= i;
        const int rr = i - c * E;
        int bx = (int)((float)rr * r12);
        bx -= bx * E12 > rr;
        bx += (bx + 1) * E12 <= rr;
        const int rem = rr - bx * E12;
        int by = (int)((float)rem * r2);
        by -= by * E2 > rem;
        by += (by + 1) * E2 <= rem;
        const int bz = rem - by * E2;
        const int iv = ibox[c * VOL + (bx * DY + by) * DZ + bz];
        if (iv == 0) continue;
        const float v = (float)iv * s_inv[c];
        const float sg = fsg[bx] * fsg[B::DX + by] * fsg[B::DX + DY + bz];
        if (sg != 0.f) atomic_add(out + c * oC + fofs[bx] + fofs[B::DX + by] + fofs[B::DX + DY + bz], v * sg);
      }
    }
  }
  if (MODE == TM_GGRAD || MODE == TM_GGRAD1) {
#pragma unroll
    for (int k = 0; k < TL_PER_THREAD; ++k)
      if (valid[k]) {
        float *gp = a.gout + (int)(n * a.gout_s[0] + px[k] * a.gout_s[1] + py[k] * a.gout_s[2] + pz[k] * a.gout_s[3]);
        const int gD = (int)a.gout_s[4];
        gp[0] = og[k][0];
        gp[gD] = og[k][1];
        gp[2 * gD] = og[k][2];
      }
  }
}

// Tiles whose box does not fit: the same tiling and bounding-box test as
// tile_kernel (bit-identical, so every tile is handled by exactly one of the
// two kernels), then the direct per-sample routine.  Kept in its own kernel
// so that its register demand does not lower the occupancy of tile_kernel.
template <int ORD, int MODE>
__global__ __launch_bounds__(NI_BLOCK) void tile_fallback(PPArgs<float> a, int ntx, int nty, int ntz,
                                                          int64_t nblocks) {
  constexpr int NT = ORD + 1;
  constexpr int FL = TileFlags<MODE>::value;
  using B = TileBox<ORD>;
  __shared__ int red[NI_BLOCK / 64][6];
  int64_t r = xcd_block(nblocks);
  const int tz = (int)(r % ntz); r /= ntz;
  const int ty = (int)(r % nty); r /= nty;
  const int tx = (int)(r % ntx);
  const int n = (int)(r / ntx);
  const int X = (int)a.tX, Y = (int)a.tY, Z = (int)a.tZ;
  const int sz[3] = {(int)a.sX, (int)a.sY, (int)a.sZ};
  int px[TL_PER_THREAD], py[TL_PER_THREAD], pz[TL_PER_THREAD];
  bool valid[TL_PER_THREAD];
  int mn[3] = {INT_MAX, INT_MAX, INT_MAX}, mx[3] = {INT_MIN, INT_MIN, INT_MIN};
  const float tol = (a.extrap & 2) ? 0.55f : 0.05f;
#pragma unroll
  for (int k = 0; k < TL_PER_THREAD; ++k) {
    const int s = threadIdx.x + k * NI_BLOCK;
    px[k] = tx * TL_X + s / (TL_Y * TL_Z);
    py[k] = ty * TL_Y + (s / TL_Z) % TL_Y;
    pz[k] = tz * TL_Z + s % TL_Z;
    valid[k] = px[k] < X && py[k] < Y && pz[k] < Z;
    if (valid[k]) {
      const float *gp = a.grid + (int)(n * a.grid_s[0] + px[k] * a.grid_s[1] + py[k] * a.grid_s[2] + pz[k] * a.grid_s[3]);
      const int gD = (int)a.grid_s[4];
      const float xs[3] = {gp[0], gp[gD], gp[2 * gD]};
      bool ok = true;
      if (!(a.extrap & 1))
#pragma unroll
        for (int d = 0; d < 3; ++d) ok = ok && (xs[d] >= -tol && xs[d] < (float)(sz[d] - 1) + tol);
      if (ok) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          float wd[NT], gd[NT];
          const int l = tl_taps<ORD>(xs[d], wd, gd);
          mn[d] = min(mn[d], l);
          mx[d] = max(mx[d], l);
        }
      }
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    mn[d] = wave_min(mn[d]);
    mx[d] = wave_max(mx[d]);
  }
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      red[threadIdx.x >> 6][d] = mn[d];
      red[threadIdx.x >> 6][3 + d] = mx[d];
    }
  }
  __syncthreads();
  bool fits = true;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    int m0 = red[0][d], m1 = red[0][3 + d];
#pragma unroll
    for (int wv = 1; wv < NI_BLOCK / 64; ++wv) {
      m0 = min(m0, red[wv][d]);
      m1 = max(m1, red[wv][3 + d]);
    }
    if (m0 > m1) return;   // no in-bound sample: handled by tile_kernel
    const int64_t e = (int64_t)m1 - (int64_t)m0 + NT;
    fits = fits && e <= (d == 0 ? B::DX : d == 1 ? B::DY : B::DZ);
  }
  if (fits) return;
#pragma unroll
  for (int k = 0; k < TL_PER_THREAD; ++k)
    if (valid[k]) pp_sample<float, int32_t, ORD, FL, true>(a, n, px[k], py[k], pz[k]);
}

template <int ORD, int MODE>
static void launch_tile(const PPArgs<float> &a, hipStream_t st) {
  const int ntx = (int)((a.tX + TL_X - 1) / TL_X);
  const int nty = (int)((a.tY + TL_Y - 1) / TL_Y);
  const int ntz = (int)((a.tZ + TL_Z - 1) / TL_Z);
  const int64_t nb = a.N * (int64_t)ntx * nty * ntz;
  const dim3 g((unsigned)nb), b(NI_BLOCK);
  constexpr int M = TileBox<ORD>::MAX_NCC;
  if constexpr (MODE == TM_COUNT) {
    hipLaunchKernelGGL((tile_kernel<ORD, MODE, 1>), g, b, 0, st, a, ntx, nty, ntz, nb);
  } else {
    if (a.C == 1)
      hipLaunchKernelGGL((tile_kernel<ORD, MODE, 1>), g, b, 0, st, a, ntx, nty, ntz, nb);
    else if (a.C == 2 || M == 2)
      hipLaunchKernelGGL((tile_kernel<ORD, MODE, (M < 2 ? M : 2)>), g, b, 0, st, a, ntx, nty, ntz, nb);
    else
      hipLaunchKernelGGL((tile_kernel<ORD, MODE, M>), g, b, 0, st, a, ntx, nty, ntz, nb);
  }
  hipLaunchKernelGGL((tile_fallback<ORD, MODE>), g, b, 0, st, a, ntx, nty, ntz, nb);
}

}  // namespace ni_amd
