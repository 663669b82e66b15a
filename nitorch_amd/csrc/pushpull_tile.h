// LDS-tiled pull / push / count / gradients for 3-D float fields (isotropic
// order ORD).
//
// One work-group owns a TX x TY x TZ tile of samples (grid points); each
// thread owns R of them.  The work-group reduces the bounding box of the
// first taps of its in-bound samples, and when the box (extended by the
// ORD + 1 taps) fits in LDS:
//  * gather modes (pull, gradients): stage the box of the source -- folded
//            by the boundary condition and signed, so the taps need no
//            folding -- in LDS, then read every tap from LDS;
//  * push / count : accumulate the tile's splats into an LDS box of FP64
//            values with FP64 LDS atomics (gfx950 runs ds_add_f64 at ~80% of
//            the rate of its 64-bit integer atomics, and ~40x faster than
//            ds_add_f32), then add the box to the output once, folded and
//            signed, with one FP32 global atomic per non-zero box element.
//            At orders >= 1 a thread's z-consecutive samples whose taps are
//            z-consecutive too first sum their splats along z in registers
//            (at most R = 4 or 2 FP32 terms a box element) and add one FP64
//            atomic a position.  Every box element is a double sum of its splats
//            (FP64 products at order 0; above, FP32 sums of at most R FP32
//            products), so each output element keeps float's relative
//            accuracy against the sum of its splats' magnitudes however
//            small it is next to the tile's other values, and NaN / Inf
//            values propagate as in the reference's float sums.
// The box has fixed dimensions per order (every tap offset is then an
// immediate of the LDS instruction); channels are staged NCC at a time.
// A tile whose box does not fit (scattered or far out-of-range
// coordinates with extrapolation) is listed for a second pass.  Pull, push,
// count and grid_grad: in calls of at least BIN_MIN_SAMPLES samples the
// listed tiles' samples are sorted by first-tap bin and every bin runs with
// a box of its own (pushpull_bin.h).  In smaller calls of pull, push and
// count, and in the entry point that allocates nothing, tile_dyn cuts each
// listed tile into sub-tiles with boxes sized to their own spread and runs
// the direct per-sample routine pp_sample for a sub-tile that still does
// not fit; tile_fallback runs pp_sample for the other gradient cases.  The
// result never depends on the data's smoothness, only the speed does.
//
// Weights, the out-of-bound rule and the splat order of each sample are
// those of pp_sample (nitorch/_C/src/impl/pushpull_common.cpp: OOB :926-974,
// interpolate3d / splat3d :1176-1290, linear :1653-1858, nearest
// :2314-2341); folding and signs follow bounds_common.h:28-219.  Gathers
// sum the taps x-outer, z-inner (pp_sample: z-outer), which differs only by
// rounding.
#pragma once
#include <climits>

#include "pushpull_impl.h"

namespace ni_amd {

#define TL_SLACK 2     // extra box extent per axis for the spread of a tile's coordinates
#define BIN_MIN_SAMPLES (int64_t(1) << 20)   // smallest call whose listed tiles take the binned pass

// Tile and box geometry of order ORD, fixed at compile time.
//  * Gathers: 256 threads, 4 x 8 x 16 tiles, two samples per thread WY
//    y-rows apart (larger tiles re-read less halo but cost occupancy, which
//    the gather needs more).  Scatters up to order 5: 512 threads,
//    8 x 16 x 16 tiles, four z-consecutive samples per thread -- every box
//    element is one global atomic in the flush and a box exceeds its tile
//    by the ORD + 1 taps plus TL_SLACK voxels per axis, so the atomics per
//    sample fall with the tile size, and the register window of the splat
//    spans four samples.  Scatters above order 5 (the window's weights of
//    four samples would not fit the registers, a box of 94 KB one CU's LDS
//    twice): 256 threads, 4 x 8 x 16, two samples per thread.
//  * Lanes: gathers put 16 z-samples x 4 x-rows in a wave; scatters
//    (ZL = TZ / R threads per z-row) R y-rows fastest, then x-rows.
//  * The box is [c][x][y][z] with x stride DYZP padded to 48 (mod 64) words
//    and, for scatters, DZ = 1 (mod R): the 32 lanes of a half-wave reading
//    or adding one tap fall on the 32 distinct LDS banks (gathers: two
//    x-rows of 16 z-samples, 16 banks apart; scatters: R y-rows in the R
//    residue classes mod R, two x-rows 16 banks apart in each class, R
//    lanes R banks apart per row).
template <int ORD, bool SCAT>
struct TileBox {
  static constexpr bool BIG = SCAT && ORD <= 5;   // the 512-thread scatter
  static constexpr int NTH = BIG ? 512 : 256;          // threads per work-group
  static constexpr int NW = NTH / 64;                  // waves
  static constexpr int R = BIG ? 4 : 2;                // samples per thread
  static constexpr int TX = BIG ? 8 : 4;
  static constexpr int TY = BIG ? 16 : 8;
  static constexpr int TZ = 16;
  static constexpr int ZL = SCAT ? TZ / R : TZ;        // threads per z-row of the tile
  static constexpr int RY = SCAT ? R : 1;              // y-rows per wave
  static constexpr int RX = 64 / ZL / RY;              // x-rows per wave
  static constexpr int WX = TX / RX;                   // waves along x
  static constexpr int WY = NW / WX;                   // waves along y
  static constexpr int SDY = SCAT ? 0 : WY;            // step between a thread's samples
  static constexpr int SDZ = SCAT ? 1 : 0;
  static constexpr int DX = TX + ORD + TL_SLACK;
  static constexpr int DY = TY + ORD + TL_SLACK;
  static constexpr int DZ0 = TZ + ORD + TL_SLACK;
  static constexpr int DZ = SCAT ? DZ0 + ((1 - DZ0) % R + R) % R : DZ0;
  static constexpr int DYZ = DY * DZ;
  static constexpr int DYZP = DYZ + ((48 - DYZ) % 64 + 64) % 64;
  static constexpr int VOL = DX * DYZP;
  // pull: 4 interleaved channels (16 B) per voxel, x stride DYZ4 = 16 (mod
  // 32) voxels, so that two x rows read in one LDS cycle fall on opposite
  // halves of the banks
  static constexpr int DYZ4 = DYZ + ((16 - DYZ) % 32 + 32) % 32;
  static constexpr int VOL4 = DX * DYZ4;
  static constexpr int NPP = (DYZ + NTH - 1) / NTH;     // y-z plane positions per thread
  // channels staged at once: 1 for the scatters (a 52 KB box of doubles
  // at order 3), 4 for the gathers (pull, grid_grad, the grid gradients of
  // the backward passes), interleaved per voxel so that a tap is one
  // 16-byte LDS read
  static constexpr int MAX_NCC = SCAT ? 1 : 4;
  // waves per SIMD asked of the register allocator for the 512-thread
  // scatter: 6 (<= 85 VGPRs) lets three 52 KB boxes share a CU's LDS; the
  // windowed scatter (orders 1-5 here) holds its samples' weights and
  // needs 97 (order 3) to 128 (order 5) VGPRs: two work-groups a CU
  static constexpr int MINW = BIG ? (ORD >= 1 ? 4 : 6) : 1;
  static_assert(TX * TY * TZ == NTH * R, "R samples per thread");
  static_assert(WX * RX == TX && WX * WY == NW, "waves tile the x rows");
  static_assert(SCAT ? WY * RY == TY : WY * R == TY, "waves tile the y rows");
};

// First sample of this thread within its tile (sample i adds i * (SDY, SDZ)).
template <class B>
__device__ __forceinline__ void tl_thread(int &sx, int &sy, int &sz0) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int row = lane / B::ZL;
  sz0 = (lane % B::ZL) * (B::SDZ ? B::R : 1);
  if (B::SDZ) {
    sx = row / B::RY + B::RX * (w % B::WX);
    sy = row % B::RY + B::RY * (w / B::WX);
  } else {
    sx = row % B::RX + B::RX * (w % B::WX);
    sy = w / B::WX;
  }
}

// The y-z plane of a box's ext[1] x ext[2] corner, split over the threads:
// position j of a thread is (by, bz), with folded source offset and sign
// from the axis tables and the LDS offset by * DZ + bz.
template <class B>
struct TilePlane {
  int yz_ofs[B::NPP];   // folded y + z offset in the tensor, elements
  int yz_lds[B::NPP];   // by * DZ + bz, or -1 when the position is outside the plane
  float yz_sg[B::NPP];  // product of the y and z signs
  __device__ __forceinline__ void init(int E1, int E2, const int *fofs, const float *fsg) {
    const float r2 = 1.f / (float)E2;
#pragma unroll
    for (int j = 0; j < B::NPP; ++j) {
      const int p = threadIdx.x + j * B::NTH;
      int by = (int)((float)p * r2);   // floor(p / E2): exact after the fix-ups (p < 2^22)
      by -= by * E2 > p;
      by += (by + 1) * E2 <= p;
      const int bz = p - by * E2;
      const bool in = by < E1;
      yz_lds[j] = in ? by * B::DZ + bz : -1;
      yz_ofs[j] = in ? fofs[B::DX + by] + fofs[B::DX + B::DY + bz] : 0;
      yz_sg[j] = in ? fsg[B::DX + by] * fsg[B::DX + B::DY + bz] : 0.f;
    }
  }
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

// Spline weight / derivative of tap T at signed distance d: the
// evaluation of spl_w / spl_g (common.h), with the piece and its Horner
// coefficients fixed at compile time (immediates, never table loads).
template <int O, int P, int I>
__device__ __forceinline__ float tl_horner_w(float u) {
  constexpr float c = (float)SplTab<O>::W[P][I];
  if constexpr (I == SplTab<O>::NDW - 1) return c;
  else return tl_horner_w<O, P, I + 1>(u) * u + c;
}
template <int O, int P, int I>
__device__ __forceinline__ float tl_horner_g(float u) {
  constexpr float c = (float)SplTab<O>::G[P][I];
  if constexpr (I == SplTab<O>::NDG - 1) return c;
  else return tl_horner_g<O, P, I + 1>(u) * u + c;
}
template <int ORD, bool DERIV, int T>
__device__ __forceinline__ void tl_weight(float x, int lo, float (&w)[ORD + 1], float (&g)[ORD + 1]) {
  if constexpr (T <= ORD) {
    if constexpr (T == 0 || T == ORD) {   // last piece, from x itself (spl_outer_s)
      float h;
      spl_outer<float, ORD>(T == 0, spl_outer_s<float, int>(ORD, T, x, lo), w[T], g[T], h);
    } else {
      constexpr int P = tap_piece<ORD>(T);
      constexpr float centre = (float)piece_centre<ORD>(P);
      const float d = x - (float)(lo + T);
      const float a = d < 0.f ? -d : d;
      w[T] = tl_horner_w<ORD, P, 0>(a - centre);
      if constexpr (DERIV) {
        const float gg = tl_horner_g<ORD, P, 0>(a - centre);
        g[T] = d < 0.f ? -gg : gg;
      }
    }
    tl_weight<ORD, DERIV, T + 1>(x, lo, w, g);
  }
}

// First tap index and the ORD + 1 weights (and, when DERIV, first
// derivatives) of coordinate x, as pp_sample computes them.
template <int ORD, bool DERIV = false>
__device__ __forceinline__ int tl_taps(float x, float (&w)[ORD + 1], float (&g)[ORD + 1]) {
  if constexpr (ORD == 0) {
    w[0] = 1.f;
    g[0] = 0.f;
    return (int)roundf(x);
  } else if constexpr (ORD == 1) {
    // weights as distances to the far tap: exact when small
    const int j0 = (int)floorf(x);
    w[0] = (float)(j0 + 1) - x;
    w[1] = x - (float)j0;
    g[0] = -1.f;
    g[1] = 1.f;
    return j0;
  } else {
    const int lo = (int)floorf(x - first_tap_offset<float>(ORD));
    tl_weight<ORD, DERIV, 0>(x, lo, w, g);
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

// p / d for 0 <= p, d < 2^22 from the float reciprocal rinv = 1 / d: the
// float quotient is within 1 of the exact one, one correction each way
// (a runtime integer division costs ~30 VALU operations)
__device__ __forceinline__ int tl_div(int p, int d, float rinv) {
  int q = (int)((float)p * rinv);
  q -= q * d > p;
  q += (q + 1) * d <= p;
  return q;
}

// Rotate a small register array left by one: a rolled tap loop reads
// element 0 and rotates, so that no array is indexed by the loop counter
// (which would put it in scratch memory); NT rotations restore the order.
template <typename T, int N>
__device__ __forceinline__ void tl_rot(T (&v)[N]) {
  const T t = v[0];
#pragma unroll
  for (int j = 0; j + 1 < N; ++j) v[j] = v[j + 1];
  v[N - 1] = t;
}

// Hide a register array's values from the optimiser for one loop
// iteration, so that products of it are not hoisted out of the loop (where
// they would all be live at once).
template <int N>
__device__ __forceinline__ void tl_opaque(float (&v)[N]) {
#pragma unroll
  for (int j = 0; j < N; ++j) asm volatile("" : "+v"(v[j]));
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
__global__ __launch_bounds__((TileBox<ORD, !TileMeta<MODE>::gather>::NTH),
                             (TileBox<ORD, !TileMeta<MODE>::gather>::MINW)) void tile_kernel(PPArgs<float> a, int ntx, int nty, int ntz,
                                                        int64_t nblocks, int *fb) {
  constexpr int NT = ORD + 1;
  constexpr bool GATHER = TileMeta<MODE>::gather, DERIV = TileMeta<MODE>::deriv;
  using B = TileBox<ORD, !GATHER>;
  constexpr int R = B::R;
  // pull and grid_grad stage their box with the NCC = 4 channels of a
  // voxel interleaved (one 16-byte LDS read per tap: gfx950's LDS serves
  // 128-bit reads at ~1.5x the bytes per clock of 32-bit ones, and 16
  // z-consecutive lanes read 16 consecutive 16-byte voxels, which no bank
  // serves twice); the other modes keep channel planes [c][x][y][z]
  constexpr bool IL = MODE == TM_PULL || (DERIV && NCC == 4);
  constexpr int DY = B::DY, DZ = B::DZ, DYZ = IL ? B::DYZ4 : B::DYZP, VOL = IL ? B::VOL4 : B::VOL;
  static_assert(NCC >= 1 && NCC <= B::MAX_NCC, "channel chunk exceeds the LDS budget");
  static_assert(!IL || NCC == 4 || NCC == 1, "pull / grad interleave 4 channels (pull may run one)");
  constexpr int NTH = B::NTH;
  // gathers: FP32 source box; scatters: FP64 box (two words per element)
  __shared__ __align__(16) float box[(GATHER ? 1 : 2) * NCC * VOL];
  __shared__ int red[B::NW][6];

  const int64_t tile = xcd_block(nblocks);
  int64_t r = tile;
  const int tz = (int)(r % ntz); r /= ntz;
  const int ty = (int)(r % nty); r /= nty;
  const int tx = (int)(r % ntx);
  const int n = (int)(r / ntx);
  const int X = (int)a.tX, Y = (int)a.tY, Z = (int)a.tZ;
  const int sz[3] = {(int)a.sX, (int)a.sY, (int)a.sZ};

  // ---- samples of this thread: in-bound flags, taps ----
  // sample i is (px, py + i * SDY, pz0 + i * SDZ)
  int sx, sy, sz0;
  tl_thread<B>(sx, sy, sz0);
  const int px = tx * B::TX + sx, py = ty * B::TY + sy, pz0 = tz * B::TZ + sz0;
  bool valid[R], inb[R];
  // pull and the windowed scatter keep every sample's weights; the other
  // modes keep the coordinates and compute one sample's weights (and
  // derivatives) when it is processed (registers: occupancy)
  constexpr bool WIN = !GATHER && ORD >= 1;
  static_assert(!WIN || (B::SDZ == 1 && B::SDY == 0), "the scatter window needs z-consecutive samples");
  constexpr int NW = DERIV ? 1 : (GATHER || WIN) ? R : 1;
  int lo[R][3];
  float w[NW][3][NT];
  float xq[R][3];
  if constexpr (WIN) {   // the window multiplies the weights of samples outside a group by 0: keep them finite
#pragma unroll
    for (int i = 0; i < NW; ++i)
#pragma unroll
      for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int t = 0; t < NT; ++t) w[i][d][t] = 0.f;
  }
  int mn[3] = {INT_MAX, INT_MAX, INT_MAX}, mx[3] = {INT_MIN, INT_MIN, INT_MIN};
  const float tol = (a.extrap & 2) ? 0.55f : 0.05f;
  const int gD = (int)a.grid_s[4];
  const int gS = (int)(B::SDY * a.grid_s[2] + B::SDZ * a.grid_s[3]);
  const float *gp0 = a.grid + (int)(n * a.grid_s[0] + px * a.grid_s[1] + py * a.grid_s[2] + pz0 * a.grid_s[3]);
#pragma unroll
  for (int i = 0; i < R; ++i) {
    valid[i] = px < X && py + i * B::SDY < Y && pz0 + i * B::SDZ < Z;
    inb[i] = false;
    if (valid[i]) {
      const float *gp = gp0 + i * gS;
      const float xs[3] = {gp[0], gp[gD], gp[2 * gD]};
      bool ok = true;
      if (!(a.extrap & 1))
#pragma unroll
        for (int d = 0; d < 3; ++d) ok = ok && (xs[d] >= -tol && xs[d] < (float)(sz[d] - 1) + tol);
      inb[i] = ok;
      if (ok) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          float wtmp[NT], gtmp[NT];
          xq[i][d] = xs[d];
          lo[i][d] = tl_taps<ORD>(xs[d], wtmp, gtmp);
          if constexpr (NW == R) {
#pragma unroll
            for (int t = 0; t < NT; ++t) w[i][d][t] = wtmp[t];
          }
          mn[d] = min(mn[d], lo[i][d]);
          mx[d] = max(mx[d], lo[i][d]);
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
    for (int wv = 1; wv < B::NW; ++wv) {
      mn[d] = min(mn[d], red[wv][d]);
      mx[d] = max(mx[d], red[wv][3 + d]);
    }
  }

  const int oC = (int)a.out_s[1], oK = (int)a.out_s[5];
  const int oS = (int)(B::SDY * a.out_s[3] + B::SDZ * a.out_s[4]);   // between a thread's samples
  if (mn[0] > mx[0]) {   // no in-bound sample in this tile
    if (MODE == TM_PULL || MODE == TM_SGRAD) {
      const int C = (int)a.C, nk = MODE == TM_SGRAD ? 3 : 1;
      float *op0 = a.out + (int)(n * a.out_s[0] + px * a.out_s[2] + py * a.out_s[3] + pz0 * a.out_s[4]);
#pragma unroll
      for (int i = 0; i < R; ++i)
        if (valid[i])
          for (int c = 0; c < C; ++c)
            for (int d = 0; d < nk; ++d) op0[i * oS + c * oC + d * oK] = 0.f;
    } else if (MODE == TM_GGRAD || MODE == TM_GGRAD1) {
      float *gq = a.gout + (int)(n * a.gout_s[0] + px * a.gout_s[1] + py * a.gout_s[2] + pz0 * a.gout_s[3]);
#pragma unroll
      for (int i = 0; i < R; ++i)
        if (valid[i])
          for (int d = 0; d < 3; ++d) gq[i * (B::SDY * a.gout_s[2] + B::SDZ * a.gout_s[3]) + d * a.gout_s[4]] = 0.f;
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
  if (!fits) {   // scattered tile: listed for the second kernel
    if (fb && threadIdx.x == 0) fb[1 + atomicAdd(fb, 1)] = (int)tile;
    return;
  }
  int off[R];
#pragma unroll
  for (int i = 0; i < R; ++i)
    off[i] = inb[i] ? (lo[i][0] - mn[0]) * DYZ + (lo[i][1] - mn[1]) * DZ + (lo[i][2] - mn[2]) : 0;

  const int C = MODE == TM_COUNT ? 1 : (int)a.C;

  // folded offsets (in elements of the staged / scattered tensor) and signs
  // of the box positions along each axis: fold() runs once per position
  constexpr int NF = B::DX + DY + DZ;
  __shared__ int fofs[NF];
  __shared__ float fsg[NF];
  {
    const int64_t *st = GATHER ? a.src_s : a.out_s;
    const int st0 = (int)st[2], st1 = (int)st[3], st2 = (int)st[4];
    for (int i = threadIdx.x; i < NF; i += NTH) {
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
  TilePlane<B> pl;
  pl.init(ext[1], ext[2], fofs, fsg);
  if (!GATHER) {   // scatter box: cleared once; each flush clears what it read
    double *dbox = (double *)box;
    for (int i = threadIdx.x; i < NCC * VOL; i += NTH) dbox[i] = 0.0;
    __syncthreads();
  }

  float og[R][3];   // grid gradient (TM_GGRAD*), summed over chunks
#pragma unroll
  for (int i = 0; i < R; ++i) og[i][0] = og[i][1] = og[i][2] = 0.f;

  for (int c0 = 0; c0 < C; c0 += NCC) {
    const int nc = (C - c0) < NCC ? (C - c0) : NCC;
    if (c0) __syncthreads();   // previous chunk fully consumed
    // keep the per-tap weight products inside the chunk loop (hoisting
    // them would hold (ORD + 1)^3 products per sample in registers)
#pragma unroll
    for (int i = 0; i < NW; ++i)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
#pragma unroll
        for (int t = 0; t < NT; ++t) asm volatile("" : "+v"(w[i][d][t]));
      }

    if (GATHER) {
      // ---- stage the folded, signed source box ----
      // only the ext[0] x ext[1] x ext[2] corner of the box is ever read:
      // each thread fills its y-z plane positions for every channel and x
      // one channel row at a time: all DX loads issued before any store
      // (positions past ext[0] re-read the last one: every folded offset is
      // in range, so no load needs a guard)
      const float *src = a.src + (int)(n * a.src_s[0] + c0 * a.src_s[1]);
      const int sC = (int)a.src_s[1];
      const int e0 = ext[0] - 1;
      if constexpr (IL && NCC == 4) {
        // one x row of the 4 channels at a time: 4 DX loads, then DX
        // 16-byte stores (channels >= nc of a partial chunk stage zeros)
#pragma unroll
        for (int j = 0; j < B::NPP; ++j) {
          if (pl.yz_lds[j] < 0) continue;
          const float *sp = src + pl.yz_ofs[j];
          float v[4][B::DX];
#pragma unroll
          for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int bx = 0; bx < B::DX; ++bx) v[c][bx] = c < nc ? sp[c * sC + fofs[bx < e0 ? bx : e0]] : 0.f;
          float4 *bq = (float4 *)box + pl.yz_lds[j];
#pragma unroll
          for (int bx = 0; bx < B::DX; ++bx) {
            const float sg = fsg[bx] * pl.yz_sg[j];
            if (bx <= e0)
              bq[bx * DYZ] = sg != 0.f ? make_float4(v[0][bx] * sg, v[1][bx] * sg, v[2][bx] * sg, v[3][bx] * sg)
                                       : make_float4(0.f, 0.f, 0.f, 0.f);
          }
        }
      } else
#pragma unroll
      for (int j = 0; j < B::NPP; ++j) {
        if (pl.yz_lds[j] < 0) continue;
#pragma unroll 1
        for (int c = 0; c < nc; ++c) {
          const float *sp = src + c * sC + pl.yz_ofs[j];
          float v[B::DX];
#pragma unroll
          for (int bx = 0; bx < B::DX; ++bx) v[bx] = sp[fofs[bx < e0 ? bx : e0]];
          float *bq = box + c * VOL + pl.yz_lds[j];
#pragma unroll
          for (int bx = 0; bx < B::DX; ++bx) {
            const float sg = fsg[bx] * pl.yz_sg[j];
            if (bx <= e0) bq[bx * DYZ] = sg != 0.f ? v[bx] * sg : 0.f;
          }
        }
      }
    }
    // scatter: values of this chunk
    float tv[R][NCC];
    if (!GATHER) {
      const float *tp0 = MODE == TM_PUSH ? a.trgt + (int)(n * a.trgt_s[0] + (int64_t)c0 * a.trgt_s[1] +
                                                          px * a.trgt_s[2] + py * a.trgt_s[3] + pz0 * a.trgt_s[4])
                                         : nullptr;
      const int tC = (int)a.trgt_s[1], tS = (int)(B::SDY * a.trgt_s[3] + B::SDZ * a.trgt_s[4]);
#pragma unroll
      for (int i = 0; i < R; ++i) {
#pragma unroll
        for (int c = 0; c < NCC; ++c) tv[i][c] = 0.f;
        if (!inb[i]) continue;
        if (MODE == TM_PUSH) {
#pragma unroll
          for (int c = 0; c < NCC; ++c) tv[i][c] = c < nc ? tp0[i * tS + c * tC] : 0.f;
        } else {
          tv[i][0] = 1.f;
        }
      }
    }
    __syncthreads();

    if (MODE == TM_PULL) {
      float acc[R][NCC];
#pragma unroll
      for (int i = 0; i < R; ++i)
#pragma unroll
        for (int c = 0; c < NCC; ++c) acc[i][c] = 0.f;
      {
#pragma unroll
        for (int i = 0; i < R; ++i) {
          if (!inb[i]) continue;
          if constexpr (NCC == 4) {
            const float4 *bp = (const float4 *)box + off[i];
#pragma unroll 1
            for (int kx = 0; kx < NT; ++kx, bp += DYZ) {
#pragma unroll
              for (int ky = 0; ky < NT; ++ky) {
                const float wxy = w[i][0][0] * w[i][1][ky];
#pragma unroll
                for (int kz = 0; kz < NT; ++kz) {
                  const float ww = wxy * w[i][2][kz];
                  const float4 v = bp[ky * DZ + kz];
                  acc[i][0] += v.x * ww;
                  acc[i][1] += v.y * ww;
                  acc[i][2] += v.z * ww;
                  acc[i][3] += v.w * ww;
                }
                __builtin_amdgcn_sched_barrier(0);
              }
              tl_rot(w[i][0]);
            }
          } else {
            const float *bp = box + off[i];
#pragma unroll 1
            for (int kx = 0; kx < NT; ++kx, bp += DYZ) {
#pragma unroll
              for (int ky = 0; ky < NT; ++ky) {
                const float wxy = w[i][0][0] * w[i][1][ky];
#pragma unroll
                for (int kz = 0; kz < NT; ++kz) {
                  const float ww = wxy * w[i][2][kz];
                  const int o = ky * DZ + kz;
#pragma unroll
                  for (int c = 0; c < NCC; ++c) acc[i][c] += bp[c * VOL + o] * ww;
                }
                __builtin_amdgcn_sched_barrier(0);
              }
              tl_rot(w[i][0]);
            }
          }
        }
      }
      float *op0 = a.out + (int)(n * a.out_s[0] + (int64_t)c0 * oC + px * a.out_s[2] + py * a.out_s[3] +
                                 pz0 * a.out_s[4]);
#pragma unroll
      for (int i = 0; i < R; ++i)
        if (valid[i])
#pragma unroll
          for (int c = 0; c < NCC; ++c)
            if (c < nc) op0[i * oS + c * oC] = acc[i][c];
    } else if (MODE == TM_SGRAD) {
      // samples one at a time (rolled): element 0 of the per-sample state
      // is the current sample, rotated after each (R rotations restore it)
#pragma unroll 1
      for (int i = 0; i < R; ++i) {
        const bool vi = valid[0], ii = inb[0];
        const int oi = off[0];
        const float q0 = xq[0][0], q1 = xq[0][1], q2 = xq[0][2];
        tl_rot(valid);
        tl_rot(inb);
        tl_rot(off);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const float t = xq[0][d];
#pragma unroll
          for (int k = 0; k + 1 < R; ++k) xq[k][d] = xq[k + 1][d];
          xq[R - 1][d] = t;
        }
        if (!vi) continue;
        float ax[NCC], ay[NCC], az[NCC];
#pragma unroll
        for (int c = 0; c < NCC; ++c) ax[c] = ay[c] = az[c] = 0.f;
        if (ii) {
          float wi[3][NT], gi[3][NT];
          tl_taps<ORD, true>(q0, wi[0], gi[0]);
          tl_taps<ORD, true>(q1, wi[1], gi[1]);
          tl_taps<ORD, true>(q2, wi[2], gi[2]);
          const float *bp = box + oi;
#pragma unroll 1
          // separable: per (z, y) tap row, the x sums with weights and
          // derivatives, then the y-z factors (x inner, y, z rolled)
          for (int kz = 0; kz < NT; ++kz, ++bp) {
            const float *bq = bp;
#pragma unroll 1
            for (int ky = 0; ky < NT; ++ky, bq += DZ) {
              tl_opaque(wi[0]);
              tl_opaque(gi[0]);
              const float wyz = wi[1][0] * wi[2][0], gyz = gi[1][0] * wi[2][0], wgz = wi[1][0] * gi[2][0];
              if constexpr (IL) {   // 4 interleaved channels: one 16-byte read a tap
                // bq - box counts voxels here (off[] and the strides are in voxels)
                const float4 *vq = (const float4 *)box + (bq - box);
                float sw[4] = {0.f, 0.f, 0.f, 0.f}, sd[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int kx = 0; kx < NT; ++kx) {
                  const float4 v = vq[kx * DYZ];
                  const float vv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                  for (int c = 0; c < 4; ++c) {
                    sw[c] += wi[0][kx] * vv[c];
                    sd[c] += gi[0][kx] * vv[c];
                  }
                }
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                  ax[c] += sd[c] * wyz;
                  ay[c] += sw[c] * gyz;
                  az[c] += sw[c] * wgz;
                }
              } else {
#pragma unroll
                for (int c = 0; c < NCC; ++c) {
                  float sw = 0.f, sd = 0.f;
#pragma unroll
                  for (int kx = 0; kx < NT; ++kx) {
                    const float v = bq[c * VOL + kx * DYZ];
                    sw += wi[0][kx] * v;
                    sd += gi[0][kx] * v;
                  }
                  ax[c] += sd * wyz;
                  ay[c] += sw * gyz;
                  az[c] += sw * wgz;
                }
              }
              tl_rot(wi[1]);
              tl_rot(gi[1]);
            }
            tl_rot(wi[2]);
            tl_rot(gi[2]);
          }
        }
        float *op = a.out + (int)(n * a.out_s[0] + (int64_t)c0 * oC + px * a.out_s[2] + py * a.out_s[3] +
                                  pz0 * a.out_s[4]) + i * oS;
#pragma unroll
        for (int c = 0; c < NCC; ++c)
          if (c < nc) {
            op[c * oC] = ax[c];
            op[c * oC + oK] = ay[c];
            op[c * oC + 2 * oK] = az[c];
          }
      }
    } else if (MODE == TM_GGRAD || MODE == TM_GGRAD1) {
      // samples one at a time (rolled), as in TM_SGRAD; og rotates with them
#pragma unroll 1
      for (int i = 0; i < R; ++i) {
        const bool ii = inb[0];
        const int oi = off[0];
        const float q0 = xq[0][0], q1 = xq[0][1], q2 = xq[0][2];
        float o0 = og[0][0], o1 = og[0][1], o2 = og[0][2];
        tl_rot(inb);
        tl_rot(off);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const float t = xq[0][d];
#pragma unroll
          for (int k = 0; k + 1 < R; ++k) {
            xq[k][d] = xq[k + 1][d];
            og[k][d] = og[k + 1][d];
          }
          xq[R - 1][d] = t;
        }
        if (ii) {
        float tvv[NCC];
        if (MODE == TM_GGRAD) {
          const float *tp = a.trgt + (int)(n * a.trgt_s[0] + (int64_t)c0 * a.trgt_s[1] + px * a.trgt_s[2] +
                                           py * a.trgt_s[3] + pz0 * a.trgt_s[4]) +
                            i * (int)(B::SDY * a.trgt_s[3] + B::SDZ * a.trgt_s[4]);
          const int tC = (int)a.trgt_s[1];
#pragma unroll
          for (int c = 0; c < NCC; ++c) tvv[c] = c < nc ? tp[c * tC] : 0.f;
        } else {
#pragma unroll
          for (int c = 0; c < NCC; ++c) tvv[c] = c < nc ? 1.f : 0.f;
        }
        float wi[3][NT], gi[3][NT];
        tl_taps<ORD, true>(q0, wi[0], gi[0]);
        tl_taps<ORD, true>(q1, wi[1], gi[1]);
        tl_taps<ORD, true>(q2, wi[2], gi[2]);
        const float *bp = box + oi;
#pragma unroll 1
        // separable, as in TM_SGRAD, on the channel dot products
        for (int kz = 0; kz < NT; ++kz, ++bp) {
          const float *bq = bp;
#pragma unroll 1
          for (int ky = 0; ky < NT; ++ky, bq += DZ) {
            tl_opaque(wi[0]);
            tl_opaque(gi[0]);
            const float wyz = wi[1][0] * wi[2][0], gyz = gi[1][0] * wi[2][0], wgz = wi[1][0] * gi[2][0];
            float sw = 0.f, sd = 0.f;
            if constexpr (IL) {   // 4 interleaved channels (zeros staged past nc): one 16-byte read a tap
              const float4 *vq = (const float4 *)box + (bq - box);   // bq - box counts voxels here
#pragma unroll
              for (int kx = 0; kx < NT; ++kx) {
                const float4 v = vq[kx * DYZ];
                const float dot = v.x * tvv[0] + v.y * tvv[1] + v.z * tvv[2] + v.w * tvv[3];
                sw += wi[0][kx] * dot;
                sd += gi[0][kx] * dot;
              }
            } else {
#pragma unroll
              for (int kx = 0; kx < NT; ++kx) {
                float dot = 0.f;
                // channels >= nc of a partial chunk hold stale LDS: select, not multiply by 0
#pragma unroll
                for (int c = 0; c < NCC; ++c) dot += c < nc ? bq[c * VOL + kx * DYZ] * tvv[c] : 0.f;
                sw += wi[0][kx] * dot;
                sd += gi[0][kx] * dot;
              }
            }
            o0 += sd * wyz;
            o1 += sw * gyz;
            o2 += sw * wgz;
            tl_rot(wi[1]);
            tl_rot(gi[1]);
          }
          tl_rot(wi[2]);
          tl_rot(gi[2]);
        }
        }
        og[R - 1][0] = o0;
        og[R - 1][1] = o1;
        og[R - 1][2] = o2;
      }
    } else {
      if constexpr (WIN) {
      // push / count, orders >= 1: a thread's R z-consecutive samples whose
      // first taps share x and y and advance by one along z with the sample
      // share a window of R + NT - 1 z positions per (x, y) tap row.  Their
      // splats are summed in registers (FP32, at most R terms a position:
      // sample i reaches positions i .. i + NT - 1 of the window) and each
      // position is added to the box by one FP64 atomic: R + NT - 1 atomics
      // a row instead of R * NT.  Sample i's taps start at box element
      // off[i], so samples with equal off[i] - i are exactly those one
      // window at base off[i] - i serves.  A thread's samples split into
      // such groups (anchored at its first sample left); the lanes of a wave
      // loop until every group is done, so any grid is handled, a smooth
      // one in one pass.  A non-finite value reaches only its own taps.
        constexpr int WS = R + NT - 1;
        double *dbox = (double *)box;
        unsigned rem = 0;
#pragma unroll
        for (int i = 0; i < R; ++i) rem |= inb[i] ? 1u << i : 0u;
        while (rem) {
          const int an = __builtin_ctz(rem);
          int base = off[0];
#pragma unroll
          for (int i = 1; i < R; ++i)
            if (an == i) base = off[i] - i;
          unsigned gm = 0;
#pragma unroll
          for (int i = 0; i < R; ++i)
            if (((rem >> i) & 1u) && off[i] - i == base) gm |= 1u << i;
          rem &= ~gm;
          const int hi = 31 - __builtin_clz(gm) + NT - 1;   // last window position reached
          // window position s is box element base + s (an <= s <= hi)
          double *bp = dbox + base;
#pragma unroll 1
          for (int kx = 0; kx < NT; ++kx, bp += DYZ) {
            float xv[R];
#pragma unroll
            for (int i = 0; i < R; ++i) xv[i] = ((gm >> i) & 1u) ? w[i][0][0] * tv[i][0] : 0.f;
#pragma unroll
            for (int ky = 0; ky < NT; ++ky) {
              float acc[WS];
#pragma unroll
              for (int s = 0; s < WS; ++s) acc[s] = 0.f;
#pragma unroll
              for (int i = 0; i < R; ++i) {
                const float cw = xv[i] * w[i][1][ky];
#pragma unroll
                for (int kz = 0; kz < NT; ++kz) acc[i + kz] = fmaf(cw, w[i][2][kz], acc[i + kz]);
              }
#pragma unroll
              for (int s = 0; s < WS; ++s)
                if (s >= an && s <= hi) atomicAdd(bp + ky * DZ + s, (double)acc[s]);
            }
#pragma unroll
            for (int i = 0; i < R; ++i) tl_rot(w[i][0]);
          }
        }
      } else {
      // push / count at order 0: FP64 splats, one sample at a time (rolled):
      // element 0 of the per-sample state is the current sample, rotated
      // after each (R rotations restore it for the next channel chunk)
#pragma unroll 1
      for (int i = 0; i < R; ++i) {
        const bool ii = inb[0];
        const int oi = off[0];
        const float q0 = xq[0][0], q1 = xq[0][1], q2 = xq[0][2];
        double vd[NCC];
#pragma unroll
        for (int c = 0; c < NCC; ++c) vd[c] = (double)tv[0][c];
        tl_rot(inb);
        tl_rot(off);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const float t = xq[0][d];
#pragma unroll
          for (int k = 0; k + 1 < R; ++k) xq[k][d] = xq[k + 1][d];
          xq[R - 1][d] = t;
        }
#pragma unroll
        for (int c = 0; c < NCC; ++c) {
          const float t = tv[0][c];
#pragma unroll
          for (int k = 0; k + 1 < R; ++k) tv[k][c] = tv[k + 1][c];
          tv[R - 1][c] = t;
        }
        if (!ii) continue;
        float wi[3][NT], gi[3][NT];
        tl_taps<ORD>(q0, wi[0], gi[0]);
        tl_taps<ORD>(q1, wi[1], gi[1]);
        tl_taps<ORD>(q2, wi[2], gi[2]);
        double wy[NT], wz[NT];
#pragma unroll
        for (int k = 0; k < NT; ++k) {
          wy[k] = (double)wi[1][k];
          wz[k] = (double)wi[2][k];
        }
        double *bp = (double *)box + oi;
#pragma unroll 1
        for (int kx = 0; kx < NT; ++kx, bp += DYZ) {
#pragma unroll
          for (int ky = 0; ky < NT; ++ky) {
            const double wxy = (double)wi[0][0] * wy[ky];
            double t[NCC];
#pragma unroll
            for (int c = 0; c < NCC; ++c) t[c] = wxy * vd[c];
#pragma unroll
            for (int kz = 0; kz < NT; ++kz) {
              const int o = ky * DZ + kz;
#pragma unroll
              for (int c = 0; c < NCC; ++c) atomicAdd(bp + c * VOL + o, t[c] * wz[kz]);
            }
          }
          tl_rot(wi[0]);
        }
      }
      }

      // ---- add the box to the output, folded and signed; clear it for
      // the next chunk (the splats never leave the ext corner) ----
      __syncthreads();
      float *out = a.out + (int)(n * a.out_s[0] + (int64_t)c0 * oC);
      double *dbox = (double *)box;
      const int e0 = ext[0] - 1;
#pragma unroll
      for (int j = 0; j < B::NPP; ++j) {
        if (pl.yz_lds[j] < 0) continue;
#pragma unroll 1
        for (int c = 0; c < nc; ++c) {
          // a whole x row of the box read (and cleared) at once, then its
          // atomics issued back to back
          double *bq = dbox + c * VOL + pl.yz_lds[j];
          float fv[B::DX];   // the row, rounded to float (0 where the box is empty)
#pragma unroll
          for (int bx = 0; bx < B::DX; ++bx) fv[bx] = (float)bq[bx * DYZ];
#pragma unroll
          for (int bx = 0; bx < B::DX; ++bx) bq[bx * DYZ] = 0.0;
          float *op = out + c * oC + pl.yz_ofs[j];
#pragma unroll
          for (int bx = 0; bx < B::DX; ++bx) {
            const float sg = fsg[bx] * pl.yz_sg[j];
            if (bx <= e0 && fv[bx] != 0.f && sg != 0.f) atomic_add(op + fofs[bx], fv[bx] * sg);
            if ((bx & 3) == 3) __builtin_amdgcn_sched_barrier(0);   // bound the addresses held at once
          }
        }
      }
    }
  }
  if (MODE == TM_GGRAD || MODE == TM_GGRAD1) {
    float *gq = a.gout + (int)(n * a.gout_s[0] + px * a.gout_s[1] + py * a.gout_s[2] + pz0 * a.gout_s[3]);
    const int gZ = (int)(B::SDY * a.gout_s[2] + B::SDZ * a.gout_s[3]), gDo = (int)a.gout_s[4];
#pragma unroll
    for (int i = 0; i < R; ++i)
      if (valid[i]) {
        gq[i * gZ] = og[i][0];
        gq[i * gZ + gDo] = og[i][1];
        gq[i * gZ + 2 * gDo] = og[i][2];
      }
  }
}

// Tiles whose box does not fit, by the direct per-sample routine.  With a
// list (fb: count, then the tile ids that tile_kernel found too scattered)
// a fixed grid walks the list; without one, every tile repeats tile_kernel's
// bounding-box test (bit-identical, so every tile is handled by exactly one
// of the two kernels).  Kept apart from tile_kernel so that its register
// demand does not lower that kernel's occupancy.
template <int ORD, int MODE>
__global__ __launch_bounds__((TileBox<ORD, !TileMeta<MODE>::gather>::NTH)) void tile_fallback(PPArgs<float> a, int ntx, int nty, int ntz,
                                                          int64_t nblocks, const int *fb) {
  constexpr int NT = ORD + 1;
  constexpr int FL = TileFlags<MODE>::value;
  using B = TileBox<ORD, !TileMeta<MODE>::gather>;
  constexpr int R = B::R;
  const int X = (int)a.tX, Y = (int)a.tY, Z = (int)a.tZ;
  int sx, sy, sz0;
  tl_thread<B>(sx, sy, sz0);
  if (fb) {
    const int cnt = fb[0];
    for (int64_t j = blockIdx.x; j < cnt; j += gridDim.x) {
      int64_t r = fb[1 + j];
      const int tz = (int)(r % ntz); r /= ntz;
      const int ty = (int)(r % nty); r /= nty;
      const int tx = (int)(r % ntx);
      const int n = (int)(r / ntx);
      const int px = tx * B::TX + sx, py = ty * B::TY + sy, pz0 = tz * B::TZ + sz0;
#pragma unroll 1
      for (int i = 0; i < R; ++i)
        if (px < X && py + i * B::SDY < Y && pz0 + i * B::SDZ < Z)
          pp_sample<float, int32_t, ORD, FL, true>(a, n, px, py + i * B::SDY, pz0 + i * B::SDZ);
    }
    return;
  }
  __shared__ int red[B::NW][6];
  int64_t r = xcd_block(nblocks);
  const int tz = (int)(r % ntz); r /= ntz;
  const int ty = (int)(r % nty); r /= nty;
  const int tx = (int)(r % ntx);
  const int n = (int)(r / ntx);
  const int sz[3] = {(int)a.sX, (int)a.sY, (int)a.sZ};
  const int px = tx * B::TX + sx, py = ty * B::TY + sy, pz0 = tz * B::TZ + sz0;
  bool valid[R], inb[R];
  int mn[3] = {INT_MAX, INT_MAX, INT_MAX}, mx[3] = {INT_MIN, INT_MIN, INT_MIN};
  const float tol = (a.extrap & 2) ? 0.55f : 0.05f;
  const int gD = (int)a.grid_s[4];
  const int gS = (int)(B::SDY * a.grid_s[2] + B::SDZ * a.grid_s[3]);
  const float *gp0 = a.grid + (int)(n * a.grid_s[0] + px * a.grid_s[1] + py * a.grid_s[2] + pz0 * a.grid_s[3]);
#pragma unroll
  for (int i = 0; i < R; ++i) {
    valid[i] = px < X && py + i * B::SDY < Y && pz0 + i * B::SDZ < Z;
    inb[i] = false;
    if (valid[i]) {
      const float *gp = gp0 + i * gS;
      const float xs[3] = {gp[0], gp[gD], gp[2 * gD]};
      bool ok = true;
      if (!(a.extrap & 1))
#pragma unroll
        for (int d = 0; d < 3; ++d) ok = ok && (xs[d] >= -tol && xs[d] < (float)(sz[d] - 1) + tol);
      inb[i] = ok;
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
    for (int wv = 1; wv < B::NW; ++wv) {
      m0 = min(m0, red[wv][d]);
      m1 = max(m1, red[wv][3 + d]);
    }
    if (m0 > m1) return;   // no in-bound sample: handled by tile_kernel
    const int64_t e = (int64_t)m1 - (int64_t)m0 + NT;
    fits = fits && e <= (d == 0 ? B::DX : d == 1 ? B::DY : B::DZ);
  }
  if (fits) return;
#pragma unroll 1
  for (int i = 0; i < R; ++i)
    if (valid[i]) pp_sample<float, int32_t, ORD, FL, true>(a, n, px, py + i * B::SDY, pz0 + i * B::SDZ);
}

// Tiles that tile_kernel listed (taps too spread for the static box): each
// listed tile is cut into sub-tiles of
// 4 x 8 x 8 samples (one per thread of a 256-thread work-group), and each
// sub-tile stages / accumulates an LDS box sized from its own first-tap
// spread (runtime strides, up to DYN_MAXE elements per channel).  Strongly
// deformed grids (large rotations, zooms, displacements with a large
// Jacobian) keep most of the tile path's reuse this way: a sub-tile's box is
// a few times its sample count instead of the per-sample kernel's 64 taps
// x channels of scattered global traffic.  A sub-tile whose box is still
// too large runs the per-sample routine.  Same weights, out-of-bound rule,
// box fill and FP64 accumulation as tile_kernel (pull / push / count);
// other modes take the per-sample kernel for listed tiles.
#define DYN_TX 4
#define DYN_TY 8
#define DYN_TZ 8
#define DYN_NTH 256
#define DYN_MAXE 8192        // box elements per channel (32 KB of floats, 64 KB of doubles)
#define DYN_MAXAX 128        // box extent per axis

template <int ORD, int MODE>
__global__ __launch_bounds__(DYN_NTH) void tile_dyn(PPArgs<float> a, int ntx, int nty, int ntz, const int *fb) {
  constexpr int NT = ORD + 1;
  constexpr bool GATHER = MODE == TM_PULL;
  constexpr int FL = TileFlags<MODE>::value;
  using B = TileBox<ORD, !GATHER>;
  constexpr int NCC = GATHER ? 2 : 1;
  constexpr int NW = DYN_NTH / 64;
  static_assert(DYN_TX * DYN_TY * DYN_TZ == DYN_NTH, "one sample per thread");
  static_assert(B::TX % DYN_TX == 0 && B::TY % DYN_TY == 0 && B::TZ % DYN_TZ == 0, "sub-tiles tile the tile");
  constexpr int NSX = B::TX / DYN_TX, NSY = B::TY / DYN_TY, NSZ = B::TZ / DYN_TZ;
  // gathers: NCC FP32 channels; scatters: one channel of doubles
  __shared__ __align__(16) float box[2 * DYN_MAXE];
  __shared__ int fofs[3 * DYN_MAXAX];
  __shared__ float fsg[3 * DYN_MAXAX];
  __shared__ int red[NW][6];

  const int X = (int)a.tX, Y = (int)a.tY, Z = (int)a.tZ;
  const int sz[3] = {(int)a.sX, (int)a.sY, (int)a.sZ};
  const float tol = (a.extrap & 2) ? 0.55f : 0.05f;
  const int lx = threadIdx.x / (DYN_TY * DYN_TZ), ly = (threadIdx.x / DYN_TZ) % DYN_TY, lz = threadIdx.x % DYN_TZ;
  const int cnt = fb[0];
  for (int64_t j = blockIdx.x; j < (int64_t)cnt * NSX * NSY * NSZ; j += gridDim.x) {
    const int sub = (int)(j % (NSX * NSY * NSZ));
    int64_t r = fb[1 + j / (NSX * NSY * NSZ)];
    const int tz = (int)(r % ntz); r /= ntz;
    const int ty = (int)(r % nty); r /= nty;
    const int tx = (int)(r % ntx);
    const int n = (int)(r / ntx);
    const int px = tx * B::TX + (sub / (NSY * NSZ)) * DYN_TX + lx;
    const int py = ty * B::TY + ((sub / NSZ) % NSY) * DYN_TY + ly;
    const int pz = tz * B::TZ + (sub % NSZ) * DYN_TZ + lz;
    const bool valid = px < X && py < Y && pz < Z;
    bool inb = false;
    float xq[3] = {0.f, 0.f, 0.f};
    int lo[3] = {0, 0, 0};
    int mn[3] = {INT_MAX, INT_MAX, INT_MAX}, mx[3] = {INT_MIN, INT_MIN, INT_MIN};
    if (valid) {
      const float *gp = a.grid + (int64_t)n * a.grid_s[0] + (int64_t)px * a.grid_s[1] + (int64_t)py * a.grid_s[2] +
                        (int64_t)pz * a.grid_s[3];
      const int gD = (int)a.grid_s[4];
      xq[0] = gp[0]; xq[1] = gp[gD]; xq[2] = gp[2 * gD];
      bool ok = true;
      if (!(a.extrap & 1))
#pragma unroll
        for (int d = 0; d < 3; ++d) ok = ok && (xq[d] >= -tol && xq[d] < (float)(sz[d] - 1) + tol);
      inb = ok;
      if (ok) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          float wt[NT], gt[NT];
          lo[d] = tl_taps<ORD>(xq[d], wt, gt);
          mn[d] = lo[d];
          mx[d] = lo[d];
        }
      }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      mn[d] = wave_min(mn[d]);
      mx[d] = wave_max(mx[d]);
    }
    __syncthreads();   // previous sub-tile done with red / box / tables
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        red[threadIdx.x >> 6][d] = mn[d];
        red[threadIdx.x >> 6][3 + d] = mx[d];
      }
    }
    __syncthreads();
    int ext[3];
    bool none = false, fits = true;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      int m0 = red[0][d], m1 = red[0][3 + d];
#pragma unroll
      for (int wv = 1; wv < NW; ++wv) {
        m0 = min(m0, red[wv][d]);
        m1 = max(m1, red[wv][3 + d]);
      }
      mn[d] = m0;
      none = none || m0 > m1;
      const int64_t e = (int64_t)m1 - (int64_t)m0 + NT;
      fits = fits && e <= DYN_MAXAX;
      ext[d] = fits && !none ? (int)e : 1;
    }
    if (none) {   // no in-bound sample: pull writes zeros
      if (MODE == TM_PULL && valid) {
        float *op = a.out + (int64_t)n * a.out_s[0] + (int64_t)px * a.out_s[2] + (int64_t)py * a.out_s[3] +
                    (int64_t)pz * a.out_s[4];
        for (int c = 0; c < (int)a.C; ++c) op[(int64_t)c * a.out_s[1]] = 0.f;
      }
      continue;
    }
    const int DZ = ext[2], DYZ = ext[1] * ext[2];
    fits = fits && (int64_t)ext[0] * DYZ <= DYN_MAXE;
    if (!fits) {   // per-sample routine for this sub-tile
      if (valid) pp_sample<float, int32_t, ORD, FL, true>(a, n, px, py, pz);
      continue;
    }
    const int VOL = ext[0] * DYZ;
    const float rz = 1.f / (float)DZ, ry = 1.f / (float)ext[1];
    // folded offsets and signs of the box positions along each axis
    {
      const int64_t *st = GATHER ? a.src_s : a.out_s;
      for (int i = threadIdx.x; i < ext[0] + ext[1] + ext[2]; i += DYN_NTH) {
        const int d = i < ext[0] ? 0 : (i < ext[0] + ext[1] ? 1 : 2);
        const int b = i - (d == 0 ? 0 : (d == 1 ? ext[0] : ext[0] + ext[1]));
        int sg = 0;
        const int jj = fold<int>(a.bound[d], mn[d] + b, sz[d], sg);
        fofs[i] = jj * (int)st[2 + d];
        fsg[i] = (float)sg;
      }
    }
    // this sample's box offset and weights
    float w[3][NT], g[3][NT];
    int off = 0;
    if (inb) {
#pragma unroll
      for (int d = 0; d < 3; ++d) tl_taps<ORD>(xq[d], w[d], g[d]);
      off = (lo[0] - mn[0]) * DYZ + (lo[1] - mn[1]) * DZ + (lo[2] - mn[2]);
    }
    __syncthreads();
    if (GATHER) {
      const float *src = a.src + (int64_t)n * a.src_s[0];
      float *op = a.out + (int64_t)n * a.out_s[0] + (int64_t)px * a.out_s[2] + (int64_t)py * a.out_s[3] +
                  (int64_t)pz * a.out_s[4];
      for (int c0 = 0; c0 < (int)a.C; c0 += NCC) {
        const int nc = min(NCC, (int)a.C - c0);
        if (c0) __syncthreads();
        // stage the folded, signed box of channels c0 .. c0 + nc
        for (int q = threadIdx.x; q < VOL; q += DYN_NTH) {
          const int row = tl_div(q, DZ, rz), bz = q - row * DZ;
          const int bx = tl_div(row, ext[1], ry), by = row - bx * ext[1];
          const float sg = fsg[bx] * fsg[ext[0] + by] * fsg[ext[0] + ext[1] + bz];
          const int so = fofs[bx] + fofs[ext[0] + by] + fofs[ext[0] + ext[1] + bz];
#pragma unroll
          for (int c = 0; c < NCC; ++c)
            if (c < nc) box[c * DYN_MAXE + q] = sg != 0.f ? src[(int64_t)(c0 + c) * a.src_s[1] + so] * sg : 0.f;
        }
        __syncthreads();
        if (valid) {
          float acc[NCC];
#pragma unroll
          for (int c = 0; c < NCC; ++c) acc[c] = 0.f;
          if (inb) {
            const float *bp = box + off;
#pragma unroll 1
            for (int kx = 0; kx < NT; ++kx, bp += DYZ) {
#pragma unroll
              for (int ky = 0; ky < NT; ++ky) {
                const float wxy = w[0][kx] * w[1][ky];
                const float *bq = bp + ky * DZ;
#pragma unroll
                for (int kz = 0; kz < NT; ++kz) {
                  const float ww = wxy * w[2][kz];
#pragma unroll
                  for (int c = 0; c < NCC; ++c) acc[c] += bq[c * DYN_MAXE + kz] * ww;
                }
              }
            }
          }
#pragma unroll
          for (int c = 0; c < NCC; ++c)
            if (c < nc) op[(int64_t)(c0 + c) * a.out_s[1]] = acc[c];
        }
      }
    } else {
      double *dbox = (double *)box;
      for (int q = threadIdx.x; q < VOL; q += DYN_NTH) dbox[q] = 0.0;
      const float *tq = MODE == TM_PUSH ? a.trgt + (int64_t)n * a.trgt_s[0] + (int64_t)px * a.trgt_s[2] +
                                              (int64_t)py * a.trgt_s[3] + (int64_t)pz * a.trgt_s[4]
                                        : nullptr;
      float *out = a.out + (int64_t)n * a.out_s[0];
      const int C = MODE == TM_COUNT ? 1 : (int)a.C;
      for (int c = 0; c < C; ++c) {
        const float v = !inb ? 0.f : (MODE == TM_PUSH ? tq[(int64_t)c * a.trgt_s[1]] : 1.f);
        __syncthreads();   // box cleared (first channel) / flushed (later ones)
        if (inb && v != 0.f) {
          const double vd = (double)v;
          double *bp = dbox + off;
#pragma unroll 1
          for (int kx = 0; kx < NT; ++kx, bp += DYZ) {
#pragma unroll
            for (int ky = 0; ky < NT; ++ky) {
              const double t = (double)w[0][kx] * (double)w[1][ky] * vd;
              double *bq = bp + ky * DZ;
#pragma unroll
              for (int kz = 0; kz < NT; ++kz) atomicAdd(bq + kz, t * (double)w[2][kz]);
            }
          }
        }
        __syncthreads();
        // flush: one FP32 global atomic per non-zero box element; clear it
        float *oc = out + (int64_t)c * a.out_s[1];
        for (int q = threadIdx.x; q < VOL; q += DYN_NTH) {
          const double dv = dbox[q];
          dbox[q] = 0.0;
          if (dv == 0.0) continue;
          const int row = tl_div(q, DZ, rz), bz = q - row * DZ;
          const int bx = tl_div(row, ext[1], ry), by = row - bx * ext[1];
          const float sg = fsg[bx] * fsg[ext[0] + by] * fsg[ext[0] + ext[1] + bz];
          if (sg != 0.f)
            atomic_add(oc + fofs[bx] + fofs[ext[0] + by] + fofs[ext[0] + ext[1] + bz], (float)dv * sg);
        }
      }
    }
  }
}

}  // namespace ni_amd

#include "pushpull_bin.h"   // the binned second pass (uses the helpers above)

namespace ni_amd {

// tile_kernel, then tile_fallback on the tiles it listed as too scattered
// (list in stream-ordered scratch; without it, tile_fallback repeats the
// bounding-box test on every tile).
template <int ORD, int MODE>
static void launch_tile(const PPArgs<float> &a, hipStream_t st) {
  using B = TileBox<ORD, !TileMeta<MODE>::gather>;
  const int ntx = (int)((a.tX + B::TX - 1) / B::TX);
  const int nty = (int)((a.tY + B::TY - 1) / B::TY);
  const int ntz = (int)((a.tZ + B::TZ - 1) / B::TZ);
  const int64_t nb = a.N * (int64_t)ntx * nty * ntz;
  const dim3 g((unsigned)nb), b(B::NTH);
  int *fb = nullptr;
  const bool own = a.scratch == nullptr;
  if (!own) {
    fb = a.scratch;   // the caller checked nb + 1 <= a.scratch_cap
    hipLaunchKernelGGL(fill_zero<int>, dim3(1), dim3(NI_BLOCK), 0, st, fb, (int64_t)1);
  } else if (scratch_alloc((void **)&fb, sizeof(int) * (size_t)(nb + 1), st) != hipSuccess) {
    (void)hipGetLastError();
    fb = nullptr;
  } else {
    hipLaunchKernelGGL(fill_zero<int>, dim3(1), dim3(NI_BLOCK), 0, st, fb, (int64_t)1);
  }
  if constexpr (MODE == TM_COUNT) {
    hipLaunchKernelGGL((tile_kernel<ORD, MODE, 1>), g, b, 0, st, a, ntx, nty, ntz, nb, fb);
  } else {
    if (a.C == 1)
      hipLaunchKernelGGL((tile_kernel<ORD, MODE, 1>), g, b, 0, st, a, ntx, nty, ntz, nb, fb);
    else
      hipLaunchKernelGGL((tile_kernel<ORD, MODE, B::MAX_NCC>), g, b, 0, st, a, ntx, nty, ntz, nb, fb);
  }
  if (fb) {
    if constexpr (MODE == TM_PULL || MODE == TM_PUSH || MODE == TM_COUNT) {
      // listed tiles: binned by first tap (when the call may allocate its
      // bins and has at least BIN_MIN_SAMPLES samples: below that, the
      // binned pass's six launches cost more than the sub-tile kernel's
      // one), else sub-tiles with boxes sized to their spread
      const bool big = a.N * a.tX * a.tY * a.tZ >= BIN_MIN_SAMPLES;
      if (!own || !big || !launch_binned<ORD, MODE>(a, ntx, nty, ntz, nb, fb, st)) {
        constexpr int nsub = (B::TX / DYN_TX) * (B::TY / DYN_TY) * (B::TZ / DYN_TZ);
        const int64_t want = nb * nsub;
        const unsigned gd = (unsigned)(want < 4096 ? want : 4096);
        hipLaunchKernelGGL((tile_dyn<ORD, MODE>), dim3(gd), dim3(DYN_NTH), 0, st, a, ntx, nty, ntz, (const int *)fb);
      }
    } else {
      // grid_grad's listed tiles take the binned pass in calls that may
      // allocate and have BIN_MIN_SAMPLES samples; the other gradient modes,
      // and grid_grad otherwise, run pp_sample over the listed tiles
      bool done = false;
      if constexpr (MODE == TM_SGRAD)
        done = own && a.N * a.tX * a.tY * a.tZ >= BIN_MIN_SAMPLES && launch_binned<ORD, MODE>(a, ntx, nty, ntz, nb, fb, st);
      if (!done) {
        const unsigned gf = (unsigned)(nb < 2048 ? nb : 2048);
        hipLaunchKernelGGL((tile_fallback<ORD, MODE>), dim3(gf), b, 0, st, a, ntx, nty, ntz, nb, (const int *)fb);
      }
    }
    if (own) (void)hipFreeAsync(fb, st);
  } else {
    hipLaunchKernelGGL((tile_fallback<ORD, MODE>), g, b, 0, st, a, ntx, nty, ntz, nb, (const int *)nullptr);
  }
}

}  // namespace ni_amd
