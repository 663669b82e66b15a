// BF16 kernels of the learned preconditioner's 3-D U-Net (no counterpart in
// the reference): 3x3x3 convolutions with FP32 accumulation on the gfx950
// matrix cores, their backward passes, and the pooling / upsampling / concat
// / packing kernels around them.
//
// Layout: activations channels-last [N, D, H, W, C] in bf16, C a multiple of
// 16 (the 9 input channels are padded to 16 by the packing kernel).
// Weights of a layer CI -> CO: [CO][KP] bf16 with K = tap * CI + ci,
// tap = 9 dx + 3 dy + dz (dx, dy, dz in 0..2 <-> offsets -1..1), KP = 27 CI
// rounded up to 32 and zero-padded.  Zero padding of 1 voxel, stride 1.
//
// Convolution as an implicit GEMM on v_mfma_f32_16x16x32_bf16:
//   rows (M) = 16 consecutive z positions of one (n, x, y) row,
//   cols (N) = 16 output channels, depth (K) = 32 (tap, ci) pairs.
// A lane supplies row (lane % 16) and the 8 consecutive K elements
// 8 * (lane / 16) .. +7 of its operand, i.e. 8 consecutive channels of one
// tap of one voxel (one 16-byte load); accumulator element r of a lane is
// row 4 * (lane / 16) + r, column lane % 16.  A wave computes 4 row blocks
// (64 positions) x all output-channel blocks.
//
// Backward data is the same kernel run on the activation gradient with the
// flipped, transposed weights (prepared on the host side); backward weights
// reduce over positions with the same instruction (rows = output channels,
// cols = (tap, ci), depth = 32 positions), FP32 partial sums added with
// atomics.
#include <cstdint>

#include "common.h"
#include "nitorch_amd.h"

namespace ni_amd {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

#define UN_BLOCK 256
#define UN_MB 4          // 16-position row blocks per wave
#define UN_SLOPE 0.1f    // leaky ReLU negative slope

// XCD-contiguous logical block id (the 8 XCDs of an MI355X each get one
// contiguous range of blocks, so the +-1 x/y neighbour rows of a convolution
// are mostly in the same L2).
__device__ __forceinline__ int64_t un_block() {
  const int64_t nb = gridDim.x, b = blockIdx.x;
  const int64_t xcd = b % 8, local = b / 8, q = nb / 8, r = nb % 8;
  return xcd * q + (xcd < r ? xcd : r) + local;
}

template <int CI>
struct KPad {
  static constexpr int value = (27 * CI + 31) / 32 * 32;
};

__device__ __forceinline__ bf16x8 zero8() {
  bf16x8 z;
#pragma unroll
  for (int j = 0; j < 8; ++j) z[j] = (bf16)0.f;
  return z;
}

// ------------------------------------------------------------------------
// y = act(conv(x, w) + b).  ACT: 0 none, 1 leaky ReLU.  COS output channels
// are stored (COS <= CO; a 3-channel head is computed with CO = 16 and zero
// weights in the padding), as bf16 or fp32.
// ------------------------------------------------------------------------
template <int CI, int CO, int COS, int ACT, bool OUTF32>
__global__ __launch_bounds__(UN_BLOCK) void conv3_fwd(const bf16 *__restrict__ X, const bf16 *__restrict__ Wk,
                                                      const float *__restrict__ bias, void *__restrict__ Yv,
                                                      int N, int D, int H, int W) {
  constexpr int KP = KPad<CI>::value;
  constexpr int CB = CO / 16;
  static_assert(CI % 16 == 0 && CO % 16 == 0 && COS <= CO, "channel counts");
  const int WB = (W + 15) / 16;
  const int64_t total = (int64_t)N * D * H * WB;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int row = lane & 15, ks = lane >> 4;
  const int64_t pb0 = (un_block() * (UN_BLOCK / 64) + wave) * UN_MB;

  int px[UN_MB], py[UN_MB], pz[UN_MB], pn[UN_MB];
  bool pv[UN_MB];
#pragma unroll
  for (int m = 0; m < UN_MB; ++m) {
    const int64_t pb = pb0 + m;
    pv[m] = pb < total;
    int64_t t = pv[m] ? pb : 0;
    const int zb = (int)(t % WB); t /= WB;
    py[m] = (int)(t % H); t /= H;
    px[m] = (int)(t % D);
    pn[m] = (int)(t / D);
    pz[m] = zb * 16 + row;
  }

  f32x4 acc[UN_MB][CB];
#pragma unroll
  for (int m = 0; m < UN_MB; ++m)
#pragma unroll
    for (int c = 0; c < CB; ++c) acc[m][c] = (f32x4){0.f, 0.f, 0.f, 0.f};

#pragma unroll 2
  for (int kc = 0; kc < KP / 32; ++kc) {
    const int k0 = kc * 32 + ks * 8;
    const int tap = k0 / CI, ci = k0 - tap * CI;
    const int dx = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dz = tap % 3 - 1;
    bf16x8 b[CB];
#pragma unroll
    for (int c = 0; c < CB; ++c) b[c] = *(const bf16x8 *)(Wk + (size_t)(c * 16 + row) * KP + k0);
#pragma unroll
    for (int m = 0; m < UN_MB; ++m) {
      const int xx = px[m] + dx, yy = py[m] + dy, zz = pz[m] + dz;
      bf16x8 a = zero8();
      if (pv[m] && tap < 27 && xx >= 0 && xx < D && yy >= 0 && yy < H && zz >= 0 && zz < W)
        a = *(const bf16x8 *)(X + ((((size_t)pn[m] * D + xx) * H + yy) * W + zz) * CI + ci);
#pragma unroll
      for (int c = 0; c < CB; ++c) acc[m][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[c], acc[m][c], 0, 0, 0);
    }
  }

#pragma unroll
  for (int m = 0; m < UN_MB; ++m) {
    if (!pv[m]) continue;
    const int zb0 = pz[m] - row;
#pragma unroll
    for (int c = 0; c < CB; ++c) {
      const int co = c * 16 + row;
      if (co >= COS) continue;
      const float bb = bias ? bias[co] : 0.f;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int z = zb0 + ks * 4 + r;
        if (z >= W) continue;
        float v = acc[m][c][r] + bb;
        if (ACT == 1) v = v > 0.f ? v : UN_SLOPE * v;
        const size_t o = ((((size_t)pn[m] * D + px[m]) * H + py[m]) * W + z) * COS + co;
        if (OUTF32) ((float *)Yv)[o] = v;
        else ((bf16 *)Yv)[o] = (bf16)v;
      }
    }
  }
}

// ------------------------------------------------------------------------
// The same convolution with the input staged in LDS.  A work-group owns a
// tile of one (n, x) plane: UN_TY = 8 rows of y by UN_TZ = 32 positions of
// z; wave w computes rows 2w and 2w + 1, each as two 16-position row blocks
// (m = 2 * (row - 2w) + z block).  The input voxels the tile's taps reach,
// x - 1 .. x + 1 by (UN_TY + 2) x (UN_TZ + 2), are staged in LDS (zero
// outside the volume, which is the convolution's zero padding) and the 27
// taps read them from LDS: each input voxel leaves L2 about 4 times per
// layer instead of 27.  A K-slice -> LDS-offset table makes each 32-deep K
// chunk one table read and one add; the four A fragments of a chunk are LDS
// reads at immediate offsets.  Fragments, K order and accumulation order
// are those of conv3_fwd, so the kernels give identical results.
// ------------------------------------------------------------------------
#define UN_TY 8
#define UN_TZ 32

// ------------------------------------------------------------------------
// Staged convolution for wide inputs (CI a multiple of 32, 64 and 96 here): the
// K order is tap-major with x outermost (tap = 9 dx + 3 dy + dz), so the 9 CI
// K values of one dx are whole chunks and need only one x-plane of the box.
// The work-group stages the three planes one after the other, each
// (UN_TY + 2) x (UN_TZ + 2) voxels, the next one's loads in flight while
// the current one is consumed; tiles, fragments and accumulation order are
// those of conv3_fwd (identical results).
// ------------------------------------------------------------------------
template <int CI>
struct PlaneBytes {
  static constexpr int box = (UN_TY + 2) * (UN_TZ + 2) * CI * 2;
  static constexpr int value = box + 9 * CI / 8 * 4;    // + the offset table of one plane's K slices
};

template <int CI, int CO, int COS, int ACT, bool OUTF32>
__global__ __launch_bounds__(UN_BLOCK) void conv3_fwd_plane(const bf16 *__restrict__ X, const bf16 *__restrict__ Wk,
                                                            const float *__restrict__ bias, void *__restrict__ Yv,
                                                            int N, int D, int H, int W) {
  extern __shared__ __align__(16) unsigned char un_smem[];
  bf16 *box = (bf16 *)un_smem;
  int *tab = (int *)(un_smem + PlaneBytes<CI>::box);
  constexpr int KP = KPad<CI>::value, NKP = 9 * CI / 32;    // K chunks per plane
  constexpr int CB = CO / 16;
  constexpr int BY = UN_TY + 2, BZ = UN_TZ + 2, V8 = CI / 8;
  constexpr int RC = BZ * V8, NCH = BY * RC;
  constexpr int NIT = (NCH + UN_BLOCK - 1) / UN_BLOCK;
  static_assert(CI % 32 == 0 && KP == 27 * CI && CO % 16 == 0 && COS <= CO, "channel counts");
  const int YT = (H + UN_TY - 1) / UN_TY, ZT = (W + UN_TZ - 1) / UN_TZ;
  int64_t t = un_block();
  const int zt = (int)(t % ZT); t /= ZT;
  const int yt = (int)(t % YT); t /= YT;
  const int x = (int)(t % D);
  const int n = (int)(t / D);
  if (n >= N) return;    // uniform over the work-group
  const int y0 = yt * UN_TY - 1, z0 = zt * UN_TZ - 1;

  // K slice of one plane (tap 3 dy + dz, channel ci) -> plane offset from (-1, -1)
  for (int i = threadIdx.x; i < 9 * CI / 8; i += UN_BLOCK) {
    const int tap = i * 8 / CI, ci = i * 8 - tap * CI;
    tab[i] = (tap / 3 * BZ + tap % 3) * CI + ci;
  }
  uint4 v[NIT];
  auto fetch = [&](int dx) {
    const int xx = x - 1 + dx;
    const bool xok = xx >= 0 && xx < D;
#pragma unroll
    for (int u = 0; u < NIT; ++u) {
      const int i = u * UN_BLOCK + threadIdx.x;
      const int by = i / RC, j = i - by * RC;
      const int yy = y0 + by, zz = z0 + j / V8;
      v[u] = (uint4){0u, 0u, 0u, 0u};
      if (xok && i < NCH && yy >= 0 && yy < H && zz >= 0 && zz < W)
        v[u] = *(const uint4 *)(X + ((((int64_t)n * D + xx) * H + yy) * W + z0) * CI + j * 8);
    }
  };
  fetch(0);

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int row = lane & 15, ks = lane >> 4;
  const bf16 *abase = box + ((2 * wave) * BZ + row) * CI;
  const int *tl = tab + ks;
  const bf16 *wbase = Wk + (size_t)row * KP + ks * 8;
  f32x4 acc[UN_MB][CB];
#pragma unroll
  for (int m = 0; m < UN_MB; ++m)
#pragma unroll
    for (int c = 0; c < CB; ++c) acc[m][c] = (f32x4){0.f, 0.f, 0.f, 0.f};

#pragma unroll 1
  for (int dx = 0; dx < 3; ++dx) {
    if (dx) __syncthreads();    // the previous plane is consumed
#pragma unroll
    for (int u = 0; u < NIT; ++u) {
      const int i = u * UN_BLOCK + threadIdx.x;
      if (i < NCH) *(uint4 *)(box + i * 8) = v[u];
    }
    __syncthreads();
    if (dx < 2) fetch(dx + 1);
    const bf16 *wp = wbase + dx * NKP * 32;
#pragma unroll
    for (int kp = 0; kp < NKP; ++kp) {
      const bf16 *ap = abase + tl[kp * 4];
      bf16x8 b[CB];
#pragma unroll
      for (int c = 0; c < CB; ++c) b[c] = *(const bf16x8 *)(wp + (size_t)c * 16 * KP + kp * 32);
#pragma unroll
      for (int m = 0; m < UN_MB; ++m) {
        const bf16x8 a = *(const bf16x8 *)(ap + ((m >> 1) * BZ + (m & 1) * 16) * CI);
#pragma unroll
        for (int c = 0; c < CB; ++c)
          acc[m][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[c], acc[m][c], 0, 0, 0);
      }
    }
  }

  const int yb = yt * UN_TY + 2 * wave, zb = zt * UN_TZ + ks * 4;
  const size_t obase = ((((size_t)n * D + x) * H + yb) * W + zb) * COS;
#pragma unroll
  for (int c = 0; c < CB; ++c) {
    const int co = c * 16 + row;
    if (co >= COS) continue;
    const float bb = bias ? bias[co] : 0.f;
#pragma unroll
    for (int m = 0; m < UN_MB; ++m) {
      if (yb + (m >> 1) >= H) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if (zb + (m & 1) * 16 + r >= W) continue;
        float val = acc[m][c][r] + bb;
        if (ACT == 1) val = val > 0.f ? val : UN_SLOPE * val;
        const size_t o = obase + ((size_t)(m >> 1) * W + (m & 1) * 16 + r) * COS + co;
        if (OUTF32) ((float *)Yv)[o] = val;
        else ((bf16 *)Yv)[o] = (bf16)val;
      }
    }
  }
}

// ------------------------------------------------------------------------
// Staged convolution for inputs of up to 48 channels, in two x-plane slots
// (16 | CI): the K values of x-offsets 0 and 1 end at chunk 9 CI / 16, so
// the work-group stages planes x - 1 and x into slots 0 and 1, consumes
// those chunks, then stages plane x + 1 (its loads in flight meanwhile)
// into slot 0 for the rest: two thirds of the whole box's LDS (65 KB at
// CI = 48), so twice the work-groups per CU at CI = 48.
// ------------------------------------------------------------------------
template <int CI>
struct SlotBytes {
  static constexpr int plane = (UN_TY + 2) * (UN_TZ + 2) * CI;    // elements
  static constexpr int value = 2 * plane * 2 + KPad<CI>::value / 8 * 4;
};

// C1 > 0: the input is concat(nearest 2x upsampling of X [N, D/2, H/2, W/2,
// C1], X2 [N, D, H, W, CI - C1]) (up2_concat_fwd's output, D, H, W even),
// staged from the two sources without materialising it.
template <int CI, int CO, int COS, int ACT, bool OUTF32, int C1 = 0>
__global__ __launch_bounds__(UN_BLOCK) void conv3_fwd_2slot(const bf16 *__restrict__ X, const bf16 *__restrict__ Wk,
                                                            const float *__restrict__ bias, void *__restrict__ Yv,
                                                            int N, int D, int H, int W,
                                                            const bf16 *__restrict__ X2 = nullptr) {
  extern __shared__ __align__(16) unsigned char un_smem[];
  constexpr int PE = SlotBytes<CI>::plane;
  bf16 *box = (bf16 *)un_smem;
  int *tab = (int *)(un_smem + 2 * PE * 2);
  constexpr int KP = KPad<CI>::value, NKC = KP / 32, NKA = 9 * CI / 16;    // chunks of x-offsets 0, 1
  constexpr int CB = CO / 16;
  constexpr int BY = UN_TY + 2, BZ = UN_TZ + 2, V8 = CI / 8;
  constexpr int RC = BZ * V8, NCH = BY * RC;
  constexpr int NIT = (NCH + UN_BLOCK - 1) / UN_BLOCK;
  constexpr bool PAD = KP != 27 * CI;
  static_assert(CI % 16 == 0 && CO % 16 == 0 && COS <= CO, "channel counts");
  static_assert(!PAD || KP - 27 * CI == 16, "K padding is half a chunk");
  const int YT = (H + UN_TY - 1) / UN_TY, ZT = (W + UN_TZ - 1) / UN_TZ;
  int64_t t = un_block();
  const int zt = (int)(t % ZT); t /= ZT;
  const int yt = (int)(t % YT); t /= YT;
  const int x = (int)(t % D);
  const int n = (int)(t / D);
  if (n >= N) return;    // uniform over the work-group
  const int y0 = yt * UN_TY - 1, z0 = zt * UN_TZ - 1;

  // K slice -> slot (x-offset 0, 2: slot 0; 1: slot 1) and plane offset
  for (int i = threadIdx.x; i < KP / 8; i += UN_BLOCK) {
    const int tap = i * 8 / CI, ci = i * 8 - tap * CI;
    tab[i] = tap < 27 ? (tap / 9 & 1) * PE + ((tap / 3) % 3 * BZ + tap % 3) * CI + ci : 0;
  }
  uint4 v[NIT];
  auto fetch = [&](int dx) {
    const int xx = x - 1 + dx;
    const bool xok = xx >= 0 && xx < D;
#pragma unroll
    for (int u = 0; u < NIT; ++u) {
      const int i = u * UN_BLOCK + threadIdx.x;
      const int by = i / RC, j = i - by * RC;
      const int yy = y0 + by, zz = z0 + j / V8;
      v[u] = (uint4){0u, 0u, 0u, 0u};
      if (xok && i < NCH && yy >= 0 && yy < H && zz >= 0 && zz < W) {
        if constexpr (C1 == 0) {
          v[u] = *(const uint4 *)(X + ((((int64_t)n * D + xx) * H + yy) * W + z0) * CI + j * 8);
        } else {
          const int c = (j % V8) * 8;
          if (c < C1)
            v[u] = *(const uint4 *)(X + ((((int64_t)n * (D / 2) + xx / 2) * (H / 2) + yy / 2) * (W / 2) + zz / 2) * C1 + c);
          else
            v[u] = *(const uint4 *)(X2 + ((((int64_t)n * D + xx) * H + yy) * W + zz) * (CI - C1) + (c - C1));
        }
      }
    }
  };
  auto store = [&](int slot) {
#pragma unroll
    for (int u = 0; u < NIT; ++u) {
      const int i = u * UN_BLOCK + threadIdx.x;
      if (i < NCH) *(uint4 *)(box + slot * PE + i * 8) = v[u];
    }
  };
  fetch(0);
  store(0);
  fetch(1);
  store(1);
  __syncthreads();
  fetch(2);

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int row = lane & 15, ks = lane >> 4;
  const bf16 *abase = box + ((2 * wave) * BZ + row) * CI;
  const int *tl = tab + ks;
  const bf16 *wbase = Wk + (size_t)row * KP + ks * 8;
  f32x4 acc[UN_MB][CB];
#pragma unroll
  for (int m = 0; m < UN_MB; ++m)
#pragma unroll
    for (int c = 0; c < CB; ++c) acc[m][c] = (f32x4){0.f, 0.f, 0.f, 0.f};

  auto chunk = [&](int kc) {
    const bf16 *ap = abase + tl[kc * 4];
    bf16x8 b[CB];
#pragma unroll
    for (int c = 0; c < CB; ++c) b[c] = *(const bf16x8 *)(wbase + (size_t)c * 16 * KP + kc * 32);
#pragma unroll
    for (int m = 0; m < UN_MB; ++m) {
      bf16x8 a = *(const bf16x8 *)(ap + ((m >> 1) * BZ + (m & 1) * 16) * CI);
      if (PAD && kc == NKC - 1 && ks >= 2) a = zero8();
#pragma unroll
      for (int c = 0; c < CB; ++c) acc[m][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[c], acc[m][c], 0, 0, 0);
    }
  };
#pragma unroll
  for (int kc = 0; kc < NKA; ++kc) chunk(kc);
  __syncthreads();    // slot 0 (plane x - 1) is consumed
  store(0);
  __syncthreads();
#pragma unroll
  for (int kc = NKA; kc < NKC; ++kc) chunk(kc);

  const int yb = yt * UN_TY + 2 * wave, zb = zt * UN_TZ + ks * 4;
  const size_t obase = ((((size_t)n * D + x) * H + yb) * W + zb) * COS;
#pragma unroll
  for (int c = 0; c < CB; ++c) {
    const int co = c * 16 + row;
    if (co >= COS) continue;
    const float bb = bias ? bias[co] : 0.f;
#pragma unroll
    for (int m = 0; m < UN_MB; ++m) {
      if (yb + (m >> 1) >= H) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if (zb + (m & 1) * 16 + r >= W) continue;
        float val = acc[m][c][r] + bb;
        if (ACT == 1) val = val > 0.f ? val : UN_SLOPE * val;
        const size_t o = obase + ((size_t)(m >> 1) * W + (m & 1) * 16 + r) * COS + co;
        if (OUTF32) ((float *)Yv)[o] = val;
        else ((bf16 *)Yv)[o] = (bf16)val;
      }
    }
  }
}

// ------------------------------------------------------------------------
// dW[co][tap * CI + ci] += sum_p dY[p][co] * X[p + off(tap)][ci]  (fp32, [CO][KP])
// One wave: one 16-channel block of CO, UN_NT column blocks of (tap, ci),
// `rows` consecutive (n, x, y) rows, all z.
// ------------------------------------------------------------------------
#define UN_NT 4

template <int CI, int CO>
__global__ __launch_bounds__(UN_BLOCK) void conv3_bwdw(const bf16 *__restrict__ X, const bf16 *__restrict__ dY,
                                                       float *__restrict__ dW, int N, int D, int H, int W,
                                                       int rows) {
  constexpr int KP = KPad<CI>::value;
  constexpr int NCB = 27 * CI / 16;
  constexpr int NG = (NCB + UN_NT - 1) / UN_NT;
  constexpr int COB = CO / 16;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int l16 = lane & 15, ks = lane >> 4;
  const int64_t nrows = (int64_t)N * D * H;
  const int64_t nranges = (nrows + rows - 1) / rows;
  int64_t task = (int64_t)blockIdx.x * (UN_BLOCK / 64) + wave;
  if (task >= nranges * COB * NG) return;
  const int grp = (int)(task % NG); task /= NG;
  const int cob = (int)(task % COB);
  const int64_t rr = task / COB;
  const int co = cob * 16 + l16;

  int tap[UN_NT], ci[UN_NT], ox[UN_NT], oy[UN_NT], oz[UN_NT];
  bool cv[UN_NT];
#pragma unroll
  for (int t = 0; t < UN_NT; ++t) {
    const int col = (grp * UN_NT + t) * 16 + l16;
    cv[t] = col < 27 * CI;
    tap[t] = cv[t] ? col / CI : 0;
    ci[t] = cv[t] ? col - tap[t] * CI : 0;
    ox[t] = tap[t] / 9 - 1; oy[t] = (tap[t] / 3) % 3 - 1; oz[t] = tap[t] % 3 - 1;
  }
  f32x4 acc[UN_NT];
#pragma unroll
  for (int t = 0; t < UN_NT; ++t) acc[t] = (f32x4){0.f, 0.f, 0.f, 0.f};

  const int64_t r_end = (rr + 1) * rows < nrows ? (rr + 1) * rows : nrows;
  for (int64_t r = rr * rows; r < r_end; ++r) {
    const int y = (int)(r % H), x = (int)((r / H) % D), n = (int)(r / ((int64_t)H * D));
    for (int z0 = 0; z0 < W; z0 += 32) {
      bf16x8 a;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int z = z0 + ks * 8 + j;
        a[j] = z < W ? dY[((((size_t)n * D + x) * H + y) * W + z) * CO + co] : (bf16)0.f;
      }
#pragma unroll
      for (int t = 0; t < UN_NT; ++t) {
        bf16x8 b;
        const int xx = x + ox[t], yy = y + oy[t];
        const bool rowok = cv[t] && xx >= 0 && xx < D && yy >= 0 && yy < H;
        const size_t base = (((size_t)n * D + (rowok ? xx : 0)) * H + (rowok ? yy : 0)) * W;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int zz = z0 + ks * 8 + j + oz[t];
          const bool ok = rowok && (z0 + ks * 8 + j) < W && zz >= 0 && zz < W;
          b[j] = ok ? X[(base + zz) * CI + ci[t]] : (bf16)0.f;
        }
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int t = 0; t < UN_NT; ++t) {
    const int col = (grp * UN_NT + t) * 16 + l16;
    if (col >= 27 * CI) continue;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int cr = cob * 16 + ks * 4 + q;
      atomicAdd(dW + (size_t)cr * KP + col, acc[t][q]);
    }
  }
}

// db[co] += sum_p dY[p][co]
template <int CO>
__global__ __launch_bounds__(UN_BLOCK) void bias_grad(const bf16 *__restrict__ dY, float *__restrict__ db, int64_t P,
                                                      int64_t per_block) {
  __shared__ float part[UN_BLOCK];
  const int co = threadIdx.x % CO, lanes = UN_BLOCK / CO;
  const int64_t p0 = (int64_t)blockIdx.x * per_block;
  const int64_t p1 = p0 + per_block < P ? p0 + per_block : P;
  float s = 0.f;
  for (int64_t p = p0 + threadIdx.x / CO; p < p1; p += lanes) s += (float)dY[p * CO + co];
  part[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x < CO) {
    float t = 0.f;
    for (int k = threadIdx.x; k < lanes * CO; k += CO) t += part[k];
    atomicAdd(db + threadIdx.x, t);
  }
}

// ------------------------------------------------------------------------
// elementwise kernels (bf16, channels-last).  One thread per 8 consecutive
// channels of one voxel (16-byte loads and stores); channel counts are
// multiples of 8 and element counts fit 32 bits.
// ------------------------------------------------------------------------
__device__ __forceinline__ bf16x8 ld8v(const bf16 *p) { return *(const bf16x8 *)p; }
__device__ __forceinline__ void st8v(bf16 *p, bf16x8 v) { *(bf16x8 *)p = v; }

// dpre = dy * (y > 0 ? 1 : slope)
__global__ void lrelu_bwd(const bf16 *__restrict__ dy, const bf16 *__restrict__ y, bf16 *__restrict__ out, int64_t n) {
  const int64_t i = ((int64_t)blockIdx.x * UN_BLOCK + threadIdx.x) * 8;
  if (i >= n) return;
  const bf16x8 g = ld8v(dy + i), a = ld8v(y + i);
  bf16x8 o;
#pragma unroll
  for (int j = 0; j < 8; ++j) o[j] = (bf16)((float)a[j] > 0.f ? (float)g[j] : UN_SLOPE * (float)g[j]);
  st8v(out + i, o);
}

// 2x2x2 average pool: x [N, D, H, W, C] -> y [N, D/2, H/2, W/2, C]
__global__ void avgpool2_fwd(const bf16 *__restrict__ x, bf16 *__restrict__ y, int N, int D, int H, int W, int C) {
  const int d2 = D / 2, h2 = H / 2, w2 = W / 2, C8 = C / 8;
  const int i = blockIdx.x * UN_BLOCK + threadIdx.x;
  if (i >= N * d2 * h2 * w2 * C8) return;
  int t = i;
  const int c8 = t % C8; t /= C8;
  const int z = t % w2; t /= w2;
  const int yy = t % h2; t /= h2;
  const int xx = t % d2;
  const int n = t / d2;
  float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const bf16x8 v = ld8v(x + (((((size_t)n * D + 2 * xx + (q >> 2)) * H + 2 * yy + ((q >> 1) & 1)) * W + 2 * z + (q & 1)) * C + c8 * 8));
#pragma unroll
    for (int j = 0; j < 8; ++j) s[j] += (float)v[j];
  }
  bf16x8 o;
#pragma unroll
  for (int j = 0; j < 8; ++j) o[j] = (bf16)(s[j] * 0.125f);
  st8v(y + (size_t)i * 8, o);
}

// dx[child] = dy[parent] / 8
__global__ void avgpool2_bwd(const bf16 *__restrict__ dy, bf16 *__restrict__ dx, int N, int D, int H, int W, int C) {
  const int C8 = C / 8;
  const int i = blockIdx.x * UN_BLOCK + threadIdx.x;
  if (i >= N * D * H * W * C8) return;
  int t = i;
  const int c8 = t % C8; t /= C8;
  const int z = t % W; t /= W;
  const int yy = t % H; t /= H;
  const int xx = t % D;
  const int n = t / D;
  const bf16x8 v = ld8v(dy + ((((size_t)n * (D / 2) + xx / 2) * (H / 2) + yy / 2) * (W / 2) + z / 2) * C + c8 * 8);
  bf16x8 o;
#pragma unroll
  for (int j = 0; j < 8; ++j) o[j] = (bf16)((float)v[j] * 0.125f);
  st8v(dx + (size_t)i * 8, o);
}

// out [N, D, H, W, C1 + C2] = concat(nearest_up2(a [N, D/2, H/2, W/2, C1]), skip [N, D, H, W, C2])
__global__ void up2_concat_fwd(const bf16 *__restrict__ a, const bf16 *__restrict__ skip, bf16 *__restrict__ out,
                               int N, int D, int H, int W, int C1, int C2) {
  const int C8 = (C1 + C2) / 8;
  const int i = blockIdx.x * UN_BLOCK + threadIdx.x;
  if (i >= N * D * H * W * C8) return;
  int t = i;
  const int c8 = t % C8; t /= C8;
  const int p = t;
  const int z = t % W; t /= W;
  const int yy = t % H; t /= H;
  const int xx = t % D;
  const int n = t / D;
  const int c = c8 * 8;
  bf16x8 v;
  if (c < C1)
    v = ld8v(a + ((((size_t)n * (D / 2) + xx / 2) * (H / 2) + yy / 2) * (W / 2) + z / 2) * C1 + c);
  else
    v = ld8v(skip + (size_t)p * C2 + (c - C1));
  st8v(out + (size_t)i * 8, v);
}

// backward of up2_concat_fwd: da = sum over the 8 children, dskip = slice
__global__ void up2_concat_bwd(const bf16 *__restrict__ dout, bf16 *__restrict__ da, bf16 *__restrict__ dskip,
                               int N, int D, int H, int W, int C1, int C2) {
  const int C = C1 + C2, d2 = D / 2, h2 = H / 2, w2 = W / 2;
  const int na = N * d2 * h2 * w2 * (C1 / 8), ns = N * D * H * W * (C2 / 8);
  const int i = blockIdx.x * UN_BLOCK + threadIdx.x;
  if (i < na) {
    int t = i;
    const int c8 = t % (C1 / 8); t /= C1 / 8;
    const int z = t % w2; t /= w2;
    const int yy = t % h2; t /= h2;
    const int xx = t % d2;
    const int n = t / d2;
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const bf16x8 v = ld8v(dout + (((((size_t)n * D + 2 * xx + (q >> 2)) * H + 2 * yy + ((q >> 1) & 1)) * W + 2 * z + (q & 1)) * C + c8 * 8));
#pragma unroll
      for (int j = 0; j < 8; ++j) s[j] += (float)v[j];
    }
    bf16x8 o;
#pragma unroll
    for (int j = 0; j < 8; ++j) o[j] = (bf16)s[j];
    st8v(da + (size_t)i * 8, o);
  } else if (i < na + ns) {
    const int j = i - na;
    const int c8 = j % (C2 / 8), p = j / (C2 / 8);
    st8v(dskip + (size_t)j * 8, ld8v(dout + (size_t)p * C + C1 + c8 * 8));
  }
}

// network input: [N, V, 16] bf16 = (g * sg (3 ch), h * sh (CC ch), 0 ...) from
// channel-first fp32 g [N, 3, V], h [N, CC, V]
__global__ void pack_input(const float *__restrict__ g, const float *__restrict__ h, int CC, bf16 *__restrict__ out,
                           int N, int64_t V, float sg, float sh) {
  const int64_t i = (int64_t)blockIdx.x * UN_BLOCK + threadIdx.x;
  if (i >= (int64_t)N * V) return;
  const int64_t n = i / V, v = i - n * V;
  bf16 o[16];
#pragma unroll
  for (int c = 0; c < 16; ++c) o[c] = (bf16)0.f;
#pragma unroll
  for (int c = 0; c < 3; ++c) o[c] = (bf16)(g[(n * 3 + c) * V + v] * sg);
  for (int c = 0; c < CC && c < 6; ++c) o[3 + c] = (bf16)(h[(n * CC + c) * V + v] * sh);
  bf16x8 *dst = (bf16x8 *)(out + i * 16);
  bf16x8 lo, hi;
#pragma unroll
  for (int c = 0; c < 8; ++c) { lo[c] = o[c]; hi[c] = o[8 + c]; }
  dst[0] = lo;
  dst[1] = hi;
}

// x [N, 3, V] (fp32, channel-first) = scale * y [N, V, 3] (+ x if accumulate)
__global__ void unpack_output(const float *__restrict__ y, float *__restrict__ x, int N, int64_t V, float scale,
                              int accumulate) {
  const int64_t i = (int64_t)blockIdx.x * UN_BLOCK + threadIdx.x;
  if (i >= (int64_t)N * 3 * V) return;
  const int64_t n = i / (3 * V), r = i - n * 3 * V, c = r / V, v = r - c * V;
  const float val = scale * y[(n * V + v) * 3 + c];
  x[i] = accumulate ? x[i] + val : val;
}

static inline unsigned nblk(int64_t n) { return (unsigned)((n + UN_BLOCK - 1) / UN_BLOCK); }

// ---------------------------- dispatch ----------------------------------
#define UN_FWD_CASES(X_)                                                      \
  X_(16, 16) X_(16, 32) X_(32, 32) X_(32, 64) X_(64, 64) X_(96, 32) X_(48, 16) \
  X_(32, 16) X_(64, 32) X_(32, 96) X_(16, 48)

// Raise the dynamic-LDS limit of one staged-convolution instantiation to
// `bytes` (once per device); false if that does not fit.
template <int CI, int CO, int COS, int ACT, bool OUTF32, bool PLANE, int C1 = 0>    // conv3_fwd_plane, else conv3_fwd_2slot
static bool lds_ready() {
  constexpr int bytes = PLANE ? PlaneBytes<CI>::value : SlotBytes<CI>::value;
  static int state[64] = {0};    // 0 unknown, 1 usable, 2 not
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return false;
  if (state[dev] == 0) {
    int v = 0;
    bool ok = bytes <= 65536;
    const void *fn;
    if constexpr (PLANE) fn = (const void *)conv3_fwd_plane<CI, CO, COS, ACT, OUTF32>;
    else fn = (const void *)conv3_fwd_2slot<CI, CO, COS, ACT, OUTF32, C1>;
    if (!ok && hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess &&
        v >= bytes)
      ok = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) == hipSuccess;
    (void)hipGetLastError();
    state[dev] = ok ? 1 : 2;
  }
  return state[dev] == 1;
}

template <int CI, int CO, int COS, int ACT, bool OUTF32>
static void launch_fwd1(const bf16 *x, const bf16 *w, const float *b, void *y, int N, int D, int H, int W,
                        hipStream_t st) {
  // staged kernels: two plane slots up to 48 input channels (<= 66 KB of
  // LDS), one plane at a time for 64 and 96 (<= 66 KB at CI = 96)
  const int64_t tiles = (int64_t)N * D * ((H + UN_TY - 1) / UN_TY) * ((W + UN_TZ - 1) / UN_TZ);
  if constexpr (CI <= 48) {
    if (tiles < (int64_t(1) << 31) && lds_ready<CI, CO, COS, ACT, OUTF32, false>()) {
      hipLaunchKernelGGL((conv3_fwd_2slot<CI, CO, COS, ACT, OUTF32>), dim3((unsigned)tiles), dim3(UN_BLOCK),
                         SlotBytes<CI>::value, st, x, w, b, y, N, D, H, W);
      return;
    }
  } else if constexpr (CI % 32 == 0) {
    if (tiles < (int64_t(1) << 31) && lds_ready<CI, CO, COS, ACT, OUTF32, true>()) {
      hipLaunchKernelGGL((conv3_fwd_plane<CI, CO, COS, ACT, OUTF32>), dim3((unsigned)tiles), dim3(UN_BLOCK),
                         PlaneBytes<CI>::value, st, x, w, b, y, N, D, H, W);
      return;
    }
  }
  const int64_t pbs = (int64_t)N * D * H * ((W + 15) / 16);
  const int64_t waves = (pbs + UN_MB - 1) / UN_MB;
  hipLaunchKernelGGL((conv3_fwd<CI, CO, COS, ACT, OUTF32>), dim3((unsigned)((waves + 3) / 4)), dim3(UN_BLOCK), 0,
                     st, x, w, b, y, N, D, H, W);
}

template <int CI, int CO>
static void launch_fwd(const bf16 *x, const bf16 *w, const float *b, void *y, int N, int D, int H, int W, int act,
                       int out_f32, int cos, hipStream_t st) {
  if (out_f32) {
    if constexpr (CO == 16)
      if (cos == 3) {
        launch_fwd1<CI, 16, 3, 0, true>(x, w, b, y, N, D, H, W, st);
        return;
      }
    launch_fwd1<CI, CO, CO, 0, true>(x, w, b, y, N, D, H, W, st);
  } else if (act) {
    launch_fwd1<CI, CO, CO, 1, false>(x, w, b, y, N, D, H, W, st);
  } else {
    launch_fwd1<CI, CO, CO, 0, false>(x, w, b, y, N, D, H, W, st);
  }
}

template <int CI, int CO>
static void launch_bwdw(const bf16 *x, const bf16 *dy, float *dw, int N, int D, int H, int W, hipStream_t st) {
  constexpr int NG = (27 * CI / 16 + UN_NT - 1) / UN_NT;
  const int64_t nrows = (int64_t)N * D * H;
  // rows per wave: enough work per wave to amortise the atomics, enough waves to fill the chip
  int rows = (int)(nrows * CO / 16 * NG / 4096);
  rows = rows < 1 ? 1 : (rows > 64 ? 64 : rows);
  const int64_t tasks = (nrows + rows - 1) / rows * (CO / 16) * NG;
  hipLaunchKernelGGL((conv3_bwdw<CI, CO>), dim3((unsigned)((tasks + 3) / 4)), dim3(UN_BLOCK), 0, st, x, dy, dw, N, D,
                     H, W, rows);
}

}  // namespace ni_amd

using namespace ni_amd;

extern "C" {

int ni_unet_conv3_fwd(int64_t N, const int64_t *shape, int32_t CI, int32_t CO, const void *x, const void *w,
                      const float *bias, void *y, int32_t act, int32_t out_f32, int32_t out_channels, void *stream) {
  set_error("");
  NI_CHECK(x && w && y && shape, "unet_conv3_fwd: missing argument");
  NI_CHECK(N >= 1 && shape[0] >= 1 && shape[1] >= 1 && shape[2] >= 1, "unet_conv3_fwd: bad shape");
  NI_CHECK(out_channels == CO || (out_f32 && CO == 16 && out_channels == 3),
           "unet_conv3_fwd: out_channels must equal CO (or 3 with CO = 16, fp32 output)");
  NI_CHECK(!(out_f32 && act), "unet_conv3_fwd: fp32 output is for the linear head");
  const int D = (int)shape[0], H = (int)shape[1], W = (int)shape[2];
  hipStream_t st = (hipStream_t)stream;
  bool done = false;
#define UN_X(ci, co)                                                                                   \
  if (!done && CI == ci && CO == co) {                                                                 \
    launch_fwd<ci, co>((const bf16 *)x, (const bf16 *)w, bias, y, (int)N, D, H, W, act, out_f32,       \
                       out_channels, st);                                                              \
    done = true;                                                                                       \
  }
  UN_FWD_CASES(UN_X)
#undef UN_X
  NI_CHECK(done, "unet_conv3_fwd: unsupported channel pair %d -> %d", CI, CO);
  NI_HIP_CHECK(hipGetLastError());
  return 0;
}

int ni_unet_conv3_fwd_upcat(int64_t N, const int64_t *shape, int32_t C1, int32_t C2, int32_t CO, const void *a,
                            const void *skip, const void *w, const float *bias, void *y, int32_t act, void *stream) {
  set_error("");
  NI_CHECK(a && skip && w && y && shape, "unet_conv3_fwd_upcat: missing argument");
  NI_CHECK(C1 == 32 && C2 == 16 && CO == 16, "unet_conv3_fwd_upcat: only 32 + 16 -> 16 channels (got %d + %d -> %d)",
           C1, C2, CO);
  NI_CHECK(N >= 1 && shape[0] >= 2 && shape[1] >= 2 && shape[2] >= 2, "unet_conv3_fwd_upcat: bad shape");
  NI_CHECK(shape[0] % 2 == 0 && shape[1] % 2 == 0 && shape[2] % 2 == 0, "unet_conv3_fwd_upcat: even sizes required");
  const int D = (int)shape[0], H = (int)shape[1], W = (int)shape[2];
  const int64_t tiles = N * D * ((H + UN_TY - 1) / UN_TY) * ((W + UN_TZ - 1) / UN_TZ);
  NI_CHECK(tiles < (int64_t(1) << 31), "unet_conv3_fwd_upcat: too many tiles");
  hipStream_t st = (hipStream_t)stream;
  if (act) {
    NI_CHECK((lds_ready<48, 16, 16, 1, false, false, 32>()), "unet_conv3_fwd_upcat: LDS limit not raised");
    hipLaunchKernelGGL((conv3_fwd_2slot<48, 16, 16, 1, false, 32>), dim3((unsigned)tiles), dim3(UN_BLOCK),
                       SlotBytes<48>::value, st, (const bf16 *)a, (const bf16 *)w, bias, y, (int)N, D, H, W,
                       (const bf16 *)skip);
  } else {
    NI_CHECK((lds_ready<48, 16, 16, 0, false, false, 32>()), "unet_conv3_fwd_upcat: LDS limit not raised");
    hipLaunchKernelGGL((conv3_fwd_2slot<48, 16, 16, 0, false, 32>), dim3((unsigned)tiles), dim3(UN_BLOCK),
                       SlotBytes<48>::value, st, (const bf16 *)a, (const bf16 *)w, bias, y, (int)N, D, H, W,
                       (const bf16 *)skip);
  }
  NI_HIP_CHECK(hipGetLastError());
  return 0;
}

int ni_unet_conv3_bwd_weight(int64_t N, const int64_t *shape, int32_t CI, int32_t CO, const void *x, const void *dy,
                             float *dw, float *db, void *stream) {
  set_error("");
  NI_CHECK(x && dy && dw && shape, "unet_conv3_bwd_weight: missing argument");
  const int D = (int)shape[0], H = (int)shape[1], W = (int)shape[2];
  hipStream_t st = (hipStream_t)stream;
  bool done = false;
#define UN_X(ci, co)                                                                                   \
  if (!done && CI == ci && CO == co) {                                                                 \
    launch_bwdw<ci, co>((const bf16 *)x, (const bf16 *)dy, dw, (int)N, D, H, W, st);                   \
    done = true;                                                                                       \
  }
  UN_FWD_CASES(UN_X)
#undef UN_X
  NI_CHECK(done, "unet_conv3_bwd_weight: unsupported channel pair %d -> %d", CI, CO);
  if (db) {
    const int64_t P = N * shape[0] * shape[1] * shape[2];
    const int64_t per = 4096;
    const unsigned nb = (unsigned)((P + per - 1) / per);
    switch (CO) {
      case 16: hipLaunchKernelGGL(bias_grad<16>, dim3(nb), dim3(UN_BLOCK), 0, st, (const bf16 *)dy, db, P, per); break;
      case 32: hipLaunchKernelGGL(bias_grad<32>, dim3(nb), dim3(UN_BLOCK), 0, st, (const bf16 *)dy, db, P, per); break;
      case 48: hipLaunchKernelGGL(bias_grad<48>, dim3(nb), dim3(UN_BLOCK), 0, st, (const bf16 *)dy, db, P, per); break;
      case 64: hipLaunchKernelGGL(bias_grad<64>, dim3(nb), dim3(UN_BLOCK), 0, st, (const bf16 *)dy, db, P, per); break;
      case 96: hipLaunchKernelGGL(bias_grad<96>, dim3(nb), dim3(UN_BLOCK), 0, st, (const bf16 *)dy, db, P, per); break;
      default: NI_CHECK(false, "unet_conv3_bwd_weight: unsupported CO %d for the bias gradient", CO);
    }
  }
  NI_HIP_CHECK(hipGetLastError());
  return 0;
}

int ni_unet_lrelu_bwd(int64_t n, const void *dy, const void *y, void *out, void *stream) {
  set_error("");
  NI_CHECK(dy && y && out, "unet_lrelu_bwd: missing argument");
  NI_CHECK(n % 8 == 0, "unet_lrelu_bwd: element count must be a multiple of 8");
  if (n > 0)
    hipLaunchKernelGGL(lrelu_bwd, dim3(nblk(n / 8)), dim3(UN_BLOCK), 0, (hipStream_t)stream, (const bf16 *)dy,
                       (const bf16 *)y, (bf16 *)out, n);
  NI_HIP_CHECK(hipGetLastError());
  return 0;
}

int ni_unet_avgpool2(int32_t backward, int64_t N, const int64_t *shape, int32_t C, const void *in, void *out,
                     void *stream) {
  set_error("");
  NI_CHECK(in && out && shape, "unet_avgpool2: missing argument");
  const int D = (int)shape[0], H = (int)shape[1], W = (int)shape[2];
  NI_CHECK(D % 2 == 0 && H % 2 == 0 && W % 2 == 0, "unet_avgpool2: even sizes required");
  NI_CHECK(C % 8 == 0 && N * D * H * W * C < (1LL << 31), "unet_avgpool2: C % 8 == 0 and < 2^31 elements required");
  hipStream_t st = (hipStream_t)stream;
  if (!backward)
    hipLaunchKernelGGL(avgpool2_fwd, dim3(nblk(N * D * H * W * C / 64)), dim3(UN_BLOCK), 0, st, (const bf16 *)in,
                       (bf16 *)out, (int)N, D, H, W, C);
  else
    hipLaunchKernelGGL(avgpool2_bwd, dim3(nblk(N * D * H * W * C / 8)), dim3(UN_BLOCK), 0, st, (const bf16 *)in,
                       (bf16 *)out, (int)N, D, H, W, C);
  NI_HIP_CHECK(hipGetLastError());
  return 0;
}

int ni_unet_up2_concat(int32_t backward, int64_t N, const int64_t *shape, int32_t C1, int32_t C2, const void *a,
                       const void *skip, void *out, void *stream) {
  set_error("");
  NI_CHECK(a && skip && out && shape, "unet_up2_concat: missing argument");
  const int D = (int)shape[0], H = (int)shape[1], W = (int)shape[2];
  NI_CHECK(D % 2 == 0 && H % 2 == 0 && W % 2 == 0, "unet_up2_concat: even sizes required");
  NI_CHECK(C1 % 8 == 0 && C2 % 8 == 0 && N * D * H * W * (C1 + C2) < (1LL << 31),
           "unet_up2_concat: channel counts multiple of 8 and < 2^31 elements required");
  hipStream_t st = (hipStream_t)stream;
  if (!backward)
    hipLaunchKernelGGL(up2_concat_fwd, dim3(nblk(N * D * H * W * (C1 + C2) / 8)), dim3(UN_BLOCK), 0, st,
                       (const bf16 *)a, (const bf16 *)skip, (bf16 *)out, (int)N, D, H, W, C1, C2);
  else  // a <- d(low-res input), skip <- d(skip), out = incoming gradient
    hipLaunchKernelGGL(up2_concat_bwd, dim3(nblk(N * D * H * W * C1 / 64 + N * D * H * W * C2 / 8)), dim3(UN_BLOCK),
                       0, st, (const bf16 *)out, (bf16 *)a, (bf16 *)skip, (int)N, D, H, W, C1, C2);
  NI_HIP_CHECK(hipGetLastError());
  return 0;
}

int ni_unet_pack_input(int64_t N, int64_t V, const float *g, const float *h, int32_t CC, void *out, float sg,
                       float sh, void *stream) {
  set_error("");
  NI_CHECK(g && out, "unet_pack_input: missing argument");
  NI_CHECK(h == nullptr || CC == 1 || CC == 3 || CC == 6, "unet_pack_input: CC must be 1, 3 or 6");
  hipLaunchKernelGGL(pack_input, dim3(nblk(N * V)), dim3(UN_BLOCK), 0, (hipStream_t)stream, g, h, h ? CC : 0,
                     (bf16 *)out, (int)N, V, sg, sh);
  NI_HIP_CHECK(hipGetLastError());
  return 0;
}

int ni_unet_unpack_output(int64_t N, int64_t V, const float *y, float *x, float scale, int32_t accumulate,
                          void *stream) {
  set_error("");
  NI_CHECK(y && x, "unet_unpack_output: missing argument");
  hipLaunchKernelGGL(unpack_output, dim3(nblk(N * 3 * V)), dim3(UN_BLOCK), 0, (hipStream_t)stream, y, x, (int)N, V,
                     scale, accumulate);
  NI_HIP_CHECK(hipGetLastError());
  return 0;
}

}  // extern "C"
