// Binned second pass for the tiles that tile_kernel listed (pull / push /
// count / grid_grad of strongly deformed grids).
//
// A tile of samples that a rotation, a zoom or a displacement with a large
// Jacobian spreads over a wide region of the source has a bounding box far
// larger than its sample count.  Here the listed tiles' samples are sorted
// by the bin of their first tap instead: BIN_S^3 first-tap positions per
// bin, in source space, so that a bin's samples touch exactly a
// (BIN_S + ORD)^3 box of voxels whatever the deformation.  One work-group
// per bin then stages that box (pull, grid_grad: 4 channels interleaved per
// voxel) or accumulates into it (push / count: FP64, as tile_kernel) and
// flushes it folded and signed with one FP32 global atomic per non-zero
// element.
//
// Counting sort in three launches over the listed samples: count per bin,
// exclusive scan of the counts, scatter of the sample positions.  A pull or
// grid_grad writes the zeros of its out-of-bound samples in the count pass; samples
// whose first tap lies outside the binned range (far out-of-range
// coordinates with extrapolation) are listed there and run the direct
// per-sample routine pp_sample in a kernel of their own (its registers
// would otherwise weigh on the sort).
//
// Sample positions are packed as (x * Y + y) << 12 | z, which needs Z <=
// 4096 and X * Y < 2^19; the host checks both.
#pragma once

namespace ni_amd {

#define BIN_S 8          // first-tap positions per bin and axis
#define BIN_OFF 8        // first taps >= -BIN_OFF are binned
#define BIN_NTH 256
#define BIN_ZBITS 12

template <int ORD>
struct BinBox {
  static constexpr int E = BIN_S + ORD;      // box extent per axis
  static constexpr int VOL = E * E * E;
};

// Listed sample j -> (n, x, y, z) and whether it is a real sample.
template <class B>
__device__ __forceinline__ bool bin_sample(const PPArgs<float> &a, const int *fb, int64_t j, int ntx, int nty,
                                           int ntz, int &n, int &px, int &py, int &pz) {
  constexpr int TS = B::TX * B::TY * B::TZ;
  int64_t r = fb[1 + j / TS];
  const int w = (int)(j % TS);
  const int tz = (int)(r % ntz); r /= ntz;
  const int ty = (int)(r % nty); r /= nty;
  const int tx = (int)(r % ntx);
  n = (int)(r / ntx);
  px = tx * B::TX + w / (B::TY * B::TZ);
  py = ty * B::TY + (w / B::TZ) % B::TY;
  pz = tz * B::TZ + w % B::TZ;
  return px < a.tX && py < a.tY && pz < a.tZ;
}

// In-bound test and first taps of the sample at (n, x, y, z); false when
// the sample is out of bounds.
template <int ORD>
__device__ __forceinline__ bool bin_taps(const PPArgs<float> &a, int n, int px, int py, int pz, float (&xq)[3],
                                         int (&lo)[3]) {
  const float *gp = a.grid + (int64_t)n * a.grid_s[0] + (int64_t)px * a.grid_s[1] + (int64_t)py * a.grid_s[2] +
                    (int64_t)pz * a.grid_s[3];
  const int gD = (int)a.grid_s[4];
  xq[0] = gp[0];
  xq[1] = gp[gD];
  xq[2] = gp[2 * gD];
  const float tol = (a.extrap & 2) ? 0.55f : 0.05f;
  const int sz[3] = {(int)a.sX, (int)a.sY, (int)a.sZ};
  bool ok = true;
  if (!(a.extrap & 1))
#pragma unroll
    for (int d = 0; d < 3; ++d) ok = ok && (xq[d] >= -tol && xq[d] < (float)(sz[d] - 1) + tol);
  if (ok) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      float w[ORD + 1], g[ORD + 1];
      lo[d] = tl_taps<ORD>(xq[d], w, g);
    }
  }
  return ok;
}

// Bin of first taps lo, or -1 when lo is outside the binned range.
__device__ __forceinline__ int64_t bin_of(const int (&lo)[3], int n, int nbx, int nby, int nbz) {
  const int bx = lo[0] + BIN_OFF, by = lo[1] + BIN_OFF, bz = lo[2] + BIN_OFF;
  if (bx < 0 || by < 0 || bz < 0 || bx >= nbx * BIN_S || by >= nby * BIN_S || bz >= nbz * BIN_S) return -1;
  return (((int64_t)n * nbx + bx / BIN_S) * nby + by / BIN_S) * nbz + bz / BIN_S;
}

// Rank of this lane's sample among the samples of bin b (b >= 0; lanes with
// b < 0 take no part): lanes of the wave with the same bin -- neighbouring
// samples mostly share one -- are served by one global atomic, issued by
// the lowest of them.  Called by every active lane of the wave.
__device__ __forceinline__ int wave_rank(int *cnt, int b) {
  const int lane = (int)(threadIdx.x & 63);
  const unsigned long long below = (1ull << lane) - 1ull;
  unsigned long long pending = __ballot(b >= 0);
  int rank = 0;
  while (pending) {
    const int leader = __ffsll((long long)pending) - 1;
    const int lb = __shfl(b, leader, 64);
    const unsigned long long same = __ballot(b == lb) & pending;
    int base = 0;
    if (lane == leader) base = atomicAdd(cnt + lb, __popcll(same));
    base = __shfl(base, leader, 64);
    if (b == lb) rank = base + __popcll(same & below);
    pending &= ~same;
  }
  return rank;
}

// Pass 1 (SCATTER = false): count the listed samples per bin; a pull writes
// the zeros of its out-of-bound samples, and in-bound samples whose first
// tap lies outside the binned range are listed (listed-sample index, from
// the end of idx, count in *rest) for bin_rest.  Pass 2 (SCATTER = true):
// write each binned sample's packed position at start[bin] + (its rank in
// the bin); `cnt` was zeroed by the scan and counts again.
template <int ORD, int MODE, bool SCATTER>
__global__ __launch_bounds__(BIN_NTH) void bin_sort(PPArgs<float> a, int ntx, int nty, int ntz, const int *fb,
                                                    int nbx, int nby, int nbz, int *cnt, const int *start,
                                                    int *idx, int *rest, int nsamp) {
  using B = TileBox<ORD, !TileMeta<MODE>::gather>;
  constexpr int TS = B::TX * B::TY * B::TZ;
  const int64_t total = (int64_t)fb[0] * TS;
  for (int64_t j = (int64_t)blockIdx.x * BIN_NTH + threadIdx.x; j < total; j += (int64_t)gridDim.x * BIN_NTH) {
    int n, px, py, pz;
    const bool valid = bin_sample<B>(a, fb, j, ntx, nty, ntz, n, px, py, pz);
    float xq[3];
    int lo[3];
    const bool ok = valid && bin_taps<ORD>(a, n, px, py, pz, xq, lo);
    const int b = ok ? (int)bin_of(lo, n, nbx, nby, nbz) : -1;
    if (!SCATTER && (MODE == TM_PULL || MODE == TM_SGRAD) && valid && !ok) {   // out of bounds: zero
      float *op = a.out + (int64_t)n * a.out_s[0] + (int64_t)px * a.out_s[2] + (int64_t)py * a.out_s[3] +
                  (int64_t)pz * a.out_s[4];
      for (int c = 0; c < (int)a.C; ++c)
        for (int d = 0; d < (MODE == TM_SGRAD ? 3 : 1); ++d) op[(int64_t)c * a.out_s[1] + (int64_t)d * a.out_s[5]] = 0.f;
    }
    if (!SCATTER && ok && b < 0) idx[nsamp - 1 - atomicAdd(rest, 1)] = (int)j;
    const int rank = wave_rank(cnt, b);
    if (SCATTER && b >= 0) idx[start[b] + rank] = ((px * (int)a.tY + py) << BIN_ZBITS) | pz;
  }
}

// The listed samples that no bin holds, by the direct per-sample routine.
template <int ORD, int MODE>
__global__ __launch_bounds__(BIN_NTH) void bin_rest(PPArgs<float> a, int ntx, int nty, int ntz, const int *fb,
                                                    const int *idx, const int *rest, int nsamp) {
  using B = TileBox<ORD, !TileMeta<MODE>::gather>;
  if (fb[0] == 0) return;
  const int nr = *rest;
  for (int r = blockIdx.x * BIN_NTH + threadIdx.x; r < nr; r += gridDim.x * BIN_NTH) {
    int n, px, py, pz;
    if (bin_sample<B>(a, fb, (int64_t)idx[nsamp - 1 - r], ntx, nty, ntz, n, px, py, pz))
      pp_sample<float, int32_t, ORD, TileFlags<MODE>::value, true>(a, n, px, py, pz);
  }
}

// Exclusive scan of cnt[0, nb) into start[0, nb], then cnt = 0: one
// work-group of 1024 threads walks the bins in chunks of 4096 (4
// consecutive bins a thread, coalesced), scanning each chunk by wave
// shuffles and one pass over the 16 wave totals.  Returns at once when no
// tile was listed (fb[0] == 0: nothing to sort).
static __global__ __launch_bounds__(1024) void bin_scan(const int *fb, int *cnt, int *start, int64_t nb) {
  __shared__ int wsum[16];
  __shared__ int carry_s;
  if (fb[0] == 0) return;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  if (t == 0) carry_s = 0;
  __syncthreads();
  for (int64_t c0 = 0; c0 < nb; c0 += 4096) {
    const int64_t b = c0 + 4 * (int64_t)t;
    int v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = b + k < nb ? cnt[b + k] : 0;
    const int own = v[0] + v[1] + v[2] + v[3];
    int inc = own;   // inclusive scan over the wave
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int u = __shfl_up(inc, o, 64);
      if (lane >= o) inc += u;
    }
    if (lane == 63) wsum[wv] = inc;
    __syncthreads();
    int before = carry_s;
    for (int w = 0; w < wv; ++w) before += wsum[w];
    int run = before + inc - own;
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (b + k < nb) {
        start[b + k] = run;
        cnt[b + k] = 0;
        run += v[k];
      }
    __syncthreads();   // every thread has read carry_s and wsum
    if (t == 1023) carry_s = run;   // the chunk's end
    __syncthreads();
  }
  if (t == 0) start[nb] = carry_s;
}

// Folded offsets (source / output strides) and signs of one bin's box.
template <int ORD>
__device__ __forceinline__ void bin_tables(const PPArgs<float> &a, const int64_t *st, const int (&o)[3],
                                           int *fofs, float *fsg) {
  constexpr int E = BinBox<ORD>::E;
  const int sz[3] = {(int)a.sX, (int)a.sY, (int)a.sZ};
  for (int i = threadIdx.x; i < 3 * E; i += BIN_NTH) {
    const int d = i / E, k = i - d * E;
    int sg = 0;
    const int j = fold<int>(a.bound[d], o[d] + k, sz[d], sg);
    fofs[i] = j * (int)st[2 + d];
    fsg[i] = (float)sg;
  }
}

// bin -> (n, box origin)
__device__ __forceinline__ void bin_decode(int64_t b, int nbx, int nby, int nbz, int &n, int (&o)[3]) {
  int64_t r = b;
  const int bz = (int)(r % nbz); r /= nbz;
  const int by = (int)(r % nby); r /= nby;
  const int bx = (int)(r % nbx);
  n = (int)(r / nbx);
  o[0] = bx * BIN_S - BIN_OFF;
  o[1] = by * BIN_S - BIN_OFF;
  o[2] = bz * BIN_S - BIN_OFF;
}

// A batch of up to SPT * BIN_NTH samples of a bin: each thread keeps its
// samples' weights, box offsets and positions in registers across the
// channel chunks (pull: 4 samples a thread; push: 2, whose FP64 splats
// need more registers).
template <int ORD, int SPT>
struct BinSamples {
  static constexpr int NT = ORD + 1;
  float w[SPT][3][NT];
  int off[SPT];     // box offset of the first tap, or -1 (no sample)
  int pos[SPT];     // x * X-stride + y * Y-stride + z * Z-stride of `st`
  // samples i0 + k * BIN_NTH + threadIdx.x of the bin's list
  __device__ __forceinline__ void load(const PPArgs<float> &a, const int *idx, int s0, int cnt, int i0, int n,
                                       const int (&o)[3], const int64_t *st) {
    constexpr int E = BinBox<ORD>::E;
    const float rY = 1.f / (float)a.tY;
#pragma unroll
    for (int k = 0; k < SPT; ++k) {
      const int i = i0 + k * BIN_NTH + (int)threadIdx.x;
      off[k] = -1;
      pos[k] = 0;
      if (i >= cnt) continue;
      const int s = idx[s0 + i];
      const int pz = s & ((1 << BIN_ZBITS) - 1), row = s >> BIN_ZBITS;
      const int px = tl_div(row, (int)a.tY, rY), py = row - px * (int)a.tY;
      const float *gp = a.grid + (int64_t)n * a.grid_s[0] + (int64_t)px * a.grid_s[1] + (int64_t)py * a.grid_s[2] +
                        (int64_t)pz * a.grid_s[3];
      const int gD = (int)a.grid_s[4];
      float g[NT];
      const int l0 = tl_taps<ORD>(gp[0], w[k][0], g) - o[0];
      const int l1 = tl_taps<ORD>(gp[gD], w[k][1], g) - o[1];
      const int l2 = tl_taps<ORD>(gp[2 * gD], w[k][2], g) - o[2];
      off[k] = (l0 * E + l1) * E + l2;
      pos[k] = px * (int)st[2] + py * (int)st[3] + pz * (int)st[4];
    }
  }
};

// Pull: one work-group per bin; 4 channels interleaved per box voxel.
template <int ORD>
__global__ __launch_bounds__(BIN_NTH) void bin_pull(PPArgs<float> a, int nbx, int nby, int nbz, const int *fb,
                                                    const int *start, const int *idx, int64_t nbins) {
  constexpr int NT = ORD + 1, E = BinBox<ORD>::E, VOL = BinBox<ORD>::VOL;
  __shared__ __align__(16) float4 box[VOL];
  __shared__ int fofs[3 * E];
  __shared__ float fsg[3 * E];
  if (fb[0] == 0) return;   // no listed tile: start[] was not written
  const int64_t b = xcd_block(nbins);
  const int s0 = start[b], cnt = start[b + 1] - s0;
  if (cnt == 0) return;
  int n, o[3];
  bin_decode(b, nbx, nby, nbz, n, o);
  bin_tables<ORD>(a, a.src_s, o, fofs, fsg);
  const int C = (int)a.C, sC = (int)a.src_s[1], oC = (int)a.out_s[1];
  const float *src = a.src + (int64_t)n * a.src_s[0];
  float *out = a.out + (int64_t)n * a.out_s[0];
  constexpr int SPT = 4;
  BinSamples<ORD, SPT> sm;
  for (int i0 = 0; i0 < cnt; i0 += SPT * BIN_NTH) {
    sm.load(a, idx, s0, cnt, i0, n, o, a.out_s);
    for (int c0 = 0; c0 < C; c0 += 4) {
      const int nc = min(4, C - c0);
      __syncthreads();   // tables ready / previous chunk's taps done
      for (int q = threadIdx.x; q < VOL; q += BIN_NTH) {
        const int bx = q / (E * E), by = (q / E) % E, bz = q % E;
        const float sg = fsg[bx] * fsg[E + by] * fsg[2 * E + bz];
        const float *sp = src + (int64_t)c0 * sC + fofs[bx] + fofs[E + by] + fofs[2 * E + bz];
        float v[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) v[c] = (c < nc && sg != 0.f) ? sp[c * sC] * sg : 0.f;
        box[q] = make_float4(v[0], v[1], v[2], v[3]);
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < SPT; ++k) {
        if (sm.off[k] < 0) continue;
        const float4 *bp = box + sm.off[k];
        float acc[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
        for (int kx = 0; kx < NT; ++kx, bp += E * E) {
#pragma unroll
          for (int ky = 0; ky < NT; ++ky) {
            const float wxy = sm.w[k][0][0] * sm.w[k][1][ky];
#pragma unroll
            for (int kz = 0; kz < NT; ++kz) {
              const float ww = wxy * sm.w[k][2][kz];
              const float4 v = bp[ky * E + kz];
              acc[0] += v.x * ww;
              acc[1] += v.y * ww;
              acc[2] += v.z * ww;
              acc[3] += v.w * ww;
            }
          }
          tl_rot(sm.w[k][0]);
        }
        float *op = out + (int64_t)c0 * oC + sm.pos[k];
#pragma unroll
        for (int c = 0; c < 4; ++c)
          if (c < nc) op[c * oC] = acc[c];
      }
    }
  }
}

// grid_grad: one work-group per bin, 4 channels interleaved per box voxel
// (as bin_pull); a sample's weights and derivatives are evaluated when it
// is processed (registers), and its taps summed separably as in the tile
// kernel: per (z, y) tap row the x sums with weights and derivatives, then
// the y-z factors.
template <int ORD>
__global__ __launch_bounds__(BIN_NTH) void bin_grad(PPArgs<float> a, int nbx, int nby, int nbz, const int *fb,
                                                    const int *start, const int *idx, int64_t nbins) {
  constexpr int NT = ORD + 1, E = BinBox<ORD>::E, VOL = BinBox<ORD>::VOL;
  __shared__ __align__(16) float4 box[VOL];
  __shared__ int fofs[3 * E];
  __shared__ float fsg[3 * E];
  if (fb[0] == 0) return;   // no listed tile: start[] was not written
  const int64_t b = xcd_block(nbins);
  const int s0 = start[b], cnt = start[b + 1] - s0;
  if (cnt == 0) return;
  int n, o[3];
  bin_decode(b, nbx, nby, nbz, n, o);
  bin_tables<ORD>(a, a.src_s, o, fofs, fsg);
  const int C = (int)a.C, sC = (int)a.src_s[1], oC = (int)a.out_s[1], oK = (int)a.out_s[5];
  const float *src = a.src + (int64_t)n * a.src_s[0];
  float *out = a.out + (int64_t)n * a.out_s[0];
  const float rY = 1.f / (float)a.tY;
  const int gD = (int)a.grid_s[4];
  constexpr int SPT = 2;
  for (int i0 = 0; i0 < cnt; i0 += SPT * BIN_NTH) {
    float q[SPT][3];
    int off[SPT], pos[SPT];
#pragma unroll
    for (int k = 0; k < SPT; ++k) {
      const int i = i0 + k * BIN_NTH + (int)threadIdx.x;
      off[k] = -1;
      pos[k] = 0;
      q[k][0] = q[k][1] = q[k][2] = 0.f;
      if (i >= cnt) continue;
      const int s = idx[s0 + i];
      const int pz = s & ((1 << BIN_ZBITS) - 1), row = s >> BIN_ZBITS;
      const int px = tl_div(row, (int)a.tY, rY), py = row - px * (int)a.tY;
      const float *gp = a.grid + (int64_t)n * a.grid_s[0] + (int64_t)px * a.grid_s[1] + (int64_t)py * a.grid_s[2] +
                        (int64_t)pz * a.grid_s[3];
      int l[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        float wt[NT], gt[NT];
        q[k][d] = gp[d * gD];
        l[d] = tl_taps<ORD>(q[k][d], wt, gt) - o[d];
      }
      off[k] = (l[0] * E + l[1]) * E + l[2];
      pos[k] = px * (int)a.out_s[2] + py * (int)a.out_s[3] + pz * (int)a.out_s[4];
    }
    for (int c0 = 0; c0 < C; c0 += 4) {
      const int nc = min(4, C - c0);
      __syncthreads();   // tables ready / previous chunk's taps done
      for (int qv = threadIdx.x; qv < VOL; qv += BIN_NTH) {
        const int bx = qv / (E * E), by = (qv / E) % E, bz = qv % E;
        const float sg = fsg[bx] * fsg[E + by] * fsg[2 * E + bz];
        const float *sp = src + (int64_t)c0 * sC + fofs[bx] + fofs[E + by] + fofs[2 * E + bz];
        float v[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) v[c] = (c < nc && sg != 0.f) ? sp[c * sC] * sg : 0.f;
        box[qv] = make_float4(v[0], v[1], v[2], v[3]);
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < SPT; ++k) {
        if (off[k] < 0) continue;
        float wi[3][NT], gi[3][NT];
#pragma unroll
        for (int d = 0; d < 3; ++d) tl_taps<ORD, true>(q[k][d], wi[d], gi[d]);
        float ax[4] = {0.f, 0.f, 0.f, 0.f}, ay[4] = {0.f, 0.f, 0.f, 0.f}, az[4] = {0.f, 0.f, 0.f, 0.f};
        const float4 *bp = box + off[k];
#pragma unroll 1
        for (int kz = 0; kz < NT; ++kz, ++bp) {
          const float4 *bq = bp;
#pragma unroll 1
          for (int ky = 0; ky < NT; ++ky, bq += E) {
            const float wyz = wi[1][0] * wi[2][0], gyz = gi[1][0] * wi[2][0], wgz = wi[1][0] * gi[2][0];
            float sw[4] = {0.f, 0.f, 0.f, 0.f}, sd[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kx = 0; kx < NT; ++kx) {
              const float4 v = bq[kx * E * E];
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
            tl_rot(wi[1]);
            tl_rot(gi[1]);
          }
          tl_rot(wi[2]);
          tl_rot(gi[2]);
        }
        float *op = out + (int64_t)c0 * oC + pos[k];
#pragma unroll
        for (int c = 0; c < 4; ++c)
          if (c < nc) {
            op[c * oC] = ax[c];
            op[c * oC + oK] = ay[c];
            op[c * oC + 2 * oK] = az[c];
          }
      }
    }
  }
}

// Push / count: one work-group per bin; FP64 box, one channel at a time.
template <int ORD, int MODE>
__global__ __launch_bounds__(BIN_NTH, 4) void bin_push(PPArgs<float> a, int nbx, int nby, int nbz, const int *fb,
                                                       const int *start, const int *idx, int64_t nbins) {
  constexpr int NT = ORD + 1, E = BinBox<ORD>::E, VOL = BinBox<ORD>::VOL;
  __shared__ double box[VOL];
  __shared__ int fofs[3 * E];
  __shared__ float fsg[3 * E];
  if (fb[0] == 0) return;   // no listed tile: start[] was not written
  const int64_t b = xcd_block(nbins);
  const int s0 = start[b], cnt = start[b + 1] - s0;
  if (cnt == 0) return;
  int n, o[3];
  bin_decode(b, nbx, nby, nbz, n, o);
  bin_tables<ORD>(a, a.out_s, o, fofs, fsg);
  for (int q = threadIdx.x; q < VOL; q += BIN_NTH) box[q] = 0.0;
  const int C = MODE == TM_COUNT ? 1 : (int)a.C;
  const float *trgt = MODE == TM_COUNT ? nullptr : a.trgt + (int64_t)n * a.trgt_s[0];
  float *out = a.out + (int64_t)n * a.out_s[0];
  constexpr int SPT = 2;
  BinSamples<ORD, SPT> sm;
  for (int i0 = 0; i0 < cnt; i0 += SPT * BIN_NTH) {
    sm.load(a, idx, s0, cnt, i0, n, o, MODE == TM_COUNT ? a.out_s : a.trgt_s);
    for (int c = 0; c < C; ++c) {
      __syncthreads();   // tables and box ready / previous flush done
#pragma unroll
      for (int k = 0; k < SPT; ++k) {
        if (sm.off[k] < 0) continue;
        const float v = MODE == TM_COUNT ? 1.f : trgt[(int64_t)c * a.trgt_s[1] + sm.pos[k]];
        if (v == 0.f) continue;
        const double vd = (double)v;
        double *bp = box + sm.off[k];
        // lanes visit the z taps in rotated orders (lane mod NT first):
        // samples of a compressed region share first taps, and the lanes
        // would otherwise hit the same LDS address in the same instruction
        const int rot = (int)(threadIdx.x % NT);
        float wz[NT];
#pragma unroll
        for (int kz = 0; kz < NT; ++kz) wz[kz] = sm.w[k][2][kz];
        for (int r = 0; r < rot; ++r) tl_rot(wz);
#pragma unroll 1
        for (int kx = 0; kx < NT; ++kx, bp += E * E) {
#pragma unroll
          for (int ky = 0; ky < NT; ++ky) {
            const double t = (double)sm.w[k][0][0] * (double)sm.w[k][1][ky] * vd;
#pragma unroll
            for (int kz = 0; kz < NT; ++kz) {
              const int z = kz + rot < NT ? kz + rot : kz + rot - NT;
              atomicAdd(bp + ky * E + z, t * (double)wz[kz]);
            }
          }
          tl_rot(sm.w[k][0]);   // NT rotations restore it for the next channel
        }
      }
      __syncthreads();
      // flush: one FP32 global atomic per non-zero box element; clear it
      float *oc = out + (int64_t)c * a.out_s[1];
      for (int q = threadIdx.x; q < VOL; q += BIN_NTH) {
        const double dv = box[q];
        box[q] = 0.0;
        if (dv == 0.0) continue;
        const int bx = q / (E * E), by = (q / E) % E, bz = q % E;
        const float sg = fsg[bx] * fsg[E + by] * fsg[2 * E + bz];
        if (sg != 0.f) atomic_add(oc + fofs[bx] + fofs[E + by] + fofs[2 * E + bz], (float)dv * sg);
      }
    }
  }
}

// Host: the binned pass over the tiles listed in fb (count, scan, scatter,
// one work-group per bin).  Returns false (nothing launched) when the
// volume does not fit the packed sample positions or the scratch cannot be
// allocated; the caller then runs its other second pass.
template <int ORD, int MODE>
static bool launch_binned(const PPArgs<float> &a, int ntx, int nty, int ntz, int64_t ntiles, const int *fb,
                          hipStream_t st) {
  using B = TileBox<ORD, !TileMeta<MODE>::gather>;
  constexpr int TS = B::TX * B::TY * B::TZ;
  if (a.tZ > (1 << BIN_ZBITS) || a.tX * a.tY >= (int64_t(1) << 19)) return false;
  const int nbx = (int)((a.sX + 2 * BIN_OFF + BIN_S - 1) / BIN_S);
  const int nby = (int)((a.sY + 2 * BIN_OFF + BIN_S - 1) / BIN_S);
  const int nbz = (int)((a.sZ + 2 * BIN_OFF + BIN_S - 1) / BIN_S);
  const int64_t nbins = a.N * (int64_t)nbx * nby * nbz;
  const int64_t nsamp = ntiles * TS;   // listed samples, at most
  if (nbins >= (int64_t(1) << 31) || nsamp >= (int64_t(1) << 31)) return false;
  int *mem = nullptr;
  if (scratch_alloc((void **)&mem, sizeof(int) * (size_t)(2 * nbins + 2 + nsamp), st) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  // [rest count | cnt: nbins | start: nbins + 1 | idx: nsamp]
  int *rest = mem, *cnt = mem + 1, *start = cnt + nbins, *idx = start + nbins + 1;
  const int64_t gz = (nbins + NI_BLOCK - 1) / NI_BLOCK, gsamp = (nsamp + BIN_NTH - 1) / BIN_NTH;
  hipLaunchKernelGGL(fill_zero<int>, dim3((unsigned)(gz < 65536 ? gz : 65536)), dim3(NI_BLOCK), 0, st, mem, nbins + 1);
  const unsigned gs = (unsigned)(gsamp < 8192 ? gsamp : 8192);
  hipLaunchKernelGGL((bin_sort<ORD, MODE, false>), dim3(gs), dim3(BIN_NTH), 0, st, a, ntx, nty, ntz, fb, nbx, nby,
                     nbz, cnt, (const int *)start, idx, rest, (int)nsamp);
  hipLaunchKernelGGL(bin_scan, dim3(1), dim3(1024), 0, st, fb, cnt, start, nbins);
  hipLaunchKernelGGL((bin_sort<ORD, MODE, true>), dim3(gs), dim3(BIN_NTH), 0, st, a, ntx, nty, ntz, fb, nbx, nby,
                     nbz, cnt, (const int *)start, idx, rest, (int)nsamp);
  hipLaunchKernelGGL((bin_rest<ORD, MODE>), dim3(gs < 1024 ? gs : 1024), dim3(BIN_NTH), 0, st, a, ntx, nty, ntz, fb,
                     (const int *)idx, (const int *)rest, (int)nsamp);
  if constexpr (MODE == TM_PULL)
    hipLaunchKernelGGL((bin_pull<ORD>), dim3((unsigned)nbins), dim3(BIN_NTH), 0, st, a, nbx, nby, nbz, fb,
                       (const int *)start, (const int *)idx, nbins);
  else if constexpr (MODE == TM_SGRAD)
    hipLaunchKernelGGL((bin_grad<ORD>), dim3((unsigned)nbins), dim3(BIN_NTH), 0, st, a, nbx, nby, nbz, fb,
                       (const int *)start, (const int *)idx, nbins);
  else
    hipLaunchKernelGGL((bin_push<ORD, MODE>), dim3((unsigned)nbins), dim3(BIN_NTH), 0, st, a, nbx, nby, nbz, fb,
                       (const int *)start, (const int *)idx, nbins);
  (void)hipFreeAsync(mem, st);
  return true;
}

}  // namespace ni_amd
