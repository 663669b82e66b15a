/*
 * ORACLE — test infrastructure only.  Not part of the product: only tests/,
 * the smoke check and the bench's CPU-baseline leg may load it, and only as
 * the checker.
 *
 * Plain-C restatement of nitorch's spline push/pull algorithm
 * (nitorch/_C/src/impl/pushpull_common.cpp, PushPullImpl) for 1D/2D/3D,
 * orders 0-7, all bound types and all mode flags (pull, push, count, grad,
 * sgrad, with trgt_K = 0 or 3).
 *
 * Compiled twice (see oracle/Makefile):
 *   OR_REAL=double, suffix _f64 : the FP64 oracle;
 *   OR_REAL=float,  suffix _f32 : the reference's FP32 semantics (float data,
 *                                 float coordinate arithmetic, spline weights
 *                                 evaluated in higher precision then rounded).
 *
 * Spline values are NOT taken from the reference's hand-written polynomials:
 * they come from the truncated-power form of the centred B-spline,
 *   B_n(x) = 1/n! sum_{k=0}^{n+1} (-1)^k C(n+1,k) (x + (n+1)/2 - k)_+^n,
 * with B_n' (x) = B_{n-1}(x+1/2) - B_{n-1}(x-1/2) and
 *      B_n''(x) = B_{n-2}(x+1) - 2 B_{n-2}(x) + B_{n-2}(x-1),
 * evaluated in long double.  Conventions at the break points where a
 * derivative is discontinuous follow the reference
 * (nitorch/_C/src/impl/interpolation_common.h:100-170, :230-255).
 *
 * Each function below cites the reference lines whose semantics it restates.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef OR_REAL
#define OR_REAL double
#endif
#ifndef OR_SUFFIX
#define OR_SUFFIX _f64
#endif
#define OR_CAT2(a, b) a##b
#define OR_CAT(a, b) OR_CAT2(a, b)
#define OR_FN(name) OR_CAT(name, OR_SUFFIX)

typedef OR_REAL real;
typedef long double ldbl;

/* Bound codes: the reference's BoundType enum (nitorch/_C/src/bounds.h:7-17) */
enum { B_REPLICATE = 0, B_DCT1 = 1, B_DCT2 = 2, B_DST1 = 3, B_DST2 = 4,
       B_DFT = 5, B_SLIDING = 6, B_ZERO = 7, B_NOCHECK = 8 };

/* ------------------------------------------------------------------------ */
/* Boundary folding: bounds_common.h:28-96 (index) and :105-138 (sign).      */
/* Written from the periodic-extension definitions of each transform.       */
/* ------------------------------------------------------------------------ */
static int64_t pmod(int64_t a, int64_t m) { int64_t r = a % m; return r < 0 ? r + m : r; }

static int64_t fold_index(int b, int64_t i, int64_t n)
{
  switch (b) {
  case B_REPLICATE: return i < 0 ? 0 : (i >= n ? n - 1 : i);
  case B_DCT1: {                       /* mirror about the centre of edge voxels, period 2(n-1) */
    if (n == 1) return 0;
    int64_t p = 2 * (n - 1), k = pmod(i < 0 ? -i : i, p);
    return k >= n ? p - k : k;
  }
  case B_DCT2: case B_DST2: {          /* mirror about the voxel edge, period 2n */
    int64_t p = 2 * n, k = pmod(i, p);
    return k >= n ? p - 1 - k : k;
  }
  case B_DST1: {                       /* antisymmetric about the first outside voxel, period 2(n+1) */
    if (n == 1) return 0;
    int64_t p = 2 * (n + 1);
    int64_t k = (i == -1) ? 0 : (i < 0 ? -i - 2 : i);
    k = pmod(k, p);
    if (k == n) return n - 1;
    return k > n ? p - k - 2 : k;
  }
  case B_DFT: return pmod(i, n);
  default: return i;                   /* zero / nocheck: identity (sign masks it) */
  }
}

static int fold_sign(int b, int64_t i, int64_t n)
{
  switch (b) {
  case B_DST1: {
    if (n == 1) return 1;
    int64_t p = 2 * (n + 1);
    int64_t k = i < 0 ? n - i - 1 : i;
    k = k % p;
    if (k % (n + 1) == n) return 0;
    return ((k / (n + 1)) % 2) ? -1 : 1;
  }
  case B_DST2: {
    int64_t k = i < 0 ? n - i - 1 : i;
    return ((k / n) % 2) ? -1 : 1;
  }
  case B_ZERO: case B_NOCHECK: case B_SLIDING:
    return (i < 0 || i >= n) ? 0 : 1;
  default: return 1;
  }
}

/* ------------------------------------------------------------------------ */
/* Centred B-splines (independent statement, see header).                   */
/* ------------------------------------------------------------------------ */
static ldbl binom(int n, int k)
{
  ldbl r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

static ldbl bspl(int n, ldbl x)
{
  if (n < 0) return 0;
  ldbl half = (ldbl)(n + 1) / 2;
  if (fabsl(x) >= half) return 0;
  if (n == 0) return fabsl(x) < 0.5L ? 1.0L : 0.0L;
  ldbl s = 0, fact = 1;
  for (int i = 2; i <= n; ++i) fact *= i;
  /* last piece, (half - |x|)^n / n!: the truncated-power sum below cancels
   * there, and would leave tiny weights with an absolute error only */
  if (fabsl(x) >= half - 1) return powl(half - fabsl(x), n) / fact;
  for (int k = 0; k <= n + 1; ++k) {
    ldbl t = x + half - k;
    if (t > 0) s += ((k & 1) ? -1 : 1) * binom(n + 1, k) * powl(t, n);
  }
  return s / fact;
}

/* weight of the tap at distance x (x inside the support) */
static ldbl spl_w(int order, ldbl x)
{
  if (order == 0) return 1.0L;            /* fastweight0: tap chosen by rounding */
  return bspl(order, x);
}

/* derivative, with the reference's convention at the order-1 kink (x == 0 -> -1) */
static ldbl spl_g(int order, ldbl x)
{
  if (order == 0) return 0;
  if (order == 1) return x < 0 ? 1.0L : -1.0L;
  return bspl(order - 1, x + 0.5L) - bspl(order - 1, x - 0.5L);
}

/* second derivative; order 2 at |x| == 1/2 takes the outer piece (= 1),
 * orders 5-7 are evaluated (the reference returns 0 there, see the
 * product's documentation of that deviation). */
static ldbl spl_h(int order, ldbl x)
{
  if (order <= 1) return 0;
  if (order == 2) { ldbl a = fabsl(x); return a < 0.5L ? -2.0L : 1.0L; }
  return bspl(order - 2, x + 1) - 2 * bspl(order - 2, x) + bspl(order - 2, x - 1);
}

/* first tap: interpolation_common.h bounds0..bounds7 */
static int64_t spl_first(int order, real x)
{
  if (order == 0) return (int64_t)round(x);          /* half away from zero */
  return (int64_t)floor((double)x - (double)(order - 1) / 2.0);
}

/* ------------------------------------------------------------------------ */
/* Per-axis tap tables                                                      */
/* ------------------------------------------------------------------------ */
typedef struct {
  int ntap;
  int64_t idx[8];
  int sgn[8];
  real w[8], g[8], h[8];
} taps_t;

static void make_taps(taps_t *t, int order, int bound, real x, int64_t n,
                      int need_g, int need_h)
{
  int64_t lo = spl_first(order, x);
  t->ntap = order + 1;
  for (int k = 0; k <= order; ++k) {
    int64_t b = lo + k;
    real d = x - (real)b;                 /* pushpull_common.cpp:1127 (in real) */
    t->w[k] = (real)spl_w(order, (ldbl)d);
    t->g[k] = need_g ? (real)spl_g(order, (ldbl)d) : 0;
    t->h[k] = need_h ? (real)spl_h(order, (ldbl)d) : 0;
    t->sgn[k] = fold_sign(bound, b, n);
    t->idx[k] = fold_index(bound, b, n);
  }
}

static void unit_taps(taps_t *t)
{
  t->ntap = 1; t->idx[0] = 0; t->sgn[0] = 1; t->w[0] = 1; t->g[0] = 0; t->h[0] = 0;
}

/* ------------------------------------------------------------------------ */
/* Main entry                                                               */
/* ------------------------------------------------------------------------ */
/*
 * Layouts (contiguous, dims beyond `dim` have size 1):
 *   src  [N, C, X, Y, Z]            (source of pull / sgrad / grad)
 *   grid [N, Xt, Yt, Zt, dim]
 *   trgt [N, C, Xt, Yt, Zt (, K)]   (K = 0 or dim; NULL for count backward)
 *   out  pull  [N, C, Xt, Yt, Zt]       sgrad [N, C, Xt, Yt, Zt, dim]
 *        push  [N, C, X, Y, Z]          count [N, 1, X, Y, Z]
 *   gout [N, Xt, Yt, Zt, dim]       (grid gradient, when do_grad)
 * `out` and `gout` must be allocated by the caller; this function
 * zero-initialises push/count/grad outputs like the reference's allocator
 * (pushpull_common.cpp:475-571).
 *
 * quirks, a bit set:
 *   1  reproduces pushpull_common.cpp:591 (the generic path uses the y-axis
 *      order for z when orders are not all equal);
 *   2  decides out-of-bound samples in float arithmetic, as the reference
 *      does for float32 data (inbounds<float>, bounds_common.h:168-171, with
 *      tol = (float)TINY, pushpull_common.cpp:900); weights and sums keep
 *      `real`.
 */
int OR_FN(or_pushpull)(
    int dim, const int *bound, const int *order, int extrapolate,
    int do_pull, int do_push, int do_count, int do_grad, int do_sgrad,
    int64_t N, int64_t C, const int64_t *src_size, const int64_t *trgt_size,
    int trgt_K, const real *src, const real *trgt, const real *grid,
    real *out, real *gout, int quirks)
{
  int64_t X = src_size[0], Y = dim > 1 ? src_size[1] : 1, Z = dim > 2 ? src_size[2] : 1;
  int64_t Xt = trgt_size[0], Yt = dim > 1 ? trgt_size[1] : 1, Zt = dim > 2 ? trgt_size[2] : 1;
  int b[3], o[3];
  /* bound/order arrays always hold 3 entries, padded by repeating the last
   * given entry (pushpull_common.cpp:96-110) */
  for (int d = 0; d < 3; ++d) { b[d] = bound[d]; o[d] = order[d]; }
  int iso = o[0] == o[1] && o[0] == o[2];
  int ogen[3] = {o[0], o[1], o[2]};
  if ((quirks & 1) && !iso && dim == 3) ogen[2] = o[1];
  int64_t vin = X * Y * Z, vout = Xt * Yt * Zt;
  int64_t nout_c = do_sgrad ? (int64_t)dim : 1;

  /* output initialisation rules (allocator) */
  if (do_push) memset(out, 0, sizeof(real) * N * C * vin);
  if (do_count) memset(out, 0, sizeof(real) * N * vin);
  if (do_grad) memset(gout, 0, sizeof(real) * N * vout * dim);
  if (do_sgrad && iso && (o[0] == 0 || (o[0] == 1 && dim == 1)))
    memset(out, 0, sizeof(real) * N * C * vout * dim);

  real tol = (real)((extrapolate & 2) ? 0.5 + 5e-2 : 5e-2);
  real *target = (real *)malloc(sizeof(real) * (C > 0 ? C : 1) * 3);

  for (int64_t n = 0; n < N; ++n)
  for (int64_t i = 0; i < vout; ++i) {
    const real *gp = grid + (n * vout + i) * dim;
    real xs[3] = {gp[0], dim > 1 ? gp[1] : 0, dim > 2 ? gp[2] : 0};
    int64_t sz[3] = {X, Y, Z};
    /* check{1,2,3}d: pushpull_common.cpp:925-1071 */
    int inb = 1;
    for (int d = 0; d < dim; ++d) {
      if (quirks & 2) {
        const float xf = (float)xs[d], tf = (float)tol;
        inb = inb && (xf >= -tf && xf < (float)(sz[d] - 1) + tf);
      } else {
        inb = inb && (xs[d] >= -tol && xs[d] < (real)(sz[d] - 1) + tol);
      }
    }
    if (!((extrapolate & 1) || inb)) {
      if (do_pull || do_sgrad)
        for (int64_t c = 0; c < C; ++c)
          for (int64_t k = 0; k < nout_c; ++k)
            out[((n * C + c) * vout + i) * nout_c + k] = 0;
      if (do_grad)
        for (int d = 0; d < dim; ++d) gout[(n * vout + i) * dim + d] = 0;
      continue;
    }
    if (do_grad || do_push)
      for (int64_t c = 0; c < C; ++c)
        for (int k = 0; k < (trgt_K ? 3 : 1); ++k)
          target[c + k * C] = (trgt && (k < trgt_K || k == 0))
              ? trgt[((n * C + c) * vout + i) * (trgt_K ? trgt_K : 1) + (trgt_K ? k : 0)]
              : 0;

    /* ---------------- nearest (iso order 0): :2313-2444 ---------------- */
    if (iso && o[0] == 0) {
      int64_t off = 0; int s = 1;
      int64_t strides[3] = {Y * Z, Z, 1};
      for (int d = 0; d < dim; ++d) {
        int64_t j = (int64_t)round(xs[d]);
        s *= fold_sign(b[d], j, sz[d]);
        off += fold_index(b[d], j, sz[d]) * strides[d];
      }
      for (int64_t c = 0; c < C; ++c) {
        if (do_pull) {
          real v = s == 0 ? 0 : (s < 0 ? -src[(n * C + c) * vin + off] : src[(n * C + c) * vin + off]);
          out[(n * C + c) * vout + i] = v;
        } else if (do_push && trgt_K == 0) {
          if (s) out[(n * C + c) * vin + off] += s < 0 ? -target[c] : target[c];
        }
      }
      if (do_count && !do_pull && !do_push && s)
        out[n * vin + off] += (real)s;
      continue;
    }

    /* ---------------- tap tables ---------------- */
    taps_t T[3];
    int lin = iso && o[0] == 1;
    for (int d = 0; d < 3; ++d) {
      if (d >= dim) { unit_taps(&T[d]); continue; }
      if (lin) {
        /* linear fast path weights: dx1 = x - floor(x), dx0 = 1 - dx1 (:1658-1667) */
        int64_t j0 = (int64_t)floor(xs[d]);
        real d1 = xs[d] - (real)j0;
        real d0 = (real)(1. - d1);
        T[d].ntap = 2;
        T[d].w[0] = d0; T[d].w[1] = d1;
        T[d].g[0] = -1; T[d].g[1] = 1;
        T[d].h[0] = 0; T[d].h[1] = 0;
        for (int k = 0; k < 2; ++k) {
          T[d].sgn[k] = fold_sign(b[d], j0 + k, sz[d]);
          T[d].idx[k] = fold_index(b[d], j0 + k, sz[d]);
        }
      } else {
        make_taps(&T[d], ogen[d], b[d], xs[d], sz[d],
                  do_grad || do_sgrad || (do_push && trgt_K), do_grad && trgt_K);
      }
    }

    if (lin && do_grad && trgt_K == 0) {
      /* linear fast path, grid gradient: channel-outer accumulation (:1710-1760) */
      real gg[3] = {0, 0, 0};
      for (int64_t c = 0; c < C; ++c) {
        real tv = trgt ? target[c] : (real)1;
        for (int kz = 0; kz < T[2].ntap; ++kz)
        for (int ky = 0; ky < T[1].ntap; ++ky)
        for (int kx = 0; kx < T[0].ntap; ++kx) {
          int s = T[0].sgn[kx] * T[1].sgn[ky] * T[2].sgn[kz];
          int64_t off = (T[0].idx[kx] * Y + T[1].idx[ky]) * Z + T[2].idx[kz];
          real v = s == 0 ? 0 : (s < 0 ? -src[(n * C + c) * vin + off] : src[(n * C + c) * vin + off]);
          if (trgt) v *= tv;
          gg[0] += (T[0].g[kx] * T[1].w[ky]) * T[2].w[kz] * v;
          if (dim > 1) gg[1] += (T[0].w[kx] * T[1].g[ky]) * T[2].w[kz] * v;
          if (dim > 2) gg[2] += (T[0].w[kx] * T[1].w[ky]) * T[2].g[kz] * v;
        }
      }
      for (int d = 0; d < dim; ++d) gout[(n * vout + i) * dim + d] = gg[d];
    }

    /* ---------------- generic tap loop (z outer, x inner): :1176-1290 ---------------- */
    if (do_pull || do_sgrad)
      for (int64_t c = 0; c < C; ++c)
        for (int64_t k = 0; k < nout_c; ++k)
          out[((n * C + c) * vout + i) * nout_c + k] = 0;
    real og[3] = {0, 0, 0};
    int grad_generic = do_grad && !(lin && trgt_K == 0);
    for (int kz = 0; kz < T[2].ntap; ++kz)
    for (int ky = 0; ky < T[1].ntap; ++ky)
    for (int kx = 0; kx < T[0].ntap; ++kx) {
      int s = T[0].sgn[kx] * T[1].sgn[ky] * T[2].sgn[kz];
      int64_t off = (T[0].idx[kx] * Y + T[1].idx[ky]) * Z + T[2].idx[kz];
      real wx = T[0].w[kx], wy = T[1].w[ky], wz = T[2].w[kz];
      real gx = T[0].g[kx], gy = T[1].g[ky], gz = T[2].g[kz];
      real hx = T[0].h[kx], hy = T[1].h[ky], hz = T[2].h[kz];
#define SRC(c) (s == 0 ? (real)0 : (s < 0 ? -src[((n) * C + (c)) * vin + off] : src[((n) * C + (c)) * vin + off]))
      if (do_pull) {
        for (int64_t c = 0; c < C; ++c)
          out[(n * C + c) * vout + i] += SRC(c) * (wx * wy * wz);
      } else if (do_sgrad) {
        for (int64_t c = 0; c < C; ++c) {
          real v = SRC(c);
          real *op = out + ((n * C + c) * vout + i) * dim;
          op[0] += v * (gx * wy * wz);
          if (dim > 1) op[1] += v * (wx * gy * wz);
          if (dim > 2) op[2] += v * (wx * wy * gz);
        }
      } else if (do_push) {
        if (s) for (int64_t c = 0; c < C; ++c) {
          real v;
          if (trgt_K == 0) v = (wx * wy * wz) * target[c];
          else {
            v = (gx * wy * wz) * target[c];
            if (dim > 1) v += (wx * gy * wz) * target[c + C];
            if (dim > 2) v += (wx * wy * gz) * target[c + 2 * C];
          }
          out[(n * C + c) * vin + off] += s < 0 ? -v : v;
        }
      } else if (do_count) {
        if (s) out[n * vin + off] += s < 0 ? -(wx * wy * wz) : (wx * wy * wz);
      }
      if (grad_generic) {
        if (trgt_K == 0) {
          real dot = 0;
          for (int64_t c = 0; c < C; ++c) { real v = SRC(c); dot += trgt ? v * target[c] : v; }
          og[0] += (gx * wy * wz) * dot;
          if (dim > 1) og[1] += (wx * gy * wz) * dot;
          if (dim > 2) og[2] += (wx * wy * gz) * dot;
        } else {
          real d0 = 0, d1 = 0, d2 = 0;
          for (int64_t c = 0; c < C; ++c) {
            real v = SRC(c);
            d0 += v * target[c];
            if (dim > 1) d1 += v * target[c + C];
            if (dim > 2) d2 += v * target[c + 2 * C];
          }
          og[0] += (hx * wy * wz) * d0 + (gx * gy * wz) * d1 + (gx * wy * gz) * d2;
          if (dim > 1) og[1] += (gx * gy * wz) * d0 + (wx * hy * wz) * d1 + (wx * gy * gz) * d2;
          if (dim > 2) og[2] += (gx * wy * gz) * d0 + (wx * gy * gz) * d1 + (wx * wy * hz) * d2;
        }
      }
#undef SRC
    }
    if (grad_generic)
      for (int d = 0; d < dim; ++d) gout[(n * vout + i) * dim + d] = og[d];
  }
  free(target);
  return 0;
}

/* Exposed for unit tests of the basis (closed-form / invariant checks). */
double OR_FN(or_spline_weight)(int order, double x) { return (double)(order == 0 ? (fabs(x) < 0.5) : bspl(order, x)); }
double OR_FN(or_spline_grad)(int order, double x) { return (double)spl_g(order, x); }
double OR_FN(or_spline_hess)(int order, double x) { return (double)spl_h(order, x); }
int64_t OR_FN(or_bound_index)(int b, int64_t i, int64_t n) { return fold_index(b, i, n); }
int OR_FN(or_bound_sign)(int b, int64_t i, int64_t n) { return fold_sign(b, i, n); }
