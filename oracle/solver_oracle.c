/*
 * ORACLE — test infrastructure only (see oracle/__init__.py).
 *
 * FP64 restatement of nitorch's displacement-field solver path for 3-D
 * grids [N, 3, X, Y, Z] (contiguous):
 *   or_regulariser_grid : out = (H + L) x
 *       (nitorch/_C/src/impl/regulariser_grid_common.cpp, set_kernel :265-298,
 *        vel2mom3d_all :823-923, matvec3d_{sym,diag,eye} :675-740)
 *   or_relax_grid       : multi-colour Gauss-Seidel sweeps
 *       (relax_grid_common.cpp, set_kernel :268-307, bandwidth/folds :309-332,
 *        :454-476, loops :672-781, relax3d_all :1064-1166, invert3d_* :829-976)
 *   or_prolong / or_restrict : FMG transfer operators
 *       (multires_common.cpp: resize3d :430-499, restrict3d :586-669,
 *        shapes :800-839)
 *   or_precond_grid     : block-Jacobi solve (H + diag L)^-1 g
 *       (precond_grid_common.cpp, set_kernel :266-298, invert3d_* :620-700)
 *   or_pcg_grid         : preconditioned CG (pcg_common.cpp:38-85, :125-163)
 *   or_fmg_grid         : full multigrid (fmg_common.cpp:45-222, :356-533),
 *                         Gauss-Seidel or PCG smoother (:375-387)
 *   or_cg_grid          : plain CG on (H + L) x = g (exact solutions for tests)
 * Every entry point takes an optional voxel-wise weight map w [N, 1, X, Y, Z]
 * (RLS weights; absolute and membrane energies only, as the reference):
 *   regulariser_grid_common.cpp:1171-1248 (vel2mom3d_rls_membrane/absolute),
 *   relax_grid_common.cpp:1415-1500 (relax3d_rls_*),
 *   precond_grid_common.cpp:800-840 (precond3d_rls_*),
 *   fmg_common.cpp:176-181 (weights restricted like gradients, unscaled).
 *
 * Quirk switches reproduce the reference's own behaviour where it departs
 * from the intended algorithm, so that the restatement can be compared
 * exactly with the compiled reference:
 *   quirk_fold  : 1 = fold start indices of banded relax use `% Y`, `% Z`
 *                 (relax_grid_common.cpp:469-470), and the sign typos of
 *                 relax3d_all's third component (:1154-1158) are kept;
 *                 0 = the reference's colour order without those quirks;
 *                 2 = the product's colour classes (lin_colouring, else
 *                 axis_colours), 3 = the same visited in reverse within each
 *                 class (equal to 2 exactly when every class is independent);
 *   quirk_zscale: FMG rescales Z-slices instead of displacement channels
 *                 (fmg_common.cpp:392-489).
 * With both off, the restatement follows the intended semantics (per-channel
 * rescaling as in nitorch/spatial/_solvers.py:700-740, all colour classes
 * visited), which is what the product implements.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ONE_PLUS_TINY 1.000001

enum { B_REPLICATE = 0, B_DCT1 = 1, B_DCT2 = 2, B_DST1 = 3, B_DST2 = 4,
       B_DFT = 5, B_ZERO = 7 };

static int64_t pmod(int64_t a, int64_t m) { int64_t r = a % m; return r < 0 ? r + m : r; }

static int64_t fidx(int b, int64_t i, int64_t n)
{
  switch (b) {
  case B_REPLICATE: return i < 0 ? 0 : (i >= n ? n - 1 : i);
  case B_DCT1: { if (n == 1) return 0; int64_t p = 2 * (n - 1), k = pmod(i < 0 ? -i : i, p); return k >= n ? p - k : k; }
  case B_DCT2: case B_DST2: { int64_t p = 2 * n, k = pmod(i, p); return k >= n ? p - 1 - k : k; }
  case B_DST1: { if (n == 1) return 0; int64_t p = 2 * (n + 1); int64_t k = (i == -1) ? 0 : (i < 0 ? -i - 2 : i);
                 k = pmod(k, p); if (k == n) return n - 1; return k > n ? p - k - 2 : k; }
  case B_DFT: return pmod(i, n);
  default: return (i < 0 || i >= n) ? 0 : i;
  }
}

static int fsgn(int b, int64_t i, int64_t n)
{
  switch (b) {
  case B_DST1: {
    if (n == 1) return 1;
    int64_t p = 2 * (n + 1);
    int64_t k = i < 0 ? n - i - 1 : i;
    k %= p;
    if (k % (n + 1) == n) return 0;
    return ((k / (n + 1)) % 2) ? -1 : 1;
  }
  case B_DST2: { int64_t k = i < 0 ? n - i - 1 : i; return ((k / n) % 2) ? -1 : 1; }
  case B_ZERO: return (i < 0 || i >= n) ? 0 : 1;
  default: return 1;
  }
}

typedef struct {
  double vx[3];      /* 1 / voxel_size^2 */
  double lam0, lam1, lam2, mu, lam;
  double w000, w100, w010, w001, w200, w020, w002, w110, w101, w011;
  double wd[3][4];   /* per component: centre, x, y, z (Lame + scaled bending/membrane) */
  double w2;
} weights_t;

/* regulariser_grid_common.cpp:265-298 / relax_grid_common.cpp:268-307 */
static void set_weights(weights_t *W, const double *vs, double absolute, double membrane,
                        double bending, double lame_shear, double lame_div)
{
  for (int d = 0; d < 3; ++d) W->vx[d] = 1.0 / (vs[d] * vs[d]);
  double v0 = W->vx[0], v1 = W->vx[1], v2 = W->vx[2];
  double l0 = absolute, l1 = membrane, l2 = bending, mu = lame_shear, la = lame_div;
  W->lam0 = l0; W->lam1 = l1; W->lam2 = l2; W->mu = mu; W->lam = la;
  W->w000 = l2 * (6.0 * (v0 * v0 + v1 * v1 + v2 * v2) + 8.0 * (v0 * v1 + v0 * v2 + v1 * v2)) + l1 * 2.0 * (v0 + v1 + v2) + l0;
  W->w100 = l2 * (-4.0 * v0 * (v0 + v1 + v2)) - l1 * v0;
  W->w010 = l2 * (-4.0 * v1 * (v0 + v1 + v2)) - l1 * v1;
  W->w001 = l2 * (-4.0 * v2 * (v0 + v1 + v2)) - l1 * v2;
  W->w200 = l2 * v0 * v0; W->w020 = l2 * v1 * v1; W->w002 = l2 * v2 * v2;
  W->w110 = l2 * 2.0 * v0 * v1; W->w101 = l2 * 2.0 * v0 * v2; W->w011 = l2 * 2.0 * v1 * v2;
  double v[3] = {v0, v1, v2};
  double w1[3] = {W->w100, W->w010, W->w001};
  for (int d = 0; d < 3; ++d) {
    /* centre: 2 mu (sum_e v_e + v_d) / v_d + 2 lam + w000 / v_d */
    W->wd[d][0] = 2.0 * mu * (v0 + v1 + v2 + v[d]) / v[d] + 2.0 * la + W->w000 / v[d];
    for (int e = 0; e < 3; ++e)
      W->wd[d][1 + e] = (e == d ? (-2.0 * mu - la) : (-mu * v[e] / v[d])) + w1[e] / v[d];
  }
  W->w2 = 0.25 * mu + 0.25 * la;
}

typedef struct { int64_t X, Y, Z; int b[3]; } geom_t;

static inline double getv(const double *p, const geom_t *G, int64_t x, int64_t y, int64_t z)
{
  int s = fsgn(G->b[0], x, G->X) * fsgn(G->b[1], y, G->Y) * fsgn(G->b[2], z, G->Z);
  if (!s) return 0.0;
  double v = p[(fidx(G->b[0], x, G->X) * G->Y + fidx(G->b[1], y, G->Y)) * G->Z + fidx(G->b[2], z, G->Z)];
  return s < 0 ? -v : v;
}

/* value at (x, y, z) multiplied by the sign computed at (sx, sy, sz) */
static inline double getv_s(const double *p, const geom_t *G, int64_t x, int64_t y, int64_t z,
                            int64_t sx, int64_t sy, int64_t sz)
{
  int s = fsgn(G->b[0], sx, G->X) * fsgn(G->b[1], sy, G->Y) * fsgn(G->b[2], sz, G->Z);
  if (!s) return 0.0;
  double v = p[(fidx(G->b[0], x, G->X) * G->Y + fidx(G->b[1], y, G->Y)) * G->Z + fidx(G->b[2], z, G->Z)];
  return s < 0 ? -v : v;
}

/* (L x)_d at voxel (x, y, z) for component d; x3 = [3][X*Y*Z].
 * quirk_sign reproduces the sign typos of relax3d_all for the third
 * component (relax_grid_common.cpp:1154-1158): three diagonal taps take the
 * sign of the mirrored tap. */
static double lx_component_q(const double *x3, const geom_t *G, const weights_t *W,
                             int d, int64_t x, int64_t y, int64_t z, int quirk_sign);

static double lx_component(const double *x3, const geom_t *G, const weights_t *W,
                           int d, int64_t x, int64_t y, int64_t z)
{
  return lx_component_q(x3, G, W, d, x, y, z, 0);
}

static double lx_component_q(const double *x3, const geom_t *G, const weights_t *W,
                             int d, int64_t x, int64_t y, int64_t z, int quirk_sign)
{
  int64_t V = G->X * G->Y * G->Z;
  const double *u = x3 + d * V;
  double c = u[(x * G->Y + y) * G->Z + z];
#define U(dx, dy, dz) (getv(u, G, x + (dx), y + (dy), z + (dz)) - c)
  double lame = W->wd[d][1] * (U(-1, 0, 0) + U(1, 0, 0))
              + W->wd[d][2] * (U(0, -1, 0) + U(0, 1, 0))
              + W->wd[d][3] * (U(0, 0, -1) + U(0, 0, 1));
  /* cross-component Lame terms: w2 * (mixed second difference) */
  double cross = 0;
  int o1 = (d + 1) % 3, o2 = (d + 2) % 3;
  int others[2] = {d == 0 ? 1 : 0, d == 2 ? 1 : 2};
  (void)o1; (void)o2;
  for (int k = 0; k < 2; ++k) {
    int e = others[k];
    const double *q = x3 + e * V;
    int64_t ax[3] = {0, 0, 0}, bx[3] = {0, 0, 0};
    /* +d-1e  -d+1e... sign pattern: f(+1d,-1e) - f(+1d,+1e) + f(-1d,+1e) - f(-1d,-1e) */
    ax[d] = 1; bx[e] = 1;
    cross += getv(q, G, x + ax[0] - bx[0], y + ax[1] - bx[1], z + ax[2] - bx[2])
           - getv(q, G, x + ax[0] + bx[0], y + ax[1] + bx[1], z + ax[2] + bx[2])
           + getv(q, G, x - ax[0] + bx[0], y - ax[1] + bx[1], z - ax[2] + bx[2])
           - getv(q, G, x - ax[0] - bx[0], y - ax[1] - bx[1], z - ax[2] - bx[2]);
  }
  double q110 = U(-1, 1, 0), q101 = U(-1, 0, 1), q011 = U(0, -1, 1);
  if (quirk_sign && d == 2) {
    q110 = getv_s(u, G, x - 1, y + 1, z, x + 1, y + 1, z) - c;
    q101 = getv_s(u, G, x - 1, y, z + 1, x + 1, y, z + 1) - c;
    q011 = getv_s(u, G, x, y - 1, z + 1, x, y + 1, z + 1) - c;
  }
  double bend = W->lam0 * c
      + W->w110 * (U(-1, -1, 0) + U(1, -1, 0) + q110 + U(1, 1, 0))
      + W->w101 * (U(-1, 0, -1) + U(1, 0, -1) + q101 + U(1, 0, 1))
      + W->w011 * (U(0, -1, -1) + U(0, 1, -1) + q011 + U(0, 1, 1))
      + W->w200 * (U(-2, 0, 0) + U(2, 0, 0))
      + W->w020 * (U(0, -2, 0) + U(0, 2, 0))
      + W->w002 * (U(0, 0, -2) + U(0, 0, 2));
#undef U
  return lame + W->w2 * cross + bend / W->vx[d];
}

/* H at voxel: CC = 0 (none), 1 (scalar), 3 (diagonal), 6 (h00 h11 h22 h01 h02 h12) */
static void hmat(const double *h, int CC, int64_t V, int64_t i, double M[3][3])
{
  memset(M, 0, sizeof(double) * 9);
  if (CC == 0) return;
  if (CC == 1) { M[0][0] = M[1][1] = M[2][2] = h[i]; return; }
  M[0][0] = h[i]; M[1][1] = h[V + i]; M[2][2] = h[2 * V + i];
  if (CC == 6) {
    M[0][1] = M[1][0] = h[3 * V + i];
    M[0][2] = M[2][0] = h[4 * V + i];
    M[1][2] = M[2][1] = h[5 * V + i];
  }
}

/* RLS face weights at voxel (x, y, z) of the weight map wn: for the six
 * first-order neighbours (x-1, x+1, y-1, y+1, z-1, z+1), the membrane weight
 * of the axis (-membrane / voxel_size^2) times (w_centre + w_neighbour), the
 * neighbour's weight read with the bound's index and sign like any field
 * value.  Returns the centre weight. */
static double rls_faces(const double *wn, const geom_t *G, const weights_t *W,
                        int64_t x, int64_t y, int64_t z, double fw[6])
{
  double wc = wn[(x * G->Y + y) * G->Z + z];
  const int64_t off[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
  for (int k = 0; k < 6; ++k) {
    double m = -W->lam1 * W->vx[k / 2];
    fw[k] = m * (wc + getv(wn, G, x + off[k][0], y + off[k][1], z + off[k][2]));
  }
  return wc;
}

/* ((L_w x)_d at (x, y, z), and the diagonal of L_w for component d):
 *   (absolute w_c x_c + 1/2 sum_faces fw (x_nb - x_c)) / vx_d
 *   (absolute w_c     - 1/2 sum_faces fw              ) / vx_d            */
static double rls_component(const double *x3, const double *wn, const geom_t *G, const weights_t *W,
                            int d, int64_t x, int64_t y, int64_t z, double *diag)
{
  int64_t V = G->X * G->Y * G->Z;
  double fw[6];
  double wc = rls_faces(wn, G, W, x, y, z, fw);
  const double *xd = x3 + d * V;
  double c = xd[(x * G->Y + y) * G->Z + z];
  const int64_t off[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
  double acc = 0, fs = 0;
  for (int k = 0; k < 6; ++k) {
    acc += fw[k] * (getv(xd, G, x + off[k][0], y + off[k][1], z + off[k][2]) - c);
    fs += fw[k];
  }
  if (diag) *diag = (W->lam0 * wc - 0.5 * fs) / W->vx[d];
  return (W->lam0 * wc * c + 0.5 * acc) / W->vx[d];
}

void or_regulariser_grid(const double *x, double *out, const double *h, int CC,
                         int64_t N, int64_t X, int64_t Y, int64_t Z, const int *bound,
                         const double *voxel_size, double absolute, double membrane,
                         double bending, double lame_shear, double lame_div, const double *w)
{
  weights_t W; set_weights(&W, voxel_size, absolute, membrane, bending, lame_shear, lame_div);
  geom_t G = {X, Y, Z, {bound[0], bound[1], bound[2]}};
  int64_t V = X * Y * Z;
  for (int64_t n = 0; n < N; ++n) {
    const double *xn = x + n * 3 * V;
    const double *hn = h ? h + n * CC * V : NULL;
    const double *wn = w ? w + n * V : NULL;
    double *on = out + n * 3 * V;
    for (int64_t i = 0; i < V; ++i) {
      int64_t xx = i / (Y * Z), yy = (i / Z) % Y, zz = i % Z;
      double M[3][3];
      hmat(hn, hn ? CC : 0, V, i, M);
      for (int d = 0; d < 3; ++d) {
        double v = wn ? rls_component(xn, wn, &G, &W, d, xx, yy, zz, NULL) : lx_component(xn, &G, &W, d, xx, yy, zz);
        for (int e = 0; e < 3; ++e) v += M[d][e] * xn[e * V + i];
        on[d * V + i] = v;
      }
    }
  }
}

/* One Gauss-Seidel update of voxel i (relax3d_all + invert3d_*). */
static void relax_voxel(double *x3, const double *g3, const double *h, int CC, const geom_t *G,
                        const weights_t *W, int64_t V, int64_t xx, int64_t yy, int64_t zz, int quirk,
                        const double *wn)
{
  int64_t i = (xx * G->Y + yy) * G->Z + zz;
  double r[3], M[3][3], dg[3];
  for (int d = 0; d < 3; ++d) {
    if (wn) {
      r[d] = g3[d * V + i] - rls_component(x3, wn, G, W, d, xx, yy, zz, &dg[d]);
    } else {
      r[d] = g3[d * V + i] - lx_component_q(x3, G, W, d, xx, yy, zz, quirk);
      dg[d] = W->wd[d][0] * ONE_PLUS_TINY;
    }
  }
  hmat(h, h ? CC : 0, V, i, M);
  for (int d = 0; d < 3; ++d)
    for (int e = 0; e < 3; ++e) r[d] -= M[d][e] * x3[e * V + i];
  double A[3][3];
  for (int d = 0; d < 3; ++d)
    for (int e = 0; e < 3; ++e) A[d][e] = M[d][e] * (d == e ? ONE_PLUS_TINY : 1.0);
  for (int d = 0; d < 3; ++d) A[d][d] += dg[d];
  /* 3x3 symmetric solve by the adjugate (relax_grid_common.cpp:840-852) */
  double a = A[0][0], b = A[1][1], c = A[2][2], p = A[0][1], q = A[0][2], s = A[1][2];
  double det = a * b * c - a * s * s - b * q * q - c * p * p + 2 * p * q * s;
  double idt = 1.0 / det;
  double d0 = idt * (r[0] * (b * c - s * s) + r[1] * (q * s - p * c) + r[2] * (p * s - q * b));
  double d1 = idt * (r[0] * (q * s - p * c) + r[1] * (a * c - q * q) + r[2] * (p * q - a * s));
  double d2 = idt * (r[0] * (p * s - q * b) + r[1] * (p * q - a * s) + r[2] * (a * b - p * p));
  x3[i] += d0; x3[V + i] += d1; x3[2 * V + i] += d2;
}

/* The product's colour classes (mode 2).
 *
 * Linear colouring: voxel (x, y, z) has colour (a x + b y + c z) mod m.  A
 * class is an independent set when no coupling offset v of the stencil
 * (bending: |v|_1 <= 2; Lame only: |v|_1 <= 2 and |v|_inf <= 1; membrane
 * only: |v|_1 = 1) has a v0 + b v1 + c v2 = 0 (mod m), where on a circulant
 * axis v is taken modulo its length (then coefficient * length = 0 (mod m)
 * is needed for the colour to be defined on the torus); on the other bounds
 * a folded neighbour is never farther than the unfolded one.  The smallest m
 * (up to 32) wins, ties by the smallest gcd(c, m), then c, a, b; the class
 * order is colour 0, 1, ..., m - 1.  Shapes without such a colouring (odd
 * circulant lengths, for instance) use the per-axis period classes below. */
typedef struct { int m, a, b, c; } lincol_t;

static int gcd_i(int u, int v) { while (v) { int t = u % v; u = v; v = t; } return u; }

static int lin_colouring(const int64_t sz[3], const int *bound, int bw, lincol_t *out)
{
  int off[124][3], no = 0;
  for (int i = -2; i <= 2; ++i)
    for (int j = -2; j <= 2; ++j)
      for (int k = -2; k <= 2; ++k) {
        int v[3] = {i, j, k}, l1 = abs(i) + abs(j) + abs(k), li = 0;
        for (int d = 0; d < 3; ++d) if (abs(v[d]) > li) li = abs(v[d]);
        int coupled = bw == 3 ? l1 <= 2 : (bw == 2 ? (l1 <= 2 && li <= 1) : (bw == 0 ? l1 == 1 : 0));
        if (!l1 || !coupled) continue;
        int ok = 1, nz = 0;
        for (int d = 0; d < 3; ++d) {
          int64_t n = sz[d];
          if (n == 1) v[d] = 0;
          else if (bound[d] == B_DFT) { int64_t r = pmod(v[d], n); if (r > n / 2) r -= n; v[d] = (int)r; }
          else if (abs(v[d]) > n - 1) ok = 0;
          nz |= v[d] != 0;
        }
        if (ok && nz) { off[no][0] = v[0]; off[no][1] = v[1]; off[no][2] = v[2]; ++no; }
      }
  for (int m = 1; m <= 32; ++m) {
    int found = 0, bg = 0;
    for (int c = 0; c < m; ++c) {
      int g = gcd_i(c, m);
      if (found && g >= bg) continue;
      for (int a = 0; a < m; ++a)
        for (int b = 0; b < m; ++b) {
          int co[3] = {a, b, c}, ok = 1;
          for (int d = 0; d < 3 && ok; ++d)
            if (bound[d] == B_DFT && sz[d] > 1 && (co[d] * sz[d]) % m) ok = 0;
          for (int o = 0; o < no && ok; ++o)
            if (pmod((int64_t)a * off[o][0] + (int64_t)b * off[o][1] + (int64_t)c * off[o][2], m) == 0) ok = 0;
          if (ok && (!found || g < bg)) {
            found = 1; bg = g; out->m = m; out->a = a; out->b = b; out->c = c;
          }
        }
    }
    if (found) return 1;
  }
  return 0;
}

/* Colour classes along one axis of length n for period F: class c covers
 * start + F*i.  With circulant bounds and n % F != 0 the wrap-around would put
 * two coupled voxels in one class, so the last n % F slices get classes of
 * their own (the product's ordering; the reference uses plain x % F). */
typedef struct { int64_t start[8], step[8], count[8]; int n; } colours_t;

static void axis_colours(colours_t *c, int64_t n, int64_t F, int bound, int safe)
{
  int64_t Fe = n < F ? n : F;
  int64_t r = (safe && bound == B_DFT && n > F) ? n % F : 0;
  int64_t nF = n - r;
  c->n = 0;
  for (int64_t k = 0; k < Fe; ++k) {
    c->start[c->n] = k; c->step[c->n] = F; c->count[c->n] = (nF - k + F - 1) / F; c->n++;
  }
  for (int64_t k = 0; k < r; ++k) {
    c->start[c->n] = nF + k; c->step[c->n] = F; c->count[c->n] = 1; c->n++;
  }
}

/* order: 0 = the reference's folds (quirk_fold reproduces its start indices),
 *        1 = the same with the reference's quirks,
 *        2 = the product's colour classes (lin_colouring, else axis_colours),
 *        3 = as 2 with every class visited in reverse voxel order (a class is
 *            an independent set exactly when this changes nothing). */
void or_relax_grid(const double *g, const double *h, int CC, double *x,
                   int64_t N, int64_t X, int64_t Y, int64_t Z, const int *bound,
                   const double *voxel_size, double absolute, double membrane,
                   double bending, double lame_shear, double lame_div, int nb_iter,
                   int quirk_fold, const double *w)
{
  weights_t W; set_weights(&W, voxel_size, absolute, membrane, bending, lame_shear, lame_div);
  geom_t G = {X, Y, Z, {bound[0], bound[1], bound[2]}};
  int64_t V = X * Y * Z;
  int safe = quirk_fold >= 2;           /* 2, 3: product ordering */
  int rev = quirk_fold == 3;
  int quirk = quirk_fold == 1;
  /* bandwidth: bending 3, Lame 2, membrane 0 (red/black), else 1 */
  int bw = bending ? 3 : ((lame_shear || lame_div) ? 2 : (membrane ? 0 : 1));
  lincol_t L = {0, 0, 0, 0};
  int lin = 0, F = bw == 3 ? 3 : (bw == 1 ? 1 : 2);
  colours_t cx, cy, cz;
  if (safe) {
    const int64_t sz[3] = {X, Y, Z};
    axis_colours(&cx, X, F, bound[0], 1);
    axis_colours(&cy, Y, F, bound[1], 1);
    axis_colours(&cz, Z, F, bound[2], 1);
    lin = lin_colouring(sz, bound, bw, &L) && L.m <= cx.n * cy.n * cz.n;
  }
  for (int64_t n = 0; n < N; ++n) {
    double *xn = x + n * 3 * V;
    const double *gn = g + n * 3 * V;
    const double *hn = h ? h + n * CC * V : NULL;
    const double *wn = w ? w + n * V : NULL;
    for (int it = 0; it < nb_iter; ++it) {
      if (safe && lin) {
        for (int col = 0; col < L.m; ++col)
          for (int64_t t = 0; t < V; ++t) {
            int64_t i = rev ? V - 1 - t : t;
            int64_t xx = i / (Y * Z), yy = (i / Z) % Y, zz = i % Z;
            if (pmod(L.a * xx + L.b * yy + L.c * zz, L.m) == col)
              relax_voxel(xn, gn, hn, CC, &G, &W, V, xx, yy, zz, 0, wn);
          }
        continue;
      }
      if (safe) {
        for (int a = 0; a < cx.n; ++a)
          for (int b = 0; b < cy.n; ++b)
            for (int c = 0; c < cz.n; ++c) {
              int64_t cnt = cx.count[a] * cy.count[b] * cz.count[c];
              for (int64_t t = 0; t < cnt; ++t) {
                int64_t u = rev ? cnt - 1 - t : t;
                int64_t i = u / (cy.count[b] * cz.count[c]), j = (u / cz.count[c]) % cy.count[b], k = u % cz.count[c];
                relax_voxel(xn, gn, hn, CC, &G, &W, V, cx.start[a] + cx.step[a] * i,
                            cy.start[b] + cy.step[b] * j, cz.start[c] + cz.step[c] * k, 0, wn);
              }
            }
        continue;
      }
      if (bw == 0) {
        for (int rb = 0; rb < 2; ++rb)
          for (int64_t xx = 0; xx < X; ++xx)
            for (int64_t yy = 0; yy < Y; ++yy)
              for (int64_t zz = 0; zz < Z; ++zz)
                if ((xx + yy + zz) % 2 == rb) relax_voxel(xn, gn, hn, CC, &G, &W, V, xx, yy, zz, quirk, wn);
        continue;
      }
      int64_t Fx = X < bw ? X : bw, Fy = Y < bw ? Y : bw, Fz = Z < bw ? Z : bw;
      for (int64_t f = 0; f < Fx * Fy * Fz; ++f) {
        int64_t fx = f / (Fy * Fz);
        int64_t fy = quirk ? (f / Fz) % Y : (f / Fz) % Fy;
        int64_t fz = quirk ? f % Z : f % Fz;
        for (int64_t xx = fx; xx < X; xx += Fx)
          for (int64_t yy = fy; yy < Y; yy += Fy)
            for (int64_t zz = fz; zz < Z; zz += Fz)
              relax_voxel(xn, gn, hn, CC, &G, &W, V, xx, yy, zz, quirk, wn);
      }
    }
  }
}

/* Colouring the product uses for a shape, bounds and bandwidth: returns m
 * (linear colouring, coefficients in abc) or 0 (per-axis period classes,
 * their class count in abc[0]). */
int or_relax_plan(int64_t X, int64_t Y, int64_t Z, const int *bound, int bw, int *abc)
{
  const int64_t sz[3] = {X, Y, Z};
  int F = bw == 3 ? 3 : (bw == 1 ? 1 : 2);
  colours_t cx, cy, cz;
  axis_colours(&cx, X, F, bound[0], 1);
  axis_colours(&cy, Y, F, bound[1], 1);
  axis_colours(&cz, Z, F, bound[2], 1);
  lincol_t L;
  if (lin_colouring(sz, bound, bw, &L) && L.m <= cx.n * cy.n * cz.n) {
    abc[0] = L.a; abc[1] = L.b; abc[2] = L.c;
    return L.m;
  }
  abc[0] = cx.n * cy.n * cz.n; abc[1] = abc[2] = 0;
  return 0;
}

/* ---------------------------------------------------------------------- */
/* Transfer operators (multires_common.cpp).                              */
/* ---------------------------------------------------------------------- */
static double w_lin(double x) { x = fabs(x); return x < 1 ? 1 - x : 0; }
static double w_quad(double x) { x = fabs(x); if (x < 0.5) return 0.75 - x * x; if (x < 1.5) { x = 1.5 - x; return 0.5 * x * x; } return 0; }

/* prolongation (quadratic, 3 taps/dim): in [N,C,Xi,Yi,Zi] -> out [N,C,Xo,Yo,Zo] */
void or_prolong(const double *in, double *out, int64_t N, int64_t C,
                const int64_t *isz, const int64_t *osz, const int *bound)
{
  geom_t G = {isz[0], isz[1], isz[2], {bound[0], bound[1], bound[2]}};
  double sc[3];
  for (int d = 0; d < 3; ++d) sc[d] = (double)isz[d] / (double)osz[d];
  int64_t Vi = isz[0] * isz[1] * isz[2], Vo = osz[0] * osz[1] * osz[2];
  for (int64_t nc = 0; nc < N * C; ++nc)
    for (int64_t o = 0; o < Vo; ++o) {
      int64_t w[3] = {o / (osz[1] * osz[2]), (o / osz[2]) % osz[1], o % osz[2]};
      int64_t i0[3]; double wt[3][3];
      for (int d = 0; d < 3; ++d) {
        double x = (w[d] + 0.5) * sc[d] - 0.5;
        int64_t c = (int64_t)floor(x + 0.5);
        i0[d] = c - 1;
        for (int k = 0; k < 3; ++k) wt[d][k] = w_quad(x - (double)(c - 1 + k));
      }
      double acc = 0;
      for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) for (int c = 0; c < 3; ++c)
        acc += wt[0][a] * wt[1][b] * wt[2][c] * getv(in + nc * Vi, &G, i0[0] + a, i0[1] + b, i0[2] + c);
      out[nc * Vo + o] = acc;
    }
}

/* restriction (adjoint of linear, 4 taps/dim, weights w_lin(d/s)/s) */
void or_restrict(const double *in, double *out, int64_t N, int64_t C,
                 const int64_t *isz, const int64_t *osz, const int *bound)
{
  geom_t G = {isz[0], isz[1], isz[2], {bound[0], bound[1], bound[2]}};
  double sc[3];
  for (int d = 0; d < 3; ++d) sc[d] = (double)isz[d] / (double)osz[d];
  int64_t Vi = isz[0] * isz[1] * isz[2], Vo = osz[0] * osz[1] * osz[2];
  for (int64_t nc = 0; nc < N * C; ++nc)
    for (int64_t o = 0; o < Vo; ++o) {
      int64_t w[3] = {o / (osz[1] * osz[2]), (o / osz[2]) % osz[1], o % osz[2]};
      int64_t i0[3]; double wt[3][4];
      for (int d = 0; d < 3; ++d) {
        double x = (w[d] + 0.5) * sc[d] - 0.5;
        int64_t c = (int64_t)floor(x);
        i0[d] = c - 1;
        for (int k = 0; k < 4; ++k) wt[d][k] = w_lin((x - (double)(c - 1 + k)) / sc[d]) / sc[d];
      }
      double acc = 0;
      for (int a = 0; a < 4; ++a) for (int b = 0; b < 4; ++b) for (int c = 0; c < 4; ++c)
        acc += wt[0][a] * wt[1][b] * wt[2][c] * getv(in + nc * Vi, &G, i0[0] + a, i0[1] + b, i0[2] + c);
      out[nc * Vo + o] = acc;
    }
}

/* ---------------------------------------------------------------------- */
/* FMG (fmg_common.cpp:45-222, :356-533), relax smoother.                  */
/* ---------------------------------------------------------------------- */
#define MAXL 16

/* per-level rescaling of displacement-like (C=3) or Hessian-like (C=6) fields */
static void scale_fields(double *t, int64_t N, int64_t C, const int64_t *sz, const double *f,
                         int hess, int quirk)
{
  int64_t V = sz[0] * sz[1] * sz[2];
  for (int64_t n = 0; n < N; ++n) {
    double *tn = t + n * C * V;
    if (!quirk) {
      if (!hess) {
        for (int d = 0; d < 3 && d < C; ++d) for (int64_t i = 0; i < V; ++i) tn[d * V + i] *= f[d];
      } else {
        double s[6] = {f[0] * f[0], f[1] * f[1], f[2] * f[2], f[0] * f[1], f[0] * f[2], f[1] * f[2]};
        if (C == 1) s[0] = f[0] * f[0];
        for (int d = 0; d < C; ++d) for (int64_t i = 0; i < V; ++i) tn[d * V + i] *= s[d];
      }
    } else {
      /* reference: o.slice(-1, k, k+1) scales the k-th plane along Z of every channel */
      if (!hess) {
        for (int k = 0; k < 3; ++k)
          for (int64_t c = 0; c < C; ++c)
            for (int64_t xy = 0; xy < sz[0] * sz[1]; ++xy)
              if (k < sz[2]) tn[c * V + xy * sz[2] + k] *= f[k];
      } else {
        double s[6] = {f[0] * f[0], f[1] * f[1], f[2] * f[2], 1, f[0] * f[2], f[1] * f[2]};
        /* order of the reference's switch: plane 2 (f2^2), 5 (f1 f2), 4 (f0 f2), 1 (f1^2), 3 (f0 f1), 0 (f0^2) */
        int planes[6] = {2, 5, 4, 1, 3, 0};
        double fac[6] = {f[2] * f[2], f[1] * f[2], f[0] * f[2], f[1] * f[1], f[0] * f[1], f[0] * f[0]};
        (void)s;
        for (int j = 0; j < 6; ++j) {
          int k = planes[j];
          if (C == 1 && k != 0) continue;
          if (C <= 3 && k >= 3) continue;
          for (int64_t c = 0; c < C; ++c)
            for (int64_t xy = 0; xy < sz[0] * sz[1]; ++xy)
              if (k < sz[2]) tn[c * V + xy * sz[2] + k] *= fac[j];
        }
      }
    }
  }
}

/* (H (1+tiny) + diag(L) (1+tiny))^-1 g per voxel: the symmetric 3x3 solve by
 * the adjugate for CC = 6, divisions for CC in {0, 1, 3}
 * (precond_grid_common.cpp:620-700). */
void or_precond_grid(const double *h, const double *g, double *out, int CC,
                     int64_t N, int64_t X, int64_t Y, int64_t Z, const int *bound,
                     const double *voxel_size, double absolute, double membrane, double bending,
                     double lame_shear, double lame_div, const double *wmap)
{
  weights_t W; set_weights(&W, voxel_size, absolute, membrane, bending, lame_shear, lame_div);
  geom_t G = {X, Y, Z, {bound[0], bound[1], bound[2]}};
  int64_t V = X * Y * Z;
  double w[3];
  for (int64_t n = 0; n < N; ++n) {
    const double *hn = h ? h + n * CC * V : NULL;
    const double *gn = g + n * 3 * V;
    const double *wn = wmap ? wmap + n * V : NULL;
    double *on = out + n * 3 * V;
    for (int64_t i = 0; i < V; ++i) {
      for (int d = 0; d < 3; ++d) {
        if (wn) {   /* RLS diagonal (no 1 + tiny on the regulariser part) */
          int64_t xx = i / (Y * Z), yy = (i / Z) % Y, zz = i % Z;
          double fw[6], fs = 0, wc = rls_faces(wn, &G, &W, xx, yy, zz, fw);
          for (int k = 0; k < 6; ++k) fs += fw[k];
          w[d] = (W.lam0 * wc - 0.5 * fs) / W.vx[d];
        } else {
          w[d] = W.wd[d][0] * ONE_PLUS_TINY;
        }
      }
      double v0 = gn[i], v1 = gn[V + i], v2 = gn[2 * V + i];
      if (!hn) {
        on[i] = v0 / w[0]; on[V + i] = v1 / w[1]; on[2 * V + i] = v2 / w[2];
      } else if (CC == 1) {
        double h0 = hn[i] * ONE_PLUS_TINY;
        on[i] = v0 / (h0 + w[0]); on[V + i] = v1 / (h0 + w[1]); on[2 * V + i] = v2 / (h0 + w[2]);
      } else if (CC == 3) {
        on[i] = v0 / (hn[i] * ONE_PLUS_TINY + w[0]);
        on[V + i] = v1 / (hn[V + i] * ONE_PLUS_TINY + w[1]);
        on[2 * V + i] = v2 / (hn[2 * V + i] * ONE_PLUS_TINY + w[2]);
      } else {
        double a = hn[i] * ONE_PLUS_TINY + w[0], b = hn[V + i] * ONE_PLUS_TINY + w[1];
        double c = hn[2 * V + i] * ONE_PLUS_TINY + w[2];
        double p = hn[3 * V + i], q = hn[4 * V + i], t = hn[5 * V + i];
        double idt = 1.0 / (a * b * c - a * t * t - b * q * q - c * p * p + 2 * p * q * t);
        on[i]         = idt * (v0 * (b * c - t * t) + v1 * (q * t - p * c) + v2 * (p * t - q * b));
        on[V + i]     = idt * (v0 * (q * t - p * c) + v1 * (a * c - q * q) + v2 * (p * q - a * t));
        on[2 * V + i] = idt * (v0 * (p * t - q * b) + v1 * (p * q - a * t) + v2 * (a * b - p * p));
      }
    }
  }
}

/* Per-batch dot products (pcg_common.cpp:24-36). */
static void batch_dot(const double *a, const double *b, int64_t N, int64_t n3, double *out)
{
  for (int64_t n = 0; n < N; ++n) {
    double s = 0;
    for (int64_t i = 0; i < n3; ++i) s += a[n * n3 + i] * b[n * n3 + i];
    out[n] = s;
  }
}

/* do_pcg (pcg_common.cpp:38-85); x in/out; returns the iterations done. */
int or_pcg_grid(const double *h, const double *g, double *x, int CC,
                int64_t N, int64_t X, int64_t Y, int64_t Z, const int *bound,
                const double *voxel_size, double absolute, double membrane, double bending,
                double lame_shear, double lame_div, int nb_iter, double tol, const double *w)
{
  int64_t V = X * Y * Z, n3 = 3 * V, tot = N * n3;
  double *r = malloc(sizeof(double) * tot), *z = malloc(sizeof(double) * tot);
  double *p = malloc(sizeof(double) * tot);
  double *rz = malloc(sizeof(double) * N), *rz0 = malloc(sizeof(double) * N);
  double *pAp = malloc(sizeof(double) * N), *alpha = malloc(sizeof(double) * N);
#define FWD(in, out) or_regulariser_grid(in, out, h, CC, N, X, Y, Z, bound, voxel_size, absolute, \
                                         membrane, bending, lame_shear, lame_div, w)
#define PRE(in, out) or_precond_grid(h, in, out, CC, N, X, Y, Z, bound, voxel_size, absolute, membrane, \
                                     bending, lame_shear, lame_div, w)
  FWD(x, r);
  for (int64_t i = 0; i < tot; ++i) r[i] = g[i] - r[i];
  PRE(r, z);
  batch_dot(r, z, N, n3, rz);
  memcpy(p, z, sizeof(double) * tot);
  int it = 0;
  for (; it < nb_iter; ++it) {
    FWD(p, z);
    batch_dot(p, z, N, n3, pAp);
    double obj = 0;
    for (int64_t n = 0; n < N; ++n) {
      alpha[n] = rz[n] / (pAp[n] > 1e-12 ? pAp[n] : 1e-12);
      obj += alpha[n] * (alpha[n] * pAp[n] + 2 * rz[n]);
    }
    if (tol != 0 && obj < tol * (double)(N * V)) break;
    for (int64_t n = 0; n < N; ++n)
      for (int64_t i = n * n3; i < (n + 1) * n3; ++i) {
        x[i] += alpha[n] * p[i];
        r[i] -= alpha[n] * z[i];
      }
    PRE(r, z);
    memcpy(rz0, rz, sizeof(double) * N);
    batch_dot(r, z, N, n3, rz);
    for (int64_t n = 0; n < N; ++n) {
      double beta = rz[n] / (rz0[n] > 1e-12 ? rz0[n] : 1e-12);
      for (int64_t i = n * n3; i < (n + 1) * n3; ++i) p[i] = z[i] + beta * p[i];
    }
  }
#undef FWD
#undef PRE
  free(r); free(z); free(p); free(rz); free(rz0); free(pAp); free(alpha);
  return it;
}

void or_fmg_grid(const double *h, const double *g, double *x, int CC,
                 int64_t N, int64_t X, int64_t Y, int64_t Z, const int *bound,
                 const double *voxel_size, double absolute, double membrane, double bending,
                 double lame_shear, double lame_div, int nb_cycles, int nb_iter, int max_levels,
                 int quirk_fold, int quirk_zscale, int use_cg, const double *w)
{
  int64_t sz[MAXL][3];
  double vs[MAXL][3];
  double *H[MAXL], *Gr[MAXL], *Xs[MAXL], *R[MAXL], *Wm[MAXL];
  Wm[0] = (double *)w;
  sz[0][0] = X; sz[0][1] = Y; sz[0][2] = Z;
  for (int d = 0; d < 3; ++d) vs[0][d] = voxel_size[d];
  int64_t V0 = X * Y * Z;
  H[0] = h ? (double *)h : NULL;
  Gr[0] = (double *)malloc(sizeof(double) * N * 3 * V0);
  memcpy(Gr[0], g, sizeof(double) * N * 3 * V0);
  Xs[0] = x;
  R[0] = (double *)malloc(sizeof(double) * N * 3 * V0);
  int L = 1;
  for (; L < max_levels && L < MAXL; ++L) {
    int all1 = 1;
    for (int d = 0; d < 3; ++d) {
      sz[L][d] = 1 + (sz[L - 1][d] - 1) / 2;
      all1 = all1 && sz[L][d] == 1;
    }
    if (all1) break;
    double f[3];
    for (int d = 0; d < 3; ++d) {
      vs[L][d] = vs[L - 1][d] * ((double)sz[L - 1][d] / (double)sz[L][d]);
      f[d] = (double)sz[L - 1][d] / (double)sz[L][d];
    }
    int64_t V = sz[L][0] * sz[L][1] * sz[L][2];
    if (h) {
      H[L] = (double *)malloc(sizeof(double) * N * CC * V);
      or_restrict(H[L - 1], H[L], N, CC, sz[L - 1], sz[L], bound);
      scale_fields(H[L], N, CC, sz[L], f, 1, quirk_zscale);
    } else H[L] = NULL;
    Gr[L] = (double *)malloc(sizeof(double) * N * 3 * V);
    or_restrict(Gr[L - 1], Gr[L], N, 3, sz[L - 1], sz[L], bound);
    scale_fields(Gr[L], N, 3, sz[L], f, 0, quirk_zscale);
    if (w) {   /* weights: restricted, never rescaled (fmg_common.cpp:176-181, 409-414) */
      Wm[L] = (double *)malloc(sizeof(double) * N * V);
      or_restrict(Wm[L - 1], Wm[L], N, 1, sz[L - 1], sz[L], bound);
    } else Wm[L] = NULL;
    R[L] = (double *)malloc(sizeof(double) * N * 3 * V);
    Xs[L] = (double *)malloc(sizeof(double) * N * 3 * V);
    or_restrict(Xs[L - 1], Xs[L], N, 3, sz[L - 1], sz[L], bound);
    scale_fields(Xs[L], N, 3, sz[L], f, 0, quirk_zscale);
  }
  int NL = L;
#define RELAX(l) do { \
    if (use_cg) or_pcg_grid(H[l], Gr[l], Xs[l], CC, N, sz[l][0], sz[l][1], sz[l][2], bound, vs[l], \
                            absolute, membrane, bending, lame_shear, lame_div, nb_iter, 0.0, Wm[l]); \
    else or_relax_grid(Gr[l], H[l], CC, Xs[l], N, sz[l][0], sz[l][1], sz[l][2], bound, vs[l], \
                       absolute, membrane, bending, lame_shear, lame_div, nb_iter, quirk_fold, Wm[l]); \
  } while (0)
#define FAC(l, f) do { for (int d_ = 0; d_ < 3; ++d_) f[d_] = (double)sz[l][d_] / (double)sz[(l) + 1][d_]; } while (0)
  RELAX(NL - 1);
  for (int base = NL - 2; base >= 0; --base) {
    double f[3];
    or_prolong(Xs[base + 1], Xs[base], N, 3, sz[base + 1], sz[base], bound);
    FAC(base, f);
    scale_fields(Xs[base], N, 3, sz[base], f, 0, quirk_zscale);
    for (int cyc = 0; cyc < nb_cycles; ++cyc) {
      for (int l = base; l < NL - 1; ++l) {
        RELAX(l);
        int64_t V = sz[l][0] * sz[l][1] * sz[l][2];
        or_regulariser_grid(Xs[l], R[l], H[l], CC, N, sz[l][0], sz[l][1], sz[l][2], bound, vs[l],
                            absolute, membrane, bending, lame_shear, lame_div, Wm[l]);
        for (int64_t i = 0; i < N * 3 * V; ++i) R[l][i] = Gr[l][i] - R[l][i];
        or_restrict(R[l], Gr[l + 1], N, 3, sz[l], sz[l + 1], bound);
        FAC(l, f);
        scale_fields(Gr[l + 1], N, 3, sz[l + 1], f, 0, quirk_zscale);
        memset(Xs[l + 1], 0, sizeof(double) * N * 3 * sz[l + 1][0] * sz[l + 1][1] * sz[l + 1][2]);
      }
      RELAX(NL - 1);
      for (int l = NL - 2; l >= base; --l) {
        int64_t V = sz[l][0] * sz[l][1] * sz[l][2];
        or_prolong(Xs[l + 1], R[l], N, 3, sz[l + 1], sz[l], bound);
        FAC(l, f);
        scale_fields(R[l], N, 3, sz[l], f, 0, quirk_zscale);
        for (int64_t i = 0; i < N * 3 * V; ++i) Xs[l][i] += R[l][i];
        RELAX(l);
      }
    }
  }
#undef RELAX
#undef FAC
  for (int l = 0; l < NL; ++l) {
    free(Gr[l]); free(R[l]);
    if (l > 0) { free(Xs[l]); if (H[l]) free(H[l]); if (Wm[l]) free(Wm[l]); }
  }
}

/* Plain CG on (H + L) x = g, from x (in/out); returns iterations used. */
int or_cg_grid(const double *h, const double *g, double *x, int CC,
               int64_t N, int64_t X, int64_t Y, int64_t Z, const int *bound,
               const double *voxel_size, double absolute, double membrane, double bending,
               double lame_shear, double lame_div, int max_iter, double tol, const double *w)
{
  int64_t n = N * 3 * X * Y * Z;
  double *r = malloc(sizeof(double) * n), *p = malloc(sizeof(double) * n), *q = malloc(sizeof(double) * n);
  or_regulariser_grid(x, q, h, CC, N, X, Y, Z, bound, voxel_size, absolute, membrane, bending, lame_shear, lame_div, w);
  double rr = 0, gg = 0;
  for (int64_t i = 0; i < n; ++i) { r[i] = g[i] - q[i]; p[i] = r[i]; rr += r[i] * r[i]; gg += g[i] * g[i]; }
  int it = 0;
  for (; it < max_iter && sqrt(rr) > tol * sqrt(gg); ++it) {
    or_regulariser_grid(p, q, h, CC, N, X, Y, Z, bound, voxel_size, absolute, membrane, bending, lame_shear, lame_div, w);
    double pq = 0;
    for (int64_t i = 0; i < n; ++i) pq += p[i] * q[i];
    double a = rr / pq, rr2 = 0;
    for (int64_t i = 0; i < n; ++i) { x[i] += a * p[i]; r[i] -= a * q[i]; rr2 += r[i] * r[i]; }
    double b = rr2 / rr;
    rr = rr2;
    for (int64_t i = 0; i < n; ++i) p[i] = r[i] + b * p[i];
  }
  free(r); free(p); free(q);
  return it;
}
