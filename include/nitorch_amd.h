/*
 * nitorch_amd — C ABI of the MI355X (gfx950) spline push/pull and FMG
 * solver library (libnitorch_amd.so).
 *
 * Conventions shared by every entry point
 *  - all pointers are device pointers (HIP), element strides are int64 and
 *    counted in elements, never bytes;
 *  - `dtype` is NI_FLOAT32 or NI_FLOAT64;
 *  - `bound` and `order` always hold 3 entries, padded by repeating the last
 *    given value, with the reference's enum values: BoundType replicate=0,
 *    dct1=1, dct2=2, dst1=3, dst2=4, dft=5, zero=7 (nitorch/_C/src/bounds.h:7-17)
 *    and InterpolationType 0..7 (nitorch/_C/src/interpolation.h:17-19);
 *  - tensor layouts are those of the reference's C++ layer: image-like
 *    tensors [N, C, *spatial] and grids [N, *spatial, dim] with the last dim
 *    holding (x, y, z) voxel coordinates of the source;
 *  - outputs are allocated by the caller (with the shapes the reference would
 *    allocate) and are fully written by the call; push/count outputs must be
 *    contiguous;
 *  - `stream` is a hipStream_t (NULL = legacy default stream); every call is
 *    asynchronous with respect to the host;
 *  - the return value is 0 on success; on failure it is non-zero and
 *    ni_last_error() returns a thread-local message (argument errors return
 *    1, HIP errors return 2).
 */
#ifndef NITORCH_AMD_H
#define NITORCH_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NI_FLOAT32 0
#define NI_FLOAT64 1

/* Mode flags of ni_pushpull (the reference's do_pull/do_push/do_count/
 * do_grad/do_sgrad booleans, nitorch/_C/src/impl/pushpull_common.cpp:91-97). */
#define NI_MODE_PULL 1
#define NI_MODE_PUSH 2
#define NI_MODE_COUNT 4
#define NI_MODE_GRAD 8
#define NI_MODE_SGRAD 16
#define NI_MODE_TRGT 32   /* a target tensor is given (push values, grad weights) */
#define NI_MODE_KTRGT 64  /* target has a trailing dim of size `dim` (grad backward) */

/* Message of the last failing call on this thread ("" if none). */
const char *ni_last_error(void);

/* General entry: one pass of the reference's PushPullImpl with explicit
 * mode flags (nitorch/_C/src/impl/pushpull_common.cpp:580-794, dispatch
 * :2494-2609).  src: [N, C, *src_size]; trgt: [N, C, *trgt_size(, dim)];
 * grid: [N, *trgt_size, dim]; out: pull/sgrad [N, C, *trgt_size(, dim)],
 * push [N, C, *src_size], count [N, 1, *src_size]; gout: [N, *trgt_size, dim]. */
int ni_pushpull(int32_t dtype, int32_t dim, const int32_t *bound, const int32_t *order,
                int32_t extrapolate, int32_t mode, int64_t N, int64_t C,
                const int64_t *src_size, const int64_t *trgt_size,
                const void *src, const int64_t *src_stride,
                const void *trgt, const int64_t *trgt_stride,
                const void *grid, const int64_t *grid_stride,
                void *out, const int64_t *out_stride,
                void *gout, const int64_t *gout_stride, void *stream);

/* Replaces ni::grid_pull (nitorch/_C/src/pushpull.h:10-14, pushpull.cpp:75-107).
 * input [N, C, *src_size], grid [N, *out_size, dim] -> out [N, C, *out_size]. */
int ni_grid_pull(int32_t dtype, int32_t dim, const int32_t *bound, const int32_t *order,
                 int32_t extrapolate, int64_t N, int64_t C, const int64_t *src_size,
                 const int64_t *out_size, const void *input, const int64_t *input_stride,
                 const void *grid, const int64_t *grid_stride, void *out,
                 const int64_t *out_stride, void *stream);

/* Replaces ni::grid_grad (pushpull.h:46-50, pushpull.cpp:268-300).
 * -> out [N, C, *out_size, dim] (spatial gradients of the spline at the grid). */
int ni_grid_grad(int32_t dtype, int32_t dim, const int32_t *bound, const int32_t *order,
                 int32_t extrapolate, int64_t N, int64_t C, const int64_t *src_size,
                 const int64_t *out_size, const void *input, const int64_t *input_stride,
                 const void *grid, const int64_t *grid_stride, void *out,
                 const int64_t *out_stride, void *stream);

/* Replaces ni::grid_push (pushpull.h:22-27, pushpull.cpp:131-184).
 * input [N, C, *in_size], grid [N, *in_size, dim] -> out [N, C, *out_size]
 * (contiguous; zero-filled by the call, then scatter-added). */
int ni_grid_push(int32_t dtype, int32_t dim, const int32_t *bound, const int32_t *order,
                 int32_t extrapolate, int64_t N, int64_t C, const int64_t *out_size,
                 const int64_t *in_size, const void *input, const int64_t *input_stride,
                 const void *grid, const int64_t *grid_stride, void *out,
                 const int64_t *out_stride, void *stream);

/* No counterpart in the reference.  ni_grid_push for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): same arguments and result, but
 * the list of scattered tiles that the tile path hands to its second
 * kernel lives in a device-resident array of the library (2^20 ints per
 * GPU, held until the process ends), so the call allocates nothing.
 * Limit: N * ceil(X/4) * ceil(Y/8) * ceil(Z/16) <= 2^20 - 1 over in_size
 * (X, Y, Z = 1 for missing dims); beyond it the call returns an error.
 * Every call on a GPU shares that GPU's array: calls that overlap on one
 * GPU (concurrent streams, threads) race. */
int ni_grid_push_noalloc(int32_t dtype, int32_t dim, const int32_t *bound, const int32_t *order,
                         int32_t extrapolate, int64_t N, int64_t C, const int64_t *out_size,
                         const int64_t *in_size, const void *input, const int64_t *input_stride,
                         const void *grid, const int64_t *grid_stride, void *out,
                         const int64_t *out_stride, void *stream);

/* Replaces ni::grid_count (pushpull.h:34-38, pushpull.cpp:204-247).
 * grid [N, *in_size, dim] -> out [N, 1, *out_size] (contiguous). */
int ni_grid_count(int32_t dtype, int32_t dim, const int32_t *bound, const int32_t *order,
                  int32_t extrapolate, int64_t N, const int64_t *out_size,
                  const int64_t *in_size, const void *grid, const int64_t *grid_stride,
                  void *out, const int64_t *out_stride, void *stream);

/* Replaces ni::grid_pull_backward (pushpull.h:16-20, pushpull.cpp:109-128):
 * fused push of `grad` (-> grad_input [N, C, *src_size], contiguous) and grid
 * gradient (-> grad_grid [N, *out_size, dim]).  Pass NULL for an output that
 * is not needed (the reference's requires_grad tests). */
int ni_grid_pull_backward(int32_t dtype, int32_t dim, const int32_t *bound, const int32_t *order,
                          int32_t extrapolate, int64_t N, int64_t C, const int64_t *src_size,
                          const int64_t *out_size, const void *grad, const int64_t *grad_stride,
                          const void *input, const int64_t *input_stride, const void *grid,
                          const int64_t *grid_stride, void *grad_input,
                          const int64_t *grad_input_stride, void *grad_grid,
                          const int64_t *grad_grid_stride, void *stream);

/* Replaces ni::grid_push_backward (pushpull.h:28-32, pushpull.cpp:186-202):
 * grad [N, C, *out_size] (push-output space), input [N, C, *in_size]
 * -> grad_input = pull(grad) [N, C, *in_size], grad_grid [N, *in_size, dim]. */
int ni_grid_push_backward(int32_t dtype, int32_t dim, const int32_t *bound, const int32_t *order,
                          int32_t extrapolate, int64_t N, int64_t C, const int64_t *out_size,
                          const int64_t *in_size, const void *grad, const int64_t *grad_stride,
                          const void *input, const int64_t *input_stride, const void *grid,
                          const int64_t *grid_stride, void *grad_input,
                          const int64_t *grad_input_stride, void *grad_grid,
                          const int64_t *grad_grid_stride, void *stream);

/* Replaces ni::grid_count_backward (pushpull.h:40-44, pushpull.cpp:249-265):
 * grad [N, 1, *out_size] -> grad_grid [N, *in_size, dim]. */
int ni_grid_count_backward(int32_t dtype, int32_t dim, const int32_t *bound, const int32_t *order,
                           int32_t extrapolate, int64_t N, const int64_t *out_size,
                           const int64_t *in_size, const void *grad, const int64_t *grad_stride,
                           const void *grid, const int64_t *grid_stride, void *grad_grid,
                           const int64_t *grad_grid_stride, void *stream);

/* Replaces ni::grid_grad_backward (pushpull.h:52-56, pushpull.cpp:303-327):
 * grad [N, C, *out_size, dim] -> grad_input [N, C, *src_size] (contiguous),
 * grad_grid [N, *out_size, dim]. */
int ni_grid_grad_backward(int32_t dtype, int32_t dim, const int32_t *bound, const int32_t *order,
                          int32_t extrapolate, int64_t N, int64_t C, const int64_t *src_size,
                          const int64_t *out_size, const void *grad, const int64_t *grad_stride,
                          const void *input, const int64_t *input_stride, const void *grid,
                          const int64_t *grid_stride, void *grad_input,
                          const int64_t *grad_input_stride, void *grad_grid,
                          const int64_t *grad_grid_stride, void *stream);

/* Interpolating B-spline coefficients, in place.  Replaces interpol.coeff's
 * spline_coeff / spline_coeff_nd, which nitorch/_C/grid.py:4 imports and
 * SplineCoeff / SplineCoeffND (nitorch/_C/grid.py:335-390) call (restated
 * in the reference's nitorch/_C/_ts/coeff.py:251-344).  `data` is a
 * contiguous [nbatch, *shape] tensor with `ndim` (1..8) filtered axes;
 * `bound[d]` / `order[d]` per axis (one entry each, no padding).  Orders 0
 * and 1 and length-1 axes are left unchanged; other axes need bound zero,
 * dct1 (mirror rules), replicate, dct2 (half-sample rules) or dft. */
int ni_spline_coeff(int32_t dtype, int64_t nbatch, int32_t ndim, const int64_t *shape, const int32_t *bound,
                    const int32_t *order, void *data, void *stream);

/* Replaces ni::resize (nitorch/_C/src/resize.cpp:25-59; bound as `resize`,
 * `prolongation`, `restriction` and their backwards, spatial.cpp:99-108).
 * input [N, C, *in_shape], output [N, C, *out_shape], `dim` (1..3) spatial
 * dims, both contiguous and caller-allocated.  `factor`, `bound`, `order`
 * and `mode` (GridAlignType edge 0, center 1, first 2, last 3) hold 3
 * entries.  source = do_adjoint ? output : input, target = the other; each
 * target voxel w samples the source at scale * w + shift (prepare_affine,
 * resize_common.cpp:1548-1567) without out-of-bound masking; the adjoint
 * splats into the zeroed source.  With `accumulate` the generic kernels add
 * into the given output (the isotropic nearest / linear, and 3-D quadratic,
 * ones overwrite it, as the reference's).  `normalize` scales the result by
 * the scales with the reference's fall-through product. */
int ni_resize(int32_t dtype, int32_t dim, int64_t N, int64_t C, const int64_t *in_shape, const int64_t *out_shape,
              const double *factor, const int32_t *bound, const int32_t *order, const int32_t *mode,
              int32_t do_adjoint, int32_t normalize, int32_t accumulate, const void *input, void *output,
              void *stream);

/* ------------------------------------------------------------------------
 * Displacement-field solver (3-D).  Fields are [N, 3, *shape] and Hessians
 * [N, CC, *shape], contiguous, with CC in {1, 3, 6} (6 = h00 h11 h22 h01 h02
 * h12) or hessian == NULL.  `penalty` = {absolute, membrane, bending,
 * lame_shear, lame_div}; `voxel_size` has 3 entries.  `weight` (NULL for
 * none): voxel-wise RLS weights [N, 1, *shape], contiguous, of the absolute
 * and membrane energies (the reference's `weight` argument; bending and Lame
 * penalties with weights are an error, as there).
 * ---------------------------------------------------------------------- */

/* Replaces ni::regulariser_grid (nitorch/_C/src/regulariser_grid.h:8-11):
 * output = (H + L) input. */
int ni_regulariser_grid(int32_t dtype, int64_t N, const int64_t *shape, const void *input, void *output,
                        const void *hessian, int32_t CC, const void *weight, const int32_t *bound,
                        const double *voxel_size, const double *penalty, void *stream);

/* Residual g - (H + L) x (the reference's residuals_grid,
 * nitorch/_C/src/impl/fmg_common.cpp:334-352).  If `norms` (device, 2 doubles)
 * is given it receives ||g - (H+L)x||^2 and ||g||^2. */
int ni_residual_grid(int32_t dtype, int64_t N, const int64_t *shape, const void *solution, const void *gradient,
                     void *residual, const void *hessian, int32_t CC, const void *weight, const int32_t *bound,
                     const double *voxel_size, const double *penalty, double *norms, void *stream);

/* Replaces ni::relax_grid (nitorch/_C/src/relax_grid.h:8-13): nb_iter
 * multi-colour Gauss-Seidel sweeps on (H + L) x = g, x updated in place. */
int ni_relax_grid(int32_t dtype, int64_t N, const int64_t *shape, const void *hessian, int32_t CC,
                  const void *weight, const void *gradient, void *solution, const int32_t *bound,
                  const double *voxel_size, const double *penalty, int32_t nb_iter, void *stream);

/* The colour classes ni_relax_grid and the relax smoother of ni_fmg_grid /
 * ni_fmg_solve use on a lattice (3-D shape, bounds, penalties as there):
 * plan[0] = 1 for a linear colouring (a x + b y + c z) mod m, 0 for per-axis
 * period classes; plan[1] = the number of classes (m); plan[2..4] = a, b, c
 * (0 for per-axis classes).  Host only, no device work.  No counterpart in the
 * reference, whose colour classes are fixed per-axis periods
 * (relax_grid_common.cpp:454-476); for callers that need the relaxation
 * order, such as tests that restate it. */
int ni_relax_colouring(const int64_t *shape, const int32_t *bound, const double *penalty, int32_t *plan);

/* Replace ni::fmg_restriction / ni::fmg_prolongation (nitorch/_C/src/multires.h):
 * [N, C, *in_shape] -> [N, C, *out_shape] (linear-adjoint restriction,
 * quadratic prolongation). */
int ni_fmg_restriction(int32_t dtype, int64_t N, int64_t C, const int64_t *in_shape, const int64_t *out_shape,
                       const void *input, void *output, const int32_t *bound, void *stream);
int ni_fmg_prolongation(int32_t dtype, int64_t N, int64_t C, const int64_t *in_shape, const int64_t *out_shape,
                        const void *input, void *output, const int32_t *bound, void *stream);

/* Bytes of device workspace ni_fmg_grid needs for this problem (use_cg:
 * with the PCG smoother's scratch; has_weight: with the RLS weight pyramid). */
int64_t ni_fmg_grid_workspace(int32_t dtype, int64_t N, const int64_t *shape, int32_t CC, int32_t max_levels,
                              int32_t use_cg, int32_t has_weight);

/* Replaces ni::fmg_grid (nitorch/_C/src/fmg.h:22-37): full multigrid with
 * nb_cycles V-cycles per level and nb_iter smoothing steps (use_cg = 0:
 * Gauss-Seidel sweeps, solver='relax'; use_cg = 1: PCG iterations,
 * solver='cg'), then `extra_vcycles` further V-cycles on the finest level.
 * `solution` is the initial guess (zeros for none) and receives the result. */
int ni_fmg_grid(int32_t dtype, int64_t N, const int64_t *shape, const void *hessian, int32_t CC,
                const void *weight, const void *gradient, void *solution, const int32_t *bound,
                const double *voxel_size, const double *penalty, int32_t nb_cycles, int32_t nb_iter,
                int32_t max_levels, int32_t use_cg, int32_t extra_vcycles, void *workspace, int64_t workspace_bytes,
                void *stream);

/* Time-to-tolerance solve (no reference counterpart: the benchmark's "GN step
 * to 1e-4 relative residual"): ni_fmg_grid's full multigrid (skipped when
 * nb_cycles <= 0: `solution` is then the starting iterate), then V-cycles on
 * the finest level until ||g - (H+L)x|| / ||g|| <= tol or max_vcycles.  The
 * workspace is ni_fmg_grid_workspace's; the call synchronises `stream` once
 * per V-cycle to read the residual.  Reports the final relative residual
 * and the number of extra V-cycles. */
int ni_fmg_solve(int32_t dtype, int64_t N, const int64_t *shape, const void *hessian, int32_t CC,
                 const void *weight, const void *gradient, void *solution, const int32_t *bound,
                 const double *voxel_size, const double *penalty, int32_t nb_cycles, int32_t nb_iter,
                 int32_t max_levels, int32_t use_cg, double tol, int32_t max_vcycles, void *workspace,
                 int64_t workspace_bytes, double *out_rel, int32_t *out_vcycles, void *stream);

/* Bytes of device workspace ni_pcg_grid needs (r, z, p and the partial sums
 * of its scalars, 4 x 256 doubles per batch element). */
int64_t ni_pcg_grid_workspace(int32_t dtype, int64_t N, const int64_t *shape);

/* Replaces ni::pcg_grid (nitorch/_C/src/pcg.h:20-32, pcg_common.cpp:38-85,
 * :125-163): nb_iter conjugate-gradient iterations on (H + L) x = g with the
 * block-Jacobi preconditioner of ni_precond_grid; x updated in place.  tol = 0
 * runs all iterations without host synchronisation; tol > 0 reads two
 * scalars back per iteration and stops when
 * sum_n alpha_n (alpha_n p'Ap + 2 r'z) < tol * N * prod(shape).
 * out_iters (optional, host) receives the iterations done. */
int ni_pcg_grid(int32_t dtype, int64_t N, const int64_t *shape, const void *hessian, int32_t CC,
                const void *weight, const void *gradient, void *solution, const int32_t *bound,
                const double *voxel_size, const double *penalty, int32_t nb_iter, double tol, void *workspace,
                int64_t workspace_bytes, int32_t *out_iters, void *stream);

/* Replaces ni::precond_grid (nitorch/_C/src/precond_grid.h:8-12,
 * precond_grid_common.cpp:266-298, :620-660):
 * output = (H (1 + 1e-6) + diag(L) (1 + 1e-6))^-1 gradient, voxel by voxel
 * (with RLS weights, the weighted operator's diagonal, which reads the
 * neighbours' weights through `bound`). */
int ni_precond_grid(int32_t dtype, int64_t N, const int64_t *shape, const void *hessian, int32_t CC,
                    const void *weight, const void *gradient, void *output, const int32_t *bound,
                    const double *voxel_size, const double *penalty, void *stream);

/* ------------------------------------------------------------------------
 * Learned preconditioner (3-D U-Net on BF16 matrix-core kernels).  No
 * counterpart in the reference: used by the learned-PCG Gauss-Newton solve.
 * Activations are channels-last [N, *shape, C] bf16 with C a multiple of 16;
 * weights of a CI -> CO layer are [CO][KP] bf16, K = tap * CI + ci with
 * tap = 9 dx + 3 dy + dz (offsets -1..1 <-> 0..2), KP = 27 CI rounded up to
 * 32 (zero-padded).  Zero padding of 1, stride 1, FP32 accumulation.
 * ---------------------------------------------------------------------- */

/* y = act(conv3x3x3(x, w) + bias): act 0 none / 1 leaky ReLU (slope 0.1);
 * out_f32 = 1 writes fp32 (linear head; out_channels = 3 with CO = 16 keeps
 * the first 3 channels), else bf16 with out_channels = CO.  bias may be NULL.
 * Supported CI -> CO: 16->16, 16->32, 32->32, 32->64, 64->64, 96->32, 48->16
 * and (backward data) 32->16, 64->32, 32->96, 16->48. */
int ni_unet_conv3_fwd(int64_t N, const int64_t *shape, int32_t CI, int32_t CO, const void *x, const void *w,
                      const float *bias, void *y, int32_t act, int32_t out_f32, int32_t out_channels, void *stream);

/* ni_unet_conv3_fwd on concat(nearest 2x upsampling of a [N, *shape/2, C1],
 * skip [N, *shape, C2]) without materialising the concatenation (the result
 * ni_unet_up2_concat + ni_unet_conv3_fwd give, bit for bit); bf16 output,
 * act 0 / 1 as there, bias may be NULL, shape even.  Supported: 32 + 16 ->
 * 16 (the decoder's first full-resolution layer). */
int ni_unet_conv3_fwd_upcat(int64_t N, const int64_t *shape, int32_t C1, int32_t C2, int32_t CO, const void *a,
                            const void *skip, const void *w, const float *bias, void *y, int32_t act, void *stream);

/* dw[CO][KP] += sum over positions of dy (x) shifted x (fp32 accumulators,
 * caller zeroes them); db[CO] += sum of dy if db is not NULL. */
int ni_unet_conv3_bwd_weight(int64_t N, const int64_t *shape, int32_t CI, int32_t CO, const void *x, const void *dy,
                             float *dw, float *db, void *stream);

/* out = dy * (y > 0 ? 1 : 0.1), n bf16 elements. */
int ni_unet_lrelu_bwd(int64_t n, const void *dy, const void *y, void *out, void *stream);

/* 2x2x2 average pooling of [N, *shape, C] (backward: in = gradient of the
 * pooled tensor, out = gradient at full resolution). */
int ni_unet_avgpool2(int32_t backward, int64_t N, const int64_t *shape, int32_t C, const void *in, void *out,
                     void *stream);

/* out [N, *shape, C1 + C2] = concat(nearest 2x upsampling of a
 * [N, *shape/2, C1], skip [N, *shape, C2]); backward: out is the incoming
 * gradient, a and skip receive theirs. */
int ni_unet_up2_concat(int32_t backward, int64_t N, const int64_t *shape, int32_t C1, int32_t C2, const void *a,
                       const void *skip, void *out, void *stream);

/* out [N, V, 16] bf16 = (sg * g, sh * h, 0...) from channel-first fp32
 * g [N, 3, V] and h [N, CC, V] (h may be NULL). */
int ni_unet_pack_input(int64_t N, int64_t V, const float *g, const float *h, int32_t CC, void *out, float sg,
                       float sh, void *stream);

/* x [N, 3, V] (channel-first fp32) = scale * y [N, V, 3] (+ x if accumulate). */
int ni_unet_unpack_output(int64_t N, int64_t V, const float *y, float *x, float scale, int32_t accumulate,
                          void *stream);

/* ---- SVF exponentiation backward: fields of 3 x 3 matrices ----
 * Matrix fields are addressed by element strides {per voxel, per row, per
 * column}, vector fields by {per voxel, per component}, over `nvox`
 * voxels; outputs are dense.  These replace the torch products of the
 * reference's exp_backward / _composition_jac (nitorch/spatial/_svf.py:
 * 271-284, linalg.matvec and jac.det() per squaring, and :480,
 * torch.matmul(jac_left, jac)). */

/* out [nvox, 3, 3] = a @ b */
int ni_svf_mat3_mul(int32_t dtype, int64_t nvox, const void *a, const int64_t *a_stride, const void *b,
                    const int64_t *b_stride, void *out, void *stream);

/* out [nvox, 3] = g0 + det(jac) * jac^T @ p  (one squaring's push-forward) */
int ni_svf_grad_step(int32_t dtype, int64_t nvox, const void *jac, const int64_t *jac_stride, const void *p,
                     const int64_t *p_stride, const void *g0, const int64_t *g0_stride, void *out, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* NITORCH_AMD_H */
