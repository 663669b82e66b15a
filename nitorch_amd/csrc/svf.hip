// Fields of 3 x 3 matrices in the scaling-and-squaring backward of a
// stationary velocity field (nitorch/spatial/_svf.py:271-284 and :480):
// the product of two Jacobian fields, and the push-forward of a gradient
// through one squaring, g0 + det(J) J^T p, fused into one pass.
//
// Every field is addressed by element strides: a vector field by (per
// voxel, per component), a matrix field by (per voxel, per row, per
// column), so transposed and channel-first views are read in place.  The
// voxel index runs over the N x D x H x W voxels of a view whose leading
// dims collapse to one stride (the caller checks).  Outputs are dense,
// [V, 3, 3] or [V, 3].  Each lane handles one voxel; the bound is HBM.
#include <cstdint>

#include "common.h"
#include "nitorch_amd.h"

namespace ni_amd {

#define SV3_BLOCK 256

struct Mat3 {
  int64_t v, r, c;
};

template <typename T>
__device__ __forceinline__ void load3x3(const T *p, const Mat3 &s, int64_t i, T m[3][3]) {
  const T *q = p + i * s.v;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) m[r][c] = q[r * s.r + c * s.c];
}

// O[i] = A[i] B[i]
template <typename T>
__global__ __launch_bounds__(SV3_BLOCK) void mat3_mul(int64_t V, const T *__restrict__ A, Mat3 sa,
                                                      const T *__restrict__ B, Mat3 sb, T *__restrict__ O) {
  const int64_t i = (int64_t)blockIdx.x * SV3_BLOCK + threadIdx.x;
  if (i >= V) return;
  T a[3][3], b[3][3];
  load3x3(A, sa, i, a);
  load3x3(B, sb, i, b);
  T *o = O + i * 9;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      // summed in k order, as the rank-one accumulation it replaces
      T s = a[r][0] * b[0][c];
      s = s + a[r][1] * b[1][c];
      s = s + a[r][2] * b[2][c];
      o[r * 3 + c] = s;
    }
}

// O[i] = G0[i] + det(J[i]) * J[i]^T P[i]
template <typename T>
__global__ __launch_bounds__(SV3_BLOCK) void grad_step(int64_t V, const T *__restrict__ J, Mat3 sj,
                                                       const T *__restrict__ P, int64_t pv, int64_t pc,
                                                       const T *__restrict__ G0, int64_t gv, int64_t gc,
                                                       T *__restrict__ O) {
  const int64_t i = (int64_t)blockIdx.x * SV3_BLOCK + threadIdx.x;
  if (i >= V) return;
  T m[3][3];
  load3x3(J, sj, i, m);
  const T det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
                m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
  const T p0 = P[i * pv], p1 = P[i * pv + pc], p2 = P[i * pv + 2 * pc];
  T *o = O + i * 3;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    T s = m[0][c] * p0;
    s = s + m[1][c] * p1;
    s = s + m[2][c] * p2;
    o[c] = G0[i * gv + c * gc] + s * det;
  }
}

static inline Mat3 mat3(const int64_t *s) { return Mat3{s[0], s[1], s[2]}; }

}  // namespace ni_amd

using namespace ni_amd;

extern "C" {

int ni_svf_mat3_mul(int32_t dtype, int64_t nvox, const void *a, const int64_t *a_stride, const void *b,
                    const int64_t *b_stride, void *out, void *stream) {
  set_error("");
  NI_CHECK(dtype == NI_FLOAT32 || dtype == NI_FLOAT64, "svf_mat3_mul: unsupported dtype %d", dtype);
  NI_CHECK(nvox >= 0, "svf_mat3_mul: negative voxel count");
  if (nvox == 0) return 0;
  NI_CHECK(a && b && out && a_stride && b_stride, "svf_mat3_mul: missing argument");
  hipStream_t st = (hipStream_t)stream;
  const unsigned nb = (unsigned)((nvox + SV3_BLOCK - 1) / SV3_BLOCK);
  NI_CHECK((nvox + SV3_BLOCK - 1) / SV3_BLOCK < (int64_t(1) << 31), "svf_mat3_mul: too many voxels");
  if (dtype == NI_FLOAT32)
    hipLaunchKernelGGL(mat3_mul<float>, dim3(nb), dim3(SV3_BLOCK), 0, st, nvox, (const float *)a,
                       mat3(a_stride), (const float *)b, mat3(b_stride), (float *)out);
  else
    hipLaunchKernelGGL(mat3_mul<double>, dim3(nb), dim3(SV3_BLOCK), 0, st, nvox, (const double *)a,
                       mat3(a_stride), (const double *)b, mat3(b_stride), (double *)out);
  NI_HIP_CHECK(hipGetLastError());
  return 0;
}

int ni_svf_grad_step(int32_t dtype, int64_t nvox, const void *jac, const int64_t *jac_stride, const void *p,
                     const int64_t *p_stride, const void *g0, const int64_t *g0_stride, void *out, void *stream) {
  set_error("");
  NI_CHECK(dtype == NI_FLOAT32 || dtype == NI_FLOAT64, "svf_grad_step: unsupported dtype %d", dtype);
  NI_CHECK(nvox >= 0, "svf_grad_step: negative voxel count");
  if (nvox == 0) return 0;
  NI_CHECK(jac && p && g0 && out && jac_stride && p_stride && g0_stride, "svf_grad_step: missing argument");
  NI_CHECK((nvox + SV3_BLOCK - 1) / SV3_BLOCK < (int64_t(1) << 31), "svf_grad_step: too many voxels");
  hipStream_t st = (hipStream_t)stream;
  const unsigned nb = (unsigned)((nvox + SV3_BLOCK - 1) / SV3_BLOCK);
  if (dtype == NI_FLOAT32)
    hipLaunchKernelGGL(grad_step<float>, dim3(nb), dim3(SV3_BLOCK), 0, st, nvox, (const float *)jac,
                       mat3(jac_stride), (const float *)p, p_stride[0], p_stride[1], (const float *)g0,
                       g0_stride[0], g0_stride[1], (float *)out);
  else
    hipLaunchKernelGGL(grad_step<double>, dim3(nb), dim3(SV3_BLOCK), 0, st, nvox, (const double *)jac,
                       mat3(jac_stride), (const double *)p, p_stride[0], p_stride[1], (const double *)g0,
                       g0_stride[0], g0_stride[1], (double *)out);
  NI_HIP_CHECK(hipGetLastError());
  return 0;
}

}  // extern "C"
