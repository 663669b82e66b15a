// Spline pull / push / count / grad and their backward passes on gfx950.
//
// Semantics follow nitorch's compiled backend
// (nitorch/_C/src/impl/pushpull_common.cpp, PushPullImpl): one work item per
// *target* sample (the output voxel for pull/sgrad, the input sample for
// push/count), out-of-bound rule x in [-tol, n-1+tol) with tol = 0.05
// (0.55 when `extrapolate & 2`), zero-filled outputs outside, mode flags as
// in nitorch/_C/src/pushpull.cpp.
//
// Design (not a translation of the reference's kernel):
//  * kernels are specialised at compile time on the isotropic order and the
//    mode for the 3-D float path, so that tap tables, signs and offsets live
//    in registers and dead modes vanish;
//  * a runtime-generic kernel covers 1-D/2-D, float64, anisotropic orders and
//    64-bit offsets;
//  * weights come from exact piecewise tables (spline_tables.h): tap k of an
//    order-O sample always lies on the same polynomial piece;
//  * channels are processed in register blocks of NI_CB, the tap tables are
//    built once per sample;
//  * block -> sample-range mapping is XCD-aware: the 8 XCDs of an MI355X
//    each take a contiguous range of samples (so their L2s see contiguous
//    regions of the volume) instead of a round-robin interleave;
//  * push/count scatter with hardware FP32/FP64 atomics (no-return).
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <type_traits>

#include "pushpull_impl.h"
#include "nitorch_amd.h"

namespace ni_amd {

static thread_local char g_err[1024] = {0};

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
const char *get_error() { return g_err; }

// ---------------- dispatch ----------------
template <typename T>
static bool dispatch_fast(const PPArgs<T> &a, int64_t total, hipStream_t st) {
  if constexpr (!std::is_same<T, float>::value) {
    return false;
  } else {
    switch (a.order[0]) {
      case 0: return dispatch_order<0>(a, total, st);
      case 1: return dispatch_order<1>(a, total, st);
      case 2: return dispatch_order<2>(a, total, st);
      case 3: return dispatch_order<3>(a, total, st);
      case 4: return dispatch_order<4>(a, total, st);
      case 5: return dispatch_order<5>(a, total, st);
      case 6: return dispatch_order<6>(a, total, st);
      case 7: return dispatch_order<7>(a, total, st);
      default: return false;
    }
  }
}

static bool fits32(const int64_t *size, const int64_t *stride, int nd) {
  int64_t m = 0;
  for (int i = 0; i < nd; ++i) {
    if (size[i] <= 0) continue;
    int64_t s = stride[i] < 0 ? -stride[i] : stride[i];
    m += (size[i] - 1) * s;
  }
  return m < (int64_t(1) << 31) - 1;
}

template <typename T>
static int run_pushpull(PPArgs<T> a, hipStream_t st) {
  NI_CHECK(a.dim >= 1 && a.dim <= 3, "pushpull: dim must be 1, 2 or 3 (got %d)", a.dim);
  for (int k = 0; k < 3; ++k) {
    NI_CHECK(a.order[k] >= 0 && a.order[k] <= 7, "pushpull: interpolation order must be in 0..7 (got %d)", a.order[k]);
    NI_CHECK(a.bound[k] >= 0 && a.bound[k] <= 8, "pushpull: unknown bound code %d", a.bound[k]);
  }
  const int64_t total = a.N * a.tX * a.tY * a.tZ;
  const bool scatter = a.flags & (F_PUSH | F_COUNT);
  if (scatter) {
    const int64_t Cout = (a.flags & F_PUSH) ? a.C : 1;
    const int64_t nvox = a.sX * a.sY * a.sZ;
    // contiguous [N, Cout, sX, sY, sZ]; strides of singleton dims are free
    const int64_t sizes[5] = {a.N, Cout, a.sX, a.sY, a.sZ};
    bool contig = true;
    int64_t expect = 1;
    for (int k = 4; k >= 0; --k) {
      if (sizes[k] > 1 && a.out_s[k] != expect) contig = false;
      expect *= sizes[k];
    }
    NI_CHECK(contig, "pushpull: push/count output must be contiguous");
    if (int rc = zero_fill<T>(a.out, a.N * Cout * nvox, st)) return rc;
  }
  if (total == 0 || a.C == 0) return 0;

  // grid_grad of nearest-neighbour splines is zero everywhere (the reference
  // zero-fills it, pushpull_common.cpp:513-516): a contiguous output is
  // cleared without sampling anything
  if (a.flags == F_SGRAD && a.order[0] == 0 && (a.dim < 2 || a.order[1] == 0) && (a.dim < 3 || a.order[2] == 0)) {
    const int64_t osz[6] = {a.N, a.C, a.tX, a.tY, a.tZ, a.dim};
    bool contig = true;
    int64_t expect = 1;
    for (int k = 5; k >= 0; --k) {
      if (osz[k] > 1 && a.out_s[k] != expect) contig = false;
      expect *= osz[k];
    }
    if (contig) return zero_fill<T>(a.out, expect, st);
  }

  // 32-bit offsets when every tensor allows it
  const int64_t ssz[5] = {a.N, a.C, a.sX, a.sY, a.sZ};
  const int64_t tsz[6] = {a.N, a.C, a.tX, a.tY, a.tZ, (a.flags & F_KTRGT) ? a.dim : 1};
  const int64_t gsz[5] = {a.N, a.tX, a.tY, a.tZ, a.dim};
  bool ok32 = total < (int64_t(1) << 31) - NI_BLOCK * NI_NXCD && fits32(gsz, a.grid_s, 5);
  if (a.src) ok32 = ok32 && fits32(ssz, a.src_s, 5);
  if (a.trgt) ok32 = ok32 && fits32(tsz, a.trgt_s, 6);
  if (a.out) {
    if (scatter) ok32 = ok32 && fits32(ssz, a.out_s, 5);
    else ok32 = ok32 && fits32(tsz, a.out_s, 6);
  }
  if (a.gout) ok32 = ok32 && fits32(gsz, a.gout_s, 5);

  const bool iso = a.order[0] == a.order[1] && a.order[0] == a.order[2];
  bool done = false;
  if (ok32 && iso && a.dim == 3) done = dispatch_fast<T>(a, total, st);
  if (!done) {
    int64_t nb = (total + NI_BLOCK - 1) / NI_BLOCK;
    if (ok32)
      hipLaunchKernelGGL((pp_kernel<T, int32_t, -1, -1, false>), dim3((unsigned)nb), dim3(NI_BLOCK), 0, st, a, total, nb);
    else
      hipLaunchKernelGGL((pp_kernel<T, int64_t, -1, -1, false>), dim3((unsigned)nb), dim3(NI_BLOCK), 0, st, a, total, nb);
  }
  NI_HIP_CHECK(hipGetLastError());
  return 0;
}

template <typename T>
static int pushpull_entry(int dim, const int32_t *bound, const int32_t *order, int32_t extrapolate,
                          int32_t flags, int64_t N, int64_t C, const int64_t *src_size,
                          const int64_t *trgt_size, const void *src, const int64_t *src_stride,
                          const void *trgt, const int64_t *trgt_stride, const void *grid,
                          const int64_t *grid_stride, void *out, const int64_t *out_stride,
                          void *gout, const int64_t *gout_stride, void *stream,
                          int *scratch = nullptr, int64_t scratch_cap = 0) {
  PPArgs<T> a;
  std::memset(&a, 0, sizeof(a));
  a.scratch = scratch;
  a.scratch_cap = scratch_cap;
  a.src = (const T *)src;
  a.trgt = (const T *)trgt;
  a.grid = (const T *)grid;
  a.out = (T *)out;
  a.gout = (T *)gout;
  a.N = N;
  a.C = C;
  a.sX = src_size[0]; a.sY = dim > 1 ? src_size[1] : 1; a.sZ = dim > 2 ? src_size[2] : 1;
  a.tX = trgt_size[0]; a.tY = dim > 1 ? trgt_size[1] : 1; a.tZ = dim > 2 ? trgt_size[2] : 1;
  // strides: (N, C, spatial[dim], [K]); pad missing spatial dims with 0
  auto unpack5 = [&](const int64_t *s, int64_t *o, bool has_c) {
    if (!s) return;
    int i = 0;
    o[0] = s[i++];
    o[1] = has_c ? s[i++] : 0;
    for (int k = 0; k < 3; ++k) o[2 + k] = k < dim ? s[i++] : 0;
  };
  unpack5(src_stride, a.src_s, true);
  if (trgt_stride) {
    unpack5(trgt_stride, a.trgt_s, true);
    a.trgt_s[5] = (flags & F_KTRGT) ? trgt_stride[2 + dim] : 0;
  }
  if (grid_stride) {
    a.grid_s[0] = grid_stride[0];
    for (int k = 0; k < 3; ++k) a.grid_s[1 + k] = k < dim ? grid_stride[1 + k] : 0;
    a.grid_s[4] = grid_stride[1 + dim];
  }
  if (out_stride) {
    unpack5(out_stride, a.out_s, true);
    a.out_s[5] = (flags & F_SGRAD) ? out_stride[2 + dim] : 0;
  }
  if (gout_stride) {
    a.gout_s[0] = gout_stride[0];
    for (int k = 0; k < 3; ++k) a.gout_s[1 + k] = k < dim ? gout_stride[1 + k] : 0;
    a.gout_s[4] = gout_stride[1 + dim];
  }
  for (int k = 0; k < 3; ++k) {
    a.bound[k] = bound[k];
    a.order[k] = order[k];
  }
  a.dim = dim;
  a.extrap = extrapolate;
  a.flags = flags;
  return run_pushpull<T>(a, (hipStream_t)stream);
}

// Device-resident scratch of ni_grid_push_noalloc: the tile path's list of
// scattered tiles (a count, then tile ids), so that the call allocates
// nothing.  Each GPU holds its own copy for the life of the process; its
// address is looked up once per device.
#define NI_SCRATCH_INTS (1 << 20)   // 4 MiB: up to 2^20 - 1 tiles
static __device__ __align__(256) int scratch[NI_SCRATCH_INTS];
static std::mutex g_scratch_mu;
static int *g_scratch_ptr[64] = {nullptr};

static int scratch_for_device(int **out) {
  int dev = -1;
  NI_HIP_CHECK(hipGetDevice(&dev));
  NI_CHECK(dev >= 0 && dev < 64, "grid_push_noalloc: device index %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    NI_CHECK(rc == hipSuccess && p, "grid_push_noalloc: hipGetSymbolAddress failed (%s)", hipGetErrorString(rc));
    g_scratch_ptr[dev] = (int *)p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

}  // namespace ni_amd

using namespace ni_amd;

// ============================ C ABI ============================

extern "C" {

const char *ni_last_error(void) { return get_error(); }

int ni_pushpull(int32_t dtype, int32_t dim, const int32_t *bound, const int32_t *order,
                int32_t extrapolate, int32_t mode, int64_t N, int64_t C,
                const int64_t *src_size, const int64_t *trgt_size,
                const void *src, const int64_t *src_stride,
                const void *trgt, const int64_t *trgt_stride,
                const void *grid, const int64_t *grid_stride,
                void *out, const int64_t *out_stride,
                void *gout, const int64_t *gout_stride, void *stream) {
  set_error("");
  NI_CHECK(bound && order && src_size && trgt_size && grid && grid_stride,
           "ni_pushpull: missing argument");
  if (dtype == NI_FLOAT32)
    return pushpull_entry<float>(dim, bound, order, extrapolate, mode, N, C, src_size, trgt_size,
                                 src, src_stride, trgt, trgt_stride, grid, grid_stride, out,
                                 out_stride, gout, gout_stride, stream);
  if (dtype == NI_FLOAT64)
    return pushpull_entry<double>(dim, bound, order, extrapolate, mode, N, C, src_size, trgt_size,
                                  src, src_stride, trgt, trgt_stride, grid, grid_stride, out,
                                  out_stride, gout, gout_stride, stream);
  set_error("ni_pushpull: unsupported dtype %d", dtype);
  return 1;
}

int ni_grid_pull(int32_t dtype, int32_t dim, const int32_t *bound, const int32_t *order,
                 int32_t extrapolate, int64_t N, int64_t C, const int64_t *src_size,
                 const int64_t *out_size, const void *input, const int64_t *input_stride,
                 const void *grid, const int64_t *grid_stride, void *out,
                 const int64_t *out_stride, void *stream) {
  return ni_pushpull(dtype, dim, bound, order, extrapolate, F_PULL, N, C, src_size, out_size,
                     input, input_stride, nullptr, nullptr, grid, grid_stride, out, out_stride,
                     nullptr, nullptr, stream);
}

int ni_grid_grad(int32_t dtype, int32_t dim, const int32_t *bound, const int32_t *order,
                 int32_t extrapolate, int64_t N, int64_t C, const int64_t *src_size,
                 const int64_t *out_size, const void *input, const int64_t *input_stride,
                 const void *grid, const int64_t *grid_stride, void *out,
                 const int64_t *out_stride, void *stream) {
  return ni_pushpull(dtype, dim, bound, order, extrapolate, F_SGRAD, N, C, src_size, out_size,
                     input, input_stride, nullptr, nullptr, grid, grid_stride, out, out_stride,
                     nullptr, nullptr, stream);
}

int ni_grid_push(int32_t dtype, int32_t dim, const int32_t *bound, const int32_t *order,
                 int32_t extrapolate, int64_t N, int64_t C, const int64_t *out_size,
                 const int64_t *in_size, const void *input, const int64_t *input_stride,
                 const void *grid, const int64_t *grid_stride, void *out,
                 const int64_t *out_stride, void *stream) {
  return ni_pushpull(dtype, dim, bound, order, extrapolate, F_PUSH | F_TRGT, N, C, out_size,
                     in_size, nullptr, nullptr, input, input_stride, grid, grid_stride, out,
                     out_stride, nullptr, nullptr, stream);
}

int ni_grid_push_noalloc(int32_t dtype, int32_t dim, const int32_t *bound, const int32_t *order,
                         int32_t extrapolate, int64_t N, int64_t C, const int64_t *out_size,
                         const int64_t *in_size, const void *input, const int64_t *input_stride,
                         const void *grid, const int64_t *grid_stride, void *out,
                         const int64_t *out_stride, void *stream) {
  set_error("");
  NI_CHECK(bound && order && out_size && in_size && grid && grid_stride, "grid_push_noalloc: missing argument");
  NI_CHECK(dim >= 1 && dim <= 3, "grid_push_noalloc: dim must be 1, 2 or 3 (got %d)", dim);
  // the most tiles the tile path can list: 4 x 8 x 16 samples per tile
  int64_t tiles = N;
  const int64_t per[3] = {4, 8, 16};
  for (int k = 0; k < 3; ++k) tiles *= k < dim ? (in_size[k] + per[k] - 1) / per[k] : 1;
  NI_CHECK(tiles + 1 <= NI_SCRATCH_INTS,
           "grid_push_noalloc: %lld tiles of 4x8x16 samples exceed the scratch capacity of %d (at most %d "
           "tiles; use ni_grid_push)", (long long)tiles, NI_SCRATCH_INTS, NI_SCRATCH_INTS - 1);
  int *scr = nullptr;
  if (int rc = scratch_for_device(&scr)) return rc;
  if (dtype == NI_FLOAT32)
    return pushpull_entry<float>(dim, bound, order, extrapolate, F_PUSH | F_TRGT, N, C, out_size, in_size, nullptr,
                                 nullptr, input, input_stride, grid, grid_stride, out, out_stride, nullptr, nullptr,
                                 stream, scr, NI_SCRATCH_INTS);
  if (dtype == NI_FLOAT64)
    return pushpull_entry<double>(dim, bound, order, extrapolate, F_PUSH | F_TRGT, N, C, out_size, in_size, nullptr,
                                  nullptr, input, input_stride, grid, grid_stride, out, out_stride, nullptr, nullptr,
                                  stream, scr, NI_SCRATCH_INTS);
  set_error("grid_push_noalloc: unsupported dtype %d", dtype);
  return 1;
}

int ni_grid_count(int32_t dtype, int32_t dim, const int32_t *bound, const int32_t *order,
                  int32_t extrapolate, int64_t N, const int64_t *out_size,
                  const int64_t *in_size, const void *grid, const int64_t *grid_stride,
                  void *out, const int64_t *out_stride, void *stream) {
  return ni_pushpull(dtype, dim, bound, order, extrapolate, F_COUNT, N, 1, out_size, in_size,
                     nullptr, nullptr, nullptr, nullptr, grid, grid_stride, out, out_stride,
                     nullptr, nullptr, stream);
}

int ni_grid_pull_backward(int32_t dtype, int32_t dim, const int32_t *bound, const int32_t *order,
                          int32_t extrapolate, int64_t N, int64_t C, const int64_t *src_size,
                          const int64_t *out_size, const void *grad, const int64_t *grad_stride,
                          const void *input, const int64_t *input_stride, const void *grid,
                          const int64_t *grid_stride, void *grad_input,
                          const int64_t *grad_input_stride, void *grad_grid,
                          const int64_t *grad_grid_stride, void *stream) {
  int32_t mode = F_TRGT | (grad_input ? F_PUSH : 0) | (grad_grid ? F_GRAD : 0);
  if (!(mode & (F_PUSH | F_GRAD))) return 0;
  return ni_pushpull(dtype, dim, bound, order, extrapolate, mode, N, C, src_size, out_size,
                     input, input_stride, grad, grad_stride, grid, grid_stride, grad_input,
                     grad_input_stride, grad_grid, grad_grid_stride, stream);
}

int ni_grid_push_backward(int32_t dtype, int32_t dim, const int32_t *bound, const int32_t *order,
                          int32_t extrapolate, int64_t N, int64_t C, const int64_t *out_size,
                          const int64_t *in_size, const void *grad, const int64_t *grad_stride,
                          const void *input, const int64_t *input_stride, const void *grid,
                          const int64_t *grid_stride, void *grad_input,
                          const int64_t *grad_input_stride, void *grad_grid,
                          const int64_t *grad_grid_stride, void *stream) {
  if (!grad_input && !grad_grid) return 0;
  // the target only weights the grid gradient; a pull of `grad` alone needs none
  int32_t mode = grad_grid ? (F_TRGT | F_GRAD | (grad_input ? F_PULL : 0)) : F_PULL;
  return ni_pushpull(dtype, dim, bound, order, extrapolate, mode, N, C, out_size, in_size, grad,
                     grad_stride, input, input_stride, grid, grid_stride, grad_input,
                     grad_input_stride, grad_grid, grad_grid_stride, stream);
}

int ni_grid_count_backward(int32_t dtype, int32_t dim, const int32_t *bound, const int32_t *order,
                           int32_t extrapolate, int64_t N, const int64_t *out_size,
                           const int64_t *in_size, const void *grad, const int64_t *grad_stride,
                           const void *grid, const int64_t *grid_stride, void *grad_grid,
                           const int64_t *grad_grid_stride, void *stream) {
  return ni_pushpull(dtype, dim, bound, order, extrapolate, F_GRAD, N, 1, out_size, in_size, grad,
                     grad_stride, nullptr, nullptr, grid, grid_stride, nullptr, nullptr,
                     grad_grid, grad_grid_stride, stream);
}

int ni_grid_grad_backward(int32_t dtype, int32_t dim, const int32_t *bound, const int32_t *order,
                          int32_t extrapolate, int64_t N, int64_t C, const int64_t *src_size,
                          const int64_t *out_size, const void *grad, const int64_t *grad_stride,
                          const void *input, const int64_t *input_stride, const void *grid,
                          const int64_t *grid_stride, void *grad_input,
                          const int64_t *grad_input_stride, void *grad_grid,
                          const int64_t *grad_grid_stride, void *stream) {
  int32_t mode = F_TRGT | F_KTRGT | (grad_input ? F_PUSH : 0) | (grad_grid ? F_GRAD : 0);
  if (!(mode & (F_PUSH | F_GRAD))) return 0;
  return ni_pushpull(dtype, dim, bound, order, extrapolate, mode, N, C, src_size, out_size,
                     input, input_stride, grad, grad_stride, grid, grid_stride, grad_input,
                     grad_input_stride, grad_grid, grad_grid_stride, stream);
}

}  // extern "C"
