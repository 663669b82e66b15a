"""ctypes binding of libnitorch_amd.so (the C ABI declared in
include/nitorch_amd.h).

There is no fallback: if the library is missing or cannot be loaded the
import of the ops raises, naming the file and how to build it.
"""
import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libnitorch_amd.so")

NI_FLOAT32 = 0
NI_FLOAT64 = 1

MODE_PULL, MODE_PUSH, MODE_COUNT, MODE_GRAD, MODE_SGRAD = 1, 2, 4, 8, 16
MODE_TRGT, MODE_KTRGT = 32, 64

_lib = None

_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_PI32 = ctypes.POINTER(ctypes.c_int32)
_PI64 = ctypes.POINTER(ctypes.c_int64)

# name -> argtypes (all return int32 status)
_SIGS = {
    "ni_pushpull": [_I32, _I32, _PI32, _PI32, _I32, _I32, _I64, _I64, _PI64, _PI64,
                    _P, _PI64, _P, _PI64, _P, _PI64, _P, _PI64, _P, _PI64, _P],
    "ni_grid_pull": [_I32, _I32, _PI32, _PI32, _I32, _I64, _I64, _PI64, _PI64,
                     _P, _PI64, _P, _PI64, _P, _PI64, _P],
    "ni_grid_grad": [_I32, _I32, _PI32, _PI32, _I32, _I64, _I64, _PI64, _PI64,
                     _P, _PI64, _P, _PI64, _P, _PI64, _P],
    "ni_grid_push": [_I32, _I32, _PI32, _PI32, _I32, _I64, _I64, _PI64, _PI64,
                     _P, _PI64, _P, _PI64, _P, _PI64, _P],
    "ni_grid_count": [_I32, _I32, _PI32, _PI32, _I32, _I64, _PI64, _PI64,
                      _P, _PI64, _P, _PI64, _P],
    "ni_grid_pull_backward": [_I32, _I32, _PI32, _PI32, _I32, _I64, _I64, _PI64, _PI64,
                              _P, _PI64, _P, _PI64, _P, _PI64, _P, _PI64, _P, _PI64, _P],
    "ni_grid_push_backward": [_I32, _I32, _PI32, _PI32, _I32, _I64, _I64, _PI64, _PI64,
                              _P, _PI64, _P, _PI64, _P, _PI64, _P, _PI64, _P, _PI64, _P],
    "ni_grid_count_backward": [_I32, _I32, _PI32, _PI32, _I32, _I64, _PI64, _PI64,
                               _P, _PI64, _P, _PI64, _P, _PI64, _P],
    "ni_grid_grad_backward": [_I32, _I32, _PI32, _PI32, _I32, _I64, _I64, _PI64, _PI64,
                              _P, _PI64, _P, _PI64, _P, _PI64, _P, _PI64, _P, _PI64, _P],
    "ni_spline_coeff": [_I32, _I64, _I32, _PI64, _PI32, _PI32, _P, _P],
}

# optional symbols (solver / learned preconditioner), registered when present
_OPTIONAL_SIGS = {}


def register_signatures(sigs):
    """Called by the solver / network modules to declare their entry points."""
    _OPTIONAL_SIGS.update(sigs)
    if _lib is not None:
        _apply(sigs)


def _apply(sigs):
    for name, args in sigs.items():
        fn = getattr(_lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int32


def lib():
    """Load the library (once). Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"nitorch_amd: HIP library not found at {LIB_PATH}; build it with "
                "`python -m nitorch_amd.build` (hipcc, gfx950)")
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.ni_last_error.restype = ctypes.c_char_p
        _lib.ni_last_error.argtypes = []
        _apply(_SIGS)
        _apply(_OPTIONAL_SIGS)
    return _lib


def last_error():
    return lib().ni_last_error().decode()


def check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: {last_error()}")


def i32arr(vals):
    vals = [int(v) for v in vals]
    return (ctypes.c_int32 * len(vals))(*vals)


def i64arr(vals):
    vals = [int(v) for v in vals]
    return (ctypes.c_int64 * max(len(vals), 1))(*vals)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def strides(t):
    return i64arr(t.stride()) if t is not None else None


def dtype_code(dtype):
    if dtype == torch.float32:
        return NI_FLOAT32
    if dtype == torch.float64:
        return NI_FLOAT64
    raise TypeError(f"nitorch_amd: unsupported dtype {dtype} (float32 and float64 are supported)")


def stream_of(t):
    """hipStream_t of torch's current stream on the tensor's device."""
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def require_device(*tensors):
    for t in tensors:
        if t is not None and not t.is_cuda:
            raise RuntimeError(
                "nitorch_amd: tensors must live on a ROCm GPU device "
                f"(got a tensor on {t.device}); there is no CPU path")
