"""ctypes binding of liblaker_hip.so (the C ABI declared in include/laker_hip.h).

The library is mandatory: importing this module raises ``ImportError`` when the
shared object is missing, and every entry point raises when no ROCm GPU is
visible.  There is no CPU fallback on the product path.
"""

from __future__ import annotations

import ctypes
import os

import torch

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "liblaker_hip.so")

LAKER_OK = 0
LAKER_EINVAL = -1
LAKER_EHIP = -2
LAKER_ENOMEM = -3
LAKER_EBREAKDOWN = -4
LAKER_ESTATE = -5
LAKER_ENUMERIC = -6

LAKER_F32 = 0
LAKER_F64 = 1

PREC_FP32 = 0
PREC_MIXED = 1
PREC_FP64 = 2
PREC_BF16 = 3
PREC_HILO = 4

CLAMP_AUTO = 0
CLAMP_ALWAYS = 1
CLAMP_NEVER = 2

c_int32 = ctypes.c_int32
c_int64 = ctypes.c_int64
c_double = ctypes.c_double
c_void_p = ctypes.c_void_p
c_uint64 = ctypes.c_uint64


class PcgResult(ctypes.Structure):
    _fields_ = [
        ("iterations", c_int32),
        ("status", c_int32),
        ("converged", c_int32),
        ("zero_rhs", c_int32),
        ("residual_norm", c_double),
        ("rhs_norm", c_double),
    ]


class MixedResult(ctypes.Structure):
    _fields_ = [
        ("outer_iterations", c_int32),
        ("inner_iterations", c_int32),
        ("status", c_int32),
        ("converged", c_int32),
        ("residual_norm", c_double),
        ("rhs_norm", c_double),
    ]


def _load():
    if not os.path.exists(_LIB_PATH):
        raise ImportError(
            f"liblaker_hip.so not found at {_LIB_PATH}; build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'` (hipcc, gfx950)"
        )
    # torch first: its HIP runtime (same SONAME) is the one the library binds to
    lib = ctypes.CDLL(_LIB_PATH, mode=ctypes.RTLD_GLOBAL)
    sig = {
        "laker_last_error": (ctypes.c_char_p, []),
        "laker_version": (ctypes.c_char_p, []),
        "laker_operator_create": (c_int32, [c_void_p, c_int64, c_int32, c_int32, c_double, c_double,
                                            c_int32, c_void_p, ctypes.POINTER(c_void_p)]),
        "laker_operator_destroy": (c_int32, [c_void_p]),
        "laker_operator_info": (c_int32, [c_void_p, ctypes.POINTER(c_int32), ctypes.POINTER(c_double)]),
        "laker_operator_set_lambda": (c_int32, [c_void_p, c_double]),
        "laker_matvec": (c_int32, [c_void_p, c_void_p, c_int32, c_int64, c_void_p, c_int64, c_int32,
                                   c_void_p]),
        "laker_matvec_noalloc": (c_int32, [c_void_p, c_void_p, c_int32, c_int64, c_void_p, c_int64, c_int32,
                                           c_void_p]),
        "laker_residual_fp64": (c_int32, [c_void_p, c_void_p, c_void_p, c_int32, c_void_p, c_void_p]),
        "laker_cross_matvec": (c_int32, [c_void_p, c_void_p, c_int64, c_void_p, c_int32, c_int64,
                                         c_void_p, c_int64, c_int32, c_int32, c_void_p]),
        "laker_cross_matvec_t": (c_int32, [c_void_p, c_void_p, c_int64, c_void_p, c_int32, c_int64,
                                           c_void_p, c_int64, c_int32, c_void_p]),
        "laker_kernel_eval": (c_int32, [c_void_p, c_int64, c_void_p, c_int64, c_int32, c_int32,
                                        c_double, c_int32, c_void_p, c_int64, c_void_p]),
        "laker_diagonal": (c_int32, [c_void_p, c_void_p, c_void_p]),
        "laker_to_dense": (c_int32, [c_void_p, c_void_p, c_void_p]),
        "laker_matvec_backward": (c_int32, [c_void_p, c_void_p, c_void_p, c_int32, c_void_p, c_void_p,
                                            c_void_p]),
        "laker_cross_matvec_backward": (c_int32, [c_void_p, c_void_p, c_int64, c_void_p, c_int32,
                                                  c_void_p, c_void_p, c_void_p, c_void_p, c_int32,
                                                  c_void_p]),
        "laker_precond_build": (c_int32, [c_void_p, c_int32, c_double, c_double, c_double, c_int32,
                                          c_double, c_void_p, c_uint64, c_int32, c_void_p,
                                          ctypes.POINTER(c_void_p), ctypes.POINTER(c_int32)]),
        "laker_precond_build_from_responses": (c_int32, [c_void_p, c_int64, c_int32, c_int32, c_double,
                                                         c_double, c_double, c_int32, c_double, c_void_p,
                                                         ctypes.POINTER(c_void_p),
                                                         ctypes.POINTER(c_int32)]),
        "laker_precond_copy_out": (c_int32, [c_void_p, c_int32, c_void_p, c_void_p]),
        "laker_precond_destroy": (c_int32, [c_void_p]),
        "laker_precond_info": (c_int32, [c_void_p, ctypes.POINTER(c_double), ctypes.POINTER(c_int32),
                                         ctypes.POINTER(c_double), ctypes.POINTER(c_void_p),
                                         ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p),
                                         ctypes.POINTER(c_void_p)]),
        "laker_precond_apply": (c_int32, [c_void_p, c_void_p, c_int32, c_int64, c_void_p, c_int64,
                                          c_void_p]),
        "laker_pcg_solve": (c_int32, [c_void_p, c_void_p, c_void_p, c_int32, c_void_p, c_void_p,
                                      c_double, c_int32, c_double, c_int32, c_int32, c_int32, c_void_p,
                                      ctypes.POINTER(PcgResult)]),
        "laker_solve_mixed": (c_int32, [c_void_p, c_void_p, c_void_p, c_int32, c_void_p, c_double,
                                        c_int32, c_double, c_int32, c_int32, c_void_p,
                                        ctypes.POINTER(MixedResult)]),
        "laker_coldot": (c_int32, [c_void_p, c_void_p, c_int64, c_int32, c_int64, c_int32, c_void_p,
                                   c_void_p]),
        "laker_set_graphs": (c_int32, [c_int32]),
        "laker_batch_create": (c_int32, [c_int32, c_void_p, ctypes.POINTER(c_void_p)]),
        "laker_batch_destroy": (c_int32, [c_void_p]),
        "laker_batch_add": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_void_p, c_double, c_int32,
                                      c_double, c_int32, ctypes.POINTER(c_int32)]),
        "laker_batch_step": (c_int32, [c_void_p, c_int32, ctypes.POINTER(c_int32), ctypes.POINTER(c_int32),
                                       ctypes.POINTER(c_int32)]),
        "laker_batch_result": (c_int32, [c_void_p, c_int32, ctypes.POINTER(MixedResult)]),
        "laker_precond_create_jacobi": (c_int32, [c_void_p, c_int64, c_int32, c_void_p, ctypes.POINTER(c_void_p)]),
        "laker_precond_from_factors": (c_int32, [c_void_p, c_int64, c_int32, c_int32, c_double, c_double, c_void_p,
                                                 c_void_p, c_void_p, c_void_p, ctypes.POINTER(c_void_p)]),
        "laker_orthonormalize": (c_int32, [c_void_p, c_int64, c_int32, c_int32, c_void_p]),
        "laker_randn": (c_int32, [c_void_p, c_int64, c_int32, ctypes.c_uint64, c_void_p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()

EXPORTED = [
    "laker_last_error", "laker_version", "laker_operator_create", "laker_operator_destroy",
    "laker_operator_info", "laker_operator_set_lambda", "laker_matvec", "laker_matvec_noalloc",
    "laker_residual_fp64",
    "laker_cross_matvec", "laker_cross_matvec_t", "laker_kernel_eval", "laker_diagonal",
    "laker_to_dense", "laker_matvec_backward", "laker_cross_matvec_backward", "laker_precond_build",
    "laker_precond_build_from_responses", "laker_precond_copy_out", "laker_precond_destroy",
    "laker_precond_info", "laker_precond_apply", "laker_pcg_solve", "laker_solve_mixed",
    "laker_set_graphs", "laker_coldot", "laker_precond_create_jacobi", "laker_precond_from_factors",
    "laker_orthonormalize", "laker_randn", "laker_batch_create", "laker_batch_destroy", "laker_batch_add",
    "laker_batch_step", "laker_batch_result",
]


def last_error() -> str:
    msg = lib.laker_last_error()
    return msg.decode() if msg else ""


def check(rc: int) -> None:
    """Map a status code to the reference's exception types."""
    if rc == LAKER_OK:
        return
    msg = last_error()
    if rc == LAKER_EINVAL:
        raise ValueError(msg)
    raise RuntimeError(msg)


def require_gpu() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("the LAKER HIP backend needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def stream_ptr(device: torch.device | None = None) -> c_void_p:
    return c_void_p(torch.cuda.current_stream(device).cuda_stream)


def ptr(t: torch.Tensor | None) -> c_void_p:
    return c_void_p(0 if t is None else t.data_ptr())


def dtype_code(dtype: torch.dtype) -> int:
    if dtype == torch.float32:
        return LAKER_F32
    if dtype == torch.float64:
        return LAKER_F64
    raise ValueError(f"unsupported dtype {dtype}; the HIP solver supports float32 and float64")
