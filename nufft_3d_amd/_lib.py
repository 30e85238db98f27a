"""ctypes binding of the C ABI declared in ``include/nufft3d.h``.

The library is the in-tree ``libnufft3d.so`` built by ``nufft_3d_amd.build``.
There is no fallback: if it is missing, importing the operator fails loudly.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libnufft3d.so")

NUFFT3D_OK = 0
NUFFT3D_ERR_INVALID = -1
NUFFT3D_ERR_NOMEM = -2
NUFFT3D_ERR_DEVICE = -3
NUFFT3D_ERR_UNSUPPORTED = -4


class Opts(C.Structure):
    _fields_ = [("J", C.c_double), ("u", C.c_double), ("alpha", C.c_double),
                ("radial", C.c_int32), ("precision", C.c_int32),
                ("d", C.POINTER(C.c_double)), ("nd", C.c_int64),
                ("tile", C.c_int32), ("chunk", C.c_int32), ("device", C.c_int32),
                ("verbose", C.c_int32), ("compensated", C.c_int32),
                ("fp64_compute", C.c_int32)]


class PrunoOpts(C.Structure):
    _fields_ = [("width", C.c_int32), ("radial", C.c_int32), ("noise", C.c_double),
                ("tol", C.c_double), ("literal_first_nv", C.c_int32)]


class SakeOpts(C.Structure):
    _fields_ = [("width", C.c_int32), ("radial", C.c_int32), ("loraks", C.c_int32),
                ("tol", C.c_double), ("maxit", C.c_int32), ("noise", C.c_double),
                ("center", C.c_int32 * 3), ("damp", C.c_double), ("sparsity", C.c_double),
                ("dc_iters", C.c_int32), ("output", C.c_int32)]


class SakeInfo(C.Structure):
    _fields_ = [("iters", C.c_int32), ("rank", C.c_int32), ("noise", C.c_double),
                ("sigma", C.c_double), ("seconds", C.c_double)]


class Info(C.Structure):
    _fields_ = [("nrhs", C.c_int32), ("iters", C.c_int32 * 64),
                ("relres", C.c_double * 64), ("seconds", C.c_double)]


# name -> (restype, argtypes)
SIGNATURES = {
    "nufft3d_opts_default": (C.c_int, [C.POINTER(Opts)]),
    "nufft3d_plan_create": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(C.c_int32),
                                      C.POINTER(Opts), C.POINTER(C.c_void_p)]),
    "nufft3d_plan_destroy": (C.c_int, [C.c_void_p]),
    "nufft3d_plan_warnings": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int64]),
    "nufft3d_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]),
    "nufft3d_forward_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]),
    "nufft3d_adjoint": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]),
    "nufft3d_inverse": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_double,
                                  C.c_void_p, C.c_int64, C.c_double, C.c_void_p,
                                  C.POINTER(Info)]),
    "nufft3d_inverse_phase": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_double,
                                        C.c_void_p, C.c_int64, C.c_double, C.c_double, C.c_void_p,
                                        C.POINTER(Info)]),
    "nufft3d_inverse_cs": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_double,
                                     C.c_void_p, C.c_int64, C.c_double, C.c_double, C.c_void_p,
                                     C.POINTER(Info)]),
    "nufft3d_normal": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_double, C.c_void_p]),
    "nufft3d_discrep": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_double, C.c_int32,
                                  C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_void_p]),
    "nufft3d_density": (C.c_int, [C.c_void_p, C.c_void_p]),
    "nufft3d_deap": (C.c_int, [C.c_void_p, C.c_void_p]),
    "nufft3d_nnz": (C.c_int64, [C.c_void_p]),
    "nufft3d_get_params": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                     C.POINTER(C.c_double), C.POINTER(C.c_double),
                                     C.POINTER(C.c_double), C.POINTER(C.c_int32),
                                     C.POINTER(C.c_int32)]),
    "nufft3d_set_stream": (C.c_int, [C.c_void_p, C.c_void_p]),
    "nufft3d_set_graphs": (C.c_int, [C.c_void_p, C.c_int32]),
    "nufft3d_graph_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "nufft3d_device_bytes": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "nufft3d_compute_precision": (C.c_int32, [C.c_void_p]),
    "nufft3d_autotune": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_char_p, C.c_int64]),
    "nufft3d_release_workspaces": (C.c_int, [C.c_void_p]),
    "nufft3d_profile_enable": (C.c_int, [C.c_void_p, C.c_int32]),
    "nufft3d_profile_read": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int64]),
    "nufft3d_pruno_opts_default": (C.c_int, [C.POINTER(PrunoOpts)]),
    "nufft3d_gram": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_int32), C.c_int32,
                               C.c_int32, C.c_void_p, C.POINTER(C.c_int32)]),
    "nufft3d_gram_loraks": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_int32), C.c_int32,
                                      C.c_int32, C.c_void_p, C.c_void_p, C.POINTER(C.c_int32)]),
    "nufft3d_pruno": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(PrunoOpts),
                                C.POINTER(C.c_int32), C.c_void_p, C.POINTER(C.c_double),
                                C.c_void_p]),
    "nufft3d_inverse_pruno": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_double,
                                        C.c_void_p, C.c_int64, C.c_double, C.POINTER(PrunoOpts),
                                        C.c_int32, C.c_void_p, C.POINTER(Info)]),
    "nufft3d_sake_opts_default": (C.c_int, [C.POINTER(SakeOpts)]),
    "nufft3d_sake": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(SakeOpts),
                               C.c_void_p, C.c_void_p, C.POINTER(SakeInfo)]),
    "nufft3d_last_error": (C.c_char_p, []),
}

_lib = None


def load():
    """Load the HIP library (once). Raises if it has not been built.

    torch (when installed) is imported first so that one HIP runtime serves
    the process: the library's libamdhip64 dependency then binds to the copy
    torch already loaded, and torch tensors and streams can be passed in.
    """
    global _lib
    if _lib is not None:
        return _lib
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIBPATH):
        raise RuntimeError(
            "libnufft3d.so is not built (expected at %s); run "
            "`python -m nufft_3d_amd.build`" % LIBPATH)
    lib = C.CDLL(LIBPATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class NufftError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("nufft3d error %d: %s" % (code, msg))
        self.code = code


def check(code):
    if code != NUFFT3D_OK:
        raise NufftError(code, load().nufft3d_last_error().decode(errors="replace"))
    return code
