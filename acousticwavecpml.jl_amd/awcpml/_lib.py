"""ctypes binding of libawcpml.so (the C ABI declared in include/awcpml.h).

The shared library is built in-tree (``acousticwavecpml.jl_amd/lib``) by
``acousticwavecpml.jl_amd/build.py``.  There is no fallback: if the library
is missing or cannot be loaded, every entry point raises.
"""
import ctypes
import os

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libawcpml.so")

OK = 0
ERR_ARG = -1
ERR_HIP = -2
ERR_NOMEM = -3
ERR_STATE = -4
ERR_RANGE = -5

F64 = 0
F32 = 1

_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p

# name -> (restype, argtypes); every symbol include/awcpml.h declares
SIGNATURES = {
    "awcpml_plan_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int64, ctypes.c_int64,
                                          ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.POINTER(_vp)]),
    "awcpml_plan_destroy": (ctypes.c_int, [_vp]),
    "awcpml_set_model": (ctypes.c_int, [_vp, _dp, _dp, _dp]),
    "awcpml_set_cpml": (ctypes.c_int, [_vp, ctypes.c_int] + [_dp] * 8),
    "awcpml_set_sources": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _i64p,
                                          ctypes.c_int64, _dp]),
    "awcpml_set_receivers": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _i64p]),
    "awcpml_reset": (ctypes.c_int, [_vp]),
    "awcpml_advance": (ctypes.c_int, [_vp, ctypes.c_int64]),
    "awcpml_forward": (ctypes.c_int, [_vp, ctypes.c_int64]),
    "awcpml_steps_done": (ctypes.c_int, [_vp, _i64p]),
    "awcpml_get_traces": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "awcpml_get_field": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _dp]),
    "awcpml_run": (ctypes.c_int, [_vp, _dp, _dp]),
    "awcpml_time_stencil": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "awcpml_time_advance": (ctypes.c_int, [_vp, ctypes.c_int64, _dp]),
    "awcpml_set_tiling": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "awcpml_get_tiling": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_int)]),
    "awcpml_set_shot_group": (ctypes.c_int, [_vp, ctypes.c_int]),
    "awcpml_get_shot_group": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int)]),
    "awcpml_set_time_block": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_int]),
    "awcpml_get_time_block": (ctypes.c_int, [_vp] + [ctypes.POINTER(ctypes.c_int)] * 4),
    "awcpml_set_two_step_align": (ctypes.c_int, [_vp, ctypes.c_int]),
    "awcpml_get_two_step_align": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int)]),
    "awcpml_time_stencil2": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "awcpml_set_graph": (ctypes.c_int, [_vp, ctypes.c_int]),
    "awcpml_set_wide_offsets": (ctypes.c_int, [_vp, ctypes.c_int]),
    "awcpml_batch_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, ctypes.POINTER(_vp)]),
    "awcpml_batch_destroy": (ctypes.c_int, [_vp]),
    "awcpml_batch_reset": (ctypes.c_int, [_vp]),
    "awcpml_batch_advance": (ctypes.c_int, [_vp, ctypes.c_int64]),
    "awcpml_batch_set_graph": (ctypes.c_int, [_vp, ctypes.c_int]),
    "awcpml_batch_time_advance": (ctypes.c_int, [_vp, ctypes.c_int64, _dp]),
    "awcpml_batch_time_stencil": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "awcpml_batch_set_time_block": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "awcpml_batch_get_time_block": (ctypes.c_int, [_vp] + [ctypes.POINTER(ctypes.c_int)] * 3),
    "awcpml_batch_time_stencil2": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "awcpml_checkpoint_run": (ctypes.c_int, [_vp, ctypes.c_int64]),
    "awcpml_adjoint": (ctypes.c_int, [_vp, _dp, ctypes.c_int64, _dp]),
    "awcpml_set_gradient_fold": (ctypes.c_int, [_vp, ctypes.c_int64]),
    "awcpml_set_adjoint_group": (ctypes.c_int, [_vp, ctypes.c_int]),
    "awcpml_get_adjoint_group": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int)]),
    "awcpml_run_noalloc": (ctypes.c_int, [ctypes.c_int, ctypes.c_int64, ctypes.c_int64,
                                          ctypes.c_int64, ctypes.c_int, ctypes.c_int, _vp, _dp,
                                          _dp, ctypes.POINTER(_vp), ctypes.POINTER(ctypes.c_int),
                                          ctypes.c_int, _i64p, ctypes.c_int64, _vp, ctypes.c_int,
                                          _i64p, _vp, _vp, _vp]),
    "awcpml_last_error": (ctypes.c_char_p, []),
    "awcpml_version": (ctypes.c_char_p, []),
}

_lib = None


class AwcpmlError(RuntimeError):
    """A libawcpml call returned a negative status."""

    def __init__(self, code, message):
        super().__init__(f"[{code}] {message}")
        self.code = code


def load():
    """Load libawcpml.so (raises if it is missing: no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"libawcpml.so not found at {LIB_PATH}; build it with "
                "`python acousticwavecpml.jl_amd/build.py` (or __graft_entry__.build())")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(code):
    if code != OK:
        msg = load().awcpml_last_error().decode("utf-8", "replace")
        raise AwcpmlError(code, msg)
    return code


def dptr(a):
    return a.ctypes.data_as(_dp)


def i64ptr(a):
    return a.ctypes.data_as(_i64p)
