"""ctypes binding of libsgldfea.so (the HIP C ABI declared in include/sgldfea.h).

The library must have been built (sgldbench_amd/build.py or
__graft_entry__.build()); there is no fallback path: a missing library raises.
"""
import ctypes as C
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libsgldfea.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_bp = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_i64 = C.c_int64
_vp = C.c_void_p

# name -> argtypes (restype int unless listed in _RESTYPE)
SIGNATURES = {
    "sgld_last_error": [],
    "sgld_version": [],
    "sgld_kbyu": [_dp, _i64, _ip, _i64, _dp, _i64, _i64, _dp, _i64, _ip, _lp, C.c_int, _dp],
    "sgld_kbyu8x8": [_dp, _i64, _ip, _i64, _dp, _i64, _i64, _ip, _lp, C.c_int, _dp],
    "sgld_restrict": [_dp, _i64, _ip, _i64, _dp, _i64, _i64, _ip, _i64, _i64, _i64, _ip, _dp,
                      _ip, _lp, C.c_int, _dp],
    "sgld_restrict_noalloc": [_vp, _i64, _vp, _i64, _vp, _i64, _i64, _vp, _i64, _i64, _i64, _vp, _vp,
                              _vp, _lp, C.c_int, _vp],
    "sgld_interpolate": [_dp, _i64, _ip, _i64, _dp, _i64, _i64, _ip, _i64, _i64, _i64, _ip, _dp,
                         _ip, _lp, C.c_int, _dp],
    "sgld_assemble_finest": [_dp, _i64, _i64, _dp, _i64, _ip, _i64, _i64, _dp],
    "sgld_assemble_nonfinest": [_dp, _i64, _ip, _i64, _i64, _dp],
    "sgld_unit_compliance": [_dp, _i64, _ip, _i64, _dp, _i64, _i64, _dp],
    "sgld_element_stiffness": [C.c_double, C.c_double, _dp],
    "sgld_ctx_create": [C.POINTER(_vp), C.c_int, C.c_int, C.c_int, _vp, C.c_int, C.c_int,
                        C.c_double, C.c_double],
    "sgld_ctx_destroy": [_vp],
    "sgld_ctx_info": [_vp, C.c_int, _lp],
    "sgld_ctx_set_fixed": [_vp, _bp],
    "sgld_ctx_set_modulus": [_vp, _dp],
    "sgld_ctx_assemble": [_vp, C.c_double],
    "sgld_ctx_kbyu": [_vp, _dp, _dp],
    "sgld_ctx_vcycle": [_vp, _dp, _dp],
    "sgld_ctx_pcg": [_vp, _dp, _dp, C.c_double, C.c_int, C.c_int, C.POINTER(C.c_int),
                     C.POINTER(C.c_double)],
    "sgld_ctx_pcg_device": [_vp, _vp, _vp, C.c_double, C.c_int, C.c_int, C.POINTER(C.c_int),
                            C.POINTER(C.c_double)],
    "sgld_ctx_unit_compliance": [_vp, _dp, _dp],
    "sgld_ctx_kbyu_device": [_vp, _vp, _vp, C.c_int],
    "sgld_ctx_bench_kbyu": [_vp, C.c_int, C.c_int, C.POINTER(C.c_double)],
    "sgld_ctx_pcg_batched": [_vp, C.c_int, _dp, _dp, C.c_double, C.c_int, C.c_int, _ip, _dp],
    "sgld_ctx_pcg_batched_device": [_vp, C.c_int, _vp, _vp, C.c_double, C.c_int, C.c_int, _ip, _dp],
    "sgld_ctx_kbyu_batched_device": [_vp, C.c_int, _vp, _vp],
    "sgld_ctx_bench_kbyu_batched": [_vp, C.c_int, C.POINTER(C.c_double)],
    "sgld_ctx_sensitivity": [_vp, _dp, _dp, C.c_double, C.c_double, C.c_double, C.c_double, _dp,
                             C.POINTER(C.c_double)],
    "sgld_ctx_sensitivity_device": [_vp, _vp, _vp, C.c_double, C.c_double, C.c_double, C.c_double, _vp,
                                    C.POINTER(C.c_double)],
    "sgld_ctx_autotune": [_vp, C.c_int, _dp],
    "sgld_ctx_set_schedule": [_vp, C.c_int, C.c_int, C.c_int, C.c_int],
    "sgld_multi_pcg_device": [C.POINTER(_vp), C.c_int, C.POINTER(_vp), C.POINTER(_vp), C.c_double, C.c_int,
                              C.c_int, C.c_int, _ip, _dp, _dp],
    "sgld_multi_sensitivity_device": [C.POINTER(_vp), C.c_int, C.POINTER(_vp), C.POINTER(_vp), C.c_double,
                                      C.c_double, C.c_double, C.c_double, C.POINTER(_vp), _dp],
    "sgld_density_filter_setup": [C.c_double, _i64, _ip, C.c_int, C.c_int, C.c_int, _dp],
    "sgld_density_filter": [_dp, C.c_double, _i64, _ip, C.c_int, C.c_int, C.c_int, _dp],
    "sgld_ctx_density_filter_device": [_vp, _vp, C.c_double, _vp],
    "sgld_ctx_stress": [_vp, _dp, _dp, _dp],
    "sgld_ctx_stress_device": [_vp, _vp, _vp, _vp],
    "sgld_ctx_pde_filter": [_vp, _dp, C.c_double, C.c_double, C.c_int, _dp, C.POINTER(C.c_int)],
    "sgld_ctx_pde_filter_device": [_vp, _vp, C.c_double, C.c_double, C.c_int, _vp, C.POINTER(C.c_int)],
    "sgld_ctx_stats": [_vp, _dp],
    "sgld_ctx_kernel_timing": [_vp, C.c_int],
    "sgld_ctx_kernel_times": [_vp, _dp],
    "sgld_ctx_kernel_name": [_vp, C.c_int],
}
_RESTYPE = {"sgld_last_error": C.c_char_p, "sgld_version": C.c_char_p, "sgld_ctx_kernel_name": C.c_char_p}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libsgldfea.so is not built (run __graft_entry__.build()); "
                               "there is no CPU fallback")
        # torch ships its own libamdhip64.so.7.  Loading it first lets this
        # library bind to that same HIP runtime (same soname), so a process
        # that also uses torch tensors holds a single runtime; loading this
        # library first would pull in a second copy from /opt/rocm and torch
        # would then fail to initialise the device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RESTYPE.get(name, C.c_int)
        _lib = L
    return _lib


class SGLDError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code
        self.identifier = msg.split(":")[0] + ":" + msg.split(":")[1] if msg.count(":") >= 2 else msg


OK, ERR_ARG, ERR_HIP, ERR_NOT_SPD, ERR_MAXIT, ERR_STATE, ERR_NOMEM = range(7)  # include/sgldfea.h


def check(rc, allow=()):
    if rc != 0 and rc not in allow:
        raise SGLDError(rc, lib().sgld_last_error().decode())
    return rc
