"""libgint_amd: MI355X (gfx950) truncated-Coulomb exchange (K) build behind the
libgint_* C-ABI (include/libGint.h).

The product is the native library libgint_amd/libgint.so (C++ host engine +
HIP kernels).  This module only loads it with ctypes and declares the entry
points, for the tests, the smoke check and the benchmark; loading fails
loudly when the library has not been built.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgint.so")
DRV_PATH = os.path.join(HERE, "libgintdrv.so")
TCG_PATH = os.path.join(HERE, "data", "tcg_c0.bin")

_dp = C.POINTER(C.c_double)
_vp = C.c_void_p
_i = C.c_int
_u = C.c_uint
_b = C.c_bool

# name -> (restype, argtypes); the reference entry points first
ABI = {
    "libgint_create_handle": (_vp, []),
    "libgint_init": (None, [_vp]),
    "libgint_set_Potential_Truncated": (None, [_vp, C.c_double, _dp, _i, _i]),
    "libgint_set_hf_fac": (None, [_vp, C.c_double]),
    "libgint_set_max_mem": (None, [_vp, _i]),
    "libgint_set_P": (None, [_vp, _dp, _i]),
    "libgint_set_P_polarized": (None, [_vp, _dp, _dp, _i]),
    "libgint_set_K": (None, [_vp, _dp, _i]),
    "libgint_set_K_polarized": (None, [_vp, _dp, _dp, _i]),
    "libgint_get_K": (None, [_vp, _dp]),
    "libgint_get_K_polarized": (None, [_vp, _dp, _dp]),
    "libgint_set_cell": (None, [_vp, _b, _dp, _dp]),
    "libgint_set_neighs": (None, [_vp, _dp, _i]),
    "libgint_set_Atom": (None, [_vp, _i, _dp, _dp, _i]),
    "compute_norm_psi": (C.c_double, [_dp, _dp, _i, _i]),
    "libgint_set_Atom_L": (None, [_vp, _i, _i, _i, _dp]),
    "libgint_set_AtomInfo": (None, [_vp, _i, _dp, _dp, _i, _i, _i, C.POINTER(_i), _dp]),
    "libgint_add_prm": (None, [_vp, _i, _i, _i, _i]),
    "libgint_add_shell": (None, [_vp] + [_i] * 6),
    "libgint_add_cell": (None, [_vp]),
    "libgint_add_qrt": (None, [_vp] + [_i] * 8),
    "libgint_add_qrtt": (None, [_vp, C.c_double] + [_i] * 12 + [_u] * 4 + [_b] * 4),
    "libgint_add_set": (None, [_vp]),
    "libgint_dispatch": (None, [_vp]),
    # this library's own
    "libgint_amd_destroy_handle": (None, [_vp]),
    "libgint_amd_synchronize": (None, [_vp]),
    "libgint_amd_enable_stats": (None, [_vp, _i]),
    "libgint_amd_stats": (None, [_vp, C.POINTER(C.c_uint64)]),
    "libgint_amd_enable_timing": (None, [_vp, _i]),
    "libgint_amd_timing": (None, [_vp, _dp]),
    "libgint_amd_reset_counters": (None, [_vp]),
    "libgint_amd_timing_detail": (None, [_vp, _i, _dp]),
    "libgint_amd_set_debug": (None, [_vp, _i]),
    "libgint_amd_plan_begin": (None, [_vp]),
    "libgint_amd_plan_end": (None, [_vp]),
    "libgint_amd_plan_execute": (None, [_vp]),
    "libgint_amd_plan_clear": (None, [_vp]),
    "libgint_amd_plan_info": (None, [_vp, C.POINTER(C.c_uint64)]),
    "libgint_amd_queue_info": (None, [_vp, C.POINTER(C.c_uint64)]),
    "libgint_amd_math_selftest": (_i, [_dp, _dp, _i, _dp]),
    "libgint_amd_math_host": (None, [_dp, _dp, _i, _dp]),
}

_lib = None


def load():
    """the product library; raises if it is not built"""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libgint_amd/libgint.so is not built: run __graft_entry__.build()")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in ABI.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def reference_symbols():
    """the 24 entry points the reference exports (src/extern_functions.cpp)"""
    return [n for n in ABI if not n.startswith("libgint_amd_")]
