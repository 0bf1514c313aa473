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
    "libgint_amd_set_output": (None, [_vp, _dp, C.c_uint64]),
    "libgint_amd_add_tra": (None, [_vp, _i, _i, _i, _i, C.c_uint64]),
    "libgint_amd_get_output": (None, [_vp]),
    "libgint_amd_destroy_handle": (None, [_vp]),
    "libgint_amd_synchronize": (None, [_vp]),
    "libgint_amd_enable_stats": (None, [_vp, _i]),
    "libgint_amd_stats": (None, [_vp, C.POINTER(C.c_uint64)]),
    "libgint_amd_enable_timing": (None, [_vp, _i]),
    "libgint_amd_timing": (None, [_vp, _dp]),
    "libgint_amd_timing_eco": (None, [_vp, _dp]),
    "libgint_amd_reset_counters": (None, [_vp]),
    "libgint_amd_timing_detail": (None, [_vp, _i, _dp]),
    "libgint_amd_set_debug": (None, [_vp, _i]),
    "libgint_amd_set_graph": (None, [_vp, _i]),
    "libgint_amd_set_autotune": (None, [_vp, _i]),
    "libgint_amd_set_mixed_precision": (None, [_vp, C.c_double]),
    "libgint_amd_mixed_info": (None, [_vp, _dp]),
    "libgint_amd_tune_info": (None, [_vp, _dp]),
    "libgint_amd_tune_detail": (None, [_vp, _dp, C.POINTER(_i)]),
    "libgint_amd_graph_info": (None, [_vp, C.POINTER(C.c_uint64)]),
    "libgint_amd_plan_begin": (None, [_vp]),
    "libgint_amd_plan_end": (None, [_vp]),
    "libgint_amd_plan_execute": (None, [_vp]),
    "libgint_amd_plan_execute_noalloc": (_i, [_vp, _vp, _vp, _vp, _vp, C.c_uint64]),
    "libgint_amd_plan_clear": (None, [_vp]),
    "libgint_amd_plan_info": (None, [_vp, C.POINTER(C.c_uint64)]),
    "libgint_amd_queue_info": (None, [_vp, C.POINTER(C.c_uint64)]),
    "libgint_amd_queue_types": (_i, [_vp, C.POINTER(C.c_uint64), _i]),
    "libgint_amd_add_set_blocks": (None, [_vp, C.c_double] + [_i] * 4 + [_u] * 4 + [_b] * 4),
    "libgint_amd_mfma_selftest": (_i, [_dp, _dp, _dp, _i]),
    "libgint_amd_math_selftest": (_i, [_dp, _dp, _i, _dp]),
    "libgint_amd_math_host": (None, [_dp, _dp, _i, _dp]),
    "libgint_amd_server_create": (_vp, [_i, _i]),
    "libgint_amd_server_add_client": (_i, [_vp, _vp, _i]),
    "libgint_amd_server_submit": (C.c_int64, [_vp, _i, _dp, _dp]),
    "libgint_amd_server_wait": (_i, [_vp, C.c_int64, _dp]),
    "libgint_amd_server_stats": (None, [_vp, _dp]),
    "libgint_amd_server_destroy": (None, [_vp]),
    "libgint_amd_server_client_bytes": (C.c_int64, [_vp, _i]),
    "libgint_amd_server_set_budget": (None, [_vp, _i]),
    "libgint_amd_server_state": (None, [_vp, C.POINTER(C.c_int64)]),
    "libgint_amd_read_c0": (_i, [C.c_char_p, _i, _dp, _i, C.POINTER(_i)]),
    "libgint_amd_write_c0": (_i, [C.c_char_p, _dp, _i]),
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


def read_c0(path, nder=21):
    """CP2K's t_c_g.dat coefficient file -> (C0 as 207 x (nder+1)*105 float64,
    ld_C0), through the library's reader; raises on a malformed file"""
    import numpy as np

    ld = (nder + 1) * 105
    c0 = np.empty((207, ld))
    got = C.c_int(0)
    n = load().libgint_amd_read_c0(os.fsencode(path), nder, c0.ctypes.data_as(_dp), c0.size, C.byref(got))
    if n < 0:
        raise ValueError("libgint_amd_read_c0(%s, nder=%d) failed with %d" % (path, nder, n))
    return c0, got.value


def write_c0(path, c0):
    """writes a full 207 x 2310 table in the t_c_g.dat format"""
    import numpy as np

    c0 = np.ascontiguousarray(c0, dtype=np.float64)
    if c0.shape != (207, 2310):
        raise ValueError("write_c0 takes the full 207 x 2310 table")
    rc = load().libgint_amd_write_c0(os.fsencode(path), c0.ctypes.data_as(_dp), 2310)
    if rc != 0:
        raise OSError("libgint_amd_write_c0(%s) failed with %d" % (path, rc))


def reference_symbols():
    """the 24 entry points the reference exports (src/extern_functions.cpp)"""
    return [n for n in ABI if not n.startswith("libgint_amd_")]
