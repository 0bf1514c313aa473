"""Loader of the CPU oracle (oracle/_build/liboracle.so) -- TEST
INFRASTRUCTURE: used only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker."""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "_build", "liboracle.so")

_dp = C.POINTER(C.c_double)
_vp = C.c_void_p

SIG = {
    "oracle_create_handle": (_vp, []),
    "oracle_destroy_handle": (None, [_vp]),
    "oracle_init": (None, [_vp]),
    "oracle_set_Potential_Truncated": (None, [_vp, C.c_double, _dp, C.c_int, C.c_int]),
    "oracle_set_hf_fac": (None, [_vp, C.c_double]),
    "oracle_set_max_mem": (None, [_vp, C.c_int]),
    "oracle_set_P": (None, [_vp, _dp, C.c_int]),
    "oracle_set_P_polarized": (None, [_vp, _dp, _dp, C.c_int]),
    "oracle_set_K": (None, [_vp, _dp, C.c_int]),
    "oracle_set_K_polarized": (None, [_vp, _dp, _dp, C.c_int]),
    "oracle_get_K": (None, [_vp, _dp]),
    "oracle_get_K_polarized": (None, [_vp, _dp, _dp]),
    "oracle_set_cell": (None, [_vp, C.c_bool, _dp, _dp]),
    "oracle_set_neighs": (None, [_vp, _dp, C.c_int]),
    "oracle_set_Atom": (None, [_vp, C.c_int, _dp, _dp, C.c_int]),
    "oracle_set_Atom_L": (None, [_vp, C.c_int, C.c_int, C.c_int, _dp]),
    "oracle_add_prm": (None, [_vp, C.c_int, C.c_int, C.c_int, C.c_int]),
    "oracle_add_shell": (None, [_vp] + [C.c_int] * 6),
    "oracle_add_cell": (None, [_vp]),
    "oracle_add_qrt": (None, [_vp] + [C.c_int] * 8),
    "oracle_add_qrtt": (None, [_vp, C.c_double] + [C.c_int] * 12 + [C.c_uint] * 4 + [C.c_bool] * 4),
    "oracle_add_set": (None, [_vp]),
    "oracle_add_set_blocks": (None, [_vp, C.c_double] + [C.c_int] * 4 + [C.c_uint] * 4 + [C.c_bool] * 4),
    "oracle_dispatch": (None, [_vp]),
    "oracle_set_output": (None, [_vp, _dp, C.c_uint64]),
    "oracle_add_tra": (None, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64]),
    "oracle_get_output": (None, [_vp]),
    "oracle_norm_psi": (C.c_double, [_dp, _dp, C.c_int, C.c_int]),
    "oracle_stats": (None, [_vp, C.POINTER(C.c_uint64)]),
    "oracle_exp": (C.c_double, [C.c_double]),
    "oracle_boys": (None, [C.c_int, C.c_double, _dp, C.c_double]),
    "oracle_tcg": (C.c_int, [_vp, C.c_double, C.c_double, C.c_int, _dp]),
    "oracle_patch_grid": (None, [C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "oracle_boys_table": (None, [_dp]),
    "oracle_patch_bw": (None, [_dp]),
    "oracle_c2s": (None, [_dp]),
    "oracle_slot_key": (C.c_uint64, [C.c_int] * 15),
    "oracle_row_cart": (None, [_vp] + [C.c_int] * 11 + [C.POINTER(C.c_uint32), _dp]),
}

_lib = None


def load():
    global _lib
    if _lib is None:
        lib = C.CDLL(SO)
        for name, (res, args) in SIG.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib
