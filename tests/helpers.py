"""Shared test helpers: locate and load the oracle (the checker), the
product library and the table; nothing here is product code."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libref.so")
TCG_BIN = os.path.join(ROOT, "libgint_amd", "data", "tcg_c0.bin")

dp = C.POINTER(C.c_double)
vp = C.c_void_p


def as_dp(a):
    return a.ctypes.data_as(dp)


_oracle = None


def oracle():
    global _oracle
    if _oracle is not None:
        return _oracle
    lib = C.CDLL(ORACLE_SO)
    sig = {
        "oracle_create_handle": (vp, []),
        "oracle_destroy_handle": (None, [vp]),
        "oracle_init": (None, [vp]),
        "oracle_set_Potential_Truncated": (None, [vp, C.c_double, dp, C.c_int, C.c_int]),
        "oracle_set_hf_fac": (None, [vp, C.c_double]),
        "oracle_set_max_mem": (None, [vp, C.c_int]),
        "oracle_set_P": (None, [vp, dp, C.c_int]),
        "oracle_set_P_polarized": (None, [vp, dp, dp, C.c_int]),
        "oracle_set_K": (None, [vp, dp, C.c_int]),
        "oracle_set_K_polarized": (None, [vp, dp, dp, C.c_int]),
        "oracle_get_K": (None, [vp, dp]),
        "oracle_get_K_polarized": (None, [vp, dp, dp]),
        "oracle_set_cell": (None, [vp, C.c_bool, dp, dp]),
        "oracle_set_neighs": (None, [vp, dp, C.c_int]),
        "oracle_set_Atom": (None, [vp, C.c_int, dp, dp, C.c_int]),
        "oracle_set_Atom_L": (None, [vp, C.c_int, C.c_int, C.c_int, dp]),
        "oracle_add_prm": (None, [vp, C.c_int, C.c_int, C.c_int, C.c_int]),
        "oracle_add_shell": (None, [vp] + [C.c_int] * 6),
        "oracle_add_cell": (None, [vp]),
        "oracle_add_qrt": (None, [vp] + [C.c_int] * 8),
        "oracle_add_qrtt": (None, [vp, C.c_double] + [C.c_int] * 12 + [C.c_uint] * 4 + [C.c_bool] * 4),
        "oracle_add_set": (None, [vp]),
        "oracle_dispatch": (None, [vp]),
        "oracle_norm_psi": (C.c_double, [dp, dp, C.c_int, C.c_int]),
        "oracle_stats": (None, [vp, C.POINTER(C.c_uint64)]),
        "oracle_exp": (C.c_double, [C.c_double]),
        "oracle_boys": (None, [C.c_int, C.c_double, dp, C.c_double]),
        "oracle_tcg": (C.c_int, [vp, C.c_double, C.c_double, C.c_int, dp]),
        "oracle_patch_grid": (None, [C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "oracle_boys_table": (None, [dp]),
        "oracle_slot_key": (C.c_uint64, [C.c_int] * 15),
        "oracle_row_cart": (None, [vp] + [C.c_int] * 11 + [C.POINTER(C.c_uint32), dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _oracle = lib
    return lib


def tcg_table(nder=21):
    """the C0 table (207 x (nder+1)*105), as numpy, and ld_C0"""
    full = np.fromfile(TCG_BIN, dtype="<f8").reshape(207, 22 * 105)
    return np.ascontiguousarray(full[:, : (nder + 1) * 105]), (nder + 1) * 105
