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


def oracle():
    """the CPU oracle (checker)"""
    import oracle as _o

    return _o.load()


def tcg_table(nder=21):
    """the C0 table (207 x (nder+1)*105), as numpy, and ld_C0"""
    full = np.fromfile(TCG_BIN, dtype="<f8").reshape(207, 22 * 105)
    return np.ascontiguousarray(full[:, : (nder + 1) * 105]), (nder + 1) * 105
