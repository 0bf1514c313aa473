"""How precise can a K-build be?  A one-part-in-1e15 change of every
exponent (the rounding of the inputs) moves the oracle's own K by a few
1e-12 of max|K| in periodic f-shell cases.  This sets the relative part of
the GPU-vs-oracle tolerance (1e-11 max|K|) used by the parity tests."""
import ctypes as C

import numpy as np

from helpers import oracle


def test_kbuild_sensitivity_to_input_rounding():
    from libgint_amd import driver as D

    o = oracle()
    s = D.System(n_water=1, box_A=6.0, image_shells=1, basis=4, np_max=3, seed=2)
    st = s.record()
    K1, h1 = D.kbuild(o, "oracle_", s, st)
    o.oracle_destroy_handle(h1)
    h, table = D.setup_handle(o, "oracle_", s)
    SA = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int)

    def set_atom(hh, i, R, Z, n):
        z = np.array([Z[k] * (1 + 1e-15) for k in range(n)])
        o.oracle_set_Atom(hh, i, R, z.ctypes.data_as(C.POINTER(C.c_double)), n)

    cb = SA(set_atom)
    table.set_Atom = C.cast(cb, C.c_void_p).value
    D.drv().gdrv_emit_basis(s.h, C.byref(table), h)
    P = s.P()
    o.oracle_set_P(h, D._ptr(P), s.P_size)
    st.replay(table, h)
    K2 = np.zeros(s.P_size)
    o.oracle_get_K(h, D._ptr(K2))
    o.oracle_destroy_handle(h)
    rel = np.max(np.abs(K1 - K2)) / np.max(np.abs(K1))
    print("relative change of K under 1e-15 exponent perturbation:", rel)
    assert 1e-12 < rel < 1e-11
