"""Mixed precision for resident plans: set quartets whose calibrated bound
times max|P| is below tau take the single-precision matrix-core contraction.
The calibration execution is FP64 (parity bar against the oracle); the
routed executions must stay within the metric's 1e-8 Ha of the FP64 oracle
and must route some set quartets at the looser tau."""
import ctypes as C

import numpy as np
import pytest

from helpers import oracle

pytestmark = pytest.mark.gpu
KTOL_ABS, KTOL_REL = 1e-10, 1e-11


@pytest.mark.parametrize("tau", [1e-7, 1e-4], ids=["tau1e-7", "tau1e-4"])
def test_mixed_precision_routes_and_bounds_the_error(tau):
    from libgint_amd import driver as D, load

    lib = load()
    o = oracle()
    s = D.System(n_water=2, n_tm=1, box_A=7.0, image_shells=1, np_max=3, np_tm=3, seed=31)
    st = s.record()
    P = s.P()
    Ko, ho = D.kbuild(o, "oracle_", s, st, P=P)
    o.oracle_destroy_handle(ho)
    kmax = np.max(np.abs(Ko))
    h, table = D.setup_handle(lib, "libgint_", s, P=P)
    lib.libgint_amd_plan_begin(h)
    st.replay(table, h)
    lib.libgint_amd_plan_end(h)
    lib.libgint_amd_set_mixed_precision(h, tau)
    info = (C.c_double * 4)()
    errs = []
    for it in range(3):
        lib.libgint_set_P(h, D._ptr(P), s.P_size)
        lib.libgint_amd_plan_execute(h)
        K = np.zeros(s.P_size)
        lib.libgint_get_K(h, D._ptr(K))
        lib.libgint_amd_mixed_info(h, info)
        err = float(np.max(np.abs(K - Ko)))
        errs.append(err)
        print("tau", tau, "execution", it, "routed set quartets", info[2], "of", info[3], "max|dK|", err,
              "max|K|", kmax, flush=True)
        if it == 0:
            # the calibration execution is FP64
            assert err <= KTOL_ABS + KTOL_REL * kmax
        else:
            assert info[1] == 1.0
            assert err <= 1e-8, (tau, it, err)
    if tau >= 1e-4:
        assert info[2] > 0
    lib.libgint_amd_destroy_handle(h)
