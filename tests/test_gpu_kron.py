"""The Kronecker HRR of the per-class kernel (the HRR of each side unrolled
into a dense map per row, the spherical transform folded in, two FP64
matrix-core products per contraction and row), forced for every class that
can take it (debug bit 3: with the matrix-core contraction where the class
has a place in the contraction buffer, the classes with many contractions
of an spdf centre; bit 4: with the class's own contraction where it fits,
every class of small and mid size; maps in LDS or in HBM), against the
oracle and against the same plan on the schedules the autotuner chose, to
the parity bar."""
import numpy as np
import pytest

from helpers import oracle

pytestmark = pytest.mark.gpu
KTOL_ABS, KTOL_REL = 1e-10, 1e-11


@pytest.mark.parametrize("n_tm,seed", [(1, 31), (2, 7)], ids=["1tm", "2tm"])
def test_kron_path_against_oracle(n_tm, seed):
    from libgint_amd import driver as D, load

    lib = load()
    o = oracle()
    s = D.System(n_water=2, n_tm=n_tm, box_A=7.0, image_shells=1, np_max=3, np_tm=3, seed=seed)
    st = s.record()
    P = s.P()
    Ko, ho = D.kbuild(o, "oracle_", s, st, P=P)
    o.oracle_destroy_handle(ho)
    kmax = np.max(np.abs(Ko))
    Ks = []
    for debug in (0, 8, 16):
        h, table = D.setup_handle(lib, "libgint_", s, P=P)
        lib.libgint_amd_plan_begin(h)
        st.replay(table, h)
        lib.libgint_amd_plan_end(h)
        lib.libgint_amd_set_debug(h, debug)
        for it in range(2):
            lib.libgint_set_P(h, D._ptr(P), s.P_size)
            lib.libgint_amd_plan_execute(h)
            K = np.zeros(s.P_size)
            lib.libgint_get_K(h, D._ptr(K))
            err = float(np.max(np.abs(K - Ko)))
            print("debug", debug, "execution", it, "max|dK|", err, "max|K|", kmax, flush=True)
            assert err <= KTOL_ABS + KTOL_REL * kmax, (debug, it, err)
        Ks.append(K)
        lib.libgint_amd_destroy_handle(h)
    for Kf in Ks[1:]:
        assert np.max(np.abs(Ks[0] - Kf)) <= KTOL_ABS + KTOL_REL * kmax


def test_kron_path_integral_output():
    """the integral output (the reference's TRA layout) of every class of a
    call stream with an spdf centre, through the Kronecker path forced as
    above (bits 3 and 4), against the oracle element by element"""
    from libgint_amd import driver as D, load

    lib = load()
    o = oracle()
    s = D.System(n_water=1, n_tm=1, box_A=6.0, image_shells=1, np_max=2, np_tm=2, seed=26)
    st = s.record()
    Io, ho = D.integrals(o, "oracle_", s, st)
    o.oracle_destroy_handle(ho)
    scale = np.max(np.abs(Io))
    for debug in (8, 16):
        Ip, hp = D.integrals(lib, "libgint_", s, st, debug=debug)
        lib.libgint_amd_destroy_handle(hp)
        err = np.max(np.abs(Ip - Io))
        print("debug", debug, "integrals", Io.size, "max|I|", scale, "max|dI|", err, flush=True)
        assert Io.size > 0 and scale > 0
        assert err <= KTOL_ABS + KTOL_REL * scale, (debug, err)
