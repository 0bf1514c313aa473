"""The OpenMP team boundary, as CP2K drives the library: one handle per
thread, libgint_init / set_P / get_K called by every thread of the team
(collective; the device P and K are shared by the team, reference
src/libGint.cpp:37-60,441-447,613-635), each thread sending a disjoint
slice of the set quartets.  The team's K must be the oracle's K of the
whole stream, to the parity bar."""
import numpy as np
import pytest

from helpers import oracle

pytestmark = pytest.mark.gpu
KTOL_ABS, KTOL_REL = 1e-10, 1e-11


@pytest.mark.parametrize("nthreads,plan,blocks", [(4, False, False), (4, True, False), (3, True, True)],
                         ids=["4_threads_stream", "4_threads_plan", "3_threads_plan_blocks"])
def test_team_kbuild(nthreads, plan, blocks):
    from libgint_amd import driver as D, load

    lib = load()
    o = oracle()
    s = D.System(n_water=2, n_tm=1, box_A=7.0, image_shells=1, np_max=3, np_tm=2, seed=27)
    st = s.record()
    K, (t_ing, t_all) = D.team_kbuild(lib, s, st, nthreads, plan=plan, blocks=blocks)
    Ko, ho = D.kbuild(o, "oracle_", s, st)
    o.oracle_destroy_handle(ho)
    err = np.max(np.abs(K - Ko))
    tol = KTOL_ABS + KTOL_REL * np.max(np.abs(Ko))
    print("team", nthreads, "plan", plan, "sets", st.n_sets, "max|dK|", err, "tol", tol,
          "ingestion %.3fs k-build %.3fs" % (t_ing, t_all), flush=True)
    assert np.max(np.abs(Ko)) > 0 and err <= tol


def test_team_ingestion_speedup():
    """per-call ingestion of a larger stream (8 waters) by 1 and by 4
    threads of a team: the same K, and the slowest thread's ingestion time
    of each (printed; the host engine has no shared state on this path)"""
    from libgint_amd import driver as D, load

    lib = load()
    s = D.System(0)
    st = s.record()
    K1, (i1, a1) = D.team_kbuild(lib, s, st, 1, plan=True)
    K4, (i4, a4) = D.team_kbuild(lib, s, st, 4, plan=True)
    err = np.max(np.abs(K4 - K1))
    print("8 waters: sets %d qrtt %d; ingestion 1 thread %.3fs, 4 threads %.3fs (x%.2f); k-build %.3fs / %.3fs; "
          "max|dK| %.2e" % (st.n_sets, st.n_qrtt, i1, i4, i1 / max(i4, 1e-9), a1, a4, err), flush=True)
    assert err <= KTOL_ABS + KTOL_REL * np.max(np.abs(K1))
    assert i4 < i1
