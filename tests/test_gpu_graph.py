"""Resident plans replayed as one HIP graph: every replay against the eager
execution of the same plan and density, with eager executions and host
synchronisations between the replays, and one replay against the oracle."""
import ctypes as C

import numpy as np
import pytest

from helpers import oracle

pytestmark = pytest.mark.gpu
KTOL_ABS, KTOL_REL = 1e-10, 1e-11


def _k(lib, D, s, h, P):
    lib.libgint_set_P(h, D._ptr(P), s.P_size)
    lib.libgint_amd_plan_execute(h)
    K = np.zeros(s.P_size)
    lib.libgint_get_K(h, D._ptr(K))
    return K


def _info(lib, h):
    out = (C.c_uint64 * 4)()
    lib.libgint_amd_graph_info(h, out)
    return list(out)


@pytest.mark.parametrize("name,sysargs", [
    ("water2_pbc27", dict(n_water=2, box_A=7.0, image_shells=1, np_max=3, seed=12)),
    ("water1_tm1_pbc27_spdf", dict(n_water=1, n_tm=1, box_A=6.0, image_shells=1, np_max=2, np_tm=2, seed=5)),
], ids=["water2_pbc27", "water1_tm1_pbc27_spdf"])
def test_graph_replays_match_eager(name, sysargs):
    from libgint_amd import driver as D, load

    lib = load()
    s = D.System(**sysargs)
    st = s.record()
    rng = np.random.default_rng(11)
    P1 = s.P()
    P2 = P1 * (1.0 + 0.3 * rng.standard_normal(P1.size))
    h, table = D.setup_handle(lib, "libgint_", s, P=P1)
    lib.libgint_amd_plan_begin(h)
    st.replay(table, h)
    lib.libgint_amd_plan_end(h)
    lib.libgint_amd_set_graph(h, 0)
    eager = {1: _k(lib, D, s, h, P1), 2: _k(lib, D, s, h, P2)}
    assert _info(lib, h)[1] == 0
    lib.libgint_amd_set_graph(h, 1)
    order = [1, 2, "eager", 1, 2, "eager", 2, 1]
    n_graph = 0
    for x in order:
        if x == "eager":
            # eager work and a synchronisation between replays
            lib.libgint_amd_set_graph(h, 0)
            Ke = _k(lib, D, s, h, P2)
            lib.libgint_amd_synchronize(h)
            lib.libgint_amd_set_graph(h, 1)
            assert np.max(np.abs(Ke - eager[2])) <= KTOL_ABS + KTOL_REL * np.max(np.abs(eager[2]))
            continue
        K = _k(lib, D, s, h, P1 if x == 1 else P2)
        n_graph += 1
        ref = eager[x]
        err = np.max(np.abs(K - ref))
        print(name, "replay", n_graph, "density", x, "max|dK| vs eager", err, flush=True)
        assert np.max(np.abs(ref)) > 0 and err <= KTOL_ABS + KTOL_REL * np.max(np.abs(ref)), (n_graph, err)
    info = _info(lib, h)
    print(name, "graph info", info, flush=True)
    assert info[1] == 1 and info[2] == n_graph and info[3] == 1
    # the graph's K against the oracle
    o = oracle()
    Ko, ho = D.kbuild(o, "oracle_", s, st, P=P2)
    o.oracle_destroy_handle(ho)
    K = _k(lib, D, s, h, P2)
    assert np.max(np.abs(K - Ko)) <= KTOL_ABS + KTOL_REL * np.max(np.abs(Ko))
    lib.libgint_amd_destroy_handle(h)
