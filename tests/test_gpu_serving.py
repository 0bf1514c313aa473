"""The device-resident request loop: K-build requests of several clients
(independent periodic systems, each with a resident plan and its density and
exchange matrix kept on the device) arriving at random times, admitted as
worker slots and the device budget allow, with evictions of idle clients.
Each request's K in the mix must equal the same request's K computed alone
(the client's plan executed on its own, eagerly) to the parity bar, and one
request per client is checked against the oracle."""
import numpy as np
import pytest

from helpers import oracle

pytestmark = pytest.mark.gpu
KTOL_ABS, KTOL_REL = 1e-10, 1e-11


def _alone(lib, D, c, P):
    """the same K-build on a handle of its own: set_P, plan, get_K"""
    h, table = D.setup_handle(lib, "libgint_", c.s, P=P)
    lib.libgint_amd_plan_begin(h)
    c.st.replay(table, h, 0, c.st.n_sets, 1, True)
    lib.libgint_amd_plan_end(h)
    lib.libgint_amd_set_graph(h, 0)
    lib.libgint_amd_plan_execute(h)
    K = np.zeros(c.s.P_size)
    lib.libgint_get_K(h, D._ptr(K))
    lib.libgint_amd_destroy_handle(h)
    return K


def test_requests_alone_and_in_the_mix():
    from libgint_amd import driver as D, load

    lib = load()
    systems = D.serving_systems(n_systems=4, atoms=(9, 30), seed=5, np_max=2)
    clients = [D.ServingClient(lib, s, steps=3, seed=i) for i, s in enumerate(systems)]
    for c in clients:
        print("client", c.s.describe(), "sets", c.st.n_sets, "prm", c.st.n_prm, flush=True)
    # a budget that holds about two clients at a time: idle ones are evicted
    done, st = D.serve(lib, clients, n_requests=24, rate=400.0, slots=3, max_mem_mib=65536, seed=2, keep=True,
                       resident_clients=2)
    print("served", st[0], "p50 %.4f p99 %.4f s" % (st[1], st[2]), "peak resident MiB %.1f" % (st[5] / 2**20),
          "evictions", st[6], "admissions", st[7], flush=True)
    assert st[0] == 24 and st[6] > 0
    o = oracle()
    checked = set()
    for i, (c, P, K, rc, lat, wait, run) in enumerate(done):
        assert rc == 0, (i, rc)
        Ka = _alone(lib, D, clients[c], P)
        err = np.max(np.abs(K - Ka))
        tol = KTOL_ABS + KTOL_REL * np.max(np.abs(Ka))
        print("request", i, "client", c, "latency %.4f" % lat, "max|K| %.3g" % np.max(np.abs(Ka)),
              "max|dK| mix vs alone %.3g" % err, flush=True)
        assert np.max(np.abs(Ka)) > 0 and err <= tol, (i, err, tol)
        if c not in checked:
            Ko, ho = D.kbuild(o, "oracle_", clients[c].s, clients[c].st, P=P)
            o.oracle_destroy_handle(ho)
            assert np.max(np.abs(K - Ko)) <= KTOL_ABS + KTOL_REL * np.max(np.abs(Ko))
            checked.add(c)
    for c in clients:
        lib.libgint_amd_destroy_handle(c.h)
