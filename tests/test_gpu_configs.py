"""Parity at the benchmark workloads' own sizes: the HIP K-build against the
CPU oracle on the full stream of the 8-water cluster and on reproducible hash
samples of the set quartets of the periodic workloads (driver: 1 in N set
quartets (i,j,k,l) by index hash, all their rows and primitive quartets),
replayed identically into both libraries through the per-call ABI.

Bar, as for the small cases: screening decisions bit-exact (count and
order-independent checksum of the surviving slots) and
max|K_gpu - K_oracle| <= 1e-10 Ha + 1e-11 max|K|; the max|dK| is also
printed against the metric's 1e-8 Ha."""
import time

import numpy as np
import pytest

from helpers import oracle

pytestmark = pytest.mark.gpu
KTOL_ABS, KTOL_REL = 1e-10, 1e-11

CASES = [
    # 8 H2O cluster, every set quartet (24 k set quartets, 1.07 M primitive quartets)
    ("water8_cluster_full", 0, 1),
    # 32 H2O, 12.4 A box, 27 cells: 1 in 500 set quartets
    ("water32_pbc27_sample500", 1, 500),
    # 64 H2O + 8 spdf centres (f shells, 16 primitives, 4 contractions),
    # 15.6 A box, 125 cells: 1 in 2000 set quartets, about a sixth of them
    # with an spdf centre
    ("water64_tm8_pbc125_sample2000", 2, 2000),
    # 128 H2O with O 4s3p2d1f / H 3s2p1d (up to 16 primitives), 15.6 A box,
    # 256 cells: 1 in 20000 set quartets
    ("water128_spdf_pbc256_sample20000", 4, 20000),
]


@pytest.mark.parametrize("name,config,sample", CASES, ids=[c[0] for c in CASES])
def test_workload_parity(name, config, sample):
    from libgint_amd import driver as D, load

    lib = load()
    o = oracle()
    s = D.System(config)
    st = s.record(sample=sample)
    sets = st.sets()
    _, kind, _ = s.atoms()
    n_tm_sets = int(np.sum(np.any(kind[sets] == 2, axis=1)))
    print(name, s.describe(), "| sample 1 in", sample, "sets", st.n_sets, "rows", st.n_rows, "prm", st.n_prm,
          "qrtt", st.n_qrtt, "sets with an spdf centre", n_tm_sets, flush=True)
    if config == 2:
        assert n_tm_sets > 0
    t = time.perf_counter()
    Kp, hp = D.kbuild(lib, "libgint_", s, st, stats=True)
    tg = time.perf_counter() - t
    sp = D.stats(lib, "libgint_", hp)
    lib.libgint_amd_destroy_handle(hp)
    t = time.perf_counter()
    Ko, ho = D.kbuild(o, "oracle_", s, st)
    to = time.perf_counter() - t
    so = D.stats(o, "oracle_", ho)
    o.oracle_destroy_handle(ho)
    err = float(np.max(np.abs(Kp - Ko)))
    kmax = float(np.max(np.abs(Ko)))
    tol = KTOL_ABS + KTOL_REL * kmax
    print(name, "surviving slots", so[1], "(gpu", sp[1], ") max|K|", kmax, "max|dK|", err, "tol", tol,
          "metric bar 1e-8:", err <= 1e-8, "| gpu path %.2fs, oracle %.2fs" % (tg, to), flush=True)
    assert sp[1] == so[1], (name, "surviving slots differ", sp[1], so[1])
    assert sp[4] == so[4], (name, "survivor checksum differs")
    assert kmax > 0 and err <= tol, (name, err, tol)
