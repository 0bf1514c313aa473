"""Pins the oracle's integral output (the reference's TRA layout,
src/compute_TRA.cpp:36-47: [ila, isa, ilb, isb, ilc, isc, ild, isd] per
(set quartet, class), classes in shell order) against spherical integrals
built here from an independent McMurchie-Davidson evaluation
(tests/md_ref.py), contracted with the normalised coefficients and
transformed with the c2s rows.  A one-cell system with no Schwarz
screening, so that every primitive quartet is in the stream."""
import numpy as np
import pytest

import md_ref
from helpers import as_dp, oracle
from test_kbuild_dense import _boys_aux, _c2s


@pytest.mark.parametrize("truncated", [False, True], ids=["coulomb", "truncated"])
def test_tra_output_matches_md(truncated):
    from libgint_amd import driver as D

    lib = oracle()
    s = D.System(n_water=1, box_A=4.0, image_shells=0, basis=0, np_min=1, np_max=2, seed=22, eps=0.0,
                 rcut_A=1.2 if truncated else 6.0)
    st = s.record()
    out, h = D.integrals(lib, "oracle_", s, st, truncated=truncated)
    if truncated:
        Rcut = s.rcut()
        G = np.zeros(13)

        def aux(rho, T, L):
            lib.oracle_tcg(h, Rcut * np.sqrt(rho), T, L, as_dp(G))
            return G[: L + 1].copy()
    else:
        aux = _boys_aux
    c2s = _c2s()
    R, kind, _ = s.atoms()
    basis = {k: s.kind_basis(k) for k in set(kind.tolist())}

    def shells(a):
        Z, sh = basis[int(kind[a])]
        return [(l, Z, np.array([K[n] / lib.oracle_norm_psi(as_dp(np.ascontiguousarray(K[n])), as_dp(Z), len(Z), l)
                                 for n in range(nl)])) for (l, nl, K) in sh]

    pos = 0
    worst = 0.0
    for q in st.sets():
        sh = [shells(int(a)) for a in q]
        for (la, Za, Ka) in sh[0]:
            for (lb, Zb, Kb) in sh[1]:
                for (lc, Zc, Kc) in sh[2]:
                    for (ld, Zd, Kd) in sh[3]:
                        cart = 0.0
                        for ia, za in enumerate(Za):
                            for ib, zb in enumerate(Zb):
                                for ic, zc in enumerate(Zc):
                                    for idd, zd in enumerate(Zd):
                                        prim = md_ref.eri_prim(la, lb, lc, ld, za, zb, zc, zd, R[q[0]], R[q[1]],
                                                               R[q[2]], R[q[3]], aux)
                                        w = np.einsum("a,b,c,d->abcd", Ka[:, ia], Kb[:, ib], Kc[:, ic], Kd[:, idd])
                                        cart = cart + np.einsum("abcd,ijkl->aibjckdl", w, prim)
                        sph = np.einsum("aibjckdl,pi,qj,rk,sl->apbqcrds", cart, c2s[la], c2s[lb], c2s[lc], c2s[ld])
                        got = out[pos:pos + sph.size]
                        pos += sph.size
                        err = np.max(np.abs(got - sph.ravel()))
                        worst = max(worst, err / (1.0 + np.max(np.abs(sph))))
                        assert err <= 1e-9 + 1e-10 * np.max(np.abs(sph)), (q, la, lb, lc, ld, err)
    assert pos == out.size
    print("set quartets", st.n_sets, "integrals", out.size, "worst relative error", worst)
    lib.oracle_destroy_handle(h)
