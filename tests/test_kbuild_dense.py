"""Pins the exchange bookkeeping of the oracle's K-build: the symmetric call
stream (unique atom quartets with symm_fac, transposed blocks, block
offsets) followed by the caller's symmetrisation must equal the dense
definition K_xy = -sum_ac (ax|cy) P_ac over every ordered quartet, with the
integrals from an independent McMurchie-Davidson evaluation (tests/md_ref.py)
and the spherical transform applied here."""
import numpy as np
import pytest

import md_ref
from helpers import as_dp, oracle


def _c2s():
    v = np.zeros(110)
    oracle().oracle_c2s(as_dp(v))
    off = [0, 1, 10, 40]
    return [v[off[l]: off[l] + (2 * l + 1) * (l + 1) * (l + 2) // 2].reshape(2 * l + 1, (l + 1) * (l + 2) // 2)
            for l in range(4)]


def _boys_aux(rho, T, L):
    from scipy.special import hyp1f1

    return np.array([hyp1f1(n + 0.5, n + 1.5, -T) / (2 * n + 1) for n in range(L + 1)])


@pytest.mark.parametrize("truncated", [False, True], ids=["coulomb", "truncated"])
def test_dense_exchange(truncated):
    from libgint_amd import driver as D

    lib = oracle()
    # a short cut-off (1.2 A) so that the truncation changes the integrals
    s = D.System(n_water=1, box_A=4.0, image_shells=0, np_min=1, np_max=1, seed=22, eps=0.0,
                 rcut_A=1.2 if truncated else 6.0)
    st = s.record()
    K, h = D.kbuild(lib, "oracle_", s, st, truncated=truncated)
    Kd = s.dense(s.symmetrize(K))
    Pd = s.dense(s.P())

    if truncated:
        Rcut = s.rcut()
        G = np.zeros(13)

        def aux(rho, T, L):
            lib.oracle_tcg(h, Rcut * np.sqrt(rho), T, L, as_dp(G))
            return G[: L + 1].copy()
    else:
        aux = _boys_aux
    c2s = _c2s()
    R, kind, f0 = s.atoms()
    basis = {k: s.kind_basis(k) for k in set(kind.tolist())}
    # normalised contractions per atom: list of (l, Z, Kn [nl x np], first function)
    shells = []
    for a in range(s.natom):
        Z, sh = basis[int(kind[a])]
        fun = int(f0[a])
        lst = []
        for (l, nl, Kraw) in sh:
            Kn = np.array([Kraw[n] / lib.oracle_norm_psi(as_dp(np.ascontiguousarray(Kraw[n])), as_dp(Z), len(Z), l)
                           for n in range(nl)])
            lst.append((l, Z, Kn, fun))
            fun += nl * (2 * l + 1)
        shells.append(lst)
    nbf = s.nbf
    I = np.zeros((nbf, nbf, nbf, nbf))
    for a in range(s.natom):
        for b in range(s.natom):
            for c in range(s.natom):
                for d in range(s.natom):
                    for (la, Za, Ka, fa) in shells[a]:
                        for (lb, Zb, Kb, fb) in shells[b]:
                            for (lc, Zc, Kc, fc) in shells[c]:
                                for (ld, Zd, Kd_, fd) in shells[d]:
                                    cart = 0.0
                                    for ia, za in enumerate(Za):
                                        for ib, zb in enumerate(Zb):
                                            for ic, zc in enumerate(Zc):
                                                for idd, zd in enumerate(Zd):
                                                    prim = md_ref.eri_prim(la, lb, lc, ld, za, zb, zc, zd, R[a], R[b],
                                                                           R[c], R[d], aux)
                                                    w = np.einsum("a,b,c,d->abcd", Ka[:, ia], Kb[:, ib], Kc[:, ic],
                                                                  Kd_[:, idd])
                                                    cart = cart + np.einsum("abcd,ijkl->aibjckdl", w, prim)
                                    sph = np.einsum("aibjckdl,pi,qj,rk,sl->apbqcrds", cart, c2s[la], c2s[lb],
                                                    c2s[lc], c2s[ld])
                                    na, nb_, nc, nd = (Ka.shape[0] * (2 * la + 1), Kb.shape[0] * (2 * lb + 1),
                                                       Kc.shape[0] * (2 * lc + 1), Kd_.shape[0] * (2 * ld + 1))
                                    I[fa:fa + na, fb:fb + nb_, fc:fc + nc, fd:fd + nd] = sph.reshape(na, nb_, nc, nd)
    Kref = -np.einsum("axcy,ac->xy", I, Pd)
    err = np.max(np.abs(Kd - Kref))
    print("max|K| %.3e max|dK| %.3e" % (np.max(np.abs(Kref)), err))
    assert np.max(np.abs(Kref)) > 1e-3
    assert err <= 1e-9 + 1e-10 * np.max(np.abs(Kref))
    lib.oracle_destroy_handle(h)
