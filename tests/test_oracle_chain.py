"""Pins the oracle's integral chain (prepare -> G_m/F_m -> OS VRR -> contraction
-> HGP HRR) against an independent McMurchie-Davidson evaluation."""
import ctypes as C
import math

import numpy as np
import pytest

import md_ref
from helpers import as_dp, oracle, tcg_table

BIG = 1000.0  # cubic cell edge (bohr) large enough that the minimum image is the identity


def _new_handle(truncated, R_cut=11.3384):
    lib = oracle()
    h = lib.oracle_create_handle()
    cell = np.eye(3) * BIG
    celli = np.eye(3) / BIG
    lib.oracle_set_cell(h, True, as_dp(np.asfortranarray(cell).ravel(order="F")), as_dp(np.asfortranarray(celli).ravel(order="F")))
    lib.oracle_set_neighs(h, as_dp(np.zeros(3)), 1)
    if truncated:
        c0, ld = tcg_table(21)
        lib.oracle_set_Potential_Truncated(h, R_cut, as_dp(c0.ravel()), ld, c0.size)
    return lib, h


def _sets(lib, h, shells, rng):
    """shells: list of (l, np, nl, R); returns exponents, normalised K"""
    data = []
    for i, (l, npr, nl, R) in enumerate(shells):
        Z = np.sort(10 ** rng.uniform(-0.8, 0.8, npr))[::-1].copy()
        K = rng.normal(size=(nl, npr))
        lib.oracle_set_Atom(h, i, as_dp(np.asarray(R, float)), as_dp(Z), npr)
        lib.oracle_set_Atom_L(h, i, l, nl, as_dp(np.ascontiguousarray(K).ravel()))
        Kn = np.array([K[n] / lib.oracle_norm_psi(as_dp(np.ascontiguousarray(K[n])), as_dp(Z), npr, l) for n in range(nl)])
        data.append((l, Z, Kn, np.asarray(R, float)))
    return data


def _md_contracted(data, aux):
    (la, Za, Ka, A), (lb, Zb, Kb, B), (lc, Zc, Kc, Cc), (ld, Zd, Kd, D) = data
    nla, nlb, nlc, nld = len(Ka), len(Kb), len(Kc), len(Kd)
    Nc = len(md_ref.cart_list(la)) * len(md_ref.cart_list(lb)) * len(md_ref.cart_list(lc)) * len(md_ref.cart_list(ld))
    out = np.zeros((nla * nlb * nlc * nld, Nc))
    for ia, za in enumerate(Za):
        for ib, zb in enumerate(Zb):
            for ic, zc in enumerate(Zc):
                for idd, zd in enumerate(Zd):
                    I = md_ref.eri_prim(la, lb, lc, ld, za, zb, zc, zd, A, B, Cc, D, aux).ravel()
                    lab = 0
                    for a in range(nla):
                        for b in range(nlb):
                            for c in range(nlc):
                                for d in range(nld):
                                    out[lab] += Ka[a, ia] * Kb[b, ib] * Kc[c, ic] * Kd[d, idd] * I
                                    lab += 1
    return out


CLASSES = [(0, 0, 0, 0), (1, 0, 0, 0), (0, 0, 1, 0), (1, 1, 1, 0), (1, 1, 1, 1), (2, 0, 2, 0),
           (2, 1, 0, 1), (0, 2, 1, 2), (2, 2, 2, 2), (3, 0, 0, 0), (3, 1, 2, 0), (1, 3, 0, 2), (3, 2, 1, 3)]


@pytest.mark.parametrize("truncated", [False, True])
@pytest.mark.parametrize("cls", CLASSES)
def test_row_vs_mcmurchie_davidson(cls, truncated):
    rng = np.random.default_rng(hash((cls, truncated)) % 2**32)
    lib, h = _new_handle(truncated)
    try:
        pos = rng.normal(size=(4, 3)) * 0.7  # compact: no slot falls under the 1e-12 gate
        nps = [int(rng.integers(1, 3)) for _ in range(4)]
        nls = [int(rng.integers(1, 3)) for _ in range(4)]
        shells = [(cls[t], nps[t], nls[t], pos[t]) for t in range(4)]
        data = _sets(lib, h, shells, rng)
        prm = np.array([a | (b << 8) | (c << 16) | (d << 24)
                        for a in range(nps[0]) for b in range(nps[1]) for c in range(nps[2]) for d in range(nps[3])],
                       dtype=np.uint32)
        nlabcd = int(np.prod(nls))
        Nc = int(np.prod([(l + 1) * (l + 2) // 2 for l in cls]))
        out = np.zeros(nlabcd * Nc)
        lib.oracle_row_cart(h, 0, 1, 2, 3, 0, 0, 0, 0, 0, 0, len(prm), prm.ctypes.data_as(C.POINTER(C.c_uint32)), as_dp(out))
        if truncated:
            def aux(rho, T, L):
                g = np.zeros(L + 1)
                lib.oracle_tcg(h, 11.3384 * math.sqrt(rho), T, L, as_dp(g))
                return g
        else:
            aux = md_ref.boys_aux
        ref = _md_contracted(data, aux).ravel()
        scale = np.max(np.abs(ref))
        assert scale > 0
        err = np.max(np.abs(out - ref)) / scale
        assert err < 1e-11, (cls, truncated, err)
    finally:
        lib.oracle_destroy_handle(h)
