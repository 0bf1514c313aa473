"""Pins of the oracle's tables and scalar functions.

Each test compares one function of the oracle with something it did not
produce: the reference's own tables compiled in place (oracle/_ref, where
the reference checkout exists), closed forms evaluated with mpmath (the
committed golden file tests/golden/tcg_golden.json and direct
evaluation), or exact algebra.
"""
import ctypes as C
import json
import math
import os

import numpy as np
import pytest

from helpers import REF_SO, ROOT, as_dp, oracle, tcg_table

GOLDEN = os.path.join(ROOT, "tests", "golden", "tcg_golden.json")


def _ref():
    if not os.path.exists(REF_SO):
        pytest.skip("reference tables not built here (no reference checkout)")
    lib = C.CDLL(REF_SO)
    lib.ref_locate_patch_low_R.restype = C.c_int
    lib.ref_locate_patch_low_R.argtypes = [C.c_double, C.c_double]
    lib.ref_locate_patch_high_R.restype = C.c_int
    lib.ref_locate_patch_high_R.argtypes = [C.c_double, C.c_double]
    lib.ref_bw_by_patch.restype = C.c_double
    lib.ref_bw_by_patch.argtypes = [C.c_int]
    lib.ref_c2s.restype = C.c_double
    lib.ref_c2s.argtypes = [C.c_int]
    lib.ref_c2s_ptr.restype = C.c_int
    lib.ref_c2s_ptr.argtypes = [C.c_int]
    return lib


def test_patch_grid_matches_reference_trees():
    """the 128 x 128 patch grids equal the reference's decision trees sampled
    as reference src/t_c_g0_n.cpp:53-64 does (X = i / 127)"""
    ref = _ref()
    lo = np.zeros(128 * 128, dtype=np.int32)
    hi = np.zeros(128 * 128, dtype=np.int32)
    ip = C.POINTER(C.c_int)
    oracle().oracle_patch_grid(lo.ctypes.data_as(ip), hi.ctypes.data_as(ip))
    bad = 0
    for i1 in range(128):
        x1 = i1 / 127.0
        for i2 in range(128):
            x2 = i2 / 127.0
            bad += lo[i1 * 128 + i2] != ref.ref_locate_patch_low_R(x1, x2)
            bad += hi[i1 * 128 + i2] != ref.ref_locate_patch_high_R(x1, x2)
    assert bad == 0


def test_patch_bias_width_matches_reference():
    ref = _ref()
    bw = np.zeros(207 * 4)
    oracle().oracle_patch_bw(as_dp(bw))
    want = np.array([ref.ref_bw_by_patch(i) for i in range(207 * 4)])
    assert np.array_equal(bw, want)


def _c2s():
    v = np.zeros(110)
    oracle().oracle_c2s(as_dp(v))
    off = [0, 1, 10, 40]
    return [v[off[l]: off[l] + (2 * l + 1) * (l + 1) * (l + 2) // 2].reshape(2 * l + 1, (l + 1) * (l + 2) // 2)
            for l in range(4)]


def test_c2s_matches_reference_literals():
    """bit for bit, including 0.5/sqrt(3.0) and the 9-digit l = 3 decimals"""
    ref = _ref()
    mats = _c2s()
    for l in range(4):
        base = ref.ref_c2s_ptr(l)
        want = np.array([ref.ref_c2s(base + i) for i in range(mats[l].size)]).reshape(mats[l].shape)
        assert np.array_equal(mats[l], want), l


def _cart(l):
    return [(lx, ly, l - lx - ly) for lx in range(l, -1, -1) for ly in range(l - lx, -1, -1)]


def test_c2s_solid_harmonics():
    """rows are CP2K's real solid harmonics, m = -l..l, as polynomials in the
    cartesian monomials; l <= 2 exact, l = 3 to the 9 digits kept"""
    s6, s10, s60 = math.sqrt(6.0), math.sqrt(10.0), math.sqrt(60.0)
    r3 = math.sqrt(3.0)
    exact = {
        0: [{(0, 0, 0): 1.0}],
        1: [{(0, 1, 0): 1.0}, {(0, 0, 1): 1.0}, {(1, 0, 0): 1.0}],
        2: [{(1, 1, 0): 1.0}, {(0, 1, 1): 1.0},
            {(2, 0, 0): -0.5 / r3, (0, 2, 0): -0.5 / r3, (0, 0, 2): 1.0 / r3},
            {(1, 0, 1): 1.0}, {(2, 0, 0): 0.5, (0, 2, 0): -0.5}],
        3: [{(2, 1, 0): 3 * s6 / 12, (0, 3, 0): -s6 / 12},
            {(1, 1, 1): 1.0},
            {(2, 1, 0): -s10 / 20, (0, 3, 0): -s10 / 20, (0, 1, 2): 4 * s10 / 20},
            {(2, 0, 1): -3 / s60, (0, 2, 1): -3 / s60, (0, 0, 3): 2 / s60},
            {(3, 0, 0): -s10 / 20, (1, 2, 0): -s10 / 20, (1, 0, 2): 4 * s10 / 20},
            {(2, 0, 1): 0.5, (0, 2, 1): -0.5},
            {(3, 0, 0): s6 / 12, (1, 2, 0): -3 * s6 / 12}],
    }
    mats = _c2s()
    for l in range(4):
        want = np.zeros(mats[l].shape)
        cl = _cart(l)
        for m, poly in enumerate(exact[l]):
            for mono, c in poly.items():
                want[m, cl.index(mono)] = c
        tol = 0.0 if l < 3 else 6e-10
        assert np.max(np.abs(mats[l] - want)) <= max(tol, 2.3e-16), l


def test_exp_within_two_ulp():
    lib = oracle()
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(-745, 709, 20000), rng.uniform(-1, 1, 5000), [0.0, -1e-300, 1e-300]])
    for v in x:
        got = lib.oracle_exp(float(v))
        want = math.exp(v)
        if want == 0.0:
            assert got == 0.0 or got < 5e-324 * 4
            continue
        assert abs(got - want) <= 2 * np.spacing(want), (v, got, want)


def _boys_mp(n, T):
    import mpmath as mp

    mp.mp.dps = 30
    return float(mp.hyp1f1(n + mp.mpf(1) / 2, n + mp.mpf(3) / 2, -mp.mpf(T)) / (2 * n + 1))


def test_boys_table_matches_series():
    """ftable[i][n] = F_n(0.1 i), built like reference src/fgamma.cpp:53-110:
    a series stopped when a term drops below DBL_EPSILON, so the error is
    absolute (a few 1e-16), not relative"""
    ft = np.zeros(121 * 28)
    oracle().oracle_boys_table(as_dp(ft))
    ft = ft.reshape(121, 28)
    for i in (0, 1, 7, 40, 77, 100, 120):
        for n in (0, 1, 5, 12, 20, 27):
            want = _boys_mp(n, 0.1 * i)
            assert abs(ft[i, n] - want) <= 1e-15, (i, n, ft[i, n], want)


@pytest.mark.parametrize("T", [0.0, 5e-14, 3e-4, 0.37, 2.05, 7.77, 11.81, 11.96, 12.0, 15.3, 29.0, 61.0, 300.0])
def test_boys_function(T):
    """F_n(T) of the reference formula (reference src/fgamma.cpp:196-243) against
    the series.  Below T = 12 the 6-term Taylor step about the table's grid
    point is exact to ~3e-13 relative, and the downward recursion multiplies
    the table's absolute error (<= 1e-15) by prod_k max(1, 2T/(2k+1)); above,
    the 4-term asymptotic form is good to ~1e-14.  A wrong coefficient or
    index moves the values by far more."""
    lib = oracle()
    f = np.zeros(13)
    for nmax in (0, 4, 8, 12):
        lib.oracle_boys(nmax, T, as_dp(f), 1.0)
        for n in range(nmax + 1):
            want = _boys_mp(n, T)
            if T < 1e-13:
                tol = 0.0
                want = 1.0 / (2 * n + 1)
            elif T < 12:
                amp = 1.0
                for k in range(n, nmax):
                    amp *= max(1.0, 2 * T / (2 * k + 1))
                tol = 3e-13 * want + 1e-15 * amp
            else:
                tol = 2e-14 * want
            assert abs(f[n] - want) <= tol, (T, nmax, n, f[n], want)


def test_truncated_coulomb_against_closed_form():
    """G_m(R,T), m = 0..12, from the patch tables against the closed form
    (mpmath, tests/golden/tcg_golden.json): within the fit error in the
    patch regions, the Boys function where the cut-off cannot matter, and
    exactly zero past the truncation cut, where the closed form is < 1e-12"""
    if not os.path.exists(GOLDEN):
        pytest.skip("golden file not generated")
    lib = oracle()
    h = lib.oracle_create_handle()
    c0, ld = tcg_table(21)
    lib.oracle_set_Potential_Truncated(h, 11.3384, as_dp(c0.ravel()), ld, c0.size)
    pts = json.load(open(GOLDEN))["points"]
    seen = set()
    G = np.zeros(13)
    for p in pts:
        br = lib.oracle_tcg(h, p["R"], p["T"], 12, as_dp(G))
        seen.add(br)
        want = np.array(p["G"])
        if br == 0:
            assert np.all(G == 0.0) and np.max(np.abs(want)) < 1e-12, p
        elif br == 1:
            assert np.max(np.abs(G - np.array(p["F"])) / np.array(p["F"])) < 3e-9, p
            assert np.max(np.abs(want - np.array(p["F"]))) < 1e-12, p
        else:
            assert np.max(np.abs(G - want)) < 3e-11, (p["R"], p["T"], np.max(np.abs(G - want)))
    lib.oracle_destroy_handle(h)
    assert seen == {0, 1, 2}


@pytest.mark.parametrize("l", [0, 1, 2, 3])
def test_norm_psi_by_quadrature(l):
    """compute_norm_psi (reference src/extern_functions.cpp:95-110) makes a
    contracted cartesian function whose exponents are all <= 1 (x^l for
    l <= 1, xy, xyz) have unit norm; checked by 1-D quadrature of the
    factorised overlap, independent of the closed form"""
    from scipy.integrate import quad

    lib = oracle()
    rng = np.random.default_rng(40 + l)
    Z = 10 ** rng.uniform(-1, 2, 4)
    K = rng.normal(size=4)
    N = lib.oracle_norm_psi(as_dp(K), as_dp(Z), 4, l)
    powers = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (1, 1, 1)][l]
    S = 0.0
    for i in range(4):
        for j in range(4):
            a = Z[i] + Z[j]
            term = K[i] * K[j]
            for p in powers:
                term *= quad(lambda x: x ** (2 * p) * math.exp(-a * x * x), -np.inf, np.inf, epsabs=0, epsrel=1e-13)[0]
            S += term
    assert abs(S / N ** 2 - 1.0) < 1e-10
