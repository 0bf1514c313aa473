"""libgint_amd_plan_execute_noalloc, the entry point that allocates nothing:
the resident plan run into caller-owned device P / K with every scratch
buffer in the library's static device array.  Checked against the
allocating entry point (set_P / libgint_amd_plan_execute / get_K) to the
parity bar, on two calls in a row, the second with a smaller input, and
for its size checks (rejected with an error code, nothing launched)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
KTOL_ABS, KTOL_REL = 1e-10, 1e-11


class Hip:
    """the caller's own device arrays, through the HIP runtime the library
    links (the caller allocates; the entry point does not)"""

    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]

    def alloc(self, n):
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), 8 * n) == 0
        return p

    def put(self, dst, a):
        assert self.rt.hipMemcpy(dst, a.ctypes.data, 8 * a.size, 1) == 0   # host to device

    def get(self, src, n):
        a = np.zeros(n)
        assert self.rt.hipMemcpy(a.ctypes.data, src, 8 * n, 2) == 0         # device to host
        return a

    def zero(self, p, n):
        assert self.rt.hipMemset(p, 0, 8 * n) == 0

    def free(self, p):
        assert self.rt.hipFree(p) == 0


def _planned(lib, D, sysargs):
    s = D.System(**sysargs)
    st = s.record()
    P = s.P()
    h, table = D.setup_handle(lib, "libgint_", s, P=P)
    lib.libgint_amd_plan_begin(h)
    st.replay(table, h)
    lib.libgint_amd_plan_end(h)
    return s, st, P, h


def _allocating(lib, D, s, h, P):
    lib.libgint_set_P(h, D._ptr(P), s.P_size)
    lib.libgint_amd_plan_execute(h)
    K = np.zeros(s.P_size)
    lib.libgint_get_K(h, D._ptr(K))
    return K


def _noalloc(lib, hip, s, h, P):
    n = s.P_size
    dP, dK = hip.alloc(n), hip.alloc(n)
    hip.put(dP, P)
    hip.zero(dK, n)
    rc = lib.libgint_amd_plan_execute_noalloc(h, dP, None, dK, None, n)
    lib.libgint_amd_synchronize(h)
    K = hip.get(dK, n)
    hip.free(dP)
    hip.free(dK)
    return rc, K


def test_noalloc_matches_allocating_twice():
    from libgint_amd import driver as D, load

    lib = load()
    hip = Hip()
    cases = [
        ("water1_pbc27", dict(n_water=1, box_A=6.0, image_shells=1, np_max=3, seed=2)),
        ("water2_pbc27_fewer_primitive_quartets", dict(n_water=2, box_A=7.0, image_shells=1, np_max=3, seed=12)),
    ]
    sizes = []
    for name, sysargs in cases:
        s, st, P, h = _planned(lib, D, sysargs)
        print(name, "sets", st.n_sets, "primitive quartets", st.n_prm, "P size", s.P_size, flush=True)
        Kref = _allocating(lib, D, s, h, P)
        rc, K = _noalloc(lib, hip, s, h, P)
        assert rc == 0, (name, rc)
        err = np.max(np.abs(K - Kref))
        tol = KTOL_ABS + KTOL_REL * np.max(np.abs(Kref))
        print(name, "max|K|", np.max(np.abs(Kref)), "max|dK| vs allocating path", err, flush=True)
        assert np.max(np.abs(Kref)) > 0 and err <= tol, (name, err, tol)
        sizes.append(st.n_prm)
        lib.libgint_amd_destroy_handle(h)
    assert sizes[1] < sizes[0]


def test_noalloc_spin_polarised_adds_into_K():
    """two spins, and K added to (not zeroed): K starts from given values"""
    from libgint_amd import driver as D, load

    lib = load()
    hip = Hip()
    s, st, P, h = _planned(lib, D, dict(n_water=1, n_tm=1, box_A=6.0, image_shells=1, np_max=2, np_tm=2, seed=24))
    Pb = 0.5 * P
    rng = np.random.default_rng(5)
    K0 = (rng.normal(size=s.P_size), rng.normal(size=s.P_size))
    lib.libgint_set_P_polarized(h, D._ptr(P), D._ptr(Pb), s.P_size)
    lib.libgint_set_K_polarized(h, D._ptr(K0[0]), D._ptr(K0[1]), s.P_size)
    lib.libgint_amd_plan_execute(h)
    Ka, Kb = np.zeros(s.P_size), np.zeros(s.P_size)
    lib.libgint_get_K_polarized(h, D._ptr(Ka), D._ptr(Kb))
    n = s.P_size
    d = [hip.alloc(n) for _ in range(4)]
    hip.put(d[0], P)
    hip.put(d[1], Pb)
    hip.put(d[2], K0[0])
    hip.put(d[3], K0[1])
    rc = lib.libgint_amd_plan_execute_noalloc(h, d[0], d[1], d[2], d[3], n)
    lib.libgint_amd_synchronize(h)
    assert rc == 0
    for ref, dk in ((Ka, d[2]), (Kb, d[3])):
        got = hip.get(dk, n)
        err = np.max(np.abs(got - ref))
        print("spin max|dK|", err, flush=True)
        assert err <= KTOL_ABS + KTOL_REL * np.max(np.abs(ref))
    for p in d:
        hip.free(p)
    lib.libgint_amd_destroy_handle(h)


def test_noalloc_rejects_sizes_beyond_its_scratch():
    """a set quartet whose primitive blocks exceed the static array's block
    region (an spdf centre with 16 primitives: 16^4 primitive quartets of
    7056-double blocks), and arrays shorter than the plan's offsets: error
    codes, nothing launched; a valid call afterwards still matches"""
    from libgint_amd import driver as D, load

    lib = load()
    hip = Hip()
    s, st, P, h = _planned(lib, D, dict(n_water=1, n_tm=1, box_A=5.0, image_shells=0, np_max=2, np_tm=16, seed=3))
    info = (C.c_uint64 * 6)()
    lib.libgint_amd_plan_info(h, info)
    print("large set quartets: primitive quartets", st.n_prm, "resident bytes", info[0], flush=True)
    n = s.P_size
    dP, dK = hip.alloc(n), hip.alloc(n)
    assert lib.libgint_amd_plan_execute_noalloc(h, dP, None, dK, None, n) == -4
    lib.libgint_amd_destroy_handle(h)
    s2, st2, P2, h2 = _planned(lib, D, dict(n_water=1, box_A=4.0, image_shells=0, np_max=3, seed=8))
    assert lib.libgint_amd_plan_execute_noalloc(h2, dP, None, dK, None, 1) == -3
    assert lib.libgint_amd_plan_execute_noalloc(h2, None, None, dK, None, s2.P_size) == -2
    Kref = _allocating(lib, D, s2, h2, P2)
    rc, K = _noalloc(lib, hip, s2, h2, P2)
    assert rc == 0
    assert np.max(np.abs(K - Kref)) <= KTOL_ABS + KTOL_REL * np.max(np.abs(Kref))
    hip.free(dP)
    hip.free(dK)
    lib.libgint_amd_destroy_handle(h2)
