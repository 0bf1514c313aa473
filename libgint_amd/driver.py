"""Python side of the caller-side driver (libgint_amd/libgintdrv.so).

Builds the synthetic systems of the workloads, records their screened
libgint_* call streams, and replays a stream into any library that exports
the libgint_* protocol: the product (libgint.so) or, in the tests, the CPU
oracle (same call protocol, oracle_* names).
"""
import ctypes as C
import os

import numpy as np

from . import DRV_PATH, TCG_PATH, load

_dp = C.POINTER(C.c_double)
_vp = C.c_void_p
_i64p = C.POINTER(C.c_int64)


class Params(C.Structure):
    _fields_ = [
        ("n_water", C.c_int), ("n_tm", C.c_int), ("box_A", C.c_double), ("image_shells", C.c_int),
        ("basis", C.c_int), ("np_min", C.c_int), ("np_max", C.c_int), ("seed", C.c_uint64),
        ("rcut_A", C.c_double), ("eps", C.c_double), ("min_oo_A", C.c_double), ("symmetric", C.c_int),
        ("cell_A", C.c_double), ("np_tm", C.c_int), ("max_cells", C.c_int), ("ks_mix", C.c_int),
        ("zmin", C.c_double), ("zmax", C.c_double), ("shift_atom", C.c_int), ("shift_cells", C.c_int * 3), ("replicate", C.c_int),
    ]


class AbiTable(C.Structure):
    _fields_ = [(n, _vp) for n in ("add_prm", "add_shell", "add_qrt", "add_qrtt", "add_set", "set_Atom", "set_Atom_L",
                                   "add_set_blocks", "add_tra")]


# Workloads.  Config 0 has one cell; its cell edge (100 A) only has to make
# the minimum image the identity for molecules placed in a 10 A cube.
CONFIGS = {
    0: dict(n_water=8, n_tm=0, box_A=10.0, image_shells=0, basis=0, np_min=1, np_max=6, seed=0),
    1: dict(n_water=32, n_tm=0, box_A=12.4, image_shells=1, basis=0, np_min=1, np_max=8, seed=1),
    2: dict(n_water=64, n_tm=8, box_A=15.6, image_shells=2, basis=0, np_min=1, np_max=8, seed=2),
    4: dict(n_water=128, n_tm=0, box_A=15.6, image_shells=3, basis=4, np_min=1, np_max=16, seed=4),
}

_drv = None


def drv():
    global _drv
    if _drv is None:
        if not os.path.exists(DRV_PATH):
            raise RuntimeError("libgint_amd/libgintdrv.so is not built: run __graft_entry__.build()")
        d = C.CDLL(DRV_PATH)
        sig = {
            "gdrv_system_new": (_vp, [C.POINTER(Params)]),
            "gdrv_system_free": (None, [_vp]),
            "gdrv_system_info": (None, [_vp, _i64p]),
            "gdrv_system_P": (None, [_vp, _dp]),
            "gdrv_system_cell": (None, [_vp, _dp, _dp]),
            "gdrv_system_neighs": (None, [_vp, _dp]),
            "gdrv_system_rcut": (C.c_double, [_vp]),
            "gdrv_system_atoms": (None, [_vp, _dp, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
            "gdrv_kind_basis": (C.c_int, [_vp, C.c_int, _dp, C.POINTER(C.c_int), _dp]),
            "gdrv_emit_basis": (None, [_vp, C.POINTER(AbiTable), _vp]),
            "gdrv_stream_record": (_vp, [_vp, C.c_int, C.c_int64]),
            "gdrv_stream_info": (None, [_vp, _i64p]),
            "gdrv_stream_free": (None, [_vp]),
            "gdrv_stream_replay": (None, [_vp, _vp, C.POINTER(AbiTable), _vp, C.c_int64, C.c_int64, C.c_int64,
                                          C.c_int]),
            "gdrv_stream_sets": (None, [_vp, C.POINTER(C.c_int32)]),
            "gdrv_stream_replay_tra": (None, [_vp, _vp, C.POINTER(AbiTable), _vp, C.c_int64, C.c_int64, C.c_int64,
                                              C.POINTER(C.c_uint64)]),
            "gdrv_team_kbuild": (C.c_int, [_vp, _vp, C.POINTER(AbiTable), C.POINTER(TeamAbi), C.c_int, C.c_int,
                                           C.c_int, _dp, C.c_int, C.c_int, C.c_int, _dp, _dp]),
            "gdrv_symmetrize": (None, [_vp, _dp]),
            "gdrv_dense": (None, [_vp, _dp, _dp]),
        }
        for n, (r, a) in sig.items():
            f = getattr(d, n)
            f.restype = r
            f.argtypes = a
        _drv = d
    return _drv


def _ptr(a):
    return a.ctypes.data_as(_dp)


def abi_table(lib, prefix="libgint_"):
    """the call-protocol entry points of `lib` (product or oracle); the
    set-level extension is libgint_amd_add_set_blocks / oracle_add_set_blocks"""
    t = AbiTable()
    for name, _ in AbiTable._fields_:
        own = name in ("add_set_blocks", "add_tra")   # this library's extensions
        full = ("libgint_amd_" + name) if (own and prefix == "libgint_") else prefix + name
        setattr(t, name, C.cast(getattr(lib, full), _vp).value)
    return t


def tcg_table(nder=21):
    """truncated-Coulomb coefficient table (207 x (nder+1)*105) and its ld"""
    full = np.fromfile(TCG_PATH, dtype="<f8").reshape(207, 22 * 105)
    c0 = np.ascontiguousarray(full[:, : (nder + 1) * 105])
    return c0, (nder + 1) * 105


class System:
    def __init__(self, config=None, **kw):
        p = dict(n_water=1, n_tm=0, box_A=10.0, image_shells=0, basis=0, np_min=1, np_max=4, seed=0,
                 rcut_A=6.0, eps=1e-11, min_oo_A=2.4, symmetric=1, cell_A=0.0, np_tm=16, max_cells=0, ks_mix=0,
                 zmin=0.1, zmax=5e3, shift_atom=-1, shift_cells=(0, 0, 0), replicate=1)
        if config is not None:
            p.update(CONFIGS[config])
        p.update(kw)
        self.params = dict(p)
        self.periodic = p["image_shells"] > 0
        if not self.periodic and p["cell_A"] == 0.0:
            # molecules in a box_A cube, cell edge 100 A so the minimum image is the identity
            p["cell_A"] = 100.0
        self.params = dict(p)
        vals = {k: p[k] for k, _ in Params._fields_}
        vals["shift_cells"] = (C.c_int * 3)(*p["shift_cells"])
        self._p = Params(**vals)
        self.h = drv().gdrv_system_new(C.byref(self._p))
        info = (C.c_int64 * 8)()
        drv().gdrv_system_info(self.h, info)
        self.natom, self.nbf, self.P_size, self.Ng = info[0], info[1], info[2], info[3]
        self.np_kinds = (info[5], info[6], info[7])

    def describe(self):
        """the workload in words, with its sizes"""
        p = self.params
        basis = {0: "O 3s2p1d / H 2s1p", 1: "O 2s1p / H 1s", 4: "O 4s3p2d1f / H 3s2p1d"}[p["basis"]]
        tm = f" + {p['n_tm']} spdf centres" if p["n_tm"] else ""
        box = f"{p['box_A']} A periodic box" if self.periodic else f"{p['box_A']} A cluster"
        return (f"{p['n_water']} H2O{tm} in a {box}, {self.Ng} cells, {basis}, "
                f"truncated Coulomb R_c {p['rcut_A']} A, {self.nbf} basis functions")

    def __del__(self):
        try:
            drv().gdrv_system_free(self.h)
        except Exception:
            pass

    def P(self):
        out = np.zeros(self.P_size)
        drv().gdrv_system_P(self.h, _ptr(out))
        return out

    def cell(self):
        h = np.zeros(9)
        hi = np.zeros(9)
        drv().gdrv_system_cell(self.h, _ptr(h), _ptr(hi))
        return h, hi

    def neighs(self):
        g = np.zeros(3 * self.Ng)
        drv().gdrv_system_neighs(self.h, _ptr(g))
        return g

    def rcut(self):
        return drv().gdrv_system_rcut(self.h)

    def record(self, nthreads=0, sample=1):
        """the screened call stream; sample > 1 keeps the set quartets
        (i,j,k,l) whose index hash is a multiple of `sample` (about 1 in
        `sample`, the same ones on every machine)"""
        return Stream(self, drv().gdrv_stream_record(self.h, nthreads, sample))

    def symmetrize(self, K):
        K = np.ascontiguousarray(K, dtype=np.float64).copy()
        drv().gdrv_symmetrize(self.h, _ptr(K))
        return K

    def dense(self, M):
        D = np.zeros((self.nbf, self.nbf))
        drv().gdrv_dense(self.h, _ptr(np.ascontiguousarray(M)), _ptr(D))
        return D

    def atoms(self):
        R = np.zeros(3 * self.natom)
        kind = np.zeros(self.natom, dtype=np.int32)
        f0 = np.zeros(self.natom, dtype=np.int32)
        drv().gdrv_system_atoms(self.h, _ptr(R), kind.ctypes.data_as(C.POINTER(C.c_int)), f0.ctypes.data_as(C.POINTER(C.c_int)))
        return R.reshape(-1, 3), kind, f0

    def kind_basis(self, kind):
        Z = np.zeros(16)
        lnl = np.zeros(16, dtype=np.int32)
        K = np.zeros(16 * 4 * 8)
        ns = drv().gdrv_kind_basis(self.h, kind, _ptr(Z), lnl.ctypes.data_as(C.POINTER(C.c_int)), _ptr(K))
        np_ = self.np_kinds[kind]
        shells = []
        off = 0
        for s in range(ns):
            l, nl = int(lnl[2 * s]), int(lnl[2 * s + 1])
            shells.append((l, nl, K[off: off + nl * np_].reshape(nl, np_).copy()))
            off += nl * np_
        return Z[:np_].copy(), shells


def tra_size(system, sets):
    """output doubles of the integral-output replay of the set quartets `sets`
    (rows of (i, j, k, l)): every class of each, all contractions"""
    _, kind, _ = system.atoms()
    shells = {k: system.kind_basis(k)[1] for k in set(kind.tolist())}
    n = 0
    for q in sets:
        per = [[(l, nl) for (l, nl, _) in shells[int(kind[a])]] for a in q]
        for (la, na) in per[0]:
            for (lb, nb) in per[1]:
                for (lc, nc) in per[2]:
                    for (ld, nd) in per[3]:
                        n += na * nb * nc * nd * (2 * la + 1) * (2 * lb + 1) * (2 * lc + 1) * (2 * ld + 1)
    return n


class Stream:
    def __init__(self, system, handle):
        self.system = system
        self.h = handle
        info = (C.c_int64 * 5)()
        drv().gdrv_stream_info(handle, info)
        self.words, self.n_sets, self.n_rows, self.n_prm, self.n_qrtt = list(info)

    def __del__(self):
        try:
            drv().gdrv_stream_free(self.h)
        except Exception:
            pass

    def sets(self):
        """(i, j, k, l) of every set quartet, in stream order"""
        out = np.zeros((self.n_sets, 4), dtype=np.int32)
        drv().gdrv_stream_sets(self.h, out.ctypes.data_as(C.POINTER(C.c_int32)))
        return out

    def replay_tra(self, table, handle, begin=0, end=None, stride=1, dest=0):
        """the set quartets' integrals of every class into the output array
        (libgint_amd_add_tra), class after class from `dest`; returns the end"""
        if end is None:
            end = self.n_sets
        d = C.c_uint64(dest)
        drv().gdrv_stream_replay_tra(self.system.h, self.h, C.byref(table), handle, begin, end, stride, C.byref(d))
        return d.value

    def replay(self, table, handle, begin=0, end=None, stride=1, blocks=False):
        """set quartets [begin, end) with stride; blocks=True sends each set
        quartet's K layout as one add_set_blocks call instead of its
        add_qrt / add_qrtt calls"""
        if end is None:
            end = self.n_sets
        drv().gdrv_stream_replay(self.system.h, self.h, C.byref(table), handle, begin, end, stride, 1 if blocks else 0)


def setup_handle(lib, prefix, system, hf_fac=1.0, max_mem_mib=4096, truncated=True, P=None, Pb=None,
                 stats=False, timing=False, debug=0, c0=None):
    """create + configure a handle of `lib` (product or oracle) for `system`;
    c0 = (table, ld_C0), e.g. from a t_c_g.dat via libgint_amd.read_c0, in
    place of the generated coefficient table"""
    f = lambda n: getattr(lib, prefix + n)
    h = f("create_handle")()
    f("init")(h)
    if stats and prefix == "libgint_":
        lib.libgint_amd_enable_stats(h, 1)
    if timing and prefix == "libgint_":
        lib.libgint_amd_enable_timing(h, 1)
    if debug and prefix == "libgint_":
        lib.libgint_amd_set_debug(h, debug)
    if truncated:
        c0, ld = tcg_table() if c0 is None else c0
        c0 = np.ascontiguousarray(c0, dtype=np.float64)
        f("set_Potential_Truncated")(h, system.rcut(), _ptr(c0.ravel()), ld, c0.size)
    f("set_hf_fac")(h, hf_fac)
    f("set_max_mem")(h, max_mem_mib)
    hm, hi = system.cell()
    f("set_cell")(h, system.periodic, _ptr(hm), _ptr(hi))
    g = system.neighs()
    f("set_neighs")(h, _ptr(g), system.Ng)
    table = abi_table(lib, prefix)
    drv().gdrv_emit_basis(system.h, C.byref(table), h)
    if P is None:
        P = system.P()
    if Pb is None:
        f("set_P")(h, _ptr(P), len(P))
    else:
        f("set_P_polarized")(h, _ptr(P), _ptr(Pb), len(P))
    return h, table


def kbuild(lib, prefix, system, stream, begin=0, end=None, stride=1, K0=None, plan=False, blocks=False, **kw):
    """one K-build of (a sample of) the stream; returns K in block layout
    (as accumulated by the library, before symmetrisation) and the handle.
    With Pb in kw the build is spin-polarised and K is (Ka, Kb).  With
    plan=True the stream is recorded as a resident plan and executed."""
    h, table = setup_handle(lib, prefix, system, **kw)
    polar = kw.get("Pb") is not None
    if K0 is not None:
        if polar:
            getattr(lib, prefix + "set_K_polarized")(h, _ptr(K0[0]), _ptr(K0[1]), system.P_size)
        else:
            getattr(lib, prefix + "set_K")(h, _ptr(K0), system.P_size)
    if plan:
        lib.libgint_amd_plan_begin(h)
        stream.replay(table, h, begin, end, stride, blocks)
        lib.libgint_amd_plan_end(h)
        lib.libgint_amd_plan_execute(h)
    else:
        stream.replay(table, h, begin, end, stride, blocks)
    if polar:
        Ka = np.zeros(system.P_size)
        Kb = np.zeros(system.P_size)
        getattr(lib, prefix + "get_K_polarized")(h, _ptr(Ka), _ptr(Kb))
        return (Ka, Kb), h
    K = np.zeros(system.P_size)
    getattr(lib, prefix + "get_K")(h, _ptr(K))
    return K, h


class TeamAbi(C.Structure):
    _fields_ = [(n, _vp) for n in ("create_handle", "init", "set_Potential_Truncated", "set_hf_fac", "set_max_mem",
                                   "set_cell", "set_neighs", "set_P", "get_K", "destroy_handle", "plan_begin",
                                   "plan_end", "plan_execute")]


def team_kbuild(lib, system, stream, nthreads, blocks=False, plan=False, truncated=True, max_mem_mib=4096):
    """one K-build by an OpenMP team of `nthreads` threads, one handle each,
    collective init / set_P / get_K, thread t sending set quartets
    t, t + nthreads, ...  (the product library only); returns K and
    (ingestion seconds of the slowest thread, K-build seconds)"""
    t = TeamAbi()
    own = {"destroy_handle", "plan_begin", "plan_end", "plan_execute"}
    for name, _ in TeamAbi._fields_:
        full = ("libgint_amd_" if name in own else "libgint_") + name
        setattr(t, name, C.cast(getattr(lib, full), _vp).value)
    table = abi_table(lib)
    K = np.zeros(system.P_size)
    secs = np.zeros(2)
    c0 = ld = None
    if truncated:
        c0, ld = tcg_table()
        c0 = np.ascontiguousarray(c0)
    rc = drv().gdrv_team_kbuild(system.h, stream.h, C.byref(table), C.byref(t), nthreads, 1 if blocks else 0,
                                1 if plan else 0, _ptr(c0.ravel()) if truncated else None, ld or 0,
                                c0.size if truncated else 0, max_mem_mib, _ptr(K), _ptr(secs))
    if rc != 0:
        raise RuntimeError("gdrv_team_kbuild failed: %d" % rc)
    return K, (float(secs[0]), float(secs[1]))


def integrals(lib, prefix, system, stream, begin=0, end=None, stride=1, **kw):
    """the spherical integrals of every class of the sampled set quartets,
    through the integral-output extension, in the reference's TRA layout,
    class after class; returns (array, handle)"""
    h, table = setup_handle(lib, prefix, system, **kw)
    if end is None:
        end = stream.n_sets
    n = tra_size(system, stream.sets()[begin:end:stride])
    out = np.zeros(max(n, 1))
    own = (lambda x: getattr(lib, "libgint_amd_" + x)) if prefix == "libgint_" else (lambda x: getattr(lib, prefix + x))
    own("set_output")(h, _ptr(out), n)
    got = stream.replay_tra(table, h, begin, end, stride)
    assert got == n, (got, n)
    own("get_output")(h)
    return out[:n], h


def stats(lib, prefix, h):
    out = (C.c_uint64 * 8)()
    if prefix == "libgint_":
        lib.libgint_amd_stats(h, out)
    else:
        lib.oracle_stats(h, out)
    return list(out)


def product():
    return load()


# ---------------------------------------------------------------------------
# serving workload: independent periodic systems sending K-build requests to
# one device-resident request loop (libgint_amd_server_*)
# ---------------------------------------------------------------------------
WATER_VOLUME_A3 = 29.9   # A^3 per water molecule at 1 g/cm^3


def serving_systems(n_systems=16, atoms=(20, 200), seed=3, image_shells=1, np_max=3, zmin=0.6):
    """n_systems periodic water-like systems, sizes log-uniform in `atoms`,
    each in a cubic box at liquid density, O 3s2p1d / H 2s1p with 1..np_max
    primitives (exponents log-uniform in [zmin, 5e3])"""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_systems):
        a = float(np.exp(rng.uniform(np.log(atoms[0]), np.log(atoms[1]))))
        nw = max(1, int(round(a / 3)))
        box = (nw * WATER_VOLUME_A3) ** (1.0 / 3.0)
        out.append(System(n_water=nw, box_A=box, image_shells=image_shells, basis=0, np_min=1, np_max=np_max,
                          zmin=zmin, seed=1000 * seed + i))
    return out


class ServingClient:
    """one system as a client of the server: its plan recorded into a handle
    of its own, its SCF-like density sequence"""

    def __init__(self, lib, system, steps=10, seed=0):
        self.s = system
        self.st = system.record()
        self.P0 = system.P()
        self.h, table = setup_handle(lib, "libgint_", system, P=self.P0, max_mem_mib=1024)
        lib.libgint_amd_plan_begin(self.h)
        self.st.replay(table, self.h, 0, self.st.n_sets, 1, True)
        lib.libgint_amd_plan_end(self.h)
        self.rng = np.random.default_rng(seed)
        self.steps = steps
        self.k = 0
        self.P = self.P0.copy()

    def next_density(self):
        """the density of the next SCF-like step: the previous one plus
        1e-3 N(0,1), back to the start after `steps` steps"""
        if self.k % self.steps == 0:
            self.P = self.P0.copy()
        else:
            self.P = self.P + 1e-3 * self.rng.standard_normal(self.P.size)
        self.k += 1
        return self.P.copy()


def serve(lib, clients, n_requests, rate, slots, max_mem_mib, seed=7, keep=False, resident_clients=None,
          progress=None, progress_s=10.0):
    """Poisson arrivals at `rate` requests/s, each to a client drawn at
    random, through one server; returns per-request (client, P, K or None,
    rc, latency, queueing, execution) and the server statistics.  With
    resident_clients = n the budget is set to hold the n largest clients.
    progress: a function of one string, called every progress_s seconds
    with the server's state (queued, running, done, refused, resident and
    busy clients, resident MiB) while requests are outstanding"""
    import threading
    import time

    srv = lib.libgint_amd_server_create(max_mem_mib, slots)
    ids = [lib.libgint_amd_server_add_client(srv, c.h, c.s.P_size) for c in clients]
    if resident_clients:
        sizes = sorted((lib.libgint_amd_server_client_bytes(srv, i) for i in ids), reverse=True)
        lib.libgint_amd_server_set_budget(srv, int(sum(sizes[:resident_clients]) / 2**20) + 1)
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, n_requests)
    who = rng.integers(0, len(clients), n_requests)
    reqs = [None] * n_requests
    done = [None] * n_requests

    def waiter():
        out = np.zeros(3)
        for i in range(n_requests):
            while reqs[i] is None:
                time.sleep(0.0005)
            rid, c, P, K = reqs[i]
            rc = lib.libgint_amd_server_wait(srv, rid, _ptr(out))
            done[i] = (c, P if keep else None, K if keep else None, rc, float(out[0]), float(out[1]), float(out[2]))
            if not keep:
                reqs[i] = (rid, c, None, None)   # the server is done with P and K

    th = threading.Thread(target=waiter)
    th.start()
    finished = threading.Event()

    def watch():
        state = (C.c_int64 * 8)()
        while not finished.wait(progress_s):
            lib.libgint_amd_server_state(srv, state)
            progress(f"server: queued {state[0]} running {state[1]} done {state[2]} refused {state[3]}; "
                     f"resident clients {state[4]} busy {state[5]}; resident {state[6] / 2**20:.0f} of "
                     f"{state[7] / 2**20:.0f} MiB")

    wt = threading.Thread(target=watch, daemon=True) if progress else None
    if wt:
        wt.start()
    t0 = time.perf_counter()
    t_next = 0.0
    for i in range(n_requests):
        t_next += gaps[i]
        c = int(who[i])
        P = clients[c].next_density()
        K = np.zeros(clients[c].s.P_size)
        while time.perf_counter() - t0 < t_next:
            time.sleep(0.0002)
        rid = lib.libgint_amd_server_submit(srv, ids[c], _ptr(P), _ptr(K))
        reqs[i] = (rid, c, P, K)
    th.join()
    finished.set()
    if wt:
        wt.join()
    st = np.zeros(8)
    lib.libgint_amd_server_stats(srv, _ptr(st))
    lib.libgint_amd_server_destroy(srv)
    return done, st
