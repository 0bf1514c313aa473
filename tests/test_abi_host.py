"""Host-side checks of the C-ABI (no GPU call is made): every declared entry
point is exported, the reference's entry points are all there, the
normalisation formula, the host engine's ingestion of a call stream and
its compression of add_qrtt calls, and the reference's Fortran binding
compiling and linking against the library."""
import ctypes as C
import math
import os
import re
import shutil
import sys
import subprocess

import numpy as np
import pytest

from helpers import ROOT

HEADER = os.path.join(ROOT, "include", "libGint.h")
REF_SRC = "/root/reference/src"


def _declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return re.findall(r"^\s*(?:void|double|int64_t|int)\s*\*?\s*(\w+)\s*\(", txt, flags=re.M)


def test_exports_every_declared_symbol():
    from libgint_amd import ABI, LIB_PATH

    lib = C.CDLL(LIB_PATH)
    names = _declared()
    assert len(names) >= 24
    for n in names:
        assert hasattr(lib, n), n
        assert n in ABI, n


def test_reference_entry_points_present():
    """the 24 functions of reference src/extern_functions.cpp, same names"""
    src = os.path.join(REF_SRC, "extern_functions.cpp")
    if not os.path.exists(src):
        pytest.skip("no reference checkout")
    txt = open(src).read()
    ref = set(re.findall(r"^\s*(?:void\s*\*|void|double)\s+(\w+)\s*\(", txt, flags=re.M))
    ours = set(_declared())
    assert ref and ref <= ours, sorted(ref - ours)
    assert len(ref) == 24


def test_compute_norm_psi_closed_form():
    """one primitive: N = sqrt(pi^(3/2) / 2^l / (2 z)^(3/2 + l))"""
    from libgint_amd import load

    lib = load()
    dp = C.POINTER(C.c_double)
    for l in range(4):
        for z in (0.1, 1.7, 250.0):
            K = np.array([1.0])
            Z = np.array([z])
            got = lib.compute_norm_psi(K.ctypes.data_as(dp), Z.ctypes.data_as(dp), 1, l)
            want = math.sqrt(math.pi ** 1.5 / 2 ** l / (2 * z) ** (1.5 + l))
            assert abs(got - want) <= 1e-15 * want


def _queue(lib, h):
    out = (C.c_uint64 * 8)()
    lib.libgint_amd_queue_info(h, out)
    return list(out)


def test_ingestion_counts_and_affine_groups():
    """a recorded stream lands as one row per add_shell, one group per set
    quartet; the driver's CP2K-like block offsets compress completely"""
    from libgint_amd import driver as D, load

    lib = load()
    s = D.System(0)
    st = s.record()
    h = lib.libgint_create_handle()
    lib.libgint_set_max_mem(h, 65536)
    hm, hi = s.cell()
    lib.libgint_set_cell(h, False, D._ptr(hm), D._ptr(hi))
    g = s.neighs()
    lib.libgint_set_neighs(h, D._ptr(g), s.Ng)
    table = D.abi_table(lib)
    D.drv().gdrv_emit_basis(s.h, C.byref(table), h)
    lib.libgint_amd_plan_begin(h)   # recording: nothing reaches the device
    st.replay(table, h)
    q = _queue(lib, h)
    assert q[1] == st.n_sets and q[2] == st.n_rows and q[3] == st.n_prm
    assert q[0] >= 1 and q[4] > 0 and q[5] == 0
    lib.libgint_amd_destroy_handle(h)


def test_non_affine_calls_kept_explicit():
    from libgint_amd import load

    lib = load()
    dp = C.POINTER(C.c_double)
    h = lib.libgint_create_handle()
    R = np.zeros(3)
    Z = np.array([1.0, 0.3])
    K = np.array([0.7, 0.3, 0.2, 0.9])
    lib.libgint_set_Atom(h, 0, R.ctypes.data_as(dp), Z.ctypes.data_as(dp), 2)
    lib.libgint_set_Atom_L(h, 0, 0, 2, K.ctypes.data_as(dp))
    lib.libgint_amd_plan_begin(h)

    def one_set(offsets):
        lib.libgint_add_prm(h, 0, 0, 0, 0)
        lib.libgint_add_shell(h, 0, 0, 0, 0, 0, 0)
        lib.libgint_add_qrt(h, 0, 0, 0, 0, 2, 2, 2, 2)
        for l in range(16):
            a, b, c, d = l >> 3, (l >> 2) & 1, (l >> 1) & 1, l & 1
            lib.libgint_add_qrtt(h, 0.25, 0, 0, 0, 0, a, b, c, d, 4, 4, 4, 4, *offsets(a, b, c, d),
                                 False, False, False, False)
        lib.libgint_add_set(h)

    one_set(lambda a, b, c, d: (a + 2 * c, a + 2 * d, b + 2 * c, b + 2 * d))       # affine
    one_set(lambda a, b, c, d: (a * c, a + 2 * d, b + 2 * c, b + 2 * d))          # not affine in (a, c)
    q = _queue(lib, h)
    assert q[1] == 2 and q[4] == 2 and q[5] == 1
    lib.libgint_amd_destroy_handle(h)


def test_fortran_binding_links(tmp_path):
    """the reference's Fortran module (src/interface_libgint.F90, compiled
    where it lies) and a caller of it link against libgint.so"""
    from libgint_amd import LIB_PATH

    fc = shutil.which("amdflang") or "/opt/rocm/llvm/bin/amdflang"
    mod = os.path.join(REF_SRC, "interface_libgint.F90")
    if not os.path.exists(mod) or not os.path.exists(fc):
        pytest.skip("no reference checkout or no amdflang")
    drive = os.path.join(ROOT, "tests", "fortran", "drive_libgint.F90")
    out = tmp_path / "drive"
    cmd = [fc, "-O1", "-J", str(tmp_path), "-o", str(out), mod, drive, LIB_PATH,
           "-Wl,-rpath," + os.path.dirname(LIB_PATH)]
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 0, r.stderr[-2000:]
    nm = subprocess.run(["nm", "-u", str(out)], capture_output=True, text=True).stdout
    used = set(re.findall(r"\b(libgint_\w+)", nm))
    assert len(used) >= 20, used


def test_synthetic_cell_list_cap():
    """max_cells keeps the nearest lattice vectors of the full list, in
    order: 27 = the 3x3x3 block, 33 = that block plus the six |g|^2 = 4 cells"""
    from libgint_amd import driver as D

    full = D.System(n_water=1, box_A=6.0, image_shells=2, basis=1, np_max=2, seed=1)
    gf = full.neighs().reshape(-1, 3)
    assert full.Ng == 125
    edge = np.max(np.abs(gf))
    for mc in (27, 32, 33):
        s = D.System(n_water=1, box_A=6.0, image_shells=2, basis=1, np_max=2, seed=1, max_cells=mc)
        g = s.neighs().reshape(-1, 3)
        assert s.Ng == mc
        np.testing.assert_array_equal(g, gf[:mc])
    r2 = np.rint(np.sum((gf / (edge / 2)) ** 2, axis=1)).astype(int)
    assert np.all(np.diff(r2) >= 0)
    assert np.all(r2[:27] <= 3) and np.all(r2[27:33] == 4)


def _recorded_queue(lib, s, st, blocks, stride=1):
    from libgint_amd import driver as D

    h = lib.libgint_create_handle()
    lib.libgint_set_max_mem(h, 65536)
    hm, hi = s.cell()
    lib.libgint_set_cell(h, s.periodic, D._ptr(hm), D._ptr(hi))
    g = s.neighs()
    lib.libgint_set_neighs(h, D._ptr(g), s.Ng)
    table = D.abi_table(lib)
    D.drv().gdrv_emit_basis(s.h, C.byref(table), h)
    lib.libgint_amd_plan_begin(h)   # recording: nothing reaches the device
    st.replay(table, h, stride=stride, blocks=blocks)
    q = _queue(lib, h)
    lib.libgint_amd_destroy_handle(h)
    return q


def test_set_level_records_from_calls_and_blocks():
    """the driver's CP2K-like block offsets: every set quartet reduces to one
    32-byte set-level record, whether its K layout arrives as add_qrt /
    add_qrtt calls or as one libgint_amd_add_set_blocks call"""
    from libgint_amd import driver as D, load

    lib = load()
    s = D.System(n_water=2, n_tm=1, box_A=6.0, image_shells=1, np_max=2, np_tm=2, seed=5)
    st = s.record()
    q_calls = _recorded_queue(lib, s, st, blocks=False)
    q_blocks = _recorded_queue(lib, s, st, blocks=True)
    assert q_calls[1] == st.n_sets and q_calls[6] == st.n_sets, q_calls
    assert q_calls == q_blocks


def test_oracle_blocks_equal_calls():
    """the oracle expands a set-level record into the same add_qrt / add_qrtt
    calls as the per-call stream, so both give the same K bit for bit"""
    from libgint_amd import driver as D

    o = oracle_lib()
    s = D.System(n_water=1, n_tm=1, box_A=5.0, image_shells=0, np_max=2, np_tm=1, seed=3)
    st = s.record()
    Ka, ha = D.kbuild(o, "oracle_", s, st, blocks=False)
    Kb, hb = D.kbuild(o, "oracle_", s, st, blocks=True)
    o.oracle_destroy_handle(ha)
    o.oracle_destroy_handle(hb)
    assert np.max(np.abs(Ka)) > 0
    np.testing.assert_array_equal(Ka, Kb)


def oracle_lib():
    from helpers import oracle

    return oracle()


def _abort_script(body):
    """runs `body` in a fresh interpreter (the library aborts on a refused
    argument); returns (exit code, stderr)"""
    code = "import sys; sys.path.insert(0, %r)\n" % ROOT + body
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    return r.returncode, r.stderr


_C0_SCRIPT = """
import ctypes as C, numpy as np
from libgint_amd import load
from libgint_amd.driver import tcg_table
lib = load()
dp = C.POINTER(C.c_double)
c0, ld = tcg_table(8)            # derivative orders 0..8
assert ld == 9 * 105
h = lib.libgint_create_handle()
lib.libgint_set_Potential_Truncated(h, 11.3, c0.ctypes.data_as(dp), ld, c0.size)
R = np.zeros(3); Z = np.array([1.0, 0.4]); K = np.array([0.6, 0.5])
for i, l in enumerate(({L})):
    lib.libgint_set_Atom(h, i, R.ctypes.data_as(dp), Z.ctypes.data_as(dp), 2)
    lib.libgint_set_Atom_L(h, i, l, 1, K.ctypes.data_as(dp))
lib.libgint_amd_plan_begin(h)
lib.libgint_add_prm(h, 0, 0, 0, 0)
lib.libgint_add_shell(h, 0, 1, 2, 3, 0, 0)
print("accepted", flush=True)
"""


def test_c0_table_of_fewer_orders():
    """a C0 table of 9 derivative orders (e.g. sized for an spd basis) is
    accepted; quartets up to L = 8 use it, a quartet of L = 9 is refused
    with a message instead of reading past the table"""
    rc, err = _abort_script(_C0_SCRIPT.replace("{L}", "2, 2, 2, 2"))
    assert rc == 0, err
    rc, err = _abort_script(_C0_SCRIPT.replace("{L}", "3, 2, 2, 2"))
    assert rc != 0 and "fewer derivative orders" in err, (rc, err[-500:])


def test_config4_lattice_is_the_256_shortest():
    """the large periodic workload uses 256 lattice vectors, the reference's
    cell limit: every vector with |g|^2 <= 14 (in cell edges) and 5 of |g|^2 = 17"""
    from libgint_amd import driver as D

    s = D.System(n_water=1, box_A=6.0, image_shells=3, basis=1, np_max=2, seed=1)
    assert s.Ng == 256
    g = s.neighs().reshape(-1, 3)
    edge = s.cell()[0][0]
    n2 = np.rint(np.sum((g / edge) ** 2, axis=1)).astype(int)
    assert np.all(np.diff(n2) >= 0) and n2[0] == 0
    assert np.sum(n2 <= 14) == 251 and np.all(n2[251:] == 17)
    assert len({tuple(x) for x in np.rint(g / edge).astype(int)}) == 256


def test_mixed_layouts_reach_every_record_kind():
    """ks_mix: the set quartets' calls arrive as set-level layouts, per-class
    affine layouts and explicit lists, a third each"""
    from libgint_amd import driver as D, load

    lib = load()
    s = D.System(n_water=2, box_A=6.0, image_shells=1, np_max=2, seed=5, ks_mix=1)
    st = s.record()
    q = _recorded_queue(lib, s, st, blocks=False)
    assert 0 < q[6] < st.n_sets and q[5] > 0 and q[4] > q[5] + q[6], q
