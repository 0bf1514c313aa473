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
import subprocess

import numpy as np
import pytest

from helpers import ROOT

HEADER = os.path.join(ROOT, "include", "libGint.h")
REF_SRC = "/root/reference/src"


def _declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return re.findall(r"^\s*(?:void|double|int)\s*\*?\s*(\w+)\s*\(", txt, flags=re.M)


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
    out = (C.c_uint64 * 6)()
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
