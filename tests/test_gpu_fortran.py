"""The reference's Fortran binding (module libgint, the interface CP2K calls)
driving this library on the GPU: oracle/_ref/drive_libgint, built from the
reference's src/interface_libgint.F90 and tests/fortran/drive_libgint.F90
where the reference checkout exists, prints K for one set quartet; the CPU
oracle run on the same calls must agree."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from helpers import ROOT, TCG_BIN, as_dp, oracle, tcg_table

pytestmark = pytest.mark.gpu
DRIVE = os.path.join(ROOT, "oracle", "_ref", "drive_libgint")


def _oracle_K():
    lib = oracle()
    h = lib.oracle_create_handle()
    c0, ld = tcg_table(21)
    lib.oracle_set_Potential_Truncated(h, 11.3384, as_dp(c0.ravel()), ld, c0.size)
    lib.oracle_set_hf_fac(h, 1.0)
    cell = np.eye(3).ravel() * 100.0
    celli = np.eye(3).ravel() * 0.01
    lib.oracle_set_cell(h, False, as_dp(cell), as_dp(celli))
    lib.oracle_set_neighs(h, as_dp(np.zeros(3)), 1)
    lib.oracle_set_Atom(h, 0, as_dp(np.zeros(3)), as_dp(np.array([1.3, 0.4])), 2)
    lib.oracle_set_Atom_L(h, 0, 0, 2, as_dp(np.array([0.6, 0.5, 0.2, 0.9])))
    P = 0.01 * np.arange(1, 17, dtype=float)
    lib.oracle_set_P(h, as_dp(P), 16)
    lib.oracle_add_prm(h, 0, 0, 0, 0)
    lib.oracle_add_prm(h, 1, 1, 1, 1)
    lib.oracle_add_shell(h, 0, 0, 0, 0, 0, 0)
    lib.oracle_add_qrt(h, 0, 0, 0, 0, 2, 2, 2, 2)
    for i in range(16):
        a, b, c, d = i // 8, (i // 4) % 2, (i // 2) % 2, i % 2
        lib.oracle_add_qrtt(h, 0.25, 0, 0, 0, 0, a, b, c, d, 4, 4, 4, 4, a + 4 * c, a + 4 * d, b + 4 * c, b + 4 * d,
                            False, False, False, False)
    lib.oracle_add_set(h)
    K = np.zeros(16)
    lib.oracle_get_K(h, as_dp(K))
    lib.oracle_destroy_handle(h)
    return K


def test_fortran_binding_drives_the_gpu_library():
    if not os.path.exists(DRIVE):
        pytest.skip("Fortran caller not built (needs the reference checkout at build time)")
    print("running", DRIVE, flush=True)
    r = subprocess.run([DRIVE, TCG_BIN], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    Kg = np.array([float(x) for x in r.stdout.split()])
    Ko = _oracle_K()
    print("max|K|", np.max(np.abs(Ko)), "max|dK|", np.max(np.abs(Kg - Ko)), flush=True)
    assert Kg.shape == (16,) and np.max(np.abs(Ko)) > 0
    assert np.max(np.abs(Kg - Ko)) <= 1e-10 + 1e-11 * np.max(np.abs(Ko))
