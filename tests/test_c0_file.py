"""CP2K's t_c_g.dat coefficient file (reference src/t_c_g0_n.cpp:29-51,
read_c0): 207 patches of 22 x 105 numbers, patch after patch; a reader for
orders 0..nder keeps the first (nder+1)*105 of each patch.  Host only."""
import numpy as np
import pytest

from helpers import tcg_table


def test_round_trip_is_exact(tmp_path):
    from libgint_amd import read_c0, write_c0

    full, ld = tcg_table(21)
    assert ld == 2310
    path = tmp_path / "t_c_g.dat"
    write_c0(path, full)
    lines = path.read_text().splitlines()
    assert len(lines) == 207 and all(len(x.split()) == 2310 for x in lines[:3])
    back, ld21 = read_c0(path, 21)
    assert ld21 == 2310 and np.array_equal(back, full)
    for nder in (0, 5, 12):
        part, ldp = read_c0(path, nder)
        assert ldp == (nder + 1) * 105
        assert np.array_equal(part, full[:, :ldp])


def test_order_of_numbers_and_free_format(tmp_path):
    """value i*2310+j at patch i, position j, written in mixed number formats
    and line breaks: the reader's layout follows the file's order alone"""
    from libgint_amd import read_c0

    vals = np.arange(207 * 2310, dtype=float).reshape(207, 2310)
    fmts = ["%.1f", "%.6E", "%g", "%.8e", "%d"]
    chunks = []
    for i in range(207):
        for j in range(2310):
            v = vals[i, j]
            chunks.append((fmts[(i + j) % len(fmts)] % (int(v) if (i + j) % len(fmts) == 4 else v)))
            chunks.append("\n" if (j % 7 == 6) else " ")
    path = tmp_path / "free.dat"
    path.write_text("".join(chunks))
    c, ld = read_c0(path, 1)
    assert ld == 210
    assert np.array_equal(c, vals[:, :210])
    c, ld = read_c0(path, 21)
    assert np.array_equal(c, vals)


def test_malformed_files_are_errors(tmp_path):
    import ctypes as C

    from libgint_amd import load

    lib = load()
    buf = np.zeros(207 * 2310)
    dp = buf.ctypes.data_as(C.POINTER(C.c_double))
    ld = C.c_int(-7)
    short = tmp_path / "short.dat"
    short.write_text(" ".join(["1.0"] * (207 * 2310 - 1)))
    assert lib.libgint_amd_read_c0(bytes(short), 21, dp, buf.size, C.byref(ld)) == -4
    bad = tmp_path / "bad.dat"
    bad.write_text("1.0 2.0D+00 " + " ".join(["1.0"] * (207 * 2310)))
    assert lib.libgint_amd_read_c0(bytes(bad), 0, dp, buf.size, C.byref(ld)) == -4
    assert ld.value == -7
    assert lib.libgint_amd_read_c0(bytes(tmp_path / "missing.dat"), 3, dp, buf.size, C.byref(ld)) == -1
    assert lib.libgint_amd_read_c0(bytes(short), 22, dp, buf.size, C.byref(ld)) == -2
    assert lib.libgint_amd_read_c0(bytes(short), 3, dp, 207 * 420 - 1, C.byref(ld)) == -3
    with pytest.raises(ValueError):
        from libgint_amd import read_c0

        read_c0(short, 21)
