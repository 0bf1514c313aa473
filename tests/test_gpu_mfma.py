"""The matrix-core tiles of the contraction kernel (v_mfma_f64_16x16x4f64
and, for the mixed-precision path, v_mfma_f32_16x16x4f32) with the lane
layout the kernel assumes, against a host product: a wrong layout
assumption (A or B transposed, D rows and columns swapped) fails.  The
entries are small integers and halves, exact in single precision too."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [0, 1], ids=["f64", "f32"])
def test_mfma_tile_layout(prec):
    from libgint_amd import load

    lib = load()
    rng = np.random.default_rng(3)
    dp = C.POINTER(C.c_double)
    for trial in range(3):
        # distinct, exactly representable values so that any permutation shows
        A = np.ascontiguousarray(rng.integers(-64, 64, size=(16, 4)).astype(np.float64) + 0.5 * trial)
        B = np.ascontiguousarray(rng.integers(-64, 64, size=(4, 16)).astype(np.float64))
        D = np.zeros((16, 16))
        assert lib.libgint_amd_mfma_selftest(A.ctypes.data_as(dp), B.ctypes.data_as(dp), D.ctypes.data_as(dp),
                                             prec) == 0
        ref = A @ B
        assert np.array_equal(D, ref), (trial, np.argwhere(D != ref)[:5])
        assert not np.array_equal(D, ref.T)
