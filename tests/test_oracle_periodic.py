"""Independent pins of the oracle's periodic lattice sums (minimum-image
wrap of B - A and D - C, the n1 / n2 images, the recentring of the ket on
the image nearest P and the n3 sum; reference src/compute_Fm.cpp:79-136,
src/compute_ECO.cpp:216-272), by properties the periodic exchange matrix
must have whatever the implementation:

* moving one atom by a lattice vector describes the same crystal: K is
  unchanged (every pair difference is wrapped back);
* a cell doubled along x, holding two copies of the system and the density
  of every copy pair equal to the original pair's (a Gamma-point density),
  describes the same crystal: its exchange matrix, summed over the copies
  of its second index, is the original one, for either copy of the first
  index.  Image sums that were incomplete, double counted or recentred on
  the wrong cell would make the two lattices disagree;
* with a cell much longer than the interaction range the periodic K-build
  is the molecular one.

Two regimes, each with two systems whose molecules interact across the x
faces of the cell (up to 3% of max|K| between the copies of the doubled
cell):

* tight: exponents 2.5 .. 5e3 bohr^-2, a 1.2 A cut-off, 27 cells of 7 A.
  Every overlap and truncated interaction stays inside one neighbouring
  cell (a contribution at |g| needs one reaching |g| / 2 = 3.5 A, where
  both are below 1e-20): pins the minimum-image wraps and the recentring;
* diffuse: exponents 0.6 .. 4, a 2.5 A cut-off, 125 cells of 6 A.  Second
  images of B and D and ket cells beyond the nearest one contribute, and
  the sums converge inside +-2 cells (each reaches at most 6.6 A, under
  1.5 cells): pins the n1 / n2 images and the n3 sum.

Tolerance: the parity bar, 1e-10 + 1e-11 max|K|: the two lattices order
their sums differently, and slots within rounding of the 1e-12 screening
threshold may be kept in one and dropped in the other."""
import numpy as np
import pytest

from helpers import oracle

TIGHT = dict(n_water=8, box_A=7.0, image_shells=1, basis=1, np_min=1, np_max=2, zmin=2.5, rcut_A=1.2, eps=1e-28,
             min_oo_A=2.2)
DIFFUSE = dict(n_water=3, box_A=6.0, image_shells=2, basis=1, np_min=1, np_max=2, zmin=0.6, zmax=4.0, rcut_A=2.5,
               eps=1e-28, min_oo_A=2.2)
SYSTEMS = [("tight", TIGHT, 40), ("tight", TIGHT, 46), ("diffuse", DIFFUSE, 40), ("diffuse", DIFFUSE, 41)]
IDS = ["%s_%d" % (n, sd) for n, _, sd in SYSTEMS]


def _tol(scale):
    return 1e-10 + 1e-11 * scale


def _K(**kw):
    from libgint_amd import driver as D

    o = oracle()
    s = D.System(**kw)
    st = s.record()
    K, h = D.kbuild(o, "oracle_", s, st)
    o.oracle_destroy_handle(h)
    return s, s.dense(s.symmetrize(K)), st


@pytest.mark.parametrize("name,base,seed", SYSTEMS, ids=IDS)
def test_lattice_vector_moves_nothing(name, base, seed):
    s0, K0, st0 = _K(seed=seed, **base)
    scale = np.max(np.abs(K0))
    assert scale > 1e-3
    for atom, cells in ((1, (1, 0, -1)), (3, (0, -1, 2)), (0, (-1, 1, 1))):
        s1, K1, st1 = _K(seed=seed, shift_atom=atom, shift_cells=cells, **base)
        err = np.max(np.abs(K1 - K0))
        print("atom", atom, "moved by", cells, "max|dK|", err, "max|K|", scale)
        assert err <= _tol(scale), (atom, cells, err)


@pytest.mark.parametrize("name,base,seed", SYSTEMS, ids=IDS)
def test_doubled_cell_folds_to_the_cell(name, base, seed):
    s1, K1, _ = _K(seed=seed, **base)
    s2, K2, st2 = _K(seed=seed, replicate=2, **base)
    n = s1.nbf
    assert s2.nbf == 2 * n and s2.natom == 2 * s1.natom
    scale = np.max(np.abs(K1))
    assert scale > 1e-3
    for c1 in range(2):
        rows = K2[c1 * n:(c1 + 1) * n]
        folded = rows[:, :n] + rows[:, n:]
        err = np.max(np.abs(folded - K1))
        # the block between the two copies carries the interactions across
        # the x faces, which the original cell sums over its images
        part = np.max(np.abs(rows[:, (1 - c1) * n:(2 - c1) * n]))
        print("copy", c1, "max|folded - K|", err, "max|K|", scale, "max|other copy block|", part)
        assert part > 1e-4 * scale
        assert err <= _tol(scale), (c1, err)


def test_long_cell_is_the_molecule():
    kw = dict(n_water=2, box_A=6.0, basis=1, np_max=2, seed=45)
    s0, K0, _ = _K(image_shells=0, cell_A=100.0, **kw)
    s1, K1, _ = _K(image_shells=1, cell_A=100.0, **kw)
    scale = np.max(np.abs(K0))
    err = np.max(np.abs(K1 - K0))
    print("27 cells of 100 A vs one cell: max|dK|", err, "max|K|", scale)
    assert scale > 1e-3 and err <= 1e-13 * scale
