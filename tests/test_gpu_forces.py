"""GPU parity of the force kernels against the oracle on identical inputs.

Bar: FP64 build, per-atom forces within 1e-10 of max|F| and energies within
1e-10 relative; mixed build within 1e-5 of max|F|.
"""
import numpy as np
import pytest

import fmd_amd as F
from tests.pairsys import build, lattice_fcc, random_separated, rel_err

pytestmark = pytest.mark.gpu

AR = dict(sig=3.405, eps=0.0104, rc=2.5 * 3.405, a=5.26, m=39.948)


def _compare(o, d, tol_f, tol_e, ag=-1):
    o.forces(ag)
    d.forces(ag)
    eo = o.energies()
    got = d.download()
    ref = o.atoms()
    assert np.array_equal(got["cell_count"], ref["cell_count"])
    assert np.array_equal(got["atomid"], ref["atomid"])
    ef = rel_err(got["f"], ref["f"])
    assert ef < tol_f, "force error %.3e" % ef
    t = d.read()
    assert t.err == 0
    pe_err = abs(t.pe - eo["pe"]) / max(abs(eo["pe"]), 1e-300)
    assert pe_err < tol_e, "pe error %.3e (%r vs %r)" % (pe_err, t.pe, eo["pe"])
    return got, ref


def test_lj_argon_4000(rng):
    x = lattice_fcc((10, 10, 10), AR["a"], rng, jitter=0.15)
    L = 10 * AR["a"]
    o, d = build([L] * 3, [1, 1, 1], [AR["m"]], ["Ar"], [("lj", 0, 0, AR["sig"], AR["eps"], AR["rc"])], x)
    _compare(o, d, 1e-10, 1e-10)


def test_morse_cu_mixed_pbc(rng):
    x = lattice_fcc((8, 8, 9), 3.615, rng, jitter=0.05)
    L = [8 * 3.615, 8 * 3.615, 9 * 3.615 + 12.0]
    o, d = build(L, [1, 1, 0], [63.546], ["Cu"], [("morse", 0, 0, 0.3429, 1.3588, 2.866, 7.0)], x)
    _compare(o, d, 1e-10, 1e-10)


def test_eam_cu01_fixture(rng, cu01):
    x = lattice_fcc((8, 8, 8), 3.615, rng, jitter=0.04)
    L = [8 * 3.615] * 3
    o, d = build(L, [1, 1, 1], [63.546], ["Cu"], [("eam", cu01, [(0, 0)])], x)
    got, ref = _compare(o, d, 1e-10, 1e-10)
    # pass 1 leaves F'(rho) per atom in aux, as the reference leaves it in vaream
    assert rel_err(got["aux"], ref["aux"]) < 1e-10


def test_eam_synthetic_setfl(rng, synth_setfl):
    x = lattice_fcc((7, 7, 7), 3.615, rng, jitter=0.02)
    L = [7 * 3.615] * 3
    o, d = build(L, [1, 1, 1], [63.546], ["Cu"], [("eam", synth_setfl, [(0, 0)])], x)
    _compare(o, d, 1e-10, 1e-10)


@pytest.mark.parametrize("precision", [F.FMDX_PREC_FP64, F.FMDX_PREC_MIXED])
def test_hybrid_eam_lj_two_kinds_with_group(rng, cu01, precision):
    # Cu (EAM) and Ar (LJ Ar-Ar, LJ Cu-Ar) as in examples/02_collision.c
    xcu = lattice_fcc((5, 5, 5), 3.6316, rng, jitter=0.03, origin=(2.0, 2.0, 2.0))
    xar = lattice_fcc((4, 4, 4), 5.26, rng, jitter=0.05, origin=(22.0, 2.0, 2.0))
    x = np.vstack([xcu, xar])
    kind = np.r_[np.zeros(len(xcu)), np.ones(len(xar))].astype(np.int32)
    group = np.r_[np.zeros(len(xcu)), np.ones(len(xar))].astype(np.int32)
    L = [48.0, 30.0, 30.0]
    pots = [("eam", cu01, [(0, 0)]), ("lj", 1, 1, 3.40, 0.0104, 2.5 * 3.40), ("lj", 0, 1, 2.87, 0.0652, 2.5 * 2.87)]
    o, d = build(L, [1, 1, 1], [63.546, 39.948], ["Cu", "Ar"], pots, x, group=group, kind=kind,
                 precision=precision)
    tol = (1e-10, 1e-10) if precision == F.FMDX_PREC_FP64 else (1e-5, 1e-6)
    _compare(o, d, *tol)
    _compare(o, d, *tol, ag=0)
    _compare(o, d, *tol, ag=1)


@pytest.mark.parametrize("precision", [F.FMDX_PREC_FP64, F.FMDX_PREC_MIXED])
def test_eam_two_elements(rng, tmp_path, precision):
    """eam/alloy with two elements (general-category kernel: per-element density functions,
    shared unlike-pair phi) on a random 40/60 mixture."""
    from tests.pairsys import write_setfl_two
    setfl = write_setfl_two(str(tmp_path / "two.eam.alloy"))
    x = lattice_fcc((7, 7, 7), 3.65, rng, jitter=0.04)
    kind = (rng.random(len(x)) < 0.4).astype(np.int32)
    L = [7 * 3.65] * 3
    o, d = build(L, [1, 1, 1], [63.546, 58.69], ["Cu", "Xx"], [("eam", setfl, [(0, 0), (0, 1), (1, 1)])], x,
                 kind=kind, precision=precision)
    if precision == F.FMDX_PREC_FP64:
        got, ref = _compare(o, d, 1e-10, 1e-10)
        assert rel_err(got["aux"], ref["aux"]) < 1e-10
    else:
        _compare(o, d, 1e-5, 1e-6)


@pytest.mark.parametrize("pot", ["lj", "eam"])
@pytest.mark.parametrize("precision", [F.FMDX_PREC_FP64, F.FMDX_PREC_MIXED])
def test_pairs_at_the_cutoff(pot, precision, synth_setfl):
    """Dimers just inside and just outside the cutoff, 2e-13 of rc away from it.  The
    candidate scans select with a margin above rc^2 (1e-12 relative in FP64, 2.5e-4 in
    FP32) and the pair evaluation tests r^2 < rc^2 again; a dimer outside rc must
    contribute nothing, and LJ and the synthetic table are both nonzero at rc.  The
    mixed build evaluates pairs within 1e-5 of rc^2 (and, for the synthetic table,
    over the steep run of spline intervals below rc^2 that its missing taper leaves)
    in FP64 from the FP64 positions, so it keeps the same pair set and meets its 1e-5
    bar on these dimers too, which get all of their force from those intervals."""
    rel = 2e-13
    if pot == "lj":
        rc, pots, names, mass = AR["rc"], [("lj", 0, 0, AR["sig"], AR["eps"], AR["rc"])], ["Ar"], [AR["m"]]
    else:
        rc, pots, names, mass = 5.5, [("eam", synth_setfl, [(0, 0)])], ["Cu"], [63.546]
    L = 6.0 * rc
    x = []
    # along x, y, z and a diagonal, inside and outside by rel
    dirs = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in ((1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 1))]
    k = 0
    for dvec in dirs:
        for sgn in (-1.0, 1.0):
            a = np.array([0.7 + 1.5 * rc * (k % 4), 0.9 + 3.0 * rc * (k // 4), 1.1 + 1.5 * rc * (k % 2)]) % L
            x += [a, (a + dvec * rc * (1.0 + sgn * rel)) % L]
            k += 1
    x = np.array(x)
    o, d = build([L] * 3, [1, 1, 1], mass, names, pots, x, precision=precision)
    o.forces()
    d.forces()
    got, ref = d.download(), o.atoms()
    assert np.array_equal(got["atomid"], ref["atomid"])
    fmax = np.max(np.abs(ref["f"]))
    assert fmax > 0
    # the outside dimers carry exactly zero force in the oracle
    ids = ref["atomid"]
    outside = np.isin(ids, [2 * j + b for j in range(8) if j % 2 == 1 for b in (0, 1)])
    assert np.all(ref["f"][outside] == 0.0)
    assert np.all(got["f"][outside] == 0.0)
    tol = 1e-10 if precision == F.FMDX_PREC_FP64 else 1e-5
    assert rel_err(got["f"], ref["f"], scale=fmax) < tol


def test_mixed_precision_eam(rng, cu01):
    x = lattice_fcc((8, 8, 8), 3.615, rng, jitter=0.04)
    L = [8 * 3.615] * 3
    o, d = build(L, [1, 1, 1], [63.546], ["Cu"], [("eam", cu01, [(0, 0)])], x, precision=F.FMDX_PREC_MIXED)
    _compare(o, d, 1e-5, 1e-6)


def test_mixed_precision_lj(rng):
    x = lattice_fcc((10, 10, 10), AR["a"], rng, jitter=0.15)
    L = 10 * AR["a"]
    o, d = build([L] * 3, [1, 1, 1], [AR["m"]], ["Ar"], [("lj", 0, 0, AR["sig"], AR["eps"], AR["rc"])], x,
                 precision=F.FMDX_PREC_MIXED)
    _compare(o, d, 1e-5, 1e-6)


def test_mixed_precision_morse(rng):
    # the FP32 Morse pass: r and 1/r from one reciprocal square root, the
    # exponential from the hardware's base-2 instruction; pairs from the 1st to
    # the 7th neighbour shell inside rc = 7 A, open along z
    x = lattice_fcc((8, 8, 8), 3.615, rng, jitter=0.1)
    L = [8 * 3.615, 8 * 3.615, 8 * 3.615 + 10.0]
    o, d = build(L, [1, 1, 0], [63.546], ["Cu"], [("morse", 0, 0, 0.3429, 1.3588, 2.866, 7.0)], x,
                 precision=F.FMDX_PREC_MIXED)
    _compare(o, d, 1e-5, 1e-6)


@pytest.mark.parametrize("precision", [F.FMDX_PREC_FP64, F.FMDX_PREC_MIXED])
def test_degenerate_boxes_1_and_2_cells(rng, precision):
    # 1, 2 and 3 cells along x, y, z: the stencil meets the same cell through
    # several images, and an atom pairs with its own images (pinned on the
    # oracle by an all-images brute force in test_oracle_pins.py)
    rc = AR["rc"]
    L = [1.5 * rc, 2.3 * rc, 3.2 * rc]
    x = random_separated(60, L, 3.0, rng)
    o, d = build(L, [1, 1, 1], [AR["m"]], ["Ar"], [("lj", 0, 0, AR["sig"], AR["eps"], rc)], x, precision=precision)
    assert list(o.nc) == [1, 2, 3]
    tol = (1e-10, 1e-10) if precision == F.FMDX_PREC_FP64 else (1e-5, 1e-6)
    _compare(o, d, *tol)


@pytest.mark.parametrize("precision", [F.FMDX_PREC_FP64, F.FMDX_PREC_MIXED])
@pytest.mark.parametrize("ucells,a,lo,hi", [(7, 3.615, 33, 64), (6, 3.45, 65, 256)])
def test_large_cells_split_into_home_blocks(rng, precision, ucells, a, lo, hi):
    # Morse copper (rc 7 A) on 7^3 FCC unit cells (3 cells per edge of 8.4 A, 35 to 62
    # atoms per cell) and on 6^3 compressed ones (2 cells per edge of 10.4 A, 108 atoms
    # per cell): the kernels take cells of 33 to 64 atoms in a block of 32 and a block
    # of the rest, and larger cells in blocks of 64
    x = lattice_fcc((ucells,) * 3, a, rng, jitter=0.2)
    L = [ucells * a] * 3
    o, d = build(L, [1, 1, 1], [63.546], ["Cu"], [("morse", 0, 0, 0.3429, 1.3588, 2.866, 7.0)], x,
                 precision=precision)
    counts = o.atoms()["cell_count"]
    assert counts.min() >= lo and counts.max() <= hi, (counts.min(), counts.max())
    tol = (1e-10, 1e-10) if precision == F.FMDX_PREC_FP64 else (1e-5, 1e-6)
    _compare(o, d, *tol)


@pytest.mark.parametrize("precision", [F.FMDX_PREC_FP64, F.FMDX_PREC_MIXED])
def test_eam_cells_split_into_home_blocks(rng, cu01, precision):
    # Cu01 EAM on 4^3 FCC unit cells: 2 cells per edge of 7.2 A, 32 atoms per cell before
    # the jitter moves some across (the lattice planes are shifted onto the cell faces):
    # cells of more than 32 atoms are taken in two blocks,
    # whose density pass records no hit words, so that the force pass scans them again
    a = 3.615
    L = [4 * a] * 3
    x = np.mod(lattice_fcc((4, 4, 4), a, rng, jitter=0.25, origin=(-0.2 * a,) * 3), L[0])
    o, d = build(L, [1, 1, 1], [63.546], ["Cu"], [("eam", cu01, [(0, 0)])], x, precision=precision)
    counts = o.atoms()["cell_count"]
    assert counts.max() > 32 and counts.min() <= 32, (counts.min(), counts.max())
    tol = (1e-10, 1e-10) if precision == F.FMDX_PREC_FP64 else (1e-5, 1e-6)
    got, ref = _compare(o, d, *tol)
    assert rel_err(got["aux"], ref["aux"]) < tol[0]
