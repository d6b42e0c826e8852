"""fmdx_update_forces_noalloc against fmdx_update_forces.

The allocation-free entry point keeps the call's scratch in one fixed device
array per GPU and skips the density pass's hit words (the force pass rescans);
its forces, F'(rho) and energies must be the same bits as the existing entry
point's on identical systems.  Two calls in a row on the same GPU, the second on
a smaller system, so the second reads a scratch array holding the first call's
values; an EAM system in both precisions and an LJ system (no density pass).  A
system beyond the array's cell limit is refused with FMDX_ESIZE before
anything runs.
"""
import ctypes as C

import numpy as np
import pytest

import fmd_amd as F
from fmd_amd.workloads import maxwell_velocities
from tests.pairsys import build, lattice_fcc

pytestmark = pytest.mark.gpu


def _cu(seed, n, setfl, precision):
    rng = np.random.default_rng(seed)
    x = lattice_fcc(n, 3.615, rng, jitter=0.05)
    v = maxwell_velocities(len(x), 63.546, 600.0, rng)
    L = [n[0] * 3.615, n[1] * 3.615, n[2] * 3.615]
    return build(L, [1, 1, 1], [63.546], ["Cu"], [("eam", setfl, [(0, 0)])], x, v=v, precision=precision)[1]


def _ar(seed, n):
    rng = np.random.default_rng(seed)
    x = lattice_fcc(n, 5.26, rng, jitter=0.1)
    L = [n[0] * 5.26, n[1] * 5.26, n[2] * 5.26]
    return build(L, [1, 1, 1], [39.948], ["Ar"], [("lj", 0, 0, 3.405, 0.0104, 2.5 * 3.405)], x,
                 precision=F.FMDX_PREC_MIXED)[1]


def _same(a, b):
    ga, gb = a.download(), b.download()
    for k in ("cell_count", "atomid", "f", "aux"):
        assert np.array_equal(ga[k], gb[k]), k
    ta, tb = a.read(), b.read()
    assert ta.pe == tb.pe and ta.embed == tb.embed
    assert ta.err == 0 and tb.err == 0


@pytest.mark.parametrize("precision", [F.FMDX_PREC_FP64, F.FMDX_PREC_MIXED])
def test_noalloc_matches_update_forces_two_calls(cu01, synth_setfl, precision):
    lib = F.load_fmdx()
    for seed, n, table in ((1, (12, 12, 12), synth_setfl), (2, (7, 7, 8), cu01)):   # larger, then smaller
        print("noalloc: EAM %s, precision %d" % (n, precision), flush=True)
        a, b = _cu(seed, n, table, precision), _cu(seed, n, table, precision)
        a.forces()
        F._ck(lib.fmdx_update_forces_noalloc(b.h, -1, 0), "fmdx_update_forces_noalloc")
        _same(a, b)
        # and after a step, where atoms changed cells
        for d, noalloc in ((a, False), (b, True)):
            d.thermo()
            d.vv_start(2e-3)
            if noalloc:
                F._ck(lib.fmdx_update_forces_noalloc(d.h, -1, 0), "fmdx_update_forces_noalloc")
            else:
                d.forces()
            d.vv_finish(2e-3)
        _same(a, b)
        a.close()
        b.close()


def test_noalloc_lj_and_size_limit():
    lib = F.load_fmdx()
    print("noalloc: LJ argon 8^3 mixed", flush=True)
    a, b = _ar(3, (8, 8, 8)), _ar(3, (8, 8, 8))
    a.forces()
    F._ck(lib.fmdx_update_forces_noalloc(b.h, -1, 0), "fmdx_update_forces_noalloc")
    _same(a, b)
    a.close()
    b.close()
    print("noalloc: size limit", flush=True)
    big = F.DeviceSystem([1000.0] * 3, [1, 1, 1], [100, 100, 100], [10.0] * 3, [1.0])
    assert big.ncell > F.FMDX_NOALLOC_MAX_CELLS
    assert lib.fmdx_update_forces_noalloc(big.h, -1, 0) == F.FMDX_ESIZE
    big.close()
