"""Device state checks and the guards around batch membership.

* fmdx_audit counts the atoms over the cells, finds the fullest cell and the
  atoms with a non-finite position, velocity or force; on a healthy system
  it agrees with the downloaded state, in both precisions and after steps
  that move atoms between cells.
* The force pass replays the density pass's hit words only for tile rounds
  that staged the same number of candidates; on these systems every round
  matches (the mismatch counter stays 0) and no bad-force or bad-position
  cell is recorded.
* While a system is a member of a batch, the calls that free or replace its
  device arrays refuse it (the batch's captured step reads them); after
  fmdx_batch_remove they work again.  The batch is never stepped here.
"""
import ctypes as C

import numpy as np
import pytest

import fmd_amd as F
from fmd_amd.workloads import maxwell_velocities
from tests.pairsys import build, lattice_fcc

pytestmark = pytest.mark.gpu


def _cu(rng, setfl, precision, n=(8, 8, 8), T=900.0):
    a = 3.615
    x = lattice_fcc(n, a, rng, jitter=0.05)
    v = maxwell_velocities(len(x), 63.546, T, rng)
    L = [n[0] * a, n[1] * a, n[2] * a]
    return build(L, [1, 1, 1], [63.546], ["Cu"], [("eam", setfl, [(0, 0)])], x, v=v, precision=precision)


@pytest.mark.parametrize("precision", [F.FMDX_PREC_FP64, F.FMDX_PREC_MIXED])
def test_audit_matches_the_downloaded_state(rng, cu01, precision):
    print("audit: Cu01 8^3 FCC, precision", precision, flush=True)
    o, d = _cu(rng, cu01, precision)
    d.forces()
    d.thermo()
    for _ in range(40):   # 900 K, 2 fs: atoms cross cells
        d.vv_start(2e-3)
        d.forces()
        d.vv_finish(2e-3)
    t = d.audit()
    got = d.download()
    assert t.err == 0
    assert t.audit_atoms == t.natoms == len(got["x"]) == 2048
    assert t.audit_max_cell == int(np.max(got["cell_count"]))
    assert t.audit_max_cell <= t.cap
    assert t.audit_nonfinite == 0
    assert t.replay_mismatch == 0 and t.exact_unmatched == 0
    assert t.bad_force_cell == 0 and t.bad_position_cell == 0
    # the atoms did move between cells, so the check saw re-binned cells
    assert not np.array_equal(got["atomid"], o.atoms()["atomid"])


def test_audit_on_the_synthetic_table_mixed(rng, synth_setfl):
    print("audit: synthetic setfl, mixed", flush=True)
    _, d = _cu(rng, synth_setfl, F.FMDX_PREC_MIXED, n=(7, 7, 7), T=600.0)
    d.forces()
    d.thermo()
    for _ in range(20):
        d.vv_start(2e-3, thermostat=1, tau=0.1, T0=600.0)
        d.forces()
        d.vv_finish(2e-3)
    t = d.audit()
    assert t.err == 0 and t.audit_nonfinite == 0 and t.replay_mismatch == 0 and t.exact_unmatched == 0
    assert t.audit_atoms == t.natoms == 4 * 7 ** 3


def test_batch_member_guards(rng):
    print("member guards: LJ argon system in a batch, never stepped", flush=True)
    x = lattice_fcc((5, 5, 5), 5.26, rng, jitter=0.02)
    v = maxwell_velocities(len(x), 39.948, 100.0, rng)
    L = [5 * 5.26] * 3
    o, d = build(L, [1, 1, 1], [39.948], ["Ar"], [("lj", 0, 0, 3.405, 0.0104, 2.5 * 3.405)], x, v=v)
    lib = F.load_fmdx()
    st = o.atoms()
    b = F.Batch(F.FMDX_PREC_FP64)
    b.add(d, 1e-3)
    assert lib.fmdx_system_destroy(d.h) == F.FMDX_ESTATE
    assert lib.fmdx_set_capacity(d.h, 96) == F.FMDX_ESTATE
    cc = np.ascontiguousarray(st["cell_count"], dtype=np.uint32)
    xx = np.ascontiguousarray(st["x"]).reshape(-1)
    assert lib.fmdx_upload(d.h, len(st["x"]), cc.ctypes.data_as(C.c_void_p), xx.ctypes.data_as(C.c_void_p),
                           None, None, None, None) == F.FMDX_ESTATE
    # a second batch may not take it either
    b2 = F.Batch(F.FMDX_PREC_FP64)
    assert lib.fmdx_batch_add(b2.h, d.h, 1e-3, 0, 1.0, 0.0) == F.FMDX_ESTATE
    b2.close()
    b.remove(d)
    d.set_capacity(96)
    assert d.read().cap == 96
    b.add(d, 1e-3)
    b.close()          # the batch lets its members go
    d.set_capacity(64)
    d.close()


@pytest.mark.parametrize("precision", [F.FMDX_PREC_FP64, F.FMDX_PREC_MIXED])
@pytest.mark.parametrize("table", ["cu01", "lj"])
def test_force_call_without_energy(rng, cu01, precision, table):
    """fmdx_update_forces_ex without FMDX_FORCES_ENERGY: the same force and F'(rho) bits as with
    it, the recorded potential energy left as the last energy call set it."""
    print("forces without energy:", table, precision, flush=True)
    lib = F.load_fmdx()
    if table == "cu01":
        o, d = _cu(rng, cu01, precision, n=(7, 7, 7))
    else:
        x = lattice_fcc((7, 7, 7), 5.26, rng, jitter=0.1)
        o, d = build([7 * 5.26] * 3, [1, 1, 1], [39.948], ["Ar"], [("lj", 0, 0, 3.405, 0.0104, 2.5 * 3.405)], x,
                     precision=precision)
    d.forces()
    pe0 = d.read().pe
    ref = d.download()
    d.vv_start(1e-3)
    F._ck(lib.fmdx_update_forces_ex(d.h, -1, 0, 0), "fmdx_update_forces_ex")
    t = d.read()
    assert t.pe == pe0 and t.err == 0
    a = d.download()
    F._ck(lib.fmdx_update_forces_ex(d.h, -1, 0, F.FMDX_FORCES_ENERGY), "fmdx_update_forces_ex")
    b = d.download()
    assert np.array_equal(a["f"], b["f"]) and np.array_equal(a["aux"], b["aux"])
    assert d.read().pe != pe0 and not np.array_equal(a["x"], ref["x"])
