"""GPU parity of the integrators and the exact-order re-binning.

* With zero forces a velocity-Verlet start step is exact arithmetic on both
  sides, so positions, velocities, cell counts and the within-cell order after
  re-binning (including periodic migration in all three dimensions and
  deletion at open faces) must match the oracle bit for bit.
* Along real trajectories the forces differ in the last bits only; the cell
  order must still match step for step and the state within 1e-9.
"""
import numpy as np
import pytest

import fmd_amd as F
from tests.pairsys import build, lattice_fcc, rel_err

pytestmark = pytest.mark.gpu

AR = dict(sig=3.405, eps=0.0104, rc=2.5 * 3.405, a=5.26, m=39.948)


def _state_equal(d, o):
    got, ref = d.download(), o.atoms()
    assert np.array_equal(got["cell_count"], ref["cell_count"])
    assert np.array_equal(got["atomid"], ref["atomid"])
    assert np.array_equal(got["x"], ref["x"])
    assert np.array_equal(got["v"], ref["v"])
    assert np.array_equal(got["group"], ref["group"])
    return got


@pytest.mark.parametrize("pbc", [(1, 1, 1), (1, 0, 1), (0, 0, 0)])
def test_rebin_bitexact(rng, pbc):
    x = lattice_fcc((6, 6, 6), AR["a"], rng, jitter=0.2)
    L = [6 * AR["a"]] * 3
    # displacements up to ~0.9 cell per step: many movers, corner migrations, deletions
    v = rng.uniform(-0.9, 0.9, size=x.shape) * (L[0] / 3) / 1e-2
    o, d = build(L, list(pbc), [AR["m"]], ["Ar"], [("lj", 0, 0, AR["sig"], AR["eps"], AR["rc"])], x, v=v)
    for step in range(3):
        o.vv_start(1e-2)
        d.vv_start(1e-2)
        got = _state_equal(d, o)
        assert d.read().err == 0
    if pbc != (1, 1, 1):
        assert len(got["x"]) < len(x)  # atoms left through the open faces


def test_rebin_two_groups_inactive_frozen(rng):
    x = lattice_fcc((6, 6, 6), AR["a"], rng, jitter=0.2)
    L = [6 * AR["a"]] * 3
    v = rng.uniform(-0.5, 0.5, size=x.shape) * (L[0] / 3) / 1e-2
    g = (np.arange(len(x)) % 2).astype(np.int32)
    o, d = build(L, [1, 1, 1], [AR["m"]], ["Ar"], [("lj", 0, 0, AR["sig"], AR["eps"], AR["rc"])], x, v=v, group=g)
    for _ in range(2):
        o.vv_start(1e-2, ag=1)
        d.vv_start(1e-2, active_group=1)
        _state_equal(d, o)


def test_vv_trajectory_lj(rng):
    x = lattice_fcc((10, 10, 10), AR["a"], rng, jitter=0.05)
    L = [10 * AR["a"]] * 3
    from fmd_amd.workloads import maxwell_velocities
    v = maxwell_velocities(len(x), AR["m"], 300.0, rng)
    o, d = build(L, [1, 1, 1], [AR["m"]], ["Ar"], [("lj", 0, 0, AR["sig"], AR["eps"], AR["rc"])], x, v=v)
    dt = 2e-3
    o.forces()
    d.forces()
    for step in range(25):
        o.vv_start(dt)
        d.vv_start(dt)
        o.forces()
        d.forces()
        o.vv_finish(dt)
        d.vv_finish(dt)
        got, ref = d.download(), o.atoms()
        assert np.array_equal(got["cell_count"], ref["cell_count"]), step
        assert np.array_equal(got["atomid"], ref["atomid"]), step
        assert rel_err(got["x"], ref["x"], scale=L[0]) < 1e-12
        assert rel_err(got["v"], ref["v"]) < 1e-9
        t, e = d.read(), o.energies()
        assert abs(t.temperature - e["T"]) / e["T"] < 1e-9
        assert abs(t.pe - e["pe"]) / abs(e["pe"]) < 1e-9
        assert np.allclose([t.momentum[0], t.momentum[1], t.momentum[2]], e["p"], rtol=0, atol=1e-9)


def test_berendsen_eam(rng, cu01):
    x = lattice_fcc((8, 8, 8), 3.615, rng, jitter=0.02)
    L = [8 * 3.615] * 3
    from fmd_amd.workloads import maxwell_velocities
    v = maxwell_velocities(len(x), 63.546, 600.0, rng)
    o, d = build(L, [1, 1, 1], [63.546], ["Cu"], [("eam", cu01, [(0, 0)])], x, v=v)
    dt, tau, T0 = 2e-3, 0.1, 600.0
    o.forces(); d.forces()
    o.thermo(); d.thermo()
    for step in range(20):
        o.vv_start(dt, thermostat=1, tau=tau, T0=T0)
        d.vv_start(dt, thermostat=1, tau=tau, T0=T0)
        o.forces(); d.forces()
        o.vv_finish(dt); d.vv_finish(dt)
    got, ref = d.download(), o.atoms()
    assert np.array_equal(got["atomid"], ref["atomid"])
    assert rel_err(got["x"], ref["x"], scale=L[0]) < 1e-12
    t, e = d.read(), o.energies()
    assert abs(t.temperature - e["T"]) / e["T"] < 1e-9
    assert abs(t.pe - e["pe"]) / abs(e["pe"]) < 1e-9


def test_symplectic_euler_and_thermo(rng):
    x = lattice_fcc((8, 8, 8), AR["a"], rng, jitter=0.1)
    L = [8 * AR["a"]] * 3
    from fmd_amd.workloads import maxwell_velocities
    v = maxwell_velocities(len(x), AR["m"], 200.0, rng)
    o, d = build(L, [1, 1, 0], [AR["m"]], ["Ar"], [("lj", 0, 0, AR["sig"], AR["eps"], AR["rc"])], x, v=v)
    o.forces(); d.forces()
    for _ in range(10):
        o.se_step(1e-3); d.se_step(1e-3)
        o.forces(); d.forces()
    got, ref = d.download(), o.atoms()
    assert np.array_equal(got["atomid"], ref["atomid"])
    assert rel_err(got["x"], ref["x"], scale=L[0]) < 1e-12
    t, e = d.read(), o.energies()
    assert abs(t.temperature - e["T"]) / e["T"] < 1e-9
    assert int(t.n_active) == e["n"]


def test_single_atom_crosses_every_face(rng):
    """One atom alone (no pair partner, zero force) flying through a periodic box: every step
    is exact arithmetic on both sides, and the atom migrates through faces in all three
    dimensions; the device matches the oracle bit for bit, as do a second, empty system and a
    batch holding both."""
    L = [4 * AR["a"]] * 3
    x = np.array([[0.5 * L[0], 0.5 * L[1], 0.5 * L[2]]])
    v = np.array([[1.7, -2.3, 3.1]]) * L[0] / 20 / 1e-2   # ~1/20 to 1/7 of the box per step
    pot = [("lj", 0, 0, AR["sig"], AR["eps"], AR["rc"])]
    o, d = build(L, [1, 1, 1], [AR["m"]], ["Ar"], pot, x, v=v)
    for _ in range(30):
        o.vv_start(1e-2)
        d.vv_start(1e-2)
        d.forces()
        d.vv_finish(1e-2)
        o.forces()
        o.vv_finish(1e-2)
        got = _state_equal(d, o)
        assert np.array_equal(got["f"], np.zeros((1, 3)))
        assert d.read().err == 0
    # an empty system, alone and in a batch with a one-atom system
    nc = o.nc
    e = F.DeviceSystem(L, [1, 1, 1], list(nc), list(o.cellh), [AR["m"] / F.MASS_UNIT])
    e.set_pair(0, 0, F.POT_LJ, AR["rc"] ** 2, 4.0 * AR["eps"], AR["sig"] ** 2, 0.0)
    e.upload(np.zeros(int(np.prod(nc)), np.uint32), np.zeros((0, 3)), np.zeros((0, 3)))
    e.forces()
    e.thermo()
    t = e.read()
    assert t.err == 0 and t.pe == 0.0 and t.natoms == 0
    o2, d2 = build(L, [1, 1, 1], [AR["m"]], ["Ar"], pot, x, v=v)
    b = F.Batch(F.FMDX_PREC_FP64)
    b.add(e, 1e-2)
    b.add(d2, 1e-2)
    b.step(12)
    st = b.status()
    assert st[0].err == 0 and st[0].natoms == 0 and st[1].err == 0 and st[1].natoms == 1
    b.close()
    o2.forces()
    o2.thermo()
    for _ in range(12):
        o2.vv_start(1e-2)
        o2.forces()
        o2.vv_finish(1e-2)
    _state_equal(d2, o2)
    assert e.download()["x"].shape == (0, 3)
    e.close()
