"""State files after dynamics, resume from a state file, and the laser source's
normalisation, on the GPU through the public API.

* fmd_io_saveState after device dynamics writes misc.c:326-339's format from
  device memory (cell order), byte for byte as tests/test_state_files.py
  composes it from the downloaded atoms.
* Resume: a run saved at step k, loaded into a fresh instance
  (fmd_io_loadState with its time, misc.c:121-266) and continued equals the
  same run continued without the round trip, bit for bit: the file carries
  positions and velocities in 17 significant digits (exact for FP64), the
  loader bins them in file order (the same cells and within-cell order), and
  both loops recompute forces before their first step (integrators.c:232).
  NVE and Berendsen, LJ argon and EAM copper.
* Laser: with no conduction and no coupling, the electrons of a deep slab gain
  the absorbed fluence F (1 - R) per unit area, the integral of the source
  S = I0 (1-R)/Lp exp(-(t-t0)^2/2s^2) exp(-(z-front)/Lp) over t and z
  (ttm.c:260-275, 794-816) -- the electron energy is sum over cells of
  gamma/2 (Te^2 - Te0^2) V with Ce = gamma Te.  The sum over cell centres and
  sub-steps differs from the integral by (h/Lp)^2/24 and the slab's depth
  (14 Lp), both well under the 1e-3 bar; a missing (1-R), 1/Lp or Gaussian
  normalisation sqrt(4 ln 2/pi) would be off by 6 % or more.
"""
import math

import numpy as np
import pytest

from fmd_amd import api
from fmd_amd.workloads import maxwell_velocities
from tests.pairsys import lattice_fcc
from tests.test_state_files import expected_state

pytestmark = pytest.mark.gpu


def _system(kind, rng):
    f = api.Fmd()
    if kind == "lj":
        a, n, names, m = 5.26, (6, 6, 6), ["Ar"], [39.948]
    else:
        a, n, names, m = 3.615, (7, 7, 7), ["Cu"], [63.546]
    L = [n[0] * a, n[1] * a, n[2] * a]
    f.fmd_box_setSize(*L)
    f.fmd_box_setPBC(True, True, True)
    f.set_kinds(names, m)
    return f, a, n, L, names, m


def _pots(f, kind, cu01):
    if kind == "lj":
        f.fmd_pot_lj_apply(0, 0, 3.405, 0.0104, 2.5 * 3.405)
    else:
        f.fmd_pot_apply(0, 0, f.fmd_pot_eam_alloy_load(cu01.encode()))


@pytest.mark.parametrize("kind", ["lj", "eam"])
@pytest.mark.parametrize("thermostat", [False, True])
def test_resume_equals_uninterrupted(rng, cu01, tmp_path, kind, thermostat):
    print("resume:", kind, "Berendsen" if thermostat else "NVE", flush=True)
    f, a, n, L, names, m = _system(kind, rng)
    _pots(f, kind, cu01)
    x = np.mod(lattice_fcc(n, a, rng, jitter=0.05), L[0])
    v = maxwell_velocities(len(x), m[0], 400.0, rng)
    f.load_atoms(x, v)
    dt = 2e-3 if kind == "eam" else 1e-3

    def run(g, steps):
        if thermostat:
            g.fmd_dync_equilibrate(api.GROUP_ALL, (steps - 0.5) * dt, dt, 0.05, 300.0)
        else:
            g.fmd_dync_integrate(api.GROUP_ALL, (steps - 0.5) * dt, dt)

    run(f, 15)
    f.fmd_io_setSaveDirectory((str(tmp_path) + "/").encode())
    f.fmd_io_saveState(b"k.stt")
    # the file is the reference's format, written from device memory
    atoms = f.atoms()
    assert open(tmp_path / "k.stt").read() == expected_state(f.fmd_dync_getTime(), L, [1, 1, 1], names, atoms)
    g, *_ = _system(kind, rng)
    _pots(g, kind, cu01)
    g.fmd_io_loadState(str(tmp_path / "k.stt").encode(), True)
    assert g.fmd_dync_getTime() == f.fmd_dync_getTime()
    run(f, 25)
    run(g, 25)
    af, ag = f.atoms(), g.atoms()
    assert g.fmd_dync_getTime() == f.fmd_dync_getTime()
    assert np.array_equal(af["x"], ag["x"]) and np.array_equal(af["v"], ag["v"]) and np.array_equal(af["f"], ag["f"])
    assert f.fmd_matt_getTotalEnergy() == g.fmd_matt_getTotalEnergy()
    f.free()
    g.free()


def test_laser_deposits_the_absorbed_fluence(rng, cu01):
    a = 3.615
    Lp, R, F, tau, t0 = 10e-9, 0.3, 0.05, 100e-15, 300e-15   # m, -, J/m^2, s, s
    nz_cells, h = 400, 4.0                                   # turi cells along z, A each
    nx = 5
    depth_cells = 388                                        # FCC cells along z: 14 Lp
    front = 40.0                                             # a turi cell boundary
    L = [nx * a, nx * a, nz_cells * h]
    x = lattice_fcc((nx, nx, depth_cells), a, origin=(0.0, 0.0, front - 0.25 * a))
    assert abs(np.min(x[:, 2]) - front) < 1e-12 and np.max(x[:, 2]) < L[2] - 10.0
    v = maxwell_velocities(len(x), 63.546, 10.0, rng)   # Ti > 0: xi = G V (Te - Ti) / (3 kB n Ti) with G = 0
    print("laser pin: %d atoms, slab depth %.0f A, Lp %.0f A" % (len(x), depth_cells * a, Lp * 1e10), flush=True)
    f = api.Fmd()
    f.fmd_box_setSize(*L)
    f.fmd_box_setPBC(True, True, False)
    f.set_kinds(["Cu"], [63.546])
    f.fmd_pot_apply(0, 0, f.fmd_pot_eam_alloy_load(cu01.encode()))
    f.load_atoms(x, v)
    turi = f.fmd_turi_add(api.TURI_TTM_TYPE1, 1, 1, nz_cells, 0.0, -1.0)
    gamma, Te0 = 96.6, 300.0
    f._fmd_ttm_setHeatCapacity_linear(turi, api.HeatCapLinear(gamma))
    f._fmd_ttm_setHeatConductivity_constant1(turi, 0.0)
    f._fmd_ttm_setCouplingFactor_constant1(turi, 0.0)
    f.fmd_ttm_setElectronTemperature(turi, Te0)
    f.fmd_ttm_setTimestepRatio(turi, 50)
    f.fmd_ttm_setCellActivationFraction(turi, 0.1)
    f._fmd_ttm_setLaserSource_gaussian(turi, api.LaserGaussian(F, R, t0, tau, Lp))
    f.fmd_dync_integrate(api.GROUP_ALL, 0.7, 1e-3)            # the source acts for |t - t0| < 5.66 sigma
    te = f.field_array(turi, f.fmd_field_find(turi, api.FIELD_TTM_TE))
    act = te >= 0
    assert np.sum(act) >= depth_cells * a / h - 2
    vol = L[0] * L[1] * h * 1e-30                              # m^3
    energy = float(np.sum(0.5 * gamma * (te[act] ** 2 - Te0 ** 2))) * vol   # J
    absorbed = F * (1 - R) * L[0] * L[1] * 1e-20                # J
    # the sum over cell centres: (h/Lp) / (2 sinh(h / 2Lp)) of the integral, and exp(-depth/Lp) lost
    assert energy / absorbed == pytest.approx(1.0, abs=1e-3), energy / absorbed
    assert np.max(te) > Te0 + 10.0
    f.free()
