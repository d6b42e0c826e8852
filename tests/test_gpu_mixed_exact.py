"""The mixed build (FP32 pair arithmetic, FP64 integration) against the FP64
oracle where FP32 separations are not good enough on their own.

The synthetic setfl of the EAM Cu workloads ends at rc = 5.5 A without a taper:
the reference appends two zero knots after the last point (eam.c:444-471), so
the r^2 splines ring through the intervals below rc^2 with curvatures up to
1e5 x their median.  The mixed build evaluates the pairs that fall there, and
any pair within 1e-5 of rc^2, in FP64 from the FP64 positions.  Bar: the mixed
build's 1e-5 of max|F|.

* a 12^3 crystal jittered by 0.1 A, which puts pairs all over the last
  intervals (395 of 6912 atoms were over the bar before the exact shell);
* the TTM-MD slab in the form of its workload: mixed forces on the synthetic
  table, a 2 x 2 x 20 grid turi with kappa = K0 Te / Tl heated by a laser, and
  the friction term (forces.c:292-330) in the force pass, 10 symplectic Euler
  steps through the public API against the oracle (cell counts, Te, x, v), and
  the mixed force pass with friction against the FP64 one on the final state.
"""
import numpy as np
import pytest

import fmd_amd as F
from fmd_amd import api
from fmd_amd.workloads import maxwell_velocities
from oracle.oracle import Oracle
from tests.pairsys import build, lattice_fcc, rel_err
from tests.test_oracle_ttm3d import cfg

pytestmark = pytest.mark.gpu


def test_mixed_synthetic_table_jittered_crystal(rng, synth_setfl):
    print("mixed exact shell: 12^3 Cu, synthetic setfl, jitter 0.1 A", flush=True)
    x = lattice_fcc((12, 12, 12), 3.615, rng, jitter=0.1)
    L = [12 * 3.615] * 3
    o, d = build(L, [1, 1, 1], [63.546], ["Cu"], [("eam", synth_setfl, [(0, 0)])], x, precision=F.FMDX_PREC_MIXED)
    o.forces()
    d.forces()
    got, ref = d.download(), o.atoms()
    assert np.array_equal(got["atomid"], ref["atomid"])
    scale = np.max(np.abs(ref["f"]))
    err = np.max(np.abs(got["f"] - ref["f"]), axis=1) / scale
    assert np.max(err) < 1e-5, "max %.3e, %d atoms over 1e-5" % (np.max(err), np.sum(err > 1e-5))
    assert rel_err(got["aux"], ref["aux"]) < 1e-6
    t, e = d.read(), o.energies()
    assert t.err == 0 and t.exact_unmatched == 0
    assert abs(t.pe - e["pe"]) < 1e-6 * abs(e["pe"])


def test_mixed_ttm_slab_synthetic_table_with_friction(rng, synth_setfl):
    print("mixed TTM slab: synthetic setfl, 2x2x20 grid turi, kappa0 Te/Tl, laser, 10 SE steps", flush=True)
    lp = 3.615
    dims = (2, 2, 20)
    L = [6 * lp, 6 * lp, 40 * lp]
    x = lattice_fcc((6, 6, 30), lp, rng, jitter=0.1, origin=(0.0, 0.0, 5.0 * lp))
    x[:, 0] %= L[0]
    x[:, 1] %= L[1]
    v = maxwell_velocities(len(x), 63.546, 300.0, rng)
    laser = dict(fluence=4e4, R=0.85, duration=100e-15, t0=5e-15, depth=14e-9)
    f = api.Fmd()
    f.fmd_box_setSize(*L)
    f.fmd_box_setPBC(True, True, False)
    f.set_kinds(["Cu"], [63.546])
    f.fmd_pot_apply(0, 0, f.fmd_pot_eam_alloy_load(synth_setfl.encode()))
    f.load_atoms(x, v)
    f.fmd_ext_setPrecision(F.FMDX_PREC_MIXED)
    assert f.fmd_ext_distribute()
    turi = f.fmd_turi_add(api.TURI_TTM_TYPE1, dims[0], dims[1], dims[2], 0.0, -1.0)
    f._fmd_ttm_setHeatCapacity_linear(turi, api.HeatCapLinear(96.6))
    f.fmd_ext_ttm_setHeatConductivity_proportional(turi, 401.0)
    f._fmd_ttm_setCouplingFactor_constant1(turi, 1e17)
    f.fmd_ttm_setElectronTemperature(turi, 300.0)
    f.fmd_ttm_setTimestepRatio(turi, 200)
    f.fmd_ttm_setCellActivationFraction(turi, 0.1)
    f._fmd_ttm_setLaserSource_gaussian(turi, api.LaserGaussian(laser["fluence"], laser["R"], laser["t0"],
                                                                laser["duration"], laser["depth"]))
    f.fmd_dync_integrate(api.GROUP_ALL, 0.01, 1e-3)
    o = Oracle(L, [1, 1, 0], [63.546], ["Cu"])
    o.load_eam(synth_setfl)
    o.apply_eam(0, 0)
    o.make_grid()
    o.add_atoms(x, v)
    o.distribute()
    o.thermo()
    o.ttm_init(cfg(dims[2], ratio=200, K=401.0, ncx=dims[0], ncy=dims[1], kmode=1, laser=laser), 300.0)
    steps, _ = o.integrate_se(0.0, 0.01, 1e-3)
    assert steps == 10
    nx, ny, nz = dims
    r = o.ttm_arrays(nx * ny * (nz + 2))
    rte = r["Te"].reshape(nx * ny, nz + 2)[:, 1:nz + 1].ravel()
    rn = r["num"].reshape(nx * ny, nz + 2)[:, 1:nz + 1].ravel()
    rxi = r["xi"].reshape(nx * ny, nz + 2)[:, 1:nz + 1].ravel()
    te = f.field_array(turi, f.fmd_field_find(turi, api.FIELD_TTM_TE))
    n = f.field_array(turi, f.fmd_field_find(turi, api.FIELD_NUMBER), np.uint32)
    assert np.array_equal(n, rn)
    act = rte >= 0
    assert np.array_equal(te < 0, ~act)
    assert np.max(rte[act]) > 400.0 and np.max(np.abs(rxi)) > 0   # heated electrons, friction on
    assert rel_err(te[act], rte[act]) < 1e-6
    ga, ra = f.atoms(), o.atoms()
    assert np.array_equal(ga["atomid"], ra["atomid"])
    assert rel_err(ga["x"], ra["x"], scale=L[2]) < 1e-9
    assert rel_err(ga["v"], ra["v"]) < 1e-5
    # Forces on one state: the mixed force pass with the friction term (fields of the last TTM
    # update) against the FP64 pass on the same atoms and fields, which the FP64 tests pin to the
    # oracle.  The two trajectories above are not compared by force: on this table a pair in the
    # last intervals changes its force by more than the bar for a 1e-7 A move.
    fx = F.load_fmdx()
    dev = f.fmd_ext_getDevice()
    F._ck(fx.fmdx_update_forces(dev, -1, 1), "forces")
    fm = f.atoms()["f"]
    f.fmd_ext_setPrecision(F.FMDX_PREC_FP64)
    F._ck(fx.fmdx_update_forces(dev, -1, 1), "forces")
    f64 = f.atoms()["f"]
    assert np.array_equal(fm, ga["f"])   # the loop's last force pass, reproduced
    assert rel_err(fm, f64) < 1e-5
    f.free()
