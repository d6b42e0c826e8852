"""The device TTM grid solver (ncx x ncy x ncz type-1 turies, this build's
extension of the reference's 1-D turi) against the oracle, through the
public API: a 2 x 2 x 20 turi with kappa = K0 Te / Tl heated by the laser,
and the grid solver on a 1 x 1 x 20 turi against the reference 1-D scheme."""
import numpy as np
import pytest

from fmd_amd import api
from fmd_amd.workloads import maxwell_velocities
from oracle.oracle import Oracle
from tests.pairsys import lattice_fcc, rel_err
from tests.test_oracle_ttm3d import cfg

pytestmark = pytest.mark.gpu

LASER = dict(fluence=4e4, reflectance=0.85, duration=100e-15, t0=5e-15, AbsorptionDepth=14e-9)


def _run(rng, cu01, dims, proportional, grid_solver=False):
    lp = 3.6316
    L = [6 * lp, 6 * lp, 40 * lp]
    x = lattice_fcc((6, 6, 30), lp, rng, jitter=0.02, origin=(0.0, 0.0, 5.0 * lp))
    v = maxwell_velocities(len(x), 63.546, 300.0, rng)
    f = api.Fmd()
    f.fmd_box_setSize(*L)
    f.fmd_box_setPBC(True, True, False)
    f.set_kinds(["Cu"], [63.546])
    pot = f.fmd_pot_eam_alloy_load(cu01.encode())
    f.fmd_pot_apply(0, 0, pot)
    f.load_atoms(x, v)
    f.fmd_ext_distribute()
    turi = f.fmd_turi_add(api.TURI_TTM_TYPE1, dims[0], dims[1], dims[2], 0.0, -1.0)
    assert turi >= 0
    f._fmd_ttm_setHeatCapacity_linear(turi, api.HeatCapLinear(96.6))
    if proportional:
        f.fmd_ext_ttm_setHeatConductivity_proportional(turi, 401.0)
    else:
        f._fmd_ttm_setHeatConductivity_constant1(turi, 400.0)
    if grid_solver:
        f.fmd_ext_ttm_useGridSolver(turi, True)
    f._fmd_ttm_setCouplingFactor_constant1(turi, 1e17)
    f.fmd_ttm_setElectronTemperature(turi, 300.0)
    # kappa0 / (gamma Tl) does not fall as Te rises: 200 sub-steps keep the explicit step stable
    f.fmd_ttm_setTimestepRatio(turi, 200 if proportional else 20)
    f.fmd_ttm_setCellActivationFraction(turi, 0.1)
    f._fmd_ttm_setLaserSource_gaussian(turi, api.LaserGaussian(LASER["fluence"], LASER["reflectance"],
                                                                LASER["t0"], LASER["duration"],
                                                                LASER["AbsorptionDepth"]))
    f.fmd_dync_integrate(api.GROUP_ALL, 0.01, 1e-3)
    o = Oracle(L, [1, 1, 0], [63.546], ["Cu"])
    o.load_eam(cu01)
    o.apply_eam(0, 0)
    o.make_grid()
    o.add_atoms(x, v)
    o.distribute()
    o.thermo()
    las = dict(fluence=LASER["fluence"], R=LASER["reflectance"], duration=LASER["duration"], t0=LASER["t0"],
               depth=LASER["AbsorptionDepth"])
    o.ttm_init(cfg(dims[2], ratio=200 if proportional else 20, K=401.0 if proportional else 400.0, ncx=dims[0], ncy=dims[1],
                   kmode=1 if proportional else 0, solver3d=1 if grid_solver else 0, laser=las), 300.0)
    steps, _ = o.integrate_se(0.0, 0.01, 1e-3)
    assert steps == 10
    return f, turi, o, L


def _compare(f, turi, o, L, dims):
    nx, ny, nz = dims
    r = o.ttm_arrays(nx * ny * (nz + 2))
    rte = r["Te"].reshape(nx * ny, nz + 2)[:, 1:nz + 1].ravel()
    rn = r["num"].reshape(nx * ny, nz + 2)[:, 1:nz + 1].ravel()
    te = f.field_array(turi, f.fmd_field_find(turi, api.FIELD_TTM_TE))
    n = f.field_array(turi, f.fmd_field_find(turi, api.FIELD_NUMBER), np.uint32)
    assert np.array_equal(n, rn)
    act = rte >= 0
    assert np.array_equal(te < 0, ~act)
    assert rel_err(te[act], rte[act]) < 1e-10
    assert np.all(np.isfinite(rte)) and np.max(rte[act]) > 400.0   # the laser heated the electrons
    ga, ra = f.atoms(), o.atoms()
    assert np.array_equal(ga["atomid"], ra["atomid"])
    assert rel_err(ga["x"], ra["x"], scale=L[2]) < 1e-12
    assert rel_err(ga["v"], ra["v"]) < 1e-9


def test_ttm_grid_proportional_conductivity(rng, cu01):
    dims = (2, 2, 20)
    f, turi, o, L = _run(rng, cu01, dims, proportional=True)
    _compare(f, turi, o, L, dims)
    f.free()


def test_ttm_grid_solver_reduces_to_1d(rng, cu01):
    dims = (1, 1, 20)
    f, turi, o, L = _run(rng, cu01, dims, proportional=False, grid_solver=True)
    _compare(f, turi, o, L, dims)
    f.free()


def test_device_activation_order_matches_oracle(rng):
    """The device presolve (one lane, cells in order) activates and deactivates
    turi cells exactly as ttm.c:194-230: atoms placed so that the six cells hold
    20, 30, 40, 2, 3 and 60 atoms, Te injected as 1000, -, -, 2000, -, - (K = G
    = 0, no source, so the sub-step keeps what the presolve set); expected
    1000, 1000, 1062.5, -1, -1, Ti of cell 6 (hand-derived in
    test_oracle_order.py), and the oracle on the device's own fields agrees."""
    import fmd_amd as F
    from tests.pairsys import build
    counts = [20, 30, 40, 2, 3, 60]
    xs = []
    for k, n in enumerate(counts):
        for _ in range(n):
            xs.append([rng.uniform(0.5, 9.5), rng.uniform(0.5, 9.5), k * 10.0 + rng.uniform(0.5, 9.5)])
    x = np.array(xs)
    v = rng.normal(0.0, 5.0, size=x.shape)
    _, d = build([10.0, 10.0, 60.0], [1, 1, 0], [63.546], ["Cu"], [("lj", 0, 0, 2.0, 0.0001, 2.5)], x, v=v)
    c = F.TtmCfg()
    c.type, c.ncz, c.ratio = 1, 6, 1
    c.tcellh_z, c.tcell_volume = 10.0, 1000.0
    c.C_gamma, c.K, c.G = 96.6 * 6.2415091259e-12, 0.0, 0.0
    c.laser_t0, c.laser_tdiff = 1e9, -1.0
    c.frontsurf, c.min_atoms = 0.0, 5
    c.ncx = c.ncy = 1
    te0 = np.array([-1.0, 1000.0, -1.0, -1.0, 2000.0, -1.0, -1.0, -1.0])
    d.ttm_setup(c, te0)
    d.ttm_update(0.0, 1e-3)
    r = d.ttm_download(8)
    assert list(r["num"][1:7]) == counts
    assert list(r["Te"][1:7]) == [1000.0, 1000.0, 1062.5, -1.0, -1.0, r["Ti"][6]]
    oc = cfg(6, ratio=1, K=0.0, G=0.0)
    o = Oracle([10.0, 10.0, 60.0], [1, 1, 0], [63.546], ["Cu"])
    o.ttm_init(oc, 300.0)
    o.ttm_set_fields(te0, r["Ti"], r["num"], frontsurf=0.0, min_atoms=5)
    o.ttm_presolve_solve(0.0, 1e-3)
    assert np.array_equal(o.ttm_arrays(8)["Te"][1:7], r["Te"][1:7])
