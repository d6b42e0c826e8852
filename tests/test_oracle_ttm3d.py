"""Pins of the oracle's ncx x ncy x ncz TTM solver (this build's extension of
the reference's 1-D turi; the reference has no 3-D solver, ttm.c:451-452):

* on a 1 x 1 x N turi the extension reproduces the reference 1-D type-1
  path bit for bit, atoms and fields, along a symplectic-Euler run;
* columns of an x,y-uniform ncx x ncy x N field evolve exactly like the
  1 x 1 x N field (both conductivity modes);
* kappa = K Te / Ti with Ce = gamma Te is the reference's type-2 scheme with
  a linear Ce table, a constant G table and Zhigilei A = 0
  (kappa = Ce v^2 / (3 B Ti)); the two agree to rounding;
* with G = 0 and no source the flux form conserves sum(gamma Te dTe) to
  rounding (telescoping face fluxes, periodic x,y, inactive z boundaries).

and of the reference's 1-D type-2 solver with Zhigilei conductivity (A != 0),
using example 04's Ce and G tables:

* one sub-step balances sum Ce dTe against the coupling and the laser source
  (the face fluxes cancel), with Ce and G from scipy's natural splines;
* two cells exchange heat at the mean of the kinetic-theory conductivities
  Ce v^2 / (3 (A Te^2 + B Ti)).
"""
import math

import numpy as np
import pytest

from oracle.oracle import Oracle, TtmCfg
from tests.pairsys import lattice_fcc
from fmd_amd.workloads import maxwell_velocities

J_M3K2, W_MK, W_M3K = 6.2415091259e-12, 6.2415091259e-4, 6.2415091259e-24


def cfg(ncz, ratio=10, gamma=96.6, K=400.0, G=1e17, ncx=1, ncy=1, kmode=0, solver3d=0, laser=None):
    c = TtmCfg()
    c.type, c.ncz, c.ratio = 1, ncz, ratio
    c.C_gamma, c.K, c.G = gamma * J_M3K2, K * W_MK, G * W_M3K
    c.act_frac, c.start, c.stop = 0.1, 0.0, -1.0
    c.ncx, c.ncy, c.kmode, c.solver3d = ncx, ncy, kmode, solver3d
    if laser:
        Lp = laser["depth"] * 1e10
        I0 = math.sqrt(4 * math.log(2) / math.pi) * laser["fluence"] / laser["duration"] * 6.2415091259e-14
        sig = laser["duration"] / math.sqrt(8 * math.log(2)) * 1e12
        c.laser_t0 = laser["t0"] * 1e12
        c.laser_tdiff = 4 * math.sqrt(2) * sig
        c.laser_2sig2_inv = -0.5 / (sig * sig)
        c.laser_absdepth_inv = -1.0 / Lp
        c.laser_factor = I0 * (1 - laser["R"]) / Lp
    else:
        c.laser_t0, c.laser_tdiff = 1e9, 0.0
    return c


LASER = dict(fluence=4e4, R=0.85, duration=100e-15, t0=5e-15, depth=14e-9)


def _slab(rng, cu01):
    lp = 3.6316
    L = [5 * lp, 5 * lp, 30 * lp]
    x = lattice_fcc((5, 5, 22), lp, rng, jitter=0.02, origin=(0.0, 0.0, 4.0 * lp))
    v = maxwell_velocities(len(x), 63.546, 300.0, rng)
    return L, x, v


def _oracle(L, x, v, cu01):
    o = Oracle(L, [1, 1, 0], [63.546], ["Cu"])
    o.load_eam(cu01)
    o.apply_eam(0, 0)
    o.make_grid()
    o.add_atoms(x, v)
    o.distribute()
    o.thermo()
    return o


def test_extension_reduces_to_reference_1d(rng, cu01):
    L, x, v = _slab(rng, cu01)
    res = []
    for forced in (0, 1):
        o = _oracle(L, x, v, cu01)
        o.ttm_init(cfg(15, laser=LASER, solver3d=forced), 300.0)
        steps, _ = o.integrate_se(0.0, 0.006, 1e-3)
        assert steps == 6
        res.append((o.ttm_arrays(17), o.atoms()))
    (ra, aa), (rb, ab) = res
    for k in ("Te", "xi", "num"):
        assert np.array_equal(ra[k][1:16], rb[k][1:16]), k
    for k in ("x", "v", "f"):
        assert np.array_equal(aa[k], ab[k]), k
    assert np.any(ra["Te"][1:16] > 300.0)   # the laser heated the electrons


def _inject_solve(o, Te, Ti, num, ratio_dt=1e-3):
    o.ttm_set_fields(Te, Ti, num, frontsurf=10.0, min_atoms=5)
    o.ttm_solve_only(0.0, ratio_dt)


def _profile(nz, rng):
    Te = np.full(nz + 2, -1.0)
    Te[1:nz + 1] = rng.uniform(300.0, 3000.0, nz)
    Te[1] = Te[nz] = -1.0            # vacuum ends
    Ti = np.zeros(nz + 2)
    Ti[1:nz + 1] = rng.uniform(250.0, 600.0, nz)
    num = np.zeros(nz + 2, np.uint32)
    num[2:nz] = rng.integers(20, 40, nz - 2)
    return Te, Ti, num


@pytest.mark.parametrize("kmode", [0, 1])
def test_uniform_columns_equal_one_column(rng, kmode):
    nz, nx, ny = 12, 3, 2
    Te, Ti, num = _profile(nz, rng)
    L = [20.0, 18.0, 240.0]          # dz = 20 A: the explicit sub-step is stable
    o1 = Oracle(L, [1, 1, 0], [63.546], ["Cu"])
    o1.ttm_init(cfg(nz, ratio=25, kmode=kmode, solver3d=1, laser=LASER), 300.0)
    _inject_solve(o1, Te, Ti, num)
    r1 = o1.ttm_arrays(nz + 2)
    o3 = Oracle([L[0] * nx, L[1] * ny, L[2]], [1, 1, 0], [63.546], ["Cu"])   # same turi-cell volume
    o3.ttm_init(cfg(nz, ratio=25, kmode=kmode, ncx=nx, ncy=ny, laser=LASER), 300.0)
    _inject_solve(o3, np.tile(Te, nx * ny), np.tile(Ti, nx * ny), np.tile(num, nx * ny))
    r3 = o3.ttm_arrays(nx * ny * (nz + 2))
    for col in range(nx * ny):
        s = slice(col * (nz + 2), (col + 1) * (nz + 2))
        assert np.array_equal(r3["Te"][s], r1["Te"]), col
        assert np.array_equal(r3["xi"][s], r1["xi"]), col
    assert np.all(np.isfinite(r1["Te"])) and np.any(r1["Te"][2:nz] != Te[2:nz])


def test_proportional_conductivity_is_type2_special_case(rng):
    nz = 14
    Te, Ti, num = _profile(nz, rng)
    L = [20.0, 20.0, 280.0]
    gamma, K0, G = 96.6, 401.0, 1e17
    # type 2: Ce(T) = gamma T as a table, G(T) = G, Zhigilei Ke = Ce v2 / (3 (A Te^2 + B Ti)) with A = 0
    T = np.linspace(10.0, 60000.0, 400)
    B, v2 = 1.0, 3.0 * K0 * W_MK / (gamma * J_M3K2)       # then Ke = K0 Te / Ti
    c2 = cfg(nz, ratio=30, laser=LASER)
    c2.type = 2
    c2.Kzh_A, c2.Kzh_B, c2.Kzh_v2 = 0.0, B, v2
    Ct, Gt = np.ascontiguousarray(gamma * J_M3K2 * T), np.full(len(T), G * W_M3K)
    Tc = np.ascontiguousarray(T)
    keep = (Tc, Ct, Gt)
    c2.nC = c2.nG = len(T)
    c2.T_C = c2.T_G = Tc.ctypes.data
    c2.Ct, c2.Gt = Ct.ctypes.data, Gt.ctypes.data
    o2 = Oracle(L, [1, 1, 0], [63.546], ["Cu"])
    o2.ttm_init(c2, 300.0)
    _inject_solve(o2, Te, Ti, num)
    r2 = o2.ttm_arrays(nz + 2)
    op = Oracle(L, [1, 1, 0], [63.546], ["Cu"])
    op.ttm_init(cfg(nz, ratio=30, gamma=gamma, K=K0, G=G, kmode=1, laser=LASER), 300.0)
    _inject_solve(op, Te, Ti, num)
    rp = op.ttm_arrays(nz + 2)
    act = r2["Te"] >= 0
    assert np.array_equal(act, rp["Te"] >= 0)
    d = np.abs(rp["Te"][act] - r2["Te"][act]) / np.abs(r2["Te"][act])
    assert d.max() < 1e-12
    assert np.all(np.isfinite(rp["Te"])) and np.max(np.abs(rp["Te"][act] - Te[act])) > 1.0
    del keep


@pytest.mark.parametrize("kmode", [0, 1])
def test_flux_form_conserves_electron_energy(rng, kmode):
    nx, ny, nz = 3, 4, 9
    n = nx * ny * (nz + 2)
    Te = rng.uniform(300.0, 5000.0, n)
    Ti = rng.uniform(280.0, 400.0, n)
    num = np.full(n, 30, np.uint32)
    for c in range(nx * ny):      # inactive z ghosts: zero-flux faces
        Te[c * (nz + 2)] = Te[c * (nz + 2) + nz + 1] = -1.0
        num[c * (nz + 2)] = num[c * (nz + 2) + nz + 1] = 0
    o = Oracle([30.0, 40.0, 90.0], [1, 1, 0], [63.546], ["Cu"])
    o.ttm_init(cfg(nz, ratio=1, G=0.0, kmode=kmode, ncx=nx, ncy=ny), 300.0)
    _inject_solve(o, Te, Ti, num, ratio_dt=2e-6)
    Te2 = o.ttm_arrays(n)["Te"]
    act = Te >= 0
    de = Te[act] * (Te2[act] - Te[act])            # gamma V / dt times the flux sum
    assert abs(de.sum()) < 1e-11 * np.abs(de).sum()
    assert np.abs(de).sum() > 0


def _table(path, tscale, vscale):
    d = np.loadtxt(path)
    return np.ascontiguousarray(d[:, 0] * tscale), np.ascontiguousarray(d[:, 1] * vscale)


@pytest.mark.parametrize("with_source", [False, True])
def test_type2_zhigilei_sub_step_conserves_energy(rng, data_dir, with_source):
    """Reference type 2 with example 04's tables and Zhigilei constants (A != 0):
    the face conductivities are shared by both cells of a face (ttm.c:341-348)
    and inactive neighbours give zero flux, so one explicit sub-step satisfies
        sum_i Ce_i (Te'_i - Te_i) = dt sum_i (-G_i (Te_i - Ti_i) + S_i)
    whatever Ke is; Ce_i and G_i are evaluated here with scipy's natural cubic
    splines of the tables, independently of the oracle.  A face conductivity
    taken from one side only, a flux with a wrong index or a term divided by
    the wrong Ce breaks the identity."""
    from scipy.interpolate import CubicSpline
    import os
    T_C, Ct = _table(os.path.join(data_dir, "Ce_Cu.dat"), 1e4, 1e5 * J_M3K2)
    T_G, Gt = _table(os.path.join(data_dir, "G_Cu.dat"), 1e4, 1e17 * W_M3K)
    nz = 24
    Te = np.full(nz + 2, -1.0)
    Te[2:nz] = rng.uniform(1000.0, 30000.0, nz - 2)
    Te[nz // 2] = -1.0                          # a deactivated cell inside the slab
    Ti = np.zeros(nz + 2)
    Ti[2:nz] = rng.uniform(300.0, 900.0, nz - 2)
    num = np.zeros(nz + 2, np.uint32)
    num[2:nz] = 30
    num[nz // 2] = 0
    c = cfg(nz, ratio=1, laser=LASER if with_source else None)
    c.type = 2
    c.Kzh_A, c.Kzh_B, c.Kzh_v2 = 2.66e6 / 1e12, 2.41e11 / 1e12, (1.57e6 * 1e-2) ** 2
    c.nC, c.nG = len(T_C), len(T_G)
    c.T_C, c.Ct, c.T_G, c.Gt = T_C.ctypes.data, Ct.ctypes.data, T_G.ctypes.data, Gt.ctypes.data
    keep = (T_C, Ct, T_G, Gt)
    o = Oracle([20.0, 20.0, 10.0 * (nz + 2)], [1, 1, 0], [63.546], ["Cu"])
    o.ttm_init(c, 300.0)
    dt, t = 2e-5, LASER["t0"] * 1e12
    o.ttm_set_fields(Te, Ti, num, frontsurf=10.0, min_atoms=5)
    o.ttm_solve_only(t if with_source else 0.0, dt)
    Te2 = o.ttm_arrays(nz + 2)["Te"]
    act = Te >= 0
    assert np.array_equal(act, Te2 >= 0)
    Ce = CubicSpline(T_C, Ct, bc_type="natural")(Te[act])
    G = CubicSpline(T_G, Gt, bc_type="natural")(Te[act])
    lhs = Ce * (Te2[act] - Te[act])
    rhs = -dt * G * (Te[act] - Ti[act])
    if with_source:
        # the reference's Gaussian pulse at the centre of each cell, ttm.c:327-337
        fs, _, h = o.ttm_params()
        z = (np.nonzero(act)[0] - 1 + 0.5) * h
        rhs = rhs + dt * c.laser_factor * np.exp(c.laser_2sig2_inv * (t - c.laser_t0) ** 2) * \
            np.exp(c.laser_absdepth_inv * (z - fs))
    scale = np.abs(lhs).sum()
    assert np.abs(lhs - rhs).max() > 1e-6 * scale      # cells exchange heat: the per-cell identity fails
    assert abs(lhs.sum() - rhs.sum()) < 1e-10 * scale  # the sum holds
    del keep


def test_type2_two_cell_exchange_follows_zhigilei_conductivity(rng, data_dir):
    """Two active cells: the heat passed across their face in one sub-step,
    (Ce_1 dTe_1 / dt + G_1 (Te_1 - Ti_1)) dz^2 / (Te_2 - Te_1), equals the mean
    of the kinetic-theory conductivities K = Ce v^2 tau / 3 of the two cells
    with the Zhigilei scattering rate 1/tau = A Te^2 + B Ti.  Checked at two
    values of A so that both scattering terms are exercised."""
    from scipy.interpolate import CubicSpline
    import os
    T_C, Ct = _table(os.path.join(data_dir, "Ce_Cu.dat"), 1e4, 1e5 * J_M3K2)
    T_G, Gt = _table(os.path.join(data_dir, "G_Cu.dat"), 1e4, 1e17 * W_M3K)
    ce = CubicSpline(T_C, Ct, bc_type="natural")
    gg = CubicSpline(T_G, Gt, bc_type="natural")
    B, v2 = 2.41e11 / 1e12, (1.57e6 * 1e-2) ** 2
    for A in (2.66e6 / 1e12, 4 * 2.66e6 / 1e12):
        Te = np.array([-1.0, 4000.0, 15000.0, -1.0])
        Ti = np.array([0.0, 400.0, 700.0, 0.0])
        num = np.array([0, 30, 30, 0], np.uint32)
        c = cfg(2, ratio=1)
        c.type = 2
        c.Kzh_A, c.Kzh_B, c.Kzh_v2 = A, B, v2
        c.nC, c.nG = len(T_C), len(T_G)
        c.T_C, c.Ct, c.T_G, c.Gt = T_C.ctypes.data, Ct.ctypes.data, T_G.ctypes.data, Gt.ctypes.data
        o = Oracle([20.0, 20.0, 40.0], [1, 1, 0], [63.546], ["Cu"])
        o.ttm_init(c, 300.0)
        dt = 1e-5
        o.ttm_set_fields(Te, Ti, num, frontsurf=0.0, min_atoms=5)
        o.ttm_solve_only(0.0, dt)
        Te2 = o.ttm_arrays(4)["Te"]
        h = o.ttm_params()[2]
        k_obs = (ce(Te[1]) * (Te2[1] - Te[1]) / dt + gg(Te[1]) * (Te[1] - Ti[1])) * h * h / (Te[2] - Te[1])
        kin = [ce(Te[i]) * v2 / 3.0 / (A * Te[i] ** 2 + B * Ti[i]) for i in (1, 2)]
        assert k_obs == pytest.approx(0.5 * (kin[0] + kin[1]), rel=1e-9)
        del c
