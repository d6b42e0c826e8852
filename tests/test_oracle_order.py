"""Pins of the oracle's re-binning order, integrators, step counts and TTM
sub-step, each derived by hand from the reference's definitions
(misc.c:268-308, cell.c:151-160, md-ghost.c:827-869, integrators.c:40-126,
ttm.c:250-312) rather than recomputed by the oracle's own code."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests.pairsys import build

L = 30.0          # 3 x 3 x 3 cells of 10 A (LJ cutoff 9 A)
H = 10.0


def _sys(x, v, pbc=(1, 1, 1)):
    o, _ = build([L] * 3, list(pbc), [39.948], ["Ar"], [("lj", 0, 0, 3.0, 0.001, 9.0)], np.array(x),
                 v=np.array(v, dtype=np.float64), device=False)
    return o


def _cells(o):
    a = o.atoms()
    out, k = {}, 0
    for c, n in enumerate(a["cell_count"]):
        out[c] = [int(i) for i in a["atomid"][k:k + n]]
        k += n
    return out


def center(ix, iy, iz, dx=0.0, dy=0.0, dz=0.0):
    return [ix * H + 5.0 + dx, iy * H + 5.0 + dy, iz * H + 5.0 + dz]


def flat(ix, iy, iz):
    return (ix * 3 + iy) * 3 + iz


def test_swap_remove_order_single_cell():
    # cell 0 holds A B C D (ids 0..3); B and D move to the centre cell 13
    x = [center(0, 0, 0, -2, 0, 0), center(0, 0, 0, -1, 0, 0), center(0, 0, 0, 0, 0, 0), center(0, 0, 0, 1, 0, 0),
         center(1, 1, 1)]
    v = np.zeros((5, 3))
    v[1] = [1000.0, 1000.0, 1000.0]    # dt = 1e-2 -> +10 A in each dimension
    v[3] = [1000.0, 1000.0, 1000.0]
    o = _sys(x, v)
    o.vv_start(1e-2)
    c = _cells(o)
    # swap-with-last: B leaves, D fills slot 1 and leaves, C fills slot 1 -> [A, C]
    assert c[0] == [0, 2]
    # arrivals are appended in detection order after the cell's own atom
    assert c[flat(1, 1, 1)] == [4, 1, 3]


def test_migration_order_two_dimensions():
    # destination cell (2,2,0) starts with atom E (id 0)
    x = [center(2, 2, 0),
         center(0, 0, 0),           # P (id 1): crosses x<0 and y<0 -> ghost corner -> (2,2,0)
         center(0, 2, 0),           # R (id 2): crosses x<0 -> (2,2,0)
         center(2, 1, 0)]           # Q (id 3): moves +y directly into (2,2,0), lower source index
    v = np.zeros((4, 3))
    v[1] = [-800.0, -800.0, 0.0]
    v[2] = [-800.0, 0.0, 0.0]
    v[3] = [0.0, 800.0, 0.0]
    o = _sys(x, v)
    o.vv_start(1e-2)
    c = _cells(o)
    # Q arrives during the scan (before migration), R in the x phase, P in the y phase
    assert c[flat(2, 2, 0)] == [0, 3, 2, 1]
    a = o.atoms()
    p = a["x"][list(a["atomid"]).index(1)]
    assert p[0] == pytest.approx(5.0 - 8.0 + L) and p[1] == pytest.approx(5.0 - 8.0 + L)


def test_open_faces_delete():
    x = [center(1, 1, 0), center(1, 1, 2), center(1, 1, 1)]
    v = np.zeros((3, 3))
    v[0] = [0, 0, -800.0]
    v[1] = [0, 0, 800.0]
    o = _sys(x, v, pbc=(1, 1, 0))
    o.vv_start(1e-2)
    assert o.natoms() == 1
    assert _cells(o)[flat(1, 1, 1)] == [2]


def test_free_flight_and_berendsen_lambda(rng):
    # atoms farther apart than the cutoff: F = 0, so one start step is x + dt v exactly
    x = np.array([center(i % 3, (i // 3) % 3, i // 9) for i in range(27)])
    v = rng.normal(0, 3.0, size=(27, 3))
    o = _sys(x, v)
    o.forces()
    assert np.all(o.atoms()["f"] == 0.0)
    o.thermo()
    m = 39.948 / O.MASS_UNIT
    T = m * np.sum(v * v) / (3 * 27 * O.KB)
    assert o.energies()["T"] == pytest.approx(T, rel=1e-14)
    dt, tau, T0 = 2e-3, 0.05, 300.0
    lam = math.sqrt(1 + dt / tau * (T0 / T - 1))
    a0 = o.atoms()
    o.vv_start(dt, thermostat=1, tau=tau, T0=T0)
    a1 = o.atoms()
    order = np.argsort(a0["atomid"])
    v0 = a0["v"][order]
    v1 = a1["v"][np.argsort(a1["atomid"])]
    x1 = a1["x"][np.argsort(a1["atomid"])]
    assert np.allclose(v1, lam * v0, rtol=1e-15, atol=0)
    assert np.allclose(x1, a0["x"][order] + dt * lam * v0, rtol=1e-15, atol=1e-13)


@pytest.mark.parametrize("duration,dt,steps", [(1.0, 1e-3, 1000), (2.0, 1e-3, 2001), (4.0, 1e-3, 4001),
                                               (1.0, 2e-3, 500), (0.1, 2e-3, 50)])
def test_step_count_from_accumulated_time(duration, dt, steps):
    o = _sys([center(1, 1, 1)], [[0.0, 0.0, 0.0]])
    n, t = o.integrate_vv(0.0, duration, dt)
    # closed form: the loop runs while sum of dt (IEEE double accumulation) < duration
    acc, k = 0.0, 0
    while acc < duration:
        acc += dt
        k += 1
    assert n == k == steps


def test_timer_predicate():
    assert O.timer_is_its_time(0.05, 0.001, 0.0, 0.05)
    assert O.timer_is_its_time(0.0999, 0.001, 0.0, 0.05)
    assert not O.timer_is_its_time(0.0975, 0.001, 0.0, 0.05)


def test_ttm_type1_single_substep(rng):
    # slab of atoms filling turi cells 2..3 of 4 along z; cells 1 and 4 empty (deactivated)
    Lz = 40.0
    xs = []
    for iz in (1, 2):
        for k in range(60):
            xs.append([rng.uniform(0, L), rng.uniform(0, L), iz * 10.0 + rng.uniform(0.5, 9.5)])
    x = np.array(xs)
    v = rng.normal(0, 5.0, size=x.shape)
    o, _ = build([L, L, Lz], [1, 1, 0], [63.546], ["Cu"], [("lj", 0, 0, 2.0, 0.0001, 2.5)], x, v=v, device=False)
    cfg = O.TtmCfg()
    cfg.type, cfg.ncz, cfg.ratio = 1, 4, 1
    cfg.C_gamma, cfg.K, cfg.G = 96.6 * 6.2415091259e-12, 400.0 * 6.2415091259e-4, 1e17 * 6.2415091259e-24
    cfg.act_frac, cfg.start, cfg.stop = 0.1, 0.0, -1.0
    cfg.laser_tdiff = -1.0
    o.ttm_init(cfg, 500.0)
    o.thermo()
    o.ttm_get_ready()
    dt = 1e-3
    o.ttm_update(dt, dt)
    r = o.ttm_arrays(6)
    m = 63.546 / O.MASS_UNIT
    tcz = Lz / 4
    Te2 = {}
    for i in (2, 3):
        sel = ((x[:, 2] / tcz).astype(int) + 1) == i
        vc = v[sel].mean(axis=0)
        Ti = m * np.sum((v[sel] - vc) ** 2) / (3 * O.KB * sel.sum())
        assert r["Ti"][i] == pytest.approx(Ti, rel=1e-12)
        # neighbours outside the slab are inactive (zero-flux), ttm.c:279-285
        lo = i if i == 2 else i - 1
        hi = i if i == 3 else i + 1
        Te = 500.0
        Te2[i] = Te + dt / (cfg.C_gamma * Te) * (cfg.K * (Te - 2 * Te + Te) / tcz ** 2 - cfg.G * (Te - Ti))
        assert r["Te"][i] == pytest.approx(Te2[i], rel=1e-13)
        xi = cfg.G * (L * L * tcz) * (Te2[i] - Ti) / (3 * O.KB * sel.sum() * Ti)
        assert r["xi"][i] == pytest.approx(xi, rel=1e-11)
    assert r["Te"][1] == -1.0 and r["Te"][4] == -1.0


def test_ttm_activation_order_hand_derived():
    """Cell activation and deactivation of the 1-D presolve (ttm.c:194-230), on
    injected fields, with K = G = 0 and no source so that the sub-step leaves Te
    as the presolve set it.  The loop runs upwards: a cell activated in this pass
    lends its new Te to the next cell, while the next cell's old Te is used
    even if that cell is deactivated later in the same pass.  Hand-derived:
      cell 1: active, kept                         1000
      cell 2: inactive, 30 atoms: from cell 1      (1000*20)/20 = 1000
      cell 3: inactive, 40 atoms: cells 2 (new) and 4 (old, 2 atoms)
                                                   (1000*30 + 2000*2)/32 = 1062.5
      cell 4: 2 atoms < 5: deactivated             -1
      cell 5: 3 atoms: stays inactive              -1
      cell 6: inactive, 60 atoms, no active neighbour: Ti = 350"""
    cfg = O.TtmCfg()
    cfg.type, cfg.ncz, cfg.ratio = 1, 6, 1
    cfg.C_gamma, cfg.K, cfg.G = 96.6 * 6.2415091259e-12, 0.0, 0.0
    cfg.act_frac, cfg.start, cfg.stop = 0.1, 0.0, -1.0
    cfg.laser_tdiff = -1.0
    cfg.laser_t0 = 1e9
    o = O.Oracle([10.0, 10.0, 60.0], [1, 1, 0], [63.546], ["Cu"])
    o.ttm_init(cfg, 300.0)
    Te = np.array([-1.0, 1000.0, -1.0, -1.0, 2000.0, -1.0, -1.0, -1.0])
    Ti = np.array([0.0, 300.0, 310.0, 320.0, 330.0, 340.0, 350.0, 0.0])
    num = np.array([0, 20, 30, 40, 2, 3, 60, 0], np.uint32)
    o.ttm_set_fields(Te, Ti, num, frontsurf=0.0, min_atoms=5)
    o.ttm_presolve_solve(0.0, 1e-3)
    got = o.ttm_arrays(8)["Te"]
    assert list(got[1:7]) == [1000.0, 1000.0, 1062.5, -1.0, -1.0, 350.0]
