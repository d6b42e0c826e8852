"""The batch engine on the GPU: every member advances exactly as it does on
its own through the single-system entry points (same kernels bodies, same
summation orders), whatever else is in the batch, with or without the
captured graph, across capacity growth, and along the oracle's trajectory."""
import numpy as np
import pytest

import fmd_amd as F
from fmd_amd.workloads import maxwell_velocities
from tests.pairsys import build, lattice_fcc, rel_err

pytestmark = pytest.mark.gpu

AR = dict(sig=3.405, eps=0.0104, rc=2.5 * 3.405, a=5.26, m=39.948)
CU_MORSE = dict(D0=0.3429, alpha=1.3588, r0=2.866, rc=7.0, a=3.615, m=63.546)


def _lj(rng, n=(6, 6, 6), T=300.0, precision=F.FMDX_PREC_FP64, pbc=(1, 1, 1)):
    x = lattice_fcc(n, AR["a"], rng, jitter=0.05)
    v = maxwell_velocities(len(x), AR["m"], T, rng)
    L = [n[0] * AR["a"], n[1] * AR["a"], n[2] * AR["a"]]
    spec = (L, list(pbc), [AR["m"]], ["Ar"], [("lj", 0, 0, AR["sig"], AR["eps"], AR["rc"])], x, v)
    return spec


def _morse(rng, n=(5, 5, 6), T=500.0, pbc=(1, 1, 0)):
    a = CU_MORSE["a"]
    x = lattice_fcc(n, a, rng, jitter=0.03, origin=(0.0, 0.0, 4.0))
    v = maxwell_velocities(len(x), CU_MORSE["m"], T, rng)
    L = [n[0] * a, n[1] * a, n[2] * a + 8.0]
    c = CU_MORSE
    spec = (L, list(pbc), [c["m"]], ["Cu"], [("morse", 0, 0, c["D0"], c["alpha"], c["r0"], c["rc"])], x, v)
    return spec


def _eam(rng, cu01, n=(6, 6, 6), T=600.0):
    a = 3.615
    x = lattice_fcc(n, a, rng, jitter=0.03)
    v = maxwell_velocities(len(x), 63.546, T, rng)
    L = [n[0] * a] * 3
    return (L, [1, 1, 1], [63.546], ["Cu"], [("eam", cu01, [(0, 0)])], x, v)


def _make(spec, precision):
    L, pbc, m, names, pots, x, v = spec
    o, d = build(L, pbc, m, names, pots, x, v=v, precision=precision)
    return o, d


def _same(a, b):
    ga, gb = a.download(), b.download()
    for k in ("cell_count", "atomid", "x", "v", "f"):
        assert np.array_equal(ga[k], gb[k]), k
    return ga


def _run_single(d, steps, dt, thermostat=0, tau=1.0, T0=0.0):
    d.forces()
    d.thermo()
    for _ in range(steps):
        d.vv_start(dt, thermostat=thermostat, tau=tau, T0=T0)
        d.forces()
        d.vv_finish(dt)


@pytest.mark.parametrize("precision", [F.FMDX_PREC_FP64, F.FMDX_PREC_MIXED])
def test_batch_equals_single_system(rng, cu01, precision):
    specs = [_lj(rng), _morse(rng), _eam(rng, cu01)]
    params = [(2e-3, 0, 1.0, 0.0), (1e-3, 1, 0.05, 300.0), (2e-3, 1, 0.1, 600.0)]
    batch = F.Batch(precision)
    members, singles = [], []
    for sp, pr in zip(specs, params):
        _, d = _make(sp, precision)
        _, e = _make(sp, precision)
        batch.add(d, *pr)
        members.append(d)
        singles.append(e)
    batch.step(12)
    for e, pr in zip(singles, params):
        _run_single(e, 12, *pr)
    st = batch.status()
    for d, e, t in zip(members, singles, st):
        _same(d, e)
        r = e.read()
        assert t.temperature == r.temperature and t.pe == r.pe and t.err == 0
    batch.close()


def test_batch_alone_vs_mix(rng, cu01):
    spec_a = _lj(rng, n=(7, 7, 7))
    others = [_morse(rng), _eam(rng, cu01), _lj(rng, n=(5, 5, 5), T=80.0)]
    _, a1 = _make(spec_a, F.FMDX_PREC_MIXED)
    _, a2 = _make(spec_a, F.FMDX_PREC_MIXED)
    alone = F.Batch(F.FMDX_PREC_MIXED)
    alone.add(a1, 1e-3)
    mix = F.Batch(F.FMDX_PREC_MIXED)
    keep = []
    for sp in others[:2]:
        _, d = _make(sp, F.FMDX_PREC_MIXED)
        mix.add(d, 1e-3, 1, 0.1, 400.0)
        keep.append(d)
    mix.add(a2, 1e-3)
    alone.step(10)
    mix.step(5)
    _, d = _make(others[2], F.FMDX_PREC_MIXED)   # a job joins mid-run
    mix.add(d, 2e-3)
    keep.append(d)
    mix.step(3)
    mix.remove(keep[0])                          # and another one leaves
    mix.step(2)
    _same(a1, a2)
    assert alone.status()[0].temperature == mix.status()[1].temperature
    alone.close(); mix.close()


def test_batch_graph_replays_match_eager(rng, cu01):
    specs = [_lj(rng), _eam(rng, cu01, n=(5, 5, 5))]
    g, e = F.Batch(F.FMDX_PREC_FP64), F.Batch(F.FMDX_PREC_FP64)
    e.use_graph(False)
    dg, de = [], []
    for sp in specs:
        _, d1 = _make(sp, F.FMDX_PREC_FP64)
        _, d2 = _make(sp, F.FMDX_PREC_FP64)
        g.add(d1, 2e-3, 1, 0.1, 500.0)
        e.add(d2, 2e-3, 1, 0.1, 500.0)
        dg.append(d1); de.append(d2)
    for _ in range(4):   # several replays, with synchronising host work in between
        g.step(3)
        e.step(3)
        for a, b in zip(dg, de):
            _same(a, b)
    g.close(); e.close()


def test_batch_grows_capacity(rng):
    # a hot, compressed LJ crystal: cells gain atoms as it relaxes
    spec = _lj(rng, n=(6, 6, 6), T=2500.0)
    _, d = _make(spec, F.FMDX_PREC_FP64)
    _, e = _make(spec, F.FMDX_PREC_FP64)
    mx = int(np.max(d.download()["cell_count"]))
    d.set_capacity(mx + 4)
    assert d.read().cap == mx + 4
    b = F.Batch(F.FMDX_PREC_FP64)
    b.set_check_interval(1)
    b.add(d, 2e-3)
    b.step(40)
    _run_single(e, 40, 2e-3)
    t = b.status()[0]
    assert t.err == 0
    assert t.cap > mx + 4 and t.cap % 32 == 0   # grown by the batch, to whole cache lines
    _same(d, e)
    b.close()


def test_batch_member_follows_oracle(rng):
    spec = _lj(rng, n=(8, 8, 8), T=120.0)
    o, d = _make(spec, F.FMDX_PREC_FP64)
    b = F.Batch(F.FMDX_PREC_FP64)
    b.add(d, 1e-3)
    b.step(25)
    steps, _ = o.integrate_vv(0.0, 0.025, 1e-3)
    assert steps == 25
    ga, ra = d.download(), o.atoms()
    assert np.array_equal(ga["cell_count"], ra["cell_count"])
    assert np.array_equal(ga["atomid"], ra["atomid"])
    assert rel_err(ga["x"], ra["x"], scale=spec[0][0]) < 1e-12
    assert rel_err(ga["v"], ra["v"]) < 1e-9
    en = o.energies()
    t = b.status()[0]
    assert t.temperature == pytest.approx(en["T"], rel=1e-10)
    assert t.pe == pytest.approx(en["pe"], rel=1e-10)
    b.close()


def _cluster(rng, pots, m, a, R, L, T, drift, name="Cu"):
    """an FCC sphere of radius R at the centre of an open L^3 box, moving at `drift` (A/ps)"""
    n = int(np.ceil(R / a)) + 1
    x = lattice_fcc((2 * n, 2 * n, 2 * n), a, rng, jitter=0.03, origin=(-n * a, -n * a, -n * a))
    x = x[np.linalg.norm(x, axis=1) <= R] + L / 2.0
    v = maxwell_velocities(len(x), m, T, rng) + np.asarray(drift)
    return ([L, L, L], [0, 0, 0], [m], [name], pots, x, v)


@pytest.mark.parametrize("precision", [F.FMDX_PREC_FP64, F.FMDX_PREC_MIXED])
def test_compacted_cell_lists_same_bits(rng, cu01, precision):
    """Clusters in vacuum (a few percent of the cells occupied): a step over the occupied
    cells only, from lists the device rebuilds after each re-binning, gives the same bits as
    a step over every cell, also with grids exactly as long as the lists at the last status
    read, so that cells occupied since then are taken by the looping launch and the step is
    captured again after a read that finds longer lists.  Members: LJ, Morse and EAM
    clusters, and a two-kind EAM + LJ pair of clusters (the general pair category)."""
    c = CU_MORSE
    specs = [
        _cluster(rng, [("morse", 0, 0, c["D0"], c["alpha"], c["r0"], c["rc"])], c["m"], c["a"], 11.0, 64.0,
                 900.0, (30.0, -10.0, 5.0)),
        _cluster(rng, [("eam", cu01, [(0, 0)])], 63.546, 3.615, 9.0, 56.0, 1200.0, (0.0, 25.0, 0.0)),
        _cluster(rng, [("lj", 0, 0, AR["sig"], AR["eps"], AR["rc"])], AR["m"], AR["a"], 14.0, 70.0, 60.0,
                 (-8.0, 0.0, 12.0), name="Ar"),
    ]
    # a Cu (EAM) cluster and an Ar (LJ) cluster in one open box: two kinds, the general category
    cu = _cluster(rng, [], 63.546, 3.615, 8.0, 60.0, 900.0, (10.0, 0.0, 0.0))[5] - np.array([12.0, 0.0, 0.0])
    ar = _cluster(rng, [], AR["m"], AR["a"], 9.0, 60.0, 60.0, (-10.0, 0.0, 0.0))[5] + np.array([12.0, 0.0, 0.0])
    x = np.vstack([cu, ar])
    v = np.vstack([maxwell_velocities(len(cu), 63.546, 900.0, rng) + [10.0, 0.0, 0.0],
                   maxwell_velocities(len(ar), AR["m"], 60.0, rng) + [-10.0, 0.0, 0.0]])
    kind = np.r_[np.zeros(len(cu)), np.ones(len(ar))].astype(np.int32)
    hyb = ([60.0] * 3, [0, 0, 0], [63.546, AR["m"]], ["Cu", "Ar"],
           [("eam", cu01, [(0, 0)]), ("lj", 1, 1, AR["sig"], AR["eps"], AR["rc"]), ("lj", 0, 1, 2.87, 0.0652, 2.5 * 2.87)],
           x, v, kind)
    runs = {}
    for mode in (F.FMDX_COMPACT_OFF, F.FMDX_COMPACT_AUTO, F.FMDX_COMPACT_TIGHT):
        b = F.Batch(precision)
        b.set_compaction(mode)
        ds = []
        for sp in specs:
            _, d = _make(sp, precision)
            b.add(d, 2e-3)
            ds.append(d)
        L, pbc, m, names, pots, xh, vh, kh = hyb
        _, d = build(L, pbc, m, names, pots, xh, v=vh, kind=kh, precision=precision)
        b.add(d, 2e-3)
        ds.append(d)
        for _ in range(3):
            b.step(16)
        st = [(t.temperature, t.pe, t.err) for t in b.status()]
        runs[mode] = (ds, st)
        b.close()
    ds0, st0 = runs[F.FMDX_COMPACT_OFF]
    assert all(t[2] == 0 for t in st0)
    for mode in (F.FMDX_COMPACT_AUTO, F.FMDX_COMPACT_TIGHT):
        ds, st = runs[mode]
        assert st == st0, mode
        for a, b in zip(ds0, ds):
            _same(a, b)
