"""Pins of the oracle (CPU) by checks that do not restate its own code.

* forces / energies: O(N^2) minimum-image brute force in numpy, built from the
  textbook pair formulas and scipy's natural cubic splines;
* two-atom closed forms for LJ and Morse;
* spline preparation: scipy.interpolate.CubicSpline(bc_type='natural');
* EAM r^2 tables: the analytic functions they tabulate;
* symmetry: a perfect periodic FCC lattice has zero force and equal density;
* conservation: total force is zero with full periodicity; translation invariance;
* the EAM + LJ mixture: brute force composed from its three pair families.
"""
import math

import numpy as np
import pytest
from scipy.interpolate import CubicSpline

from oracle import oracle as O
from tests.pairsys import build, lattice_fcc, random_separated, rel_err

MU = O.MASS_UNIT


def minimg(d, L, pbc):
    d = d.copy()
    for k in range(3):
        if pbc[k]:
            d[..., k] -= L[k] * np.round(d[..., k] / L[k])
    return d


def brute_pairwise(x, L, pbc, pair):
    """pair(r2) -> (U, dU/dr2); returns forces and energy over all minimum-image pairs."""
    n = len(x)
    f = np.zeros_like(x)
    e = 0.0
    for i in range(n):
        d = minimg(x[i] - x, L, pbc)
        r2 = np.einsum("ij,ij->i", d, d)
        m = (r2 > 0)
        u, du = pair(r2[m])
        e += 0.5 * np.sum(u)
        f[i] = -2.0 * np.sum(du[:, None] * d[m], axis=0)
    return f, e


def test_lj_two_atoms_closed_form():
    sig, eps, rc = 3.4, 0.0104, 8.5
    for r in (3.2, 3.8, 5.0, 8.4):
        o, _ = build([40.0] * 3, [0, 0, 0], [39.948], ["Ar"], [("lj", 0, 0, sig, eps, rc)],
                     np.array([[10.0, 10.0, 10.0], [10.0 + r, 10.0, 10.0]]), device=False)
        o.forces()
        a = o.atoms()
        fx = 24 * eps / r * (2 * (sig / r) ** 12 - (sig / r) ** 6)
        u = 4 * eps * ((sig / r) ** 12 - (sig / r) ** 6)
        i0 = int(np.argmin(a["x"][:, 0]))
        assert a["f"][i0, 0] == pytest.approx(-fx, rel=1e-13)
        assert a["f"][1 - i0, 0] == pytest.approx(fx, rel=1e-13)
        assert o.energies()["pe"] == pytest.approx(u, rel=1e-13)


def test_morse_two_atoms_closed_form():
    D0, al, r0, rc = 0.3429, 1.3588, 2.866, 7.0
    for r in (2.3, 2.866, 4.1, 6.9):
        o, _ = build([40.0] * 3, [0, 0, 0], [63.546], ["Cu"], [("morse", 0, 0, D0, al, r0, rc)],
                     np.array([[10.0, 10.0, 10.0], [10.0, 10.0 + r, 10.0]]), device=False)
        o.forces()
        a = o.atoms()
        e1 = math.exp(-al * (r - r0))
        fr = 2 * al * D0 * (e1 * e1 - e1)
        i0 = int(np.argmin(a["x"][:, 1]))
        assert a["f"][i0, 1] == pytest.approx(-fr, rel=1e-12, abs=1e-15)
        assert o.energies()["pe"] == pytest.approx(D0 * (e1 * e1 - 2 * e1), rel=1e-13)


@pytest.mark.parametrize("pbc", [(1, 1, 1), (1, 0, 1)])
def test_lj_brute_force(rng, pbc):
    sig, eps, rc = 3.405, 0.0104, 8.5125
    x = lattice_fcc((6, 6, 6), 5.26, rng, jitter=0.2)
    L = [6 * 5.26] * 3
    if not pbc[1]:
        L[1] += 10.0
    o, _ = build(L, list(pbc), [39.948], ["Ar"], [("lj", 0, 0, sig, eps, rc)], x, device=False)
    o.forces()
    a = o.atoms()

    def lj(r2):
        m = r2 < rc * rc
        s6 = (sig * sig / r2) ** 3
        u = np.where(m, 4 * eps * (s6 * s6 - s6), 0.0)
        du = np.where(m, 4 * eps * (-6 * s6 * s6 + 3 * s6) / r2, 0.0)
        return u, du

    f, e = brute_pairwise(a["x"], L, pbc, lj)
    assert rel_err(a["f"], f) < 1e-11
    assert o.energies()["pe"] == pytest.approx(e, rel=1e-12)
    if all(pbc):
        assert np.max(np.abs(a["f"].sum(axis=0))) < 1e-11 * np.max(np.abs(f))


def _scipy_tables(setfl):
    """Re-derive the reference's r^2 tables with scipy: spline in r, sample at sqrt(i dr2)."""
    with open(setfl) as fp:
        lines = fp.read().split("\n")
    head = lines[4].split()
    nrho, drho, nr, dr, rc = int(head[0]), float(head[1]), int(head[2]), float(head[3]), float(head[4])
    vals = np.array(" ".join(lines[6:]).split(), dtype=np.float64)
    F = vals[:nrho]
    rho = vals[nrho:nrho + nr]
    rphi = vals[nrho + nr:nrho + 2 * nr]
    Nr = nr + 2
    r = np.arange(Nr) * dr
    rho_ext = np.r_[rho, 0.0, 0.0]
    phi_ext = np.r_[np.finfo(np.float32).max, rphi[1:] / (np.arange(1, nr) * dr), 0.0, 0.0]
    dr2 = ((Nr - 1) * dr) ** 2 / (Nr - 1)
    s = np.sqrt(np.arange(Nr) * dr2)
    rho2 = CubicSpline(r, rho_ext, bc_type="natural")(s)
    phi2 = CubicSpline(r, phi_ext, bc_type="natural")(s)
    return dict(F=F, drho=drho, rho2=rho2, phi2=phi2, dr2=dr2, rc=rc, Nr2=Nr)


def test_eam_tables_match_scipy_resampling(synth_setfl):
    o, _ = build([30.0] * 3, [1, 1, 1], [63.546], ["Cu"], [("eam", synth_setfl, [(0, 0)])],
                 np.array([[1.0, 1.0, 1.0]]), device=False)
    t = o.eam_tables()
    s = _scipy_tables(synth_setfl)
    assert t["nr2"] == s["Nr2"] and t["dr2"] == pytest.approx(s["dr2"], rel=1e-15)
    # the reference zero-pads two knots: compare inside the cutoff, away from the FLT_MAX knot
    k = np.arange(1, t["nr2"])
    k = k[k * t["dr2"] < s["rc"] ** 2]
    assert rel_err(t["rho"][0][k], s["rho2"][k]) < 1e-12
    kk = k[k > 50]
    assert rel_err(t["phi"][0, 0][kk], s["phi2"][kk]) < 1e-9
    # second derivatives: the natural spline on the r^2 grid
    cs = CubicSpline(np.arange(t["nr2"]) * t["dr2"], t["rho"][0], bc_type="natural")
    assert rel_err(t["rhoDD"][0][1:-1], cs(np.arange(t["nr2"]) * t["dr2"], 2)[1:-1]) < 1e-7


def test_eam_tables_match_analytic_forms(synth_setfl):
    from fmd_amd import workloads as W
    o, _ = build([30.0] * 3, [1, 1, 1], [63.546], ["Cu"], [("eam", synth_setfl, [(0, 0)])],
                 np.array([[1.0, 1.0, 1.0]]), device=False)
    t = o.eam_tables()
    r2 = np.linspace(2.0 ** 2, 5.4 ** 2, 200)
    f_ref, phi_ref = W.eam_analytic(np.sqrt(r2))
    f = [O.spline_val(t["dr2"], t["rho"][0], t["rhoDD"][0], q) for q in r2]
    phi = [O.spline_val(t["dr2"], t["phi"][0, 0], t["phiDD"][0, 0], q) for q in r2]
    assert rel_err(f, f_ref) < 1e-9
    assert rel_err(phi, phi_ref) < 1e-8


@pytest.mark.parametrize("which", ["synth", "cu01"])
def test_eam_brute_force(rng, which, synth_setfl, cu01):
    setfl = synth_setfl if which == "synth" else cu01
    x = lattice_fcc((5, 5, 5), 3.615, rng, jitter=0.05)
    L = [5 * 3.615] * 3
    o, _ = build(L, [1, 1, 1], [63.546], ["Cu"], [("eam", setfl, [(0, 0)])], x, device=False)
    o.forces()
    a = o.atoms()
    t = o.eam_tables()
    rr = np.arange(t["nr2"]) * t["dr2"]
    rho_s = CubicSpline(rr, t["rho"][0], bc_type=((2, 0.0), (2, 0.0)))
    phi_s = CubicSpline(rr, t["phi"][0, 0], bc_type=((2, 0.0), (2, 0.0)))
    Fs = CubicSpline(np.arange(t["nrho"]) * t["drho"], t["F"][0], bc_type=((2, 0.0), (2, 0.0)))
    rc2 = t["rc2"]
    xs = a["x"]
    n = len(xs)
    rho = np.zeros(n)
    pairs = []
    for i in range(n):
        d = minimg(xs[i] - xs, L, (1, 1, 1))
        r2 = np.einsum("ij,ij->i", d, d)
        m = (r2 > 0) & (r2 < rc2)
        rho[i] = np.sum(rho_s(r2[m]))
        pairs.append((m, d, r2))
    Fp = Fs(rho, 1)
    f = np.zeros_like(xs)
    e = np.sum(Fs(rho))
    for i in range(n):
        m, d, r2 = pairs[i]
        mag = 2 * (Fp[i] * rho_s(r2[m], 1) + Fp[m] * rho_s(r2[m], 1) + phi_s(r2[m], 1))
        f[i] = -np.sum(mag[:, None] * d[m], axis=0)
        e += 0.5 * np.sum(phi_s(r2[m]))
    assert rel_err(a["aux"], Fp) < 1e-9
    assert rel_err(a["f"], f) < 1e-9
    assert o.energies()["pe"] == pytest.approx(e, rel=1e-11)


def test_eam_two_elements_brute_force(rng, tmp_path):
    """eam/alloy with two elements: rho_i sums the density function of each neighbour's element,
    the pair force uses F'_i rho'_e(j) + F'_j rho'_e(i) + phi'_e(i)e(j) (eam.h:176-228, 270-310),
    phi of the unlike pair is the one table both orders share (eam.c:455-474)."""
    from tests.pairsys import write_setfl_two
    setfl = write_setfl_two(str(tmp_path / "two.eam.alloy"))
    x = lattice_fcc((5, 5, 5), 3.65, rng, jitter=0.05)
    kind = (rng.random(len(x)) < 0.4).astype(np.int32)
    L = [5 * 3.65] * 3
    o, _ = build(L, [1, 1, 1], [63.546, 58.69], ["Cu", "Xx"], [("eam", setfl, [(0, 0), (0, 1), (1, 1)])], x,
                 kind=kind, device=False)
    o.forces()
    a = o.atoms()
    t = o.eam_tables()
    assert t["nelem"] == 2
    rr = np.arange(t["nr2"]) * t["dr2"]
    nat = ((2, 0.0), (2, 0.0))
    rho_s = [CubicSpline(rr, t["rho"][e], bc_type=nat) for e in range(2)]
    phi_s = [[CubicSpline(rr, t["phi"][e, g], bc_type=nat) for g in range(2)] for e in range(2)]
    F_s = [CubicSpline(np.arange(t["nrho"]) * t["drho"], t["F"][e], bc_type=nat) for e in range(2)]
    # the tables themselves: one phi for both orders of the unlike pair, the mixing rule of the file
    assert np.array_equal(t["phi"][0, 1], t["phi"][1, 0])
    q = np.linspace(2.3 ** 2, 5.4 ** 2, 50)
    assert rel_err(phi_s[0][1](q), 0.5 * (phi_s[0][0](q) + phi_s[1][1](q))) < 1e-9
    rc2 = t["rc2"]
    xs, ks = a["x"], a["kind"].astype(int)
    n = len(xs)
    rho = np.zeros(n)
    pairs = []
    for i in range(n):
        d = minimg(xs[i] - xs, L, (1, 1, 1))
        r2 = np.einsum("ij,ij->i", d, d)
        m = np.nonzero((r2 > 0) & (r2 < rc2))[0]
        rho[i] = sum(np.sum(rho_s[e](r2[m][ks[m] == e])) for e in range(2))
        pairs.append((m, d[m], r2[m]))
    Fp = np.array([F_s[ks[i]](rho[i], 1) for i in range(n)])
    e = sum(F_s[ks[i]](rho[i]) for i in range(n))
    f = np.zeros_like(xs)
    for i in range(n):
        m, d, r2 = pairs[i]
        ei, ej = ks[i], ks[m]
        rho_j = np.where(ej == 0, rho_s[0](r2, 1), rho_s[1](r2, 1))   # d rho_e(j) / d r2
        rho_i = rho_s[ei](r2, 1)
        phid = np.where(ej == 0, phi_s[ei][0](r2, 1), phi_s[ei][1](r2, 1))
        mag = 2 * (Fp[i] * rho_j + Fp[m] * rho_i + phid)
        f[i] = -np.sum(mag[:, None] * d, axis=0)
        e += 0.5 * np.sum(np.where(ej == 0, phi_s[ei][0](r2), phi_s[ei][1](r2)))
    assert rel_err(a["aux"], Fp) < 1e-9
    assert rel_err(a["f"], f) < 1e-9
    assert o.energies()["pe"] == pytest.approx(e, rel=1e-11)
    # a brute force that swapped the density functions of the pair (rho_e(i) into rho_i) must fail
    rho_bad = np.zeros(n)
    for i in range(n):
        m, d, r2 = pairs[i]
        rho_bad[i] = np.sum(rho_s[ks[i]](r2))
    assert rel_err(np.array([F_s[ks[i]](rho_bad[i], 1) for i in range(n)]), Fp) > 1e-3


def test_perfect_lattice_symmetry(synth_setfl):
    x = lattice_fcc((6, 6, 6), 3.615)
    L = [6 * 3.615] * 3
    o, _ = build(L, [1, 1, 1], [63.546], ["Cu"], [("eam", synth_setfl, [(0, 0)])], x, device=False)
    o.forces()
    a = o.atoms()
    assert np.max(np.abs(a["f"])) < 1e-11
    assert np.ptp(a["aux"]) < 1e-12 * abs(a["aux"][0])


def test_translation_invariance(rng):
    x = lattice_fcc((6, 6, 6), 5.26, rng, jitter=0.2)
    L = [6 * 5.26] * 3
    o1, _ = build(L, [1, 1, 1], [39.948], ["Ar"], [("lj", 0, 0, 3.405, 0.0104, 8.5125)], x, device=False)
    o2, _ = build(L, [1, 1, 1], [39.948], ["Ar"], [("lj", 0, 0, 3.405, 0.0104, 8.5125)],
                  np.mod(x + np.array([7.3, -2.1, 11.9]), L[0]), device=False)
    o1.forces(); o2.forces()
    a1, a2 = o1.atoms(), o2.atoms()
    f1 = a1["f"][np.argsort(a1["atomid"])]
    f2 = a2["f"][np.argsort(a2["atomid"])]
    assert rel_err(f2, f1) < 1e-10


def test_spline_prepare_matches_scipy(rng):
    y = np.cumsum(rng.normal(size=200))
    h = 0.37
    dd = O.spline_prepare(h, y)
    cs = CubicSpline(np.arange(200) * h, y, bc_type="natural")
    assert rel_err(dd, cs(np.arange(200) * h, 2)) < 1e-10
    xs = np.sort(rng.uniform(0, 199 * h * 0.999, 50))
    assert rel_err([O.spline_val(h, y, dd, q) for q in xs], cs(xs)) < 1e-12
    assert rel_err([O.spline_deriv(h, y, dd, q) for q in xs], cs(xs, 1)) < 1e-11
    xk = np.sort(rng.uniform(0, 10, 60))
    yk = np.sin(xk)
    assert rel_err(O.spline_prepare_adv(xk, yk), CubicSpline(xk, yk, bc_type="natural")(xk, 2)) < 1e-9


def test_hybrid_eam_lj_brute_force(rng, cu01):
    """Example 02's mixture (Cu EAM, Ar-Ar LJ, Cu-Ar LJ; examples/02_collision.c:78-85)
    against an O(N^2) brute force that composes the three parts: densities of Cu
    atoms from Cu neighbours only, EAM forces between Cu atoms, LJ forces with the
    pair's own parameters for Ar-Ar and Cu-Ar, energy F(rho) + phi/2 + LJ/2."""
    xcu = lattice_fcc((4, 4, 4), 3.6316, rng, jitter=0.03, origin=(2.0, 2.0, 2.0))
    xar = lattice_fcc((3, 3, 3), 5.26, rng, jitter=0.05, origin=(17.0, 2.0, 2.0))
    x = np.vstack([xcu, xar])
    kind = np.r_[np.zeros(len(xcu)), np.ones(len(xar))].astype(np.int32)
    L = [36.0, 26.0, 26.0]
    lj_ar, lj_cuar = (3.40, 0.0104, 2.5 * 3.40), (2.87, 0.0652, 2.5 * 2.87)
    pots = [("eam", cu01, [(0, 0)]), ("lj", 1, 1) + lj_ar, ("lj", 0, 1) + lj_cuar]
    o, _ = build(L, [1, 1, 1], [63.546, 39.948], ["Cu", "Ar"], pots, x, kind=kind, device=False)
    o.forces()
    a = o.atoms()
    t = o.eam_tables()
    rr = np.arange(t["nr2"]) * t["dr2"]
    rho_s = CubicSpline(rr, t["rho"][0], bc_type=((2, 0.0), (2, 0.0)))
    phi_s = CubicSpline(rr, t["phi"][0, 0], bc_type=((2, 0.0), (2, 0.0)))
    Fs = CubicSpline(np.arange(t["nrho"]) * t["drho"], t["F"][0], bc_type=((2, 0.0), (2, 0.0)))
    xs, ks = a["x"], a["kind"]
    n = len(xs)
    cu = ks == 0

    def lj(r2, p):
        sig, eps, rc = p
        m = r2 < rc * rc
        s6 = (sig * sig / r2) ** 3
        return np.where(m, 4 * eps * (s6 * s6 - s6), 0.0), np.where(m, 4 * eps * (-6 * s6 * s6 + 3 * s6) / r2, 0.0)

    geo = []
    rho = np.zeros(n)
    for i in range(n):
        d = minimg(xs[i] - xs, L, (1, 1, 1))
        r2 = np.einsum("ij,ij->i", d, d)
        r2[i] = np.inf
        geo.append((d, r2))
        if cu[i]:
            m = cu & (r2 < t["rc2"])
            rho[i] = np.sum(rho_s(r2[m]))
    Fp = np.where(cu, Fs(rho, 1), 0.0)
    e = np.sum(Fs(rho[cu]))
    f = np.zeros_like(xs)
    for i in range(n):
        d, r2 = geo[i]
        if cu[i]:
            m = cu & (r2 < t["rc2"])
            mag = 2 * (Fp[i] * rho_s(r2[m], 1) + Fp[m] * rho_s(r2[m], 1) + phi_s(r2[m], 1))
            f[i] -= np.sum(mag[:, None] * d[m], axis=0)
            e += 0.5 * np.sum(phi_s(r2[m]))
            other = ~cu
            u, du = lj(r2[other], lj_cuar)
        else:
            u1, du1 = lj(r2[~cu], lj_ar)
            u2, du2 = lj(r2[cu], lj_cuar)
            f[i] -= 2.0 * np.sum(du1[:, None] * d[~cu], axis=0)
            e += 0.5 * np.sum(u1)
            other = cu
            u, du = u2, du2
        f[i] -= 2.0 * np.sum(du[:, None] * d[other], axis=0)
        e += 0.5 * np.sum(u)
    assert rel_err(a["f"], f) < 1e-9
    assert rel_err(a["aux"][cu], Fp[cu]) < 1e-9
    assert o.energies()["pe"] == pytest.approx(e, rel=1e-11)
    assert np.max(np.abs(a["f"].sum(axis=0))) < 1e-9 * np.max(np.abs(f))


def brute_images(x, L, pair):
    """all pairs with j's periodic images shifted by n L, n in {-1,0,1}^3 (self only for n != 0);
    exact whenever every L >= rc, since farther images lie beyond the cutoff"""
    n = len(x)
    f = np.zeros_like(x)
    e = 0.0
    shifts = np.array([[a, b, c] for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)], dtype=float) * L
    for i in range(n):
        for s in shifts:
            d = x[i] - (x + s)
            r2 = np.einsum("ij,ij->i", d, d)
            m = r2 > 0
            u, du = pair(r2[m])
            e += 0.5 * np.sum(u)
            f[i] -= 2.0 * np.sum(du[:, None] * d[m], axis=0)
    return f, e


def test_lj_degenerate_boxes_all_images(rng):
    """Boxes with 1 and 2 cells along a dimension (md-ghost.c:85-107 copies the
    single cell to both ghost faces): an atom pairs with several images of the
    same neighbour and with its own images."""
    sig, eps, rc = 3.405, 0.0104, 8.5125
    L = [1.5 * rc, 2.3 * rc, 3.2 * rc]
    n = 60
    x = random_separated(n, L, 3.0, rng)
    o, _ = build(L, [1, 1, 1], [39.948], ["Ar"], [("lj", 0, 0, sig, eps, rc)], x, device=False)
    assert list(o.nc) == [1, 2, 3]
    o.forces()
    a = o.atoms()

    def lj(r2):
        m = r2 < rc * rc
        s6 = (sig * sig / np.where(m, r2, 1.0)) ** 3
        return np.where(m, 4 * eps * (s6 * s6 - s6), 0.0), np.where(m, 4 * eps * (-6 * s6 * s6 + 3 * s6) / r2, 0.0)

    f, e = brute_images(a["x"], np.array(L), lj)
    assert rel_err(a["f"], f) < 1e-11
    assert o.energies()["pe"] == pytest.approx(e, rel=1e-12)
