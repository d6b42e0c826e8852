"""Build the same system in the oracle and on the device, from identical inputs.

The oracle bins the atoms exactly as the reference loads a state file
(misc.c:242-259) and hands its cell-ordered arrays to the device, so both
sides start from the same cells, the same within-cell order and the same bits.
"""
import numpy as np

import fmd_amd as F
from oracle.oracle import Oracle, MASS_UNIT


def lattice_fcc(n, a, rng=None, jitter=0.0, origin=(0.0, 0.0, 0.0)):
    from fmd_amd.workloads import fcc_positions
    x = fcc_positions(n[0], n[1], n[2], a, origin)
    if jitter and rng is not None:
        x = x + rng.normal(0.0, jitter, size=x.shape)
    return x


def random_separated(n, L, dmin, rng):
    """n uniform points in the periodic box L, at least dmin apart under the minimum image"""
    L = np.asarray(L, dtype=np.float64)
    pts = []
    while len(pts) < n:
        p = rng.uniform(0.0, 1.0, 3) * L
        if pts:
            d = np.asarray(pts) - p
            d -= L * np.round(d / L)
            if np.min(np.einsum("ij,ij->i", d, d)) < dmin * dmin:
                continue
        pts.append(p)
    return np.array(pts)


def build(l, pbc, masses_amu, names, pots, x, v=None, group=None, kind=None,
          precision=F.FMDX_PREC_FP64, device=True):
    """pots: ('lj', k1, k2, sigma, eps, rc) | ('morse', k1, k2, D0, alpha, r0, rc) | ('eam', path, [(k1, k2), ...])"""
    o = Oracle(l, pbc, masses_amu, names)
    eam_pairs = []
    for p in pots:
        if p[0] == "lj":
            o.set_lj(*p[1:])
        elif p[0] == "morse":
            o.set_morse(*p[1:])
        elif p[0] == "eam":
            o.load_eam(p[1])
            for k1, k2 in p[2]:
                o.apply_eam(k1, k2)
                eam_pairs.append((k1, k2))
    o.make_grid()
    x = np.asarray(x, dtype=np.float64)
    n = len(x)
    if v is None:
        v = np.zeros((n, 3))
    o.add_atoms(x, v, group, kind)
    o.distribute()
    if not device:
        return o, None
    st = o.atoms()
    masses = np.asarray(masses_amu, dtype=np.float64) / MASS_UNIT
    d = F.DeviceSystem(l, pbc, o.nc, o.cellh, list(masses), precision)
    tables = None
    for p in pots:
        if p[0] == "lj":
            _, k1, k2, sig, eps, rc = p
            d.set_pair(k1, k2, F.POT_LJ, rc * rc, 4.0 * eps, sig * sig, 0.0)
        elif p[0] == "morse":
            _, k1, k2, D0, alpha, r0, rc = p
            d.set_pair(k1, k2, F.POT_MORSE, rc * rc, D0, alpha, r0)
        elif p[0] == "eam":
            tables = o.eam_tables()
    if tables is not None:
        elem_names = _setfl_elements(pots)
        eok = [elem_names.index(nm) if nm in elem_names else -1 for nm in names]
        d.set_eam(tables, eok)
        for k1, k2 in eam_pairs:
            d.set_pair(k1, k2, F.POT_EAM, tables["rc2"], float(eok[k1]), float(eok[k2]), 0.0)
    d.upload(st["cell_count"], st["x"], st["v"], st["group"], st["atomid"], st["kind"])
    return o, d


def _setfl_elements(pots):
    for p in pots:
        if p[0] == "eam":
            with open(p[1]) as fp:
                for _ in range(3):
                    fp.readline()
                parts = fp.readline().split()
                return parts[1:1 + int(parts[0])]
    return []


def rel_err(a, b, scale=None):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    s = scale if scale is not None else max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / s)


# second element of the two-element setfl: the same analytic forms as the
# synthetic Cu table, other amplitudes and decay lengths
XX = dict(re=2.62, fe=1.21, beta=4.6, A=0.251, alpha=7.1, B=0.402, kappa=4.6, F0=1.83, a=3.70, cutoff=5.5)


def write_setfl_two(path, nr=3000, nrho=3000):
    """A two-element eam/alloy setfl ("Cu" with the synthetic Cu forms, "Xx" with XX):
    per element F(rho) then rho(r), then r*phi for the pairs (0,0), (1,0), (1,1) with
    phi_XxCu = (phi_CuCu + phi_XxXx) / 2 -- the layout eam.c:422-474 reads."""
    from fmd_amd import workloads as W
    ps = [W.CU, XX]
    rc = W.CU["cutoff"]
    dr = rc / (nr - 1)
    r = np.arange(nr) * dr
    rho_e = [W.fcc_shell_density(p) for p in ps]
    drho = 2.0 * max(rho_e) / (nrho - 1)
    rho = np.arange(nrho) * drho
    f = [W.eam_analytic(r, p)[0] for p in ps]
    phi = [W.eam_analytic(r, p)[1] for p in ps]
    rphi = {(0, 0): r * phi[0], (1, 0): r * 0.5 * (phi[0] + phi[1]), (1, 1): r * phi[1]}

    def block(fp, arr):
        for i in range(0, len(arr), 5):
            fp.write(" ".join("%.16e" % v for v in arr[i:i + 5]) + "\n")

    with open(path, "w") as fp:
        fp.write("two-element synthetic eam/alloy (test fixture)\n\n\n")
        fp.write("2 Cu Xx\n")
        fp.write("%d %.16e %d %.16e %.16e\n" % (nrho, drho, nr, dr, rc))
        for e, (name, mass) in enumerate((("29", 63.546), ("99", 58.69))):
            fp.write("%s %.3f %.4f fcc\n" % (name, mass, ps[e]["a"]))
            block(fp, -ps[e]["F0"] * np.sqrt(rho / rho_e[e]))
            block(fp, f[e])
        for key in ((0, 0), (1, 0), (1, 1)):
            block(fp, rphi[key])
    return path
