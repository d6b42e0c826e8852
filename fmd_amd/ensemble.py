"""Job and ensemble workloads built through the public C API (libfmd.so).

Two synthetic workloads, both seeded:

* ensemble: independent periodic FCC systems, sizes log-uniform in
  [lo, hi] atoms, each either LJ argon or Morse copper, temperatures
  uniform in [50, 800] K;
* collisions: two Morse-copper spherical FCC clusters in a vacuum box,
  radii uniform in [15, 50] A, a 20 A surface gap between them, relative
  speed uniform in [10, 50] A/ps along a random axis, impact parameter
  uniform in [0, R1 + R2], 300 K internal motion, 1000-4000 steps of 1 fs.

Each system is an fmd_t set up with the reference's calls (box, kinds,
potential, atoms) and distributed to the device; ``Job.dev`` is its
fmdx_system_t handle for the batch engine.
"""
import math

import numpy as np

from . import api
from .workloads import fcc_positions, maxwell_velocities

LJ_AR = dict(name="Ar", mass=39.948, a=5.26, sigma=3.405, eps=0.0104, cutoff=2.5 * 3.405)
MORSE_CU = dict(name="Cu", mass=63.546, a=3.615, D0=0.3429, alpha=1.3588, r0=2.866, cutoff=7.0)


class Job:
    """One fmd_t with its atoms on the device."""

    def __init__(self, L, pbc, pot, x, v, precision, nsteps=0, dt=1e-3, tag=None):
        self.f = api.Fmd()
        f = self.f
        f.fmd_io_setShowErrorMessages(False)
        f.fmd_box_setSize(*L)
        f.fmd_box_setPBC(*[bool(p) for p in pbc])
        p = LJ_AR if pot == "lj" else MORSE_CU
        f.set_kinds([p["name"]], [p["mass"]])
        if pot == "lj":
            f.fmd_pot_lj_apply(0, 0, p["sigma"], p["eps"], p["cutoff"])
        else:
            f.fmd_pot_morse_apply(0, 0, p["D0"], p["alpha"], p["r0"], p["cutoff"])
        f.load_atoms(x, v)
        f.fmd_ext_setPrecision(precision)
        if not f.fmd_ext_distribute():
            raise RuntimeError("distribute failed")
        self.dev = f.fmd_ext_getDevice()
        self.natoms = len(x)
        self.nsteps = nsteps
        self.dt = dt
        self.tag = tag

    def free(self):
        self.f.free()


def ensemble_specs(n=512, seed=3, lo=500, hi=50000):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        target = math.exp(rng.uniform(math.log(lo), math.log(hi)))
        pot = "lj" if rng.random() < 0.5 else "morse"
        T = float(rng.uniform(50.0, 800.0))
        cells = max(3, int(round((target / 4.0) ** (1.0 / 3.0))))
        out.append(dict(index=i, pot=pot, cells=cells, T=T, seed=int(rng.integers(1 << 31)), target=target))
    return out


def ensemble_job(spec, precision, dt=1e-3, nsteps=2000):
    p = LJ_AR if spec["pot"] == "lj" else MORSE_CU
    n = spec["cells"]
    rng = np.random.default_rng(spec["seed"])
    x = fcc_positions(n, n, n, p["a"])
    v = maxwell_velocities(len(x), p["mass"], spec["T"], rng)
    L = [n * p["a"]] * 3
    return Job(L, (1, 1, 1), spec["pot"], x, v, precision, nsteps=nsteps, dt=dt, tag=spec["index"])


def _sphere(R, a, rng):
    n = int(math.ceil(R / a)) + 1
    x = fcc_positions(2 * n, 2 * n, 2 * n, a, origin=(-n * a, -n * a, -n * a))
    x = x[np.linalg.norm(x, axis=1) <= R]
    return x


def collision_specs(n, seed=4, rate=50.0, steps=(1000, 4000)):
    """n jobs with Poisson arrival times (s) at `rate` jobs/s.  The jobs and the arrivals come
    from separate streams: the same seed gives the same n jobs at every rate, arriving at the
    times of one unit-rate Poisson stream divided by the rate."""
    rng = np.random.default_rng(seed)
    unit = np.cumsum(np.random.default_rng([seed, 1]).exponential(1.0, size=n))
    out = []
    for i in range(n):
        out.append(dict(index=i, arrival=float(unit[i] / rate), R1=float(rng.uniform(15.0, 50.0)),
                        R2=float(rng.uniform(15.0, 50.0)), speed=float(rng.uniform(10.0, 50.0)),
                        b=float(rng.uniform(0.0, 1.0)), nsteps=int(rng.integers(steps[0], steps[1] + 1)),
                        seed=int(rng.integers(1 << 31))))
    return out


def collision_atoms(spec):
    """positions, velocities and box of a two-cluster collision (vacuum, no PBC)"""
    rng = np.random.default_rng(spec["seed"])
    p = MORSE_CU
    R1, R2 = spec["R1"], spec["R2"]
    x1, x2 = _sphere(R1, p["a"], rng), _sphere(R2, p["a"], rng)
    axis = rng.normal(size=3)
    axis /= np.linalg.norm(axis)
    perp = np.cross(axis, rng.normal(size=3))
    perp /= np.linalg.norm(perp)
    d = R1 + R2 + 20.0
    bimp = spec["b"] * (R1 + R2)
    c2 = d * axis + bimp * perp
    n1, n2 = len(x1), len(x2)
    m1, m2 = float(n1), float(n2)
    vrel = spec["speed"] * axis
    v1 = np.tile(vrel * (m2 / (m1 + m2)), (n1, 1))
    v2 = np.tile(-vrel * (m1 / (m1 + m2)), (n2, 1))
    v1 += maxwell_velocities(n1, p["mass"], 300.0, rng)
    v2 += maxwell_velocities(n2, p["mass"], 300.0, rng)
    x = np.concatenate([x1, x2 + c2])
    v = np.concatenate([v1, v2])
    lo, hi = x.min(axis=0), x.max(axis=0)
    margin = 25.0
    x = x - lo + margin
    L = list(hi - lo + 2 * margin)
    return x, v, L


def collision_job(spec, precision, dt=1e-3):
    x, v, L = collision_atoms(spec)
    return Job(L, (0, 0, 0), "morse", x, v, precision, nsteps=spec["nsteps"], dt=dt, tag=spec["index"])
