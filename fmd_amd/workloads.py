"""Synthetic inputs for the step benchmarks and parity tests.

* FCC lattices in the reference's generation order (structure.c:53-136: unit
  cells in x-outer / z-inner order, four basis points each, quarter-cell
  offset), optional Gaussian jitter, Maxwell velocities with zero net
  momentum, all from a numpy seed;
* a synthetic setfl EAM file for Cu tabulated from analytic forms
      f(r)   = fe * exp(-beta (r/re - 1))
      phi(r) = A exp(-alpha (r/re - 1)) - B exp(-kappa (r/re - 1))
      F(rho) = -F0 * sqrt(rho / rho_e)
  with rho_e the FCC equilibrium density at a = 3.615 A inside the cutoff.
"""
import ctypes as C
import math
import os
import time

import numpy as np

import fmd_amd as F

MASS_UNIT = 9.6485332907e+03
KB = 8.6173303e-5

# analytic Cu parameters (exponents and amplitudes of the generalized EAM of
# Zhou, Johnson and Wadley for Cu, without their cutoff denominators)
CU = dict(re=2.556162, fe=1.554485, beta=4.090619, A=0.327584, alpha=7.669911, B=0.468735,
          kappa=4.090619, F0=2.19, a=3.615, cutoff=5.5)


def fcc_positions(nx, ny, nz, a, origin=(0.0, 0.0, 0.0)):
    """Positions in generation order (x outer, z inner, 4 basis points per cell)."""
    basis = np.array([[0.0, 0.0, 0.0], [0.0, 0.5, 0.5], [0.5, 0.0, 0.5], [0.5, 0.5, 0.0]])
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    cells = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1).astype(np.float64)
    pos = (cells[:, None, :] + 0.25 + basis[None, :, :]) * a + np.asarray(origin)[None, None, :]
    return pos.reshape(-1, 3)


def maxwell_velocities(n, mass_amu, temperature, rng):
    """Gaussian velocities (A/ps) at the given temperature, zero net momentum."""
    m = mass_amu / MASS_UNIT
    v = rng.normal(0.0, math.sqrt(KB * temperature / m), size=(n, 3))
    v -= v.mean(axis=0)
    return v


def eam_analytic(r, p=CU):
    x = r / p["re"] - 1.0
    f = p["fe"] * np.exp(-p["beta"] * x)
    phi = p["A"] * np.exp(-p["alpha"] * x) - p["B"] * np.exp(-p["kappa"] * x)
    return f, phi


def fcc_shell_density(p=CU):
    """rho_e: density of a perfect FCC site with neighbours inside the cutoff."""
    a = p["a"]
    rng = np.arange(-3, 4)
    basis = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]])
    pts = (np.stack(np.meshgrid(rng, rng, rng, indexing="ij"), -1).reshape(-1, 1, 3) + basis[None]).reshape(-1, 3)
    d = np.linalg.norm(pts, axis=1) * a
    d = d[(d > 1e-9) & (d < p["cutoff"])]
    return float(np.sum(eam_analytic(d, p)[0]))


def write_setfl(path, nr=5000, nrho=5000, p=CU):
    """A one-element setfl file (eam/alloy) tabulating the analytic forms."""
    rho_e = fcc_shell_density(p)
    dr = p["cutoff"] / (nr - 1)
    drho = 2.0 * rho_e / (nrho - 1)
    r = np.arange(nr) * dr
    f, phi = eam_analytic(r, p)
    rho = np.arange(nrho) * drho
    F = -p["F0"] * np.sqrt(rho / rho_e)
    rphi = r * phi
    with open(path, "w") as fp:
        fp.write("synthetic Cu EAM: f = fe exp(-beta(r/re-1)), phi = A exp(-alpha(r/re-1)) - B exp(-kappa(r/re-1))\n")
        fp.write("F(rho) = -F0 sqrt(rho/rho_e), rho_e = %.10f\n" % rho_e)
        fp.write("tabulated by fmd_amd.workloads\n")
        fp.write("1 Cu\n")
        fp.write("%d %.16e %d %.16e %.16e\n" % (nrho, drho, nr, dr, p["cutoff"]))
        fp.write("29 63.546 %.6f fcc\n" % p["a"])
        for arr in (F, f, rphi):
            for i in range(0, len(arr), 5):
                fp.write(" ".join("%.16e" % v for v in arr[i:i + 5]) + "\n")
    return dict(rho_e=rho_e, dr=dr, drho=drho)


def data_dir():
    return os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests", "golden", "data")


# --- the EAM Cu step through the public API (used by bench.py and fmd_amd.autotune) ---

def fmd_lib():
    """libfmd.so with the argument types of the calls below."""
    lib = F.load_fmd()
    P = C.c_void_p
    lib.fmd_create.restype = P
    lib.fmd_free.argtypes = [P]
    lib.fmd_box_setSize.argtypes = [P, C.c_double, C.c_double, C.c_double]
    lib.fmd_box_setPBC.argtypes = [P, C.c_bool, C.c_bool, C.c_bool]
    lib.fmd_matt_setAtomKinds.argtypes = [P, C.c_uint, P, P]
    lib.fmd_pot_eam_alloy_load.restype = P
    lib.fmd_pot_eam_alloy_load.argtypes = [P, C.c_char_p]
    lib.fmd_pot_apply.argtypes = [P, C.c_uint, C.c_uint, P]
    lib.fmd_ext_loadAtoms.argtypes = [P, C.c_uint, P, P, P, P]
    lib.fmd_ext_distribute.restype = C.c_bool
    lib.fmd_ext_distribute.argtypes = [P]
    lib.fmd_ext_getDevice.restype = P
    lib.fmd_ext_getDevice.argtypes = [P]
    lib.fmd_ext_setPrecision.argtypes = [P, C.c_int]
    return lib


def eam_cu_system(cells, seed, setfl, precision):
    """The EAM Cu workload (FCC cells^3, a = 3.615 A, jitter 0.02 A, 600 K, periodic) on the device
    through the public API; returns (libfmd, fmd_t handle, device handle, natoms)."""
    lib = fmd_lib()
    a = CU["a"]
    rng = np.random.default_rng(seed)
    x = fcc_positions(cells, cells, cells, a)
    x = x + rng.normal(0.0, 0.02, size=x.shape)
    L = cells * a
    x = np.mod(x, L)
    v = maxwell_velocities(len(x), 63.546, 600.0, rng)
    md = lib.fmd_create()
    lib.fmd_box_setSize(md, L, L, L)
    lib.fmd_box_setPBC(md, True, True, True)
    names = (C.c_char_p * 1)(b"Cu")
    mass = (C.c_double * 1)(63.546)
    lib.fmd_matt_setAtomKinds(md, 1, C.cast(names, C.c_void_p), C.cast(mass, C.c_void_p))
    pot = lib.fmd_pot_eam_alloy_load(md, setfl.encode())
    if not pot:
        raise RuntimeError("setfl load failed")
    lib.fmd_pot_apply(md, 0, 0, pot)
    x = np.ascontiguousarray(x)
    v = np.ascontiguousarray(v)
    lib.fmd_ext_loadAtoms(md, len(x), x.ctypes.data_as(C.c_void_p), v.ctypes.data_as(C.c_void_p), None, None)
    lib.fmd_ext_setPrecision(md, precision)
    if not lib.fmd_ext_distribute(md):
        raise RuntimeError("upload failed")
    return lib, md, lib.fmd_ext_getDevice(md), len(x)


def choose_cap(maxc):
    """Slot capacity for a fullest cell of maxc atoms: the rule of the product's time loop
    (1.5 x + 8, in multiples of 32, at most 256)."""
    c = maxc + maxc // 2 + 8
    return min(256, (c + 31) & ~31)


def check_device(fx, dev):
    """The periodic check of the product's time loops: audit the resident state, read the thermo
    record, raise on a device error word or a broken invariant (naming where the device first saw
    it), and grow the slot capacity by the loop's rule when the fullest cell nears it."""
    F._ck(fx.fmdx_audit(dev), "audit")
    th = F.Thermo()
    F._ck(fx.fmdx_read_thermo(dev, C.byref(th)), "read_thermo")
    if th.err:
        raise RuntimeError("device error word %d (first non-finite force in cell %d, first unbinnable "
                           "position in cell %d)" % (th.err, th.bad_force_cell - 1, th.bad_position_cell - 1))
    if th.audit_atoms != th.natoms or th.audit_nonfinite or th.audit_max_cell > th.cap:
        raise RuntimeError("device state check: %d atoms in the cells for %d resident, fullest cell %d of %d "
                           "slots, %d atoms with non-finite x, v or F"
                           % (th.audit_atoms, th.natoms, th.audit_max_cell, th.cap, th.audit_nonfinite))
    if th.max_cell_count + th.max_cell_count // 4 + 4 > th.cap and th.cap < 256:
        F._ck(fx.fmdx_set_capacity(dev, choose_cap(th.max_cell_count)), "set_capacity")
    return th


def vv_berendsen_steps(fx, dev, nsteps, dt, tau, T0):
    """nsteps velocity-Verlet steps with Berendsen rescaling through the C-ABI seams, as
    fmd_dync_equilibrate runs them without a timer: the potential energy is summed on the last
    step only (the thermostat needs the temperature, integrators.c:44-50)."""
    for k in range(nsteps):
        F._ck(fx.fmdx_vv_start(dev, -1, dt, 1, tau, T0), "vv_start")
        flags = F.FMDX_FORCES_ENERGY if k == nsteps - 1 else 0
        F._ck(fx.fmdx_update_forces_ex(dev, -1, 0, flags), "update_forces")
        F._ck(fx.fmdx_vv_finish(dev, -1, dt), "vv_finish")


def time_eam_cu_steps(cells, steps, warmup, precision, setfl, seed=1):
    """HIP-event time of `steps` EAM Cu steps after `warmup`, with the two force passes' own times."""
    fx = F.load_fmdx()
    lib, md, dev, n = eam_cu_system(cells, seed, setfl, precision)
    dt, tau, T0 = 2e-3, 0.1, 600.0
    F._ck(fx.fmdx_update_forces(dev, -1, 0), "update_forces")
    F._ck(fx.fmdx_thermo(dev, -1), "thermo")
    vv_berendsen_steps(fx, dev, warmup, dt, tau, T0)
    th = F.Thermo()
    F._ck(fx.fmdx_read_thermo(dev, C.byref(th)), "read_thermo")
    F._ck(fx.fmdx_profile_enable(dev, 1), "profile")
    F._ck(fx.fmdx_event_record(dev, 0), "event")
    t0 = time.perf_counter()
    vv_berendsen_steps(fx, dev, steps, dt, tau, T0)
    F._ck(fx.fmdx_event_record(dev, 1), "event")
    F._ck(fx.fmdx_sync(dev), "sync")
    wall = time.perf_counter() - t0
    ms = C.c_float(0)
    F._ck(fx.fmdx_event_elapsed_ms(dev, 0, 1, C.byref(ms)), "elapsed")
    prof = (C.c_double * 8)()
    cnt = C.c_uint(0)
    F._ck(fx.fmdx_profile_read(dev, prof, C.byref(cnt)), "profile_read")
    F._ck(fx.fmdx_profile_enable(dev, 0), "profile")
    th = check_device(fx, dev)
    res = dict(natoms=n, steps=steps, wall_s=wall, gpu_ms=ms.value,
               pass1_ms=prof[0] / max(cnt.value, 1), pass0_ms=prof[1] / max(cnt.value, 1),
               T=th.temperature, pe=th.pe, cap=th.cap, max_cell=th.max_cell_count)
    lib.fmd_free(md)
    return res
