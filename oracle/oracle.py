"""ctypes wrapper of the CPU oracle (liboracle.so, built from fmd_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check of
__graft_entry__ and by the CPU-baseline leg of bench.py, never by the product.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")

MASS_UNIT = 9.6485332907e+03
KB = 8.6173303e-5
POT_LJ, POT_EAM, POT_MORSE = 1, 2, 3

_lib = None


class TtmCfg(C.Structure):
    _fields_ = [("type", C.c_int), ("ncz", C.c_int), ("ratio", C.c_int),
                ("C_gamma", C.c_double), ("K", C.c_double), ("G", C.c_double),
                ("Kzh_A", C.c_double), ("Kzh_B", C.c_double), ("Kzh_v2", C.c_double),
                ("nC", C.c_int), ("nG", C.c_int),
                ("T_C", C.c_void_p), ("Ct", C.c_void_p), ("T_G", C.c_void_p), ("Gt", C.c_void_p),
                ("act_frac", C.c_double),
                ("laser_factor", C.c_double), ("laser_absdepth_inv", C.c_double),
                ("laser_2sig2_inv", C.c_double), ("laser_t0", C.c_double), ("laser_tdiff", C.c_double),
                ("start", C.c_double), ("stop", C.c_double),
                ("ncx", C.c_int), ("ncy", C.c_int), ("kmode", C.c_int), ("solver3d", C.c_int)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("oracle not built: run __graft_entry__.build()")
        L = C.CDLL(LIB)
        P = C.c_void_p
        L.orc_create.restype = P
        L.orc_create.argtypes = [P, P, C.c_int, P, P]
        L.orc_free.argtypes = [P]
        L.orc_set_lj.argtypes = [P, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double]
        L.orc_set_morse.argtypes = [P, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]
        L.orc_load_eam.argtypes = [P, C.c_char_p]
        L.orc_apply_eam.argtypes = [P, C.c_int, C.c_int]
        L.orc_eam_cutoff.argtypes = [P]
        L.orc_eam_cutoff.restype = C.c_double
        L.orc_eam_sizes.argtypes = [P] + [P] * 6
        L.orc_eam_tables.argtypes = [P, C.c_int, P, P, P, P]
        L.orc_eam_phi.argtypes = [P, C.c_int, C.c_int, P, P]
        L.orc_make_grid.argtypes = [P]
        L.orc_grid_info.argtypes = [P, P, P]
        L.orc_add_atoms.argtypes = [P, C.c_int, P, P, P, P]
        L.orc_distribute.argtypes = [P]
        L.orc_natoms.argtypes = [P]
        L.orc_get_atoms.argtypes = [P] + [P] * 8
        L.orc_cell_counts.argtypes = [P, P]
        L.orc_refresh.argtypes = [P]
        L.orc_forces.argtypes = [P, C.c_int]
        L.orc_vv_start.argtypes = [P, C.c_int, C.c_double, C.c_int, C.c_double, C.c_double]
        L.orc_vv_finish.argtypes = [P, C.c_int, C.c_double]
        L.orc_se_step.argtypes = [P, C.c_int, C.c_double]
        L.orc_thermo.argtypes = [P, C.c_int]
        L.orc_get_energies.argtypes = [P, P]
        L.orc_error_count.argtypes = [P, P]
        L.orc_integrate_vv.argtypes = [P, C.c_int, P, C.c_double, C.c_double, C.c_int, C.c_double, C.c_double]
        L.orc_timer_is_its_time.argtypes = [C.c_double] * 4
        L.orc_ttm_init.argtypes = [P, C.POINTER(TtmCfg), C.c_double]
        L.orc_ttm_get_ready.argtypes = [P, C.c_int]
        L.orc_ttm_update.argtypes = [P, C.c_double, C.c_double]
        L.orc_ttm_arrays.argtypes = [P] + [P] * 6
        L.orc_ttm_params.argtypes = [P, P, P, P]
        L.orc_ttm_set_fields.argtypes = [P, P, P, P, C.c_double, C.c_uint]
        L.orc_ttm_solve_only.argtypes = [P, C.c_double, C.c_double]
        L.orc_ttm_presolve_solve.argtypes = [P, C.c_double, C.c_double]
        L.orc_integrate_se.argtypes = [P, C.c_int, P, C.c_double, C.c_double]
        L.orc_spline_prepare.argtypes = [C.c_double, P, C.c_int, P]
        L.orc_spline_prepare_adv.argtypes = [P, P, C.c_int, P]
        L.orc_spline_val.argtypes = [C.c_double, P, P, C.c_double]
        L.orc_spline_val.restype = C.c_double
        L.orc_spline_deriv.argtypes = [C.c_double, P, P, C.c_double]
        L.orc_spline_deriv.restype = C.c_double
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Oracle:
    """One reference-algorithm system on the CPU."""

    def __init__(self, l, pbc, masses_amu, names=None):
        L = lib()
        n = len(masses_amu)
        names = names or ["X%d" % i for i in range(n)]
        self._names = [C.c_char_p(s.encode()) for s in names]
        arr = (C.c_char_p * n)(*self._names)
        l = np.ascontiguousarray(l, dtype=np.float64)
        pbc = np.ascontiguousarray(pbc, dtype=np.int32)
        m = np.ascontiguousarray(masses_amu, dtype=np.float64)
        self.h = L.orc_create(_p(l), _p(pbc), n, _p(m), C.cast(arr, C.c_void_p))
        self.L = L
        self.mass_internal = m / MASS_UNIT

    def close(self):
        if self.h:
            self.L.orc_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_lj(self, k1, k2, sigma, eps, cutoff):
        self.L.orc_set_lj(self.h, k1, k2, sigma, eps, cutoff)

    def set_morse(self, k1, k2, D0, alpha, r0, cutoff):
        self.L.orc_set_morse(self.h, k1, k2, D0, alpha, r0, cutoff)

    def load_eam(self, path):
        r = self.L.orc_load_eam(self.h, path.encode())
        if r != 0:
            raise RuntimeError("oracle EAM load failed (%d)" % r)

    def apply_eam(self, k1, k2):
        if self.L.orc_apply_eam(self.h, k1, k2) != 0:
            raise RuntimeError("EAM element not found")

    def eam_tables(self):
        ne, nrho, nr2 = C.c_int(), C.c_int(), C.c_int()
        drho, dr2, rc2 = C.c_double(), C.c_double(), C.c_double()
        self.L.orc_eam_sizes(self.h, C.byref(ne), C.byref(nrho), C.byref(nr2), C.byref(drho), C.byref(dr2),
                             C.byref(rc2))
        ne, nrho, nr2 = ne.value, nrho.value, nr2.value
        F = np.zeros((ne, nrho)); FDD = np.zeros((ne, nrho))
        rho = np.zeros((ne, nr2)); rhoDD = np.zeros((ne, nr2))
        phi = np.zeros((ne, ne, nr2)); phiDD = np.zeros((ne, ne, nr2))
        for e in range(ne):
            a, b, c, d = np.zeros(nrho), np.zeros(nrho), np.zeros(nr2), np.zeros(nr2)
            self.L.orc_eam_tables(self.h, e, _p(a), _p(b), _p(c), _p(d))
            F[e], FDD[e], rho[e], rhoDD[e] = a, b, c, d
            for f in range(ne):
                p, q = np.zeros(nr2), np.zeros(nr2)
                self.L.orc_eam_phi(self.h, e, f, _p(p), _p(q))
                phi[e, f], phiDD[e, f] = p, q
        return dict(nelem=ne, nrho=nrho, nr2=nr2, drho=drho.value, dr2=dr2.value, rc2=rc2.value,
                    F=F, FDD=FDD, rho=rho, rhoDD=rhoDD, phi=phi, phiDD=phiDD)

    def make_grid(self):
        if self.L.orc_make_grid(self.h) != 0:
            raise RuntimeError("oracle grid failed")
        nc = np.zeros(3, np.int32)
        h = np.zeros(3)
        self.L.orc_grid_info(self.h, _p(nc), _p(h))
        self.nc, self.cellh = nc, h
        return nc, h

    def add_atoms(self, x, v=None, group=None, kind=None):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        n = len(x)
        v = np.ascontiguousarray(v if v is not None else np.zeros((n, 3)), dtype=np.float64)
        g = np.ascontiguousarray(group if group is not None else np.zeros(n), dtype=np.int32)
        k = np.ascontiguousarray(kind if kind is not None else np.zeros(n), dtype=np.int32)
        if self.L.orc_add_atoms(self.h, n, _p(x), _p(v), _p(g), _p(k)) != 0:
            raise RuntimeError("atom outside the grid")

    def distribute(self):
        self.L.orc_distribute(self.h)

    def natoms(self):
        return self.L.orc_natoms(self.h)

    def atoms(self):
        n = self.natoms()
        out = dict(x=np.zeros((n, 3)), v=np.zeros((n, 3)), f=np.zeros((n, 3)), aux=np.zeros(n),
                   group=np.zeros(n, np.int32), atomid=np.zeros(n, np.uint32), kind=np.zeros(n, np.int32),
                   cell=np.zeros(n, np.int32))
        self.L.orc_get_atoms(self.h, _p(out["x"]), _p(out["v"]), _p(out["f"]), _p(out["aux"]), _p(out["group"]),
                             _p(out["atomid"]), _p(out["kind"]), _p(out["cell"]))
        ncell = int(np.prod(self.nc))
        cc = np.zeros(ncell, np.int32)
        self.L.orc_cell_counts(self.h, _p(cc))
        out["cell_count"] = cc
        return out

    def refresh(self):
        return self.L.orc_refresh(self.h)

    def forces(self, ag=-1):
        self.L.orc_forces(self.h, ag)

    def vv_start(self, dt, ag=-1, thermostat=0, tau=1.0, T0=0.0):
        self.L.orc_vv_start(self.h, ag, dt, thermostat, tau, T0)

    def vv_finish(self, dt, ag=-1):
        self.L.orc_vv_finish(self.h, ag, dt)

    def se_step(self, dt, ag=-1):
        self.L.orc_se_step(self.h, ag, dt)

    def thermo(self, ag=-1):
        self.L.orc_thermo(self.h, ag)

    def energies(self):
        o = np.zeros(8)
        self.L.orc_get_energies(self.h, _p(o))
        return dict(pe=o[0], ke=o[1], T=o[2], p=o[3:6].copy(), n=int(o[6]), embed=o[7])

    def errors(self):
        last = C.c_int(0)
        n = self.L.orc_error_count(self.h, C.byref(last))
        return n, last.value

    def integrate_vv(self, time, duration, dt, ag=-1, thermostat=0, tau=1.0, T0=0.0):
        t = C.c_double(time)
        steps = self.L.orc_integrate_vv(self.h, ag, C.byref(t), duration, dt, thermostat, tau, T0)
        return steps, t.value

    def ttm_init(self, cfg, te):
        self._ttm_keep = cfg
        self.L.orc_ttm_init(self.h, C.byref(cfg), te)

    def ttm_get_ready(self, ag=-1):
        self.L.orc_ttm_get_ready(self.h, ag)

    def ttm_update(self, time, dt):
        self.L.orc_ttm_update(self.h, time, dt)

    def ttm_arrays(self, n):
        Te, Ti, xi, vz, mass = (np.zeros(n) for _ in range(5))
        num = np.zeros(n, np.uint32)
        self.L.orc_ttm_arrays(self.h, _p(Te), _p(Ti), _p(xi), _p(vz), _p(num), _p(mass))
        return dict(Te=Te, Ti=Ti, xi=xi, vcmz=vz, num=num, mass=mass)

    def ttm_set_fields(self, Te, Ti, num, frontsurf=0.0, min_atoms=1):
        Te = np.ascontiguousarray(Te, dtype=np.float64)
        Ti = np.ascontiguousarray(Ti, dtype=np.float64)
        num = np.ascontiguousarray(num, dtype=np.uint32)
        self.L.orc_ttm_set_fields(self.h, _p(Te), _p(Ti), _p(num), frontsurf, min_atoms)

    def ttm_solve_only(self, time, dt):
        self.L.orc_ttm_solve_only(self.h, time, dt)

    def ttm_presolve_solve(self, time, dt):
        self.L.orc_ttm_presolve_solve(self.h, time, dt)

    def ttm_params(self):
        fs, h = C.c_double(), C.c_double()
        mn = C.c_uint()
        self.L.orc_ttm_params(self.h, C.byref(fs), C.byref(mn), C.byref(h))
        return fs.value, mn.value, h.value

    def integrate_se(self, time, duration, dt, ag=-1):
        t = C.c_double(time)
        steps = self.L.orc_integrate_se(self.h, ag, C.byref(t), duration, dt)
        return steps, t.value


def spline_prepare(h, y):
    y = np.ascontiguousarray(y, dtype=np.float64)
    dd = np.zeros_like(y)
    lib().orc_spline_prepare(h, _p(y), len(y), _p(dd))
    return dd


def spline_prepare_adv(x, y):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    dd = np.zeros_like(y)
    lib().orc_spline_prepare_adv(_p(x), _p(y), len(y), _p(dd))
    return dd


def spline_val(h, y, dd, x):
    return lib().orc_spline_val(h, _p(y), _p(dd), x)


def spline_deriv(h, y, dd, x):
    return lib().orc_spline_deriv(h, _p(y), _p(dd), x)


def timer_is_its_time(t, dt_half, start, interval):
    return bool(lib().orc_timer_is_its_time(t, dt_half, start, interval))
