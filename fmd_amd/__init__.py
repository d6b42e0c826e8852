"""fmd_amd -- MI355X-native FMD step engine.

The product is two in-tree shared libraries:

* ``lib/libfmd.so``  -- the reference's public C API (``include/fmd.h``), C host code;
* ``lib/libfmdx.so`` -- the gfx950 kernels behind a C-ABI (``include/fmdx.h``).

This module loads them with ctypes for tests and benchmarks.  There is no
Python or CPU fallback: if a library is missing, loading raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# the built libraries: fmd_amd/lib, or another build of them (FMD_AMD_LIB_DIR)
LIB_DIR = os.environ.get("FMD_AMD_LIB_DIR") or os.path.join(HERE, "lib")
ROOT = os.path.dirname(HERE)

FMDX_OK = 0
FMDX_EHIP, FMDX_EARG, FMDX_ENOMEM, FMDX_ESTATE, FMDX_ENODEV, FMDX_ESIZE = -1, -2, -3, -4, -5, -6
FMDX_NOALLOC_MAX_CELLS = 590000
FMDX_FORCES_ENERGY = 1
FMDX_PREC_FP64 = 0
FMDX_PREC_MIXED = 1
FMDX_COMPACT_OFF, FMDX_COMPACT_AUTO, FMDX_COMPACT_TIGHT = 0, 1, 2
POT_LJ, POT_EAM, POT_MORSE = 1, 2, 3
MASS_UNIT = 9.6485332907e+03
KB = 8.6173303e-5

ERR_POSITION, ERR_RHO_RANGE, ERR_OVERFLOW, ERR_TTM_CE, ERR_TTM_G, ERR_FORCE = 1, 2, 4, 8, 16, 32


class FmdxError(RuntimeError):
    pass


class Thermo(C.Structure):
    _fields_ = [("pe", C.c_double), ("embed", C.c_double), ("mv2", C.c_double), ("ke", C.c_double),
                ("temperature", C.c_double), ("momentum", C.c_double * 3), ("n_active", C.c_double),
                ("natoms", C.c_uint), ("max_cell_count", C.c_uint), ("err", C.c_uint), ("cap", C.c_uint),
                ("bad_force_cell", C.c_uint), ("bad_position_cell", C.c_uint), ("replay_mismatch", C.c_uint), ("exact_unmatched", C.c_uint),
                ("audit_atoms", C.c_uint), ("audit_max_cell", C.c_uint), ("audit_nonfinite", C.c_uint)]


class TtmCfg(C.Structure):
    _fields_ = [("type", C.c_int), ("ncz", C.c_int), ("ratio", C.c_int), ("tcellh_z", C.c_double),
                ("tcell_volume", C.c_double), ("C_gamma", C.c_double), ("K", C.c_double), ("G", C.c_double),
                ("Kzh_A", C.c_double), ("Kzh_B", C.c_double), ("Kzh_v2", C.c_double),
                ("nC", C.c_int), ("nG", C.c_int),
                ("T_C", C.c_void_p), ("Ct", C.c_void_p), ("C_DD", C.c_void_p),
                ("T_G", C.c_void_p), ("Gt", C.c_void_p), ("G_DD", C.c_void_p),
                ("laser_factor", C.c_double), ("laser_absdepth_inv", C.c_double),
                ("laser_2sig2_inv", C.c_double), ("laser_t0", C.c_double), ("laser_tdiff", C.c_double),
                ("frontsurf", C.c_double), ("min_atoms", C.c_uint),
                ("ncx", C.c_int), ("ncy", C.c_int), ("tcellh_x", C.c_double), ("tcellh_y", C.c_double),
                ("kmode", C.c_int), ("solver3d", C.c_int)]


_fmdx = None
_fmd = None


def lib_path(name):
    return os.path.join(LIB_DIR, name)


def load_fmdx():
    """The kernel library; raises if it has not been built."""
    global _fmdx
    if _fmdx is None:
        p = lib_path("libfmdx.so")
        if not os.path.exists(p):
            raise FmdxError("libfmdx.so is not built (run __graft_entry__.build())")
        lib = C.CDLL(p, mode=C.RTLD_GLOBAL)
        P = C.c_void_p
        d3 = C.POINTER(C.c_double)
        lib.fmdx_system_create.argtypes = [C.POINTER(P), d3, C.POINTER(C.c_int), C.POINTER(C.c_int), d3, C.c_int,
                                           d3, C.c_int]
        lib.fmdx_system_destroy.argtypes = [P]
        lib.fmdx_system_destroy.restype = C.c_int
        lib.fmdx_set_precision.argtypes = [P, C.c_int]
        lib.fmdx_set_pair.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]
        lib.fmdx_set_eam.argtypes = [P, C.c_int, C.c_int, C.c_double, C.c_int, C.c_double, C.c_double,
                                     P, P, P, P, P, P, P]
        lib.fmdx_upload.argtypes = [P, C.c_uint, P, P, P, P, P, P]
        lib.fmdx_download.argtypes = [P, C.POINTER(C.c_uint), P, P, P, P, P, P, P, P]
        lib.fmdx_natoms.argtypes = [P, C.POINTER(C.c_uint)]
        lib.fmdx_ncells.argtypes = [P, C.POINTER(C.c_uint)]
        lib.fmdx_update_forces.argtypes = [P, C.c_int, C.c_int]
        lib.fmdx_update_forces_noalloc.argtypes = [P, C.c_int, C.c_int]
        lib.fmdx_update_forces_ex.argtypes = [P, C.c_int, C.c_int, C.c_int]
        lib.fmdx_vv_start.argtypes = [P, C.c_int, C.c_double, C.c_int, C.c_double, C.c_double]
        lib.fmdx_vv_finish.argtypes = [P, C.c_int, C.c_double]
        lib.fmdx_se_step.argtypes = [P, C.c_int, C.c_double]
        lib.fmdx_refresh.argtypes = [P]
        lib.fmdx_thermo.argtypes = [P, C.c_int]
        lib.fmdx_add_velocity.argtypes = [P, C.c_int, C.c_double, C.c_double, C.c_double]
        lib.fmdx_read_thermo.argtypes = [P, C.POINTER(Thermo)]
        lib.fmdx_clear_errors.argtypes = [P]
        lib.fmdx_audit.argtypes = [P]
        lib.fmdx_sync.argtypes = [P]
        lib.fmdx_set_capacity.argtypes = [P, C.c_uint]
        lib.fmdx_ttm_setup.argtypes = [P, C.POINTER(TtmCfg), P]
        lib.fmdx_ttm_set_frontsurf.argtypes = [P, C.c_double, C.c_uint]
        lib.fmdx_ttm_update.argtypes = [P, C.c_double, C.c_double]
        lib.fmdx_ttm_fields.argtypes = [P]
        lib.fmdx_ttm_presolve.argtypes = [P]
        lib.fmdx_ttm_solve.argtypes = [P, C.c_double, C.c_double]
        lib.fmdx_ttm_download.argtypes = [P, P, P, P, P, P, P]
        lib.fmdx_ttm_upload_te.argtypes = [P, P]
        lib.fmdx_find_limits.argtypes = [P, C.c_int, d3, d3]
        lib.fmdx_event_record.argtypes = [P, C.c_int]
        lib.fmdx_event_elapsed_ms.argtypes = [P, C.c_int, C.c_int, C.POINTER(C.c_float)]
        lib.fmdx_profile_enable.argtypes = [P, C.c_int]
        lib.fmdx_profile_read.argtypes = [P, d3, C.POINTER(C.c_uint)]
        lib.fmdx_device_count.argtypes = [C.POINTER(C.c_int)]
        lib.fmdx_strerror.argtypes = [C.c_int]
        lib.fmdx_strerror.restype = C.c_char_p
        lib.fmdx_batch_create.argtypes = [C.POINTER(P), C.c_int]
        lib.fmdx_batch_destroy.argtypes = [P]
        lib.fmdx_batch_destroy.restype = None
        lib.fmdx_batch_add.argtypes = [P, P, C.c_double, C.c_int, C.c_double, C.c_double]
        lib.fmdx_batch_remove.argtypes = [P, P]
        lib.fmdx_batch_size.argtypes = [P, C.POINTER(C.c_int)]
        lib.fmdx_batch_use_graph.argtypes = [P, C.c_int]
        lib.fmdx_batch_set_check_interval.argtypes = [P, C.c_int]
        lib.fmdx_batch_set_compaction.argtypes = [P, C.c_int]
        lib.fmdx_batch_step.argtypes = [P, C.c_int]
        lib.fmdx_batch_status.argtypes = [P, C.POINTER(Thermo)]
        lib.fmdx_batch_sync.argtypes = [P]
        lib.fmdx_memory.argtypes = [P, C.POINTER(C.c_size_t)]
        _fmdx = lib
    return _fmdx


def load_fmd():
    """The public C API library (links libfmdx.so)."""
    global _fmd
    if _fmd is None:
        load_fmdx()
        p = lib_path("libfmd.so")
        if not os.path.exists(p):
            raise FmdxError("libfmd.so is not built (run __graft_entry__.build())")
        _fmd = C.CDLL(p)
    return _fmd


def device_count():
    n = C.c_int(0)
    load_fmdx().fmdx_device_count(C.byref(n))
    return n.value


def _ck(r, what):
    if r != FMDX_OK:
        msg = load_fmdx().fmdx_strerror(r).decode()
        raise FmdxError("%s failed: %s (%d)" % (what, msg, r))


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _d3(v):
    return (C.c_double * 3)(*v)


def _i3(v):
    return (C.c_int * 3)(*[int(x) for x in v])


class DeviceSystem:
    """One system resident in HBM, driven through the C-ABI of libfmdx.so."""

    def __init__(self, l, pbc, nc, cellh, masses_internal, precision=FMDX_PREC_FP64):
        self.lib = load_fmdx()
        self.h = C.c_void_p()
        m = (C.c_double * len(masses_internal))(*masses_internal)
        _ck(self.lib.fmdx_system_create(C.byref(self.h), _d3(l), _i3(pbc), _i3(nc), _d3(cellh),
                                        len(masses_internal), m, precision), "fmdx_system_create")
        self.nc = tuple(int(x) for x in nc)
        self.ncell = int(np.prod(self.nc))

    def close(self):
        if self.h:
            _ck(self.lib.fmdx_system_destroy(self.h), "fmdx_system_destroy")
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_precision(self, p):
        _ck(self.lib.fmdx_set_precision(self.h, p), "fmdx_set_precision")

    def set_pair(self, k1, k2, cat, rc2, p0=0.0, p1=0.0, p2=0.0):
        _ck(self.lib.fmdx_set_pair(self.h, k1, k2, cat, rc2, p0, p1, p2), "fmdx_set_pair")

    def set_eam(self, tables, elem_of_kind):
        """tables: dict with nelem, nrho, drho, nr2, dr2, rc2, F, FDD, rho, rhoDD, phi, phiDD (float64 arrays)."""
        t = {k: (np.ascontiguousarray(v, dtype=np.float64) if isinstance(v, np.ndarray) else v)
             for k, v in tables.items()}
        eok = np.full(8, -1, dtype=np.int32)
        eok[:len(elem_of_kind)] = elem_of_kind
        self._eam_keep = (t, eok)
        _ck(self.lib.fmdx_set_eam(self.h, t["nelem"], t["nrho"], t["drho"], t["nr2"], t["dr2"], t["rc2"],
                                  _ptr(t["F"]), _ptr(t["FDD"]), _ptr(t["rho"]), _ptr(t["rhoDD"]),
                                  _ptr(t["phi"]), _ptr(t["phiDD"]), _ptr(eok)), "fmdx_set_eam")

    def upload(self, cell_count, x, v, group=None, atomid=None, kind=None):
        n = int(np.sum(cell_count))
        cc = np.ascontiguousarray(cell_count, dtype=np.uint32)
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
        g = np.ascontiguousarray(group if group is not None else np.zeros(n), dtype=np.int32)
        a = np.ascontiguousarray(atomid if atomid is not None else np.arange(n), dtype=np.uint32)
        k = np.ascontiguousarray(kind if kind is not None else np.zeros(n), dtype=np.uint32)
        _ck(self.lib.fmdx_upload(self.h, n, _ptr(cc), _ptr(x), _ptr(v), _ptr(g), _ptr(a), _ptr(k)), "fmdx_upload")

    def download(self):
        n = C.c_uint(0)
        _ck(self.lib.fmdx_download(self.h, C.byref(n), None, None, None, None, None, None, None, None),
            "fmdx_download")
        n = n.value
        out = dict(cell_count=np.zeros(self.ncell, np.uint32), x=np.zeros((n, 3)), v=np.zeros((n, 3)),
                   f=np.zeros((n, 3)), aux=np.zeros(n), group=np.zeros(n, np.int32),
                   atomid=np.zeros(n, np.uint32), kind=np.zeros(n, np.uint32))
        nn = C.c_uint(0)
        _ck(self.lib.fmdx_download(self.h, C.byref(nn), _ptr(out["cell_count"]), _ptr(out["x"]), _ptr(out["v"]),
                                   _ptr(out["f"]), _ptr(out["aux"]), _ptr(out["group"]), _ptr(out["atomid"]),
                                   _ptr(out["kind"])), "fmdx_download")
        return out

    def forces(self, active_group=-1, ttm=0):
        _ck(self.lib.fmdx_update_forces(self.h, active_group, ttm), "fmdx_update_forces")

    def vv_start(self, dt, active_group=-1, thermostat=0, tau=1.0, T0=0.0):
        _ck(self.lib.fmdx_vv_start(self.h, active_group, dt, thermostat, tau, T0), "fmdx_vv_start")

    def vv_finish(self, dt, active_group=-1):
        _ck(self.lib.fmdx_vv_finish(self.h, active_group, dt), "fmdx_vv_finish")

    def se_step(self, dt, active_group=-1):
        _ck(self.lib.fmdx_se_step(self.h, active_group, dt), "fmdx_se_step")

    def refresh(self):
        _ck(self.lib.fmdx_refresh(self.h), "fmdx_refresh")

    def thermo(self, active_group=-1):
        _ck(self.lib.fmdx_thermo(self.h, active_group), "fmdx_thermo")

    def read(self):
        t = Thermo()
        _ck(self.lib.fmdx_read_thermo(self.h, C.byref(t)), "fmdx_read_thermo")
        return t

    def audit(self):
        """the state check (atoms over the cells, fullest cell, non-finite x, v, F); returns the record"""
        _ck(self.lib.fmdx_audit(self.h), "fmdx_audit")
        return self.read()

    def sync(self):
        _ck(self.lib.fmdx_sync(self.h), "fmdx_sync")

    def set_capacity(self, cap):
        _ck(self.lib.fmdx_set_capacity(self.h, cap), "fmdx_set_capacity")

    def event(self, slot):
        _ck(self.lib.fmdx_event_record(self.h, slot), "fmdx_event_record")

    def elapsed_ms(self, a, b):
        ms = C.c_float(0)
        _ck(self.lib.fmdx_event_elapsed_ms(self.h, a, b, C.byref(ms)), "fmdx_event_elapsed_ms")
        return ms.value

    def profile(self, on):
        _ck(self.lib.fmdx_profile_enable(self.h, 1 if on else 0), "fmdx_profile_enable")

    def profile_read(self):
        ms = (C.c_double * 8)()
        n = C.c_uint(0)
        _ck(self.lib.fmdx_profile_read(self.h, ms, C.byref(n)), "fmdx_profile_read")
        return list(ms), n.value

    # TTM
    def ttm_setup(self, cfg, te_init):
        self._ttm_keep = te_init = np.ascontiguousarray(te_init, dtype=np.float64)
        _ck(self.lib.fmdx_ttm_setup(self.h, C.byref(cfg), _ptr(te_init)), "fmdx_ttm_setup")

    def ttm_update(self, time, dt):
        _ck(self.lib.fmdx_ttm_update(self.h, time, dt), "fmdx_ttm_update")

    def ttm_download(self, n):
        Te, Ti, xi, vcm = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(3 * n)
        num, mass = np.zeros(n, np.uint32), np.zeros(n)
        _ck(self.lib.fmdx_ttm_download(self.h, _ptr(Te), _ptr(Ti), _ptr(xi), _ptr(vcm), _ptr(num), _ptr(mass)),
            "fmdx_ttm_download")
        return dict(Te=Te, Ti=Ti, xi=xi, vcm=vcm.reshape(-1, 3), num=num, mass=mass)


def _handle(s):
    """a system handle: DeviceSystem, or a raw fmdx_system_t pointer"""
    return s.h if isinstance(s, DeviceSystem) else C.c_void_p(s if isinstance(s, int) else s.value)


class Batch:
    """Independent systems advanced together by the batch engine of libfmdx.so."""

    def __init__(self, precision=FMDX_PREC_MIXED):
        self.lib = load_fmdx()
        self.h = C.c_void_p()
        _ck(self.lib.fmdx_batch_create(C.byref(self.h), precision), "fmdx_batch_create")
        self.members = []

    def close(self):
        if self.h:
            self.lib.fmdx_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, s, dt, thermostat=0, tau=1.0, T0=0.0):
        _ck(self.lib.fmdx_batch_add(self.h, _handle(s), dt, thermostat, tau, T0), "fmdx_batch_add")
        self.members.append(s)

    def remove(self, s):
        _ck(self.lib.fmdx_batch_remove(self.h, _handle(s)), "fmdx_batch_remove")
        self.members.remove(s)

    def size(self):
        n = C.c_int(0)
        _ck(self.lib.fmdx_batch_size(self.h, C.byref(n)), "fmdx_batch_size")
        return n.value

    def use_graph(self, on):
        _ck(self.lib.fmdx_batch_use_graph(self.h, 1 if on else 0), "fmdx_batch_use_graph")

    def set_check_interval(self, steps):
        _ck(self.lib.fmdx_batch_set_check_interval(self.h, steps), "fmdx_batch_set_check_interval")

    def set_compaction(self, mode):
        """FMDX_COMPACT_OFF, FMDX_COMPACT_AUTO (default) or FMDX_COMPACT_TIGHT"""
        _ck(self.lib.fmdx_batch_set_compaction(self.h, mode), "fmdx_batch_set_compaction")

    def step(self, n=1):
        _ck(self.lib.fmdx_batch_step(self.h, n), "fmdx_batch_step")

    def status(self):
        n = self.size()
        out = (Thermo * max(n, 1))()
        _ck(self.lib.fmdx_batch_status(self.h, out), "fmdx_batch_status")
        return list(out)[:n]

