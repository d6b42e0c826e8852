"""ctypes bindings of the public C API (libfmd.so, include/fmd.h + fmd_ext.h).

Thin wrappers with the reference's function names, for tests, benchmarks and
Python drivers.  Every call goes to the C library; nothing is computed here.
"""
import ctypes as C

import numpy as np

from . import load_fmd, load_fmdx

P = C.c_void_p
D = C.c_double

EVENT_TIMER_TICK, EVENT_FIELD_UPDATE, EVENT_ERROR = 0, 1, 2
TURI_CUSTOM, TURI_TTM_TYPE1, TURI_TTM_TYPE2 = 0, 1, 2
FIELD_MASS, FIELD_VCM, FIELD_TEMPERATURE, FIELD_NUMBER, FIELD_NUMBER_DENSITY, FIELD_TTM_TE, FIELD_TTM_XI = range(7)
GROUP_ALL = -1
SCM_XYZ_ATOMSNUM, SCM_XYZ_SEPARATE, SCM_CSV, SCM_VTF = range(4)   # fmd_SaveConfigMode_t, fmd.h:31-37

EventHandler = C.CFUNCTYPE(None, P, C.c_int, P, P)


class LaserGaussian(C.Structure):
    _fields_ = [("fluence", D), ("reflectance", D), ("t0", D), ("duration", D), ("AbsorptionDepth", D)]


class Zhigilei(C.Structure):
    _fields_ = [("v", D), ("A", D), ("B", D)]


class HeatCapLinear(C.Structure):
    _fields_ = [("gamma", D)]


class TimerTick(C.Structure):
    _fields_ = [("timer", C.c_int)]


class FieldUpdate(C.Structure):
    _fields_ = [("turi", C.c_int), ("field", C.c_int)]


_sigs = {
    "fmd_create": ([], P), "fmd_free": ([P], None),
    "fmd_box_setSize": ([P, D, D, D], None), "fmd_box_setPBC": ([P, C.c_bool, C.c_bool, C.c_bool], None),
    "fmd_box_setSubdomains": ([P, C.c_int, C.c_int, C.c_int], C.c_bool),
    "fmd_matt_setAtomKinds": ([P, C.c_uint, P, P], None),
    "fmd_pot_lj_apply": ([P, C.c_uint, C.c_uint, D, D, D], P),
    "fmd_pot_morse_apply": ([P, C.c_uint, C.c_uint, D, D, D, D], P),
    "fmd_pot_eam_alloy_load": ([P, C.c_char_p], P), "fmd_pot_apply": ([P, C.c_uint, C.c_uint, P], None),
    "fmd_pot_eam_getCutoffRadius": ([P, P], D), "fmd_pot_eam_getLatticeParameter": ([P, P, C.c_char_p], D),
    "fmd_matt_makeCuboidFCC": ([P, D, D, D, C.c_int, C.c_int, C.c_int, D, C.c_uint, C.c_int, D], None),
    "fmd_matt_makeCuboidBCC": ([P, D, D, D, C.c_int, C.c_int, C.c_int, D, C.c_uint, C.c_int, D], None),
    "fmd_matt_makeCuboidSC": ([P, D, D, D, C.c_int, C.c_int, C.c_int, D, C.c_uint, C.c_int, D], None),
    "fmd_matt_makeCuboidFCC_mix": ([P, D, D, D, C.c_int, C.c_int, C.c_int, D, P, C.c_int, D], None),
    "fmd_matt_makeCuboidBCC_mix": ([P, D, D, D, C.c_int, C.c_int, C.c_int, D, P, C.c_int, D], None),
    "fmd_matt_makeCuboidSC_mix": ([P, D, D, D, C.c_int, C.c_int, C.c_int, D, P, C.c_int, D], None),
    "fmd_matt_translate": ([P, C.c_int, D, D, D], None),
    "fmd_matt_addVelocity": ([P, C.c_int, D, D, D], None),
    "fmd_matt_getTotalEnergy": ([P], D), "fmd_matt_getKineticEnergy": ([P], D),
    "fmd_matt_getTemperature": ([P], D), "fmd_matt_getMomentum": ([P, P], None),
    "fmd_matt_findLimits": ([P, C.c_int, P, P], None), "fmd_matt_changeGroupID": ([P, C.c_int, C.c_int], None),
    "fmd_matt_giveMaxwellDistribution": ([P, C.c_int, D], None),
    "fmd_matt_saveConfiguration": ([P], None),
    "fmd_dync_integrate": ([P, C.c_int, D, D], None), "fmd_dync_equilibrate": ([P, C.c_int, D, D, D, D], None),
    "fmd_dync_getTime": ([P], D), "fmd_dync_getTimestep": ([P], D),
    "fmd_io_saveState": ([P, C.c_char_p], None), "fmd_io_loadState": ([P, C.c_char_p, C.c_bool], None),
    "fmd_io_setSaveDirectory": ([P, C.c_char_p], None), "fmd_io_setSaveConfigMode": ([P, C.c_int], None),
    "fmd_io_setShowErrorMessages": ([P, C.c_bool], None),
    "fmd_setEventHandler": ([P, P, EventHandler], None),
    "fmd_timer_makeSimple": ([P, D, D, D], C.c_int),
    "fmd_turi_add": ([P, C.c_int, C.c_int, C.c_int, C.c_int, D, D], C.c_int),
    "fmd_field_find": ([P, C.c_int, C.c_int], C.c_int), "fmd_field_add": ([P, C.c_int, C.c_int, D], C.c_int),
    "fmd_field_getArray": ([P, C.c_int, C.c_int, P, P], P), "fmd_array3s_free": ([P], None),
    "_fmd_ttm_setHeatCapacity_linear": ([P, C.c_int, HeatCapLinear], None),
    "_fmd_ttm_setHeatCapacity_file": ([P, C.c_int, C.c_char_p], None),
    "_fmd_ttm_setHeatConductivity_constant1": ([P, C.c_int, D], None),
    "_fmd_ttm_setHeatConductivity_zhigilei": ([P, C.c_int, Zhigilei], None),
    "_fmd_ttm_setCouplingFactor_constant1": ([P, C.c_int, D], None),
    "_fmd_ttm_setCouplingFactor_file": ([P, C.c_int, C.c_char_p], None),
    "fmd_ttm_setElectronTemperature": ([P, C.c_int, D], None),
    "fmd_ttm_setTimestepRatio": ([P, C.c_int, C.c_int], None),
    "fmd_ttm_setCellActivationFraction": ([P, C.c_int, D], None),
    "_fmd_ttm_setLaserSource_gaussian": ([P, C.c_int, LaserGaussian], None),
    "fmd_ext_loadAtoms": ([P, C.c_uint, P, P, P, P], None), "fmd_ext_distribute": ([P], C.c_bool),
    "fmd_ext_getDevice": ([P], P), "fmd_ext_setPrecision": ([P, C.c_int], None),
    "fmd_ext_getAtoms": ([P, P, P, P, P, P, P], C.c_uint), "fmd_ext_getPotentialEnergy": ([P], D),
    "fmd_ext_ttm_setHeatConductivity_proportional": ([P, C.c_int, D], None),
    "fmd_ext_ttm_useGridSolver": ([P, C.c_int, C.c_bool], None),
    "fmd_version_getString": ([], C.c_char_p),
}


def lib():
    L = load_fmd()
    if not getattr(L, "_sigs_done", False):
        for name, (args, res) in _sigs.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        L._sigs_done = True
    return L


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Fmd:
    """One fmd_t instance."""

    def __init__(self):
        load_fmdx()
        self.L = lib()
        self.md = self.L.fmd_create()
        self._keep = []

    def free(self):
        if self.md:
            self.L.fmd_free(self.md)
            self.md = None

    def __getattr__(self, name):
        f = getattr(self.L, name)
        return lambda *a: f(self.md, *a)

    def set_kinds(self, names, masses):
        arr = (C.c_char_p * len(names))(*[n.encode() for n in names])
        m = (C.c_double * len(masses))(*masses)
        self._keep += [arr, m]
        self.L.fmd_matt_setAtomKinds(self.md, len(names), C.cast(arr, P), C.cast(m, P))

    def load_atoms(self, x, v=None, group=None, kind=None):
        x = np.ascontiguousarray(x, dtype=np.float64)
        n = len(x)
        v = np.ascontiguousarray(v if v is not None else np.zeros((n, 3)), dtype=np.float64)
        g = None if group is None else np.ascontiguousarray(group, dtype=np.int32)
        k = None if kind is None else np.ascontiguousarray(kind, dtype=np.uint32)
        self.L.fmd_ext_loadAtoms(self.md, n, _p(x), _p(v), _p(g), _p(k))

    def atoms(self):
        n = self.L.fmd_ext_getAtoms(self.md, None, None, None, None, None, None)
        x, v, f = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        g, a, k = np.zeros(n, np.int32), np.zeros(n, np.uint32), np.zeros(n, np.uint32)
        self.L.fmd_ext_getAtoms(self.md, _p(x), _p(v), _p(f), _p(g), _p(a), _p(k))
        return dict(x=x, v=v, f=f, group=g, atomid=a, kind=k)

    def momentum(self):
        p = (C.c_double * 3)()
        self.L.fmd_matt_getMomentum(self.md, p)
        return np.array(list(p))

    def field_array(self, turi, field, dtype=np.float64):
        arr = C.c_void_p()
        dims = (C.c_uint * 3)()
        a3s = self.L.fmd_field_getArray(self.md, turi, field, C.byref(arr), dims)
        if not a3s:
            return None
        n = dims[0] * dims[1] * dims[2]
        ptr = C.cast(arr, C.POINTER(C.POINTER(C.POINTER(C.c_uint if dtype == np.uint32 else C.c_double))))
        vals = np.array([ptr[0][0][i] for i in range(n)], dtype=dtype)
        self.L.fmd_array3s_free(a3s)
        return vals
