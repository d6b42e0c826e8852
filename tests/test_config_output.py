"""fmd_matt_saveConfiguration in its four modes (matter.c:650-829, misc.c:591-600).

The expected files are written here from the oracle's cell-ordered atoms and the
reference's formats: atoms grouped by kind in kind order, each kind in cell order;
positions rounded to float and printed "%.2f"; the group id as the per-atom value
("%.4f"); XYZ_ATOMSNUM appends to "<dir><N>.xyz" while N is unchanged; the other
modes number their files "%05d" from one shared counter.  Python's "%" and glibc's
printf both round the float's exact value, so the texts must agree byte for byte.
Before the first dynamics call the configuration comes from the host's cells; the
GPU variant (test_gpu_api.py) writes it after integration, from the device.
"""
import os

import numpy as np
import pytest

from fmd_amd import api
from oracle.oracle import Oracle
from tests.pairsys import lattice_fcc

NAMES, MASSES = ["Ar", "Kr"], [39.948, 83.798]
LJ = {(0, 0): (3.405, 0.0104), (0, 1): (3.55, 0.0140), (1, 1): (3.65, 0.0170)}
RC = 8.6


def system(rng):
    x = lattice_fcc((5, 5, 5), 5.6, rng, jitter=0.1)
    L = [5 * 5.6] * 3
    kind = (rng.random(len(x)) < 0.35).astype(np.int32)
    group = rng.integers(0, 3, len(x)).astype(np.int32)
    return x, L, kind, group


def expected(atoms, mode, n_calls_before=0):
    f32 = atoms["x"].astype(np.float32).astype(np.float64)
    order = [k for i in range(len(NAMES)) for k in np.nonzero(atoms["kind"] == i)[0]]
    n = len(order)
    if mode in (api.SCM_XYZ_ATOMSNUM, api.SCM_XYZ_SEPARATE):
        return "%d\n\n" % n + "".join("%s\t%.2f\t%.2f\t%.2f\n" % (NAMES[atoms["kind"][k]], *f32[k]) for k in order)
    if mode == api.SCM_CSV:
        return "".join("%.2f, %.2f, %.2f, %d, %.4f\n" % (*f32[k], atoms["kind"][k], float(atoms["group"][k]))
                       for k in order)
    head, aid = "", 0
    for i in range(len(NAMES)):
        c = int(np.sum(atoms["kind"] == i))
        head += "atom %d:%d\tname %s\n" % (aid, aid + c - 1, NAMES[i])
        aid += c
    beta = "".join("%d\tbeta %.4f\n" % (j, float(atoms["group"][k])) for j, k in enumerate(order))
    pos = "".join("%.2f\t%.2f\t%.2f\n" % tuple(f32[k]) for k in order)
    return head + beta + "timestep\n" + pos


def oracle_atoms(x, L, kind, group):
    o = Oracle(L, [1, 1, 1], MASSES, NAMES)
    for (a, b), (s, e) in LJ.items():
        o.set_lj(a, b, s, e, RC)
    o.make_grid()
    o.add_atoms(x, None, group, kind)
    o.distribute()
    return o.atoms()


def api_system(x, L, kind, group, d, v=None):
    f = api.Fmd()
    f.fmd_box_setSize(*L)
    f.fmd_box_setPBC(True, True, True)
    f.set_kinds(NAMES, MASSES)
    for (a, b), (s, e) in LJ.items():
        f.fmd_pot_lj_apply(a, b, s, e, RC)
    f.load_atoms(x, v, group, kind)
    f.fmd_io_setSaveDirectory((str(d) + os.sep).encode())
    return f


def test_four_modes_from_host_cells(rng, tmp_path):
    x, L, kind, group = system(rng)
    ref = oracle_atoms(x, L, kind, group)
    f = api_system(x, L, kind, group, tmp_path)
    n = len(x)
    f.fmd_io_setSaveConfigMode(api.SCM_XYZ_ATOMSNUM)
    f.fmd_matt_saveConfiguration()
    f.fmd_matt_saveConfiguration()          # appended to the same file
    f.fmd_io_setSaveConfigMode(api.SCM_XYZ_SEPARATE)
    f.fmd_matt_saveConfiguration()          # 00000.xyz
    f.fmd_io_setSaveConfigMode(api.SCM_CSV)
    f.fmd_matt_saveConfiguration()          # 00001.csv
    f.fmd_io_setSaveConfigMode(api.SCM_VTF)
    f.fmd_matt_saveConfiguration()          # 00002.vtf
    f.free()                                # closes the XYZ_ATOMSNUM file
    xyz = expected(ref, api.SCM_XYZ_ATOMSNUM)
    assert (tmp_path / ("%d.xyz" % n)).read_text() == xyz + xyz
    assert (tmp_path / "00000.xyz").read_text() == xyz
    assert (tmp_path / "00001.csv").read_text() == expected(ref, api.SCM_CSV)
    assert (tmp_path / "00002.vtf").read_text() == expected(ref, api.SCM_VTF)
    assert sorted(os.listdir(tmp_path)) == sorted(["%d.xyz" % n, "00000.xyz", "00001.csv", "00002.vtf"])
