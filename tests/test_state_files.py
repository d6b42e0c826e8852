"""State files (fmd_io_saveState) byte for byte against the reference's format.

misc.c:326-339 writes, for atoms still on the host grid: the time ("%.16e"),
the atom count ("%u"), the box ("%.16e\\t%.16e\\t%.16e"), the periodicity
("%d %d %d"), then per atom, cell by cell in the grid's storage order: the
kind's name and the group ("%s %d"), the position and the velocity (each
"%.16e\\t%.16e\\t%.16e").  The expected text is composed here from the atoms the
library holds (fmd_ext_getAtoms, cell order) and compared with the file.  The
same format after dynamics, written from device memory, is checked in
tests/test_gpu_resume.py.
"""
import ctypes as C

import numpy as np

from fmd_amd import api


def expected_state(time, L, pbc, names, atoms):
    out = ["%.16e" % time, "%u" % len(atoms["x"]), "%.16e\t%.16e\t%.16e" % tuple(L), "%d %d %d" % tuple(pbc)]
    for i in range(len(atoms["x"])):
        out.append("%s %d" % (names[atoms["kind"][i]], atoms["group"][i]))
        out.append("%.16e\t%.16e\t%.16e" % tuple(atoms["x"][i]))
        out.append("%.16e\t%.16e\t%.16e" % tuple(atoms["v"][i]))
    return "\n".join(out) + "\n"


def test_state_file_from_host_grid(tmp_path):
    f = api.Fmd()
    L = [30.0, 25.0, 40.0]
    f.fmd_box_setSize(*L)
    f.fmd_box_setPBC(True, False, True)
    f.set_kinds(["Cu", "Ar"], [63.546, 39.948])
    f.fmd_pot_lj_apply(0, 0, 2.3, 0.4, 5.5)
    f.fmd_pot_lj_apply(1, 1, 3.405, 0.0104, 8.5)
    f.fmd_pot_lj_apply(0, 1, 2.87, 0.0652, 7.0)
    f.fmd_matt_makeCuboidFCC(0.0, 0.0, 0.0, 4, 3, 5, 3.615, 0, 2, 300.0)
    f.fmd_matt_makeCuboidFCC(15.0, 4.0, 5.0, 2, 2, 3, 5.26, 1, 7, 80.0)
    f.fmd_io_setSaveDirectory((str(tmp_path) + "/").encode())
    f.fmd_io_saveState(b"host.stt")
    got = open(tmp_path / "host.stt").read()
    a = f.atoms()
    assert len(a["x"]) == 4 * (4 * 3 * 5 + 2 * 2 * 3)
    assert got == expected_state(0.0, L, [1, 0, 1], ["Cu", "Ar"], a)
    f.free()
