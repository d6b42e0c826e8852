"""CPU checks of the drop-in boundary: the libraries load, export every symbol
the public headers declare, and programs written against the reference's
fmd.h compile and link unchanged."""
import ctypes as C
import os
import re
import subprocess

import pytest

import fmd_amd as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_EXAMPLES = "/root/reference/examples"


def declared(header, prefix_re):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = set(re.findall(r"\b(" + prefix_re + r"\w*)\s*\(", txt))
    return sorted(n for n in names if not n.startswith("fmd_ttm_set") or n.startswith("fmd_ttm_set"))


def exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", F.lib_path(lib)], capture_output=True, text=True).stdout
    return {l.split()[-1] for l in out.splitlines() if " T " in l}


def test_libfmdx_exports_every_fmdx_symbol():
    names = declared("fmdx.h", "fmdx_")
    assert len(names) > 30
    missing = [n for n in names if n not in exported("libfmdx.so")]
    assert not missing, missing


def test_libfmd_exports_every_fmd_symbol():
    # every function of fmd.h (macros fmd_ttm_setHeatCapacity etc. expand to _fmd_ttm_* functions)
    names = [n for n in declared("fmd.h", "_?fmd_") if n not in
             ("fmd_ttm_setHeatCapacity", "fmd_ttm_setHeatConductivity", "fmd_ttm_setCouplingFactor",
              "fmd_ttm_setLaserSource")]
    names += declared("fmd_ext.h", "fmd_ext_")
    assert len(names) >= 69 + 6   # the reference's fmd.h functions + the extensions
    missing = [n for n in names if n not in exported("libfmd.so")]
    assert not missing, missing


def test_libraries_load_without_gpu():
    lib = F.load_fmdx()
    n = C.c_int(-1)
    assert lib.fmdx_device_count(C.byref(n)) == 0
    assert n.value >= 0
    F.load_fmd()


def test_system_create_reports_no_device_cleanly():
    if F.device_count() > 0:
        pytest.skip("a GPU is visible")
    lib = F.load_fmdx()
    h = C.c_void_p()
    d3 = (C.c_double * 3)(10, 10, 10)
    i3 = (C.c_int * 3)(1, 1, 1)
    nc = (C.c_int * 3)(3, 3, 3)
    m = (C.c_double * 1)(1.0)
    assert lib.fmdx_system_create(C.byref(h), d3, i3, nc, d3, 1, m, 0) == -5   # FMDX_ENODEV


@pytest.mark.skipif(not os.path.isdir(REF_EXAMPLES), reason="reference examples not present")
@pytest.mark.parametrize("ex", ["01_argon.c", "02_collision.c", "03_ttm1.c", "04_ttm2.c"])
def test_reference_examples_compile_unchanged(ex, tmp_path):
    from fmd_amd.build import build_program
    out = build_program(os.path.join(REF_EXAMPLES, ex), str(tmp_path / (ex + ".x")))
    assert os.path.exists(out)
