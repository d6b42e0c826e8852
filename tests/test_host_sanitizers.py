"""The host library's CPU paths under AddressSanitizer and UndefinedBehaviorSanitizer.

The host sources of libfmd.so are compiled with gcc -fsanitize=address,undefined
together with tests/host/asan_driver.c (linked to the device library, whose
entry points these paths never call) and run on the setfl tables of the tests
and workloads.  Any out-of-bounds access or undefined behaviour aborts the run.

It pins the setfl loader's r -> r^2 resampling (EAM_convert_r_to_r2,
eam.c:324-341): its last sample sits on the last knot, where the spline
evaluation (cspline.c:76-89) reads the knot after it with weight zero.  The
loader's arrays carry one zero entry past the table for that read; without it
the read takes whatever follows the array, and a NaN there made every density
and force of a run NaN, depending on the heap's history.  The state file written
from the builders is read back and written again byte for byte.
"""
import os
import subprocess

import pytest

from fmd_amd import build as B
from fmd_amd import workloads as W
from tests.pairsys import write_setfl_two

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def asan_driver(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("asan") / "asan_driver")
    srcs = [os.path.join(B.HOST_SRC, f) for f in B.HOST_FILES] + [os.path.join(HERE, "host", "asan_driver.c")]
    r = subprocess.run(["gcc", "-std=gnu11", "-O1", "-g", "-fno-omit-frame-pointer",
                        "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                        "-I" + B.INC, "-I" + B.HOST_SRC] + srcs +
                       ["-L" + B.LIB, "-lfmdx", "-Wl,-rpath," + B.LIB, "-lm", "-o", out],
                       capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail("sanitizer build failed:\n" + r.stderr[-4000:])
    return out


def test_host_paths_clean_under_sanitizers(asan_driver, tmp_path, cu01):
    synth = str(tmp_path / "Cu_synth.eam.alloy")
    W.write_setfl(synth, nr=5000, nrho=5000)
    two = write_setfl_two(str(tmp_path / "two.eam.alloy"))
    save = str(tmp_path / "out") + "/"
    os.makedirs(save)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    # freed heap memory filled with 0xFF bytes (a NaN as a double): a read past an array that lands
    # in such memory turns into NaN tables even where the sanitizer's redzones would not see it
    env["MALLOC_PERTURB_"] = "255"
    r = subprocess.run([asan_driver, save, synth, cu01, two], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-6000:]
    assert "host paths ok" in r.stdout
    a = open(os.path.join(save, "asan.stt")).read()
    b = open(os.path.join(save, "asan2.stt")).read()
    assert a == b and a.count("\n") > 864 * 2
