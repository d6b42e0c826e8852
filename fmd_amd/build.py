"""Build recipe for the in-tree libraries.

    libfmdx.so  hand-written gfx950 HIP kernels + C-ABI (include/fmdx.h), hipcc
    libfmd.so   C11 host library exporting include/fmd.h, gcc, linked to libfmdx.so
    oracle      the CPU checker (oracle/), gcc, built into oracle/_build/

Run ``python -m fmd_amd.build`` or call :func:`build_all`.  Outputs stay in
the tree so that they travel with the repository snapshot to a GPU host.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "fmd_amd")
INC = os.path.join(ROOT, "include")
DEV_SRC = os.path.join(PKG, "csrc", "device")
HOST_SRC = os.path.join(PKG, "csrc", "host")
LIB = os.path.join(PKG, "lib")
OBJ = os.path.join(ROOT, "_build", "obj")
ORACLE = os.path.join(ROOT, "oracle")
ORACLE_OUT = os.path.join(ORACLE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FMD_OFFLOAD_ARCH", "gfx950")
# packed FP32 instructions are kept out of every kernel (-packed-fp32-ops)
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-munsafe-fp-atomics",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", "-Wno-unused-value", "-I" + INC, "-I" + DEV_SRC]
DEV_FILES = ["fmdx_api.hip", "fmdx_batch.hip", "fmdx_noalloc.hip", "k_step.hip", "k_force.hip", "k_ttm.hip"]
HOST_FILES = ["core.c", "potential.c", "matter.c", "state.c", "dync.c", "turi.c", "ext.c"]


def _run(cmd, cwd=ROOT):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + "\n")
        raise RuntimeError("build step failed: " + os.path.basename(cmd[-1]))
    return r.stdout


def build_device():
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)

    def one(f):
        out = os.path.join(OBJ, f.replace(".hip", ".o"))
        _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(DEV_SRC, f), "-o", out])
        return out

    with ThreadPoolExecutor(max_workers=4) as ex:
        objs = list(ex.map(one, DEV_FILES))
    tmp = os.path.join(LIB, "libfmdx.so.tmp")
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC"] + objs + ["-o", tmp])
    os.replace(tmp, os.path.join(LIB, "libfmdx.so"))


def build_host():
    os.makedirs(LIB, exist_ok=True)
    srcs = [os.path.join(HOST_SRC, f) for f in HOST_FILES]
    tmp = os.path.join(LIB, "libfmd.so.tmp")
    _run(["gcc", "-std=gnu11", "-O2", "-fPIC", "-shared", "-Wall", "-Wno-unused-function",
          "-I" + INC, "-I" + HOST_SRC] + srcs +
         ["-L" + LIB, "-lfmdx", "-Wl,-rpath,$ORIGIN", "-lm", "-o", tmp])
    os.replace(tmp, os.path.join(LIB, "libfmd.so"))


def build_oracle():
    """The checker: FP64, no FMA contraction, single thread."""
    os.makedirs(ORACLE_OUT, exist_ok=True)
    tmp = os.path.join(ORACLE_OUT, "liboracle.so.tmp")
    _run(["gcc", "-std=c11", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-Wall",
          os.path.join(ORACLE, "fmd_oracle.c"), "-lm", "-o", tmp])
    os.replace(tmp, os.path.join(ORACLE_OUT, "liboracle.so"))


def build_program(src, out):
    """Compile a C program against include/fmd.h and link it to libfmd.so (gcc X.c -lfmd)."""
    os.makedirs(os.path.dirname(out), exist_ok=True)
    _run(["gcc", "-std=gnu11", "-O2", src, "-I" + INC, "-L" + LIB, "-lfmd",
          "-Wl,-rpath," + LIB, "-lm", "-o", out])
    return out


REF_EXAMPLES = "/root/reference/examples"
EXAMPLES_OUT = os.path.join(ORACLE, "_ref", "examples")


def build_reference_examples():
    """The reference's example programs compiled unchanged against include/fmd.h and
    linked to libfmd.so (gcc X.c -lfmd, examples/01_argon.c:9), when the reference
    tree is present; the binaries go to oracle/_ref/examples/ (not committed) and
    travel with the tree to the GPU host, where tests/test_gpu_examples.py runs them."""
    if not os.path.isdir(REF_EXAMPLES):
        return []
    out = []
    for name in sorted(os.listdir(REF_EXAMPLES)):
        if name.endswith(".c"):
            out.append(build_program(os.path.join(REF_EXAMPLES, name), os.path.join(EXAMPLES_OUT, name[:-2] + ".x")))
    return out


def build_all():
    build_device()
    build_host()
    build_oracle()
    build_reference_examples()


if __name__ == "__main__":
    build_all()
    print("built:", ", ".join(sorted(os.listdir(LIB))))
