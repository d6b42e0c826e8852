"""Build-time tuner of the pair kernels' compile-time shape (fmdx_tuned.h).

The waves per home cell (FMDX_WAVES_PER_CELL), the LDS candidate tile per
precision and pair category (FMDX_TILE64 / FMDX_TILE32 / FMDX_TILE32_EAM /
FMDX_TILE32_LJ / FMDX_TILE32_MORSE) and the pairs per iteration of the
single-element EAM passes (FMDX_EAM_NP{64,32}_{0,1}) are constants of
k_force.hip.  They are chosen at build time, not per run: the waves and the tile
decide how a cell's candidates split into staging rounds and stripes, and with
it the order in which each atom's pair terms are summed, so one build gives the
same bits for a system alone, in any batch and on every run.  The cell size
(the binning grid) is fixed by design, not tuned.

Procedure: coordinate descent over PARAMS in order.  For each candidate value,
with the other parameters at their current choice, k_force.hip is compiled
with -D overrides and linked into a complete libfmdx.so / libfmd.so pair in a
scratch directory (the candidates of one parameter compile in parallel).  A
child process loads that pair and times the workload the parameter belongs
to, best of `repeats` runs: the 4M-atom EAM Cu step in FP64 or mixed precision
(HIP events around `steps` velocity-Verlet steps after `warmup`), or, for the
LJ and Morse tiles, the 512-system mixed-precision ensemble step of the batch
engine (wall time of `steps` captured steps).  A candidate whose potential
energy differs from the incumbent's by more than `pe_tol` (relative) is
rejected as wrong, not slow.  The fastest candidate wins; one within `tol` of
the incumbent does not replace it.  With --write the choice is
written to fmdx_tuned.h with the measured table as a comment, and
`python -m fmd_amd.build` compiles it in.

    python -m fmd_amd.autotune [--cells 100] [--steps 20] [--params TILE64,NP32_1] [--write]

Needs a GPU and hipcc; a full pass over PARAMS takes a few minutes.
"""
import argparse
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

from . import build as B

FP64, MIXED, ENSEMBLE = 0, 1, 2
WORKLOAD_NAME = {FP64: "fp64 EAM step", MIXED: "mixed EAM step", ENSEMBLE: "mixed ensemble step"}
TUNED_H = os.path.join(B.DEV_SRC, "fmdx_tuned.h")
TUNED_FILE = "k_force.hip"   # the only translation unit the parameters reach

# (macro, candidates, workload whose step time decides)
PARAMS = [
    ("FMDX_WAVES_PER_CELL", (1, 2), FP64),
    ("FMDX_TILE64", (192, 224, 256), FP64),
    ("FMDX_EAM_NP64_0", (2, 3), FP64),
    ("FMDX_EAM_NP64_1", (2, 3), FP64),
    ("FMDX_TILE32", (320, 384, 448), MIXED),
    ("FMDX_TILE32_EAM", (280, 312, 344), MIXED),
    ("FMDX_EAM_NP32_0", (2, 3, 4), MIXED),
    ("FMDX_EAM_NP32_1", (2, 4, 6), MIXED),
    ("FMDX_TILE32_LJ", (320, 384, 448), ENSEMBLE),
    ("FMDX_TILE32_MORSE", (320, 384, 448), ENSEMBLE),
]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def read_tuned(path=TUNED_H):
    """The values fmdx_tuned.h defines."""
    vals = {}
    with open(path) as fp:
        for m in re.finditer(r"^#define (FMDX_\w+) (\d+)\s*$", fp.read(), re.M):
            vals[m.group(1)] = int(m.group(2))
    missing = [p[0] for p in PARAMS if p[0] not in vals]
    if missing:
        raise RuntimeError("fmdx_tuned.h does not define %s" % ", ".join(missing))
    return vals


def write_tuned(vals, table, path=TUNED_H):
    """Rewrite the defines of fmdx_tuned.h, keeping its header comment, and append the table."""
    with open(path) as fp:
        text = fp.read()
    head = text[:text.index("#ifndef FMDX_TUNED_H")]
    head = re.sub(r"\n \* Last tuning run:.*?(?=\n \*/)", "", head, flags=re.S)
    rows = "".join("\n *   %-16s %5d  %s" % (m, v, "  ".join("%d: %.3f ms" % (c, t) for c, t in ts))
                   for m, v, ts in table)
    head = head.replace("\n */", "\n * Last tuning run: macro, choice, then step time per candidate:" + rows + "\n */", 1)
    body = "#ifndef FMDX_TUNED_H\n#define FMDX_TUNED_H\n"
    # the tuned parameters, then the header's other switches as they were
    for m in [p[0] for p in PARAMS] + [m for m in vals if m not in {p[0] for p in PARAMS}]:
        body += "#ifndef %s\n#define %s %d\n#endif\n" % (m, m, vals[m])
    body += "#endif\n"
    with open(path, "w") as fp:
        fp.write(head + body)


class Builder:
    """Libraries for one set of parameter values, built under `scratch`."""

    def __init__(self, scratch):
        self.scratch = scratch
        self.base = os.path.join(scratch, "base")
        os.makedirs(self.base, exist_ok=True)
        others = [f for f in B.DEV_FILES if f != TUNED_FILE]

        def one(f):
            out = os.path.join(self.base, f.replace(".hip", ".o"))
            B._run([B.HIPCC] + B.HIP_FLAGS + ["-c", os.path.join(B.DEV_SRC, f), "-o", out])
            return out

        with ThreadPoolExecutor(max_workers=4) as ex:
            self.objs = list(ex.map(one, others))

    @staticmethod
    def key(vals):
        return "_".join("%s%d" % (m.replace("FMDX_", "").lower(), vals[m]) for m, _, _ in PARAMS)

    def libs(self, vals):
        d = os.path.join(self.scratch, self.key(vals))
        if os.path.exists(os.path.join(d, "libfmd.so")):
            return d
        os.makedirs(d, exist_ok=True)
        defs = ["-D%s=%d" % (m, vals[m]) for m, _, _ in PARAMS]
        obj = os.path.join(d, "k_force.o")
        B._run([B.HIPCC] + B.HIP_FLAGS + defs + ["-c", os.path.join(B.DEV_SRC, TUNED_FILE), "-o", obj])
        B._run([B.HIPCC, "--offload-arch=" + B.ARCH, "-shared", "-fPIC"] + self.objs + [obj] +
               ["-o", os.path.join(d, "libfmdx.so")])
        srcs = [os.path.join(B.HOST_SRC, f) for f in B.HOST_FILES]
        B._run(["gcc", "-std=gnu11", "-O2", "-fPIC", "-shared", "-I" + B.INC, "-I" + B.HOST_SRC] + srcs +
               ["-L" + d, "-lfmdx", "-Wl,-rpath,$ORIGIN", "-lm", "-o", os.path.join(d, "libfmd.so")])
        os.remove(obj)
        return d


def time_libs(libdir, workload, setfl, cells, steps, warmup, repeats, timeout):
    """Best-of-`repeats` ms per step of `workload` with the libraries in `libdir`."""
    cmd = [sys.executable, "-m", "fmd_amd.autotune", "--child", libdir, "--workload", str(workload),
           "--setfl", setfl, "--cells", str(cells), "--steps", str(steps), "--warmup", str(warmup),
           "--repeats", str(repeats)]
    r = subprocess.run(cmd, cwd=B.ROOT, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError("timing child failed for %s:\n%s" % (libdir, r.stderr[-2000:]))
    return json.loads(r.stdout.strip().splitlines()[-1])


def child_ensemble(args):
    """The 512-system LJ/Morse ensemble, mixed precision, batched and graph-captured."""
    import time
    import fmd_amd as F
    from . import ensemble as E
    jobs = [E.ensemble_job(sp, F.FMDX_PREC_MIXED) for sp in E.ensemble_specs(512, seed=3)]
    best = None
    for _ in range(args.repeats):
        b = F.Batch(F.FMDX_PREC_MIXED)
        for j in jobs:
            b.add(j.dev, j.dt)
        b.step(args.warmup)
        t0 = time.perf_counter()
        b.step(args.steps)
        ms = (time.perf_counter() - t0) / args.steps * 1e3
        st = b.status()
        b.close()
        if any(x.err for x in st):
            raise RuntimeError("ensemble member reported a device error")
        if best is None or ms < best["ms"]:
            best = dict(ms=ms, pass1_ms=float("nan"), pass0_ms=float("nan"), pe=sum(x.pe for x in st))
    for j in jobs:
        j.free()
    print(json.dumps(best), flush=True)


def child(args):
    """Runs in its own process so that each candidate's libraries are the ones loaded."""
    import fmd_amd as F
    F.LIB_DIR = args.child
    if args.workload == ENSEMBLE:
        return child_ensemble(args)
    from . import workloads as W
    best = None
    for _ in range(args.repeats):
        r = W.time_eam_cu_steps(args.cells, args.steps, args.warmup, args.workload, args.setfl)
        ms = r["gpu_ms"] / r["steps"]
        if best is None or ms < best["ms"]:
            best = dict(ms=ms, pass1_ms=r["pass1_ms"], pass0_ms=r["pass0_ms"], pe=r["pe"])
    print(json.dumps(best), flush=True)


def agrees(pe, ref, tol):
    """Same potential energy to `tol` relative: tiles and waves only reorder the pair sums."""
    return abs(pe - ref) <= tol * max(abs(ref), 1e-300)


def tune(args):
    from . import workloads as W
    chosen = read_tuned()
    names = {m.replace("FMDX_", "").replace("EAM_", ""): m for m, _, _ in PARAMS}
    wanted = [names[p.strip().upper()] for p in args.params.split(",")] if args.params else [m for m, _, _ in PARAMS]
    scratch = tempfile.mkdtemp(prefix="fmd_autotune_")
    setfl = os.path.join(scratch, "cu.eam.alloy")
    W.write_setfl(setfl, nr=5000, nrho=5000)
    table = []
    try:
        log("autotune: building the fixed objects in %s" % scratch)
        bld = Builder(scratch)
        for macro, cands, work in PARAMS:
            if macro not in wanted:
                continue
            trial = [dict(chosen, **{macro: c}) for c in cands]
            log("autotune: %s, candidates %s, %s" % (macro, list(cands), WORKLOAD_NAME[work]))
            with ThreadPoolExecutor(max_workers=args.jobs) as ex:
                dirs = list(ex.map(bld.libs, trial))
            runs = {}
            for c, d in zip(cands, dirs):
                r = time_libs(d, work, setfl, args.cells, args.steps, args.warmup, args.repeats, args.timeout)
                runs[c] = r
                log("autotune:   %s=%d  %.3f ms/step (density pass %.3f, force pass %.3f, pe %.9g)" %
                    (macro, c, r["ms"], r["pass1_ms"], r["pass0_ms"], r["pe"]))
            ref = runs.get(chosen[macro], runs[cands[0]])["pe"]
            times = [(c, r["ms"]) for c, r in runs.items() if agrees(r["pe"], ref, args.pe_tol)]
            for c in sorted(set(runs) - set(c for c, _ in times)):
                log("autotune:   %s=%d rejected: potential energy %.9g against %.9g" % (macro, c, runs[c]["pe"], ref))
            inc = dict(times).get(chosen[macro])
            best_c, best_t = min(times, key=lambda ct: ct[1])
            if inc is None or best_t < inc * (1.0 - args.tol):
                chosen[macro] = best_c
            log("autotune: %s -> %d" % (macro, chosen[macro]))
            table.append((macro, chosen[macro], times))
    finally:
        shutil.rmtree(scratch, ignore_errors=True)
    result = dict(chosen={m: chosen[m] for m, _, _ in PARAMS},
                  table=[dict(macro=m, choice=v, ms=dict(ts)) for m, v, ts in table],
                  workload="EAM Cu FCC %d^3, synthetic setfl, and the 512-system LJ/Morse ensemble; "
                           "%d steps after %d, best of %d" % (args.cells, args.steps, args.warmup, args.repeats))
    if args.write:
        write_tuned(chosen, table)
        log("autotune: wrote %s; rebuild with python -m fmd_amd.build" % TUNED_H)
    print(json.dumps(result), flush=True)
    return result


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--cells", type=int, default=100, help="FCC unit cells per edge (100 -> 4M atoms)")
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--repeats", type=int, default=2)
    ap.add_argument("--params", default="",
                    help="comma list of WAVES_PER_CELL, TILE64, NP64_0, NP64_1, TILE32, TILE32_EAM, "
                         "NP32_0, NP32_1, TILE32_LJ, TILE32_MORSE")
    ap.add_argument("--tol", type=float, default=0.01, help="relative gain needed to replace the incumbent")
    ap.add_argument("--pe-tol", type=float, default=1e-6,
                    help="relative potential-energy difference beyond which a candidate is rejected")
    ap.add_argument("--jobs", type=int, default=4, help="parallel compilations")
    ap.add_argument("--timeout", type=float, default=300.0, help="seconds per timing child")
    ap.add_argument("--write", action="store_true", help="write the choice to fmdx_tuned.h")
    ap.add_argument("--child", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--workload", type=int, default=FP64, help=argparse.SUPPRESS)
    ap.add_argument("--setfl", default=None, help=argparse.SUPPRESS)
    args = ap.parse_args(argv)
    if args.child:
        child(args)
    else:
        tune(args)


if __name__ == "__main__":
    main()
