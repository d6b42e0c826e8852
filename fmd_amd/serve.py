"""Continuous-batching request loop for MD jobs on one GPU.

Jobs arrive at given times (a Poisson stream in the benchmark).  The loop
keeps every running job's state resident in device memory as a member of
one batch (libfmdx.so's batch engine) and, between step quanta:

* admits arrived jobs in arrival order while the device-memory budget and
  the member limit allow (admission by footprint: an estimate from the atom
  count before the job is built, its measured footprint afterwards);
* advances all members together by a quantum of steps that never passes
  the end of any member's step count;
* retires members that reached their step count, and members whose device
  error word is set (reported as failed); the others are not disturbed.

A member's trajectory does not depend on the mix (the batch engine runs each
system's own kernels and reductions), so a job gives the same result alone
and among others.  Latency is finish time minus arrival time.

``run_plain`` is the reference-style alternative: the same jobs one at a
time in FP64 through the single-system time loop (fmd_dync_integrate),
first come first served.
"""
import collections
import ctypes as C
import time

import numpy as np

from . import Batch, FMDX_PREC_FP64, FMDX_PREC_MIXED, load_fmdx, _ck

# device bytes per atom assumed before a job is built (slots for about 3x
# the mean cell population plus mover staging)
EST_BYTES_PER_ATOM = 3 * 150 + 60


def memory_of(dev):
    b = C.c_size_t(0)
    _ck(load_fmdx().fmdx_memory(C.c_void_p(dev), C.byref(b)), "fmdx_memory")
    return b.value


class Result:
    __slots__ = ("index", "arrival", "admit", "finish", "latency", "natoms", "nsteps", "natoms_end",
                 "temperature", "pe", "err", "failed", "x")

    def __init__(self, index, arrival):
        self.index = index
        self.arrival = arrival
        self.failed = False
        self.x = None


class Server:
    def __init__(self, build, precision=FMDX_PREC_MIXED, budget_bytes=16 << 30, max_jobs=512, quantum=25,
                 keep_positions=False, check_interval=8):
        self.build = build
        self.precision = precision
        self.budget = budget_bytes
        self.max_jobs = max_jobs
        self.quantum = quantum
        self.keep_positions = keep_positions
        self.check_interval = check_interval

    def run(self, specs, clock=time.perf_counter, sleep=time.sleep):
        """specs: dicts with 'index', 'arrival' (s from the start) and 'nsteps'; returns Results in index order"""
        b = Batch(self.precision)
        b.set_check_interval(self.check_interval)
        pending = collections.deque(sorted(specs, key=lambda s: s["arrival"]))
        active = []      # [job, remaining steps, Result, bytes]
        results = {}
        used = 0
        t0 = clock()
        while pending or active:
            now = clock() - t0
            while pending and pending[0]["arrival"] <= now and len(active) < self.max_jobs:
                sp = pending[0]
                est = self._estimate(sp)
                if active and used + est > self.budget:
                    break
                pending.popleft()
                r = Result(sp["index"], sp["arrival"])
                job = self.build(sp, self.precision)
                b.add(job.dev, job.dt)
                r.admit = clock() - t0
                r.natoms, r.nsteps = job.natoms, job.nsteps
                nbytes = memory_of(job.dev)
                used += nbytes
                active.append([job, job.nsteps, r, nbytes])
                now = clock() - t0
            if not active:
                if pending:
                    sleep(max(0.0, pending[0]["arrival"] - (clock() - t0)))
                continue
            q = min(self.quantum, min(a[1] for a in active))
            if q > 0:
                b.step(q)
                for a in active:
                    a[1] -= q
            st = b.status()
            done = [i for i, a in enumerate(active) if a[1] == 0 or st[i].err]
            t_fin = clock() - t0
            for i in reversed(done):
                job, rem, r, nbytes = active[i]
                t = st[i]
                r.finish = t_fin
                r.latency = t_fin - r.arrival
                r.natoms_end, r.temperature, r.pe, r.err = t.natoms, t.temperature, t.pe, t.err
                r.failed = bool(t.err)
                if self.keep_positions:
                    r.x = job.f.atoms()
                b.remove(job.dev)
                job.free()
                used -= nbytes
                results[r.index] = r
                del active[i]
        b.close()
        return [results[k] for k in sorted(results)]

    def _estimate(self, sp):
        n = sp.get("natoms_hint")
        if n is None:
            n = int(0.0847 * 4.18879 * (sp.get("R1", 20.0) ** 3 + sp.get("R2", 20.0) ** 3))
        return n * EST_BYTES_PER_ATOM + (1 << 20)


def plain_service(specs, build, dt=1e-3, keep_positions=False):
    """Each job alone, FP64, through the single-system time loop: {index: (run time s, Result)}."""
    out = {}
    for sp in sorted(specs, key=lambda s: s["index"]):
        r = Result(sp["index"], sp["arrival"])
        job = build(sp, FMDX_PREC_FP64)
        f = job.f
        t0 = time.perf_counter()
        # step count exactly nsteps: the reference loop runs while time < duration (integrators.c:191-203)
        f.fmd_dync_integrate(-1, (job.nsteps - 0.5) * dt, dt)
        svc = time.perf_counter() - t0
        r.natoms, r.nsteps = job.natoms, job.nsteps
        r.temperature = f.fmd_matt_getTemperature()
        r.pe = f.fmd_ext_getPotentialEnergy()
        r.err = 0
        if keep_positions:
            r.x = f.atoms()
        r.natoms_end = len(f.atoms()["x"])
        job.free()
        out[sp["index"]] = (svc, r)
    return out


def run_plain(specs, build, dt=1e-3, keep_positions=False, service=None):
    """The same jobs one at a time, FP64, through the single-system time loop,
    first come first served; a job's latency is its queueing delay plus its
    own run time, both taken from the measured run times.  `service` (from
    plain_service) reuses run times measured for the same jobs, e.g. under
    another arrival stream: a one-at-a-time server's schedule depends only on
    the arrivals and the run times."""
    if service is None:
        service = plain_service(specs, build, dt, keep_positions)
    out = []
    t_free = 0.0
    for sp in sorted(specs, key=lambda s: s["arrival"]):
        svc, r0 = service[sp["index"]]
        r = Result(sp["index"], sp["arrival"])
        for k in ("natoms", "nsteps", "temperature", "pe", "err", "x", "natoms_end"):
            setattr(r, k, getattr(r0, k))
        start = max(sp["arrival"], t_free)
        t_free = start + svc
        r.admit, r.finish = start, t_free
        r.latency = t_free - sp["arrival"]
        out.append(r)
    return sorted(out, key=lambda r: r.index)


def summarize(results):
    lat = np.array([r.latency for r in results if not r.failed])
    makespan = max(r.finish for r in results) - min(r.arrival for r in results)
    ok = [r for r in results if not r.failed]
    return dict(jobs=len(results), failed=len(results) - len(ok),
                jobs_per_s=len(ok) / makespan if makespan > 0 else float("nan"),
                p50_s=float(np.percentile(lat, 50)) if len(lat) else float("nan"),
                p99_s=float(np.percentile(lat, 99)) if len(lat) else float("nan"),
                makespan_s=makespan,
                atom_steps=int(sum(r.natoms * r.nsteps for r in ok)))
