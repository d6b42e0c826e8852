"""Continuous-batching request loop for MD jobs on one GPU.

Jobs arrive at given times (a Poisson stream in the benchmark).  The loop
keeps every running job's state resident in device memory as a member of
one batch (libfmdx.so's batch engine) and, between step quanta:

* admits arrived jobs in arrival order while the device-memory budget and
  the member limit allow (admission by footprint: an estimate from the atom
  count before the job is built, its measured footprint afterwards), builds
  them on a builder thread while the batch steps, and adds each built job to
  the batch at the next quantum boundary;
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
from concurrent.futures import ThreadPoolExecutor

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


class _Built:
    """a build that already ran (the loop without a builder thread)"""

    def __init__(self, fn, *a):
        self.job, self.exc = None, None
        try:
            self.job = fn(*a)
        except BaseException as e:   # re-raised by result(), where a future would raise it
            self.exc = e

    def done(self):
        return True

    def result(self):
        if self.exc is not None:
            raise self.exc
        return self.job


def _release(job):
    try:
        job.free()
    except Exception:
        pass


class Server:
    """`builders` threads build arrived jobs (atoms, device system, upload) while the GPU runs
    the batch's steps: ctypes releases the GIL inside the library's calls, and a quantum of
    steps waits on the device most of its time.  A job joins the batch at the first quantum
    boundary after its build has finished, in arrival order.  The library works on the
    calling thread's current device; with one process per GPU that is device 0 in every
    thread.  builders=0 builds at the quantum boundary on the loop's own thread."""

    def __init__(self, build, precision=FMDX_PREC_MIXED, budget_bytes=16 << 30, max_jobs=512, quantum=25,
                 keep_positions=False, check_interval=8, builders=1):
        self.build = build
        self.precision = precision
        self.budget = budget_bytes
        self.max_jobs = max_jobs
        self.quantum = quantum
        self.keep_positions = keep_positions
        self.check_interval = check_interval
        self.builders = builders

    def run(self, specs, clock=time.perf_counter, sleep=time.sleep):
        """specs: dicts with 'index', 'arrival' (s from the start) and 'nsteps'; returns Results in index order"""
        b = Batch(self.precision)
        b.set_check_interval(self.check_interval)
        pool = ThreadPoolExecutor(max_workers=self.builders) if self.builders > 0 else None
        pending = collections.deque(sorted(specs, key=lambda s: s["arrival"]))
        building = collections.deque()   # [spec, future, estimated bytes]
        active = []      # [job, remaining steps, Result, bytes]
        results = {}
        used = 0         # measured bytes of the members plus the estimates of the jobs in build
        t0 = clock()
        try:
            while pending or building or active:
                now = clock() - t0
                while pending and pending[0]["arrival"] <= now and len(active) + len(building) < self.max_jobs:
                    sp = pending[0]
                    est = self._estimate(sp)
                    if (active or building) and used + est > self.budget:
                        break
                    pending.popleft()
                    fut = pool.submit(self.build, sp, self.precision) if pool else _Built(self.build, sp, self.precision)
                    building.append([sp, fut, est])
                    used += est
                # builds join in arrival order; with nothing to step, wait for the next one
                while building and (building[0][1].done() or not active):
                    sp, fut, est = building.popleft()
                    job = fut.result()
                    r = Result(sp["index"], sp["arrival"])
                    b.add(job.dev, job.dt)
                    r.admit = clock() - t0
                    r.natoms, r.nsteps = job.natoms, job.nsteps
                    nbytes = memory_of(job.dev)
                    used += nbytes - est
                    active.append([job, job.nsteps, r, nbytes])
                if not active:
                    if pending and not building:
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
        finally:
            if pool:
                pool.shutdown(wait=True)
            b.close()   # the members leave the batch
            # after an exception: release the jobs still running or built and not yet admitted
            for a in active:
                _release(a[0])
            for sp, fut, est in building:
                try:
                    job = fut.result()
                except BaseException:
                    continue
                _release(job)
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
