#!/usr/bin/env python3
"""Step benchmark: EAM-alloy Cu, 4,000,000 atoms, one MI355X.

Workload (the 4M-atom EAM Cu configuration): FCC 100x100x100 unit cells,
a = 3.615 A, periodic, Gaussian position jitter 0.02 A, velocities at 600 K,
synthetic setfl EAM (Nr = Nrho = 5000, cutoff 5.5 A), Berendsen thermostat
tau = 0.1 ps, dt = 2 fs.  A "step" is one reference time-loop iteration:
velocity-Verlet start (Berendsen rescale, kick, drift, exact-order re-binning),
the two EAM force passes, and the finishing kick with the group sums.

Set-up goes through the public C API of libfmd.so (box, kinds, setfl loader,
binary atom injection, upload); the timed steps call the same C-ABI seams
(libfmdx.so) that fmd_dync_equilibrate drives.  Inputs are resident in HBM
before timing starts.  The CPU baseline is the oracle (the FP64 restatement
of the reference algorithm, single thread) timed on a bounded sample of the
same workload.
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import fmd_amd as F  # noqa: E402
from fmd_amd import workloads as W  # noqa: E402

METRIC = "atom-steps/s EAM Cu 4M atoms (%HBM roofline); batched jobs/s + p99 vs plain"
HBM_PEAK_GBS = 8000.0
# algorithmic HBM bytes per atom of the EAM force kernel (pass 0): read x (24) + F' (8), write F (24);
# mixed: read the packed FP32 record x y z F' (16), write F (24)
PASS0_BYTES = {F.FMDX_PREC_FP64: 56.0, F.FMDX_PREC_MIXED: 16.0 + 24.0}
# the force pass as the timed steps run it: every step but the last skips the pair energy
# (k_cell<R, pass 0, single-element EAM, energy>)
PASS0_KERNEL = {F.FMDX_PREC_FP64: "k_cell<double,0,1,false>", F.FMDX_PREC_MIXED: "k_cell<float,0,1,false>"}
# HBM bytes per launch of the dominant kernels, measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "eam_cu_4m_traffic.json")
# algorithmic HBM bytes per atom-step of the whole EAM step (integrate 120 + bin 8 + rho 32 + force 56);
# mixed: integrate 120 + bin 8 + FP32 record (read x 24, write 16) + rho (read 16, write F' 8 + 4) + force 40
STEP_BYTES = {F.FMDX_PREC_FP64: 216.0, F.FMDX_PREC_MIXED: 120.0 + 8.0 + 40.0 + 28.0 + 40.0}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


fmd_lib = W.fmd_lib
make_system = W.eam_cu_system
run_steps = W.vv_berendsen_steps
gpu_bench = W.time_eam_cu_steps


def host_threads():
    """Host threads this process may use: its CPU affinity, capped by OMP_NUM_THREADS when set."""
    n = len(os.sched_getaffinity(0))
    try:
        n = min(n, int(os.environ.get("OMP_NUM_THREADS", n)))
    except ValueError:
        pass
    return max(1, n)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fp:
            for line in fp:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(setfl, cells=32, steps=20, seed=1, threads=1):
    """The oracle (FP64 restatement of the reference) on a bounded sample, on `threads` host
    threads at once: each thread builds and steps its own copy of the sample, and ctypes drops
    the GIL inside every oracle call, so the copies run in parallel.  The value is the copies'
    atom-steps over the wall time from a common start until the last copy ends: the host's
    throughput for this algorithm with every thread busy, which bounds what a domain-decomposed
    run of the reference could reach on the same threads."""
    import threading
    from oracle.oracle import Oracle
    a = W.CU["a"]
    rng = np.random.default_rng(seed)
    x = W.fcc_positions(cells, cells, cells, a) + rng.normal(0.0, 0.02, size=(4 * cells ** 3, 3))
    L = cells * a
    x = np.mod(x, L)
    v = W.maxwell_velocities(len(x), 63.546, 600.0, rng)
    start = threading.Barrier(threads + 1)
    ends, errs = [0.0] * threads, []

    def worker(k):
        try:
            o = Oracle([L] * 3, [1, 1, 1], [63.546], ["Cu"])
            o.load_eam(setfl)
            o.apply_eam(0, 0)
            o.make_grid()
            o.add_atoms(x, v)
            o.distribute()
            o.forces()
            o.thermo()
        except Exception as e:   # the barrier still has to be met
            errs.append(e)
            o = None
        start.wait()
        if o is not None:
            for _ in range(steps):
                o.vv_start(2e-3, thermostat=1, tau=0.1, T0=600.0)
                o.forces()
                o.vv_finish(2e-3)
            o.close()
        ends[k] = time.perf_counter()

    ts = [threading.Thread(target=worker, args=(k,)) for k in range(threads)]
    for t in ts:
        t.start()
    start.wait()
    t0 = time.perf_counter()
    for t in ts:
        t.join()
    if errs:
        raise RuntimeError("cpu baseline: %r" % errs[0])
    dt = max(ends) - t0
    return dict(value=threads * len(x) * steps / dt, natoms=len(x), steps=steps, seconds=dt, threads=threads)


def mixed_force_error(f, restore):
    """Max and RMS of |F_mixed - F_fp64| over all force components, each over max|F_fp64|, and the
    fraction of atoms whose largest component error exceeds 1e-5 of max|F_fp64|, on the current
    state of the distributed system `f` (fmd_amd.api.Fmd); leaves it at precision `restore`."""
    fx = F.load_fmdx()
    dev = C.c_void_p(f.fmd_ext_getDevice())
    f.fmd_ext_setPrecision(F.FMDX_PREC_FP64)
    F._ck(fx.fmdx_update_forces(dev, -1, 0), "forces")
    f64 = f.atoms()["f"]
    f.fmd_ext_setPrecision(F.FMDX_PREC_MIXED)
    F._ck(fx.fmdx_update_forces(dev, -1, 0), "forces")
    d = f.atoms()["f"] - f64
    f.fmd_ext_setPrecision(restore)
    F._ck(fx.fmdx_update_forces(dev, -1, 0), "forces")
    scale = max(float(np.max(np.abs(f64))), 1e-300)
    above = float(np.mean(np.max(np.abs(d), axis=1) > 1e-5 * scale))
    return float(np.max(np.abs(d))) / scale, float(np.sqrt(np.mean(d * d))) / scale, above


def device_atoms(dev):
    """cell-ordered atoms of a device system (x, f, F'(rho) as aux, atom ids) through the C-ABI"""
    fx = F.load_fmdx()
    n = C.c_uint(0)
    F._ck(fx.fmdx_download(dev, C.byref(n), None, None, None, None, None, None, None, None), "download")
    n = n.value
    x, f, aux, aid = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n), np.zeros(n, np.uint32)
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    F._ck(fx.fmdx_download(dev, C.byref(C.c_uint(0)), None, p(x), None, p(f), p(aux), None, p(aid), None), "download")
    return dict(x=x, f=f, aux=aux, atomid=aid)


def eam_parity(cells, setfl, seed=1):
    """Parity of the full-size EAM Cu state (the recipe of the timed run, before any step): the
    oracle (FP64 restatement of eam.c:29-146,225-322, one host thread) against the device's FP64
    and mixed force passes -- per-atom forces over max|F|, F'(rho) (the reference's vaream), and the
    potential energy -- and mixed against FP64 on the device (max, RMS, atoms over 1e-5)."""
    from fmd_amd import api
    from oracle.oracle import Oracle
    a = W.CU["a"]
    rng = np.random.default_rng(seed)
    x = W.fcc_positions(cells, cells, cells, a)
    x = np.mod(x + rng.normal(0.0, 0.02, size=x.shape), cells * a)
    v = W.maxwell_velocities(len(x), 63.546, 600.0, rng)
    L = cells * a
    o = Oracle([L] * 3, [1, 1, 1], [63.546], ["Cu"])
    o.load_eam(setfl)
    o.apply_eam(0, 0)
    o.make_grid()
    o.add_atoms(x, v)
    o.distribute()
    t0 = time.perf_counter()
    o.forces()
    t_oracle = time.perf_counter() - t0
    ref, eref = o.atoms(), o.energies()
    o.close()
    order_ref = np.argsort(ref["atomid"])
    fmax = float(np.max(np.abs(ref["f"])))
    f = api.Fmd()
    out = dict(natoms=len(x), oracle_seconds=t_oracle, oracle_threads=1)
    try:
        f.fmd_box_setSize(L, L, L)
        f.fmd_box_setPBC(True, True, True)
        f.set_kinds(["Cu"], [63.546])
        f.fmd_pot_apply(0, 0, f.fmd_pot_eam_alloy_load(setfl.encode()))
        f.load_atoms(x, v)
        if not f.fmd_ext_distribute():
            raise RuntimeError("distribute failed")
        fx = F.load_fmdx()
        dev = C.c_void_p(f.fmd_ext_getDevice())
        got = {}
        for name, prec in (("fp64", F.FMDX_PREC_FP64), ("mixed", F.FMDX_PREC_MIXED)):
            f.fmd_ext_setPrecision(prec)
            F._ck(fx.fmdx_update_forces(dev, -1, 0), "forces")
            th = F.Thermo()
            F._ck(fx.fmdx_read_thermo(dev, C.byref(th)), "read_thermo")
            d = device_atoms(dev)
            od = np.argsort(d["atomid"])
            if not np.array_equal(d["atomid"][od], ref["atomid"][order_ref]):
                raise RuntimeError("atom sets differ")
            df = d["f"][od] - ref["f"][order_ref]
            out[name + "_vs_oracle"] = dict(
                max_rel_force=float(np.max(np.abs(df))) / fmax,
                max_rel_Fprime=float(np.max(np.abs(d["aux"][od] - ref["aux"][order_ref])) /
                                     np.max(np.abs(ref["aux"]))),
                rel_pe=abs(th.pe - eref["pe"]) / abs(eref["pe"]),
                same_cell_order=bool(np.array_equal(d["atomid"], ref["atomid"])),
                device_err_word=int(th.err))
            got[name] = d["f"][od]
        dm = got["mixed"] - got["fp64"]
        scale = max(float(np.max(np.abs(got["fp64"]))), 1e-300)
        out["mixed_vs_fp64"] = dict(max_rel=float(np.max(np.abs(dm))) / scale,
                                    rms_rel=float(np.sqrt(np.mean(dm * dm))) / scale,
                                    frac_atoms_over_1em5=float(np.mean(np.max(np.abs(dm), axis=1) > 1e-5 * scale)))
    finally:
        f.free()
    return out


def batched_ensemble(nsys, steps, plain_sample, seed=3):
    """The batched-ensemble workload: nsys periodic systems (LJ Ar or Morse Cu, log-uniform
    500-50000 atoms, 50-800 K), all advanced together in mixed precision by the batch
    engine; against the plain FP64 one-system-at-a-time time loop on a stratified sample."""
    from fmd_amd import ensemble as E
    specs = E.ensemble_specs(nsys, seed=seed)
    jobs = [E.ensemble_job(sp, F.FMDX_PREC_MIXED) for sp in specs]
    natoms = sum(j.natoms for j in jobs)
    b = F.Batch(F.FMDX_PREC_MIXED)
    for j in jobs:
        b.add(j.dev, j.dt)
    b.step(4)
    t0 = time.perf_counter()
    b.step(steps)
    t = time.perf_counter() - t0
    st = b.status()
    errs = sum(1 for x in st if x.err)
    b.close()
    for j in jobs:
        j.free()
    # plain FP64, one system at a time, and the mixed-vs-FP64 force error of the same systems
    sample = specs[:: max(1, nsys // plain_sample)]
    per_step, ferr, frms = [], [], []
    for sp in sample:
        j = E.ensemble_job(sp, F.FMDX_PREC_FP64)
        f = j.f
        f.fmd_dync_integrate(-1, (3 - 0.5) * j.dt, j.dt)
        t1 = time.perf_counter()
        f.fmd_dync_integrate(-1, (steps - 0.5) * j.dt, j.dt)
        per_step.append((time.perf_counter() - t1) / steps)
        emax, erms, _ = mixed_force_error(f, F.FMDX_PREC_FP64)
        ferr.append(emax)
        frms.append(erms)
        j.free()
    t_step = t / steps
    jobs_s = nsys / (2000 * t_step)
    plain_job_s = 1.0 / (2000 * float(np.mean(per_step)))
    return dict(workload="%d periodic systems, LJ Ar or Morse Cu (p=1/2), FCC, log-uniform 500-50000 atoms "
                         "(%d atoms in all), 50-800 K, dt=1 fs; job = 2000 steps of one system" % (nsys, natoms),
                precision="mixed", timed_steps=steps, ms_per_step=t_step * 1e3,
                atom_steps_per_s=natoms * steps / t, jobs_per_s=jobs_s,
                plain_fp64_jobs_per_s=plain_job_s, speedup_vs_plain=jobs_s / plain_job_s,
                plain_sample="%d of %d systems, %d timed steps each through fmd_dync_integrate"
                             % (len(sample), nsys, steps),
                max_rel_force_err_mixed_vs_fp64=max(ferr), rms_rel_force_err_mixed_vs_fp64=max(frms),
                members_with_errors=errs)


def collision_batch(njobs, steps, seed=4):
    """The served leg's kernels without its request loop: njobs collision jobs (mixed) stepped
    together as one batch, for profiling (the loop's admissions and retirements re-capture the
    step; this batch replays one capture)."""
    from fmd_amd import ensemble as E
    specs = E.collision_specs(njobs, seed=seed, rate=1.0)
    jobs = [E.collision_job(sp, F.FMDX_PREC_MIXED) for sp in specs]
    b = F.Batch(F.FMDX_PREC_MIXED)
    try:
        for j in jobs:
            b.add(j.dev, j.dt)
        b.step(8)
        t0 = time.perf_counter()
        b.step(steps)
        ms = (time.perf_counter() - t0) / steps * 1e3
        st = b.status()
        natoms = sum(j.natoms for j in jobs)
        return dict(workload="%d collision jobs (mixed) as one batch" % njobs, natoms=natoms, steps=steps,
                    ms_per_step=ms, atom_steps_per_s=natoms / ms * 1e3, members_with_errors=sum(1 for t in st if t.err))
    finally:
        b.close()
        for j in jobs:
            j.free()


def served_collisions(njobs, rates, seed=4):
    """Served two-cluster Morse Cu collision jobs through the continuous-batching loop in mixed
    precision, at each Poisson arrival rate, against the plain FP64 one-job-at-a-time loop (first
    come first served) on the same arrival stream.  The same njobs jobs at every rate; the plain
    run time of each job is measured once and its FCFS schedule computed for each stream."""
    from fmd_amd import ensemble as E
    from fmd_amd import serve as S
    base = E.collision_specs(njobs, seed=seed, rate=1.0)
    t0 = time.perf_counter()
    service = S.plain_service(base, E.collision_job)
    plain_s = time.perf_counter() - t0
    per_rate = []
    for rate in rates:
        specs = E.collision_specs(njobs, seed=seed, rate=rate)
        served = S.Server(E.collision_job, precision=F.FMDX_PREC_MIXED).run(specs)
        plain = S.run_plain(specs, E.collision_job, service=service)
        a, p = S.summarize(served), S.summarize(plain)
        per_rate.append(dict(rate=rate, served_mixed=a, plain_fp64=p,
                             jobs_per_s_ratio=a["jobs_per_s"] / p["jobs_per_s"],
                             p99_ratio_plain_over_served=p["p99_s"] / a["p99_s"]))
        log("bench: served at %g jobs/s: %s" % (rate, per_rate[-1]))
    ferr = []
    for sp in base[:4]:   # mixed against FP64 forces on the first jobs' initial states
        j = E.collision_job(sp, F.FMDX_PREC_FP64)
        try:
            ferr.append(mixed_force_error(j.f, F.FMDX_PREC_FP64))
        finally:
            j.free()
    one = collision_batch(njobs, 48, seed=seed)
    return dict(workload="%d jobs per rate, Poisson arrivals; each two Morse-Cu FCC spheres R=15-50 A, 20 A gap, "
                         "10-50 A/ps, impact parameter 0-(R1+R2), vacuum box, 1000-4000 steps of 1 fs" % njobs,
                rates=per_rate, plain_run_seconds=plain_s,
                all_jobs_as_one_batch=dict(ms_per_step=one["ms_per_step"], atom_steps_per_s=one["atom_steps_per_s"],
                                           natoms=one["natoms"], steps=one["steps"]),
                max_rel_force_err_mixed_vs_fp64=max(e[0] for e in ferr),
                rms_rel_force_err_mixed_vs_fp64=max(e[1] for e in ferr),
                force_err_sample="initial states of the first %d jobs" % len(ferr))


def slab_system(setfl, precision, seed, pad):
    """The 40 x 40 x 400 Cu slab of the TTM workload with the EAM table `setfl`, distributed and
    brought from 600 K to 300 K by 0.4 ps of Berendsen dynamics; returns (fmd_amd.api.Fmd, box)."""
    from fmd_amd import api
    a = W.CU["a"]
    n = (40, 40, 400)
    rng = np.random.default_rng(seed)
    x = W.fcc_positions(*n, a, origin=(0.0, 0.0, pad)) + rng.normal(0.0, 0.02, size=(4 * n[0] * n[1] * n[2], 3))
    L = [n[0] * a, n[1] * a, n[2] * a + 2 * pad]
    x[:, 0] = np.mod(x[:, 0], L[0])
    x[:, 1] = np.mod(x[:, 1], L[1])
    v = W.maxwell_velocities(len(x), 63.546, 600.0, rng)
    f = api.Fmd()
    f.fmd_box_setSize(*L)
    f.fmd_box_setPBC(True, True, False)
    f.set_kinds(["Cu"], [63.546])
    pot = f.fmd_pot_eam_alloy_load(setfl.encode())
    f.fmd_pot_apply(0, 0, pot)
    f.load_atoms(x, v)
    f.fmd_ext_setPrecision(precision)
    if not f.fmd_ext_distribute():
        f.free()
        raise RuntimeError("distribute failed")
    f.fmd_dync_equilibrate(-1, 0.4, 2e-3, 0.05, 300.0)
    return f, L


def ttm_slab(steps, precision, setfl, seed=5, pad=320.0, pulse_to_ps=3.0):
    """The TTM-MD workload: EAM Cu slab of 40 x 40 x 400 FCC cells (2,560,000 atoms), periodic in
    x and y, free surfaces in z with `pad` A of vacuum on each side; a 4 x 4 x 200 TTM turi with
    Ce = gamma Te (96.6 J/m^3/K^2), kappa = kappa0 Te / Tl (401 W/m/K), G = 1e17 W/m^3/K and the
    Gaussian source (absorbed fluence 100 J/m^2, 15 nm depth, t0 = 2 ps, FWHM 0.8326 ps, i.e.
    exp(-((t - 2 ps) / 0.5 ps)^2)); 50 TTM sub-steps per 1 fs MD step; Te = Tl = 300 K at start:
    the lattice gets 600 K velocities and is brought to 300 K by 0.4 ps of Berendsen dynamics
    before the turi exists.  The vacuum keeps the explicit TTM step inside its stability limit
    (kappa0 / (gamma Tl) grows as the lattice cools).  Runs through the public API
    (fmd_dync_integrate: symplectic Euler with the TTM turi)."""
    from fmd_amd import api
    f, L = slab_system(setfl, precision, seed, pad)
    ferr = mixed_force_error(f, precision)   # on the thermalised slab, before the turi exists
    cu01 = os.path.join(ROOT, "tests", "golden", "data", "Cu01.eam.alloy")
    g, _ = slab_system(cu01, precision, seed, pad)
    try:
        ferr_cu01 = mixed_force_error(g, precision)
    finally:
        g.free()
    T_start = f.fmd_matt_getTemperature()
    dims = (4, 4, 200)
    turi = f.fmd_turi_add(api.TURI_TTM_TYPE1, *dims, 0.0, -1.0)
    gamma, k0, G, ratio, dt = 96.6, 401.0, 1e17, 50, 1e-3
    f._fmd_ttm_setHeatCapacity_linear(turi, api.HeatCapLinear(gamma))
    f.fmd_ext_ttm_setHeatConductivity_proportional(turi, k0)
    f._fmd_ttm_setCouplingFactor_constant1(turi, G)
    f.fmd_ttm_setElectronTemperature(turi, 300.0)
    f.fmd_ttm_setTimestepRatio(turi, ratio)
    f._fmd_ttm_setLaserSource_gaussian(turi, api.LaserGaussian(100.0, 0.0, 2e-12, 0.8326e-12, 15e-9))
    # explicit-FD stability at Tl = 300 K: 2 D dt_sub sum(1/h^2) with D = kappa0 / (gamma Tl)
    h = [L[0] / dims[0] * 1e-10, L[1] / dims[1] * 1e-10, L[2] / dims[2] * 1e-10]
    cfl = 2 * (k0 / (gamma * 300.0)) * (dt * 1e-12 / ratio) * sum(1.0 / (q * q) for q in h)
    f.fmd_dync_integrate(-1, (2 - 0.5) * dt, dt)
    t0 = time.perf_counter()
    f.fmd_dync_integrate(-1, (steps - 0.5) * dt, dt)
    wall = time.perf_counter() - t0
    te = f.field_array(turi, f.fmd_field_find(turi, api.FIELD_TTM_TE))
    ti = f.field_array(turi, f.fmd_field_find(turi, api.FIELD_TEMPERATURE))
    T = f.fmd_matt_getTemperature()
    act = te >= 0
    tl_min = float(np.min(ti[act])) if np.any(act) else float("nan")
    te_max, n_active = float(np.max(te)), int(np.sum(act))
    natoms = 4 * 40 * 40 * 400
    # untimed: carry on through the pulse (peak at 2 ps) and sample the electrons and the lattice
    pulse = dict(t_ps=[], max_Te=[], mean_Te_active=[], lattice_T=[])
    while pulse_to_ps and f.fmd_dync_getTime() < pulse_to_ps - 0.5 * dt:
        chunk = min(0.25, pulse_to_ps - f.fmd_dync_getTime())
        f.fmd_dync_integrate(-1, chunk - 0.5 * dt, dt)
        te = f.field_array(turi, f.fmd_field_find(turi, api.FIELD_TTM_TE))
        a = te >= 0
        pulse["t_ps"].append(round(f.fmd_dync_getTime(), 4))
        pulse["max_Te"].append(float(np.max(te)))
        pulse["mean_Te_active"].append(float(np.mean(te[a])) if np.any(a) else float("nan"))
        pulse["lattice_T"].append(f.fmd_matt_getTemperature())
    f.free()
    return dict(workload="EAM Cu slab 40x40x400 FCC (%d atoms), PBC x,y, free z with %g A vacuum per side; "
                         "TTM 4x4x200, Ce=96.6 Te J/m^3/K, kappa=401 Te/Tl W/m/K, G=1e17 W/m^3/K, source "
                         "100 J/m^2 absorbed, 15 nm, exp(-((t-2 ps)/0.5 ps)^2), 50 sub-steps per 1 fs step"
                         % (natoms, pad),
                precision="mixed" if precision == F.FMDX_PREC_MIXED else "fp64", timed_steps=steps,
                ms_per_step=wall / steps * 1e3, atom_steps_per_s=natoms * steps / wall,
                cfl_number_at_300K=cfl, cfl_number_coldest_cell=cfl * 300.0 / tl_min,
                lattice_T_start=T_start, lattice_T_after=T, max_Te_after=te_max,
                active_ttm_cells=n_active, through_the_pulse=pulse if pulse["t_ps"] else None,
                max_rel_force_err_mixed_vs_fp64=ferr[0], rms_rel_force_err_mixed_vs_fp64=ferr[1],
                frac_atoms_over_1em5_mixed_vs_fp64=ferr[2],
                force_err_cu01_slab=dict(max_rel=ferr_cu01[0], rms_rel=ferr_cu01[1], frac_atoms_over_1em5=ferr_cu01[2]),
                force_err_note="the synthetic setfl is untapered at rc: after thermalisation some pairs sit on "
                               "its last r^2 interval, whose slope is extreme, or within FP32 rounding of rc; "
                               "force_err_cu01_slab is the same thermalised slab with the smooth Cu01 table")


def nve_drift(cells, steps, precision, setfl, seed=1, dt=1e-3, equil_steps=1000, T0=600.0):
    """Total-energy drift of the EAM Cu crystal in NVE (eV/atom/ns), velocity Verlet at dt: the
    jittered 600 K lattice is first thermalised (Berendsen, tau 0.1 ps, `equil_steps` steps) so that
    the equipartition transient is over, then E = PE + KE is sampled every 10 NVE steps and the drift
    is the least-squares slope over the whole window; the spread of E about that line is reported."""
    lib, md, dev, n = make_system(cells, seed, setfl, precision)
    try:
        return drift_on(dev, n, steps, dt, equil_steps, T0)
    finally:
        lib.fmd_free(md)


def pair_drift(pot, cells, steps, precision, T0, seed=5, dt=1e-3, equil_steps=1000):
    """The same measurement on a periodic LJ-argon or Morse-copper FCC crystal of cells^3 unit
    cells, the systems of the batched ensemble (rc 2.5 sigma and 7 A, neither shifted at rc)."""
    from fmd_amd import ensemble as E
    job = E.ensemble_job(dict(pot=pot, cells=cells, T=T0, seed=seed, index=0), precision, dt=dt)
    try:
        return drift_on(job.dev, job.natoms, steps, dt, equil_steps, T0)
    finally:
        job.free()


def drift_on(dev, n, steps, dt, equil_steps, T0):
    """Berendsen to T0 for equil_steps, then NVE for `steps`: least-squares slope of (PE + KE) / n."""
    fx = F.load_fmdx()
    F._ck(fx.fmdx_update_forces(dev, -1, 0), "forces")
    F._ck(fx.fmdx_thermo(dev, -1), "thermo")   # the first Berendsen factor needs the temperature
    # the product loop's periodic check (audit, error word, slot growth) every 32 steps
    for k in range(equil_steps):
        F._ck(fx.fmdx_vv_start(dev, -1, dt, 1, 0.1, T0), "vv_start")
        F._ck(fx.fmdx_update_forces(dev, -1, 0), "forces")
        F._ck(fx.fmdx_vv_finish(dev, -1, dt), "vv_finish")
        if (k + 1) % 32 == 0:
            W.check_device(fx, dev)
    F._ck(fx.fmdx_thermo(dev, -1), "thermo")
    th = W.check_device(fx, dev)
    es = [th.pe + th.ke]
    for k in range(steps):
        F._ck(fx.fmdx_vv_start(dev, -1, dt, 0, 1.0, 0.0), "vv_start")
        F._ck(fx.fmdx_update_forces(dev, -1, 0), "forces")
        F._ck(fx.fmdx_vv_finish(dev, -1, dt), "vv_finish")
        if (k + 1) % 10 == 0:
            th = F.Thermo()
            F._ck(fx.fmdx_read_thermo(dev, C.byref(th)), "read_thermo")
            es.append(th.pe + th.ke)
        if (k + 1) % 32 == 0:
            W.check_device(fx, dev)
    th = W.check_device(fx, dev)
    T_end = th.temperature
    t = np.arange(len(es)) * 10 * dt * 1e-3   # ns
    e = np.array(es) / n
    slope, icpt = np.polyfit(t, e, 1)
    return dict(steps=steps, dt_ps=dt, natoms=n, thermalised_steps=equil_steps, T_end=T_end,
                drift_eV_per_atom_per_ns=float(slope),
                spread_eV_per_atom=float(np.std(e - (slope * t + icpt))),
                max_dev_eV_per_atom=float(np.max(np.abs(e - e[0]))))


def measured_counters(kernel, natoms):
    """rocprofv3 figures of `kernel` on this workload from the committed counter summary:
    HBM bytes per launch and the busy fractions of VALU, TA and TD (None when not profiled)"""
    try:
        with open(TRAFFIC_FILE) as fp:
            t = json.load(fp)
    except (OSError, ValueError):
        return None
    k = t.get("kernels", {}).get(kernel)
    if not k or k.get("natoms") != natoms:
        return None
    return k


def measured_traffic(kernel, natoms):
    k = measured_counters(kernel, natoms)
    return k.get("hbm_bytes") if k else None


def copy_calibration(mib):
    """one plain device-to-device copy of `mib` MiB, three times: the reference reading of the
    FETCH_SIZE / WRITE_SIZE counters (each copy reads and writes exactly mib MiB)"""
    import torch
    n = mib << 20
    a = torch.empty(n, dtype=torch.uint8, device="cuda")
    a.fill_(1)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    log("bench: copied %d bytes x3" % n)


def guarded(name, fn, *a, **k):
    """Run an optional leg; a failure is logged and reported in the JSON line instead of
    ending the run, so that the headline figure is still printed."""
    try:
        return fn(*a, **k)
    except Exception as e:
        log("bench: %s failed: %r" % (name, e))
        return {"error": repr(e)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cells", type=int, default=100, help="FCC unit cells per edge (100 -> 4M atoms)")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-parity", action="store_true", help="skip the full-size oracle comparison")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-ensemble measurement")
    ap.add_argument("--no-serve", action="store_true", help="skip the served-collision measurement")
    ap.add_argument("--batch-systems", type=int, default=512)
    ap.add_argument("--batch-steps", type=int, default=40)
    ap.add_argument("--plain-sample", type=int, default=16)
    ap.add_argument("--serve-jobs", type=int, default=64)
    ap.add_argument("--serve-rates", default="5,50,200", help="Poisson arrival rates (jobs/s), comma-separated")
    ap.add_argument("--no-ttm", action="store_true", help="skip the TTM slab measurement")
    ap.add_argument("--ttm-steps", type=int, default=200)
    ap.add_argument("--no-drift", action="store_true", help="skip the NVE drift measurement")
    ap.add_argument("--drift-steps", type=int, default=10000)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain MIB-MiB device copy (counter calibration) and exit")
    ap.add_argument("--collision-batch", type=int, default=0, metavar="JOBS",
                    help="only step JOBS served-collision jobs together as one batch (the served leg's kernels, "
                         "for profiling) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (use --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if F.device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.collision_batch:
        print(json.dumps(collision_batch(args.collision_batch, args.steps)), flush=True)
        return
    prec = F.FMDX_PREC_FP64 if args.precision == "fp64" else F.FMDX_PREC_MIXED
    with tempfile.TemporaryDirectory() as td:
        setfl = os.path.join(td, "Cu_synth.eam.alloy")
        W.write_setfl(setfl, nr=5000, nrho=5000)
        log("bench: EAM Cu %d^3 FCC, %s, %d warmup + %d timed steps" % (args.cells, args.precision, args.warmup,
                                                                         args.steps))
        r = gpu_bench(args.cells, args.steps, args.warmup, prec, setfl)
        log("bench: %s" % r)
        other = None
        if not args.no_mixed:
            p2 = F.FMDX_PREC_MIXED if prec == F.FMDX_PREC_FP64 else F.FMDX_PREC_FP64
            other = guarded("other precision", gpu_bench, args.cells, args.steps, args.warmup, p2, setfl)
            log("bench: other precision %s" % other)
        ens = srv = ttm = None
        drift = {}
        parity = None
        if not args.no_parity:
            parity = guarded("full-size parity", eam_parity, args.cells, setfl)
            log("bench: full-size parity %s" % parity)
        if not args.no_drift:
            cu01 = os.path.join(ROOT, "tests", "golden", "data", "Cu01.eam.alloy")
            for pname, pv in (("fp64", F.FMDX_PREC_FP64), ("mixed", F.FMDX_PREC_MIXED)):
                drift["cu01_" + pname] = guarded("NVE drift Cu01 " + pname, nve_drift, min(args.cells, 40),
                                                 args.drift_steps, pv, cu01)
                drift["synthetic_" + pname] = guarded("NVE drift synthetic " + pname, nve_drift,
                                                      min(args.cells, 40), args.drift_steps, pv, setfl)
                drift["lj_ar_" + pname] = guarded("NVE drift LJ Ar " + pname, pair_drift, "lj", 20,
                                                  args.drift_steps, pv, 100.0)
                drift["morse_cu_" + pname] = guarded("NVE drift Morse Cu " + pname, pair_drift, "morse", 20,
                                                     args.drift_steps, pv, 300.0)
            for key in ("cu01", "synthetic", "lj_ar", "morse_cu"):
                a, b = drift[key + "_mixed"], drift[key + "_fp64"]
                if "error" not in a and "error" not in b:
                    drift[key + "_mixed_minus_fp64_eV_per_atom_per_ns"] = (
                        a["drift_eV_per_atom_per_ns"] - b["drift_eV_per_atom_per_ns"])
            drift["note"] = ("Cu01 is the smooth Mishin setfl; the synthetic table is truncated at 5.5 A "
                             "without a taper, and LJ (rc 2.5 sigma, 100 K) and Morse (rc 7 A, 300 K) are "
                             "not shifted at rc, so their drifts include the cutoff jumps")
            log("bench: NVE drift %s" % drift)
        if not args.no_ttm:
            ttm = guarded("TTM slab", ttm_slab, args.ttm_steps, F.FMDX_PREC_MIXED, setfl)
            log("bench: TTM slab %s" % ttm)
        if not args.no_batch:
            ens = guarded("batched ensemble", batched_ensemble, args.batch_systems, args.batch_steps,
                          args.plain_sample)
            log("bench: batched ensemble %s" % ens)
        if not args.no_serve:
            srv = guarded("served collisions", served_collisions, args.serve_jobs,
                          [float(r) for r in args.serve_rates.split(",")])
            log("bench: served collisions %s" % srv)
        cpu = None
        if not args.no_cpu_baseline:
            cpu = guarded("cpu baseline", cpu_baseline, setfl, threads=host_threads())
            log("bench: cpu baseline %s" % cpu)
            cpu1 = guarded("cpu baseline, one thread", cpu_baseline, setfl, steps=10)
            log("bench: cpu baseline, one thread %s" % cpu1)
            if "error" not in cpu:
                cpu["one_thread"] = cpu1.get("value")
    n, K = r["natoms"], r["steps"]
    ms_step = r["gpu_ms"] / K
    value = n * K / (r["gpu_ms"] * 1e-3)
    achieved = PASS0_BYTES[prec] * n / (r["pass0_ms"] * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == F.FMDX_PREC_FP64 else "f32-forces/f64-integration",
        "data": "synthetic",
        "config": {"workload": "EAM-alloy Cu FCC %dx%dx%d (%d atoms), a=3.615 A, PBC, jitter 0.02 A, 600 K, "
                               "Berendsen tau=0.1 ps, dt=2 fs, synthetic setfl Nr=Nrho=5000 rc=5.5 A"
                               % (args.cells, args.cells, args.cells, n),
                   "natoms": n, "precision": args.precision},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(PASS0_KERNEL[prec], n),
                     "kernel": PASS0_KERNEL[prec] + " (EAM force pass without the pair energy; kernel_ms averages "
                                                    "the timed steps' force passes, the last of which sums it)",
                     "kernel_ms": r["pass0_ms"], "algorithmic_bytes_per_atom": PASS0_BYTES[prec],
                     "unit_busy": {k: v for k, v in (measured_counters(PASS0_KERNEL[prec], n) or {}).items()
                                   if k.endswith("_busy")} or None},
        "step_roofline_frac": STEP_BYTES[prec] * value / (HBM_PEAK_GBS * 1e9),
        "pass1_ms": r["pass1_ms"],
    }
    if other is not None and "error" in other:
        out["other_precision"] = other
    elif other is not None:
        out["other_precision"] = {"precision": "mixed" if prec == F.FMDX_PREC_FP64 else "fp64",
                                  "value": other["natoms"] * other["steps"] / (other["gpu_ms"] * 1e-3),
                                  "ms_per_step": other["gpu_ms"] / other["steps"],
                                  "pass0_ms": other["pass0_ms"], "pass1_ms": other["pass1_ms"]}
    if parity is not None:
        out["parity"] = parity
        if "mixed_vs_fp64" in parity:
            out["force_err_mixed_vs_fp64"] = dict(natoms=parity["natoms"], **parity["mixed_vs_fp64"])
    if drift:
        out["nve_drift"] = drift
    if ttm is not None:
        out["ttm_slab"] = ttm
    if ens is not None:
        out["batched_ensemble"] = ens
    if srv is not None:
        out["served_collisions"] = srv
    if cpu is not None and "error" in cpu:
        out["cpu_baseline"] = {"value": None, "unit": "atom-steps/s", "cores": host_threads(), "kind": "port",
                               "sample": None, "error": cpu["error"]}
    elif cpu is not None:
        out["cpu_baseline"] = {"value": cpu["value"], "unit": "atom-steps/s", "cores": cpu["threads"], "kind": "port",
                               "sample": "oracle (FP64 reference restatement) on EAM Cu: %d concurrent copies, one "
                                         "per host thread (%s), of a %d-atom sample (32^3 FCC, same setfl and "
                                         "lattice recipe), %d Berendsen VV steps each, %.1f s; one copy alone: "
                                         "%.3g atom-steps/s"
                                         % (cpu["threads"], cpu_model(), cpu["natoms"], cpu["steps"], cpu["seconds"],
                                            cpu["one_thread"] or float("nan")),
                               "one_thread": cpu["one_thread"]}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
