"""The continuous-batching request loop on the GPU.

* Served in FP64 among other jobs, with admissions and retirements between
  quanta, every job ends bit-identical to the same job run alone through the
  reference-style time loop (fmd_dync_integrate).
* A job that breaks (an atom thrown beyond the cell layer) is retired as
  failed; the jobs sharing the batch with it are not disturbed.
"""
import numpy as np
import pytest

import fmd_amd as F
from fmd_amd import ensemble as E
from fmd_amd import serve as S

pytestmark = pytest.mark.gpu


def _specs(n, seed):
    sp = E.collision_specs(n, seed=seed, rate=400.0, steps=(40, 90))
    for s in sp:
        s["R1"] = 10.0 + 4.0 * s["R1"] / 50.0
        s["R2"] = 10.0 + 4.0 * s["R2"] / 50.0
    return sp


@pytest.mark.parametrize("builders", [0, 1])
def test_served_jobs_equal_plain_runs(builders):
    """with the jobs built on the loop's thread at quantum boundaries, or on a builder thread
    while the batch steps"""
    specs = _specs(5, seed=11)
    served = S.Server(E.collision_job, precision=F.FMDX_PREC_FP64, quantum=7, keep_positions=True,
                      builders=builders).run(specs)
    plain = S.run_plain(specs, E.collision_job, keep_positions=True)
    assert len(served) == len(plain) == 5
    for a, b in zip(served, plain):
        assert not a.failed and a.err == 0
        assert a.natoms_end == b.natoms_end
        for k in ("x", "v", "atomid"):
            assert np.array_equal(a.x[k], b.x[k]), (a.index, k)
        assert a.temperature == b.temperature
    s = S.summarize(served)
    assert s["failed"] == 0 and s["jobs_per_s"] > 0 and s["p99_s"] >= s["p50_s"]


def _build_maybe_broken(sp, precision):
    if sp.get("broken"):
        x, v, L = E.collision_atoms(sp)
        v = v.copy()
        v[0] = [1.0e5, 0.0, 0.0]   # 100 A in one femtosecond step
        return E.Job(L, (0, 0, 0), "morse", x, v, precision, nsteps=sp["nsteps"], tag=sp["index"])
    return E.collision_job(sp, precision)


def test_failing_job_is_retired_alone():
    specs = _specs(3, seed=12)
    specs[1]["broken"] = True
    mixed = S.Server(_build_maybe_broken, precision=F.FMDX_PREC_MIXED, quantum=5, keep_positions=True).run(specs)
    assert mixed[1].failed and (mixed[1].err & F.ERR_POSITION)
    for i in (0, 2):
        alone = S.Server(_build_maybe_broken, precision=F.FMDX_PREC_MIXED, quantum=5,
                         keep_positions=True).run([dict(specs[i], arrival=0.0)])[0]
        assert not mixed[i].failed and not alone.failed
        for k in ("x", "v", "atomid"):
            assert np.array_equal(mixed[i].x[k], alone.x[k]), (i, k)
