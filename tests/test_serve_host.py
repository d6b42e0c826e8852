"""Host logic of fmd_amd.serve: the plain baseline's first-come-first-served schedule (run_plain)
from given run times, one job at a time, each starting when it has arrived and the previous one is
done; and the request loop (Server) over a stand-in batch, with and without the builder thread."""
import pytest

from fmd_amd import serve as S


def test_fcfs_schedule_from_run_times():
    specs = [dict(index=0, arrival=0.0), dict(index=1, arrival=0.1), dict(index=2, arrival=0.5),
             dict(index=3, arrival=5.0)]
    service = {}
    for i, t in enumerate((1.0, 1.0, 0.2, 0.3)):
        r = S.Result(i, 0.0)
        r.natoms, r.nsteps, r.temperature, r.pe, r.err, r.natoms_end = 10, 5, 1.0, 2.0, 0, 10
        service[i] = (t, r)
    out = S.run_plain(specs, None, service=service)
    assert [r.admit for r in out] == [0.0, 1.0, 2.0, 5.0]
    assert [r.finish for r in out] == [1.0, 2.0, 2.2, 5.3]
    assert [round(r.latency, 12) for r in out] == [1.0, 1.9, 1.7, 0.3]
    s = S.summarize(out)
    assert s["jobs"] == 4 and abs(s["makespan_s"] - 5.3) < 1e-12


class _FakeBatch:
    """stands in for fmd_amd.Batch: members stepped together, statuses in join order"""

    def __init__(self, precision):
        self.members = []
        self.steps = {}
        self.log = []

    def set_check_interval(self, n):
        pass

    def add(self, dev, dt):
        self.members.append(dev)
        self.steps[dev] = 0
        self.log.append(("add", dev))

    def remove(self, dev):
        self.members.remove(dev)
        self.log.append(("remove", dev, self.steps[dev]))

    def step(self, q):
        assert self.members
        for d in self.members:
            self.steps[d] += q

    def status(self):
        class T:
            pass
        out = []
        for d in self.members:
            t = T()
            t.natoms, t.temperature, t.pe = 10, 300.0, -1.0
            t.err = 1 if (d == 3 and self.steps[d] >= 20) else 0   # job 3 breaks after 20 steps
            out.append(t)
        return out

    def close(self):
        pass


class _FakeJob:
    def __init__(self, sp):
        self.dev = sp["index"]
        self.dt = 1e-3
        self.natoms = 10
        self.nsteps = sp["nsteps"]
        self.freed = False

    def free(self):
        self.freed = True


def _run_fake(monkeypatch, builders):
    made = []

    def batch(precision):
        made.append(_FakeBatch(precision))
        return made[-1]
    monkeypatch.setattr(S, "Batch", batch)
    monkeypatch.setattr(S, "memory_of", lambda dev: 1 << 20)
    specs = [dict(index=i, arrival=0.0005 * i, nsteps=n, natoms_hint=100)
             for i, n in enumerate((50, 130, 75, 90, 40, 200, 60))]
    jobs = {}

    def build(sp, precision):
        jobs[sp["index"]] = j = _FakeJob(sp)
        return j
    srv = S.Server(build, quantum=25, builders=builders, max_jobs=4)
    out = srv.run(specs)
    return specs, jobs, out, made[0]


@pytest.mark.parametrize("builders", [0, 1])
def test_request_loop_schedule(monkeypatch, builders):
    """The loop admits in arrival order within the member limit, steps in quanta that never
    pass a member's step count, retires finished members and a failed one (its error word
    set) without stopping the others, and frees every job, with jobs built on the loop's
    thread or on a builder thread."""
    specs, jobs, out, b = _run_fake(monkeypatch, builders)
    assert [r.index for r in out] == list(range(len(specs)))
    adds = [e[1] for e in b.log if e[0] == "add"]
    assert adds == sorted(adds)                       # arrival order
    live = peak = 0
    for e in b.log:
        live += 1 if e[0] == "add" else -1
        peak = max(peak, live)
    assert peak <= 4                                  # the member limit
    removed = {e[1]: e[2] for e in b.log if e[0] == "remove"}
    for sp in specs:
        if sp["index"] != 3:
            assert removed[sp["index"]] == sp["nsteps"]   # exactly its step count, never past it
    for sp, r in zip(specs, out):
        assert jobs[sp["index"]].freed
        assert r.admit >= sp["arrival"] and r.finish >= r.admit
        if sp["index"] == 3:
            assert r.failed and r.err == 1
        else:
            assert not r.failed
    s = S.summarize(out)
    assert s["jobs"] == len(specs) and s["failed"] == 1


@pytest.mark.parametrize("builders", [0, 1])
def test_failed_build_releases_every_built_job(monkeypatch, builders):
    """A build that raises ends the run with its exception, after every job already built,
    admitted or not, has been released."""
    monkeypatch.setattr(S, "Batch", _FakeBatch)
    monkeypatch.setattr(S, "memory_of", lambda dev: 1 << 20)
    specs = [dict(index=i, arrival=0.0, nsteps=50, natoms_hint=10) for i in range(5)]
    made = {}

    def build(sp, precision):
        if sp["index"] == 3:
            raise RuntimeError("broken build")
        made[sp["index"]] = j = _FakeJob(sp)
        return j
    with pytest.raises(RuntimeError, match="broken build"):
        S.Server(build, builders=builders, quantum=10).run(specs)
    assert sorted(made) == [0, 1, 2, 4] and all(j.freed for j in made.values())
