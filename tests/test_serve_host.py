"""The plain baseline's first-come-first-served schedule (fmd_amd.serve.run_plain) from given
run times: one job at a time, each starting when it has arrived and the previous one is done."""
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
