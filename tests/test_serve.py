"""Continuous-batching server logic on the CPU (a fake device backend and a
simulated clock): admission by HBM budget and replica slots in arrival
order, admission of arrivals while other jobs run, retirement, and the
never-fits error."""
import types

import numpy as np
import pytest

from remd_amd.serve import Request, Server, latency_percentiles


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += max(dt, 0.0)


class FakeBackend:
    """Each interval costs 10 ms + 1 us per resident replica."""

    def __init__(self, clock):
        self.clock = clock
        self.active = set()
        self.max_active_reps = 0
        self.max_active_mem = 0
        self.log = []

    def footprint(self, system):
        return system.nreplica * 1000

    def create(self, system):
        return types.SimpleNamespace(system=system, intervals=0)

    def add(self, job):
        self.active.add(id(job))
        self._jobs = getattr(self, "_jobs", {})
        self._jobs[id(job)] = job
        reps = sum(j.system.nreplica for k, j in self._jobs.items() if k in self.active)
        self.max_active_reps = max(self.max_active_reps, reps)
        self.max_active_mem = max(self.max_active_mem, reps * 1000)

    def remove(self, job):
        self.active.discard(id(job))

    def stage(self, job, T, P):
        pass

    def execute(self):
        reps = sum(j.system.nreplica for k, j in self._jobs.items() if k in self.active)
        for k in self.active:
            self._jobs[k].intervals += 1
        self.clock.t += 0.010 + 1e-6 * reps
        self.log.append(sorted(self.active))

    def exchange(self, a, step):
        pass

    def rng(self):
        return None

    def result(self, job):
        return np.full((job.system.nreplica, 19), float(job.intervals)), {}

    def close_job(self, job):
        pass


def _sys(nrep):
    return types.SimpleNamespace(nreplica=nrep, interval=100, T=np.ones(nrep), P=np.zeros(nrep), ntemp=nrep,
                                 npress=1, pconstant=0)


def _run(reqs, **kw):
    clock = FakeClock()
    be = FakeBackend(clock)
    srv = Server(be, clock=clock, sleep=clock.sleep, **kw)
    return srv.run(reqs), be


def test_every_job_runs_its_intervals():
    reqs = [Request(k, 0.003 * k, _sys(8 * (1 + k % 3)), intervals=5 + k) for k in range(10)]
    res, be = _run(reqs)
    assert sorted(res) == list(range(10))
    for r in reqs:
        out = res[r.job_id]
        assert out.intervals == r.intervals
        assert np.all(out.averages == r.intervals)
        assert len(out.latencies) == r.intervals
        assert out.replica_steps == r.intervals * 100 * r.system.nreplica
        assert out.admitted >= r.arrival and out.finished > out.admitted


def test_slot_limit_and_budget_hold():
    reqs = [Request(k, 0.0, _sys(64), intervals=3) for k in range(8)]
    res, be = _run(reqs, max_replicas=128)
    assert be.max_active_reps <= 128
    res, be = _run(reqs, hbm_budget=3 * 64 * 1000)
    assert be.max_active_mem <= 3 * 64 * 1000
    assert len(res) == 8


def test_arrivals_join_while_others_run():
    """Continuous batching: a job arriving while another runs is admitted at
    the next interval boundary, not after the running job finishes."""
    reqs = [Request(0, 0.0, _sys(16), intervals=20), Request(1, 0.05, _sys(16), intervals=4)]
    res, be = _run(reqs)
    assert res[1].admitted < res[0].finished
    assert res[1].finished < res[0].finished
    assert any(len(a) == 2 for a in be.log)


def test_admission_is_in_arrival_order():
    """A head request that does not fit blocks later ones (no overtaking)."""
    reqs = [Request(0, 0.0, _sys(100), intervals=10), Request(1, 0.001, _sys(100), intervals=2),
            Request(2, 0.002, _sys(8), intervals=2)]
    res, be = _run(reqs, max_replicas=128)
    assert res[2].admitted >= res[1].admitted >= res[0].finished - 1e-12


def test_never_fitting_job_is_an_error():
    with pytest.raises(ValueError):
        _run([Request(0, 0.0, _sys(300), intervals=1)], max_replicas=256)


def test_latency_percentiles():
    reqs = [Request(k, 0.0, _sys(8), intervals=10) for k in range(3)]
    res, _ = _run(reqs)
    p50, p99 = latency_percentiles(res)
    assert 0.010 <= p50 <= p99 <= 0.011
