"""Continuous-batching server logic on the CPU (a fake device backend and a
simulated clock): admission by HBM budget and replica slots in arrival
order, admission of arrivals while other jobs run, lanes of different
interval cost advancing independently, priority preemption with resume,
retirement, and the never-fits error."""
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


class FakeRng:
    def __init__(self):
        self.state = b"\x00"

    def getstate(self):
        return self.state

    def setstate(self, s):
        self.state = s


class FakeBackend:
    """A lane's interval lasts as long as its slowest job (`system.dur`
    seconds); lanes run concurrently.  A job counts the intervals it ran."""

    def __init__(self, clock):
        self.clock = clock
        self.lanes = {}
        self.end = {}
        self.max_active_reps = 0
        self.max_active_mem = 0
        self.log = []
        self.saved = 0

    def footprint(self, system):
        return system.nreplica * 1000

    @staticmethod
    def cost(system):
        return system.cost

    def create(self, system):
        return types.SimpleNamespace(system=system, intervals=0)

    def _active(self):
        return [j for L in self.lanes.values() for j in L]

    def add(self, lane, job):
        self.lanes.setdefault(lane, []).append(job)
        reps = sum(j.system.nreplica for j in self._active())
        self.max_active_reps = max(self.max_active_reps, reps)
        self.max_active_mem = max(self.max_active_mem, reps * 1000)

    def remove(self, lane, job):
        self.lanes[lane] = [j for j in self.lanes[lane] if j is not job]

    def stage(self, job, T, P):
        pass

    def launch(self, lane):
        self.end[lane] = self.clock.t + max(j.system.dur for j in self.lanes[lane])
        self.log.append((lane, sorted(id(j) for j in self.lanes[lane]), self.clock.t))

    def poll(self, lane):
        return self.clock.t >= self.end[lane]

    def wait(self, lane):
        self.clock.t = max(self.clock.t, self.end[lane])
        for j in self.lanes[lane]:
            j.intervals += 1

    def exchange(self, a, step):
        return []

    def averages(self, job):
        return np.zeros((job.system.nreplica, 19))

    def exchange_pairs(self, nt, npr, t):
        return np.zeros(0, int), np.zeros(0, int)

    def rng(self):
        return FakeRng()

    def save(self, job):
        self.saved += 1
        return ("state", job.intervals)

    def restore(self, system, blob):
        return types.SimpleNamespace(system=system, intervals=blob[1])

    def result(self, job):
        return np.full((job.system.nreplica, 19), float(job.intervals)), {}

    def close_job(self, job):
        pass


def _sys(nrep, dur=0.010, cost=1.0):
    return types.SimpleNamespace(nreplica=nrep, interval=100, T=np.ones(nrep), P=np.zeros(nrep), ntemp=nrep,
                                 npress=1, pconstant=0, dur=dur, cost=cost)


def _run(reqs, **kw):
    clock = FakeClock()
    be = FakeBackend(clock)
    kw.setdefault("lane_bounds", (10.0,))
    srv = Server(be, clock=clock, sleep=clock.sleep, **kw)
    return srv.run(reqs), be


def test_every_job_runs_its_intervals():
    reqs = [Request(k, 0.003 * k, _sys(8 * (1 + k % 3), cost=1.0 + 20 * (k % 2)), intervals=5 + k)
            for k in range(10)]
    res, be = _run(reqs)
    assert sorted(res) == list(range(10))
    for r in reqs:
        out = res[r.job_id]
        assert out.intervals == r.intervals
        assert np.all(out.averages == r.intervals)
        assert len(out.latencies) == r.intervals
        assert out.replica_steps == r.intervals * 100 * r.system.nreplica
        assert out.admitted >= r.arrival and out.finished > out.admitted
        assert out.lane == (1 if r.system.cost >= 10.0 else 0)


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
    assert any(len(ids) == 2 for _, ids, _ in be.log)


def test_light_lane_runs_many_intervals_during_a_heavy_one():
    """Lanes advance independently: a light job finishes 10 intervals of
    10 ms while a heavy job (100 ms per interval) runs its first."""
    reqs = [Request(0, 0.0, _sys(8, dur=0.100, cost=100.0), intervals=3),
            Request(1, 0.0, _sys(8, dur=0.010, cost=1.0), intervals=10)]
    res, be = _run(reqs)
    assert res[1].finished < 0.105 and res[0].finished >= 0.300
    assert res[0].lane == 1 and res[1].lane == 0
    assert max(res[1].latencies) <= 0.010 + 1e-3


def test_admission_is_in_arrival_order():
    """A head request that does not fit blocks later ones (no overtaking)."""
    reqs = [Request(0, 0.0, _sys(100), intervals=10), Request(1, 0.001, _sys(100), intervals=2),
            Request(2, 0.002, _sys(8), intervals=2)]
    res, be = _run(reqs, max_replicas=128)
    assert res[2].admitted >= res[1].admitted >= res[0].finished - 1e-12


def test_priority_request_preempts_and_the_preempted_job_resumes():
    """A high-priority request that does not fit preempts the running
    low-priority job (its state is saved and freed); the preempted job
    resumes from its saved state and still runs exactly its intervals."""
    reqs = [Request(0, 0.0, _sys(100), intervals=10), Request(1, 0.025, _sys(100), intervals=3, priority=5)]
    res, be = _run(reqs, max_replicas=128)
    assert be.saved == 1
    assert res[0].preemptions == 1 and res[1].preemptions == 0
    assert res[0].intervals == 10 and np.all(res[0].averages == 10)
    assert res[1].finished < res[0].finished
    assert res[1].admitted <= 0.031
    assert be.max_active_reps <= 128


def test_equal_priority_does_not_preempt():
    reqs = [Request(0, 0.0, _sys(100), intervals=10), Request(1, 0.025, _sys(100), intervals=3)]
    res, be = _run(reqs, max_replicas=128)
    assert be.saved == 0 and res[1].admitted >= res[0].finished - 1e-12


def test_never_fitting_job_is_an_error():
    with pytest.raises(ValueError):
        _run([Request(0, 0.0, _sys(300), intervals=1)], max_replicas=256)


def test_latency_percentiles():
    reqs = [Request(k, 0.0, _sys(8), intervals=10) for k in range(3)]
    res, _ = _run(reqs)
    p50, p99 = latency_percentiles(res)
    assert 0.010 <= p50 <= p99 <= 0.011
