"""Served stream on the device: jobs of different sizes, models and
confinements that arrive while others run are batched in lanes, and each
job's result (every replica's averages and replica 0's state after its last
interval) and its output stream are bit-identical to serving that job alone
-- also for a job preempted by a higher-priority request and resumed from
its saved state."""
import os

import numpy as np
import pytest

from remd_amd.serve import GPUBackend, Request, Server, latency_percentiles, poisson_stream

pytestmark = pytest.mark.gpu


def _serve(reqs, **kw):
    be = GPUBackend()
    try:
        return Server(be, **kw).run(reqs)
    finally:
        be.close()


def _same_outputs(da, db):
    names = sorted(os.listdir(da))
    assert names == sorted(os.listdir(db)) and any(n.endswith(".log") for n in names)
    for n in names:
        assert open(os.path.join(da, n), "rb").read() == open(os.path.join(db, n), "rb").read(), n


def test_each_job_same_alone_and_in_the_mix(tmp_path):
    reqs = poisson_stream(seed=7, njobs=5, rate=200.0, intervals=4, nmol_hi=220)
    for r in reqs:
        r.output_dir = str(tmp_path / "mix" / f"job{r.job_id}")
        print(r.job_id, r.system.model, r.system.confined, r.system.nmol, r.system.nreplica, round(r.arrival, 4))
    mix = _serve(reqs)
    # at least one job was admitted while another was still running
    assert any(mix[b].admitted < mix[a].finished and mix[b].arrival > mix[a].arrival
               for a in mix for b in mix if a != b)
    p50, p99 = latency_percentiles(mix)
    print("p50/p99 interval latency (ms)", p50 * 1e3, p99 * 1e3, "lanes", sorted({m.lane for m in mix.values()}))
    for r in reqs:
        d = str(tmp_path / "alone" / f"job{r.job_id}")
        alone = _serve([Request(r.job_id, 0.0, r.system, r.intervals, output_dir=d)])[r.job_id]
        m = mix[r.job_id]
        assert np.array_equal(m.index, alone.index)
        assert np.all(np.isfinite(m.averages[:, :15]))
        assert np.array_equal(m.averages, alone.averages), r.job_id
        for k in ("r", "v", "q", "p"):
            assert np.array_equal(m.state0[k], alone.state0[k]), (r.job_id, k)
        _same_outputs(r.output_dir, d)


def test_preempted_job_resumes_bit_for_bit(tmp_path):
    """Slots for one job only: a high-priority request preempts the running
    job, which resumes afterwards from its saved device state and ends with
    the bits (and output stream) of an uninterrupted run."""
    reqs = poisson_stream(seed=9, njobs=2, rate=1000.0, intervals=6, nmol_hi=160)
    low, high = reqs
    low.arrival, high.arrival, high.priority = 0.0, 0.0005, 5
    cap = max(low.system.nreplica, high.system.nreplica)
    low.output_dir = str(tmp_path / "low_mix")
    res = _serve([low, high], max_replicas=cap)
    assert res[low.job_id].preemptions >= 1 and res[high.job_id].preemptions == 0
    assert res[high.job_id].finished < res[low.job_id].finished
    alone = _serve([Request(low.job_id, 0.0, low.system, low.intervals, output_dir=str(tmp_path / "low_alone"))])
    a, b = res[low.job_id], alone[low.job_id]
    assert a.intervals == b.intervals == low.intervals
    assert np.array_equal(a.index, b.index) and np.array_equal(a.averages, b.averages)
    for k in ("r", "v", "q", "p"):
        assert np.array_equal(a.state0[k], b.state0[k]), k
    _same_outputs(str(tmp_path / "low_mix"), str(tmp_path / "low_alone"))
