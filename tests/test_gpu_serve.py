"""Served stream on the device: jobs of different sizes, models and
confinements that arrive while others run are batched together, and each
job's result (every replica's averages and replica 0's state after its last
interval) is bit-identical to serving that job alone."""
import numpy as np
import pytest

from remd_amd.serve import GPUBackend, Server, latency_percentiles, poisson_stream

pytestmark = pytest.mark.gpu


def _serve(reqs):
    be = GPUBackend()
    try:
        return Server(be).run(reqs)
    finally:
        be.close()


def test_each_job_same_alone_and_in_the_mix():
    reqs = poisson_stream(seed=7, njobs=5, rate=200.0, intervals=4, nmol_hi=220)
    for r in reqs:
        print(r.job_id, r.system.model, r.system.confined, r.system.nmol, r.system.nreplica, round(r.arrival, 4))
    mix = _serve(reqs)
    # at least one job was admitted while another was still running
    assert any(mix[b].admitted < mix[a].finished and mix[b].arrival > mix[a].arrival
               for a in mix for b in mix if a != b)
    p50, p99 = latency_percentiles(mix)
    print("p50/p99 interval latency (ms)", p50 * 1e3, p99 * 1e3)
    for r in reqs:
        alone = _serve([type(r)(r.job_id, 0.0, r.system, r.intervals)])[r.job_id]
        m = mix[r.job_id]
        assert np.array_equal(m.index, alone.index)
        assert np.array_equal(m.averages, alone.averages), r.job_id
        for k in ("r", "v", "q", "p"):
            assert np.array_equal(m.state0[k], alone.state0[k]), (r.job_id, k)
