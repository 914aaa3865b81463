"""The allocation-free interval entry point (remd_gpu_execute_steps_noalloc)
against remd_gpu_execute_steps, through the C ABI: the same bits, state and
averages, on two calls in a row -- the second with a smaller job -- and a
job whose launch scratch exceeds the static array is refused."""
import numpy as np
import pytest

import remd_amd as ra
from remd_amd import inputs as inp
from remd_amd.workloads import relax

pytestmark = pytest.mark.gpu


def _pair(system, path):
    return ra.REMDGPU(system, path=path), ra.REMDGPU(system, path=path)


def _same(ga, gb, nrep):
    assert np.all(np.isfinite(ga.all_averages()[:, :15]))
    assert np.array_equal(ga.all_averages(), gb.all_averages())
    for r in range(nrep):
        sa, sb = ga.state(r), gb.state(r)
        for name in ("r", "v", "q", "p", "f", "n", "id", "scal"):
            assert np.array_equal(sa[name], sb[name]), (r, name)


def test_noalloc_equals_execute_steps():
    big = relax(inp.config_tube(seed=61, nmol=144, ntemp=4, npress=2, nfwall=500))
    small = relax(inp.config_tube(seed=62, nmol=60, Lz=26.0, ntemp=2, npress=2, nfwall=500, model="TIP3P"))
    for s in (big, small):
        s.interval = 20
    print("noalloc: 144 TIP4P x 8 replicas (resident), then 60 TIP3P x 4 (split-global)", flush=True)
    for s, path in ((big, ra.PATH_RESIDENT), (small, ra.PATH_SPLIT)):
        ref, na = _pair(s, path)
        for T in (s.T.copy(), s.T[::-1].copy()):   # the second interval runs the exchange rescale
            ref.execute_steps(T, s.P)
            na.execute_steps_noalloc(T, s.P)
            _same(ref, na, s.nreplica)
        assert na.launch_path()[1] > 1, "the job should be split over workgroups (exchange rows in scratch)"
        ref.close()
        na.close()


def test_noalloc_refuses_oversized_scratch():
    """128 replicas of 1000 TIP4P split over workgroups need about 20 MB of
    exchange rows, more than the 16 MiB static scratch."""
    s = inp.config_tube(seed=63, nmol=1000, Lz=350.0, ntemp=32, npress=4, nfwall=500)
    s.interval = 2
    g = ra.REMDGPU(s)
    with pytest.raises(RuntimeError, match="bytes of launch scratch"):
        g.execute_steps_noalloc(s.T, s.P)
    g.close()
