"""Checkpoint / restart of device-resident jobs (REMDGPU::ConvertVariablesToHost,
src/remdgpu.cu:806-854; IOManager::WriteCheckPoint / ReadCheckPoint,
src/io_manager.cxx:528-634): two intervals straight give the same bits as
one interval, a checkpoint, a fresh job loaded from it and one more interval
-- on every kernel path, with an exchange rescale in the second interval."""
import numpy as np
import pytest

import remd_amd as ra
from remd_amd import inputs as inp
from remd_amd.workloads import relax

pytestmark = pytest.mark.gpu

PATHS = [ra.PATH_RESIDENT, ra.PATH_GLOBAL, ra.PATH_SPLIT]


def _system():
    s = relax(inp.config_tube(seed=71, nmol=72, Lz=28.0, ntemp=2, npress=2, nfwall=500))
    s.interval = 20
    return s


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("prec", [0, 1])
def test_checkpoint_restart_equals_straight(path, prec):
    s = _system()
    T1, T2 = s.T.copy(), s.T[[1, 0, 3, 2]].copy()   # swapped pairs: the second interval rescales
    straight = ra.REMDGPU(s, precision=prec, path=path)
    straight.execute_steps(T1, s.P)
    straight.execute_steps(T2, s.P)
    ref_av = straight.all_averages()
    ref_st = [straight.state(r) for r in range(s.nreplica)]
    straight.close()

    first = ra.REMDGPU(s, precision=prec, path=path)
    first.execute_steps(T1, s.P)
    blob = first.checkpoint_save()
    first.close()
    resumed = ra.REMDGPU(s, precision=prec, path=path)   # created afresh (new H0), then overwritten
    resumed.checkpoint_load(blob)
    resumed.execute_steps(T2, s.P)
    assert np.all(np.isfinite(ref_av[:, :15]))
    assert np.array_equal(resumed.all_averages(), ref_av)
    for r in range(s.nreplica):
        st = resumed.state(r)
        for name in ("r", "v", "f", "q", "p", "n", "id", "scal"):
            assert np.array_equal(st[name], ref_st[r][name]), (r, name)
    resumed.close()


def test_checkpoint_refuses_other_shapes():
    s = _system()
    g = ra.REMDGPU(s)
    blob = g.checkpoint_save()
    other = inp.config_tube(seed=72, nmol=64, Lz=26.0, ntemp=2, npress=2, nfwall=500)
    h = ra.REMDGPU(other)
    with pytest.raises(RuntimeError, match="shape"):
        h.checkpoint_load(blob)
    h64 = ra.REMDGPU(s, precision=1)
    with pytest.raises(RuntimeError, match="precision"):
        h64.checkpoint_load(blob)
    with pytest.raises(RuntimeError, match="not a remd_gpu checkpoint"):
        g.checkpoint_load(b"x" * len(blob))
    for j in (g, h, h64):
        j.close()
