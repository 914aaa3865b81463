"""Edge sizes on the GPU: the largest TIP5P replica that fits a workgroup's
LDS (2087, bulk) and the smallest replica (2 molecules), one replica each,
against the oracle;
and intervals of a one-replica job (no exchange pairs) through both
interval entry points.

Tolerances as in test_gpu_parity: step-0 forces and torques relative RMS
<= 1e-5 (FP32 pair math) against the FP64 oracle; windows and molecule
order bit-exact."""
import numpy as np
import pytest

import remd_amd as ra
from remd_amd import inputs as inp
from helpers import oracle_for, rel_rms

pytestmark = pytest.mark.gpu


def _bulk(n, L, model, seed=91):
    rng = np.random.RandomState(seed)
    mt = inp.molecule(model)
    T, P = inp.make_ladder(1, 1, inp.geometric(300.0, 300.0, 1), [0.1e6])
    s = inp.System(model, n, 1, 0, 0, 2.0, 9.0, 8.0, 100, 1, 1, T, P)
    s.states = inp._states(rng, mt, inp.place_lattice(n, L), T, (L, L, L), 0.5, 0.05)
    return s


def _check_step0(s, g):
    g.set_debug(True)
    g.compute_forces(do_sort=True, margin=0.0)
    o = oracle_for(s, 0)
    o.calc_force(True, 0.0)
    so = o.state()
    sg = g.state(0)
    assert np.array_equal(sg["id"], so["id"])
    js, je, full = g.window(0, with_full=True)
    assert np.array_equal(js, so["jstart"]) and np.array_equal(je, so["jend"])
    assert np.array_equal(full, so["wfull"])
    ef, en = rel_rms(sg["f"], so["f"]), rel_rms(sg["n"], so["n"])
    print(s.nmol, s.model, "F", ef, "N", en)
    assert ef <= 1e-5 and en <= 1e-5


# The largest replica is set by the LDS of one workgroup (160 KiB): on the
# global-state path the FP32 site records (72 B per TIP5P molecule), the
# window bounds (4 B) and 5184 B of reduction scratch must fit, so 2087 TIP5P.
TIP5P_MAX = 2087


@pytest.mark.parametrize("path", [ra.PATH_AUTO, ra.PATH_GLOBAL])
def test_largest_tip5p_replica(path):
    """2087 TIP5P on a 13^3 lattice (3.05 A spacing, 39.7 A box), the
    largest TIP5P replica that fits: step-0 forces against the oracle and a
    finite interval; one molecule more is refused with a message."""
    s = _bulk(TIP5P_MAX, 39.7, "TIP5P")
    g = ra.REMDGPU(s, path=path)
    assert g.launch_path()[0] == ra.PATH_GLOBAL
    _check_step0(s, g)
    s.interval = 10
    g.execute_steps(s.T, s.P)
    assert np.all(np.isfinite(g.all_averages()[:, :15]))
    g.close()
    with pytest.raises(RuntimeError, match="too large for workgroup LDS"):
        ra.REMDGPU(_bulk(TIP5P_MAX + 1, 39.7, "TIP5P"), path=path)


@pytest.mark.parametrize("path", [ra.PATH_RESIDENT, ra.PATH_GLOBAL])
def test_smallest_replica_two_molecules(path):
    """Two TIP4P molecules in a 20 A tube (the ABI's smallest replica):
    the window of each is the whole ring."""
    s = inp.config_tube(seed=95, nmol=2, Lz=20.0, ntemp=1, npress=1, nfwall=500)
    g = ra.REMDGPU(s, path=path)
    assert g.launch_path()[0] == path
    _check_step0(s, g)
    g.close()


def test_one_replica_job_intervals_without_exchange_pairs():
    """nreplica = 1: the exchange lists are empty, so the pipelined
    intervals draw no uniforms and give the plain interval loop's bits."""
    s = inp.config_tube(seed=93, nmol=64, Lz=26.0, ntemp=1, npress=1, nfwall=500)
    s.interval = 20
    s.dt_fs = 0.5
    a = ra.REMDGPU(s)
    for _ in range(3):
        a.execute_steps(s.T, s.P)
    b = ra.REMDGPU(s)
    index = np.zeros(1, np.int32)
    T, P = s.T.copy(), s.P.copy()
    rng = ra.GlibcRand(1)
    before = rng.getstate()
    nacc, acc, av, ms = ra.execute_intervals(b, 1, 1, 0, 3, index, T, P, rng)
    assert nacc == 0 and acc.size == 0 and rng.getstate() == before
    assert np.array_equal(av[-1], a.all_averages())
    for k in ("r", "v", "q", "p", "id"):
        assert np.array_equal(a.state(0)[k], b.state(0)[k]), k
    a.close()
    b.close()
