"""Edge sizes on the GPU: the ABI's largest replica (4096 TIP5P, bulk), the
TIP5P replicas either side of the LDS limit of the site records, and the
smallest replica (2 molecules), one replica each, against the oracle;
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


# The global-state path keeps the site records in LDS while they fit (FP32
# records of 72 B per TIP5P molecule + 4 B of window bounds + 5184 B of
# reduction scratch: up to 2087 TIP5P) and in global memory beyond that, up
# to the ABI's 4096 molecules.
TIP5P_LDS_MAX = 2087


@pytest.mark.parametrize("n,L", [(TIP5P_LDS_MAX, 39.7), (TIP5P_LDS_MAX + 1, 39.7)])
def test_tip5p_either_side_of_the_lds_limit(n, L):
    """2087 TIP5P (records in LDS) and 2088 (records in global memory) on
    a 13^3 lattice: step-0 forces against the oracle."""
    s = _bulk(n, L, "TIP5P")
    g = ra.REMDGPU(s)
    assert g.launch_path()[0] == ra.PATH_GLOBAL
    _check_step0(s, g)
    g.close()


@pytest.mark.parametrize("path", [ra.PATH_AUTO, ra.PATH_GLOBAL])
def test_largest_replica_4096_tip5p(path):
    """4096 TIP5P on a 16^3 lattice (3.1 A spacing, 49.7 A box), the ABI's
    largest replica (site records in global memory): step-0 forces against
    the oracle, then 5 NVE steps of 0.5 fs against the oracle's trajectory
    (positions within 1e-6 of the box, mixed precision); 4097 molecules are
    refused.  (The lattice start is far from equilibrium: with the
    Nose-Poincare thermostat its s runs away within 10 steps, in the oracle
    too.)"""
    s = _bulk(4096, 49.7, "TIP5P")
    s.tconstant, s.dt_fs, s.interval = 0, 0.5, 5
    g = ra.REMDGPU(s, path=path)
    assert g.launch_path()[0] == ra.PATH_GLOBAL
    _check_step0(s, g)
    g.execute_steps(s.T, s.P)
    o = oracle_for(s, 0)
    o.initialize_property()
    o.execute_steps(5)
    so, sg = o.state(), g.state(0)
    a_r, b_r = sg["r"][np.argsort(sg["id"])], so["r"][np.argsort(so["id"])]
    d = a_r - b_r
    d -= np.rint(d)
    print("4096 NVE 5 steps max |dr|", np.max(np.abs(d)))
    assert np.max(np.abs(d)) <= 1e-6
    g.close()
    with pytest.raises(RuntimeError, match="nmol must be in"):
        ra.REMDGPU(_bulk(4097, 49.7, "TIP5P"), path=path)


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
