"""Every benchmark configuration at its own size on the device.

- Plain vs GPU drop-in (nanotube, 144 TIP4P, 64 replicas): the reference
  build's seeded config-2 replica (fixture, FP64 build) in all 64 replica
  slots -- step-0 forces and torques and a 20-step trajectory against the
  fixture, and the 64 replicas bit-identical to each other.
- Precision study (nano-slit, 512 TIP5P, 40 x 40 A, 128 replicas): FP32
  pair math with FP64 accumulation against FP64 pair math on every replica
  (step-0 force <= 1e-4 relative RMS), both against the FP64 oracle on three
  replicas, and equal Nose-Poincare Hamiltonian drift over 10 000 steps.
- Batched uneven replicas (8 tubes R = 5..12 A, 256 replicas): the
  autotuned batch equals each job alone bit for bit, and one replica per
  tube matches the FP64 oracle at step 0.
- Served stream (nmol up to 2048, 50 intervals, 20 jobs/s): jobs in the mix
  equal the same jobs served alone."""
import numpy as np
import pytest

import remd_amd as ra
from remd_amd import inputs as inp
from remd_amd.workloads import relax
from helpers import by_id, oracle_for, rel_rms

pytestmark = pytest.mark.gpu


def _fixture_cfg2(d, nrep):
    st = d["init_state"]
    ids = d["init_ids"].astype(np.int32)
    sc0 = d["scal0"]
    T, P = float(d["T"]), float(d["P"])
    s = inp.System("TIP4P", 144, 1, 2, 1, 2.0, 9.0, 8.0, 20, nrep, 1, np.full(nrep, T), np.full(nrep, P),
                   wall_length=7.0, nfwall=1000, sgm_wall=3.277, eps_wall=0.3886, rho_wall=0.38016)
    # after the driver's velocity scaling: s = 1, Ps = 0, Pv = 0 (src/molecules.cxx:261-311)
    s.states = [inp.ReplicaState(r=st[:, 0:3].copy(), v=st[:, 3:6].copy(), q=st[:, 6:10].copy(),
                                 p=st[:, 10:14].copy(), s=1.0, Ps=0.0, Q=sc0[2], Pv=(0.0, 0.0, 0.0), W=sc0[6],
                                 L=tuple(sc0[7:10]), ids=ids) for _ in range(nrep)]
    return s


@pytest.mark.parametrize("prec,tol_f,tol_x", [(1, 1e-10, 1e-8), (0, 1e-5, 2e-3)])
def test_config2_64_replicas_against_reference_fixture(golden, prec, tol_f, tol_x):
    d = golden("ref_tube_cfg2.npz")
    s = _fixture_cfg2(d, 64)
    g = ra.REMDGPU(s, precision=prec)
    assert g.launch_path()[0] == ra.PATH_RESIDENT
    init, init_ids = d["init_state"], d["init_ids"]
    f_ref, n_ref = by_id(init[:, 14:17], init_ids), by_id(init[:, 17:21], init_ids)
    first = g.state(0)
    for k in range(64):
        st = g.state(k)
        for name in ("f", "n", "r", "id"):
            assert np.array_equal(st[name], first[name]), (k, name)
    ef = rel_rms(by_id(first["f"], first["id"]), f_ref)
    en = rel_rms(by_id(first["n"], first["id"]), n_ref)
    print("step 0: force", ef, "torque", en)
    assert ef <= tol_f and en <= tol_f
    av = g.averages(0)
    props = dict(zip(d["prop_keys"], d["init_props"]))
    for name, col in (("lj", 0), ("clmb", 1), ("wall", 2)):
        assert abs(av[col] - props[name]) <= 10 * tol_f * abs(props["pot"]), name
    g.execute_steps(s.T, s.P)
    ref, ref_ids = d["int_state"], d["int_ids"]
    for k in (0, 31, 63):
        st = g.state(k)
        for name, sl in (("r", slice(0, 3)), ("v", slice(3, 6)), ("q", slice(6, 10)), ("p", slice(10, 14))):
            a, b = by_id(st[name], st["id"]), by_id(ref[:, sl], ref_ids)
            if name == "r":
                dd = a - b
                dd -= np.rint(dd)
                err = float(np.max(np.abs(dd)))
            else:
                err = rel_rms(a, b)
            print(k, name, err)
            assert err <= tol_x, (k, name, err)
    g.close()


@pytest.fixture(scope="module")
def slit512():
    return relax(inp.config_slit())


def test_config3_slit_precision_forces_and_oracle(slit512):
    """Slit wall: this build's two 10-4 sheets, not the reference's 9-3
    table, so the slit figures compare mixed against FP64 runs of the same
    kernels and the oracle's 10-4 restatement, never the reference."""
    s = slit512
    assert s.nreplica == 128 and s.nmol == 512
    g0, g1 = ra.REMDGPU(s, precision=0), ra.REMDGPU(s, precision=1)
    ferr = []
    for k in range(s.nreplica):
        a, b = g0.state(k), g1.state(k)
        ferr.append(rel_rms(by_id(a["f"], a["id"]), by_id(b["f"], b["id"])))
    print("mixed vs FP64 step-0 force rel RMS: max", max(ferr), "median", np.median(ferr))
    assert max(ferr) <= 1e-4 and np.median(ferr) <= 2e-5
    for k in (0, 63, 127):
        o = oracle_for(s, k)
        o.initialize_property()
        so = o.state()
        fo = by_id(so["f"], so["id"])
        for g, tol in ((g1, 1e-10), (g0, 1e-5)):
            st = g.state(k)
            err = rel_rms(by_id(st["f"], st["id"]), fo)
            print("replica", k, "precision", 1 if g is g1 else 0, "force vs FP64 oracle", err)
            assert err <= tol
        sc = o.scalars()
        av0, av1 = g0.averages(k), g1.averages(k)
        assert abs((av1[0] + av1[1] + av1[2]) - sc["pot"]) <= 1e-10 * abs(sc["pot"])
        assert abs((av0[0] + av0[1] + av0[2]) - sc["pot"]) <= 1e-5 * abs(sc["pot"])
    g0.close()
    g1.close()


def test_config3_slit_drift_10k_steps_equal(slit512):
    """Nose-Poincare Hamiltonian drift max |H - H0| over 10 000 steps (100
    intervals, no exchanges) of every replica: FP32 pair math with FP64
    accumulation drifts as FP64 pair math does."""
    s = slit512
    worst = {}
    for prec in (0, 1):
        g = ra.REMDGPU(s, precision=prec)
        w = np.zeros(s.nreplica)
        for _ in range(100):
            g.execute_steps(s.T, s.P)
            av = g.all_averages()
            assert np.all(np.isfinite(av[:, :15]))
            w = np.maximum(w, np.abs(av[:, 12]))
        kin = av[:, 4] + av[:, 5]
        worst[prec] = (float(w.max()), float(np.median(w)), float(np.median(kin)))
        g.close()
    print("drift max/median |H-H0| (mixed, fp64):", worst)
    (m_max, m_med, kin), (d_max, d_med, _) = worst[0], worst[1]
    assert 0.8 <= m_max / d_max <= 1.25 and 0.8 <= m_med / d_med <= 1.25
    assert d_max <= 0.02 * kin


def test_config4_uneven_batch_equals_alone_and_oracle():
    systems = [relax(s) for s in inp.config_uneven_batch()]
    assert sum(s.nreplica for s in systems) == 256
    for s in systems:
        s.interval = 50
    conds = [(s.T.copy(), s.T[::-1].copy()) for s in systems]
    alone = []
    for s, (T1, T2) in zip(systems, conds):
        g = ra.REMDGPU(s)
        g.execute_steps(T1, s.P)
        g.execute_steps(T2, s.P)
        alone.append((g.all_averages(), g.state(0), g.state(s.nreplica - 1)))
        g.close()
    jobs = [ra.REMDGPU(s) for s in systems]
    for s, g in zip(systems, jobs):   # step 0 of one replica per tube against the FP64 oracle
        o = oracle_for(s, 0)
        o.initialize_property()
        so, st = o.state(), g.state(0)
        err = rel_rms(by_id(st["f"], st["id"]), by_id(so["f"], so["id"]))
        print(s.model, s.nmol, "R", s.wall_length, "step-0 force vs oracle", err)
        assert err <= 1e-5
    bt = ra.Batch()
    for g in jobs:
        bt.add(g)
    for i in range(2):
        for g, s, c in zip(jobs, systems, conds):
            g.set_conditions(c[i], s.P)
        bt.execute()
    print("plan", bt.parts())
    for g, s, (av, s0, s1) in zip(jobs, systems, alone):
        assert np.all(np.isfinite(av[:, :15]))
        assert np.array_equal(g.all_averages(), av)
        for rep, ref in ((0, s0), (s.nreplica - 1, s1)):
            st = g.state(rep)
            for name in ("r", "v", "q", "p", "id"):
                assert np.array_equal(st[name], ref[name]), (s.nmol, rep, name)
    bt.close()
    for g in jobs:
        g.close()


def test_config5_served_stream_at_full_bounds():
    from remd_amd.serve import GPUBackend, Request, Server, latency_percentiles, poisson_stream

    reqs = poisson_stream(seed=11, njobs=6, rate=20.0, intervals=50, nmol_hi=2048)
    print([(r.system.model, r.system.confined, r.system.nmol, r.system.nreplica) for r in reqs])
    be = GPUBackend()
    mix = Server(be, hbm_budget=200e9).run(reqs)
    be.close()
    print("p50/p99 ms", [x * 1e3 for x in latency_percentiles(mix)])
    big = max(reqs, key=lambda r: r.system.nmol * r.system.nreplica)
    for r in {id(x): x for x in (reqs[0], reqs[-1], big)}.values():
        be = GPUBackend()
        alone = Server(be).run([Request(r.job_id, 0.0, r.system, r.intervals)])[r.job_id]
        be.close()
        m = mix[r.job_id]
        assert m.intervals == 50
        assert np.array_equal(m.index, alone.index)
        assert np.array_equal(m.averages, alone.averages), r.job_id
        for k in ("r", "v", "q", "p"):
            assert np.array_equal(m.state0[k], alone.state0[k]), (r.job_id, k)
