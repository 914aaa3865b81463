"""HIP path vs the oracle and the reference fixtures (needs the GPU).

Tolerances (stated per quantity):
  * integer/index work (sort order, windows, exchange decisions): bit-exact;
  * step-0 forces and torques, FP32 pair math + FP64 accumulation:
    relative RMS <= 1e-5 against the FP64 oracle / reference FP64 build;
  * the same with FP64 pair math: relative RMS <= 1e-11;
  * energies at step 0: |dE| <= 1e-5 * sum|E| (mixed), 1e-10 (FP64);
  * one 100-step interval from the reference fixture: FP64 pair math
    <= 1e-7 relative on positions / momenta (chaotic growth of 1e-16
    summation-order differences), mixed <= 1e-3;
  * one 20-step interval of each barostat mode (none, 1D, 2D area, 3D
    isotropic) against the FP64 oracle: FP64 pair math <= 1e-9, mixed
    <= 1e-4 on positions, momenta and the thermostat / barostat scalars.
"""
import numpy as np
import pytest

import pyoracle as po
import remd_amd as ra
from remd_amd import inputs as inp
from helpers import PROP_KEYS, by_id, fixture_system, oracle_for, rel_rms

pytestmark = pytest.mark.gpu


def _props(d, key):
    return dict(zip(PROP_KEYS, d[key]))


# ------------------------------------------------------------------ b1
@pytest.mark.parametrize("R,nfw", [(7.0, 1000), (7.4, 4096), (12.0, 257)])
def test_wall_table_tube_matches_oracle(R, nfw):
    g = ra.wall_table(1, nfw, R, 3.277, 0.3886, 0.38016)
    o = po.wall_table_tube(nfw, R, 3.277, 0.3886, 0.38016)
    assert np.allclose(g, o, rtol=1e-12, atol=1e-12 * np.max(np.abs(o)))


def test_wall_table_tube_matches_reference_fixture(golden):
    d = golden("ref_wall_tables.npz")
    g = ra.wall_table(1, 1000, 7.0, 3.277, 0.3886, 0.38016)
    ref = d["cfg2"]
    assert np.allclose(g, ref, rtol=1e-12, atol=1e-12 * np.max(np.abs(ref)))


def test_wall_table_slit_matches_oracle():
    g = ra.wall_table(2, 1000, 10.0, 3.277, 0.3886, 0.38016)
    o = po.wall_table_slit(1000, 10.0, 3.277, 0.3886, 0.38016)
    assert np.allclose(g, o, rtol=1e-12, atol=1e-12 * np.max(np.abs(o)))


# ------------------------------------------------------------------ c1
@pytest.mark.parametrize("n,w,seed", [(144, 0.19, 0), (150, 0.32, 1), (512, 0.24, 2), (2048, 0.1, 3),
                                      (37, 0.45, 4), (256, 0.482, 5)])
def test_sort_window_bitexact(n, w, seed):
    rng = np.random.RandomState(seed)
    keys = rng.random_sample((3, n))
    keys[1, : n // 3] = keys[1, 0]  # ties
    keys[2] = np.round(keys[2] * 7) / 7.0  # lattice layers with many ties
    keys[2] = np.mod(keys[2], 1.0)
    order, js, je = ra.sort_window(keys, w)
    for row in range(3):
        o2, js2, je2 = po.sort_window(keys[row], w)
        assert np.array_equal(order[row], o2)
        assert np.array_equal(js[row], js2)
        assert np.array_equal(je[row], je2)


# ------------------------------------------------------------- a1..a5, b2
PATHS = [ra.PATH_RESIDENT, ra.PATH_GLOBAL, ra.PATH_SPLIT]


def bulk_large(n=600, L=26.2, seed=41, model="TIP4P", ntemp=2):
    """A bulk replica too large for the resident path's LDS (n = 600)."""
    rng = np.random.RandomState(seed)
    mt = inp.molecule(model)
    T, P = inp.make_ladder(ntemp, 1, inp.geometric(260.0, 320.0, ntemp), [0.1e6])
    s = inp.System(model, n, 1, 0, 0, 2.0, 9.0, 8.0, 100, ntemp, 1, T, P)
    s.states = inp._states(rng, mt, inp.place_lattice(n, L), T, (L, L, L), 0.5, 0.05)
    return s


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("prec,tol_f,tol_e", [(0, 1e-5, 1e-5), (1, 1e-11, 1e-10)])
def test_step0_forces_reference_fixture(golden, prec, tol_f, tol_e, path):
    d = golden("ref_tube_chk.npz")
    sysm = fixture_system(d, "f64")
    g = ra.REMDGPU(sysm, precision=prec, path=path)
    assert g.launch_path()[0] == path
    s = g.state(0)
    ref = d["f64_init_state"]
    ids_ref = d["f64_init_ids"]
    assert np.array_equal(s["id"], ids_ref)  # same sort order
    f_ref, n_ref = ref[:, 14:17], ref[:, 17:21]
    ef = rel_rms(s["f"], f_ref)
    en = rel_rms(s["n"], n_ref)
    print(f"prec={prec} force rel-RMS {ef:.3e} torque rel-RMS {en:.3e}")
    assert ef <= tol_f and en <= tol_f
    pr = _props(d, "f64_init_props")
    av = g.averages(0)
    scale = abs(pr["lj"]) + abs(pr["clmb"]) + abs(pr["wall"])
    for k, idx in (("lj", 0), ("clmb", 1), ("wall", 2)):
        print(k, av[idx], pr[k], av[idx] - pr[k])
        assert abs(av[idx] - pr[k]) <= tol_e * scale
    g.close()


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("model,conf", [("TIP3P", 1), ("TIP4P", 1), ("TIP5P", 1), ("TIP4P", 0), ("TIP5P", 2),
                                        ("TIP3P", 2), ("TIP4P", 3), ("TIP5P", 4)])
@pytest.mark.parametrize("prec", [0, 1])
def test_step0_forces_vs_oracle(model, conf, prec, path):
    """conf 3: 64 molecules on two lattice planes of a bulk box, so every
    window is the whole ring (the whole-ring window rule); conf 4: 600
    molecules, too large for the resident path (run on the global path)."""
    if conf == 4:
        if path == ra.PATH_RESIDENT:
            pytest.skip("600 TIP5P do not fit the resident path's LDS")
        sysm = bulk_large(model=model)
        conf = 0
    elif conf == 3:
        sysm = inp.config_bulk(seed=11, nmol=64, ntemp=2, model=model)
        conf = 0
    elif conf == 0:
        sysm = inp.config_bulk(seed=11, ntemp=2, model=model)
    elif conf == 1:
        sysm = inp.config_tube(seed=12, nmol=120, ntemp=2, npress=1, model=model)
    else:
        sysm = inp.config_slit(seed=13, nmol=200, ntemp=2, npress=1, model=model, Lxy=30.0)
    g = ra.REMDGPU(sysm, precision=prec, path=path)
    g.set_debug(True)
    g.compute_forces(do_sort=True, margin=0.0)
    tol = 1e-5 if prec == 0 else 1e-11
    for rep in range(sysm.nreplica):
        o = oracle_for(sysm, rep)
        o.calc_force(True, 0.0)
        so = o.state()
        sg = g.state(rep)
        assert np.array_equal(sg["id"], so["id"])
        js, je, full = g.window(rep, with_full=True)
        assert np.array_equal(js, so["jstart"]) and np.array_equal(je, so["jend"])
        assert np.array_equal(full, so["wfull"])
        ef = rel_rms(sg["f"], so["f"])
        en = rel_rms(sg["n"], so["n"])
        esf = rel_rms(g.site_forces(rep), so["site_f"])
        print(model, conf, prec, rep, f"F {ef:.2e} N {en:.2e} site {esf:.2e}")
        assert ef <= tol and en <= tol and esf <= tol
        av = g.averages(rep)
        sc = o.scalars()
        scale = abs(sc["lj"]) + abs(sc["clmb"]) + abs(sc["wall"]) + 1.0
        for k, idx in (("lj", 0), ("clmb", 1), ("wall", 2)):
            assert abs(av[idx] - sc[k]) <= 10 * tol * scale, (k, av[idx], sc[k])
    g.close()


# -------------------------------------------------------------- d1 (+a-c)
@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("prec,tol", [(1, 1e-7), (0, 1e-3)])
def test_interval_reference_fixture(golden, prec, tol, path):
    d = golden("ref_tube_chk.npz")
    sysm = fixture_system(d, "f64")
    g = ra.REMDGPU(sysm, precision=prec, path=path)
    g.execute_steps(sysm.T, sysm.P)
    s = g.state(0)
    ref = d["f64_int_state"]
    ids = d["f64_int_ids"]
    for name, sl in (("r", slice(0, 3)), ("v", slice(3, 6)), ("q", slice(6, 10)), ("p", slice(10, 14))):
        a = by_id(s[name], s["id"])
        b = by_id(ref[:, sl], ids)
        if name == "r":  # fractional, compare the periodic difference
            dd = a - b
            dd -= np.rint(dd)
            err = float(np.max(np.abs(dd)))
        else:
            err = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
        print(f"prec={prec} {name} max rel err {err:.3e}")
        assert err <= tol
    pr = _props(d, "f64_int_props")
    av = g.averages(0)
    for k, idx in (("lj", 0), ("clmb", 1), ("wall", 2), ("tot", 11)):
        print(k, av[idx], pr[k])
    assert abs(av[11] - pr["tot"]) <= 1e3 * tol * abs(pr["tot"]) + 1e-6
    g.close()


def _barostat_case(conf):
    from helpers import relax_oracle

    if conf == "tube":
        s, pmode = inp.config_tube(seed=31, nmol=96, Lz=36.0, ntemp=2, npress=1, nfwall=500), 2
    elif conf == "slit":
        s, pmode = inp.config_slit(seed=32, nmol=100, Lxy=24.0, ntemp=2, npress=1, model="TIP5P"), 3
    elif conf == "bulk":
        s, pmode = inp.config_bulk(seed=33, ntemp=2, npress=1, model="TIP3P"), 1
    else:  # "bulk-np": thermostat only
        s, pmode = inp.config_bulk(seed=34, ntemp=2, npress=1, model="TIP4P"), 0
    s = relax_oracle(s, steps=100, dt_fs=0.5)
    s.tconstant, s.pconstant, s.dt_fs, s.interval = 1, pmode, 0.5, 20
    s.P = np.full(s.nreplica, 50e6 / inp.UNIT_PRESS)
    return s


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("conf", ["tube", "slit", "bulk", "bulk-np"])
@pytest.mark.parametrize("prec,tol", [(1, 1e-9), (0, 1e-4)])
def test_interval_vs_oracle_all_barostats(conf, prec, tol, path):
    """One 20-step interval (0.5 fs) of the NP thermostat with the 1D (tube,
    pconstant 2), 2D area (TIP5P slit, pconstant 3) and 3D isotropic (TIP3P
    bulk, 256 molecules, pconstant 1) Andersen barostats, and with no
    barostat, against the FP64 oracle at 50 MPa: positions (periodic difference of fractions),
    velocities, quaternions, angular momenta, s and the box to `tol`
    relative; Ps and Pv to `tol` of max(|value|, 1e-3).  The systems are
    relaxed on the oracle first (100 steps at 0.5 fs, velocity scaling).
    The oracle's own FP32 build departs from its FP64 build by about 1e-6
    on these runs, so 1e-4 leaves room only for real differences."""
    s = _barostat_case(conf)
    g = ra.REMDGPU(s, precision=prec, path=path)
    g.execute_steps(s.T, s.P)
    for rep in range(s.nreplica):
        o = oracle_for(s, rep)
        o.initialize_property()
        o.execute_steps(s.interval)
        so, sg = o.state(), g.state(rep)
        sc = o.scalars()
        dd = by_id(sg["r"], sg["id"]) - by_id(so["r"], so["id"])
        dd -= np.rint(dd)
        errs = {"r": float(np.max(np.abs(dd)))}
        for name in ("v", "q", "p"):
            errs[name] = rel_rms(by_id(sg[name], sg["id"]), by_id(so[name], so["id"]))
        gs = sg["scal"]
        errs["s"] = abs(gs[0] - sc["s"]) / abs(sc["s"])
        errs["L"] = float(np.max(np.abs(gs[7:10] - np.array(sc["L"])) / np.abs(np.array(sc["L"]))))
        errs["Ps"] = abs(gs[1] - sc["Ps"]) / max(abs(sc["Ps"]), 1e-3)
        errs["Pv"] = float(np.max(np.abs(gs[3:6] - np.array(sc["Pv"])))) / max(np.max(np.abs(sc["Pv"])), 1e-3)
        print(conf, prec, rep, {k: f"{e:.1e}" for k, e in errs.items()}, "L", sc["L"], "Pv", sc["Pv"])
        for k, e in errs.items():
            assert e <= tol, (k, e)
        if s.pconstant != 0:
            assert not np.allclose(sc["L"], s.states[rep].L, rtol=1e-12, atol=0.0), "the box did not move"
    g.close()


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("conf,tmode", [("tube", 1), ("slit", 1), ("bulk", 1), ("tube", 2), ("bulk-np", 2)])
def test_exchanged_conditions_vs_oracle(conf, tmode, path):
    """Two intervals with new conditions between them, as after an accepted
    exchange (src/remd.cxx:526-536: ChangeTemperature, KillMomentum,
    SetPressure, CalcProperties, UpdateHamiltonian): replica 0 gets only a
    new pressure (ChangeTemperature at the same T still folds s into the
    momenta and resets s to 1), replica 1 only a new temperature.  With
    tconstant 1 (Nose-Poincare) and 2 (velocity scaling once per interval
    in the driver, src/remd.cxx:259-262, and at every step inside
    ExecuteSteps), FP64 pair math against the FP64 oracle to 1e-9."""
    s = _barostat_case(conf)
    s.tconstant = tmode
    T1 = s.T.copy()
    P1 = s.P.copy()
    T2 = np.array([T1[0], T1[0]])
    P2 = np.array([2.0 * P1[0], P1[1]])
    g = ra.REMDGPU(s, precision=1, path=path)
    g.execute_steps(T1, P1)
    g.execute_steps(T2, P2)
    for rep in range(s.nreplica):
        o = oracle_for(s, rep)
        o.initialize_property()
        for T_, P_ in ((None, None), (T2[rep], P2[rep])):
            if T_ is not None:
                o.change_temperature(float(T_))
                o.kill_momentum()
                o.sys.P = float(P_)
                o.calc_properties()
                o.update_hamiltonian()
            if tmode == 2:  # once per interval in the driver, src/remd.cxx:259-262
                o.velocity_scaling()
                o.ang_velocity_scaling()
            o.execute_steps(s.interval)
        so, sg = o.state(), g.state(rep)
        sc = o.scalars()
        dd = by_id(sg["r"], sg["id"]) - by_id(so["r"], so["id"])
        dd -= np.rint(dd)
        errs = {"r": float(np.max(np.abs(dd)))}
        for name in ("v", "p"):
            errs[name] = rel_rms(by_id(sg[name], sg["id"]), by_id(so[name], so["id"]))
        gs = sg["scal"]
        errs["s"] = abs(gs[0] - sc["s"]) / abs(sc["s"])
        errs["L"] = float(np.max(np.abs(gs[7:10] - np.array(sc["L"])) / np.abs(np.array(sc["L"]))))
        av = g.averages(rep)
        oa = o.average()
        errs["tot"] = abs(av[11] - oa[11]) / max(abs(oa[11]), 1.0)
        errs["ham"] = abs(av[12] - oa[12]) / max(abs(oa[11]), 1.0)
        print(conf, tmode, rep, {k: f"{e:.1e}" for k, e in errs.items()}, "T", av[6], "s", sc["s"])
        for k, e in errs.items():
            assert e <= 1e-9, (k, e)
    g.close()


@pytest.mark.parametrize("model", ["TIP4P", "TIP5P"])
def test_write_cdv_snapshot(tmp_path, model):
    """remd_gpu_write_cdv (REMDGPU::OutputCDV, src/remdgpu.cu:700-804): one
    file per replica, <prefix>R<rep:04d>s<step:07d>.cdv, a header line with
    the box and the replica's T (K) and P (MPa), then `mol site x y z` (A)
    for every site.  Checked against the downloaded state with invariants
    that do not depend on the rotation convention: the mass-weighted site
    centre is the molecule's centre of mass, and every intra-molecular site
    distance is the body-frame one (to the 6 significant digits of %g).
    The system is equilibrated first so that it stays inside the box."""
    from remd_amd.workloads import relax

    s = relax(inp.config_tube(seed=12, nmol=60, Lz=30.0, ntemp=2, npress=1, nfwall=500, model=model))
    g = ra.REMDGPU(s, precision=1)
    g.execute_steps(s.T, s.P)
    g.write_cdv(str(tmp_path / "snap"), 1234)
    mt = inp.molecule(model)
    ns = mt.nsite
    for rep in range(s.nreplica):
        lines = (tmp_path / f"snapR{rep:04d}s{1234:07d}.cdv").read_text().splitlines()
        head = lines[0]
        R = s.wall_length
        st = g.state(rep)
        L = st["scal"][7:10]
        assert head.startswith("'")
        assert f"box_sx={-R:g} box_sy={-R:g} box_sz={-0.5 * L[2]:g}" in head
        assert f"T={s.T[rep] * inp.UNIT_TEMP:g} K" in head
        assert f"P={s.P[rep] * inp.UNIT_PRESS / 1e6:g} MPa" in head
        rows = np.array([[float(x) for x in ln.split()] for ln in lines[1:]])
        assert rows.shape == (s.nmol * ns, 5)
        assert np.array_equal(rows[:, 0], np.repeat(np.arange(s.nmol), ns))
        assert np.array_equal(rows[:, 1], np.tile(np.arange(ns), s.nmol))
        x = rows[:, 2:].reshape(s.nmol, ns, 3)
        tol = 4 * 5e-6 * max(1.0, float(np.max(np.abs(x))))  # 6 significant digits
        com = (st["r"] - 0.5) * L
        centre = np.einsum("a,nad->nd", mt.m, x) / mt.m.sum()
        assert np.max(np.abs(centre - com)) < tol
        for a in range(ns):
            for b in range(a + 1, ns):
                d = np.linalg.norm(x[:, a] - x[:, b], axis=1)
                assert np.max(np.abs(d - np.linalg.norm(mt.r[a] - mt.r[b]))) < tol, (a, b, tol)
    g.close()


# ------------------------------------------------------------------ e1
@pytest.mark.parametrize("nt,npr", [(4, 2), (5, 3), (3, 1), (2, 5), (1, 4)])
def test_exchange_device_matches_oracle(nt, npr):
    """The device exchange (one thread per pair of the list) makes the
    oracle's serial decisions on grids of even and odd sides."""
    sysm = inp.config_tube(seed=21, nmol=64, Lz=30.0, ntemp=nt, npress=npr, nfwall=500)
    g = ra.REMDGPU(sysm, precision=0)
    pot, vol = g.pot_vol()
    rng = np.random.RandomState(5)
    n = nt * npr
    index = np.arange(n, dtype=np.int32)
    T = sysm.T.copy()
    P = sysm.P.copy()
    for step in range(8):
        u = rng.random_sample(n)
        i2, T2, P2, acc2, _ = po.replica_exchange(nt, npr, step, 2, pot, vol, index, T, P, u)
        i1, T1, P1, acc1 = g.exchange_device(nt, npr, step, index, T, P, u)
        assert np.array_equal(i1, i2) and np.array_equal(T1, T2) and np.array_equal(P1, P2)
        assert np.array_equal(acc1, acc2[: len(acc1)])
        index, T, P = i1, T1, P1
    g.close()


def test_execute_intervals_equals_host_exchange_loop():
    """remd_gpu_execute_intervals (intervals back to back, the exchange on
    the device with the uniforms drawn beforehand from the job's rand()
    stream) gives what the host loop gives: execute_steps, pot/vol,
    remd_exchange_host, interval after interval -- the same index, T, P and
    accept flags after every interval, the same per-interval averages and the
    same final state, bit for bit; and 4 + 4 intervals in two calls equal 8
    in one."""
    from remd_amd.workloads import relax

    s = relax(inp.config_tube(seed=77, ntemp=4, npress=2))
    s.interval = 20
    # a narrow ladder, so that swaps are accepted
    T0 = s.T.min() + (s.T - s.T.min()) * 0.05
    P0 = s.P.copy()
    K, nt, npr = 8, 4, 2

    g = ra.REMDGPU(s)
    rng = ra.GlibcRand(1)
    index = np.arange(s.nreplica, dtype=np.int32)
    T, P = T0.copy(), P0.copy()
    host_acc, host_av, host_cond = [], [], []
    for k in range(K):
        g.execute_steps(T, P)
        host_av.append(g.all_averages())
        pot, vol = g.pot_vol()
        _, acc, _ = ra.exchange_host(nt, npr, k, s.pconstant, pot, vol, index, T, P, rng)
        host_acc.append(acc.copy())
        host_cond.append((index.copy(), T.copy(), P.copy()))
    host_state = [g.state(r) for r in range(s.nreplica)]
    host_rng = rng.getstate()
    g.close()
    assert sum(int(a.sum()) for a in host_acc) > 0

    for split in ((K,), (4, 4)):
        g = ra.REMDGPU(s)
        rng = ra.GlibcRand(1)
        index = np.arange(s.nreplica, dtype=np.int32)
        T, P = T0.copy(), P0.copy()
        acc_all, av_all, step0 = [], [], 0
        for n in split:
            nacc, acc, av, ms = ra.execute_intervals(g, nt, npr, step0, n, index, T, P, rng)
            assert nacc == int(acc.sum()) and ms.shape == (n,) and np.all(ms > 0)
            acc_all.append(acc)
            av_all.extend(list(av))
            step0 += n
        assert np.array_equal(np.concatenate(acc_all), np.concatenate(host_acc)), split
        hi, hT, hP = host_cond[-1]
        assert np.array_equal(index, hi) and np.array_equal(T, hT) and np.array_equal(P, hP), split
        for k in range(K):
            assert np.array_equal(av_all[k], host_av[k]), (split, k)
        for r in range(s.nreplica):
            st = g.state(r)
            for name in ("r", "v", "q", "p", "id"):
                assert np.array_equal(st[name], host_state[r][name]), (split, r, name)
        assert rng.getstate() == host_rng, split  # as many rand() draws as the host loop
        g.close()


# ------------------------------------------------------------- paths
def _path_case(conf):
    if conf == "tube":
        return inp.config_tube(seed=2, ntemp=2, npress=2)
    from remd_amd.workloads import relax

    if conf == "slit":  # partners screened by COM distance on both paths
        return relax(inp.config_slit(seed=13, nmol=200, ntemp=2, npress=2, Lxy=30.0))
    return relax(inp.config_bulk(seed=11, ntemp=4))


@pytest.mark.parametrize("conf", ["tube", "slit", "bulk"])
@pytest.mark.parametrize("prec,tol", [(1, 1e-11), (0, 1e-5)])
def test_resident_path_agrees_with_global_path(prec, tol, conf):
    """The paths (resident, global-state, split-global) reduce in different
    orders; 20 steps of 4 replicas (the bench tube, a 200-TIP5P slit, the
    256-TIP4P bulk cube) agree to the stated tolerance."""
    s = _path_case(conf)
    s.interval = 20
    out = []
    for path in PATHS:
        g = ra.REMDGPU(s, precision=prec, path=path)
        g.execute_steps(s.T, s.P)
        out.append((g.all_averages(), [g.state(k) for k in range(4)]))
        g.close()
    a0, s0 = out[0]
    for a1, s1 in out[1:]:
        for k in range(4):
            assert np.array_equal(s0[k]["id"], s1[k]["id"])
            dd = s0[k]["r"] - s1[k]["r"]
            dd -= np.rint(dd)
            assert np.max(np.abs(dd)) <= tol
            assert rel_rms(s0[k]["p"], s1[k]["p"]) <= 1e3 * tol
        assert np.allclose(a0[:, :3], a1[:, :3], rtol=1e3 * tol, atol=1e-6)


def test_large_replica_takes_split_path():
    """700 TIP4P (a slit, 16 replicas) do not fit the LDS: AUTO takes the
    split-global path; its step-0 forces match the oracle and 20 steps agree
    with the global-state path."""
    from remd_amd.workloads import relax

    s = relax(inp.config_slit(seed=61, nmol=700, Lxy=55.0, ntemp=4, npress=4, model="TIP4P"))
    s.interval = 20
    g = ra.REMDGPU(s)
    path, parts = g.launch_path()
    print("path", path, "parts", parts)
    assert path == ra.PATH_SPLIT
    sg = g.state(0)
    o = oracle_for(s, 0)
    o.initialize_property()
    so = o.state()
    assert np.array_equal(sg["id"], so["id"])
    assert rel_rms(sg["f"], so["f"]) <= 1e-5 and rel_rms(sg["n"], so["n"]) <= 1e-5
    g.execute_steps(s.T, s.P)
    assert g.launch_path()[1] == 8
    h = ra.REMDGPU(s, path=ra.PATH_GLOBAL)
    h.execute_steps(s.T, s.P)
    for k in (0, s.nreplica - 1):
        a, b = g.state(k), h.state(k)
        assert np.array_equal(a["id"], b["id"])
        dd = a["r"] - b["r"]
        dd -= np.rint(dd)
        assert np.max(np.abs(dd)) <= 1e-5
    assert np.allclose(g.all_averages()[:, :3], h.all_averages()[:, :3], rtol=1e-2, atol=1e-6)
    g.close()
    h.close()


@pytest.mark.parametrize("conf", ["tube", "split"])
def test_resident_radix_sort_equals_rank_sort(conf, monkeypatch):
    """The resident path's neighbour sort gives the same order with the
    wave-ballot radix sort as with the rank sort (both stable sorts of the
    same keys), so two intervals give the same bits in state, molecule ids,
    windows and averages: a 144-TIP4P tube (state in LDS) and a 700-TIP4P
    slit (state in per-part global copies)."""
    from remd_amd.workloads import relax

    if conf == "tube":
        s = inp.config_tube(seed=71, ntemp=2, npress=2)
        s = relax(s)
    else:
        s = relax(inp.config_slit(seed=61, nmol=700, Lxy=55.0, ntemp=2, npress=2, model="TIP4P"))
    s.interval = 30
    out = []
    for radix_min in (1, 10 ** 6):
        monkeypatch.setenv("REMD_SORT_RADIX_MIN", str(radix_min))
        g = ra.REMDGPU(s)
        g.set_debug(True)
        for _ in range(2):
            g.execute_steps(s.T, s.P)
        out.append((g.launch_path()[0], g.all_averages(), [g.state(k) for k in range(s.nreplica)], g.window(0)))
        g.close()
    (pa, aa, sa, wa), (pb, ab, sb, wb) = out
    assert pa == pb == (ra.PATH_RESIDENT if conf == "tube" else ra.PATH_SPLIT)
    assert np.all(np.isfinite(aa[:, :15]))
    assert np.array_equal(aa, ab)
    assert np.array_equal(wa, wb)
    for x, y in zip(sa, sb):
        for k in ("r", "v", "q", "p", "f", "n", "id"):
            assert np.array_equal(x[k], y[k]), k


@pytest.mark.parametrize("path", [ra.PATH_RESIDENT, ra.PATH_SPLIT])
def test_parts_do_not_change_results(path):
    """A replica split over 7 workgroups (alone) gives the same bits as in a
    batch with a 254-replica filler job, over three intervals run with the
    batch's plan fixed to 1 and to 3 workgroups per replica, then chosen by
    the autotuner (state in LDS, or in per-workgroup global copies)."""
    a = inp.config_tube(seed=31, nmol=60, Lz=26.0, ntemp=2, npress=1, nfwall=500)
    a.interval = 20
    a.dt_fs = 0.5  # unrelaxed random placement: short steps keep it finite
    g = ra.REMDGPU(a, path=path)
    assert g.launch_path()[0] == path
    for _ in range(3):
        g.execute_steps(a.T, a.P)
    parts_alone = g.launch_path()[1]
    ref = (g.all_averages(), g.state(0)["r"], g.state(1)["p"])
    assert np.all(np.isfinite(ref[0][:, :15]))
    g.close()
    filler = inp.config_tube(seed=33, nmol=16, Lz=22.0, ntemp=127, npress=2, nfwall=300)
    filler.interval = 20
    ga, gf = ra.REMDGPU(a, path=path), ra.REMDGPU(filler)
    bt = ra.Batch()
    bt.add(ga)
    bt.add(gf)
    seen = []
    for plan in ([1, 1], [3, 1], None):
        bt.set_parts(plan)
        ga.set_conditions(a.T, a.P)
        gf.set_conditions(filler.T, filler.P)
        bt.execute()
        seen.append(bt.parts())
    print("parts alone", parts_alone, "batched", seen)
    assert parts_alone > 1 and seen[0] == [1, 1] and seen[1] == [3, 1]
    assert np.array_equal(ga.all_averages(), ref[0])
    assert np.array_equal(ga.state(0)["r"], ref[1])
    assert np.array_equal(ga.state(1)["p"], ref[2])
    with pytest.raises(RuntimeError):
        bt.set_parts([9, 1])
    bt.close()
    ga.close()
    gf.close()


# ------------------------------------------------------------- batching
def test_mixed_path_batch_equals_alone():
    """A resident job and a global-state job batched together (two kernel
    classes in one interval) give the bits each gives alone."""
    a = inp.config_tube(seed=31, nmol=60, Lz=26.0, ntemp=2, npress=1, nfwall=500)
    b = inp.config_tube(seed=32, nmol=48, Lz=24.0, ntemp=2, npress=1, model="TIP3P", nfwall=500)
    for s in (a, b):
        s.interval = 20
    alone = []
    for s, path in ((a, ra.PATH_RESIDENT), (b, ra.PATH_GLOBAL)):
        g = ra.REMDGPU(s, path=path)
        g.execute_steps(s.T, s.P)
        alone.append(g.all_averages())
        g.close()
    ga, gb = ra.REMDGPU(a, path=ra.PATH_RESIDENT), ra.REMDGPU(b, path=ra.PATH_GLOBAL)
    bt = ra.Batch()
    bt.add(ga)
    bt.add(gb)
    ga.set_conditions(a.T, a.P)
    gb.set_conditions(b.T, b.P)
    bt.execute()
    assert np.array_equal(ga.all_averages(), alone[0])
    assert np.array_equal(gb.all_averages(), alone[1])
    bt.close()
    ga.close()
    gb.close()


def test_mixed_precision_batch_equals_alone():
    """FP32-pair and FP64-pair jobs of three models in one batch (the
    packed launch's single- and double-precision variants side by side,
    resident and split-global) give, state and averages, the bits each job
    gives alone, over two intervals."""
    from remd_amd.workloads import relax

    specs = [("TIP4P", 0, ra.PATH_RESIDENT), ("TIP3P", 1, ra.PATH_RESIDENT), ("TIP5P", 1, ra.PATH_SPLIT),
             ("TIP4P", 1, ra.PATH_RESIDENT)]
    systems = []
    for k, (model, prec, path) in enumerate(specs):
        s = relax(inp.config_tube(seed=41 + k, nmol=56 + 8 * k, Lz=24.0 + 2 * k, ntemp=2, npress=1,
                                  nfwall=500, model=model))
        s.interval = 20
        systems.append((s, prec, path))
    alone = []
    for s, prec, path in systems:
        g = ra.REMDGPU(s, precision=prec, path=path)
        g.execute_steps(s.T, s.P)
        g.execute_steps(s.T[::-1].copy(), s.P)
        alone.append((g.all_averages(), g.state(1)))
        g.close()
    jobs = [ra.REMDGPU(s, precision=prec, path=path) for s, prec, path in systems]
    bt = ra.Batch()
    for g in jobs:
        bt.add(g)
    for T_of in (lambda s: s.T, lambda s: s.T[::-1].copy()):
        for g, (s, _, _) in zip(jobs, systems):
            g.set_conditions(T_of(s), s.P)
        bt.execute()
    for g, (av, st) in zip(jobs, alone):
        assert np.all(np.isfinite(av[:, :15]))
        assert np.array_equal(g.all_averages(), av)
        s1 = g.state(1)
        for name in ("r", "v", "q", "p", "id"):
            assert np.array_equal(s1[name], st[name]), name
    bt.close()
    for g in jobs:
        g.close()


@pytest.mark.parametrize("model,nmol", [("TIP3P", 64), ("TIP4P", 56), ("TIP5P", 64)])
def test_batched_job_with_new_conditions_equals_alone(model, nmol):
    """A batched FP32-pair job whose temperatures are swapped before its
    second interval (the exchange rescale runs inside the launch) gives
    the bits of the same job alone, for 1, 2 and 4 workgroups per replica
    and the autotuned plan."""
    from remd_amd.workloads import relax

    s = relax(inp.config_tube(seed=42, nmol=nmol, Lz=26.0, ntemp=2, npress=1, nfwall=500, model=model))
    s.interval = 20
    conds = (s.T.copy(), s.T[::-1].copy())
    g = ra.REMDGPU(s, precision=0, path=ra.PATH_RESIDENT)
    for T in conds:
        g.execute_steps(T, s.P)
    ref = g.all_averages()
    ref_states = [g.state(r) for r in range(s.nreplica)]
    assert np.all(np.isfinite(ref[:, :15]))
    g.close()
    for parts in (None, [1], [2], [4]):
        g = ra.REMDGPU(s, precision=0, path=ra.PATH_RESIDENT)
        bt = ra.Batch()
        bt.add(g)
        if parts:
            bt.set_parts(parts)
        for T in conds:
            g.set_conditions(T, s.P)
            bt.execute()
        assert np.array_equal(g.all_averages(), ref), (model, parts, bt.parts())
        for r in range(s.nreplica):
            st = g.state(r)
            for name in ("r", "v", "q", "p", "id"):
                assert np.array_equal(st[name], ref_states[r][name]), (model, parts, r, name)
        bt.close()
        g.close()


@pytest.mark.parametrize("conf", ["tube", "slit", "bulk"])
def test_fixed_ensemble_equals_runtime_modes(conf, monkeypatch):
    """The resident kernels built with the job's thermostat, barostat and
    confinement modes fixed give the bits of the instantiation that reads
    them from the job (REMD_ENSEMBLE_FIXED=0), alone and in a batch: a tube
    (NP, 1D barostat), a slit (NP, area barostat) and a bulk box (NP, 3D
    barostat), two intervals with the temperatures swapped before the
    second (the exchange rescale runs inside the launch)."""
    from remd_amd.workloads import relax

    if conf == "tube":
        s = inp.config_tube(seed=81, nmol=96, Lz=34.0, ntemp=2, npress=2, nfwall=500)
    elif conf == "slit":
        s = inp.config_slit(seed=82, nmol=120, ntemp=2, npress=2, model="TIP5P", Lxy=30.0)
    else:
        s = inp.config_bulk(seed=83, nmol=128, ntemp=4, model="TIP3P")
        s.pconstant = 1
    s = relax(s)
    s.interval = 20
    conds = (s.T.copy(), s.T[::-1].copy())
    runs = []
    for fixed in ("1", "0"):
        monkeypatch.setenv("REMD_ENSEMBLE_FIXED", fixed)
        g = ra.REMDGPU(s, precision=0)
        assert g.launch_path()[0] == ra.PATH_RESIDENT
        for T in conds:
            g.execute_steps(T, s.P)
        runs.append((g.all_averages(), [g.state(r) for r in range(s.nreplica)]))
        g.close()
        g = ra.REMDGPU(s, precision=0)
        bt = ra.Batch()
        bt.add(g)
        for T in conds:
            g.set_conditions(T, s.P)
            bt.execute()
        runs.append((g.all_averages(), [g.state(r) for r in range(s.nreplica)]))
        bt.close()
        g.close()
    ref_av, ref_st = runs[0]
    assert np.all(np.isfinite(ref_av[:, :15]))
    for k, (av, st) in enumerate(runs[1:], 1):
        assert np.array_equal(av, ref_av), (conf, k)
        for r in range(s.nreplica):
            for name in ("r", "v", "q", "p", "f", "n", "id"):
                assert np.array_equal(st[r][name], ref_st[r][name]), (conf, k, r, name)


def test_uneven_autotuned_batch_equals_alone():
    """Four jobs of different sizes, models and confinements in one batch
    for six intervals -- the autotuner times different packing plans on
    the first ones -- give the bits each gives alone; one job has FP64
    pair math."""
    from remd_amd.workloads import relax

    jobs = [inp.config_tube(seed=51, nmol=40, Lz=22.0, ntemp=4, npress=1, model="TIP3P", nfwall=400),
            inp.config_tube(seed=52, nmol=90, Lz=30.0, ntemp=4, npress=2, nfwall=400),
            inp.config_slit(seed=53, nmol=120, ntemp=2, npress=2, model="TIP5P", Lxy=30.0),
            inp.config_bulk(seed=54, nmol=64, ntemp=2, model="TIP4P")]
    precs = [0, 1, 0, 0]
    jobs = [relax(s) for s in jobs]  # random placements, equilibrated at 0.5 fs
    for s in jobs:
        s.interval = 20
    alone = []
    for s, pr in zip(jobs, precs):
        g = ra.REMDGPU(s, precision=pr)
        for _ in range(6):
            g.execute_steps(s.T, s.P)
        alone.append((g.all_averages(), g.state(0)["r"], g.state(s.nreplica - 1)["p"]))
        assert np.all(np.isfinite(alone[-1][0][:, :15]))
        g.close()
    gs = [ra.REMDGPU(s, precision=pr) for s, pr in zip(jobs, precs)]
    bt = ra.Batch()
    for g in gs:
        bt.add(g)
    plans = []
    for _ in range(6):
        for g, s in zip(gs, jobs):
            g.set_conditions(s.T, s.P)
        bt.execute()
        plans.append(tuple(bt.parts()))
    print("plans", plans)
    for g, s, ref in zip(gs, jobs, alone):
        assert np.array_equal(g.all_averages(), ref[0])
        assert np.array_equal(g.state(0)["r"], ref[1])
        assert np.array_equal(g.state(s.nreplica - 1)["p"], ref[2])
    bt.close()
    for g in gs:
        g.close()


@pytest.mark.parametrize("path", PATHS)
def test_batch_equals_alone_and_plain(path):
    a = inp.config_tube(seed=31, nmol=60, Lz=26.0, ntemp=2, npress=1, nfwall=500)
    b = inp.config_tube(seed=32, nmol=48, Lz=24.0, ntemp=2, npress=1, model="TIP3P", nfwall=500)
    for s in (a, b):
        s.interval = 20
    alone = {}
    for name, s in (("a", a), ("b", b)):
        g = ra.REMDGPU(s, path=path)
        g.execute_steps(s.T, s.P)
        alone[name] = (g.all_averages(), g.state(0)["r"])
        g.close()
    # plain (one replica per launch) gives the same bits
    g = ra.REMDGPU(a, path=path)
    g.execute_steps(a.T, a.P, plain=True)
    assert np.array_equal(g.all_averages(), alone["a"][0])
    g.close()
    ga, gb = ra.REMDGPU(a, path=path), ra.REMDGPU(b, path=path)
    bt = ra.Batch()
    bt.add(ga)
    bt.add(gb)
    ga.set_conditions(a.T, a.P)
    gb.set_conditions(b.T, b.P)
    bt.execute()
    assert np.array_equal(ga.all_averages(), alone["a"][0])
    assert np.array_equal(gb.all_averages(), alone["b"][0])
    assert np.array_equal(ga.state(0)["r"], alone["a"][1])
    bt.close()
    ga.close()
    gb.close()
