"""Benchmark workloads built and equilibrated on the device.

Random placement with random orientations leaves close site contacts that a
2 fs step cannot integrate; every generated system is first relaxed on the
GPU with velocity rescaling (tconstant = 2) at a 0.5 fs step and no
barostat, then handed to the production ensemble.
"""
from __future__ import annotations

import copy
import math

import numpy as np

from . import REMDGPU, inputs


def relax(system: "inputs.System", steps: int = 300, dt_fs: float = 0.5) -> "inputs.System":
    """Equilibrate every replica of `system` in one launch; returns a copy
    whose states are the relaxed ones (s = 1, Ps = 0, Pv = 0)."""
    rs = copy.copy(system)
    rs.tconstant, rs.pconstant = 2, 0
    rs.dt_fs = dt_fs
    rs.interval = steps
    g = REMDGPU(rs)
    g.execute_steps(rs.T, rs.P)
    out = copy.copy(system)
    out.states = []
    for k, st in enumerate(system.states):
        s = g.state(k)
        sc = s["scal"]
        out.states.append(inputs.ReplicaState(r=s["r"], v=s["v"], q=s["q"], p=s["p"], s=1.0, Ps=0.0, Q=st.Q,
                                              Pv=(0.0, 0.0, 0.0), W=st.W, L=(sc[7], sc[8], sc[9]),
                                              ids=s["id"].astype(np.int32)))
    g.close()
    return out


def count_pairs(r_frac, L, rcut, periodic):
    """Ordered molecule pairs (i != j) with COM distance < rcut, brute
    force with the minimum image along periodic axes."""
    x = (np.asarray(r_frac) - 0.5) * np.asarray(L)[None, :]
    n = x.shape[0]
    total = 0
    for a in range(0, n, 512):
        d = x[a:a + 512, None, :] - x[None, :, :]
        for k in range(3):
            if periodic[k]:
                d[..., k] -= L[k] * np.rint(d[..., k] / L[k])
        r2 = np.sum(d * d, axis=-1)
        total += int(np.sum(r2 < rcut * rcut))
    return total - n  # self pairs have r = 0


PERIODIC = {0: (True, True, True), 1: (False, False, True), 2: (True, True, False)}

# Algorithmic FLOPs per ordered in-cutoff molecule pair (1 LJ site term 32,
# Coulomb site term 27, per-pair overhead 38; FMA = 2), counted from the
# reference's tuned pair code.
FLOPS_PER_PAIR = {3: 313, 4: 313, 5: 502}


def algorithmic_flops(job: REMDGPU, steps: int) -> float:
    """313 (or 502) x ordered in-cutoff pairs of every replica at the
    current snapshot x steps."""
    s = job.system
    per = PERIODIC[s.confined]
    tot = 0
    for k in range(job.nrep):
        st = job.state(k)
        L = st["scal"][7:10]
        tot += count_pairs(st["r"], L, s.rcut, per)
    return float(tot) * FLOPS_PER_PAIR[job.nsite] * steps


def ns_per_day(replica_steps_per_s: float, nreplica: int, dt_fs: float) -> float:
    """Simulated ns per day per replica."""
    return replica_steps_per_s / nreplica * dt_fs * 1e-6 * 86400.0


# ------------------------------------------------------------------ studies
def _progress(msg):
    import sys

    print(msg, file=sys.stderr, flush=True)


def _timed_intervals(run_one, n):
    """Wall time of n calls of run_one (each one exchange interval)."""
    import time

    lat = []
    t0 = time.perf_counter()
    for _ in range(n):
        t1 = time.perf_counter()
        run_one()
        lat.append(time.perf_counter() - t1)
    return time.perf_counter() - t0, lat


def uneven_batch_study(intervals=10, warmup=16):
    """Eight nanotube systems (R = 5..12 A, Lz = 40 A, nmol ~40-420,
    TIP3P/TIP4P alternating, 32 replicas each = 256 replicas): all jobs in
    one batch (one resident launch per interval, replicas split over
    workgroups by the batch's packing plan; the warm-up intervals include
    the autotuner's timed plans) against the same jobs one at a time
    (each job's replicas in one launch)."""
    import numpy as np

    from . import Batch, GlibcRand, exchange_host

    systems = []
    for s in inputs.config_uneven_batch():
        systems.append(relax(s))
        _progress(f"relaxed tube R={s.wall_length} nmol={s.nmol} {s.model}")
    jobs = [REMDGPU(s) for s in systems]
    conds = [(s.T.copy(), s.P.copy(), np.arange(s.nreplica, dtype=np.int32), GlibcRand(1)) for s in systems]
    steps = [0]

    def exchange(k):
        s, (T, P, index, rng) = systems[k], conds[k]
        pot, vol = jobs[k].pot_vol()
        exchange_host(s.ntemp, s.npress, steps[0], s.pconstant, pot, vol, index, T, P, rng)

    bt = Batch()
    for j in jobs:
        bt.add(j)

    def batched():
        for k, j in enumerate(jobs):
            j.set_conditions(conds[k][0], conds[k][1])
        bt.execute()
        for k in range(len(jobs)):
            exchange(k)
        steps[0] += 1

    def one_at_a_time():
        for k, j in enumerate(jobs):
            j.execute_steps(conds[k][0], conds[k][1])
            exchange(k)
        steps[0] += 1

    for _ in range(warmup):
        batched()
    wb, lat_b = _timed_intervals(batched, intervals)
    _progress(f"batched: {intervals} intervals in {wb:.3f} s")
    parts = bt.parts()
    kms = bt.last_kernel_ms()
    bt.close()
    for _ in range(warmup):
        one_at_a_time()
    ws, lat_s = _timed_intervals(one_at_a_time, intervals)
    _progress(f"one job at a time: {intervals} intervals in {ws:.3f} s")
    rs = sum(s.nreplica * s.interval for s in systems) * intervals
    out = dict(batched=rs / wb, one_job_at_a_time=rs / ws, replicas=sum(s.nreplica for s in systems),
               nmol=[s.nmol for s in systems], models=[s.model for s in systems],
               paths=[j.launch_path()[0] for j in jobs], parts_batched=parts, batched_kernel_ms=kms,
               interval_latency_ms_batched={"p50": float(np.percentile(lat_b, 50)) * 1e3,
                                            "p99": float(np.percentile(lat_b, 99)) * 1e3})
    for j in jobs:
        j.close()
    return out


def served_study(njobs=24, rate=20.0, intervals=50, seed=5, nmol_hi=2048, lane_bounds=None, split=True,
                 reqs=None):
    """The Poisson stream of REMD jobs (mean `rate` jobs/s, nmol log-uniform
    64-nmol_hi, TIP3P/4P/5P, bulk/tube/slit, 8-128 replicas, `intervals`
    exchange intervals each) served with continuous batching under a 200 GB
    budget, against serving the same jobs one at a time."""
    import time

    import numpy as np

    from .serve import GPUBackend, Request, Server, latency_percentiles, poisson_stream

    if reqs is None:
        reqs = poisson_stream(seed=seed, njobs=njobs, rate=rate, intervals=intervals, nmol_hi=nmol_hi)
    steps = sum(r.intervals * r.system.interval * r.system.nreplica for r in reqs)
    _progress(f"{len(reqs)} jobs generated and relaxed")
    be = GPUBackend(split=split)
    kw = {} if lane_bounds is None else {"lane_bounds": lane_bounds}
    t0 = time.perf_counter()
    res = Server(be, hbm_budget=200e9, **kw).run(reqs)
    wall = time.perf_counter() - t0
    _progress(f"served stream: {wall:.3f} s")
    be.close()
    p50, p99 = latency_percentiles(res)
    # one job at a time: the next job starts when the previous one ends
    t0 = time.perf_counter()
    lat = []
    for r in reqs:
        be1 = GPUBackend()
        one = Server(be1).run([Request(r.job_id, 0.0, r.system, r.intervals)])
        be1.close()
        lat.extend(one[r.job_id].latencies)
    wall1 = time.perf_counter() - t0
    _progress(f"one job at a time: {wall1:.3f} s")
    turn = [res[r.job_id].finished - r.arrival for r in reqs]
    return dict(jobs=len(reqs), replica_steps=steps, served=steps / wall, one_job_at_a_time=steps / wall1,
                interval_latency_ms={"p50": p50 * 1e3, "p99": p99 * 1e3},
                one_job_at_a_time_latency_ms={"p50": float(np.percentile(lat, 50)) * 1e3,
                                              "p99": float(np.percentile(lat, 99)) * 1e3},
                job_turnaround_s={"p50": float(np.percentile(turn, 50)), "p99": float(np.percentile(turn, 99))},
                max_nmol=max(r.system.nmol for r in reqs),
                models=sorted({r.system.model for r in reqs}))


def precision_study(intervals=100, nrep_drift=None):
    """Nano-slit (H = 10 A), 512 TIP5P, 40 x 40 A, area barostat, 32 T x 4 P
    = 128 replicas: FP32 pair math with FP64 accumulation against FP64 pair
    math on the same states -- step-0 force and energy errors -- and the
    Nose-Poincare Hamiltonian drift over intervals x 100 steps of each
    precision from the same start (no exchanges)."""
    import numpy as np

    s = relax(inputs.config_slit())
    _progress("slit relaxed")
    out = {}
    # step-0 errors on every replica
    g0, g1 = REMDGPU(s, precision=0), REMDGPU(s, precision=1)
    ferr, eerr = [], []
    for k in range(s.nreplica):
        a, b = g0.state(k), g1.state(k)
        f0, f1 = a["f"][np.argsort(a["id"])], b["f"][np.argsort(b["id"])]
        ferr.append(float(np.sqrt(np.sum((f0 - f1) ** 2) / np.sum(f1 ** 2))))
        av0, av1 = g0.averages(k), g1.averages(k)
        pot0, pot1 = av0[0] + av0[1] + av0[2], av1[0] + av1[1] + av1[2]
        eerr.append(abs(pot0 - pot1) / abs(pot1))
    out["force_rel_rms_err"] = {"max": max(ferr), "median": float(np.median(ferr))}
    out["energy_rel_err"] = {"max": max(eerr), "median": float(np.median(eerr))}
    out["path"] = g0.launch_path()[0]
    # drift: |ham| = |H - H0| over the run, per precision
    for name, g in (("mixed", g0), ("fp64", g1)):
        T, P = s.T.copy(), s.P.copy()
        worst = np.zeros(s.nreplica)
        import time

        t0 = time.perf_counter()
        for _ in range(intervals):
            g.execute_steps(T, P)
            av = g.all_averages()
            worst = np.maximum(worst, np.abs(av[:, 12]))
        wall = time.perf_counter() - t0
        _progress(f"{name}: {intervals} intervals in {wall:.3f} s")
        kin = g.all_averages()[:, 4] + g.all_averages()[:, 5]
        out[f"drift_{name}"] = {"steps": intervals * s.interval, "max_abs_ham": float(worst.max()),
                                "median_abs_ham": float(np.median(worst)),
                                "max_rel_to_kinetic": float(np.max(worst / np.abs(kin))),
                                "replica_steps_per_s": s.nreplica * s.interval * intervals / wall}
    g0.close()
    g1.close()
    out.update(nmol=s.nmol, replicas=s.nreplica, model=s.model,
               wall="slit: two 10-4 sheets (this build's slit wall), not the reference's 9-3 table; "
                    "errors are against the same kernels in FP64, not comparable with the reference's slit runs")
    return out
