#!/usr/bin/env python3
"""Benchmark: replica-steps/s of the REMD interval path on one MI355X.

Workload (the plain-vs-GPU drop-in configuration): nanotube (confined=1),
144 TIP4P per replica, R = 7 A, Lz = 50 A, 1D Andersen barostat along z,
Nose-Poincare thermostat, 16 T (200-360 K geometric) x 4 P (0.1-200 MPa
log) = 64 replicas, rcut 9 A, rswitch 8 A, dt 2 fs, 100 steps per exchange
interval.  Synthetic seeded inputs, equilibrated on the device first.

A "step" of this benchmark is one exchange interval of the whole job
(64 replicas x 100 MD steps, one kernel launch) followed by the Metropolis
exchange.  The timed steps run back to back in one
remd_gpu_execute_intervals call: the exchange runs on the device with the
uniforms the host exchange would draw from the job's rand() stream (the same
decisions), so no host round trip separates the intervals.  `value` =
replica-MD-steps per second over the timed steps, inputs resident in HBM.
`host_exchange` times the same number of steps with the host exchange and a
host round trip per interval (execute_steps, pot/vol, remd_exchange_host),
and gives the per-interval latency a caller that waits on each interval
sees.

    python bench.py [--steps K] [--warmup W] [--no-cpu] [--no-plain]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3  # MI355X dense FP32 vector peak
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_resident_kernel_tube144_tip4p_64rep.json")
VALU_FILE = os.path.join(ROOT, "profiles", "valu_resident_kernel_tube144_tip4p_64rep.json")
CPU_RATIO_FILE = os.path.join(ROOT, "profiles", "cpu_baseline_reference_vs_port.json")
KERNEL_SRC = os.path.join(ROOT, "remd_amd", "csrc", "remd_kernels.hip")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE
    / WRITE_SIZE runs, each scaled by the same counter on a copy of known
    size; None when no measurement of this kernel is on file."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as f:
        m = json.load(f)
    if m.get("kernel") != kernel:
        return None
    kib = 1024.0
    true_kib = m["copy_bytes"] / kib
    fetch = np.mean(m["fetch_size_kib_per_launch"]) / (m["copy_fetch_size_kib"] / true_kib)
    write = np.mean(m["write_size_kib_per_launch"]) / (m["copy_write_size_kib"] / true_kib)
    return float((fetch + write) * kib)


def measured_valu(kernel, kernel_ms):
    """Executed FP32 / FP64 VALU FLOP per launch of `kernel` from the
    committed rocprofv3 counter run (SQ_INSTS_VALU_FLOPS_FP32 / _FP64 count
    FLOP per wave instruction; x 64 lanes), as FLOP/s over this run's
    kernel time, with the VALU-busy fraction of that run; None when no
    measurement of this kernel is on file.  `stale` marks a counter run of
    another kernel source than the one built here."""
    if not os.path.exists(VALU_FILE):
        return None
    with open(VALU_FILE) as f:
        m = json.load(f)
    if m.get("kernel") != kernel:
        return None
    import hashlib

    with open(KERNEL_SRC, "rb") as f:
        src = hashlib.sha256(f.read()).hexdigest()[:16]
    fp32 = float(np.mean(m["flops_fp32_per_launch"]))
    fp64 = float(np.mean(m["flops_fp64_per_launch"]))
    return {"executed_fp32_tflops": round(fp32 / (kernel_ms * 1e-3) / 1e12, 4),
            "executed_fp64_tflops": round(fp64 / (kernel_ms * 1e-3) / 1e12, 4),
            "executed_fp32_flop_per_launch": fp32, "valu_busy": m["valu_busy"],
            "lds_bank_conflict_cycles_per_lds_inst": m["lds_bank_conflict_per_lds_inst"],
            "counters": os.path.relpath(VALU_FILE, ROOT), "stale": src != m.get("kernel_source_sha16")}


def _parallel(worker, args_of, system, T, P, budget_s):
    """One spawned process per host core this process may use (at most
    16), each on its share of the replicas, started together; returns
    (replica-steps/s, processes).  Every worker is joined, and any that has
    not exited is terminated, before this returns."""
    import multiprocessing as mp

    nproc = min(16, len(os.sched_getaffinity(0)), system.nreplica)
    ctx = mp.get_context("spawn")
    barrier = ctx.Barrier(nproc)
    out = ctx.Queue()
    procs = []
    try:
        for w in range(nproc):
            reps = []
            for k in range(w, system.nreplica, nproc):
                st = system.states[k]
                reps.append((st.r, st.v, st.q, st.p, st.ids, st.scal(), float(T[k]), float(P[k])))
            pr = ctx.Process(target=worker, args=args_of(reps, barrier, out))
            pr.start()
            procs.append(pr)
        res = [out.get(timeout=budget_s + 120) for _ in procs]
    finally:
        for pr in procs:
            pr.join(timeout=30)
        for pr in procs:
            if pr.is_alive():
                pr.terminate()
                pr.join(timeout=10)
        out.close()
        out.join_thread()
    steps = sum(r[0] for r in res)
    return steps / max(r[1] for r in res), nproc


def cpu_baseline(system, T, P, budget_s=12.0):
    """Time the oracle's FP32 restatement of the reference's CPU path (bit
    for bit the reference's makefile.gnu FP32 build on these inputs) on the
    same relaxed replicas: one thread, one interval per replica (the
    shipped configuration), then one process per host core (at most 16)
    on shares of the replicas, which is `value`.  The reference's own
    build against this restatement, both timed in one container on the same
    replicas, is CPU_RATIO_FILE."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    from remd_amd import inputs as inp

    nrep = system.nreplica
    steps = system.interval
    done = 0
    elapsed = 0.0
    mt_text = inp.MOLTYPE[system.model]
    fw = po.wall_table_tube(system.nfwall, system.wall_length, system.sgm_wall, system.eps_wall, system.rho_wall)
    desc = system.desc()
    mt = po.parse_moltype(mt_text)
    for k in range(nrep):
        st = system.states[k]
        o = po.OracleSystem(desc, mt, st.r, st.v, st.q, st.p, st.scal(), T[k], P[k], fp32=True, fwall=fw)
        o.initialize_property()
        t0 = time.perf_counter()
        o.execute_steps(steps)
        elapsed += time.perf_counter() - t0
        done += 1
        if elapsed > budget_s:
            break
    single = done * steps / elapsed
    value, cores = _parallel(po.port_parallel_worker,
                             lambda reps, barrier, out: (mt_text, desc, reps, fw, budget_s, barrier, out),
                             system, T, P, budget_s)
    sample = (f"{nrep} replicas, oracle FP32 restatement (-O2): one thread on {done} replicas x 1 interval "
              f"({steps} steps); then {cores} processes (one per core) on shares of the {nrep} replicas, "
              f"whole intervals round-robin for about {budget_s:.0f} s")
    out = dict(value=value, unit="replica-steps/s", cores=cores, kind="port", sample=sample,
               single_thread_value=single, seconds=round(elapsed, 3))
    if os.path.exists(CPU_RATIO_FILE):
        with open(CPU_RATIO_FILE) as f:
            out["reference_over_port_one_thread"] = json.load(f)["reference_over_port"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed exchange intervals")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-studies", action="store_true",
                    help="skip the uneven-batch, served-stream and precision figures of the headline line")
    ap.add_argument("--plain-steps", type=int, default=6)
    ap.add_argument("--workload", default="tube64", choices=["tube64", "uneven", "served", "precision"],
                    help="tube64: the headline line; the others print one line of their own study")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run 3 device-to-device copies of this many MiB (PMC counter calibration)")
    args = ap.parse_args()
    if args.copy_probe_mib > 0:
        import torch
        n = args.copy_probe_mib << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n, "copies": 3}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: replicas shard across GPUs as whole jobs/ladders with no "
              "per-step communication, so only --gpus 1 is run here.", file=sys.stderr)
        sys.exit(2)

    import remd_amd as ra
    from remd_amd import inputs, workloads

    if args.workload != "tube64":
        study = {"uneven": workloads.uneven_batch_study, "served": workloads.served_study,
                 "precision": workloads.precision_study}[args.workload]
        print(json.dumps({"workload": args.workload, "n_gpus": 1, "data": "synthetic", **study()}))
        return

    system = inputs.config_tube(seed=2)
    system = workloads.relax(system)
    nrep, interval = system.nreplica, system.interval
    T = system.T.copy()
    P = system.P.copy()
    job = ra.REMDGPU(system, precision=0)
    shape = job.launch_shape()

    index = np.arange(nrep, dtype=np.int32)
    rng = ra.GlibcRand(1)
    step_no = 0

    def one_interval(plain=False):
        nonlocal step_no
        job.execute_steps(T, P, plain=plain)
        pot, vol = job.pot_vol()
        ra.exchange_host(system.ntemp, system.npress, step_no, system.pconstant, pot, vol, index, T, P, rng)
        step_no += 1

    for _ in range(args.warmup):
        one_interval()
    # one untimed pipelined call of the timed length: its device buffers and
    # events are allocated here, once per job
    ra.execute_intervals(job, system.ntemp, system.npress, step_no, args.steps, index, T, P, rng)
    step_no += args.steps
    path, parts = job.launch_path()
    # the resident instantiation with the tube NP ensemble fixed (code 19)
    kname = ("resident_kernel<256,4,float,19>" if path == ra.PATH_RESIDENT else
             f"interval_kernel<{shape['block_threads']},4,float>")

    # algorithmic work: ordered in-cutoff pairs at the interval-start snapshot
    flops_per_launch = workloads.algorithmic_flops(job, interval)

    t_start = time.perf_counter()
    _, _, _, kms = ra.execute_intervals(job, system.ntemp, system.npress, step_no, args.steps, index, T, P, rng)
    wall = time.perf_counter() - t_start
    step_no += args.steps
    value = nrep * interval * args.steps / wall
    ms_per_step = wall / args.steps * 1e3
    kernel_ms = float(np.mean(kms))

    # the same steps with a host round trip per interval
    lat = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        one_interval()
        lat.append(time.perf_counter() - t0)
    host_wall = time.perf_counter() - t_start
    host_value = nrep * interval * args.steps / host_wall

    plain = None
    if not args.no_plain:
        t0 = time.perf_counter()
        for _ in range(args.plain_steps):
            one_interval(plain=True)
        pw = time.perf_counter() - t0
        plain = nrep * interval * args.plain_steps / pw

    # force relative error of the FP32-pair/FP64-accumulate kernels against
    # the same kernels in FP64, on the current state of replica 0
    snap = inputs.System(**{k: getattr(system, k) for k in (
        "model", "nmol", "tconstant", "pconstant", "confined", "dt_fs", "rcut", "rswitch", "interval", "ntemp",
        "npress", "T", "P", "wall_length", "nfwall", "sgm_wall", "eps_wall", "rho_wall")})
    snap.ntemp, snap.npress = 1, 1
    st = job.state(0)
    snap.T, snap.P = T[:1].copy(), P[:1].copy()
    snap.states = [inputs.ReplicaState(r=st["r"], v=st["v"], q=st["q"], p=st["p"], s=st["scal"][0],
                                       Ps=st["scal"][1], Q=st["scal"][2], Pv=tuple(st["scal"][3:6]),
                                       W=st["scal"][6], L=tuple(st["scal"][7:10]), ids=st["id"])]
    fm = ra.REMDGPU(snap, precision=0)
    fd = ra.REMDGPU(snap, precision=1)
    f_mixed, f_fp64 = fm.state(0)["f"], fd.state(0)["f"]
    force_err = float(np.sqrt(np.sum((f_mixed - f_fp64) ** 2) / np.sum(f_fp64 ** 2)))
    fm.close()
    fd.close()

    achieved = flops_per_launch / (kernel_ms * 1e-3) / 1e12
    out = {
        "metric": "replica-steps/s batched vs plain at <=1e-4 force rel err; p99 job latency",
        "value": round(value, 1),
        "unit": "replica-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair math, f64 accumulation and integrator",
        "data": "synthetic (seeded placement, device-equilibrated)",
        "config": {
            "workload": "nanotube TIP4P, 144 mol/replica, R 7 A, Lz 50 A, 16T x 4P = 64 replicas, "
                        "rcut 9 A, dt 2 fs, 100 steps/interval, NP thermostat + 1D barostat; COMs placed "
                        "at r < 4.2 A (R - 2.8: the wall energy is about +500 kJ/mol at R - 2 A), min O-O "
                        "2.6 A, relaxed 300 x 0.5 fs on the device",
            "replicas": nrep, "nmol": system.nmol, "steps_per_interval": interval,
            "launch": {"path": "resident" if path == ra.PATH_RESIDENT else "global",
                       "workgroups_per_replica": parts, "block_threads": shape["block_threads"],
                       "chunks": shape["chunks"], "lds_bytes": shape["lds_bytes"]},
        },
        "plain_value": None if plain is None else round(plain, 1),
        "batched_over_plain": None if plain is None else round(value / plain, 2),
        "force_rel_err_vs_fp64": force_err,
        "ns_per_day_per_replica": round(workloads.ns_per_day(value, nrep, system.dt_fs), 3),
        "interval_latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 3),
                                "p99": round(float(np.percentile(lat, 99)) * 1e3, 3)},
        "host_exchange": {"value": round(host_value, 1), "ms_per_step": round(host_wall / args.steps * 1e3, 4)},
        "roofline": {
            "bound": "valu",
            "kernel": kname,
            "achieved": round(achieved, 4),
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP32_TFLOPS, 5),
            "traffic": measured_traffic(kname),
            "flops_per_launch": flops_per_launch,
            "kernel_ms": round(kernel_ms, 4),
        },
    }
    valu = measured_valu(kname, kernel_ms)
    if valu:
        valu["executed_over_algorithmic"] = round(valu["executed_fp32_flop_per_launch"] / flops_per_launch, 3)
        out["roofline"]["valu_counters"] = valu
    job.close()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(system, system.T, system.P)
    if not args.no_studies:
        out["studies"] = studies()
    print(json.dumps(out))


def studies():
    """The other benchmark configurations, each at its own size, after the
    timed headline loop: the uneven batch (8 tubes, 256 replicas), the served
    Poisson stream (24 jobs, nmol up to 2048, 50 intervals) and the
    precision study (512 TIP5P slit, 128 replicas, 10 000-step drift)."""
    from remd_amd import workloads

    out = {}
    for name, fn in (("uneven_batch", workloads.uneven_batch_study), ("served_stream", workloads.served_study),
                     ("precision", workloads.precision_study)):
        t0 = time.perf_counter()
        try:
            out[name] = fn()
        except Exception as e:  # one failed study must not lose the headline figures
            out[name] = {"error": f"{type(e).__name__}: {e}"}
        out[name]["wall_s"] = round(time.perf_counter() - t0, 1)
        print(f"study {name}: {out[name]['wall_s']} s", file=sys.stderr, flush=True)
    return out


if __name__ == "__main__":
    main()
