"""Served REMD jobs with continuous batching.

Requests (a REMD job: its system, initial replica states and a number of
exchange intervals) arrive over time.  At every interval boundary the server
admits, in arrival order, the requests that have arrived and fit -- the sum
of resident job footprints (remd_gpu_footprint) within the HBM budget and
the resident replicas within the slot limit -- creates them on the device
(state stays resident until the job retires), runs ONE batched interval for
every resident job (remd_batch_execute: one launch per kernel class), then
does each job's replica exchange on the host with the job's own glibc
rand() stream (REMD::ReplicaExchange, src/remd.cxx:486-538) and stages the
new conditions.  Jobs that finished their intervals retire and free their
slots and memory for the next admissions.

A job's result does not depend on what else is resident: the kernels give
the same bits alone and batched, and the exchange sees only the job's own
energies and stream.
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Request:
    job_id: int
    arrival: float            # seconds after the stream starts
    system: object            # inputs.System with states
    intervals: int = 50


@dataclass
class JobResult:
    job_id: int
    arrival: float
    admitted: float
    finished: float
    intervals: int
    replica_steps: int
    latencies: list = field(default_factory=list)   # seconds per interval
    averages: np.ndarray = None                      # (nreplica, 19) after the last interval
    state0: dict = None                              # replica 0 state after the last interval
    index: np.ndarray = None                         # condition slot -> replica


class _Active:
    def __init__(self, req, job, rng, now, fp):
        self.req = req
        self.job = job
        self.rng = rng
        s = req.system
        self.T = np.ascontiguousarray(s.T, dtype=np.float64).copy()
        self.P = np.ascontiguousarray(s.P, dtype=np.float64).copy()
        self.index = np.arange(s.nreplica, dtype=np.int32)
        self.done = 0
        self.admitted = now
        self.footprint = fp
        self.latencies = []


class GPUBackend:
    """The device: jobs, one batch, host exchange and the footprint query."""

    def __init__(self, precision=0):
        import remd_amd as ra

        self.ra = ra
        self.precision = precision
        self.batch = ra.Batch()

    def footprint(self, system):
        import ctypes as C

        d = self.ra.desc_for(system, self.precision)
        return int(self.ra.lib().remd_gpu_footprint(C.byref(d)))

    def create(self, system):
        return self.ra.REMDGPU(system, precision=self.precision)

    def add(self, job):
        self.batch.add(job)

    def remove(self, job):
        self.batch.remove(job)

    def stage(self, job, T, P):
        job.set_conditions(T, P)

    def execute(self):
        self.batch.execute()

    def exchange(self, a: "_Active", step):
        s = a.req.system
        pot, vol = a.job.pot_vol()
        self.ra.exchange_host(s.ntemp, s.npress, step, s.pconstant, pot, vol, a.index, a.T, a.P, a.rng)

    def rng(self):
        return self.ra.GlibcRand(1)

    def result(self, job):
        return job.all_averages(), job.state(0)

    def close_job(self, job):
        job.close()

    def close(self):
        self.batch.close()


class Server:
    """Continuous batching of REMD jobs under an HBM budget and a slot limit."""

    def __init__(self, backend, hbm_budget=200e9, max_replicas=2048, clock=time.perf_counter,
                 sleep=time.sleep):
        self.backend = backend
        self.hbm_budget = float(hbm_budget)
        self.max_replicas = int(max_replicas)
        self.clock = clock
        self.sleep = sleep

    def run(self, requests):
        """Serve `requests` (any order); returns {job_id: JobResult}."""
        pending = deque(sorted(requests, key=lambda r: (r.arrival, r.job_id)))
        fps = {r.job_id: self.backend.footprint(r.system) for r in pending}
        for r in pending:
            if fps[r.job_id] > self.hbm_budget or r.system.nreplica > self.max_replicas:
                raise ValueError(f"job {r.job_id} can never be admitted (footprint {fps[r.job_id]} B, "
                                 f"{r.system.nreplica} replicas)")
        active: list[_Active] = []
        results = {}
        used_mem = 0
        used_reps = 0
        t0 = self.clock()
        while pending or active:
            now = self.clock() - t0
            # admissions at the interval boundary, in arrival order
            while pending and pending[0].arrival <= now:
                r = pending[0]
                fp = fps[r.job_id]
                if used_mem + fp > self.hbm_budget or used_reps + r.system.nreplica > self.max_replicas:
                    break
                pending.popleft()
                job = self.backend.create(r.system)
                a = _Active(r, job, self.backend.rng(), self.clock() - t0, fp)
                self.backend.add(job)
                self.backend.stage(job, a.T, a.P)
                active.append(a)
                used_mem += fp
                used_reps += r.system.nreplica
            if not active:
                if pending:
                    wait = pending[0].arrival - (self.clock() - t0)
                    if wait > 0:
                        self.sleep(wait)
                continue
            t1 = self.clock()
            self.backend.execute()
            for a in active:
                self.backend.exchange(a, a.done)
                self.backend.stage(a.job, a.T, a.P)
            t2 = self.clock()
            still = []
            for a in active:
                a.latencies.append(t2 - t1)
                a.done += 1
                if a.done < a.req.intervals:
                    still.append(a)
                    continue
                self.backend.remove(a.job)
                av, st0 = self.backend.result(a.job)
                s = a.req.system
                results[a.req.job_id] = JobResult(
                    job_id=a.req.job_id, arrival=a.req.arrival, admitted=a.admitted, finished=t2 - t0,
                    intervals=a.done, replica_steps=a.done * s.interval * s.nreplica, latencies=a.latencies,
                    averages=av, state0=st0, index=a.index.copy())
                self.backend.close_job(a.job)
                used_mem -= a.footprint
                used_reps -= s.nreplica
            active = still
        return results


def latency_percentiles(results):
    lat = np.concatenate([np.asarray(r.latencies) for r in results.values()]) if results else np.zeros(1)
    return float(np.percentile(lat, 50)), float(np.percentile(lat, 99))


def poisson_stream(seed=5, njobs=40, rate=20.0, intervals=50, nmol_hi=2048, relax=True):
    """The served-stream workload: Poisson arrivals (mean `rate` jobs/s) of
    inputs.served_job systems, each relaxed on the device (outside the
    timed stream) and run for `intervals` exchange intervals."""
    from . import inputs, workloads

    rng = np.random.RandomState(seed)
    t = 0.0
    reqs = []
    for k in range(njobs):
        t += rng.exponential(1.0 / rate)
        s = inputs.served_job(rng, nmol_hi=nmol_hi)
        if relax:
            s = workloads.relax(s)
            print(f"job {k}: {s.model} confined={s.confined} nmol={s.nmol} replicas={s.nreplica}",
                  file=__import__("sys").stderr, flush=True)
        reqs.append(Request(job_id=k, arrival=t, system=s, intervals=intervals))
    return reqs
