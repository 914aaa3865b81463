"""Served REMD jobs with continuous batching.

Requests (a REMD job: its system, initial replica states and a number of
exchange intervals) arrive over time.  The server admits, in arrival order,
the requests that have arrived and fit -- the sum of resident job footprints
(remd_gpu_footprint) within the HBM budget and the resident replicas within
the slot limit -- and creates them on the device, where their state stays
until they retire.

Jobs run in lanes: one batch (remd_batch_*) per band of per-replica
interval cost (one lane by default), each on its own stream.  A lane
launches one interval of all its jobs (one kernel launch for the resident
replicas), and the host loop waits for whichever lane finishes first, does
the replica exchange of each of that lane's jobs with the job's own glibc
rand() stream (REMD::ReplicaExchange, src/remd.cxx:486-538), stages the new
conditions and relaunches the lane at once.  With several lanes, light jobs
go through many intervals while a heavy one runs a single interval, and the
lanes share the device.

A request of higher priority that does not fit preempts running jobs of
lower priority: their exact device state is saved (remd_gpu_checkpoint_save)
with their exchange state (condition index, conditions, rand() stream) and
they are freed; they resume from that state, bit for bit, when room frees
up.  Jobs that finished their intervals retire and free their slots and
memory for the next admissions.

A job's result does not depend on what else is resident, on its lane, on
preemption or on the packing plan: the kernels give the same bits alone and
batched, and the exchange sees only the job's own energies and stream.
Each request may name an output directory: the job then writes the md.out
outputs of its run there (binary records per condition, index.dat,
accept_rate.dat, tunnel_count.dat, energy.dat, index_temp.dat,
index_press.dat; src/remd.cxx:150-193, 816-879, 891-931,
src/remdinfo.cxx:762-832).
"""
from __future__ import annotations

import os
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
    priority: int = 0         # a waiting request preempts running jobs of lower priority
    output_dir: str = None    # per-job output stream (md.out files), or None


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
    preemptions: int = 0
    lane: int = 0


class JobOutput:
    """The md.out outputs of one served job: per condition slot the binary
    record of the 19 averages of the replica holding it after every
    interval (REMD::OutputBinary, src/remd.cxx:150-193, 891-931), and at the
    end index.dat, accept_rate.dat, tunnel_count.dat, energy.dat,
    index_temp.dat and index_press.dat (src/remd.cxx:816-879,
    src/remdinfo.cxx:762-832)."""

    def __init__(self, out_dir, system, exchange_pairs):
        os.makedirs(out_dir, exist_ok=True)
        self.dir = out_dir
        self.ntemp, self.npress, self.pconstant = system.ntemp, system.npress, system.pconstant
        n = system.nreplica
        self.files = []
        for c in range(n):
            T, P = float(system.T[c]), float(system.P[c])
            f = open(os.path.join(out_dir, f"P{_fmt(P)}T{_fmt(T)}.log"), "wb")
            f.write(np.array([T, P, 1.0], np.float64).tobytes())
            f.write(np.array([19], np.int32).tobytes())
            self.files.append(f)
        self.pairs = [exchange_pairs(system.ntemp, system.npress, t) for t in range(4)]
        self.acc = [np.zeros(len(p[0]), np.int64) for p in self.pairs]
        self.rej = [np.zeros(len(p[0]), np.int64) for p in self.pairs]
        self.tc = np.zeros(n, np.int64)
        self.tmax = np.zeros(n, bool)
        self.tmin = np.zeros(n, bool)

    def record(self, averages, index):
        for c, f in enumerate(self.files):
            f.write(np.ascontiguousarray(averages[index[c]], np.float64).tobytes())

    def exchange(self, step, accepted, index):
        t = step % 4
        for k, a in enumerate(accepted):
            (self.acc if a else self.rej)[t][k] += 1
        # REMDInfo::CheckTunneling (src/remdinfo.h:389-392)
        for rep, is_max in ((index[0], False), (index[-1], True)):
            (self.tmax if is_max else self.tmin)[rep] = True
            if self.tmax[rep] and self.tmin[rep]:
                self.tc[rep] += 1
                (self.tmin if is_max else self.tmax)[rep] = False

    def close(self, step, pot, vol, index, T, P):
        for f in self.files:
            f.close()
        n = len(index)
        d = self.dir
        with open(os.path.join(d, "index.dat"), "w") as f:
            for i in range(n):
                f.write("%4d %4d\n" % (i, index[i]))
        with open(os.path.join(d, "accept_rate.dat"), "w") as f:
            f.write("#Ti,Pi,Tj,Pj,AR\n")
            ll = [len(p[0]) for p in self.pairs]

            def line(t, i):
                i0, i1 = index[self.pairs[t][0][i]], index[self.pairs[t][1][i]]
                tries = self.acc[t][i] + self.rej[t][i]
                rate = self.acc[t][i] / tries if tries else float("nan")
                f.write("%f %f %f %f %f\n" % (T[i0], P[i0], T[i1], P[i1], rate))
            for p in range(self.npress):
                for i in range(ll[0] // self.npress):
                    line(0, i + p * ll[0] // self.npress)
                    if i >= ll[2] // self.npress:
                        break
                    line(2, i + p * ll[2] // self.npress)
                f.write("\n")
            for t in range(self.ntemp):
                for i in range(ll[1] // self.ntemp):
                    line(1, i + t * ll[1] // self.ntemp)
                    if i >= ll[3] // self.ntemp:
                        break
                    line(3, i + t * ll[3] // self.ntemp)
                f.write("\n")
        with open(os.path.join(d, "tunnel_count.dat"), "w") as f:
            f.write("".join(" %d" % x for x in self.tc) + "\n")
        with open(os.path.join(d, "energy.dat"), "w") as f:
            e = pot + vol * P if self.pconstant > 0 else pot
            f.write("%8d" % step + "".join(" %16s" % _g16(x) for x in e) + "\n")
        for name, vals in (("index_temp.dat", T), ("index_press.dat", P)):
            with open(os.path.join(d, name), "w" if step == 0 else "a") as f:
                f.write(" %7d" % step + "".join(" %f" % x for x in vals) + "\n")


def _fmt(x):
    """std::ostream default formatting (6 significant digits, %g style)."""
    return "%g" % x


def _g16(x):
    """std::ostream with setprecision(16) (16 significant digits, %g style)."""
    return "%.16g" % x


class _Active:
    def __init__(self, req, rng, fp, lane):
        self.req = req
        self.job = None
        self.rng = rng
        s = req.system
        self.T = np.ascontiguousarray(s.T, dtype=np.float64).copy()
        self.P = np.ascontiguousarray(s.P, dtype=np.float64).copy()
        self.index = np.arange(s.nreplica, dtype=np.int32)
        self.done = 0
        self.admitted = None
        self.footprint = fp
        self.latencies = []
        self.lane = lane
        self.blob = None          # saved device state while preempted
        self.preemptions = 0
        self.output = None
        self.pot = None
        self.vol = None


class GPUBackend:
    """The device: jobs, one batch per lane, host exchange, checkpoints and
    the footprint query.  Jobs of either pair precision batch together."""

    def __init__(self, precision=0, split=True):
        import remd_amd as ra

        self.ra = ra
        self.precision = precision
        self.split = split        # False: one workgroup per replica in every lane
        self.lanes = {}
        self._dirty = set()

    def footprint(self, system):
        import ctypes as C

        d = self.ra.desc_for(system, self.precision)
        return int(self.ra.lib().remd_gpu_footprint(C.byref(d)))

    @staticmethod
    def cost(system):
        """Relative per-replica interval cost: molecules x pair work per
        molecule (~ sites^2)."""
        return float(system.nmol) * float(system.nsite) ** 2

    def create(self, system):
        return self.ra.REMDGPU(system, precision=self.precision)

    def _lane(self, lane):
        if lane not in self.lanes:
            self.lanes[lane] = self.ra.Batch()
        return self.lanes[lane]

    def add(self, lane, job):
        self._lane(lane).add(job)
        self._dirty.add(lane)

    def remove(self, lane, job):
        self._lane(lane).remove(job)
        self._dirty.add(lane)

    def stage(self, job, T, P):
        job.set_conditions(T, P)

    def launch(self, lane):
        b = self._lane(lane)
        if not self.split and lane in self._dirty:
            b.set_parts([1] * len(b.jobs))
        self._dirty.discard(lane)
        b.launch()

    def poll(self, lane):
        return self._lane(lane).poll()

    def wait(self, lane):
        self._lane(lane).wait()

    def exchange(self, a: "_Active", step):
        s = a.req.system
        pot, vol = a.job.pot_vol()
        a.pot, a.vol = pot, vol
        _, accepted, _ = self.ra.exchange_host(s.ntemp, s.npress, step, s.pconstant, pot, vol, a.index, a.T, a.P,
                                               a.rng)
        return accepted

    def averages(self, job):
        return job.all_averages()

    def exchange_pairs(self, ntemp, npress, t):
        return self.ra.exchange_pairs(ntemp, npress, t)

    def rng(self):
        return self.ra.GlibcRand(1)

    def save(self, job):
        return job.checkpoint_save()

    def restore(self, system, blob):
        job = self.create(system)
        job.checkpoint_load(blob)
        return job

    def result(self, job):
        return job.all_averages(), job.state(0)

    def close_job(self, job):
        job.close()

    def close(self):
        for b in self.lanes.values():
            b.close()
        self.lanes = {}


class Server:
    """Continuous batching of REMD jobs under an HBM budget and a slot limit,
    with priority preemption.  `lane_bounds` (per-replica cost thresholds,
    GPUBackend.cost) splits the jobs into lanes that advance independently;
    the default is one lane.  On the served Poisson stream (24 jobs, nmol up
    to 2048) one lane gave the highest throughput, and three lanes a lower
    median interval latency but a higher 99th percentile: the light lanes'
    workgroups take CUs that the split replicas of the heavy lane wait for."""

    def __init__(self, backend, hbm_budget=200e9, max_replicas=2048, clock=time.perf_counter,
                 sleep=time.sleep, lane_bounds=(), poll_s=50e-6, max_retries=3):
        self.backend = backend
        self.hbm_budget = float(hbm_budget)
        self.max_replicas = int(max_replicas)
        self.clock = clock
        self.sleep = sleep
        self.lane_bounds = tuple(lane_bounds)
        self.poll_s = poll_s
        self.max_retries = max_retries

    def lane_of(self, system):
        c = self.backend.cost(system)
        return sum(1 for b in self.lane_bounds if c >= b)

    def run(self, requests):
        """Serve `requests` (any order); returns {job_id: JobResult}."""
        be = self.backend
        pending = deque(sorted(requests, key=lambda r: (r.arrival, r.job_id)))
        fps = {r.job_id: be.footprint(r.system) for r in pending}
        for r in pending:
            if fps[r.job_id] > self.hbm_budget or r.system.nreplica > self.max_replicas:
                raise ValueError(f"job {r.job_id} can never be admitted (footprint {fps[r.job_id]} B, "
                                 f"{r.system.nreplica} replicas)")
        waiting = deque()                  # _Active entries not on the device (new or preempted), arrival order
        for r in pending:
            waiting.append(_Active(r, None, fps[r.job_id], self.lane_of(r.system)))
        lanes = {}                         # lane -> {"jobs": [...], "flight": bool, "t0": float, "retries": int}
        results = {}
        used = {"mem": 0, "reps": 0}
        t0 = self.clock()

        def now():
            return self.clock() - t0

        def lane(k):
            return lanes.setdefault(k, {"jobs": [], "flight": False, "t0": 0.0, "retries": 0})

        def fits(a):
            s = a.req.system
            return (used["mem"] + a.footprint <= self.hbm_budget and
                    used["reps"] + s.nreplica <= self.max_replicas)

        def place(a):
            s = a.req.system
            if a.blob is None:
                a.job = be.create(s)
                a.rng = be.rng()
                a.admitted = now()
                if a.req.output_dir:
                    a.output = JobOutput(a.req.output_dir, s, be.exchange_pairs)
            else:
                a.job = be.restore(s, a.blob)
                a.rng.setstate(a.rng_state)
                a.blob = None
            try:
                be.add(a.lane, a.job)
            except Exception:
                be.close_job(a.job)
                raise
            be.stage(a.job, a.T, a.P)
            lane(a.lane)["jobs"].append(a)
            used["mem"] += a.footprint
            used["reps"] += s.nreplica

        def unplace(a):
            be.remove(a.lane, a.job)
            be.close_job(a.job)
            a.job = None
            lane(a.lane)["jobs"].remove(a)
            used["mem"] -= a.footprint
            used["reps"] -= a.req.system.nreplica

        def victims_for(head):
            """Running jobs of lower priority in idle lanes (lowest priority,
            latest admitted first) whose preemption makes room for `head`,
            or None when preemption cannot make room."""
            cands = [a for k, L in lanes.items() if not L["flight"] for a in L["jobs"]
                     if a.req.priority < head.req.priority]
            cands.sort(key=lambda a: (a.req.priority, -a.admitted))
            mem, reps = used["mem"], used["reps"]
            need_reps = head.req.system.nreplica
            chosen = []
            for a in cands:
                if mem + head.footprint <= self.hbm_budget and reps + need_reps <= self.max_replicas:
                    break
                chosen.append(a)
                mem -= a.footprint
                reps -= a.req.system.nreplica
            if mem + head.footprint <= self.hbm_budget and reps + need_reps <= self.max_replicas:
                return chosen
            return None

        def preempt(a):
            a.blob = be.save(a.job)
            a.rng_state = a.rng.getstate()
            a.preemptions += 1
            unplace(a)

        def admit():
            while waiting and waiting[0].req.arrival <= now():
                a = waiting[0]
                if lane(a.lane)["flight"]:
                    break                  # its lane takes new jobs at its next boundary
                victims = [] if fits(a) else victims_for(a)
                if victims is None:
                    break
                waiting.popleft()
                for v in victims:
                    preempt(v)
                place(a)
                if victims:                # back in the queue at their arrival positions
                    items = sorted(list(waiting) + victims, key=lambda x: (x.req.arrival, x.req.job_id))
                    waiting.clear()
                    waiting.extend(items)

        while waiting or any(L["jobs"] for L in lanes.values()):
            admit()
            for k, L in lanes.items():
                if L["jobs"] and not L["flight"]:
                    be.launch(k)
                    L["flight"] = True
                    L["t0"] = now()
            flying = [k for k, L in lanes.items() if L["flight"]]
            if not flying:
                if waiting:
                    wait_s = waiting[0].req.arrival - now()
                    if wait_s > 0:
                        self.sleep(wait_s)
                continue
            # the first lane to finish
            done = None
            while done is None:
                for k in flying:
                    if be.poll(k):
                        done = k
                        break
                else:
                    self.sleep(self.poll_s)
            L = lanes[done]
            try:
                be.wait(done)
            except RuntimeError as e:
                # a split replica whose parts were not co-resident gives up
                # without writing state: run the same interval again
                if "aborted" not in str(e) or L["retries"] >= self.max_retries:
                    raise
                L["retries"] += 1
                L["flight"] = False
                continue
            L["flight"] = False
            L["retries"] = 0
            t_end = now()
            still = []
            for a in L["jobs"]:
                s = a.req.system
                if a.output is not None:   # slot c holds replica index[c] during the interval
                    a.output.record(be.averages(a.job), a.index)
                accepted = be.exchange(a, a.done)
                if a.output is not None:
                    a.output.exchange(a.done, accepted, a.index)
                a.latencies.append(now() - L["t0"])
                a.done += 1
                if a.done < a.req.intervals:
                    be.stage(a.job, a.T, a.P)
                    still.append(a)
                    continue
                av, st0 = be.result(a.job)
                if a.output is not None:
                    a.output.close(a.done, a.pot, a.vol, a.index, a.T, a.P)
                results[a.req.job_id] = JobResult(
                    job_id=a.req.job_id, arrival=a.req.arrival, admitted=a.admitted, finished=t_end,
                    intervals=a.done, replica_steps=a.done * s.interval * s.nreplica, latencies=a.latencies,
                    averages=av, state0=st0, index=a.index.copy(), preemptions=a.preemptions, lane=a.lane)
                be.remove(a.lane, a.job)
                be.close_job(a.job)
                used["mem"] -= a.footprint
                used["reps"] -= s.nreplica
            L["jobs"] = still
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
