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
