"""Shared test helpers: build matching oracle / GPU systems from one
inputs.System or from the reference fixture."""
from __future__ import annotations

import numpy as np

import pyoracle as po
from remd_amd import inputs as inp

PROP_KEYS = ["H0", "Tk", "clmb", "gkT", "ham", "kin", "lj", "pot", "prs0", "prs1", "prs2", "prs3", "rot", "tot",
             "tra", "vir0", "vir1", "vir2", "wall"]


def fixture_system(d, tag="f64", nfwall=4096):
    """inputs.System holding the reference fixture state after the driver's
    init (velocity scaling, InitializeProperty): 150 TIP4P, R = 7.4 A."""
    st = d[f"{tag}_init_state"]
    r, v, q, p = st[:, 0:3], st[:, 3:6], st[:, 6:10], st[:, 10:14]
    sc = d[f"{tag}_init_scal"]
    T = float(d["T"])
    P = float(d["P"])
    s = inp.System("TIP4P", 150, 1, 2, 1, 2.0, 8.655, 6.655, 100, 1, 1, np.array([T]), np.array([P]),
                   wall_length=7.4, nfwall=nfwall, sgm_wall=3.277, eps_wall=0.3886, rho_wall=0.38016)
    s.states = [inp.ReplicaState(r=r.copy(), v=v.copy(), q=q.copy(), p=p.copy(), s=sc[0], Ps=sc[1], Q=sc[2],
                                 Pv=(sc[3], sc[4], sc[5]), W=sc[6], L=(sc[7], sc[8], sc[9]),
                                 ids=d[f"{tag}_init_ids"].astype(np.int32))]
    return s


def oracle_for(system: inp.System, rep=0, fp32=False, state=None):
    mt = po.parse_moltype(inp.MOLTYPE[system.model])
    fw = None
    if system.confined == 1:
        fw = po.wall_table_tube(system.nfwall, system.wall_length, system.sgm_wall, system.eps_wall,
                                system.rho_wall)
    elif system.confined == 2:
        fw = po.wall_table_slit(system.nfwall, system.wall_length, system.sgm_wall, system.eps_wall,
                                system.rho_wall)
    st = state or system.states[rep]
    o = po.OracleSystem(system.desc(), mt, st.r, st.v, st.q, st.p, st.scal(), float(system.T[rep]),
                        float(system.P[rep]), fp32=fp32, fwall=fw)
    if st.ids is not None:
        o._arr("id", 1, np.int32)[:, 0] = st.ids
    return o


def relax_oracle(system: inp.System, steps=300, dt_fs=0.5):
    """Equilibrate every replica of a freshly placed system with velocity
    scaling at a small time step on the oracle (tests only)."""
    out = []
    for k, st in enumerate(system.states):
        relax = inp.System(system.model, system.nmol, 2, 0, system.confined, dt_fs, system.rcut,
                           system.rswitch, steps, 1, 1, system.T[k:k + 1], system.P[k:k + 1],
                           wall_length=system.wall_length, nfwall=system.nfwall, sgm_wall=system.sgm_wall,
                           eps_wall=system.eps_wall, rho_wall=system.rho_wall)
        relax.states = [st]
        o = oracle_for(relax)
        o.velocity_scaling()
        o.ang_velocity_scaling()
        o.initialize_property()
        o.execute_steps(steps)
        s = o.state()
        sc = o.scalars()
        out.append(inp.ReplicaState(r=s["r"], v=s["v"], q=s["q"], p=s["p"], s=1.0, Ps=0.0, Q=st.Q, Pv=(0, 0, 0),
                                    W=st.W, L=tuple(sc["L"]), ids=s["id"].astype(np.int32)))
    system.states = out
    return system


def by_id(arr, ids):
    """Rows reordered by molecule identity."""
    out = np.empty_like(arr)
    out[ids] = arr
    return out


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(np.sum((a - b) ** 2) / max(np.sum(b * b), 1e-300)))
