"""Generate golden fixtures from the reference's own CPU units (oracle/_ref,
built by `make -C oracle ref` from /root/reference/src).

Run in the container that has /root/reference:
    python tests/golden/gen_golden.py

Writes small .npz files next to this script (no pickles):
  ref_tube_chk.npz   -- the reference fixture test/input/gpu0000.chk
                        (150 TIP4P in a 7.4 A tube) run through the
                        reference driver's init (init_scaling, thermostat
                        flush, InitializeProperty) and one 100-step interval,
                        FP64 and FP32 (makefile.gnu) builds.
  ref_wall_tables.npz -- CalcWallForce tables for the config-2 tube
                        (R = 7, nfwall = 1000) and the fixture tube
                        (R = 7.4, every 64th node of nfwall = 65536).
  ref_remdinfo.npz   -- REMDInfo ladders and the four exchange pair lists.
  ref_tube_cfg2.npz  -- seeded config-2 replica (144 TIP4P, R = 7,
                        Lz = 50): init + 20 steps, FP64 build.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import pyoracle as po  # noqa: E402
from remd_amd import inputs as inp  # noqa: E402

REF_TEST = "/root/reference/test/input"


def tube_fixture():
    st = inp.read_chk(os.path.join(REF_TEST, "gpu0000.chk"))
    mtl = os.path.join(REF_TEST, "moltype.txt")
    T = 200.0 / inp.UNIT_TEMP
    P = 4e9 / inp.UNIT_PRESS
    desc = dict(nmol=150, nsite=4, dt_fs=2.0, rcut=8.655, rswitch=6.655, tconstant=1, pconstant=2, confined=1,
                interval=100, nfwall=4096, wall_length=7.4, sgm_wall=3.277, eps_wall=0.3886, rho_wall=0.38016)
    out = dict(r0=st.r, v0=st.v, q0=st.q, p0=st.p, scal0=st.scal(), T=T, P=P,
               params=np.array([8.655, 6.655, 7.4, 3.277, 0.3886, 0.38016, 4096, 2.0]))
    for tag, fp32 in (("f64", False), ("f32", True)):
        R = po.RefLib(fp32)
        rs = po.RefSystem(R, mtl, desc, st.r, st.v, st.q, st.p, st.ids, st.scal(), T, P)
        rs.quiet(R.L.ref_velocity_scaling)
        rs.quiet(R.L.ref_ang_velocity_scaling)
        rs.quiet(R.L.ref_initialize_property)
        a = rs.state()
        pr = rs.props()
        out[f"{tag}_init_state"] = np.concatenate([a["r"], a["v"], a["q"], a["p"], a["f"], a["n"]], axis=1)
        out[f"{tag}_init_ids"] = a["id"]
        out[f"{tag}_init_site_f"] = a["site_f"]
        out[f"{tag}_init_js"] = a["jstart"]
        out[f"{tag}_init_je"] = a["jend"]
        out[f"{tag}_init_scal"] = a["scal"]
        out[f"{tag}_init_props"] = np.array([pr[k] for k in sorted(pr)])
        rs.quiet(R.L.ref_execute_steps)
        a = rs.state()
        pr = rs.props()
        out[f"{tag}_int_state"] = np.concatenate([a["r"], a["v"], a["q"], a["p"], a["f"], a["n"]], axis=1)
        out[f"{tag}_int_ids"] = a["id"]
        out[f"{tag}_int_scal"] = a["scal"]
        out[f"{tag}_int_props"] = np.array([pr[k] for k in sorted(pr)])
        out[f"{tag}_int_average"] = rs.average()
        out["wall_4096"] = rs.wall_table()
        out["prop_keys"] = np.array(sorted(pr))
    np.savez_compressed(os.path.join(HERE, "ref_tube_chk.npz"), **out)


def wall_fixture():
    mtl = os.path.join(REF_TEST, "moltype.txt")
    R = po.RefLib(False)
    out = {}
    for name, Rw, nfw in (("cfg2", 7.0, 1000), ("fixture", 7.4, 65536)):
        desc = dict(nmol=2, nsite=4, dt_fs=2.0, rcut=8.655, rswitch=6.655, tconstant=1, pconstant=2,
                    confined=1, interval=1, nfwall=nfw, wall_length=Rw, sgm_wall=3.277, eps_wall=0.3886,
                    rho_wall=0.38016)
        z = np.zeros((2, 3)) + 0.5
        q = np.zeros((2, 4)); q[:, 0] = 1
        rs = po.RefSystem(R, mtl, desc, z, z * 0, q, q * 0, np.arange(2, dtype=np.int32),
                          np.array([1, 0, 0.5, 0, 0, 0, 0.05, 40, 40, 40.0]), 1.0, 0.0)
        t = rs.wall_table()
        out[name] = t if nfw <= 1000 else t[::64].copy()
    out["cfg2_params"] = np.array([7.0, 1000, 3.277, 0.3886, 0.38016])
    out["fixture_params"] = np.array([7.4, 65536, 3.277, 0.3886, 0.38016, 64])
    np.savez_compressed(os.path.join(HERE, "ref_wall_tables.npz"), **out)


def remdinfo_fixture():
    R = po.RefLib(False)
    import ctypes as C
    out = {}
    for nt, npr in ((4, 1), (16, 4), (5, 3), (2, 2), (3, 1)):
        n = nt * npr
        T = np.zeros(n); P = np.zeros(n)
        cap = n + 1
        lists = np.zeros(4 * cap * 2, np.int32)
        ln = np.zeros(4, np.int32)
        R.L.ref_remdinfo(nt, npr, 200.0, 360.0, 0.1e6, 200e6, po.dptr(T), po.dptr(P), po.iptr(lists), cap,
                         po.iptr(ln))
        out[f"T_{nt}x{npr}"] = T
        out[f"P_{nt}x{npr}"] = P
        out[f"lists_{nt}x{npr}"] = lists.reshape(4, cap, 2)
        out[f"len_{nt}x{npr}"] = ln
    np.savez_compressed(os.path.join(HERE, "ref_remdinfo.npz"), **out)


def cfg2_fixture():
    """One seeded config-2 replica through the reference FP64 build."""
    sysm = inp.config_tube(seed=2, ntemp=1, npress=1)
    st = sysm.states[0]
    mtl = os.path.join(HERE, "_tip4p_moltype.txt")
    with open(mtl, "w") as f:
        f.write(inp.MOLTYPE["TIP4P"])
    T = float(sysm.T[0]); P = float(sysm.P[0])
    desc = dict(nmol=sysm.nmol, nsite=4, dt_fs=2.0, rcut=9.0, rswitch=8.0, tconstant=1, pconstant=2, confined=1,
                interval=20, nfwall=1000, wall_length=7.0, sgm_wall=3.277, eps_wall=0.3886, rho_wall=0.38016)
    R = po.RefLib(False)
    rs = po.RefSystem(R, mtl, desc, st.r, st.v, st.q, st.p, st.ids, st.scal(), T, P)
    os.remove(mtl)
    rs.quiet(R.L.ref_velocity_scaling)
    rs.quiet(R.L.ref_ang_velocity_scaling)
    rs.quiet(R.L.ref_initialize_property)
    a = rs.state(); pr = rs.props()
    out = dict(r0=st.r, v0=st.v, q0=st.q, p0=st.p, scal0=st.scal(), T=T, P=P)
    out["init_state"] = np.concatenate([a["r"], a["v"], a["q"], a["p"], a["f"], a["n"]], axis=1)
    out["init_ids"] = a["id"]
    out["init_props"] = np.array([pr[k] for k in sorted(pr)])
    rs.quiet(R.L.ref_execute_steps)
    a = rs.state(); pr = rs.props()
    out["int_state"] = np.concatenate([a["r"], a["v"], a["q"], a["p"], a["f"], a["n"]], axis=1)
    out["int_ids"] = a["id"]
    out["int_scal"] = a["scal"]
    out["int_props"] = np.array([pr[k] for k in sorted(pr)])
    out["prop_keys"] = np.array(sorted(pr))
    np.savez_compressed(os.path.join(HERE, "ref_tube_cfg2.npz"), **out)


if __name__ == "__main__":
    if not po.ref_available():
        raise SystemExit("oracle/_ref is not built (make -C oracle ref)")
    tube_fixture()
    wall_fixture()
    remdinfo_fixture()
    cfg2_fixture()
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))
