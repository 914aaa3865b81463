"""Times the reference's own FP32 CPU build (oracle/_ref, built from the
reference's sources by `make -C oracle ref`) against the oracle's FP32
restatement on the same replicas of the bench workload (seeded placement,
one thread each, one interval per replica) and writes the ratio to
profiles/cpu_baseline_reference_vs_port.json.  Runs where oracle/_ref was
built (the reference's sources are needed to build it).

    python oracle/time_ref_vs_port.py [nreplicas]
"""
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import pyoracle as po  # noqa: E402
from remd_amd import inputs as inp  # noqa: E402


def main():
    nrep = int(sys.argv[1]) if len(sys.argv) > 1 else 16
    if not po.ref_available():
        sys.exit("oracle/_ref is not built (make -C oracle ref)")
    system = inp.config_tube(seed=2)
    desc = system.desc()
    mt_text = inp.MOLTYPE[system.model]
    mt = po.parse_moltype(mt_text)
    fw = po.wall_table_tube(system.nfwall, system.wall_length, system.sgm_wall, system.eps_wall, system.rho_wall)
    # relax each replica with the FP64 restatement first (velocity scaling,
    # 0.5 fs, no barostat, 300 steps), as the bench relaxes on the device:
    # the seeded placement leaves contacts a 2 fs step cannot integrate
    rdesc = dict(desc, tconstant=2, pconstant=0, dt=desc["dt"] * 0.25, interval=300)
    relaxed = []
    for k in range(nrep):
        st = system.states[k]
        o = po.OracleSystem(rdesc, mt, st.r, st.v, st.q, st.p, st.scal(), system.T[k], system.P[k], fwall=fw)
        o.initialize_property()
        o.execute_steps(300)
        x = o.state()
        sc = list(st.scal())
        sc[0], sc[1], sc[3], sc[4], sc[5] = 1.0, 0.0, 0.0, 0.0, 0.0
        relaxed.append((x["r"], x["v"], x["q"], x["p"], x["id"].astype("int32"), sc))
    t_ref = t_port = 0.0
    with tempfile.TemporaryDirectory() as td:
        mtl = os.path.join(td, "moltype.txt")
        with open(mtl, "w") as f:
            f.write(mt_text)
        reflib = po.RefLib(True)
        for k in range(nrep):
            r, v, q, p, ids, sc = relaxed[k]
            T, P = system.T[k], system.P[k]
            rs = po.RefSystem(reflib, mtl, desc, r, v, q, p, ids, sc, T, P)
            rs.quiet(reflib.L.ref_initialize_property)
            t0 = time.perf_counter()
            rs.quiet(reflib.L.ref_execute_steps)
            t_ref += time.perf_counter() - t0
            del rs
            o = po.OracleSystem(desc, mt, r, v, q, p, sc, T, P, fp32=True, fwall=fw)
            o.initialize_property()
            t0 = time.perf_counter()
            o.execute_steps(system.interval)
            t_port += time.perf_counter() - t0
    steps = nrep * system.interval
    out = {"workload": f"nanotube TIP4P 144 mol, seeded placement (seed 2) relaxed by the FP64 restatement "
                       f"(300 steps of 0.5 fs, velocity scaling), {nrep} replicas x 1 interval "
                       f"({system.interval} steps), one thread each",
           "reference_replica_steps_per_s": steps / t_ref, "port_replica_steps_per_s": steps / t_port,
           "reference_over_port": (steps / t_ref) / (steps / t_port),
           "cpu": os.popen("grep -m1 'model name' /proc/cpuinfo").read().split(":")[-1].strip()}
    path = os.path.join(ROOT, "profiles", "cpu_baseline_reference_vs_port.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
