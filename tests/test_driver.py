"""The remd_md driver (md.out's REMD loop over the C ABI) on the reference's
own test case: test/input/gpu.inp (150 TIP4P in a 7.4 A tube, 200 K,
4 GPa, NP + 1D barostat, init_scaling 1) started from the fixture's copy of
test/input/gpu0000.chk.  After one 100-step interval its checkpoint must
match the reference build's state (FP32 pair math: 1e-3, the tolerance of
the interval parity test), and its " rep" line and binary log must carry
the same energies as the library's averages."""
import os
import subprocess

import numpy as np
import pytest

from remd_amd import inputs as inp
from helpers import by_id

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "remd_amd", "remd_md")

INP = """input_prefix {inp}/
output_prefix {out}/
chk_in  gpu
chk_out gpu
nmol 150
natom 600
nsite 4
dt    2.0
rcut 8.655
rswitch 6.655
tconstant  1
temp_max 400
temp_min 200
tstat_mass 0.5
pconstant  2
press_max 5000000000
press_min 4000000000
bstat_mass 0.05
ninterval {nint}
interval  100
snap_interval 1000
chk_interval 10000
nreplica 1
ntemp  1
npress 1
cond_mode 0
init_scaling 1
ngpus 1
confined 1
nfwall 65536
wall_length 7.4
sgm_wall 3.277
eps_wall 0.3886
rho_wall 0.38016
"""


def _setup(tmp_path, d, nint):
    ind = tmp_path / "in"
    out = tmp_path / "out"
    ind.mkdir()
    out.mkdir()
    sc = d["scal0"]
    st = inp.ReplicaState(r=d["r0"], v=d["v0"], q=d["q0"], p=d["p0"], s=sc[0], Ps=sc[1], Q=sc[2],
                          Pv=tuple(sc[3:6]), W=sc[6], L=tuple(sc[7:10]))
    inp.write_chk(str(ind / "gpu0000.chk"), st, inp.molecule("TIP4P"))
    (ind / "moltype.txt").write_text(inp.MOLTYPE["TIP4P"])
    (tmp_path / "gpu.inp").write_text(INP.format(inp=ind, out=out, nint=nint))
    return out


def test_driver_reference_case_one_interval(golden, tmp_path):
    assert os.path.exists(DRIVER), "build remd_amd/remd_md first"
    d = golden("ref_tube_chk.npz")
    out = _setup(tmp_path, d, 1)
    res = subprocess.run([DRIVER, str(tmp_path / "gpu.inp")], capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr
    lines = [l for l in res.stdout.splitlines() if l.startswith(" rep ")]
    assert len(lines) == 1
    back = inp.read_chk(str(out / "gpu0000.chk"))
    ref = d["f64_int_state"]
    ids = d["f64_int_ids"]
    # the checkpoint lists molecules in their original order
    for name, sl in (("r", slice(0, 3)), ("q", slice(6, 10))):
        a = getattr(back, name)
        b = by_id(ref[:, sl], ids)
        if name == "r":
            dd = a - b
            dd -= np.rint(dd)
            err = float(np.max(np.abs(dd)))
        else:
            err = float(np.max(np.abs(a - b)))
        print(name, err)
        assert err <= 1e-3
    # rep line: pot = lj + clmb + wall, matching the binary log record
    f = lines[0].split()
    pot, lj, clmb, wall = float(f[3]), float(f[5]), float(f[6]), float(f[7])
    assert abs(pot - (lj + clmb + wall)) <= 1e-4 * abs(pot)
    logs = [x for x in os.listdir(out) if x.endswith(".log")]
    assert len(logs) == 1
    raw = open(out / logs[0], "rb").read()
    head = np.frombuffer(raw[:24], np.float64)
    nelem = int(np.frombuffer(raw[24:28], np.int32)[0])
    rec = np.frombuffer(raw[28:], np.float64).reshape(-1, nelem)
    assert nelem == 19 and rec.shape[0] == 1
    assert head[0] == pytest.approx(200.0 / inp.UNIT_TEMP, rel=1e-6)
    assert rec[0, 0] == pytest.approx(lj, rel=1e-5) and rec[0, 1] == pytest.approx(clmb, rel=1e-5)
