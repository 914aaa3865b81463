"""The remd_md driver over several intervals and replicas: a run of 6
exchange intervals equals, bit for bit, a run of 3 followed by a restart
(`restart 1`) from its checkpoint for 3 more -- rep lines, binary logs,
checkpoints and the end-of-run files -- and the end-of-run observables
(index.dat, accept_rate.dat, tunnel_count.dat, energy.dat, index_temp.dat,
index_press.dat; src/remd.cxx:452-458, 816-879, src/remdinfo.cxx:762-832)
equal a host replay of the exchange sequence from the binary logs."""
import glob
import os
import subprocess

import numpy as np
import pytest

import remd_amd as ra
from remd_amd import inputs as inp
from remd_amd.workloads import relax

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "remd_amd", "remd_md")
NT, NP = 4, 2

INP = """input_prefix {inp}/
output_prefix {out}/
chk_in  {chk_in}
chk_out run
nmol {nmol}
natom {natom}
nsite 4
dt    2.0
rcut 9.0
rswitch 8.0
tconstant  1
temp_max 360
temp_min 200
tstat_mass 0.5
pconstant  2
press_max 200000000
press_min 100000
bstat_mass 0.05
ninterval {nint}
interval  20
snap_interval 1000
chk_interval {chk_interval}
nreplica {nrep}
ntemp  {nt}
npress {np}
cond_mode 0
init_scaling 1
restart {restart}
ngpus 1
confined 1
nfwall 500
wall_length 7.0
sgm_wall 3.277
eps_wall 0.3886
rho_wall 0.38016
"""


@pytest.fixture(scope="module")
def start_dir(tmp_path_factory):
    d = tmp_path_factory.mktemp("start")
    s = relax(inp.config_tube(seed=81, nmol=96, Lz=36.0, ntemp=NT, npress=NP, nfwall=500))
    mt = inp.molecule("TIP4P")
    for k, st in enumerate(s.states):
        inp.write_chk(str(d / f"init{k:04d}.chk"), st, mt)
    (d / "moltype.txt").write_text(inp.MOLTYPE["TIP4P"])
    return d, s.nmol


def _run(tmp, name, inp_dir, chk_in, out, nint, chk_interval, restart, nmol):
    out.mkdir(exist_ok=True)
    f = tmp / f"{name}.inp"
    f.write_text(INP.format(inp=inp_dir, out=out, chk_in=chk_in, nmol=nmol, natom=4 * nmol, nint=nint,
                            chk_interval=chk_interval, nrep=NT * NP, nt=NT, np=NP, restart=restart))
    if restart:  # the restart reads the molecule types from its input directory too
        (inp_dir / "moltype.txt").write_text(inp.MOLTYPE["TIP4P"])
    print("remd_md", name, flush=True)
    res = subprocess.run([DRIVER, str(f)], capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr
    return [l for l in res.stdout.splitlines() if l.startswith(" rep ")]


def _logs(out):
    recs = {}
    for f in glob.glob(str(out / "*.log")):
        raw = open(f, "rb").read()
        head = np.frombuffer(raw[:24], np.float64)
        nelem = int(np.frombuffer(raw[24:28], np.int32)[0])
        recs[(head[0], head[1])] = np.frombuffer(raw[28:], np.float64).reshape(-1, nelem)
    return recs


def test_restart_equals_straight_run_and_observables_replay(tmp_path, start_dir):
    assert os.path.exists(DRIVER), "build remd_amd/remd_md first"
    d, nmol = start_dir
    a, b = tmp_path / "a", tmp_path / "b"
    lines_a = _run(tmp_path, "a", d, "init", a, 6, 2, 0, nmol)
    lines_b = _run(tmp_path, "b", d, "init", b, 3, 100, 0, nmol)
    lines_c = _run(tmp_path, "c", b, "run", b, 3, 100, 1, nmol)
    nrep = NT * NP
    assert len(lines_a) == 6 * nrep and len(lines_b) == 3 * nrep and len(lines_c) == 3 * nrep
    # the restarted run continues the straight one bit for bit
    assert lines_b + lines_c == lines_a
    la, lb = _logs(a), _logs(b)
    assert la.keys() == lb.keys() and len(la) == nrep
    for k in la:
        assert la[k].shape == (6, 19) and np.array_equal(la[k], lb[k]), k
    for k in range(nrep):
        assert open(a / f"run{k:04d}.chk").read() == open(b / f"run{k:04d}.chk").read()
    for f in ("index.dat", "accept_rate.dat", "tunnel_count.dat", "energy.dat"):
        assert open(a / f).read() == open(b / f).read(), f
    assert open(a / "index_temp.dat").read().splitlines()[-1] == open(b / "index_temp.dat").read().splitlines()[-1]

    # host replay of the exchange sequence from the binary logs: slot c's
    # record of interval k is the averages of replica index[c]
    conds = sorted(la.keys(), key=lambda tp: (tp[1], tp[0]))   # slot = t + p * NT
    Tslot = np.array([c[0] for c in conds])
    Pslot = np.array([c[1] for c in conds])
    index = np.arange(nrep, dtype=np.int32)
    T, P = Tslot.copy(), Pslot.copy()
    rng = ra.GlibcRand(1)
    pairs = [ra.exchange_pairs(NT, NP, t) for t in range(4)]
    acc = [np.zeros(len(p[0]), int) for p in pairs]
    rej = [np.zeros(len(p[0]), int) for p in pairs]
    tc, tmax, tmin = np.zeros(nrep, int), np.zeros(nrep, bool), np.zeros(nrep, bool)
    pot, vol = np.zeros(nrep), np.zeros(nrep)
    for step in range(6):
        for c in range(nrep):
            rec = la[conds[c]][step]
            pot[index[c]] = rec[0] + rec[1] + rec[2]
            vol[index[c]] = rec[3]
        _, accepted, _ = ra.exchange_host(NT, NP, step, 2, pot, vol, index, T, P, rng)
        for k, x in enumerate(accepted):
            (acc if x else rej)[step % 4][k] += 1
        for rep, is_max in ((index[0], False), (index[nrep - 1], True)):
            (tmax if is_max else tmin)[rep] = True
            if tmax[rep] and tmin[rep]:
                tc[rep] += 1
                (tmin if is_max else tmax)[rep] = False
    got = np.loadtxt(a / "index.dat", dtype=int)
    assert np.array_equal(got[:, 1], index)
    assert [int(x) for x in open(a / "tunnel_count.dat").read().split()] == list(tc)
    e = open(a / "energy.dat").read().split()
    assert int(e[0]) == 6
    assert np.allclose(np.array([float(x) for x in e[1:]]), pot + vol * P, rtol=2e-15, atol=0)  # 16 digits
    it = open(a / "index_temp.dat").read().split()
    assert int(it[0]) == 6 and np.allclose([float(x) for x in it[1:]], T, rtol=0, atol=5e-7)
    ip = open(a / "index_press.dat").read().split()
    assert np.allclose([float(x) for x in ip[1:]], P, rtol=0, atol=5e-7)
    # accept_rate.dat: per pressure row T-even / T-odd pairs interleaved, then
    # per temperature column P-even / P-odd (src/remdinfo.cxx:762-820)
    rows = [l.split() for l in open(a / "accept_rate.dat").read().splitlines()[1:] if l.strip()]
    want = []
    ll = [len(p[0]) for p in pairs]

    def row(t, i):
        i0, i1 = index[pairs[t][0][i]], index[pairs[t][1][i]]
        n = acc[t][i] + rej[t][i]
        return (T[i0], P[i0], T[i1], P[i1], acc[t][i] / n if n else np.nan)
    for p_ in range(NP):
        for i in range(ll[0] // NP):
            want.append(row(0, i + p_ * ll[0] // NP))
            if i >= ll[2] // NP:
                break
            want.append(row(2, i + p_ * ll[2] // NP))
    for t_ in range(NT):
        for i in range(ll[1] // NT):
            want.append(row(1, i + t_ * ll[1] // NT))
            if i >= ll[3] // NT:
                break
            want.append(row(3, i + t_ * ll[3] // NT))
    assert len(rows) == len(want)
    for r, w in zip(rows, want):
        vals = [float(x) for x in r]
        for x, y in zip(vals, w):
            assert (np.isnan(x) and np.isnan(y)) or abs(x - y) <= 5e-7, (r, w)
