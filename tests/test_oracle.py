"""The oracle pinned on the CPU: against the reference's own outputs (the
golden fixtures written by tests/golden/gen_golden.py from the reference's
CPU units) and against what the mathematics fixes independently (closed-form
wall integrals, finite differences of the energy, energy conservation order,
Boltzmann weight ratios, brute-force windows).

Tolerances: fixture comparisons are bit-exact (the oracle restates the
reference expression by expression, FP64 and the FP32 build alike);
closed-form checks state theirs per test.
"""
import math

import numpy as np
import pytest
from scipy import integrate

import pyoracle as po
from remd_amd import inputs as inp
from helpers import PROP_KEYS, fixture_system, oracle_for

SGM, EPS, RHO = 3.277, 0.3886, 0.38016


def _fixture_oracle(d, fp32):
    """The reference driver's init on the fixture checkpoint, on the oracle."""
    s = inp.System("TIP4P", 150, 1, 2, 1, 2.0, 8.655, 6.655, 100, 1, 1, np.array([float(d["T"])]),
                   np.array([float(d["P"])]), wall_length=7.4, nfwall=4096, sgm_wall=SGM, eps_wall=EPS,
                   rho_wall=RHO)
    s.states = [inp.ReplicaState(r=d["r0"], v=d["v0"], q=d["q0"], p=d["p0"], s=d["scal0"][0], Ps=d["scal0"][1],
                                 Q=d["scal0"][2], Pv=tuple(d["scal0"][3:6]), W=d["scal0"][6],
                                 L=tuple(d["scal0"][7:10]), ids=np.arange(150, dtype=np.int32))]
    o = oracle_for(s, fp32=fp32)
    o.velocity_scaling()
    o.ang_velocity_scaling()
    o.initialize_property()
    return o


def _state_rows(o):
    a = o.state()
    return np.concatenate([a["r"], a["v"], a["q"], a["p"], a["f"], a["n"]], axis=1), a


def _props(o):
    sc = o.scalars()
    flat = {"H0": sc["H0"], "Tk": sc["Tk"], "clmb": sc["clmb"], "gkT": sc["gkT"], "ham": sc["ham"],
            "kin": sc["kin"], "lj": sc["lj"], "pot": sc["pot"], "rot": sc["rot"], "tot": sc["tot"],
            "tra": sc["tra"], "wall": sc["wall"]}
    for k in range(4):
        flat[f"prs{k}"] = sc["prs"][k]
    for k in range(3):
        flat[f"vir{k}"] = sc["vir"][k]
    return np.array([flat[k] for k in PROP_KEYS])


# ------------------------------------------------------- reference fixtures
@pytest.mark.parametrize("tag,fp32", [("f64", False), ("f32", True)])
def test_fixture_init_and_interval_bit_exact(golden, tag, fp32):
    """test/input/gpu0000.chk (150 TIP4P, R 7.4 A tube): the driver's init and
    one 100-step NP + 1D-Andersen interval equal the reference build's,
    bit for bit (src/remd.cxx:102-140, src/integrator.cxx:646-706)."""
    d = golden("ref_tube_chk.npz")
    assert list(d["prop_keys"]) == PROP_KEYS
    o = _fixture_oracle(d, fp32)
    rows, a = _state_rows(o)
    assert np.array_equal(rows, d[f"{tag}_init_state"])
    assert np.array_equal(a["id"], d[f"{tag}_init_ids"])
    assert np.array_equal(a["site_f"], d[f"{tag}_init_site_f"])
    assert np.array_equal(a["jstart"], d[f"{tag}_init_js"])
    assert np.array_equal(a["jend"], d[f"{tag}_init_je"])
    assert np.array_equal(_props(o), d[f"{tag}_init_props"])
    o.execute_steps(100)
    rows, a = _state_rows(o)
    assert np.array_equal(rows, d[f"{tag}_int_state"])
    assert np.array_equal(a["id"], d[f"{tag}_int_ids"])
    assert np.array_equal(_props(o), d[f"{tag}_int_props"])
    o.calc_properties()
    # the 19th slot (DUMMY) is left uninitialised by the reference
    assert np.array_equal(o.average()[:18], d[f"{tag}_int_average"][:18])


def test_fixture_tube_config_bit_exact(golden):
    """Seeded 144-TIP4P, R 7 A, Lz 50 A replica: init + 20 steps, FP64."""
    d = golden("ref_tube_cfg2.npz")
    s = inp.System("TIP4P", 144, 1, 2, 1, 2.0, 9.0, 8.0, 20, 1, 1, np.array([float(d["T"])]),
                   np.array([float(d["P"])]), wall_length=7.0, nfwall=1000, sgm_wall=SGM, eps_wall=EPS,
                   rho_wall=RHO)
    sc = d["scal0"]
    s.states = [inp.ReplicaState(r=d["r0"], v=d["v0"], q=d["q0"], p=d["p0"], s=sc[0], Ps=sc[1], Q=sc[2],
                                 Pv=tuple(sc[3:6]), W=sc[6], L=tuple(sc[7:10]), ids=np.arange(144, dtype=np.int32))]
    o = oracle_for(s)
    o.velocity_scaling()
    o.ang_velocity_scaling()
    o.initialize_property()
    rows, a = _state_rows(o)
    assert np.array_equal(rows, d["init_state"])
    assert np.array_equal(_props(o), d["init_props"])
    o.execute_steps(20)
    rows, a = _state_rows(o)
    assert np.array_equal(rows, d["int_state"])
    assert np.array_equal(a["id"], d["int_ids"])
    assert np.array_equal(_props(o), d["int_props"])


def test_fixture_wall_tables_bit_exact(golden):
    """ForceCalculator::CalcWallForce tables (src/force_calculator.cxx:117-161)."""
    d = golden("ref_wall_tables.npz")
    t = po.wall_table_tube(1000, 7.0, SGM, EPS, RHO)
    assert np.array_equal(t, d["cfg2"])
    t = po.wall_table_tube(65536, 7.4, SGM, EPS, RHO)
    assert np.array_equal(t[::64], d["fixture"])


@pytest.mark.parametrize("nt,npr", [(4, 1), (16, 4), (5, 3), (2, 2), (3, 1)])
def test_fixture_ladders_and_pair_lists(golden, nt, npr):
    """REMDInfo ladders (src/remdinfo.cxx:6-52) and the four exchange pair
    lists (src/remdinfo.cxx:54-108)."""
    d = golden("ref_remdinfo.npz")
    # REMDInfo builds the ladder on reduced values (src/remdinfo.cxx:168-171)
    T, P = po.ladder_geometric(nt, npr, 200.0 / inp.UNIT_TEMP, 360.0 / inp.UNIT_TEMP, 0.1e6 / inp.UNIT_PRESS,
                               200e6 / inp.UNIT_PRESS)
    assert np.array_equal(T, d[f"T_{nt}x{npr}"])
    assert np.array_equal(P, d[f"P_{nt}x{npr}"])
    for typ in range(4):
        lst = po.exchange_list(nt, npr, typ)
        n = int(d[f"len_{nt}x{npr}"][typ])
        assert lst.shape[0] == n
        assert np.array_equal(lst, d[f"lists_{nt}x{npr}"][typ, :n])


# ----------------------------------------------------- closed forms: walls
def _cyl_line(rho_dist):
    """Energy per unit length-angle of an infinite line of 12-6 sites at
    distance rho_dist, integrated over z in closed form:
    int (d^2+z^2)^-6 dz = 63 pi / 256 d^-11, int (d^2+z^2)^-3 dz = 3 pi / 8 d^-5."""
    s6 = SGM ** 6
    return 4 * EPS * RHO * (s6 * s6 * 63 * math.pi / 256 * rho_dist ** -11 - s6 * 3 * math.pi / 8 * rho_dist ** -5)


def _cyl_energy(r, R):
    f = lambda phi: _cyl_line(math.sqrt(R * R + r * r - 2 * R * r * math.cos(phi))) * R
    return 2 * integrate.quad(f, 0, math.pi, epsabs=0, epsrel=1e-11, limit=200)[0]


def test_tube_wall_closed_form_on_axis():
    """On the axis every wall site is at distance R, so the tube table's
    energy is 2 pi R x the line integral (float sigma^6 / sqrt in the table:
    relative 1e-6)."""
    for R in (7.0, 7.4, 10.0):
        t = po.wall_table_tube(512, R, SGM, EPS, RHO)
        assert t[0, 1] == pytest.approx(2 * math.pi * R * _cyl_line(R), rel=1e-6)


@pytest.mark.parametrize("frac", [0.2, 0.5, 0.7])
def test_tube_wall_off_axis_energy_and_force(frac):
    """Off the axis: the table against an adaptive quadrature of the
    cylinder integral, and its force column f = -dE/dr / r against a
    central difference of that integral.  The table sums 10000 angle nodes
    over the half circle including phi = pi but not phi = 0, a first-order
    rule: relative 2e-3 on E and f (a missing factor or a wrong distance
    law is off by far more)."""
    R, nfw = 7.0, 1000
    t = po.wall_table_tube(nfw, R, SGM, EPS, RHO)
    dr = R / nfw
    i = int(frac * nfw)
    r = i * dr
    e = _cyl_energy(r, R)
    assert t[i, 1] == pytest.approx(e, rel=2e-3)
    h = 1e-4
    dedr = (_cyl_energy(r + h, R) - _cyl_energy(r - h, R)) / (2 * h)
    assert t[i, 0] == pytest.approx(-dedr / r, rel=2e-3)


def test_slit_wall_is_the_integrated_plane():
    """Slit walls: two sheets at +-H/2, each the 12-6 potential integrated
    over an infinite plane, 2 pi rho eps sigma^2 [2/5 (s/h)^10 - (s/h)^4];
    the force column is -dE/dz / z."""
    H, nfw = 12.0, 600
    t = po.wall_table_slit(nfw, H, SGM, EPS, RHO)

    def plane(h):
        f = lambda s: 2 * math.pi * s * 4 * EPS * RHO * ((SGM * SGM / (h * h + s * s)) ** 6
                                                          - (SGM * SGM / (h * h + s * s)) ** 3)
        return integrate.quad(f, 0, np.inf, epsabs=0, epsrel=1e-12, limit=400)[0]

    dz = 0.5 * H / nfw
    for i in (1, 100, 300, 500):
        z = i * dz
        e = plane(0.5 * H + z) + plane(0.5 * H - z)
        assert t[i, 1] == pytest.approx(e, rel=1e-8)
        h = 1e-5
        de = (plane(0.5 * H + z + h) + plane(0.5 * H - z - h) - plane(0.5 * H + z - h) - plane(0.5 * H - z + h)) / (2 * h)
        assert t[i, 0] == pytest.approx(-de / z, rel=1e-5)
    assert t[0, 0] == t[1, 0]


# ----------------------------------------- closed forms: pair force/virial
def _bulk(n=64, seed=5, model="TIP4P"):
    s = inp.config_bulk(seed=seed, nmol=n, model=model, ntemp=1, npress=1)
    return s


def _energy_force(o):
    o.calc_force(True, 0.0)
    sc = o.scalars()
    st = o.state()
    return sc, st


@pytest.mark.parametrize("model", ["TIP3P", "TIP4P", "TIP5P"])
def test_pair_forces_sum_to_zero_in_bulk(model):
    """Newton's third law: in a periodic box with no wall the molecular
    forces sum to zero (FP64: 1e-10 of the force scale)."""
    s = _bulk(model=model)
    o = oracle_for(s)
    sc, st = _energy_force(o)
    f = st["f"]
    assert np.abs(f.sum(axis=0)).max() <= 1e-10 * np.abs(f).max()
    assert sc["wall"] == 0.0


@pytest.mark.parametrize("model", ["TIP3P", "TIP4P", "TIP5P"])
def test_force_is_minus_energy_gradient(model):
    """Displacing one molecule's centre of mass: the switched LJ + Coulomb
    energy changes by -F.dx (central difference, relative 1e-5)."""
    s = _bulk(model=model, n=48, seed=11)
    o = oracle_for(s)
    st0 = o.state()
    L = np.array(o.scalars()["L"])
    _, st = _energy_force(o)
    rng = np.random.default_rng(3)
    for k in rng.choice(48, size=4, replace=False):
        ident = st["id"][k]
        F = st["f"][k]
        for d in range(3):
            h = 1e-5
            e = []
            for sgn in (+1, -1):
                r = st["r"].copy()
                r[k, d] += sgn * h / L[d]
                o.set_state(r, st["v"], st["q"], st["p"])
                sc, _ = _energy_force(o)
                e.append(sc["lj"] + sc["clmb"])
            g = (e[0] - e[1]) / (2 * h)
            assert -g == pytest.approx(F[d], rel=1e-5, abs=1e-6 * np.abs(st["f"]).max())
        o.set_state(st["r"], st["v"], st["q"], st["p"])
        assert ident >= 0
    del st0


def test_virial_is_box_derivative():
    """Scaling the box along axis d with fixed fractional centres of mass and
    fixed orientations: dU/dL_d * L_d = -W_d, the molecular virial
    (relative 1e-5)."""
    s = _bulk(n=150, seed=7)  # four lattice planes: every axis has pairs
    o = oracle_for(s)
    sc0, st = _energy_force(o)
    L0 = list(o.scalars()["L"])
    assert min(abs(v) for v in sc0["vir"]) > 100.0
    for d in range(3):
        h = 1e-6
        e = []
        for sgn in (+1, -1):
            L = list(L0)
            L[d] = L0[d] * (1 + sgn * h)
            o.sys.L[0], o.sys.L[1], o.sys.L[2] = L
            sc, _ = _energy_force(o)
            e.append(sc["lj"] + sc["clmb"])
        o.sys.L[0], o.sys.L[1], o.sys.L[2] = L0
        dUdlnL = (e[0] - e[1]) / (2 * h)
        W = sc0["vir"][d]  # sum over unordered pairs of F_ij,d r_ij,d
        assert dUdlnL == pytest.approx(-W, rel=1e-5)


def _switched(r, rc, rl):
    """Switching function of the pair potential, sw(r) = 1 below rl, 0 above
    rc, the quintic between (src/force_calculator.cxx:874-989)."""
    if r <= rl:
        return 1.0
    if r >= rc:
        return 0.0
    x = (r - rl) / (rc - rl)
    return 1 - 10 * x ** 3 + 15 * x ** 4 - 6 * x ** 5


def test_two_molecule_energy_follows_switch():
    """Two TIP3P molecules along z: E(r) = sw(r) * E_unswitched(r) where the
    unswitched energy is the plain site-site sum (computed here in numpy)."""
    mt = inp.parse_moltype(inp.MOLTYPE["TIP3P"])
    L = 60.0
    rc, rl = 9.0, 7.0
    q = np.zeros((2, 4))
    q[:, 0] = 1.0
    for dz in (3.0, 6.5, 7.5, 8.2, 8.9, 9.5):
        s = inp.System("TIP3P", 2, 0, 0, 0, 2.0, rc, rl, 1, 1, 1, np.array([300.0 / inp.UNIT_TEMP]),
                       np.array([0.0]))
        r = np.array([[0.5, 0.5, 0.5], [0.5, 0.5, 0.5 + dz / L]])
        s.states = [inp.ReplicaState(r=r, v=np.zeros((2, 3)), q=q.copy(), p=np.zeros((2, 4)), s=1.0, Ps=0.0, Q=1.0,
                                     Pv=(0, 0, 0), W=1.0, L=(L, L, L), ids=np.arange(2, dtype=np.int32))]
        o = oracle_for(s)
        sc, _ = _energy_force(o)
        pos = [np.array(mt.r[a]) for a in range(3)]
        e_lj = 0.0
        e_cl = 0.0
        for a in range(3):
            for b in range(3):
                d = pos[a] - (pos[b] + np.array([0, 0, dz]))
                rr = float(np.sqrt(d @ d))
                if a == 0 and b == 0:
                    sg, ep = mt.sigma[0], mt.eps[0]
                    e_lj += 4 * ep * ((sg / rr) ** 12 - (sg / rr) ** 6)
                e_cl += mt.q[a] * mt.q[b] / rr
        w = _switched(dz, rc, rl)
        assert sc["lj"] == pytest.approx(w * e_lj, rel=1e-12, abs=1e-300)
        assert sc["clmb"] == pytest.approx(w * e_cl, rel=1e-12, abs=1e-300)


# ------------------------------------------------------- integrator laws
def _nve_drift(dt_fs, steps):
    d = np.load(__import__("os").path.join(__import__("os").path.dirname(__file__), "golden", "ref_tube_chk.npz"))
    s = fixture_system(d)
    s.tconstant, s.pconstant = 0, 0
    s.dt_fs = dt_fs
    o = oracle_for(s)
    o.initialize_property()
    st0 = o.state()
    o.q_norm0 = np.empty(st0["q"].shape[0])
    o.q_norm0[st0["id"]] = np.linalg.norm(st0["q"], axis=1)
    e0 = o.scalars()["tot"]
    dev = 0.0
    for _ in range(steps):
        o.execute_steps(1)
        o.calc_properties()
        dev = max(dev, abs(o.scalars()["tot"] - e0))
    return dev, o


def test_nve_energy_error_is_second_order():
    """Velocity-Verlet-type splitting: halving dt over the same time span
    divides the energy fluctuation by about 4."""
    e2, o = _nve_drift(1.0, 100)
    e1, _ = _nve_drift(0.5, 200)
    ratio = e2 / e1
    assert 2.8 < ratio < 5.5, ratio
    # the free-rotor splitting is orthogonal on the quaternion: |q| is kept
    # to rounding (the checkpoint's own |q| differ from 1 by ~1e-7)
    q = o.state()["q"]
    ids = o.state()["id"]
    assert np.abs(np.linalg.norm(q, axis=1) - o.q_norm0[ids]).max() < 1e-13


def test_nose_poincare_andersen_hamiltonian_conserved():
    """NP thermostat + 1D Andersen barostat: the extended Hamiltonian stays
    within 1e-3 of the kinetic energy scale over 100 steps, and its error
    shrinks ~4x when dt halves."""
    d = np.load(__import__("os").path.join(__import__("os").path.dirname(__file__), "golden", "ref_tube_chk.npz"))
    devs = []
    for dt, n in ((1.0, 100), (0.5, 200)):
        s = fixture_system(d)
        s.dt_fs = dt
        o = oracle_for(s)
        o.initialize_property()
        dev = 0.0
        for _ in range(n):
            o.execute_steps(1)
            o.calc_properties()
            dev = max(dev, abs(o.scalars()["ham"]))
        devs.append(dev)
        assert dev < 1e-3 * o.scalars()["kin"]
    assert 2.8 < devs[0] / devs[1] < 5.5, devs


# ------------------------------------------------ neighbour windows, sort
@pytest.mark.parametrize("n,w,seed", [(2, 0.3, 1), (17, 0.05, 2), (144, 0.19, 3), (300, 0.02, 4), (64, 0.0, 5)])
def test_sort_window_brackets_the_neighbourhood(n, w, seed):
    """Stable ascending order; js / je are the first sorted neighbours below
    / above i (cyclically, keys on the unit circle) whose key lies further
    than w from key i, so every key within w lies strictly between them
    (src/molecules.cxx:541-585)."""
    rng = np.random.default_rng(seed)
    keys = rng.random(n)
    if n > 4:
        keys[3] = keys[1]  # a tie: stability
    order, js, je = po.sort_window(keys, w)
    assert np.array_equal(order, np.argsort(keys, kind="stable"))
    z = keys[order]
    for i in range(n):
        def gap_down(k):  # key distance to the k-th neighbour below i
            j = i - k
            return z[i] - (z[j] if j >= 0 else z[j + n] - 1.0)

        def gap_up(k):
            j = i + k
            return (z[j] if j < n else z[j - n] + 1.0) - z[i]

        kd = (i - js[i]) % n
        ku = (je[i] - i) % n
        assert kd >= 1 and ku >= 1
        assert all(gap_down(k) <= w for k in range(1, kd)) and gap_down(kd) > w
        assert all(gap_up(k) <= w for k in range(1, ku)) and gap_up(ku) > w


# --------------------------------------------------------- exchange laws
def test_exchange_matches_boltzmann_ratio():
    """Each accepted/rejected pair agrees with min(1, w) where w is the ratio
    of the NPT Boltzmann weights after and before the swap of conditions,
    exp(-sum_k beta_k (U_k + P_k V_k)), built here from first principles."""
    rng = np.random.default_rng(9)
    nt, npr = 4, 3
    n = nt * npr
    T, P = po.ladder_geometric(nt, npr, 200.0 / inp.UNIT_TEMP, 360.0 / inp.UNIT_TEMP, 0.1e6 / inp.UNIT_PRESS,
                               200e6 / inp.UNIT_PRESS)
    for step in range(8):
        pot = rng.normal(-4000, 40, n)
        vol = rng.normal(3000, 30, n)
        index = np.arange(n, dtype=np.int32)
        u = rng.random(n + 1)
        lst = po.exchange_list(nt, npr, step % 4)
        idx, T2, P2, acc, isx = po.replica_exchange(nt, npr, step, 2, pot, vol, index, T.copy(), P.copy(), u)
        Tc, Pc, ic = T.copy(), P.copy(), index.copy()
        for k, (a, b) in enumerate(lst):
            i1, i2 = ic[a], ic[b]

            def weight(Ta, Pa, Tb, Pb):
                return -(pot[i1] + Pa * vol[i1]) / Ta - (pot[i2] + Pb * vol[i2]) / Tb

            logw = weight(Tc[i2], Pc[i2], Tc[i1], Pc[i1]) - weight(Tc[i1], Pc[i1], Tc[i2], Pc[i2])
            expect = math.exp(logw) > u[k]
            assert bool(acc[k]) == expect
            if expect:
                Tc[i1], Tc[i2] = Tc[i2], Tc[i1]
                Pc[i1], Pc[i2] = Pc[i2], Pc[i1]
                ic[a], ic[b] = ic[b], ic[a]
        assert np.array_equal(T2, Tc) and np.array_equal(P2, Pc) and np.array_equal(idx, ic)
        assert sorted(T2.tolist()) == sorted(T.tolist())


def test_pair_lists_are_disjoint_neighbours():
    """Every pair list joins grid neighbours (same P and adjacent T, or the
    reverse) and uses each condition slot at most once."""
    for nt, npr in ((16, 4), (5, 3), (3, 1), (1, 4)):
        for typ in range(4):
            lst = po.exchange_list(nt, npr, typ)
            flat = lst.ravel().tolist()
            assert len(flat) == len(set(flat))
            for a, b in lst:
                ta, pa = a % nt, a // nt
                tb, pb = b % nt, b // nt
                assert (abs(ta - tb) == 1 and pa == pb) or (abs(pa - pb) == 1 and ta == tb)


# --------------------------------------------- walls on a lone molecule
@pytest.mark.parametrize("conf", [1, 2])
def test_wall_force_is_minus_wall_energy_gradient(conf):
    """One molecule, no pair partner: the whole energy is the wall's.  The
    table lookup applies f(r) on the oxygen site; its COM force equals
    -dE_wall/dx of the interpolated energy up to the table's own
    first-order consistency (relative 2e-3)."""
    if conf == 1:
        s = inp.System("TIP4P", 1, 0, 0, 1, 2.0, 9.0, 8.0, 1, 1, 1, np.array([2.0]), np.array([0.0]),
                       wall_length=7.0, nfwall=20000, sgm_wall=SGM, eps_wall=EPS, rho_wall=RHO)
        L = (44.752, 44.752, 50.0)
        r0 = np.array([[0.5 + 3.9 / L[0], 0.5 - 1.2 / L[1], 0.3]])
        axes = (0, 1)
    else:
        s = inp.System("TIP5P", 1, 0, 0, 2, 2.0, 9.0, 8.0, 1, 1, 1, np.array([2.0]), np.array([0.0]),
                       wall_length=10.0, nfwall=4000, sgm_wall=SGM, eps_wall=EPS, rho_wall=RHO)
        L = (40.0, 40.0, 50.0)
        r0 = np.array([[0.3, 0.6, 0.5 + 2.2 / L[2]]])
        axes = (2,)
    q = np.array([[0.9, 0.1, -0.3, 0.2]])
    q /= np.linalg.norm(q)
    s.states = [inp.ReplicaState(r=r0, v=np.zeros((1, 3)), q=q, p=np.zeros((1, 4)), L=L,
                                 ids=np.zeros(1, np.int32))]
    o = oracle_for(s)
    o.calc_force(True, 0.0)
    F = o.state()["f"][0]
    assert o.scalars()["wall"] != 0.0
    for d in axes:
        h = 2e-4
        e = []
        for sgn in (+1, -1):
            r = r0.copy()
            r[0, d] += sgn * h / L[d]
            o.set_state(r, np.zeros((1, 3)), q, np.zeros((1, 4)))
            o.calc_force(True, 0.0)
            e.append(o.scalars()["wall"])
        g = (e[0] - e[1]) / (2 * h)
        assert -g == pytest.approx(F[d], rel=2e-3)


# ------------------------------------------------- exchange rescale laws
def test_change_temperature_scales_kinetic_energy():
    """ChangeTemperature (src/molecules.h:333-353): after a swap T1 -> T2
    both kinetic energies scale by exactly T2/T1 and s returns to 1."""
    d = np.load(__import__("os").path.join(__import__("os").path.dirname(__file__), "golden", "ref_tube_chk.npz"))
    o = oracle_for(fixture_system(d))
    o.initialize_property()
    o.execute_steps(20)
    o.calc_properties()
    a = o.scalars()
    T1 = a["T"]
    T2 = 1.37 * T1
    o.change_temperature(T2)
    o.calc_properties()
    b = o.scalars()
    assert b["s"] == 1.0 and b["T"] == T2
    assert b["tra"] == pytest.approx(a["tra"] * T2 / T1, rel=1e-12)
    # p . P_k q cancels against the checkpoint's p . q component: 1e-8
    assert b["rot"] == pytest.approx(a["rot"] * T2 / T1, rel=1e-8)


@pytest.mark.parametrize("conf", [0, 1, 2])
def test_kill_momentum_removes_the_free_momenta(conf):
    """KillMomentum (src/molecules.cxx:168-210): the momenta along the
    periodic axes vanish; in the tube also the angular momentum about the
    tube axis."""
    if conf == 0:
        s = inp.config_bulk(seed=3, nmol=64, ntemp=1, npress=1)
        free = (0, 1, 2)
    elif conf == 1:
        s = inp.config_tube(seed=3, nmol=60, ntemp=1, npress=1)
        free = (2,)
    else:
        s = inp.config_slit(seed=3, nmol=80, ntemp=1, npress=1, Lxy=30.0)
        free = (0, 1)
    o = oracle_for(s)
    st = o.state()
    v = st["v"] + 0.3  # add a drift on every axis
    o.set_state(st["r"], v, st["q"], st["p"])
    st0r = st["r"].copy()
    v0_after_linear = v.copy()
    v0_after_linear[:, 2] -= v0_after_linear[:, 2].mean()
    o.kill_momentum()
    st = o.state()
    L = np.array(o.scalars()["L"])
    vel = st["v"] / L  # s = 1
    scale = np.abs(vel).max()
    for d in free:
        assert abs(vel[:, d].sum()) <= 1e-12 * scale * len(vel)
    if conf == 1:
        # the tube's angular momentum about the axis: each molecule further
        # than 1 A from the axis takes -<l>, those inside keep theirs
        # (src/molecules.cxx:189-210), so the residual is <l> x n_inside
        x = (st["r"] - 0.5) * L
        lz = np.sum(x[:, 0] * vel[:, 1] - x[:, 1] * vel[:, 0])
        x0 = (st0r - 0.5) * L
        v0 = v0_after_linear / L
        l_mean = np.mean(x0[:, 0] * v0[:, 1] - x0[:, 1] * v0[:, 0])
        n_in = int(np.sum(x0[:, 0] ** 2 + x0[:, 1] ** 2 <= 1.0))
        assert abs(lz - l_mean * n_in) <= 1e-10 * scale * L[0] * len(vel)


# ------------------------------------------- barostats: conserved H
@pytest.mark.parametrize("conf", [0, 2])
def test_isotropic_and_area_barostats_conserve_hamiltonian(conf):
    """NP thermostat with the 3D isotropic (bulk, pconstant 1) or 2D area
    (slit, pconstant 3) Andersen barostat: the extended Hamiltonian's
    deviation shrinks ~4x when dt halves (second-order splitting), which a
    wrong barostat kinetic term or pressure force would break."""
    if conf == 0:
        base = inp.config_bulk(seed=21, nmol=125, L=16.0, ntemp=1, npress=1)
        pmode = 1
    else:
        base = inp.config_slit(seed=22, nmol=100, Lxy=24.0, ntemp=1, npress=1, model="TIP4P")
        pmode = 3
    from helpers import relax_oracle
    base = relax_oracle(base, steps=200, dt_fs=0.5)
    devs = []
    for dt, n in ((1.0, 60), (0.5, 120)):
        s = __import__("copy").copy(base)
        s.tconstant, s.pconstant, s.dt_fs = 1, pmode, dt
        s.P = np.array([50e6 / inp.UNIT_PRESS])
        o = oracle_for(s)
        o.initialize_property()
        dev = 0.0
        L0 = list(o.scalars()["L"])
        for _ in range(n):
            o.execute_steps(1)
            o.calc_properties()
            dev = max(dev, abs(o.scalars()["ham"]))
        assert list(o.scalars()["L"]) != L0  # the box moved
        devs.append(dev)
    assert 2.8 < devs[0] / devs[1] < 5.5, devs
