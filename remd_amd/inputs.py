"""Seeded synthetic inputs and the reference's text formats.

Writes and reads the reference's checkpoint format (.chk,
src/io_manager.cxx:528-634), its .inp key/value format
(src/parameter.h:174-243), condition.dat (src/remdinfo.cxx:374-398) and
moltype blocks (src/molecules.cxx:69-131), and generates the benchmark
workloads: bulk lattice, nanotube, nano-slit, uneven tube batches and the
served job stream.  State arrays use the reference's internal
representation: r fractional in [0,1), v = velocity * L * s, p = quaternion
momentum * s (what ReadCheckPoint leaves, io_manager.cxx:577-581).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# Physical constants and reduced units (src/unit.h:3-18).
NA = 6.0221367e23
E0 = 8.8542e-12
ELE = 1.6021761e-19
KB = 1.3806488e-23
UNIT_MASS = 15.9994e-3 / NA
UNIT_LENGTH = 1.0e-10
UNIT_ENERGY = 1e3 / NA
UNIT_TIME = math.sqrt(UNIT_MASS * UNIT_LENGTH * UNIT_LENGTH / UNIT_ENERGY)
UNIT_COULOMB = math.sqrt(4.0 * math.pi * E0 * UNIT_ENERGY * UNIT_LENGTH)
UNIT_TEMP = UNIT_ENERGY / KB
UNIT_PRESS = UNIT_ENERGY / (UNIT_LENGTH ** 3)

# Water models, moltype.txt format (A, g/mol, A, kJ/mol, e); parameters as in
# the reference's moltype.txt blocks.
MOLTYPE = {
    "TIP3P": """TIP3P 3
   OW  6.556274e-02  0.0000000e+00  0.0000000e+00 1.59994e+01 3.15061e+00 6.363860e-01 -8.3400e-01
  HW1 -5.203197e-01  7.5695022e-01  0.0000000e+00 1.00800e+00 0.00000e+00 0.000000e+00  4.1700e-01
  HW2 -5.203197e-01 -7.5695022e-01  0.0000000e+00 1.00800e+00 0.00000e+00 0.000000e+00  4.1700e-01
""",
    "TIP4P": """TIP4P 4
   OW  6.556274e-02  0.0000000e+00  0.0000000e+00 1.59994e+01 3.15365e+00 6.485200e-01  0.0000e+00
  HW1 -5.203197e-01  7.5695022e-01  0.0000000e+00 1.00800e+00 0.00000e+00 0.000000e+00  5.2000e-01
  HW2 -5.203197e-01 -7.5695022e-01  0.0000000e+00 1.00800e+00 0.00000e+00 0.000000e+00  5.2000e-01
   MW -8.443726e-02  0.0000000e+00  0.0000000e+00 0.00000e+00 0.00000e+00 0.000000e+00 -1.0400e+00
""",
    "TIP5P": """TIP5P 5
   OW  6.556274e-02  0.0000000e+00  0.0000000e+00 1.59994e+01 3.12000e+00 6.694400e-01  0.0000e+00
  HW1 -5.203197e-01  7.5695022e-01  0.0000000e+00 1.00800e+00 0.00000e+00 0.000000e+00  2.4100e-01
  HW2 -5.203197e-01 -7.5695022e-01  0.0000000e+00 1.00800e+00 0.00000e+00 0.000000e+00  2.4100e-01
  LP1  0.469714e-01  0.0000000e+00  5.7154322e-01 0.00000e+00 0.00000e+00 0.000000e+00 -2.4100e-01
  LP2  0.469714e-01  0.0000000e+00 -5.7154322e-01 0.00000e+00 0.00000e+00 0.000000e+00 -2.4100e-01
""",
}
NSITE = {"TIP3P": 3, "TIP4P": 4, "TIP5P": 5}


@dataclass
class MolType:
    """One molecule type in reduced units (src/molecules.cxx:98-110)."""
    name: str
    r: np.ndarray        # (nsite, 3) body-frame site coordinates
    m: np.ndarray        # (nsite,)
    sigma: np.ndarray
    eps: np.ndarray
    q: np.ndarray
    mass: float
    inertia: np.ndarray  # (3,)

    @property
    def nsite(self) -> int:
        return self.r.shape[0]


def parse_moltype(text: str, k: int = 0) -> MolType:
    """k-th block of a moltype.txt text (src/molecules.cxx:69-131)."""
    lines = text.splitlines()
    i = 0
    found = -1
    while i < len(lines):
        tok = lines[i].split()
        i += 1
        if not tok or not tok[0][0].isalpha() or len(tok) < 2:
            continue
        name, natom = tok[0], int(tok[1])
        found += 1
        rows = []
        for _ in range(natom):
            f = lines[i].split()
            i += 1
            rows.append([float(x) for x in f[1:8]])
        if found != k:
            continue
        a = np.array(rows)
        r = a[:, 0:3] * (1e-10 / UNIT_LENGTH)
        m = a[:, 3] * (1e-3 / NA / UNIT_MASS)
        s = a[:, 4] * (1e-10 / UNIT_LENGTH)
        e = a[:, 5] * (1e3 / NA / UNIT_ENERGY)
        q = a[:, 6] * (ELE / UNIT_COULOMB)
        x, y, z = r[:, 0], r[:, 1], r[:, 2]
        inertia = np.array([
            float(np.sum(m * (y * y + z * z))),
            float(np.sum(m * (z * z + x * x))),
            float(np.sum(m * (x * x + y * y))),
        ])
        return MolType(name, r, m, s, e, q, float(np.sum(m)), inertia)
    raise ValueError(f"moltype block {k} not found")


def molecule(name: str) -> MolType:
    return parse_moltype(MOLTYPE[name])


# ------------------------------------------------------------------ state
@dataclass
class ReplicaState:
    r: np.ndarray            # (n,3) fractional
    v: np.ndarray            # (n,3) velocity * L * s
    q: np.ndarray            # (n,4)
    p: np.ndarray            # (n,4) momentum * s
    s: float = 1.0
    Ps: float = 0.0
    Q: float = 0.5
    Pv: tuple = (0.0, 0.0, 0.0)
    W: float = 0.05
    L: tuple = (0.0, 0.0, 0.0)
    ids: np.ndarray | None = None

    def scal(self):
        return np.array([self.s, self.Ps, self.Q, self.Pv[0], self.Pv[1], self.Pv[2], self.W,
                         self.L[0], self.L[1], self.L[2]])


@dataclass
class System:
    """One REMD job: model, ensemble, parameters and per-replica states."""
    model: str
    nmol: int
    tconstant: int
    pconstant: int
    confined: int
    dt_fs: float
    rcut: float
    rswitch: float
    interval: int
    ntemp: int
    npress: int
    T: np.ndarray            # reduced, per replica (condition slot == replica at start)
    P: np.ndarray            # reduced
    states: list = field(default_factory=list)
    wall_length: float = 0.0
    nfwall: int = 0
    sgm_wall: float = 3.19
    eps_wall: float = 0.3135
    rho_wall: float = 0.3816
    sort_interval: int = 10
    margin: float = 0.5

    @property
    def nreplica(self) -> int:
        return self.ntemp * self.npress

    @property
    def nsite(self) -> int:
        return NSITE[self.model]

    @property
    def dt(self) -> float:
        """Reduced time step (src/molecules.h:92)."""
        return self.dt_fs * 1e-15 / UNIT_TIME

    def desc(self) -> dict:
        return dict(nmol=self.nmol, nsite=self.nsite, tconstant=self.tconstant,
                    pconstant=self.pconstant, confined=self.confined, dt=self.dt,
                    dt_fs=self.dt_fs, rcut=self.rcut, rswitch=self.rswitch,
                    interval=self.interval, wall_length=self.wall_length, nfwall=self.nfwall,
                    sgm_wall=self.sgm_wall, eps_wall=self.eps_wall, rho_wall=self.rho_wall,
                    sort_interval=self.sort_interval, margin=self.margin)


# ---------------------------------------------------------------- sampling
def random_quaternions(rng: np.random.RandomState, n: int) -> np.ndarray:
    """Uniform unit quaternions (Shoemake, Graphics Gems III, 1992)."""
    u1, u2, u3 = rng.random_sample(n), rng.random_sample(n), rng.random_sample(n)
    a = np.sqrt(1.0 - u1)
    b = np.sqrt(u1)
    q = np.stack([a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2),
                  b * np.sin(2 * np.pi * u3), b * np.cos(2 * np.pi * u3)], axis=1)
    return q


def smat(q):
    """S(q) of src/integrator.h:303-310 for an array of quaternions."""
    q0, q1, q2, q3 = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    return np.stack([
        np.stack([q0, -q1, -q2, -q3], 1),
        np.stack([q1, q0, -q3, q2], 1),
        np.stack([q2, q3, q0, -q1], 1),
        np.stack([q3, -q2, q1, q0], 1)], 1)


def thermal_momenta(rng, mt: MolType, q, T_red, L):
    """Maxwell-Boltzmann translational velocities and body angular
    velocities at reduced temperature T; returned in the internal
    representation (s = 1): v_int = v * L, p = 2 S(q) (0, I w)."""
    n = q.shape[0]
    v = rng.normal(0.0, math.sqrt(T_red / mt.mass), size=(n, 3)) * np.asarray(L)[None, :]
    w = rng.normal(0.0, 1.0, size=(n, 3)) * np.sqrt(T_red / mt.inertia)[None, :]
    L4 = np.zeros((n, 4))
    L4[:, 1:] = w * mt.inertia[None, :]
    p = 2.0 * np.einsum("nij,nj->ni", smat(q), L4)
    return v, p


def place_tube(rng, n, r_acc, Lz, Lxy, dmin=2.6, max_tries=2_000_000):
    """Random non-overlapping COM positions inside r < r_acc of a tube along
    z (periodic Lz), centred in an Lxy x Lxy box; min COM-COM distance dmin
    (minimum image along z).  Returns fractional coordinates."""
    pos = np.zeros((n, 3))
    k = 0
    tries = 0
    d2 = dmin * dmin
    while k < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError(f"tube placement failed at {k}/{n} (dmin={dmin})")
        rr = r_acc * math.sqrt(rng.random_sample())
        th = 2 * math.pi * rng.random_sample()
        c = np.array([rr * math.cos(th), rr * math.sin(th), (rng.random_sample() - 0.5) * Lz])
        if k:
            d = pos[:k] - c
            d[:, 2] -= Lz * np.rint(d[:, 2] / Lz)
            if np.min(np.sum(d * d, axis=1)) < d2:
                continue
        pos[k] = c
        k += 1
    frac = np.empty_like(pos)
    frac[:, 0] = pos[:, 0] / Lxy + 0.5
    frac[:, 1] = pos[:, 1] / Lxy + 0.5
    frac[:, 2] = pos[:, 2] / Lz + 0.5
    return np.mod(frac, 1.0)


def place_slit(rng, n, Lx, Ly, h_acc, Lz, dmin=2.6, max_tries=4_000_000):
    """Random non-overlapping COMs, periodic in x,y, |z| < h_acc/2 about the
    slit mid-plane of a box of height Lz."""
    pos = np.zeros((n, 3))
    k = 0
    tries = 0
    d2 = dmin * dmin
    while k < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError(f"slit placement failed at {k}/{n}")
        c = np.array([(rng.random_sample() - 0.5) * Lx, (rng.random_sample() - 0.5) * Ly,
                      (rng.random_sample() - 0.5) * h_acc])
        if k:
            d = pos[:k] - c
            d[:, 0] -= Lx * np.rint(d[:, 0] / Lx)
            d[:, 1] -= Ly * np.rint(d[:, 1] / Ly)
            if np.min(np.sum(d * d, axis=1)) < d2:
                continue
        pos[k] = c
        k += 1
    frac = np.empty_like(pos)
    frac[:, 0] = pos[:, 0] / Lx + 0.5
    frac[:, 1] = pos[:, 1] / Ly + 0.5
    frac[:, 2] = pos[:, 2] / Lz + 0.5
    return np.mod(frac, 1.0)


def place_lattice(n, L, m=None):
    """Simple-cubic m^3 lattice of spacing L/m, first n sites in z-y-x order."""
    if m is None:
        m = int(math.ceil(n ** (1.0 / 3.0) - 1e-9))
    pts = []
    for z in range(m):
        for y in range(m):
            for x in range(m):
                pts.append(((x + 0.5) / m, (y + 0.5) / m, (z + 0.5) / m))
    return np.array(pts[:n])


def _states(rng, mt, r_frac, T_list, L, Q, W):
    states = []
    for T in T_list:
        q = random_quaternions(rng, r_frac.shape[0])
        v, p = thermal_momenta(rng, mt, q, T, L)
        states.append(ReplicaState(r=r_frac.copy(), v=v, q=q, p=p, Q=Q, W=W, L=tuple(L),
                                   ids=np.arange(r_frac.shape[0], dtype=np.int32)))
    return states


def geometric(lo, hi, n):
    if n == 1:
        return np.array([lo])
    return lo * (hi / lo) ** (np.arange(n) / (n - 1))


def make_ladder(ntemp, npress, T_K, P_Pa):
    """Condition slot ind = i + j*ntemp (src/remdinfo.cxx:351-357)."""
    T = np.zeros(ntemp * npress)
    P = np.zeros(ntemp * npress)
    for j in range(npress):
        for i in range(ntemp):
            T[i + j * ntemp] = T_K[i] / UNIT_TEMP
            P[i + j * ntemp] = P_Pa[j] / UNIT_PRESS
    return T, P


# ---------------------------------------------------------------- configs
def config_bulk(seed=1, nmol=256, L=19.7, ntemp=4, npress=1, model="TIP4P"):
    """Plain REMD bulk: TIP4P on a 7^3 lattice truncated to 256, cube 19.7 A,
    NP thermostat, no barostat, 4 replicas geometric 260-320 K."""
    rng = np.random.RandomState(seed)
    mt = molecule(model)
    T, P = make_ladder(ntemp, npress, geometric(260.0, 320.0, ntemp), [0.1e6] * npress)
    r = place_lattice(nmol, L, m=7)
    sys_ = System(model, nmol, 1, 0, 0, 2.0, 9.0, 8.0, 100, ntemp, npress, T, P)
    sys_.states = _states(rng, mt, r, T, (L, L, L), 0.5, 0.05)
    return sys_


def config_tube(seed=2, nmol=144, R=7.0, Lz=50.0, Lxy=44.752, ntemp=16, npress=4,
                model="TIP4P", nfwall=1000, r_acc=None, dmin=2.6):
    """Nanotube confinement (confined=1), 1D Andersen barostat along z
    (pconstant=2); T geometric 200-360 K x P log-spaced 0.1-200 MPa."""
    rng = np.random.RandomState(seed)
    mt = molecule(model)
    T_K = geometric(200.0, 360.0, ntemp)
    P_Pa = geometric(0.1e6, 200e6, npress) if npress > 1 else [0.1e6]
    T, P = make_ladder(ntemp, npress, T_K, P_Pa)
    if r_acc is None:
        # COMs inside the attractive part of the wall: with sigma_wall =
        # 3.277 A the wall energy is already ~+500 kJ/mol at R - 2 A
        r_acc = R - 2.8
    r = place_tube(rng, nmol, r_acc, Lz, Lxy, dmin=dmin)
    sys_ = System(model, nmol, 1, 2, 1, 2.0, 9.0, 8.0, 100, ntemp, npress, T, P,
                  wall_length=R, nfwall=nfwall, sgm_wall=3.277, eps_wall=0.3886, rho_wall=0.38016)
    sys_.states = _states(rng, mt, r, T, (Lxy, Lxy, Lz), 0.5, 0.05)
    return sys_


def config_slit(seed=3, nmol=512, H=10.0, Lxy=40.0, ntemp=32, npress=4, model="TIP5P",
                nfwall=1000, dmin=2.5):
    """Nano-slit (confined=2), 2D area barostat (pconstant=3); 32 T 220-340 K
    x 4 P 0.1-100 MPa."""
    rng = np.random.RandomState(seed)
    mt = molecule(model)
    T_K = geometric(220.0, 340.0, ntemp)
    P_Pa = geometric(0.1e6, 100e6, npress) if npress > 1 else [0.1e6]
    T, P = make_ladder(ntemp, npress, T_K, P_Pa)
    Lz = H + 40.0
    r = place_slit(rng, nmol, Lxy, Lxy, H - 3.4, Lz, dmin=dmin)
    sys_ = System(model, nmol, 1, 3, 2, 2.0, 9.0, 8.0, 100, ntemp, npress, T, P,
                  wall_length=H, nfwall=nfwall, sgm_wall=3.277, eps_wall=0.3886, rho_wall=0.38016)
    sys_.states = _states(rng, mt, r, T, (Lxy, Lxy, Lz), 0.5, 0.05)
    return sys_


def config_uneven_batch(seed=4, radii=(5, 6, 7, 8, 9, 10, 11, 12), Lz=40.0, ntemp=8, npress=4):
    """Eight nanotube systems R = 5..12 A, Lz = 40 A, nmol at 1 g/cm^3 of the
    accessible volume pi (R-1.7)^2 Lz, TIP3P/TIP4P alternating, 32 replicas
    each (8 T geometric 240-340 K x 4 P 0.1-150 MPa)."""
    systems = []
    for k, R in enumerate(radii):
        model = "TIP3P" if k % 2 == 0 else "TIP4P"
        vol = math.pi * (R - 1.7) ** 2 * Lz
        nmol = int(round(vol / 29.9))
        rng = np.random.RandomState(seed * 1000 + k)
        mt = molecule(model)
        T, P = make_ladder(ntemp, npress, geometric(240.0, 340.0, ntemp), geometric(0.1e6, 150e6, npress))
        Lxy = 2.0 * (R + 9.0) + 4.0
        r = place_tube(rng, nmol, R - 2.8, Lz, Lxy, dmin=2.5)
        s = System(model, nmol, 1, 2, 1, 2.0, 9.0, 8.0, 100, ntemp, npress, T, P,
                   wall_length=float(R), nfwall=1000, sgm_wall=3.277, eps_wall=0.3886,
                   rho_wall=0.38016)
        s.states = _states(rng, mt, r, T, (Lxy, Lxy, Lz), 0.5, 0.05)
        systems.append(s)
    return systems


def served_job(rng: np.random.RandomState, nmol_lo=64, nmol_hi=2048, build_states=True):
    """One job of the served stream: nmol log-uniform 64-2048, model uniform
    TIP3P/4P/5P, confinement uniform bulk / tube (R 5-15 A) / slit (8-20 A),
    nreplica in {8,16,32,64,128}, geometric T ladder 200-380 K."""
    nmol = int(round(math.exp(rng.uniform(math.log(nmol_lo), math.log(nmol_hi)))))
    model = ["TIP3P", "TIP4P", "TIP5P"][rng.randint(3)]
    conf = rng.randint(3)
    nrep = [8, 16, 32, 64, 128][rng.randint(5)]
    T, P = make_ladder(nrep, 1, geometric(200.0, 380.0, nrep), [0.1e6])
    mt = molecule(model)
    if conf == 0:
        # bulk cube at 1 g/cm^3; the minimum image needs L > 2 (rcut + margin)
        nmol = max(nmol, 235)
        L = (nmol * 29.9) ** (1.0 / 3.0)
        s = System(model, nmol, 1, 1, 0, 2.0, 9.0, 8.0, 100, nrep, 1, T, P)
        r = place_lattice(nmol, L)
        box = (L, L, L)
    elif conf == 1:
        R = rng.uniform(5.0, 15.0)
        Lz = max(nmol * 29.9 / (math.pi * (R - 1.7) ** 2), 2 * 9.5 + 1.0)
        nmol = max(8, int(round(math.pi * (R - 1.7) ** 2 * Lz / 29.9)))
        Lxy = 2.0 * (R + 9.0) + 4.0
        s = System(model, nmol, 1, 2, 1, 2.0, 9.0, 8.0, 100, nrep, 1, T, P, wall_length=R,
                   nfwall=1000, sgm_wall=3.277, eps_wall=0.3886, rho_wall=0.38016)
        r = place_tube(rng, nmol, R - 1.7, Lz, Lxy, dmin=2.4) if build_states else None
        box = (Lxy, Lxy, Lz)
    else:
        H = rng.uniform(8.0, 20.0)
        Lxy = max(math.sqrt(nmol * 29.9 / (H - 3.4)), 2 * 9.5 + 1.0)
        nmol = max(8, int(round(Lxy * Lxy * (H - 3.4) / 29.9)))
        s = System(model, nmol, 1, 3, 2, 2.0, 9.0, 8.0, 100, nrep, 1, T, P, wall_length=H,
                   nfwall=1000, sgm_wall=3.277, eps_wall=0.3886, rho_wall=0.38016)
        r = place_slit(rng, nmol, Lxy, Lxy, H - 3.4, H + 40.0, dmin=2.4) if build_states else None
        box = (Lxy, Lxy, H + 40.0)
    if build_states and r is not None:
        s.states = _states(rng, mt, r, T, box, 0.5, 0.05)
    return s


# -------------------------------------------------------------- formats
def write_chk(path, st: ReplicaState, mt: MolType, prop_line="0"):
    """IOManager::WriteCheckPoint text format (src/io_manager.cxx:593-634)."""
    L = np.asarray(st.L)
    n = st.r.shape[0]
    with open(path, "w") as f:
        f.write(f"{n} {n * mt.nsite}\n")
        for i in range(n):
            rr = st.r[i] * L
            vv = st.v[i] / (L * st.s)
            pp = st.p[i] / st.s
            vals = " ".join(f"{x:.15e}" for x in (*rr, *vv, *st.q[i], *pp))
            ii = " ".join(f"{x:.15e}" for x in mt.inertia)
            f.write(f" {vals} {mt.mass:.15e} {ii} 0 {i * mt.nsite}\n")
        f.write(f"{st.s:.15e} {st.Ps:.15e} {st.Q:.15e}\n")
        f.write(" ".join(f"{x:.15e}" for x in st.Pv) + f" {st.W:.15e}\n")
        f.write(" ".join(f"{x:.15e}" for x in L) + "\n")
        f.write(prop_line + "\n")


def read_chk(path) -> ReplicaState:
    """IOManager::ReadCheckPoint (src/io_manager.cxx:528-591): returns the
    internal representation (r /= L, v *= L*s, p *= s)."""
    tok = open(path).read().split()
    it = iter(tok)
    n = int(next(it))
    _natom = int(next(it))
    r = np.zeros((n, 3)); v = np.zeros((n, 3)); q = np.zeros((n, 4)); p = np.zeros((n, 4))
    for i in range(n):
        r[i] = [float(next(it)) for _ in range(3)]
        v[i] = [float(next(it)) for _ in range(3)]
        q[i] = [float(next(it)) for _ in range(4)]
        p[i] = [float(next(it)) for _ in range(4)]
        _m = float(next(it))
        _I = [float(next(it)) for _ in range(3)]
        _type = int(next(it))
        _id = int(next(it))
    s, Ps, Q = (float(next(it)) for _ in range(3))
    Pv = tuple(float(next(it)) for _ in range(3))
    W = float(next(it))
    L = np.array([float(next(it)) for _ in range(3)])
    r = r / L
    v = v * (L * s)
    p = p * s
    return ReplicaState(r=r, v=v, q=q, p=p, s=s, Ps=Ps, Q=Q, Pv=Pv, W=W, L=tuple(L),
                        ids=np.arange(n, dtype=np.int32))


def read_inp(text: str) -> dict:
    """Parameter::read (src/parameter.h:174-243): `key value` lines; lines
    whose first token does not start with a letter are skipped, unknown keys
    are ignored by the caller."""
    out = {}
    for line in text.splitlines():
        tok = line.split()
        if not tok or not tok[0][0].isalpha():
            continue
        if len(tok) >= 2:
            out[tok[0]] = tok[1]
    return out
