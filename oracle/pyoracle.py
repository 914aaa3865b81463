"""ctypes bindings for the CPU oracle (liboracle.so) and, where it was built
in this container, the reference's own CPU units (oracle/_ref).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker -- never by the product.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_F32 = os.path.join(HERE, "_ref", "libremdref_f32.so")
REF_F64 = os.path.join(HERE, "_ref", "libremdref_f64.so")

MAXSITE = 5


class MolType(C.Structure):
    _fields_ = [
        ("nsite", C.c_int),
        ("r", (C.c_double * 3) * MAXSITE),
        ("m", C.c_double * MAXSITE),
        ("sigma", C.c_double * MAXSITE),
        ("eps", C.c_double * MAXSITE),
        ("q", C.c_double * MAXSITE),
        ("mass", C.c_double),
        ("inertia", C.c_double * 3),
    ]


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class System(C.Structure):
    _fields_ = [
        ("nmol", C.c_int),
        ("tmode", C.c_int), ("pmode", C.c_int), ("cmode", C.c_int),
        ("dt", C.c_double), ("dthalf", C.c_double),
        ("rcut", C.c_double), ("rswitch", C.c_double),
        ("wall_length", C.c_double),
        ("nfwall", C.c_int),
        ("fwall", _dp),
        ("sort_interval", C.c_int),
        ("margin", C.c_double),
        ("fp32", C.c_int),
        ("mt", MolType),
        ("r", _dp), ("v", _dp), ("f", _dp), ("q", _dp), ("p", _dp), ("n", _dp),
        ("id", _ip),
        ("site_f", _dp),
        ("s", C.c_double), ("Ps", C.c_double), ("Q", C.c_double),
        ("Pv", C.c_double * 3), ("W", C.c_double),
        ("L", C.c_double * 3),
        ("T", C.c_double), ("P", C.c_double),
        ("gkT", C.c_double), ("H0", C.c_double),
        ("lj", C.c_double), ("clmb", C.c_double), ("wall", C.c_double),
        ("pot", C.c_double), ("vir", C.c_double * 3),
        ("tra", C.c_double), ("trav", C.c_double * 3), ("rot", C.c_double),
        ("kin", C.c_double), ("tsm", C.c_double), ("tsp", C.c_double),
        ("tst", C.c_double), ("bst", C.c_double), ("tot", C.c_double),
        ("ham", C.c_double), ("Tk", C.c_double), ("prs", C.c_double * 4),
        ("jstart", _ip), ("jend", _ip),
        ("sort_axis", C.c_int),
        ("wfull", _ip),
    ]


_lib = None


def lib():
    """Load liboracle.so, building it with make on first use."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    L = C.CDLL(LIB)
    L.orc_system_alloc.argtypes = [C.POINTER(System), C.c_int]
    L.orc_system_free.argtypes = [C.POINTER(System)]
    for name in ("orc_calc_properties", "orc_initialize_property", "orc_kill_momentum",
                 "orc_kill_angular_momentum_tube", "orc_velocity_scaling",
                 "orc_ang_velocity_scaling", "orc_update_hamiltonian"):
        getattr(L, name).argtypes = [C.POINTER(System)]
    L.orc_calc_force.argtypes = [C.POINTER(System), C.c_int, C.c_double]
    L.orc_step.argtypes = [C.POINTER(System), C.c_int, C.c_double]
    L.orc_execute_steps.argtypes = [C.POINTER(System), C.c_int]
    L.orc_set_temperature.argtypes = [C.POINTER(System), C.c_double]
    L.orc_change_temperature.argtypes = [C.POINTER(System), C.c_double]
    L.orc_get_average.argtypes = [C.POINTER(System), _dp]
    L.orc_volume.argtypes = [C.POINTER(System)]
    L.orc_volume.restype = C.c_double
    L.orc_bottom_area.argtypes = [C.POINTER(System)]
    L.orc_bottom_area.restype = C.c_double
    L.orc_units.argtypes = [_dp]
    L.orc_moltype_parse.argtypes = [C.c_char_p, C.c_int, C.POINTER(MolType)]
    L.orc_sort_window.argtypes = [_dp, C.c_int, C.c_double, _ip, _ip, _ip]
    L.orc_window.argtypes = [_dp, C.c_int, C.c_double, _ip, _ip]
    L.orc_wall_table_tube.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, _dp]
    L.orc_wall_table_slit.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, _dp]
    L.orc_exchange_list.argtypes = [C.c_int, C.c_int, C.c_int, _ip, _ip]
    L.orc_exchange_list.restype = C.c_int
    L.orc_ladder_geometric.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                       C.c_double, _dp, _dp]
    L.orc_replica_exchange.argtypes = [C.c_int, C.c_int, C.c_long, C.c_int, _dp, _dp, _ip, _dp, _dp,
                                       C.POINTER(C.c_ubyte), _dp, _ip]
    L.orc_replica_exchange.restype = C.c_int
    L.orc_srand.argtypes = [C.c_uint]
    L.orc_rand.restype = C.c_int
    _lib = L
    return L


def dptr(a):
    return a.ctypes.data_as(_dp)


def iptr(a):
    return a.ctypes.data_as(_ip)


def units():
    out = np.zeros(9)
    lib().orc_units(dptr(out))
    keys = ["mass", "length", "energy", "time", "coulomb", "temp", "press", "Na", "ele"]
    return dict(zip(keys, out.tolist()))


def parse_moltype(text: str, k: int = 0) -> MolType:
    mt = MolType()
    rc = lib().orc_moltype_parse(text.encode(), k, C.byref(mt))
    if rc != 0:
        raise ValueError(f"moltype parse failed ({rc})")
    return mt


class OracleSystem:
    """Owns an orc_system plus the numpy views of its state arrays."""

    def __init__(self, desc, mt: MolType, r, v, q, p, scal, T, P, fp32=False,
                 fwall=None):
        n = int(desc["nmol"])
        self.sys = System()
        if lib().orc_system_alloc(C.byref(self.sys), n) != 0:
            raise MemoryError
        s = self.sys
        s.tmode, s.pmode, s.cmode = int(desc["tconstant"]), int(desc["pconstant"]), int(desc["confined"])
        s.dt = float(desc["dt"])
        s.dthalf = float(desc["dt"]) * 0.5
        s.rcut, s.rswitch = float(desc["rcut"]), float(desc["rswitch"])
        s.wall_length = float(desc.get("wall_length", 0.0))
        s.sort_interval = int(desc.get("sort_interval", 10))
        s.margin = float(desc.get("margin", 0.5))
        s.sort_axis = 0 if s.cmode == 2 else 2
        s.fp32 = 1 if fp32 else 0
        s.mt = mt
        self._fwall = None
        if fwall is not None:
            self._fwall = np.ascontiguousarray(fwall, dtype=np.float64)
            s.nfwall = self._fwall.shape[0]
            s.fwall = dptr(self._fwall)
        self.n = n
        self.set_state(r, v, q, p)
        (s.s, s.Ps, s.Q, s.Pv[0], s.Pv[1], s.Pv[2], s.W, s.L[0], s.L[1], s.L[2]) = [float(x) for x in scal]
        s.P = float(P)
        lib().orc_set_temperature(C.byref(self.sys), float(T))

    def __del__(self):
        try:
            lib().orc_system_free(C.byref(self.sys))
        except Exception:
            pass

    def _arr(self, name, width, dtype=np.float64):
        ptr = getattr(self.sys, name)
        ct = C.c_double if dtype == np.float64 else C.c_int
        buf = (ct * (width * self.n)).from_address(C.addressof(ptr.contents))
        return np.frombuffer(buf, dtype=dtype).reshape(self.n, width)

    def set_state(self, r, v, q, p):
        self._arr("r", 3)[:] = r
        self._arr("v", 3)[:] = v
        self._arr("q", 4)[:] = q
        self._arr("p", 4)[:] = p

    def state(self):
        out = {k: self._arr(k, w).copy() for k, w in (("r", 3), ("v", 3), ("f", 3), ("q", 4), ("p", 4), ("n", 4))}
        out["id"] = self._arr("id", 1, np.int32)[:, 0].copy()
        out["site_f"] = np.frombuffer(
            (C.c_double * (3 * MAXSITE * self.n)).from_address(C.addressof(self.sys.site_f.contents)),
            dtype=np.float64)[: 3 * self.sys.mt.nsite * self.n].reshape(self.n, self.sys.mt.nsite, 3).copy()
        out["jstart"] = self._arr("jstart", 1, np.int32)[:, 0].copy()
        out["jend"] = self._arr("jend", 1, np.int32)[:, 0].copy()
        out["wfull"] = self._arr("wfull", 1, np.int32)[:, 0].copy()
        return out

    def scalars(self):
        s = self.sys
        return dict(s=s.s, Ps=s.Ps, Q=s.Q, Pv=list(s.Pv), W=s.W, L=list(s.L), T=s.T, P=s.P,
                    gkT=s.gkT, H0=s.H0, lj=s.lj, clmb=s.clmb, wall=s.wall, pot=s.pot,
                    vir=list(s.vir), tra=s.tra, rot=s.rot, kin=s.kin, tot=s.tot, ham=s.ham,
                    Tk=s.Tk, prs=list(s.prs))

    def average(self):
        out = np.zeros(19)
        lib().orc_get_average(C.byref(self.sys), dptr(out))
        return out

    # thin method wrappers
    def initialize_property(self):
        lib().orc_initialize_property(C.byref(self.sys))

    def calc_force(self, do_sort=True, margin=0.0):
        lib().orc_calc_force(C.byref(self.sys), int(do_sort), float(margin))

    def execute_steps(self, nsteps):
        lib().orc_execute_steps(C.byref(self.sys), int(nsteps))

    def calc_properties(self):
        lib().orc_calc_properties(C.byref(self.sys))

    def velocity_scaling(self):
        lib().orc_velocity_scaling(C.byref(self.sys))

    def ang_velocity_scaling(self):
        lib().orc_ang_velocity_scaling(C.byref(self.sys))

    def kill_momentum(self):
        lib().orc_kill_momentum(C.byref(self.sys))

    def change_temperature(self, T):
        lib().orc_change_temperature(C.byref(self.sys), float(T))

    def update_hamiltonian(self):
        lib().orc_update_hamiltonian(C.byref(self.sys))

    def volume(self):
        return lib().orc_volume(C.byref(self.sys))


def sort_window(keys, w):
    keys = np.ascontiguousarray(keys, dtype=np.float64)
    n = keys.shape[0]
    order = np.zeros(n, np.int32)
    js = np.zeros(n, np.int32)
    je = np.zeros(n, np.int32)
    lib().orc_sort_window(dptr(keys), n, float(w), iptr(order), iptr(js), iptr(je))
    return order, js, je


def wall_table_tube(nfwall, R, sgm, eps, rho):
    out = np.zeros((nfwall, 2))
    lib().orc_wall_table_tube(int(nfwall), float(R), float(sgm), float(eps), float(rho), dptr(out))
    return out


def wall_table_slit(nfwall, H, sgm, eps, rho):
    out = np.zeros((nfwall, 2))
    lib().orc_wall_table_slit(int(nfwall), float(H), float(sgm), float(eps), float(rho), dptr(out))
    return out


def exchange_list(ntemp, npress, typ):
    a = np.zeros(ntemp * npress + 1, np.int32)
    b = np.zeros(ntemp * npress + 1, np.int32)
    n = lib().orc_exchange_list(ntemp, npress, typ, iptr(a), iptr(b))
    return np.stack([a[:n], b[:n]], axis=1)


def ladder_geometric(ntemp, npress, tmin, tmax, pmin, pmax):
    T = np.zeros(ntemp * npress)
    P = np.zeros(ntemp * npress)
    lib().orc_ladder_geometric(ntemp, npress, tmin, tmax, pmin, pmax, dptr(T), dptr(P))
    return T, P


def replica_exchange(ntemp, npress, step, pmode, pot, vol, index, T, P, uniforms=None):
    """One exchange round; mutates and returns (index, T, P, accepted)."""
    n = ntemp * npress
    pot = np.ascontiguousarray(pot, np.float64)
    vol = np.ascontiguousarray(vol, np.float64)
    index = np.ascontiguousarray(index, np.int32).copy()
    T = np.ascontiguousarray(T, np.float64).copy()
    P = np.ascontiguousarray(P, np.float64).copy()
    isx = np.zeros(n, np.uint8)
    acc = np.zeros(n + 1, np.int32)
    u = None if uniforms is None else np.ascontiguousarray(uniforms, np.float64)
    lib().orc_replica_exchange(ntemp, npress, int(step), int(pmode), dptr(pot), dptr(vol), iptr(index),
                               dptr(T), dptr(P), isx.ctypes.data_as(C.POINTER(C.c_ubyte)),
                               None if u is None else dptr(u), iptr(acc))
    return index, T, P, acc, isx


# ------------------------------------------------------------------ _ref
class RefLib:
    """The reference's own CPU units (FOUR_SITE tuned kernel) via
    ref_harness.cpp.  Only present where oracle/Makefile's `ref` target was
    built from /root/reference."""

    def __init__(self, fp32: bool):
        path = REF_F32 if fp32 else REF_F64
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = C.CDLL(path)
        L.ref_create.restype = C.c_void_p
        L.ref_create.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                 C.c_int, C.c_double, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int,
                                 C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                 C.c_double]
        vp = [C.c_void_p]
        for name in ("ref_destroy", "ref_initialize_property", "ref_execute_steps",
                     "ref_calc_properties", "ref_update_hamiltonian", "ref_velocity_scaling",
                     "ref_ang_velocity_scaling", "ref_kill_momentum"):
            getattr(L, name).argtypes = vp
        L.ref_set_state.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _ip, _dp]
        L.ref_get_state.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp, _ip, _dp]
        L.ref_get_site_forces.argtypes = [C.c_void_p, _dp]
        L.ref_set_temperature.argtypes = [C.c_void_p, C.c_double]
        L.ref_set_pressure.argtypes = [C.c_void_p, C.c_double]
        L.ref_change_temperature.argtypes = [C.c_void_p, C.c_double]
        L.ref_calc_force.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.ref_get_props.argtypes = [C.c_void_p, _dp]
        L.ref_get_average.argtypes = [C.c_void_p, _dp]
        L.ref_get_window.argtypes = [C.c_void_p, _ip, _ip]
        L.ref_nfwall.argtypes = vp
        L.ref_wall_table.argtypes = [C.c_void_p, _dp]
        L.ref_moltype.argtypes = [C.c_void_p, _dp, _dp]
        L.ref_dt.argtypes = vp
        L.ref_dt.restype = C.c_double
        L.ref_remdinfo.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                   _dp, _dp, _ip, C.c_int, _ip]
        self.L = L


class RefSystem:
    def __init__(self, reflib: RefLib, mtl_path, desc, r, v, q, p, ids, scal, T, P,
                 T_K=300.0, P_Pa=1.0):
        L = reflib.L
        self.L = L
        self.n = int(desc["nmol"])
        self.nsite = int(desc["nsite"])
        devnull = os.open(os.devnull, os.O_WRONLY)
        saved = os.dup(1)
        os.dup2(devnull, 1)
        try:
            self.h = L.ref_create(mtl_path.encode(), self.n, self.nsite, float(desc["dt_fs"]),
                                  float(desc["rcut"]), float(desc["rswitch"]), int(desc["tconstant"]),
                                  float(scal[2]), int(desc["pconstant"]), float(scal[6]),
                                  int(desc.get("interval", 1)), int(desc["confined"]),
                                  int(desc.get("nfwall", 0)), float(desc.get("wall_length", 0.0)),
                                  float(desc.get("sgm_wall", 3.19)), float(desc.get("eps_wall", 0.3135)),
                                  float(desc.get("rho_wall", 0.3816)), float(T_K), float(P_Pa))
        finally:
            os.dup2(saved, 1)
            os.close(saved)
            os.close(devnull)
        r = np.ascontiguousarray(r, np.float64)
        v = np.ascontiguousarray(v, np.float64)
        q = np.ascontiguousarray(q, np.float64)
        p = np.ascontiguousarray(p, np.float64)
        ids = np.ascontiguousarray(ids, np.int32)
        sc = np.ascontiguousarray(scal, np.float64)
        L.ref_set_state(self.h, dptr(r), dptr(v), dptr(q), dptr(p), iptr(ids), dptr(sc))
        L.ref_set_temperature(self.h, float(T))
        L.ref_set_pressure(self.h, float(P))

    def quiet(self, fn, *args):
        devnull = os.open(os.devnull, os.O_WRONLY)
        saved = os.dup(1)
        os.dup2(devnull, 1)
        try:
            return fn(self.h, *args)
        finally:
            os.dup2(saved, 1)
            os.close(saved)
            os.close(devnull)

    def state(self):
        n = self.n
        r = np.zeros((n, 3)); v = np.zeros((n, 3)); f = np.zeros((n, 3))
        q = np.zeros((n, 4)); p = np.zeros((n, 4)); nn = np.zeros((n, 4))
        ids = np.zeros(n, np.int32); sc = np.zeros(10)
        self.L.ref_get_state(self.h, dptr(r), dptr(v), dptr(f), dptr(q), dptr(p), dptr(nn), iptr(ids), dptr(sc))
        af = np.zeros((n * self.nsite, 3))
        self.L.ref_get_site_forces(self.h, dptr(af))
        js = np.zeros(n, np.int32); je = np.zeros(n, np.int32)
        self.L.ref_get_window(self.h, iptr(js), iptr(je))
        return dict(r=r, v=v, f=f, q=q, p=p, n=nn, id=ids, scal=sc,
                    site_f=af.reshape(n, self.nsite, 3), jstart=js, jend=je)

    def props(self):
        out = np.zeros(19)
        self.L.ref_get_props(self.h, dptr(out))
        keys = ["lj", "clmb", "wall", "pot", "vir0", "vir1", "vir2", "tra", "rot", "kin", "tot", "ham",
                "Tk", "prs0", "prs1", "prs2", "prs3", "H0", "gkT"]
        return dict(zip(keys, out.tolist()))

    def average(self):
        out = np.zeros(19)
        self.L.ref_get_average(self.h, dptr(out))
        return out

    def wall_table(self):
        n = self.L.ref_nfwall(self.h)
        out = np.zeros((n, 2))
        self.L.ref_wall_table(self.h, dptr(out))
        return out

    def __del__(self):
        try:
            self.L.ref_destroy(self.h)
        except Exception:
            pass


def ref_parallel_worker(mtl_path, desc, reps, budget_s, barrier, out):
    """One process of the replica-parallel CPU timing: the reference's FP32
    units (one thread) run this process's share of the replicas, one
    interval per call, round-robin until `budget_s` of timed work.  The
    reference runs replicas one after another (the replica loop of
    src/remd.cxx:257-267 has its OpenMP pragma commented out), so one
    process per core is the multi-core figure.  `reps` holds
    (r, v, q, p, ids, scal, T, P) per replica; puts (steps, seconds)."""
    import time

    lib = RefLib(True)
    systems = []
    for rep in reps:
        rs = RefSystem(lib, mtl_path, desc, *rep)
        rs.quiet(lib.L.ref_initialize_property)
        systems.append(rs)
    steps = int(desc["interval"])
    done = 0
    barrier.wait()
    t0 = time.perf_counter()
    while True:
        for rs in systems:
            rs.quiet(lib.L.ref_execute_steps)
            done += steps
        elapsed = time.perf_counter() - t0
        if elapsed > budget_s:
            break
    out.put((done, elapsed))


def port_parallel_worker(mt_text, desc, reps, fwall, budget_s, barrier, out):
    """As ref_parallel_worker, for the oracle's FP32 restatement of the same
    CPU path (bit-exact to the reference's FP32 build): this process's share
    of the replicas, one interval per call, round-robin until `budget_s`."""
    import time

    mt = parse_moltype(mt_text)
    systems = []
    for (r, v, q, p, ids, scal, T, P) in reps:
        o = OracleSystem(desc, mt, r, v, q, p, scal, T, P, fp32=True, fwall=fwall)
        o.initialize_property()
        systems.append(o)
    steps = int(desc["interval"])
    done = 0
    barrier.wait()
    t0 = time.perf_counter()
    while True:
        for o in systems:
            o.execute_steps(steps)
            done += steps
        elapsed = time.perf_counter() - t0
        if elapsed > budget_s:
            break
    out.put((done, elapsed))


def ref_available() -> bool:
    return os.path.exists(REF_F32) and os.path.exists(REF_F64)
