"""remd_amd -- MI355X-native REMD per-replica GPU path.

Python mirror of the reference's `class REMDGPU` (src/remdgpu.h:92-238) over
the C ABI in include/remd_gpu.h (libremd_gpu.so, built in-tree by
remd_amd.build).  There is no CPU fallback: if the HIP library is missing,
loading fails with an error.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import inputs

__all__ = ["REMDGPU", "Batch", "lib", "desc_for", "GlibcRand", "wall_table", "sort_window",
           "AVERAGE_NAMES", "inputs", "exchange_host", "exchange_pairs"]

PKG = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(PKG, "libremd_gpu.so")

# Average enum order (src/remdinfo.h:27-41)
AVERAGE_NAMES = ["LJ", "CLMB", "WALL", "VOL", "TRA", "ROT", "T", "P", "Px", "Py", "Pz", "TOT", "DRF",
                 "TSTs", "TSTv", "BSTx", "BSTy", "BSTz", "DUMMY"]

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class Desc(C.Structure):
    _fields_ = [
        ("nmol", C.c_int), ("nsite", C.c_int), ("nreplica", C.c_int),
        ("tconstant", C.c_int), ("pconstant", C.c_int), ("confined", C.c_int),
        ("interval", C.c_int), ("sort_interval", C.c_int),
        ("dt", C.c_double), ("rcut", C.c_double), ("rswitch", C.c_double), ("margin", C.c_double),
        ("wall_length", C.c_double), ("nfwall", C.c_int),
        ("sgm_wall", C.c_double), ("eps_wall", C.c_double), ("rho_wall", C.c_double),
        ("site_r", (C.c_double * 3) * 5),
        ("site_m", C.c_double * 5), ("site_sigma", C.c_double * 5), ("site_eps", C.c_double * 5),
        ("site_q", C.c_double * 5),
        ("mass", C.c_double), ("inertia", C.c_double * 3),
        ("precision", C.c_int), ("block_threads", C.c_int), ("chunks", C.c_int),
        ("path", C.c_int),
    ]


class Scalars(C.Structure):
    _fields_ = [("s", C.c_double), ("Ps", C.c_double), ("Q", C.c_double),
                ("Pv", C.c_double * 3), ("W", C.c_double), ("L", C.c_double * 3),
                ("T", C.c_double), ("P", C.c_double)]


class Rng(C.Structure):
    _fields_ = [("r", C.c_int * 34), ("k", C.c_int)]


_lib = None


def lib():
    """Load libremd_gpu.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise RuntimeError(f"{LIBPATH} is missing: build it with `python -m remd_amd.build` "
                           "(there is no CPU fallback)")
    L = C.CDLL(LIBPATH)
    L.remd_gpu_last_error.restype = C.c_char_p
    L.remd_gpu_create.argtypes = [C.POINTER(C.c_void_p), C.POINTER(Desc), C.POINTER(Scalars), _dp, _dp, _dp,
                                  _dp, _ip]
    for n in ("remd_gpu_execute_steps", "remd_gpu_execute_steps_plain", "remd_gpu_execute_steps_noalloc",
              "remd_gpu_set_conditions"):
        getattr(L, n).argtypes = [C.c_void_p, _dp, _dp]
    L.remd_gpu_execute_pending.argtypes = [C.c_void_p]
    L.remd_gpu_get_averages.argtypes = [C.c_void_p, C.c_int, _dp]
    L.remd_gpu_get_pot_vol.argtypes = [C.c_void_p, _dp, _dp]
    L.remd_gpu_download_state.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _ip,
                                          C.POINTER(Scalars)]
    L.remd_gpu_write_cdv.argtypes = [C.c_void_p, C.c_char_p, C.c_uint]
    L.remd_gpu_checkpoint_bytes.argtypes = [C.c_void_p]
    L.remd_gpu_checkpoint_bytes.restype = C.c_size_t
    L.remd_gpu_checkpoint_save.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    L.remd_gpu_checkpoint_load.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    L.remd_exchange_pairs.argtypes = [C.c_int, C.c_int, C.c_int, _ip, _ip]
    L.remd_gpu_destroy.argtypes = [C.c_void_p]
    L.remd_gpu_destroy.restype = None
    L.remd_gpu_compute_forces.argtypes = [C.c_void_p, C.c_int, C.c_double]
    L.remd_gpu_set_debug.argtypes = [C.c_void_p, C.c_int]
    L.remd_gpu_get_site_forces.argtypes = [C.c_void_p, C.c_int, _dp]
    L.remd_gpu_get_window.argtypes = [C.c_void_p, C.c_int, _ip, _ip, _ip]
    L.remd_gpu_last_kernel_ms.argtypes = [C.c_void_p, _dp]
    L.remd_gpu_launch_shape.argtypes = [C.c_void_p, _ip, _ip, _ip]
    L.remd_gpu_launch_path.argtypes = [C.c_void_p, _ip, _ip]
    L.remd_gpu_phase_timers.argtypes = [C.c_void_p, C.c_int, _dp]
    L.remd_gpu_autotune.argtypes = [C.c_void_p, C.c_int]
    L.remd_gpu_wall_table.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, _dp]
    L.remd_gpu_sort_window.argtypes = [_dp, C.c_int, C.c_int, C.c_double, _ip, _ip, _ip]
    L.remd_gpu_exchange_device.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_long, _ip, _dp, _dp, _dp, _ip,
                                           _ip]
    L.remd_batch_create.argtypes = [C.POINTER(C.c_void_p)]
    L.remd_batch_add.argtypes = [C.c_void_p, C.c_void_p]
    L.remd_batch_remove.argtypes = [C.c_void_p, C.c_void_p]
    L.remd_batch_execute.argtypes = [C.c_void_p]
    L.remd_batch_launch.argtypes = [C.c_void_p]
    L.remd_batch_poll.argtypes = [C.c_void_p]
    L.remd_batch_wait.argtypes = [C.c_void_p]
    L.remd_batch_set_autotune.argtypes = [C.c_void_p, C.c_int]
    L.remd_batch_set_parts.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.c_int]
    L.remd_batch_last_kernel_ms.argtypes = [C.c_void_p, _dp]
    L.remd_batch_destroy.argtypes = [C.c_void_p]
    L.remd_batch_destroy.restype = None
    L.remd_gpu_footprint.argtypes = [C.POINTER(Desc)]
    L.remd_gpu_footprint.restype = C.c_size_t
    L.remd_rng_seed.argtypes = [C.POINTER(Rng), C.c_uint]
    L.remd_rng_seed.restype = None
    L.remd_rng_next.argtypes = [C.POINTER(Rng)]
    L.remd_rng_next.restype = C.c_int
    L.remd_gpu_abi_version.restype = C.c_int
    L.remd_exchange_host.argtypes = [C.c_int, C.c_int, C.c_long, C.c_int, _dp, _dp, _ip, _dp, _dp,
                                     C.POINTER(C.c_ubyte), C.POINTER(Rng), _ip, _ip]
    L.remd_exchange_host.restype = C.c_int
    L.remd_gpu_execute_intervals.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_long, C.c_int, _ip, _dp, _dp,
                                             C.POINTER(Rng), _ip, _dp, _dp]
    L.remd_gpu_execute_intervals.restype = C.c_int
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise RuntimeError(f"remd_gpu error {rc}: {lib().remd_gpu_last_error().decode()}")


def _dp_of(a):
    return a.ctypes.data_as(_dp)


def _ip_of(a):
    return a.ctypes.data_as(_ip)


PATH_AUTO, PATH_RESIDENT, PATH_GLOBAL, PATH_SPLIT = 0, 1, 2, 3


def desc_for(system: "inputs.System", precision=0, block_threads=0, chunks=0, path=PATH_AUTO) -> Desc:
    """remd_gpu_desc for an inputs.System (reduced units)."""
    mt = inputs.molecule(system.model)
    d = Desc()
    d.nmol = system.nmol
    d.nsite = mt.nsite
    d.nreplica = system.nreplica
    d.tconstant, d.pconstant, d.confined = system.tconstant, system.pconstant, system.confined
    d.interval = system.interval
    d.sort_interval = system.sort_interval
    d.dt = system.dt
    d.rcut, d.rswitch, d.margin = system.rcut, system.rswitch, system.margin
    d.wall_length = system.wall_length
    d.nfwall = system.nfwall
    d.sgm_wall, d.eps_wall, d.rho_wall = system.sgm_wall, system.eps_wall, system.rho_wall
    for a in range(mt.nsite):
        for k in range(3):
            d.site_r[a][k] = mt.r[a, k]
        d.site_m[a] = mt.m[a]
        d.site_sigma[a] = mt.sigma[a]
        d.site_eps[a] = mt.eps[a]
        d.site_q[a] = mt.q[a]
    d.mass = mt.mass
    for k in range(3):
        d.inertia[k] = mt.inertia[k]
    d.precision, d.block_threads, d.chunks = precision, block_threads, chunks
    d.path = path
    return d


class REMDGPU:
    """Mirror of REMDGPU (src/remdgpu.h:92-238) for one job."""

    def __init__(self, system: "inputs.System", precision=0, block_threads=0, chunks=0, path=PATH_AUTO):
        L = lib()
        self.system = system
        self.nrep = system.nreplica
        self.nmol = system.nmol
        self.nsite = inputs.NSITE[system.model]
        self.desc = desc_for(system, precision, block_threads, chunks, path)
        R, N = self.nrep, self.nmol
        r = np.ascontiguousarray(np.stack([s.r for s in system.states]).reshape(R * N, 3))
        v = np.ascontiguousarray(np.stack([s.v for s in system.states]).reshape(R * N, 3))
        q = np.ascontiguousarray(np.stack([s.q for s in system.states]).reshape(R * N, 4))
        p = np.ascontiguousarray(np.stack([s.p for s in system.states]).reshape(R * N, 4))
        ids = np.ascontiguousarray(np.concatenate([
            s.ids if s.ids is not None else np.arange(N, dtype=np.int32) for s in system.states]).astype(np.int32))
        sc = (Scalars * R)()
        for k, s in enumerate(system.states):
            sc[k].s, sc[k].Ps, sc[k].Q = s.s, s.Ps, s.Q
            for d in range(3):
                sc[k].Pv[d] = s.Pv[d]
                sc[k].L[d] = s.L[d]
            sc[k].W = s.W
            sc[k].T = float(system.T[k])
            sc[k].P = float(system.P[k])
        h = C.c_void_p()
        _check(L.remd_gpu_create(C.byref(h), C.byref(self.desc), sc, _dp_of(r), _dp_of(v), _dp_of(q), _dp_of(p),
                                 _ip_of(ids)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().remd_gpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # REMDGPU::ExecuteSteps
    def execute_steps(self, T, P, plain=False):
        T = np.ascontiguousarray(T, np.float64)
        P = np.ascontiguousarray(P, np.float64)
        fn = lib().remd_gpu_execute_steps_plain if plain else lib().remd_gpu_execute_steps
        _check(fn(self.h, _dp_of(T), _dp_of(P)))

    def execute_steps_noalloc(self, T, P):
        """remd_gpu_execute_steps_noalloc: the same interval with the launch
        scratch in the library's static device array (no allocation)."""
        T = np.ascontiguousarray(T, np.float64)
        P = np.ascontiguousarray(P, np.float64)
        _check(lib().remd_gpu_execute_steps_noalloc(self.h, _dp_of(T), _dp_of(P)))

    def set_conditions(self, T, P):
        T = np.ascontiguousarray(T, np.float64)
        P = np.ascontiguousarray(P, np.float64)
        _check(lib().remd_gpu_set_conditions(self.h, _dp_of(T), _dp_of(P)))

    def execute_pending(self):
        _check(lib().remd_gpu_execute_pending(self.h))

    # REMDGPU::GetAverages
    def averages(self, rep) -> np.ndarray:
        out = np.zeros(19)
        _check(lib().remd_gpu_get_averages(self.h, int(rep), _dp_of(out)))
        return out

    def all_averages(self) -> np.ndarray:
        return np.stack([self.averages(r) for r in range(self.nrep)])

    def pot_vol(self):
        pot = np.zeros(self.nrep)
        vol = np.zeros(self.nrep)
        _check(lib().remd_gpu_get_pot_vol(self.h, _dp_of(pot), _dp_of(vol)))
        return pot, vol

    # REMDGPU::ConvertVariablesToHost (one replica)
    def state(self, rep):
        N = self.nmol
        r = np.zeros((N, 3)); v = np.zeros((N, 3)); f = np.zeros((N, 3))
        q = np.zeros((N, 4)); p = np.zeros((N, 4)); n = np.zeros((N, 4))
        ids = np.zeros(N, np.int32)
        sc = Scalars()
        _check(lib().remd_gpu_download_state(self.h, int(rep), _dp_of(r), _dp_of(v), _dp_of(f), _dp_of(q),
                                             _dp_of(p), _dp_of(n), _ip_of(ids), C.byref(sc)))
        return dict(r=r, v=v, f=f, q=q, p=p, n=n, id=ids,
                    scal=np.array([sc.s, sc.Ps, sc.Q, sc.Pv[0], sc.Pv[1], sc.Pv[2], sc.W, sc.L[0], sc.L[1],
                                   sc.L[2]]), T=sc.T, P=sc.P)

    def checkpoint_save(self) -> bytes:
        """remd_gpu_checkpoint_save: the job's exact device state."""
        n = int(lib().remd_gpu_checkpoint_bytes(self.h))
        buf = C.create_string_buffer(n)
        _check(lib().remd_gpu_checkpoint_save(self.h, buf, n))
        return buf.raw

    def checkpoint_load(self, blob: bytes):
        """remd_gpu_checkpoint_load: continue from a saved state."""
        buf = C.create_string_buffer(blob, len(blob))
        _check(lib().remd_gpu_checkpoint_load(self.h, buf, len(blob)))

    def write_cdv(self, prefix, step):
        _check(lib().remd_gpu_write_cdv(self.h, prefix.encode(), int(step)))

    # hooks
    def compute_forces(self, do_sort=True, margin=0.0):
        _check(lib().remd_gpu_compute_forces(self.h, int(do_sort), float(margin)))

    def set_debug(self, on=True):
        _check(lib().remd_gpu_set_debug(self.h, int(on)))

    def site_forces(self, rep):
        out = np.zeros(self.nmol * self.nsite * 3)
        _check(lib().remd_gpu_get_site_forces(self.h, int(rep), _dp_of(out)))
        return out.reshape(self.nmol, self.nsite, 3)

    def window(self, rep, with_full=False):
        js = np.zeros(self.nmol, np.int32)
        je = np.zeros(self.nmol, np.int32)
        full = np.zeros(self.nmol, np.int32)
        _check(lib().remd_gpu_get_window(self.h, int(rep), _ip_of(js), _ip_of(je), _ip_of(full)))
        return (js, je, full) if with_full else (js, je)

    def last_kernel_ms(self):
        ms = C.c_double()
        _check(lib().remd_gpu_last_kernel_ms(self.h, C.byref(ms)))
        return ms.value

    def phase_timers(self, rep=0):
        out = np.zeros(16)
        _check(lib().remd_gpu_phase_timers(self.h, int(rep), _dp_of(out)))
        return out

    def launch_shape(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        _check(lib().remd_gpu_launch_shape(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return dict(block_threads=a.value, chunks=b.value, lds_bytes=c.value)

    def launch_path(self):
        """(path, parts): PATH_RESIDENT or PATH_GLOBAL, and the workgroups per
        replica of the last launch."""
        p, g = C.c_int(), C.c_int()
        _check(lib().remd_gpu_launch_path(self.h, C.byref(p), C.byref(g)))
        return p.value, g.value

    def autotune(self, reps=3):
        _check(lib().remd_gpu_autotune(self.h, int(reps)))
        return self.launch_shape()

    def exchange_device(self, ntemp, npress, step, index, T, P, uniforms):
        index = np.ascontiguousarray(index, np.int32).copy()
        T = np.ascontiguousarray(T, np.float64).copy()
        P = np.ascontiguousarray(P, np.float64).copy()
        u = np.ascontiguousarray(uniforms, np.float64)
        acc = np.zeros(max(1, self.nrep), np.int32)
        npairs = C.c_int()
        _check(lib().remd_gpu_exchange_device(self.h, int(ntemp), int(npress), int(step), _ip_of(index),
                                              _dp_of(T), _dp_of(P), _dp_of(u), _ip_of(acc), C.byref(npairs)))
        return index, T, P, acc[: npairs.value]


def execute_intervals(g: "REMDGPU", ntemp, npress, step0, nint, index, T, P, rng: "GlibcRand"):
    """remd_gpu_execute_intervals: nint intervals with the exchange on the
    device between them.  index/T/P are updated in place (int32 / float64);
    returns (n_accepted, accepted per pair, averages nint x nrep x 19,
    kernel ms per interval)."""
    assert index.dtype == np.int32 and T.dtype == np.float64 and P.dtype == np.float64
    npairs = sum(int(exchange_pairs(ntemp, npress, (step0 + k) % 4)[0].size) for k in range(nint))
    acc = np.zeros(max(1, npairs), np.int32)
    av = np.zeros((nint, ntemp * npress, 19))
    ms = np.zeros(nint)
    rc = lib().remd_gpu_execute_intervals(g.h, int(ntemp), int(npress), int(step0), int(nint), _ip_of(index),
                                          _dp_of(T), _dp_of(P), C.byref(rng.s), _ip_of(acc), _dp_of(av),
                                          _dp_of(ms))
    if rc < 0:
        _check(rc)
    return rc, acc[:npairs], av, ms


def exchange_pairs(ntemp, npress, typ):
    """remd_exchange_pairs: (a, b) condition slots of exchange list `typ`."""
    n = ntemp * npress
    a = np.zeros(max(1, n), np.int32)
    b = np.zeros(max(1, n), np.int32)
    k = lib().remd_exchange_pairs(int(ntemp), int(npress), int(typ), _ip_of(a), _ip_of(b))
    if k < 0:
        _check(k)
    return a[:k].copy(), b[:k].copy()


class Batch:
    """Several REMDGPU jobs advanced by one launch per interval."""

    def __init__(self):
        h = C.c_void_p()
        _check(lib().remd_batch_create(C.byref(h)))
        self.h = h
        self.jobs = []

    def add(self, job: REMDGPU):
        _check(lib().remd_batch_add(self.h, job.h))
        self.jobs.append(job)

    def remove(self, job: REMDGPU):
        _check(lib().remd_batch_remove(self.h, job.h))
        self.jobs.remove(job)

    def execute(self):
        _check(lib().remd_batch_execute(self.h))

    def launch(self):
        """remd_batch_launch: enqueue one interval and return."""
        _check(lib().remd_batch_launch(self.h))

    def poll(self) -> bool:
        """remd_batch_poll: True once the launched interval is done."""
        rc = lib().remd_batch_poll(self.h)
        if rc < 0:
            _check(rc)
        return rc == 1

    def wait(self):
        """remd_batch_wait: block until the launched interval is done."""
        _check(lib().remd_batch_wait(self.h))

    def set_autotune(self, on: bool):
        """Time candidate packing plans on the first intervals after the
        batch changes (default), or keep the cost model's best plan."""
        _check(lib().remd_batch_set_autotune(self.h, int(bool(on))))

    def set_parts(self, parts=None):
        """Fix the workgroups per replica of each job (in add order); None
        restores the automatic plan."""
        if parts is None:
            _check(lib().remd_batch_set_parts(self.h, None, 0))
        else:
            arr = (C.c_int * len(parts))(*[int(p) for p in parts])
            _check(lib().remd_batch_set_parts(self.h, arr, len(parts)))

    def parts(self):
        """Workgroups per replica each job got in the last interval."""
        return [j.launch_path()[1] for j in self.jobs]

    def last_kernel_ms(self):
        ms = C.c_double()
        _check(lib().remd_batch_last_kernel_ms(self.h, C.byref(ms)))
        return ms.value

    def close(self):
        if getattr(self, "h", None):
            lib().remd_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GlibcRand:
    """Per-job glibc rand() stream (the reference's exchange RNG)."""

    def __init__(self, seed=1):
        self.s = Rng()
        lib().remd_rng_seed(C.byref(self.s), int(seed))

    def rand(self) -> int:
        return lib().remd_rng_next(C.byref(self.s))

    def uniform(self) -> float:
        return self.rand() / 2147483647.0

    def getstate(self) -> bytes:
        return bytes(self.s)

    def setstate(self, state: bytes):
        C.memmove(C.byref(self.s), state, C.sizeof(self.s))


def exchange_host(ntemp, npress, step, pconstant, pot, vol, index, T, P, rng: "GlibcRand"):
    """REMD::ReplicaExchange on the host (in place on index/T/P); returns
    (n_accepted, accepted per pair, is_exchanged per replica)."""
    n = ntemp * npress
    pot = np.ascontiguousarray(pot, np.float64)
    vol = np.ascontiguousarray(vol, np.float64)
    assert index.dtype == np.int32 and T.dtype == np.float64 and P.dtype == np.float64
    isx = np.zeros(n, np.uint8)
    acc = np.zeros(n + 1, np.int32)
    npairs = C.c_int()
    rc = lib().remd_exchange_host(int(ntemp), int(npress), int(step), int(pconstant), _dp_of(pot), _dp_of(vol),
                                  _ip_of(index), _dp_of(T), _dp_of(P), isx.ctypes.data_as(C.POINTER(C.c_ubyte)),
                                  C.byref(rng.s), _ip_of(acc), C.byref(npairs))
    if rc < 0:
        _check(rc)
    return rc, acc[: npairs.value], isx


def wall_table(confined, nfwall, R, sgm, eps, rho) -> np.ndarray:
    out = np.zeros((nfwall, 2))
    _check(lib().remd_gpu_wall_table(int(confined), int(nfwall), float(R), float(sgm), float(eps), float(rho),
                                     _dp_of(out)))
    return out


def sort_window(keys, w):
    keys = np.ascontiguousarray(np.atleast_2d(keys), np.float64)
    nrow, n = keys.shape
    order = np.zeros((nrow, n), np.int32)
    js = np.zeros((nrow, n), np.int32)
    je = np.zeros((nrow, n), np.int32)
    _check(lib().remd_gpu_sort_window(_dp_of(keys), nrow, n, float(w), _ip_of(order), _ip_of(js), _ip_of(je)))
    return order, js, je
