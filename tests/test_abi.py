"""The C-ABI library on the CPU: it loads, exports every entry point that
include/remd_gpu.h declares, and its host-only logic (RNG, host exchange,
footprint, argument validation) agrees with the oracle.  No device compute
is called here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import pyoracle as po
import remd_amd as ra
from remd_amd import inputs as inp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "remd_gpu.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(remd_\w+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_entry_point():
    L = ra.lib()
    names = _declared()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_header_compiles_as_c():
    """The header is plain C (no C++ or framework types)."""
    import subprocess
    import tempfile

    with tempfile.TemporaryDirectory() as td:
        src = os.path.join(td, "t.c")
        with open(src, "w") as f:
            f.write('#include "remd_gpu.h"\nint main(void){return REMD_GPU_ABI_VERSION - 1;}\n')
        r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.dirname(HEADER), src, "-o",
                            os.path.join(td, "t")], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


def test_abi_version():
    assert ra.lib().remd_gpu_abi_version() == 1


def test_glibc_rand_stream_matches_libc():
    """remd_rng reproduces glibc srand()/rand() (the exchange RNG of
    src/remd.cxx:465), and the oracle's rand is libc's own."""
    libc = C.CDLL("libc.so.6")
    for seed in (1, 7, 12345, 2 ** 31 - 1):
        g = ra.GlibcRand(seed)
        libc.srand(seed)
        a = [g.rand() for _ in range(1000)]
        b = [libc.rand() for _ in range(1000)]
        assert a == b
        po.lib().orc_srand(seed)
        c = [po.lib().orc_rand() for _ in range(50)]
        assert c == b[:50]


def test_rng_streams_are_independent_per_job():
    g1, g2 = ra.GlibcRand(3), ra.GlibcRand(3)
    a = [g1.rand() for _ in range(10)]
    _ = [g2.rand() for _ in range(5)]
    g3 = ra.GlibcRand(3)
    assert [g3.rand() for _ in range(10)] == a


@pytest.mark.parametrize("nt,npr,pmode", [(16, 4, 2), (4, 1, 0), (5, 3, 3), (2, 2, 1)])
def test_host_exchange_matches_oracle(nt, npr, pmode):
    """remd_exchange_host == REMD::ReplicaExchange (oracle) on the same
    energies, volumes and rand() stream, over 40 consecutive rounds."""
    rng = np.random.default_rng(nt * 10 + npr)
    n = nt * npr
    T, P = inp.make_ladder(nt, npr, inp.geometric(200.0, 360.0, nt), inp.geometric(0.1e6, 200e6, npr))
    index_g = np.arange(n, dtype=np.int32)
    Tg, Pg = T.copy(), P.copy()
    index_o, To, Po = index_g.copy(), T.copy(), P.copy()
    g = ra.GlibcRand(11)
    u_stream = ra.GlibcRand(11)
    for step in range(40):
        pot = rng.normal(-4000, 60, n)
        vol = rng.normal(3000, 40, n)
        nacc, acc, isx = ra.exchange_host(nt, npr, step, pmode, pot, vol, index_g, Tg, Pg, g)
        npairs = po.exchange_list(nt, npr, step % 4).shape[0]
        u = np.array([u_stream.uniform() for _ in range(npairs)])
        index_o, To, Po, acc_o, isx_o = po.replica_exchange(nt, npr, step, pmode, pot, vol, index_o, To, Po, u)
        assert np.array_equal(acc, acc_o[:npairs])
        assert nacc == int(acc_o[:npairs].sum())
        assert np.array_equal(isx, isx_o)
        assert np.array_equal(index_g, index_o) and np.array_equal(Tg, To) and np.array_equal(Pg, Po)


def test_footprint_grows_with_size():
    s = inp.config_tube(seed=2, ntemp=2, npress=1)
    d = ra.desc_for(s)
    f1 = ra.lib().remd_gpu_footprint(C.byref(d))
    d.nreplica = 64
    f2 = ra.lib().remd_gpu_footprint(C.byref(d))
    assert 0 < f1 < f2
    # molecule block (21 doubles + id per molecule) + scalars per replica at least
    assert f2 >= 64 * s.nmol * (21 * 8 + 4)


def _scalars(system):
    sc = (ra.Scalars * system.nreplica)()
    for k, st in enumerate(system.states):
        sc[k].s, sc[k].Ps, sc[k].Q, sc[k].W = st.s, st.Ps, st.Q, st.W
        for d in range(3):
            sc[k].Pv[d] = st.Pv[d]
            sc[k].L[d] = st.L[d]
        sc[k].T, sc[k].P = float(system.T[k]), float(system.P[k])
    return sc


@pytest.mark.parametrize("field,bad,msg", [
    ("nsite", 6, "nsite"), ("nmol", 1, "nmol"), ("nmol", 5000, "nmol"), ("interval", -1, "interval"),
    ("confined", 3, "confined"), ("pconstant", 1, "pconstant"), ("precision", 2, "precision"),
    ("block_threads", 96, "block_threads"), ("chunks", 3, "chunks"), ("rcut", 7.0, "rswitch"),
    ("nfwall", 2, "nfwall")])
def test_create_rejects_bad_descriptors(field, bad, msg):
    """Validation happens on the host, before any device call
    (REMDGPU::REMDGPU's checks, src/remdgpu.cu:178-212, as error codes)."""
    s = inp.config_tube(seed=2, ntemp=1, npress=1)
    d = ra.desc_for(s)
    setattr(d, field, bad)
    if field == "chunks":
        d.path = ra.PATH_GLOBAL  # a power of two there; any 1..64 on the resident path
    h = C.c_void_p()
    z = np.zeros(4 * s.nmol)
    rc = ra.lib().remd_gpu_create(C.byref(h), C.byref(d), _scalars(s), ra._dp_of(z), ra._dp_of(z),
                                  ra._dp_of(z), ra._dp_of(z), None)
    assert rc == -1, (field, bad, rc)
    assert msg in ra.lib().remd_gpu_last_error().decode()
    assert not h.value


def test_create_rejects_box_too_small_for_the_window():
    s = inp.config_tube(seed=2, ntemp=1, npress=1)
    st = s.states[0]
    st.L = (st.L[0], st.L[1], 18.0)  # < 2 (rcut + margin)
    d = ra.desc_for(s)
    h = C.c_void_p()
    z = np.zeros(4 * s.nmol)
    rc = ra.lib().remd_gpu_create(C.byref(h), C.byref(d), _scalars(s), ra._dp_of(z), ra._dp_of(z),
                                  ra._dp_of(z), ra._dp_of(z), None)
    assert rc == -1 and b"sort axis" in ra.lib().remd_gpu_last_error()


def test_chk_round_trip(tmp_path):
    """write_chk / read_chk (IOManager::WriteCheckPoint / ReadCheckPoint,
    src/io_manager.cxx:528-634): the text format keeps 16 significant
    digits and the reader converts back to the internal representation,
    so the state survives to 1e-14 relative."""
    s = inp.config_tube(seed=5, ntemp=1, npress=1)
    st = s.states[0]
    path = tmp_path / "r.chk"
    inp.write_chk(str(path), st, inp.molecule(s.model))
    back = inp.read_chk(str(path))
    for k in ("r", "v", "q", "p"):
        a, b = getattr(back, k), getattr(st, k)
        assert np.allclose(a, b, rtol=1e-14, atol=1e-14 * np.abs(b).max())
    assert np.allclose(back.scal(), st.scal(), rtol=1e-14)


def test_product_path_fails_loudly_without_library(monkeypatch, tmp_path):
    """No CPU fallback: a missing HIP library is an error."""
    monkeypatch.setattr(ra, "_lib", None)
    monkeypatch.setattr(ra, "LIBPATH", str(tmp_path / "missing.so"))
    with pytest.raises((OSError, RuntimeError, FileNotFoundError)):
        ra.lib()


def test_driver_rejects_unsupported_input(tmp_path):
    """remd_md stops with a message (no device call) on inputs it does not
    support."""
    import subprocess
    drv = os.path.join(ROOT, "remd_amd", "remd_md")
    for text, word in (("gen_vel 1\nchk_in gpu\n", "gen_vel"), ("ntemp 2\nnreplica 3\nchk_in a\n", "nreplica"),
                       ("rem_type 1\nchk_in a\n", "designed walk")):
        (tmp_path / "x.inp").write_text(text)
        res = subprocess.run([drv, str(tmp_path / "x.inp")], capture_output=True, text=True, timeout=60)
        assert res.returncode != 0 and word in res.stderr, res.stderr


@pytest.mark.parametrize("nt,npr", [(4, 1), (16, 4), (5, 3), (2, 2), (3, 1)])
def test_exchange_pairs_match_reference_lists(golden, nt, npr):
    """remd_exchange_pairs gives the reference's four exchange pair lists
    (REMDInfo::ExchangeList, src/remdinfo.cxx:54-108; fixture from the
    reference build) in list order."""
    d = golden("ref_remdinfo.npz")
    for typ in range(4):
        a, b = ra.exchange_pairs(nt, npr, typ)
        n = int(d[f"len_{nt}x{npr}"][typ])
        assert a.shape[0] == n
        ref = d[f"lists_{nt}x{npr}"][typ, :n]
        assert np.array_equal(np.stack([a, b], 1), ref)
