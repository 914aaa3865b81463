"""Build the HIP library (libremd_gpu.so), the reference-style driver
(remd_md) and the oracle checker for gfx950, in-tree.

    python -m remd_amd.build            # library + driver
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libremd_gpu.so")
DRIVER = os.path.join(PKG, "remd_md")

HIP_SOURCES = ["remd_kernels.hip", "remd_engine.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("REMD_OFFLOAD_ARCH", "gfx950")

# Packed FP32 VALU forms are disabled for all device code of this project.
# -ffp-contract=on: a multiply and an add fuse only where they are written in
# one expression (the front end decides, the same way wherever a routine is
# inlined), never across statements after inlining, so a routine gives the
# same bits in every kernel that runs it (single job, batch, plain).
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=on",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-value", "-Wno-unused-result",
]


def _run(cmd, cwd=None):
    res = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in res.stdout.splitlines() if "is not a recognized feature" not in l
                    and "packed-fp32-ops" not in l)
    if res.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")
    return out


def build_library(verbose=False):
    """Compile each HIP source to an object (in parallel), then link."""
    from concurrent.futures import ThreadPoolExecutor

    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)

    def compile_one(s):
        obj = os.path.join(objdir, os.path.splitext(s)[0] + ".o")
        return obj, _run([HIPCC, *HIP_FLAGS, "-c", "-o", obj, os.path.join(CSRC, s)])

    with ThreadPoolExecutor(len(HIP_SOURCES)) as ex:
        done = list(ex.map(compile_one, HIP_SOURCES))
    tmp = LIB + ".tmp"
    out = _run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, *[o for o, _ in done]])
    os.replace(tmp, LIB)
    if verbose:
        for _, o in done:
            if o.strip():
                print(o)
    return LIB


def build_driver():
    src = os.path.join(CSRC, "remd_driver.cpp")
    if not os.path.exists(src):
        return None
    tmp = DRIVER + ".tmp"
    # host-only C++ over the C ABI
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-o", tmp, src, f"-I{os.path.join(ROOT, 'include')}",
          f"-L{PKG}", "-lremd_gpu", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, DRIVER)
    return DRIVER


def build_oracle():
    """Checker only (test infrastructure): oracle/liboracle.so, and the
    reference's CPU units into oracle/_ref when /root/reference is present."""
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-C", odir, "liboracle.so"])
    if os.path.isdir("/root/reference/src") and shutil.which("g++"):
        _run(["make", "-s", "-j8", "-C", odir, "ref"])


def build_all():
    build_library()
    build_driver()
    build_oracle()


if __name__ == "__main__":
    build_all()
    print("built", LIB)
