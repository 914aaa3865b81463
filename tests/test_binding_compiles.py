"""The reference-side binding (include/remd_gpu_molecules.hpp) compiles
against the reference's own headers (class Molecules, src/molecules.h;
Average, src/remdinfo.h) and links against the built C-ABI library: every
member it reads or writes exists with a compatible type.  Needs the
reference sources (skipped where they are absent) and the library; runs
nothing on a GPU."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SRC = "/root/reference/src"
LIB = os.path.join(ROOT, "remd_amd", "libremd_gpu.so")

PROGRAM = r"""
#include "molecules.h"
#include "remdinfo.h"
#include "remd_gpu_molecules.hpp"
// every binding entry point referenced, none run
int main(int argc, char **argv) {
  if (argc > 1000) {
    Molecules **m = nullptr;
    remd_gpu *h = remd_amd::create(m, 1);
    remd_amd::to_host(h, m, 1);
    Average a = remd_amd::averages(h, 0);
    (void)a;
    remd_amd::calc_force_pot(h, *m[0], true, 0.0);
    remd_amd::execute_steps(h, *m[0]);
    remd_gpu_destroy(h);
  }
  return 0;
}
"""


@pytest.mark.skipif(not os.path.isdir(REF_SRC), reason="reference sources not present")
@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_binding_compiles_and_links_against_reference_headers(tmp_path):
    src = tmp_path / "binding.cxx"
    src.write_text(PROGRAM)
    obj = tmp_path / "binding.o"
    cc = subprocess.run(["g++", "-std=c++11", "-fopenmp", "-Wall", "-Werror=return-type", "-I", REF_SRC,
                         "-I", os.path.join(ROOT, "include"), "-c", str(src), "-o", str(obj)],
                        capture_output=True, text=True)
    assert cc.returncode == 0, cc.stderr
    if not os.path.exists(LIB):
        pytest.skip("libremd_gpu.so not built")
    ld = subprocess.run(["g++", "-fopenmp", str(obj), "-L", os.path.dirname(LIB), "-lremd_gpu",
                         "-o", str(tmp_path / "binding")], capture_output=True, text=True)
    assert ld.returncode == 0, ld.stderr
