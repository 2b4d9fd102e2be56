"""The drop-in adapter (ramses-cpp_amd/adapter/PatchHydroSolver.cpp) compiles
against the reference's own headers with RAMSES_HAS_PATCH_HYDRO defined, as
it would inside the reference's patch/ directory.  Skipped where the
reference tree is absent."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "include")), reason="reference tree not present")
def test_adapter_compiles_against_reference_headers():
    inc = os.path.join(REF, "include")
    subdirs = ["", "ramses", "ramses/core", "ramses/solvers/hydro", "ramses/solvers/mhd", "ramses/solvers/rhd",
               "ramses/solvers/rt", "ramses/solvers/gravity", "ramses/solvers/physics", "ramses/particles",
               "ramses/io"]
    cmd = ["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-DNDIM=3", "-DNPRE=8",
           "-DRAMSES_HAS_PATCH_HYDRO", "-I" + os.path.join(ROOT, "include")]
    cmd += ["-I" + os.path.join(inc, s) for s in subdirs]
    cmd.append(os.path.join(ROOT, "ramses-cpp_amd", "adapter", "PatchHydroSolver.cpp"))
    out = subprocess.run(cmd, capture_output=True, text=True)
    assert out.returncode == 0, out.stderr
    own = [l for l in out.stderr.splitlines() if "PatchHydroSolver.cpp" in l and "warning" in l]
    assert not own, own
