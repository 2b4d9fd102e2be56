"""Host-side range check of the per-level tables (no GPU).

Every index a kernel follows from a level's tables (grid ids, neighbour-oct
positions, record cells, reflux slots, coarse-target lists) is checked on the
host before the tables are uploaded, so a table built from a stale or
inconsistent binding ends the call with RH_ERR_TOPOLOGY instead of reaching
a launch.  tests/table_check_main.cpp builds the tables of a small adaptive
mesh with the library's own builder (rh_tables.cpp, compiled here with the
host compiler), checks that they pass, and plants one corruption per index
kind, each of which must be rejected."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_table_range_check_rejects_planted_corruptions(tmp_path):
    exe = tmp_path / "table_check"
    src = [os.path.join(ROOT, "tests", "table_check_main.cpp"),
           os.path.join(ROOT, "ramses-cpp_amd", "csrc", "rh_tables.cpp")]
    inc = ["-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "ramses-cpp_amd", "csrc")]
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    *inc, *src, "-o", str(exe)], check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "0 failures" in out.stdout
    assert out.stdout.count("is rejected") >= 16
