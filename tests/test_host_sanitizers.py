"""Host code under the compiler's sanitizers (no GPU).

  * AddressSanitizer + UndefinedBehaviorSanitizer: the library's host-only
    mesh construction, initial conditions and level-table builder
    (rh_driver.cpp, rh_tables.cpp) and the whole oracle (uniform steps, an
    adaptive run with its tree updates, a snapshot dump), driven by
    tests/host_sanitize_main.cpp.
  * ThreadSanitizer: the adaptive driver's worker pool (rh_pool.h) with two
    host threads passing work at once and the pool released (its workers
    joined) and created again, driven by tests/host_pool_main.cpp.
Any report ends the program with a nonzero status."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "ramses-cpp_amd", "csrc")
INC = ["-I" + os.path.join(ROOT, "include"), "-I" + CSRC, "-I" + os.path.join(ROOT, "oracle")]

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")


def _build_and_run(tmp_path, name, flags, sources, args=()):
    exe = tmp_path / name
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-ffp-contract=off", *flags, *INC, *sources, "-o", str(exe)]
    b = subprocess.run(cmd, capture_output=True, text=True)
    assert b.returncode == 0, b.stderr[-4000:]
    env = dict(os.environ)
    env.pop("LD_PRELOAD", None)
    r = subprocess.run([str(exe), *args], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, (r.stdout + r.stderr)[-6000:]
    return r.stdout


def test_host_code_and_oracle_under_asan_ubsan(tmp_path):
    src = [os.path.join(ROOT, "tests", "host_sanitize_main.cpp"), os.path.join(CSRC, "rh_tables.cpp"),
           os.path.join(CSRC, "rh_driver.cpp")] + [os.path.join(ROOT, "oracle", f) for f in
                                                  ("rh_oracle.cpp", "rh_oracle_amr.cpp", "rh_oracle_snapshot.cpp")]
    snap = tmp_path / "snap"
    snap.mkdir()
    out = _build_and_run(tmp_path, "host_asan", ["-fsanitize=address,undefined", "-fno-sanitize-recover=all"], src,
                         [str(snap)])
    assert "0 failures" in out
    assert out.count("ok   uniform") == 2 and "ok   adaptive" in out
    assert (snap / "output_00001" / "amr_00001.out00001").exists()


def test_worker_pool_under_tsan(tmp_path):
    out = _build_and_run(tmp_path, "pool_tsan", ["-fsanitize=thread", "-pthread"],
                         [os.path.join(ROOT, "tests", "host_pool_main.cpp")])
    assert "0 failures" in out
