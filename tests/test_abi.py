"""The C-ABI library loads on a CPU-only host and exports every entry point
the public headers declare; argument validation happens before any device
call."""
import ctypes as C
import glob
import os
import re

import pytest

import rhydro

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in sorted(glob.glob(os.path.join(ROOT, "include", "*.h"))):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(rh_[a-z_0-9]+)\s*\(", text))
    return names


def test_every_declared_symbol_is_exported():
    lib = rhydro.lib()
    names = _declared()
    assert "rh_godunov_fine" in names and len(names) >= 15
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert set(rhydro.EXPORTED) <= names


def test_binding_lists_every_declared_symbol():
    assert _declared() <= set(rhydro.EXPORTED)


def _params(**kw):
    p = rhydro.RhParams(1.4, 1, 2, 1, 0, 5, 0, 0, 1.0, 0, 0)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


@pytest.mark.parametrize("kw,code", [
    (dict(nvar=6), rhydro.RH_ERR_UNSUPPORTED),
    (dict(nener=1), rhydro.RH_ERR_UNSUPPORTED),
    (dict(barotropic_eos=1), rhydro.RH_ERR_UNSUPPORTED),
    (dict(riemann=3), rhydro.RH_ERR_UNSUPPORTED),
    (dict(slope_type=9), rhydro.RH_ERR_ARG),
    (dict(gamma=1.0), rhydro.RH_ERR_ARG),
    (dict(precision=7), rhydro.RH_ERR_ARG),
])
def test_create_validates_parameters(kw, code):
    ctx = C.c_void_p()
    p = _params(**kw)
    assert rhydro.lib().rh_create(C.byref(p), C.byref(ctx)) == code
    assert rhydro.lib().rh_last_error(None)


def test_riemann_name_resolution_matches_face_loop():
    # HydroSolver.cpp:456-461: case-sensitive exact strings, default LLF.
    assert rhydro.resolve_riemann("hllc") == 2
    assert rhydro.resolve_riemann("hll") == 1
    assert rhydro.resolve_riemann("acoustic") == 4
    assert rhydro.resolve_riemann("HLLC") == 0
    assert rhydro.resolve_riemann(None) == 0
    assert rhydro.resolve_riemann("roe") == 0


def test_no_device_is_reported_on_cpu_host():
    try:
        import subprocess

        has_gpu = subprocess.run(["sh", "-c", "ls /dev/kfd"], capture_output=True).returncode == 0
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    with pytest.raises(rhydro.RhError) as e:
        rhydro.Solver()
    assert e.value.code == rhydro.RH_ERR_NO_DEVICE
