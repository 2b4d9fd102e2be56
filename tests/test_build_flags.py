"""Build settings the results depend on.

The FP64 form is bit-identical to the reference's algorithm only when no
multiply-add is contracted and no fast-math rewriting is allowed, on the GPU
and in the oracle alike; the device code must also carry no packed FP32
arithmetic (the mixed form's FP32 half is compiled with the packed-FP32
feature turned off).  These tests read the two Makefiles and, when the
library is built, the gfx950 code object inside it."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "ramses-cpp_amd")
LIB = os.path.join(PKG, "librhydro.so")
LLVM = "/opt/rocm/llvm/bin"


def _var(text, name):
    m = re.search(rf"^{name}\s*[:?]?=\s*(.*)$", text, re.M)
    assert m, f"{name} is not set"
    return m.group(1)


def test_library_makefile_flags():
    text = open(os.path.join(PKG, "Makefile")).read()
    common = _var(text, "COMMON")
    assert "-ffp-contract=off" in common.split()
    assert "-fno-fast-math" in common.split()
    assert "-ffast-math" not in common.split()
    dev = _var(text, "DEVFLAGS")
    assert "-Xclang -target-feature -Xclang -packed-fp32-ops" in dev
    assert "--offload-arch=$(ARCH)" in dev
    assert _var(text, "ARCH") == "gfx950"
    # Both the device and the host objects are compiled with COMMON.
    rules = [l for l in text.splitlines() if "$(HIPCC)" in l and " -c " in l]
    assert rules and all("$(COMMON)" in l for l in rules)


def test_oracle_makefile_flags():
    text = open(os.path.join(ROOT, "oracle", "Makefile")).read()
    flags = _var(text, "CXXFLAGS").split()
    assert "-ffp-contract=off" in flags and "-fno-fast-math" in flags
    for bad in ("-ffast-math", "-march=native", "-mfma", "-Ofast"):
        assert bad not in flags


@pytest.mark.skipif(not os.path.exists(LIB), reason="librhydro.so is not built")
@pytest.mark.skipif(
    not (shutil.which("objcopy") and os.path.exists(os.path.join(LLVM, "clang-offload-bundler"))),
    reason="objcopy or the ROCm LLVM tools are missing",
)
def test_device_code_is_gfx950_without_packed_fp32(tmp_path):
    fatbin = tmp_path / "fatbin.bin"
    subprocess.run(["objcopy", "-O", "binary", "--only-section=.hip_fatbin", LIB, str(fatbin)], check=True)
    listed = subprocess.run([os.path.join(LLVM, "clang-offload-bundler"), "--list", "--type=o", f"--input={fatbin}"],
                            check=True, capture_output=True, text=True).stdout.split()
    devices = [t for t in listed if t.startswith("hipv4-")]
    assert devices == ["hipv4-amdgcn-amd-amdhsa--gfx950"]
    co = tmp_path / "dev.co"
    subprocess.run([os.path.join(LLVM, "clang-offload-bundler"), "--unbundle", "--type=o", f"--input={fatbin}",
                    f"--targets={devices[0]}", f"--output={co}"], check=True)
    asm = subprocess.run([os.path.join(LLVM, "llvm-objdump"), "-d", str(co)], check=True, capture_output=True,
                         text=True).stdout
    assert "k_godunov_oct" in asm
    packed = re.findall(r"\bv_pk_(?:add|mul|fma)_f32\b", asm)
    assert not packed, f"{len(packed)} packed FP32 instructions in the device code"
