"""The device-state setup of a shared discrete space (bembel_amd/csrc/bmi_space_dev.cpp) under
ThreadSanitizer and AddressSanitizer on the CPU, compiled against stand-ins of the HIP calls and
of the far-field fill kernel: eight threads make and read the far-field tables of several
degrees, with and without tangents, on one space, as concurrent assemblies of different
operators on handles that share a space do.  Every table must be complete when it is handed out
and keep its address for the life of the space."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "cpp", "space_race")
STUB_HIP = os.path.join(HERE, "cpp", "server_race")  # hip/hip_runtime.h stand-in
CLANG = "/opt/rocm/llvm/bin/clang++"


@pytest.mark.parametrize("sanitizer", ["thread", "address"])
def test_space_device_state_under_sanitizers(tmp_path, sanitizer):
    cxx = CLANG if sanitizer == "thread" else (shutil.which("g++") or CLANG)
    if not os.path.exists(cxx):
        pytest.skip("compiler missing")
    exe = str(tmp_path / ("space_" + sanitizer))
    cmd = [cxx, "-std=c++17", "-O1", "-g", "-fsanitize=" + sanitizer, "-I" + STUB_HIP,
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "bembel_amd", "csrc"), "-x", "c++",
           os.path.join(ROOT, "bembel_amd", "csrc", "bmi_space_dev.cpp"), "-x", "none",
           os.path.join(SRC, "stubs.cpp"), os.path.join(SRC, "driver.cpp"), "-o", exe, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1", ASAN_OPTIONS="detect_leaks=1")
    r = subprocess.run([exe, "10"], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0 and "ok" in r.stdout, (r.stdout + r.stderr)[-4000:]
    assert "WARNING" not in r.stderr, r.stderr[-4000:]
