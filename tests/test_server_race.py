"""The serving runtime's scheduler and device threads under ThreadSanitizer and AddressSanitizer
on the CPU: bmi_server.hip is compiled for the host against stand-ins of the HIP calls and of the
operator/solver C ABI (random delays, real buffers, solves that need two time slices); four
clients submit and wait concurrently over rounds with different slot counts, batch limits, slice
lengths and memory budgets (eviction).  Every result is checked, and every device call of a
running server must come from its one device thread."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "cpp", "server_race")
CLANG = "/opt/rocm/llvm/bin/clang++"


def build(tmp_path, sanitizer, cxx):
    exe = str(tmp_path / ("race_" + sanitizer))
    cmd = [cxx, "-std=c++17", "-O1", "-g", "-fsanitize=" + sanitizer, "-I" + SRC,
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "bembel_amd", "csrc"), "-x", "c++",
           os.path.join(ROOT, "bembel_amd", "csrc", "bmi_server.hip"), "-x", "none",
           os.path.join(SRC, "stubs.cpp"), os.path.join(SRC, "driver.cpp"), "-o", exe, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return exe


@pytest.mark.parametrize("sanitizer", ["thread", "address"])
def test_server_threads_under_sanitizers(tmp_path, sanitizer):
    # clang's ThreadSanitizer intercepts pthread_cond_clockwait (used by condition_variable::wait_for)
    cxx = CLANG if sanitizer == "thread" else (shutil.which("g++") or CLANG)
    if not os.path.exists(cxx):
        pytest.skip("compiler missing")
    exe = build(tmp_path, sanitizer, cxx)
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1", ASAN_OPTIONS="detect_leaks=1")
    r = subprocess.run([exe, "6"], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0 and "ok" in r.stdout, (r.stdout + r.stderr)[-4000:]
    assert "WARNING" not in r.stderr, r.stderr[-4000:]
