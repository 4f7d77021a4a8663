"""The reference-style C++ facade (include/bembel_amd/Bembel.hpp) drives the library."""
import os
import subprocess

import pytest

import bembel_amd as B

EXE = os.path.join(os.path.dirname(B.LIB_PATH), "_build", "facade_example")


def _exe():
    if not os.path.exists(EXE):
        from bembel_amd import build
        build.build_examples()
    return EXE


def test_facade_host_part():
    r = subprocess.run([_exe(), B.SPHERE_DAT], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "dofs 600" in r.stdout


@pytest.mark.gpu
def test_facade_gpu(gpu):
    r = subprocess.run([_exe(), B.SPHERE_DAT, "--gpu"], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
