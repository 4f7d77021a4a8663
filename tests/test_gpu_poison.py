"""No kernel or solver step reads device values it has not written.

The cases of tests/poison_cases.py (GMRES for Helmholtz in FP64 and mixed precision, Maxwell
GMRES, mixed-precision CG, every single-RHS sweep form and the few- and many-column products of
the symmetric kernels, a mixed-precision serving mix) run in a child process with
BMI_POISON_ALLOC=1: each floating-point buffer the library allocates starts as NaN bit patterns,
so a read of an unwritten value gives NaN on every run.  Without the fill such a read only shows
when the allocator hands back memory that happened to hold NaN patterns (a GMRES basis vector
formed as a X + 0 * Y over fresh memory did exactly that)."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def limit(case):
    if "_sweep_" in case or "_columns" in case:
        return 1e-5 if "_mixed_" in case else 1e-12  # FP32 blocks: rounding of the reduced operator
    if case == "maxwell_gmres":
        return 1e-9
    return 2e-10  # FP64 residual of the solves (tolerance 1e-10)


def test_gpu_no_kernel_reads_unwritten_values(gpu):
    env = dict(os.environ, BMI_POISON_ALLOC="1")
    r = subprocess.run([sys.executable, "-u", os.path.join(HERE, "poison_cases.py")], env=env, capture_output=True,
                       text=True, timeout=240)
    print(r.stdout[-6000:])
    assert r.returncode == 0, r.stderr[-4000:]
    rows = [json.loads(line) for line in r.stdout.splitlines() if line.startswith("{")]
    assert len(rows) == 27, len(rows)
    for row in rows:
        assert row["status"] == 0 and row["finite"], row
        assert row["err"] <= limit(row["case"]), row
