"""Serving runtime: scheduling and failure handling on the CPU (no device work is reached: the
admission check rejects, or the batch fails at its first HIP call and reports it)."""
import numpy as np
import pytest

import bembel_amd as B


def test_admission_rejects_requests_above_the_budget():
    srv = B.Server(hbm_budget_bytes=1e6, max_active=2)
    rid = srv.submit(B.LaplaceSingleLayerOperator(), 3, 1, B.DirichletTrace([[2.0, 0, 0]]), n_cheb=4)
    x, info = srv.result(rid, timeout=60)
    assert x is None and info["status"] == B.ADMISSION_REJECTED
    assert "budget" in info["error"]
    st = srv.stats()
    assert st["finished"] == 1 and st["queued"] == 0 and st["reserved_bytes"] == 0
    srv.close()


def test_estimate_grows_with_the_problem():
    g = B.Geometry()
    sizes = []
    for M in (2, 3, 4):
        sp = B.AnsatzSpace(g, M, 2)
        d = B.LaplaceSingleLayerOperator().desc(n_cheb=4)
        b = np.zeros(1, dtype=np.int64)
        B._check(B.library().bmi_h2_estimate_bytes(sp._h, B.C.byref(d), 1, b.ctypes.data_as(B.C.POINTER(B.C.c_int64))))
        sizes.append(int(b[0]))
    assert sizes[0] < sizes[1] < sizes[2]
    # a mixed-precision operator also holds the FP32 coupling copy
    d = B.LaplaceSingleLayerOperator().desc(n_cheb=4, precision=B.MIXED)
    b = np.zeros(1, dtype=np.int64)
    sp = B.AnsatzSpace(g, 4, 2)  # keep the space alive across the call
    B._check(B.library().bmi_h2_estimate_bytes(sp._h, B.C.byref(d), 1, b.ctypes.data_as(B.C.POINTER(B.C.c_int64))))
    assert int(b[0]) > sizes[2]


def test_invalid_requests_are_refused_at_submit():
    srv = B.Server()
    with pytest.raises(B.BembelError):
        # Laplace takes point sources, not plane waves
        srv.submit(B.LaplaceSingleLayerOperator(), 2, 1, B.PlaneWaves([[0, 0, 1.0]], 2.0))
    with pytest.raises(B.BembelError):
        srv.submit(B.LaplaceSingleLayerOperator(), 2, 1, B.DirichletTrace([[2.0, 0, 0]]), tol=-1.0)
    srv.close()


def test_failures_are_reported_not_hung():
    """Without a GPU every admitted batch fails at its first HIP call; each request still
    finishes with that status and the server drains and closes."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("exercised on the CPU only")
    srv = B.Server(max_active=2)
    ids = [srv.submit(B.LaplaceSingleLayerOperator(), 2, 1, B.DirichletTrace([[2.0 + k, 0, 0]]), n_cheb=4)
           for k in range(5)]
    for rid in ids:
        x, info = srv.result(rid, timeout=60)
        assert x is None and info["status"] == B.DEVICE_ERROR
    st = srv.stats()
    assert st["finished"] == 5 and st["running"] == 0 and st["reserved_bytes"] == 0
    srv.close()
