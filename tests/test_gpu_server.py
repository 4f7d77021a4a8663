"""Serving runtime on the GPU: results in a mix (batched, several batches' GMRES tasks in flight
on their own streams at once) equal the results of the same requests solved alone, resident
operators are reused, and the solutions satisfy the FP64 operator."""
import numpy as np
import pytest

import bembel_amd as B

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def requests():
    rng = np.random.default_rng(17)
    out = []
    for k in range(8):
        scale = tuple(rng.uniform(0.5, 1.5, 3))
        M = int(rng.choice([2, 3]))
        P = int(rng.choice([1, 2]))
        nrhs = int(rng.integers(1, 4))
        if k % 2 == 0:
            op = B.LaplaceSingleLayerOperator()
            pts = rng.normal(size=(nrhs, 3))
            pts = 3.0 * pts / np.linalg.norm(pts, axis=1, keepdims=True)
            data = B.DirichletTrace(pts)
        else:
            kappa = float(rng.uniform(1, 4))
            op = B.HelmholtzSingleLayerOperator()
            op.set_wavenumber(kappa)
            d = rng.normal(size=(nrhs, 3))
            data = B.PlaneWaves(d / np.linalg.norm(d, axis=1, keepdims=True), kappa)
        out.append(dict(op=op, M=M, P=P, data=data, scale=scale))
    # a repeat of the first request (same operator): reuses the resident operator or joins its batch
    out.append(dict(out[0]))
    return out


def solve_all(srv, reqs):
    ids = [srv.submit(r["op"], r["M"], r["P"], r["data"], scale=r["scale"], n_cheb=5, tol=1e-10) for r in reqs]
    return [srv.result(i, timeout=240) for i in ids]


@pytest.mark.skip(reason="the thread-per-batch design it tested is gone: its detached workers could outlive the "
                         "server (a worker's last operator reference freed device memory after bmi_server_destroy "
                         "returned; nothing joined the workers at exit), which ended once in a host segmentation "
                         "fault inside a worker thread; the server now drives all batches from one joined device "
                         "thread (test_gpu_server_overlapping_mix_equals_alone)")
def test_gpu_server_mix_equals_alone(gpu):
    reqs = requests()
    alone = []
    srv1 = B.Server(max_active=1, max_batch_rhs=1)
    for r in reqs:
        alone.append(solve_all(srv1, [r])[0])
    srv1.close()
    srv = B.Server(max_active=4, max_batch_rhs=64)
    mixed = solve_all(srv, reqs)
    st = srv.stats()
    srv.close()
    assert st["finished"] == len(reqs)
    batched = 0
    for (xa, ia), (xm, im), r in zip(alone, mixed, reqs):
        assert ia["status"] == B.OK and im["status"] == B.OK, (ia, im)
        assert ia["residual"] <= 1e-10 and im["residual"] <= 1e-10
        # same tolerance-level answer alone and in the mix (batched columns, other streams busy)
        assert rel(xm, xa) <= 1e-8, (r["M"], r["P"], rel(xm, xa))
        batched += im["batch_requests"] > 1
    # the repeated request found its operator resident or shared a batch with its twin
    assert mixed[-1][1]["cache_hit"] or mixed[-1][1]["batch_requests"] > 1
    print("batches with more than one request:", batched)


def test_gpu_server_overlapping_mix_equals_alone(gpu):
    """Several batches admitted at once run as GMRES tasks on their own streams, each enqueueing
    its next work whenever its convergence flags arrive, so their kernels are in flight on the
    device together; a task's device work is that of a solve alone, so every result equals the
    same request solved alone (one batch at a time, one request per batch) to the batching
    tolerance."""
    reqs = requests()
    srv1 = B.Server(max_active=1, max_batch_rhs=1, slice_iterations=0)
    alone = [solve_all(srv1, [r])[0] for r in reqs]
    srv1.close()
    srv = B.Server(max_active=4, max_batch_rhs=64, slice_iterations=30)
    mixed = solve_all(srv, reqs)
    st = srv.stats()
    srv.close()
    assert st["finished"] == len(reqs)
    for (xa, ia), (xm, im), r in zip(alone, mixed, reqs):
        assert ia["status"] == B.OK and im["status"] == B.OK, (ia, im)
        assert rel(xm, xa) <= 1e-8, (r["M"], r["P"], rel(xm, xa))
    assert mixed[-1][1]["cache_hit"] or mixed[-1][1]["batch_requests"] > 1
    # batches took many turns, and batches were admitted while others were running
    assert max(i["slices"] for _, i in mixed) > 1
    spans = [(i["start_s"], i["end_s"]) for _, i in mixed]
    assert any(a[0] < b[1] and b[0] < a[1] and a != b for a in spans for b in spans)


def test_gpu_server_solution_satisfies_operator(gpu):
    op = B.HelmholtzSingleLayerOperator()
    op.set_wavenumber(3.0)
    data = B.PlaneWaves([[0.0, 0.6, 0.8]], 3.0)
    srv = B.Server(max_active=1)
    x, info = srv.result(srv.submit(op, 3, 2, data, scale=(1.0, 0.7, 0.5), n_cheb=5, tol=1e-10))
    srv.close()
    sp = B.AnsatzSpace(B.Geometry(scale=(1.0, 0.7, 0.5)), 3, 2, op)
    H = B.H2Matrix(sp, op, n_cheb=5)
    b = B.DiscreteLinearForm(data, sp).compute().get_discrete_linear_form()
    assert info["status"] == B.OK and rel(H @ x, b) <= 2e-10


def test_gpu_server_stream_equals_alone(gpu):
    """One batch on the device at a time: requests queued while another is solved are admitted in
    turn, equal operators share a batch or find their operator resident, and every result equals
    the same request solved alone."""
    reqs = requests()
    srv1 = B.Server(max_active=1, max_batch_rhs=1)
    alone = [solve_all(srv1, [r])[0] for r in reqs]
    srv1.close()
    srv = B.Server(max_active=1, max_batch_rhs=64)
    mixed = solve_all(srv, reqs)
    srv.close()
    for (xa, ia), (xm, im), r in zip(alone, mixed, reqs):
        assert ia["status"] == B.OK and im["status"] == B.OK, (ia, im)
        assert rel(xm, xa) <= 1e-8, (r["M"], r["P"], rel(xm, xa))
    assert mixed[-1][1]["cache_hit"] or mixed[-1][1]["batch_requests"] > 1


@pytest.mark.parametrize("kind,M,P,prec,nrhs", [("laplace", 4, 2, "fp64", 1), ("laplace", 4, 1, "mixed", 3),
                                                 ("helmholtz", 4, 2, "mixed", 1), ("helmholtz", 3, 2, "fp64", 12),
                                                 ("maxwell", 3, 2, "fp64", 2)])
def test_gpu_admission_estimate_matches_device_bytes(kind, M, P, prec, nrhs):
    """The admission estimate (block tree on the host, nothing allocated) against the bytes the
    handle holds after an nrhs-column product, workspace included: the server admits by it, so it
    must not undershoot, nor overshoot by much (measured: 0.4 to 7.5 % above)."""
    import torch
    geo = B.Geometry(scale=(1.0, 0.7, 0.5)) if kind == "maxwell" else B.Geometry()
    if kind == "laplace":
        op = B.LaplaceSingleLayerOperator()
    elif kind == "helmholtz":
        op = B.HelmholtzSingleLayerOperator()
        op.set_wavenumber(2.0)
    else:
        op = B.MaxwellSingleLayerOperator()
        op.set_wavenumber(2.0)
    sp = B.AnsatzSpace(geo, M, P, op)
    precision = B.MIXED if prec == "mixed" else B.FP64
    d = op.desc(n_cheb=4, eta=1.6, min_cluster_level=1, precision=precision)
    est = np.zeros(1, dtype=np.int64)
    B._check(B.library().bmi_h2_estimate_bytes(sp._h, B.C.byref(d), nrhs,
                                                est.ctypes.data_as(B.C.POINTER(B.C.c_int64))))
    H = B.H2Matrix(sp, op, n_cheb=4, precision=precision)
    n = sp.number_of_dofs
    dt = torch.complex128 if H.scalar == np.complex128 else torch.float64
    x = torch.ones((n, nrhs), dtype=dt, device="cuda") if nrhs > 1 else torch.ones(n, dtype=dt, device="cuda")
    H.matvec(x)
    info = np.zeros(16, dtype=np.int64)
    B._check(B.library().bmi_h2_info(H._h, info.ctypes.data_as(B.C.POINTER(B.C.c_int64))))
    ratio = est[0] / info[8]
    print("%s M%d P%d %s nrhs %d: estimate %.1f MB, held %.1f MB, ratio %.3f" % (
        kind, M, P, prec, nrhs, est[0] / 1e6, info[8] / 1e6, ratio))
    assert 1.0 <= ratio <= 1.2
