"""Small solves and products run in a child process with BMI_POISON_ALLOC=1 (every floating-point
buffer the library allocates starts as NaN bit patterns).  A kernel or solver step that reads
values it has not written then turns its result into NaN on every run, instead of only when the
allocator hands back memory that held such bit patterns.  Prints one JSON line per case:
{"case": name, "status": ..., "finite": bool, "err": relative error}.  Run by
tests/test_gpu_poison.py; not a test module itself."""
import json
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), ".."))
import bembel_amd as B  # noqa: E402

ELLIPSOID = (1.0, 0.7, 0.5)


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def report(case, status, x, err):
    print(json.dumps(dict(case=case, status=int(status), finite=bool(np.all(np.isfinite(x))), err=err)), flush=True)


def gmres(H, b, tol=1e-10):
    g = B.GMRES()
    g.set_tolerance(tol)
    g.set_restart(100)
    g.set_max_iterations(4000)
    g.compute(H)
    x = g.solve(b)
    return x, g.info()


def helmholtz_solves():
    op = B.HelmholtzSingleLayerOperator()
    op.set_wavenumber(3.0)
    sp = B.AnsatzSpace(B.Geometry(scale=ELLIPSOID), 3, 2, op)
    b = B.DiscreteLinearForm(B.PlaneWaves([[0.0, 0.6, 0.8], [1.0, 0.0, 0.0]], 3.0), sp).compute(
    ).get_discrete_linear_form()
    H64 = B.H2Matrix(sp, op, n_cheb=5)
    for name, prec in (("fp64", B.FP64), ("mixed", B.MIXED)):
        H = H64 if prec == B.FP64 else B.H2Matrix(sp, op, n_cheb=5, precision=prec)
        x, info = gmres(H, b)
        report("helmholtz_gmres_" + name, info, x, rel(H64 @ x, b))


def maxwell_solve():
    kappa = 2.0
    op = B.MaxwellSingleLayerOperator()
    op.set_wavenumber(kappa)
    sp = B.AnsatzSpace(B.Geometry(scale=ELLIPSOID), 2, 2, op)
    H = B.H2Matrix(sp, op, n_cheb=5)
    pos, mom = np.array([0.1, -0.05, 0.08]), np.array([0.02, 0.05, -0.07])
    b = B.DiscreteLinearForm(B.RotatedTangentialTrace(pos, mom, kappa), sp).compute().get_discrete_linear_form()
    x, info = gmres(H, b)
    report("maxwell_gmres", info, x, rel(H @ x, b))


def laplace_cg_mixed():
    op = B.LaplaceSingleLayerOperator()
    sp = B.AnsatzSpace(B.Geometry(), 3, 2)
    H64 = B.H2Matrix(sp, op, n_cheb=4)
    H = B.H2Matrix(sp, op, n_cheb=4, precision=B.MIXED)
    b = B.DiscreteLinearForm(B.HarmonicFunction(), sp).compute().get_discrete_linear_form()
    g = B.ConjugateGradient()
    g.set_tolerance(1e-10)
    g.set_max_iterations(4000)
    g.compute(H)
    x = g.solve(b)
    report("laplace_cg_mixed", g.info(), x, rel(H64 @ x, b))


def sweep_forms():
    rng = np.random.default_rng(5)
    for name, kappa, np_ in (("laplace", None, 4), ("helmholtz", 4.0, 5)):
        if kappa is None:
            op = B.LaplaceSingleLayerOperator()
        else:
            op = B.HelmholtzSingleLayerOperator()
            op.set_wavenumber(kappa)
        sp = B.AnsatzSpace(B.Geometry(), 3, 2, op)
        n = sp.number_of_dofs
        X = rng.uniform(-1, 1, (n, 12))
        if kappa is not None:
            X = X + 1j * rng.uniform(-1, 1, (n, 12))
        for prec, pname in ((B.FP64, "fp64"), (B.MIXED, "mixed")):
            H = B.H2Matrix(sp, op, n_cheb=np_, precision=prec)
            H.set_sweep("full")
            ref = np.stack([H @ X[:, k] for k in range(12)], axis=1)
            for mode in ("upper", "upper_near", None):
                H.set_sweep(mode)
                y = H @ X[:, 0]
                report("%s_%s_sweep_%s" % (name, pname, mode), B.OK, y, rel(y, ref[:, 0]))
            for cols in (3, 12):
                Y = H @ X[:, :cols]
                report("%s_%s_%d_columns" % (name, pname, cols), B.OK, Y, rel(Y, ref[:, :cols]))


def server_mix():
    reqs = []
    for k, (M, P) in enumerate(((3, 2), (3, 1), (2, 2))):
        if k == 1:
            op = B.LaplaceSingleLayerOperator()
            data = B.DirichletTrace(np.array([[2.0, 1.0, -2.0]]))
        else:
            op = B.HelmholtzSingleLayerOperator()
            op.set_wavenumber(1.5 + k)
            data = B.PlaneWaves([[0.0, 0.6, 0.8]], 1.5 + k)
        reqs.append((op, M, P, data))
    srv = B.Server(max_active=4, max_batch_rhs=64, slice_iterations=30)
    ids = [srv.submit(op, M, P, data, scale=ELLIPSOID, n_cheb=5, precision=B.MIXED, tol=1e-10, restart=100,
                      maxit=6000) for op, M, P, data in reqs]
    res = [srv.result(i, timeout=200) for i in ids]
    srv.close()
    for (op, M, P, data), (x, info) in zip(reqs, res):
        sp = B.AnsatzSpace(B.Geometry(scale=ELLIPSOID), M, P, op)
        H = B.H2Matrix(sp, op, n_cheb=5)
        b = B.DiscreteLinearForm(data, sp).compute().get_discrete_linear_form()
        ok = x is not None
        report("server_mixed_M%d_P%d" % (M, P), info["status"], x if ok else np.array([np.nan]),
               rel(H @ x, b) if ok else float("nan"))


if __name__ == "__main__":
    if os.environ.get("BMI_POISON_ALLOC", "0") in ("", "0"):
        sys.exit("run with BMI_POISON_ALLOC=1")
    for fn in (helmholtz_solves, maxwell_solve, laplace_cg_mixed, sweep_forms, server_mix):
        print("# " + fn.__name__, flush=True)
        fn()
