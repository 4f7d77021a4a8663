#!/usr/bin/env python3
"""Benchmark of the H2 hot path on one MI355X.

Workload: Laplace single-layer operator on the 6-patch unit sphere, refinement level 6
(24,576 elements), B-spline degree 2 (26,136 DOFs), H2 with 4x4 Chebyshev nodes per cluster,
eta = 1.6.  A "step" is one H2 matrix-vector product (1 RHS) with the vectors resident in HBM;
`value` is DOF/s = DOFs / step time.  The line also carries the H2 assembly time, CG solve
latency (p50/p99, FP64 and mixed precision, with the mixed solution's error against FP64), a
roofline object for the dominant matvec kernel (near-field leaf sweep, HBM-bound) and for the
near-field quadrature (FP64-bound), and the CPU baseline (the oracle's OpenMP matvec on the host
cores, on the same leaves, and its assembly time from a sample calibrated against a timed full
assembly).  Further records: the Helmholtz 64-RHS workload (6,144 elements), the Helmholtz
24,576-element assembly + GMRES in FP64 and mixed precision, the Maxwell EFIE workload (1,536
elements), and the serving stream (128 requests), each with its CPU baseline; the CPU legs run
after all GPU legs, with OpenMP threads bound one per physical core.

usage: python bench.py [--gpus 1] [--steps 20] [--warmup 3]
"""
import argparse
import gc
import json
import math
import os
import sys
import time

# The OpenMP runtime (torch's libgomp, which the oracle shares) reads its placement at load time:
# the CPU baseline's threads are bound one per physical core, close together.  The runtime binds
# the loading thread to the first core as well; that binding is undone right after the load (and
# after every CPU leg), so the GPU side's host threads keep the process's whole CPU set.
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")
AFFINITY = sorted(os.sched_getaffinity(0))

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def pct(a, q):
    return float(np.percentile(np.asarray(a), q))


def nearfield_flops(classes_deg, P, pd, cplx=False):
    """Algorithmic FP64 flops of the Duffy pairs as the rules compute them, n Gauss points per
    direction.  Identical: 8 n^4 surface points, 4 n^4 kernel values, each added to the block and
    to its transpose ((P+1)^4 updates each).  Edge (factorised): 6 n^4 evaluations, each adding kernel x basis to a
    (P+1)^2 vector, one outer product per outer point (6 n^3), 8 n^3 tabulated surface points, 4 n^3
    contracted patch lines and 4 n^4 points on them.  Vertex (factorised): 4 n^4 evaluations into
    vectors, 4 n^2 outer products, 4 n^2 + 2 n^3 surface points."""
    I2 = (P + 1) ** 2
    c = 2 if cplx else 1
    Fk = 40 if cplx else 20                   # kernel value (distance, 1/r, sincos)
    Fe = 2 * I2 * I2 * c + I2                 # full block update of one evaluation
    Fv = I2 + 2 * I2 * c + 6 * (P + 1)        # basis vector and its accumulation
    Fo = 2 * I2 * I2 * c                      # outer product of two basis vectors
    Fg = 16 * pd * pd + 24 * pd + 90          # surface point with both tangents (2-D patch evaluation)
    Fl = 16 * pd * pd + 20                    # patch contracted to a line (values and cross derivatives)
    F1 = 24 * pd + 60                         # point on a contracted line
    tot = 0.0
    for cls, n, cnt in classes_deg:
        n = float(n)
        if cls == 2:
            f = 8 * n ** 4 * (Fe + Fg) + 4 * n ** 4 * Fk
        elif cls == 3:
            f = 6 * n ** 4 * (Fk + Fv) + 6 * n ** 3 * Fo + 8 * n ** 3 * Fg + 4 * n ** 3 * Fl + 4 * n ** 4 * F1
        else:
            f = 4 * n ** 4 * (Fk + Fv) + 4 * n ** 2 * Fo + (4 * n ** 2 + 2 * n ** 3) * Fg
        tot += cnt * f
    return tot


def measured_traffic(M, P, npts, upper):
    """HBM bytes per launch of the product sweep (FETCH_SIZE + WRITE_SIZE, rocprofv3 --pmc, scaled by a
    1 GiB copy) of the chosen form (every coupling block, or the upper blocks), as recorded under
    profiles/ for this workload; None for other sizes"""
    f = os.path.join(ROOT, "profiles", "k_sweep_fused_laplace_M%d_P%d_np%d_traffic.json" % (M, P, npts))
    if not os.path.exists(f):
        return None
    with open(f) as fh:
        d = json.load(fh)["upper" if upper else "full"]
    return d["fetch_bytes"] + d["write_bytes"]


def streaming_read_gbs(torch, mib=2048, reps=10):
    """HBM read rate of a plain streaming reduction (torch.sum over a resident FP64 buffer, HIP
    events): a measured reference for what a read-only sweep can reach, beside the 8 TB/s peak"""
    x = torch.ones(mib * 1024 * 1024 // 8, dtype=torch.float64, device="cuda")
    x.sum()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(reps):
        x.sum()
    ev1.record()
    ev1.synchronize()
    gbs = reps * x.numel() * 8 / (ev0.elapsed_time(ev1) * 1e-3) / 1e9
    del x
    torch.cuda.empty_cache()
    return gbs


def run_batched_profile(B, torch):
    """Helmholtz kappa 4, 6,144 elements, P 2, 5x5 Chebyshev nodes, 64 plane-wave right-hand sides:
    per precision an assembly, 8 batched products and 60 batched GMRES(30) steps (the workload of the
    64-RHS kernel statistics and matrix-core counters)"""
    op = B.HelmholtzSingleLayerOperator()
    op.set_wavenumber(4.0)
    space = B.AnsatzSpace(B.Geometry(), 5, 2, op)
    rng = np.random.default_rng(3)
    rhs = None
    for name, prec in (("fp64", B.FP64), ("mixed", B.MIXED)):
        H = B.H2Matrix(space, op, n_cheb=5, precision=prec)
        if rhs is None:
            rhs = B.DiscreteLinearForm(B.PlaneWaves(unit_directions(rng, 64), 4.0), space).compute(
                device=True).get_discrete_linear_form()
        for _ in range(3):
            H.matvec(rhs)
        B.synchronize()
        t0 = time.perf_counter()
        for _ in range(5):
            H.matvec(rhs, sync=False)
        B.synchronize()
        log(name, "64-RHS matvec %.3f ms" % ((time.perf_counter() - t0) / 5 * 1e3))
        t0 = time.perf_counter()
        gmres_solve(B, H, rhs, tol=1e-12, maxit=60)
        torch.cuda.synchronize()
        log(name, "60 batched GMRES steps %.3f s" % (time.perf_counter() - t0))
        del H


def run_assembly_profile(B, torch):
    """one H2 assembly each of the Laplace (4x4 nodes) and Helmholtz (kappa 4, 5x5 nodes) single
    layers on 24,576 elements, P 2 (the workload of the near-field quadrature counters)"""
    for name, op, npts in (("laplace", B.LaplaceSingleLayerOperator(), 4), ("helmholtz", B.HelmholtzSingleLayerOperator(), 5)):
        if name == "helmholtz":
            op.set_wavenumber(4.0)
        space = B.AnsatzSpace(B.Geometry(), 6, 2, op)
        t0 = time.perf_counter()
        H = B.H2Matrix(space, op, n_cheb=npts)
        torch.cuda.synchronize()
        log(name, "assembly %.3f s, pair classes %s" % (time.perf_counter() - t0, H.info["pair_classes"]))
        del H


def run_helmholtz_profile(B, torch):
    """Helmholtz kappa 4, 24,576 elements, P 2, 5x5 Chebyshev nodes: per precision an assembly, 23
    products and 300 GMRES(30) steps (the workload of the Helmholtz kernel statistics)"""
    op = B.HelmholtzSingleLayerOperator()
    op.set_wavenumber(4.0)
    space = B.AnsatzSpace(B.Geometry(), 6, 2, op)
    n = space.number_of_dofs
    for name, prec in (("fp64", B.FP64), ("mixed", B.MIXED)):
        t0 = time.perf_counter()
        H = B.H2Matrix(space, op, n_cheb=5, precision=prec)
        torch.cuda.synchronize()
        log(name, "assembly %.3f s" % (time.perf_counter() - t0))
        x = torch.rand(n, dtype=torch.float64, device="cuda") + 1j * torch.rand(n, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        for _ in range(3):
            H.matvec(x, sync=False)
        B.synchronize()
        t0 = time.perf_counter()
        for _ in range(20):
            H.matvec(x, sync=False)
        B.synchronize()
        log(name, "matvec %.3f ms" % ((time.perf_counter() - t0) / 20 * 1e3))
        b = B.DiscreteLinearForm(B.PlaneWaves([[0.0, 0.0, 1.0]], 4.0), space).compute(
            device=True).get_discrete_linear_form()
        t0 = time.perf_counter()
        gmres_solve(B, H, b, maxit=300)
        torch.cuda.synchronize()
        log(name, "300 GMRES steps %.3f s" % (time.perf_counter() - t0))
        del H


def run_solves(B, torch, H, space, op, npts, b, nsolves):
    """CG to 1e-10 in FP64, and with the mixed-precision operator inside FP64 iterative refinement;
    latencies and the mixed solution's error against the FP64 solution"""
    cg = B.ConjugateGradient()
    cg.set_tolerance(1e-10)
    cg.set_max_iterations(2000)
    cg.compute(H)
    lat64 = []
    for _ in range(nsolves):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        u64 = cg.solve(b)
        torch.cuda.synchronize()
        lat64.append(time.perf_counter() - t0)
    it64 = int(cg.iterations()[0])
    Hm = B.H2Matrix(space, op, n_cheb=npts, precision=B.MIXED)
    cgm = B.ConjugateGradient()
    cgm.set_tolerance(1e-10)
    cgm.set_max_iterations(2000)
    cgm.compute(Hm)
    latm = []
    for _ in range(nsolves):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        um = cgm.solve(b)
        torch.cuda.synchronize()
        latm.append(time.perf_counter() - t0)
    rel_err_mixed = float(torch.linalg.norm(um - u64) / torch.linalg.norm(u64))
    log("CG FP64: %d it, p50 %.1f ms; mixed: %d it, p50 %.1f ms, rel err %.2e" % (
        it64, 1e3 * pct(lat64, 50), int(cgm.iterations()[0]), 1e3 * pct(latm, 50), rel_err_mixed))

    return lat64, latm, it64, rel_err_mixed


def unit_directions(rng, k):
    d = rng.normal(size=(k, 3))
    return d / np.linalg.norm(d, axis=1, keepdims=True)


def timed(torch, fn, reps=1):
    import bembel_amd as B
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        out = fn()
    B.synchronize()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps, out


def gmres_solve(B, H, b, tol=1e-10, restart=30, maxit=20000):
    gm = B.GMRES()
    gm.set_tolerance(tol)
    gm.set_restart(restart)
    gm.set_max_iterations(maxit)
    gm.compute(H)
    x = gm.solve(b)
    return x, gm


def run_batched(B, torch, rng, steps):
    """Helmholtz single layer, kappa 4, 6,144 elements, P 2, 5x5 Chebyshev nodes, 64 plane-wave
    right-hand sides: batched product throughput and batched GMRES (FP64 and mixed)"""
    M, P, npts, kappa, nrhs = 5, 2, 5, 4.0, 64
    op = B.HelmholtzSingleLayerOperator()
    op.set_wavenumber(kappa)
    space = B.AnsatzSpace(B.Geometry(), M, P, op)
    n = space.number_of_dofs
    t_asm, H = timed(torch, lambda: B.H2Matrix(space, op, n_cheb=npts))
    Hm = B.H2Matrix(space, op, n_cheb=npts, precision=B.MIXED)
    rhs = B.DiscreteLinearForm(B.PlaneWaves(unit_directions(rng, nrhs), kappa), space).compute(
        device=True).get_discrete_linear_form()
    out = {"workload": "helmholtz_single_layer_h2_sphere_M5_P2_np5_kappa4_64rhs", "dofs": n, "nrhs": nrhs,
           "h2_assembly_s": t_asm}
    for name, HH in (("fp64", H), ("mixed", Hm)):
        for _ in range(2):
            HH.matvec(rhs)
        t, y = timed(torch, lambda: HH.matvec(rhs, sync=False), max(3, steps // 4))
        out[name + "_matvec_ms"] = t * 1e3
        out[name + "_matvec_dof_rhs_per_s"] = n * nrhs / t
        r1 = rhs[:, :1].contiguous()
        HH.matvec(r1)  # first single-RHS product of the handle (graph capture)
        t1, _ = timed(torch, lambda: HH.matvec(r1, sync=False), max(20, steps))
        out[name + "_matvec_1rhs_ms"] = t1 * 1e3
    log("batched 64 RHS: matvec fp64 %.2f ms, mixed %.2f ms" % (out["fp64_matvec_ms"], out["mixed_matvec_ms"]))
    # matrix-core roofline of the FP64 64-RHS block sweeps: algorithmic flops (8 per complex
    # multiply-add, per right-hand side) over the average launch time (HIP events in the library)
    B.timing.reset()
    B.timing.enable(True)
    keep = [H.matvec(rhs, sync=False) for _ in range(3)]  # queued back to back
    B.synchronize()
    B.timing.enable(False)
    del keep
    info = H.info
    sweeps = {}
    for key, name, blocks, dim in (("coupling_apply", "coupling", info["lowrank"], info["coupling_dim"]),
                                   ("matvec_nearfield", "near_field", info["dense"], info["dense_dim"])):
        ms, cnt = B.timing.get(key)
        flops = 8.0 * blocks * dim * dim * nrhs
        avg = ms / max(cnt, 1)
        sweeps[name] = {"bound": "mfma", "achieved": flops / (avg * 1e-3) / 1e12, "peak": 78.6, "unit": "TFLOP/s",
                        "frac": flops / (avg * 1e-3) / 78.6e12, "flops_per_launch": flops, "avg_launch_ms": avg,
                        "kernel": "k_multi_mfma<complex, FP64, %d> (v_mfma_f64_16x16x4f64)" % dim}
    out["roofline_64rhs"] = sweeps
    log("batched 64 RHS sweeps on the matrix cores: coupling %.1f TFLOP/s (%.3f ms), near field %.1f TFLOP/s "
        "(%.3f ms)" % (sweeps["coupling"]["achieved"], sweeps["coupling"]["avg_launch_ms"],
                       sweeps["near_field"]["achieved"], sweeps["near_field"]["avg_launch_ms"]))
    t, (x64, gm) = timed(torch, lambda: gmres_solve(B, H, rhs, maxit=8000))
    out["fp64_gmres_s"] = t
    out["fp64_gmres_iterations_max"] = int(np.max(gm.iterations()))
    log("batched 64 RHS: GMRES fp64 %.2f s (%d it)" % (t, out["fp64_gmres_iterations_max"]))
    t, (xm, gmm) = timed(torch, lambda: gmres_solve(B, Hm, rhs, maxit=8000))
    out["mixed_gmres_s"] = t
    out["mixed_gmres_iterations_max"] = int(np.max(gmm.iterations()))
    out["mixed_rel_err_vs_fp64_max"] = float(torch.max(torch.linalg.norm(xm - x64, dim=0) /
                                                   torch.linalg.norm(x64, dim=0)))
    log("batched 64 RHS: matvec fp64 %.2f ms, mixed %.2f ms (1 RHS %.2f ms); GMRES fp64 %.2f s (%d it), mixed %.2f s, "
        "err %.1e" % (out["fp64_matvec_ms"], out["mixed_matvec_ms"], out["fp64_matvec_1rhs_ms"], out["fp64_gmres_s"],
                      out["fp64_gmres_iterations_max"], out["mixed_gmres_s"], out["mixed_rel_err_vs_fp64_max"]))
    del Hm
    return out, H


def run_headline(B, torch, rng):
    """Helmholtz single layer, kappa 4, 24,576 elements, P 2, 5x5 Chebyshev nodes, one plane wave:
    H2 assembly + GMRES(30) to 1e-10, plain FP64 and mixed precision.  Returns the record and the
    FP64 operator (its leaves feed the CPU baseline)."""
    M, P, npts, kappa = 6, 2, 5, 4.0
    op = B.HelmholtzSingleLayerOperator()
    op.set_wavenumber(kappa)
    space = B.AnsatzSpace(B.Geometry(), M, P, op)
    n = space.number_of_dofs
    d = unit_directions(rng, 1)
    out = {"workload": "helmholtz_single_layer_h2_sphere_M6_P2_np5_kappa4", "dofs": n}
    sols = {}
    keep = None
    for name, prec in (("fp64", B.FP64), ("mixed", B.MIXED)):
        t0 = time.perf_counter()
        H = B.H2Matrix(space, op, n_cheb=npts, precision=prec)
        b = B.DiscreteLinearForm(B.PlaneWaves(d, kappa), space).compute(device=True).get_discrete_linear_form()
        torch.cuda.synchronize()
        t_asm = time.perf_counter() - t0
        t_sol, (x, gm) = timed(torch, lambda: gmres_solve(B, H, b))
        sols[name] = x
        out[name] = {"assembly_s": t_asm, "gmres_s": t_sol, "total_s": t_asm + t_sol,
                     "iterations": int(gm.iterations()[0]), "residual": float(gm.error()[0])}
        log("Helmholtz M6 %s: assembly %.2f s, GMRES %.2f s (%d it)" % (name, t_asm, t_sol, out[name]["iterations"]))
        if name == "fp64":
            keep = H
        else:
            del H
    out["mixed_rel_err_vs_fp64"] = float(torch.linalg.norm(sols["mixed"] - sols["fp64"]) /
                                         torch.linalg.norm(sols["fp64"]))
    out["mixed_speedup"] = out["fp64"]["total_s"] / out["mixed"]["total_s"]
    log("Helmholtz M6: FP64 assembly %.2f s + GMRES %.2f s (%d it); mixed %.2f + %.2f s; err %.1e" % (
        out["fp64"]["assembly_s"], out["fp64"]["gmres_s"], out["fp64"]["iterations"], out["mixed"]["assembly_s"],
        out["mixed"]["gmres_s"], out["mixed_rel_err_vs_fp64"]))
    return out, keep


def run_maxwell(B, torch, rng):
    """Maxwell single layer (EFIE), kappa 2, ellipsoid (1, 0.7, 0.5), 1,536 elements, P 2, 9x9
    Chebyshev nodes, tangential trace of a dipole field at a seeded interior point and moment:
    H2 assembly + GMRES(100) to 1e-8 with FP64 Arnoldi, plain FP64 and mixed precision.  Returns
    the record and the FP64 operator."""
    M, P, npts, kappa = 4, 2, 9, 2.0
    op = B.MaxwellSingleLayerOperator()
    op.set_wavenumber(kappa)
    space = B.AnsatzSpace(B.Geometry(scale=(1.0, 0.7, 0.5)), M, P, op)
    pos = 0.2 * rng.uniform(-1, 1, 3)
    mom = rng.normal(size=3)
    out = {"workload": "maxwell_single_layer_h2_ellipsoid_M4_P2_np9_kappa2", "dofs": space.number_of_dofs,
           "elements": space.number_of_elements, "dipole_position": pos.tolist(), "dipole_moment": mom.tolist()}
    sols, keep = {}, None
    for name, prec in (("fp64", B.FP64), ("mixed", B.MIXED)):
        t0 = time.perf_counter()
        H = B.H2Matrix(space, op, n_cheb=npts, precision=prec)
        b = B.DiscreteLinearForm(B.RotatedTangentialTrace(pos, mom, kappa), space).compute(
            device=True).get_discrete_linear_form()
        torch.cuda.synchronize()
        t_asm = time.perf_counter() - t0
        t_sol, (x, gm) = timed(torch, lambda: gmres_solve(B, H, b, tol=1e-8, restart=100))
        sols[name] = x
        out[name] = {"assembly_s": t_asm, "gmres_s": t_sol, "total_s": t_asm + t_sol,
                     "iterations": int(gm.iterations()[0]), "residual": float(gm.error()[0])}
        if name == "fp64":
            keep = H
        else:
            del H
    out["mixed_rel_err_vs_fp64"] = float(torch.linalg.norm(sols["mixed"] - sols["fp64"]) /
                                         torch.linalg.norm(sols["fp64"]))
    out["mixed_speedup"] = out["fp64"]["total_s"] / out["mixed"]["total_s"]
    log("Maxwell M4: FP64 assembly %.2f s + GMRES %.2f s (%d it); mixed %.2f + %.2f s (%d it); err %.1e" % (
        out["fp64"]["assembly_s"], out["fp64"]["gmres_s"], out["fp64"]["iterations"], out["mixed"]["assembly_s"],
        out["mixed"]["gmres_s"], out["mixed"]["iterations"], out["mixed_rel_err_vs_fp64"]))
    return out, keep


def serving_stream(B, rng, nreq, rate, pool):
    """The request stream: Poisson arrivals at `rate` per second; each request solves a Laplace
    or Helmholtz single layer on one of `pool` scatterers drawn once from the stream's
    distribution (6-patch ellipsoid with semi-axes uniform in [0.5, 1.5], M in {3,4,5,6}, P in
    {1,2}, Helmholtz with kappa uniform in [1, 8]), with 1-16 fresh right-hand sides (point
    sources at radius 3 / plane waves with random directions): a service answering incident
    fields on scatterers it has seen, so operators recur and requests on one operator can share
    a batch."""
    scatterers = []
    for _ in range(pool):
        helm = bool(rng.uniform() < 0.5)
        scatterers.append(dict(helm=helm, scale=tuple(rng.uniform(0.5, 1.5, 3)), M=int(rng.choice([3, 4, 5, 6])),
                               P=int(rng.choice([1, 2])), kappa=float(rng.uniform(1, 8)) if helm else 0.0))
    reqs = []
    t = 0.0
    for k in range(nreq):
        t += rng.exponential(1 / rate)
        sc = dict(scatterers[int(rng.integers(pool))])
        nrhs = int(rng.integers(1, 17))
        if sc["helm"]:
            op = B.HelmholtzSingleLayerOperator()
            op.set_wavenumber(sc["kappa"])
            dirs = unit_directions(rng, nrhs)
            data = B.PlaneWaves(dirs, sc["kappa"])
        else:
            op = B.LaplaceSingleLayerOperator()
            dirs = 3.0 * unit_directions(rng, nrhs)
            data = B.DirichletTrace(dirs)
        sc.update(at=t, op=op, data=data, nrhs=nrhs, points=dirs)
        reqs.append(sc)
    return reqs


def run_serving(B, torch, rng, nreq, rate=8.0, pool=12):
    """Serving stream (serving_stream): mixed precision, relative residual 1e-6 (GMRES restart
    100, at most 6000 iterations), 5x5 Chebyshev nodes; up to four batches admitted at once, their
    GMRES solves in flight on separate streams together.  Returns the record and the requests."""
    import threading
    reqs = serving_stream(B, rng, nreq, rate, pool)
    srv = B.Server(max_active=4, max_batch_rhs=64, slice_iterations=100)
    ids = [None] * nreq
    t0 = time.perf_counter()

    def feed():
        for k, r in enumerate(reqs):
            dt = r["at"] - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            ids[k] = srv.submit(r["op"], r["M"], r["P"], r["data"], scale=r["scale"], n_cheb=5,
                                precision=B.MIXED, tol=1e-6, restart=100, maxit=6000)

    th = threading.Thread(target=feed)
    th.start()
    th.join()
    infos = []
    for k in range(nreq):
        _, info = srv.result(ids[k])
        infos.append(info)
        if (k + 1) % 32 == 0:
            log("serving: %d of %d results in" % (k + 1, nreq))
    wall = time.perf_counter() - t0
    srv.close()
    lat = [i["latency_s"] for i in infos]
    ok = sum(1 for i in infos if i["status"] == B.OK)
    first = min(i["submit_s"] for i in infos)
    last = max(i["end_s"] for i in infos)
    out = {"requests": nreq, "ok": ok, "arrival_rate_per_s": rate, "scatterer_pool": pool, "wall_s": wall,
           "p50_latency_s": pct(lat, 50), "p99_latency_s": pct(lat, 99), "max_latency_s": max(lat),
           "solves_per_s": nreq / (last - first), "rhs_solved": int(sum(r["nrhs"] for r in reqs)),
           "batched_requests": int(sum(1 for i in infos if i["batch_requests"] > 1)),
           "operator_cache_hits": int(sum(1 for i in infos if i["cache_hit"])),
           "precision": "mixed (FP32 reduced operator, FP64 refinement)", "tol": 1e-6,
           "scheduling": "up to 4 batches admitted; each GMRES batch a task on its own stream, all driven by one "
                         "device thread, so the batches overlap on the device",
           "mix": [{"M": r["M"], "P": r["P"], "helmholtz": r["helm"], "nrhs": r["nrhs"], "latency_s": i["latency_s"],
                    "iterations": i["iterations"], "batch_requests": i["batch_requests"], "status": i["status"]}
                   for r, i in zip(reqs, infos)]}
    log("serving: %d requests (%d ok) in %.1f s, p50 %.2f s, p99 %.2f s, %.2f solves/s, %d batched, %d cache hits" % (
        nreq, ok, wall, out["p50_latency_s"], out["p99_latency_s"], out["solves_per_s"], out["batched_requests"],
        out["operator_cache_hits"]))
    return out, reqs, infos


def restore_affinity():
    try:
        os.sched_setaffinity(0, AFFINITY)
    except OSError:
        pass


def cpu_info(limit=16):
    """The baseline's host CPU: lscpu's model and core counts, the physical cores in this
    process's CPU set, and the OpenMP threads used: one per physical core of the set, at most
    `limit` (one GPU's share of the host)"""
    info = {}
    try:
        import subprocess
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in txt.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except Exception:
        pass
    cores = set()
    for c in AFFINITY:
        try:
            with open("/sys/devices/system/cpu/cpu%d/topology/core_id" % c) as f:
                core = f.read().strip()
            with open("/sys/devices/system/cpu/cpu%d/topology/physical_package_id" % c) as f:
                pkg = f.read().strip()
            cores.add((pkg, core))
        except OSError:
            cores.add(("cpu", str(c)))

    def num(key):
        try:
            return int(info.get(key, ""))
        except ValueError:
            return None
    sockets, per_socket = num("Socket(s)"), num("Core(s) per socket")
    return {"model": info.get("Model name", "unknown"), "sockets": sockets, "cores_per_socket": per_socket,
            "threads_per_core": num("Thread(s) per core"),
            "physical_cores": sockets * per_socket if sockets and per_socket else None,
            "cpus_in_process_set": len(AFFINITY), "physical_cores_in_process_set": len(cores),
            "omp_threads": max(1, min(limit, len(cores))), "OMP_PROC_BIND": os.environ.get("OMP_PROC_BIND"),
            "OMP_PLACES": os.environ.get("OMP_PLACES")}


def cpu_assembly_estimate(oh, keys, rng, threads, ns=256, nblk=2000):
    """Oracle (OpenMP) assembly time of an H2 operator estimated from a random sample: the 16
    element pairs of each of ns dense leaves and nblk coupling blocks, timed and scaled to all
    leaves"""
    dense = keys[keys[:, 0] == 2]
    lr_idx = np.nonzero(keys[:, 0] == 1)[0]
    ns, nblk = min(ns, len(dense)), min(nblk, len(lr_idx))
    pick = rng.choice(len(dense), ns, replace=False)
    pairs = np.array([(4 * c1 + i, 4 * c2 + j) for (_, _, c1, c2) in dense[pick] for i in range(4)
                      for j in range(4)], dtype=np.int32)
    okey = {tuple(r): i for i, r in enumerate(oh.leaves().tolist())}
    blocks = np.array([okey[tuple(keys[k])] for k in rng.choice(lr_idx, nblk, replace=False)], dtype=np.int32)
    t_pairs = oh.time_assembly_sample(pairs, [], threads)
    t_blocks = oh.time_assembly_sample(np.zeros((0, 2)), blocks, threads) if nblk else 0.0
    est = t_pairs * len(dense) / ns + (t_blocks * len(lr_idx) / nblk if nblk else 0.0)
    return est, "%d element pairs of %d sampled dense leaves (%.2f s) + %d coupling blocks (%.2f s), scaled to " \
                "%d dense + %d coupling leaves" % (len(pairs), ns, t_pairs, nblk, t_blocks, len(dense), len(lr_idx))


def cpu_calibration(O, rng, threads, level=3):
    """The sampled estimate against a timed full assembly: the oracle assembles the whole Laplace
    operator (P 2, 4x4 nodes) at `level` with the same threads, and the estimate is formed for the
    same operator"""
    osp = O.Space(O.Geometry(), level, 2)
    O.lib().or_set_num_threads(threads)
    t0 = time.perf_counter()
    full = O.H2(osp, O.LAPLACE, 0j, 4, assemble=True)
    t_full = time.perf_counter() - t0
    del full
    oh = O.H2(osp, O.LAPLACE, 0j, 4, assemble=False)
    est, sample = cpu_assembly_estimate(oh, oh.leaves(), rng, threads)
    restore_affinity()
    log("CPU calibration at M=%d: full assembly %.2f s, sampled estimate %.2f s" % (level, t_full, est))
    return {"operator": "laplace_single_layer_h2_sphere_M%d_P2_np4" % level, "elements": osp.nel, "full_assembly_s": t_full, "estimate_s": est, "full_over_estimate": t_full / est,
        "sample": sample}


def cpu_laplace(O, H, M, P, npts, n, rng, threads, calib):
    """The headline's CPU baseline: the oracle's OpenMP H2 product of the same operator (the GPU's
    leaves imported), and its assembly time from a sample, corrected by the calibration"""
    osp = O.Space(O.Geometry(), M, P)
    oh = O.H2(osp, O.LAPLACE, 0j, npts, assemble=False)
    keys, data = H.export_leaves()
    matched = oh.import_leaves(keys, data)
    del data
    assert matched == len(keys), (matched, len(keys))
    t_cpu = oh.time_matvec(1, 3, threads)
    est, sample = cpu_assembly_estimate(oh, keys, rng, threads)
    restore_affinity()
    cal = calib["full_over_estimate"] if calib else 1.0
    log("CPU baseline: matvec %.3f s (%d threads), assembly estimate %.1f s (x %.2f calibration)" % (
        t_cpu, threads, est, cal))
    return {"value": n / t_cpu, "unit": "DOF/s", "cores": threads, "kind": "port",
            "sample": "oracle OpenMP H2 matvec of the same 24,576-element operator (same leaves), 3 matvecs, "
                      "%.3f s each" % t_cpu,
            "assembly_s_estimate": est, "assembly_s_calibrated": est * cal, "assembly_sample": sample}


def cpu_headline(O, keep, out, rng, threads, calib):
    """CPU baseline of the Helmholtz 24,576-element assembly + GMRES: the oracle's matvec of the same
    leaves times the GMRES products, plus the calibrated sampled assembly"""
    M, P, npts, kappa = 6, 2, 5, 4.0
    osp = O.Space(O.Geometry(), M, P)
    oh = O.H2(osp, O.HELMHOLTZ, complex(kappa), npts, assemble=False)
    keys, data = keep.export_leaves()
    assert oh.import_leaves(keys, data) == len(keys)
    del data
    t_mv = oh.time_matvec(1, 2, threads)
    est, sample = cpu_assembly_estimate(oh, keys, rng, threads, ns=128)
    restore_affinity()
    cal = calib["full_over_estimate"] if calib else 1.0
    it = out["fp64"]["iterations"]
    total = est * cal + t_mv * (it + it // 30 + 1)
    log("Helmholtz M6 CPU estimate: assembly %.0f s, matvec %.3f s, total %.0f s (%d threads)" % (
        est * cal, t_mv, total, threads))
    return {"cores": threads, "kind": "port", "assembly_s": est * cal, "assembly_s_uncalibrated": est,
            "matvec_s": t_mv, "total_s": total, "gpu_fp64_speedup": total / out["fp64"]["total_s"],
            "gpu_mixed_speedup": total / out["mixed"]["total_s"],
            "sample": sample + "; oracle OpenMP matvec of the same leaves x (GMRES iterations + restarts)"}


def cpu_batched(O, H, out, rng, threads, calib):
    """CPU baseline of the 64-RHS Helmholtz workload (6,144 elements): the oracle's product of the
    same leaves on all 64 columns at once and as a loop over the 64 columns, the GMRES products,
    and the calibrated sampled assembly"""
    osp = O.Space(O.Geometry(), 5, 2)
    oh = O.H2(osp, O.HELMHOLTZ, 4.0 + 0j, 5, assemble=False)
    keys, data = H.export_leaves()
    assert oh.import_leaves(keys, data) == len(keys)
    del data
    t64 = oh.time_matvec(64, 2, threads)
    t1 = oh.time_matvec(1, 5, threads)
    est, sample = cpu_assembly_estimate(oh, keys, rng, threads, ns=128)
    restore_affinity()
    cal = calib["full_over_estimate"] if calib else 1.0
    it = out["fp64_gmres_iterations_max"]
    n = out["dofs"]
    res = {"cores": threads, "kind": "port", "matvec_64rhs_block_s": t64, "matvec_64rhs_column_loop_s": 64 * t1,
           "dof_rhs_per_s": n * 64 / min(t64, 64 * t1), "assembly_s": est * cal,
           "gmres_s_estimate": min(t64, 64 * t1) * (it + it // 30 + 1),
           "gpu_fp64_matvec_speedup": min(t64, 64 * t1) / (out["fp64_matvec_ms"] * 1e-3),
           "sample": sample + "; oracle OpenMP matvec of the same leaves, 64 columns per product and 64 single-"
                              "column products"}
    log("64-RHS CPU: block product %.3f s, column loop %.3f s, assembly %.1f s" % (t64, 64 * t1, est * cal))
    return res


def cpu_maxwell(O, keep, out, rng, threads, calib, frac=0.1):
    """CPU baseline of the Maxwell EFIE workload.  The oracle stores the reference's four
    (2 np^2)^2 component blocks per coupling leaf (36 GB at this size), so its product is timed on
    the dense leaves alone and on the dense leaves plus a random tenth of the coupling leaves, and
    the coupling part is scaled to all leaves; times the GMRES products, plus the calibrated
    sampled assembly"""
    osp = O.Space(O.Geometry(scale=(1.0, 0.7, 0.5)), 4, 2, form=1)
    keys = keep.leaves()
    dense = np.nonzero(keys[:, 0] == 2)[0]
    low = np.nonzero(keys[:, 0] == 1)[0]
    pick = rng.choice(low, max(1, int(frac * len(low))), replace=False)
    oh = O.H2(osp, O.MAXWELL, 2.0 + 0j, 9, assemble=False)
    k1, d1 = keep.export_leaf_data(dense)
    assert oh.import_leaves(k1, d1) == len(k1)
    del d1
    t_dense = oh.time_matvec(1, 3, threads)
    k2, d2 = keep.export_leaf_data(pick)
    assert oh.import_leaves(k2, d2) == len(k2)
    del d2
    t_part = oh.time_matvec(1, 3, threads)
    t_mv = t_dense + (t_part - t_dense) * len(low) / len(pick)
    est, sample = cpu_assembly_estimate(oh, keys, rng, threads, ns=128, nblk=1000)
    restore_affinity()
    cal = calib["full_over_estimate"] if calib else 1.0
    it = out["fp64"]["iterations"]
    total = est * cal + t_mv * (it + it // 100 + 1)
    log("Maxwell M4 CPU estimate: assembly %.1f s, matvec %.3f s (dense part %.3f s), total %.0f s" % (
        est * cal, t_mv, t_dense, total))
    return {"cores": threads, "kind": "port", "assembly_s": est * cal, "matvec_s": t_mv, "matvec_dense_part_s": t_dense,
            "total_s": total, "gpu_fp64_speedup": total / out["fp64"]["total_s"],
            "gpu_mixed_speedup": total / out["mixed"]["total_s"],
            "sample": sample + "; oracle OpenMP matvec on the dense leaves and on %d of %d coupling leaves, coupling "
                               "part scaled, x (GMRES iterations + restarts)" % (len(pick), len(low))}


def cpu_serving(O, reqs, infos, rng, threads, budget_s=20.0):
    """Serial-stream CPU baseline of the serving workload, on a sample: requests on the smaller
    scatterers (M <= 4) in random order, each solved alone by the oracle (OpenMP H2 assembly, then
    GMRES(100) to 1e-6 per right-hand side with the oracle's product), until `budget_s` of CPU time
    is spent"""
    from scipy.sparse.linalg import LinearOperator, gmres
    O.lib().or_set_num_threads(threads)
    cand = [k for k, r in enumerate(reqs) if r["M"] <= 4]
    rng.shuffle(cand)
    done, spent = [], 0.0
    for k in cand:
        if spent > budget_s:
            break
        r = reqs[k]
        t0 = time.perf_counter()
        osp = O.Space(O.Geometry(scale=r["scale"]), r["M"], r["P"])
        oh = O.H2(osp, O.HELMHOLTZ if r["helm"] else O.LAPLACE, complex(r["kappa"]), 5, assemble=True)
        n = osp.ndof
        dt = np.complex128 if r["helm"] else np.float64
        A = LinearOperator((n, n), matvec=lambda v: oh.matvec(v).reshape(-1), dtype=dt)
        for c in range(r["nrhs"]):
            b = osp.rhs(1, r["points"][c], r["kappa"]) if r["helm"] else osp.rhs(0, r["points"][c])
            x, info = gmres(A, b, rtol=1e-6, restart=100, maxiter=60)
        t = time.perf_counter() - t0
        spent += t
        done.append({"M": r["M"], "P": r["P"], "helmholtz": r["helm"], "nrhs": r["nrhs"], "cpu_s": t,
                     "gpu_latency_s": infos[k]["latency_s"]})
        del oh
    restore_affinity()
    if not done:
        return None
    tot = sum(d["cpu_s"] for d in done)
    log("serving CPU sample: %d requests, %.1f s serial (%.2f solves/s)" % (len(done), tot, len(done) / tot))
    return {"cores": threads, "kind": "port", "requests": len(done), "serial_s": tot, "solves_per_s": len(done) / tot,
            "p50_request_s": pct([d["cpu_s"] for d in done], 50), "sampled": done,
            "sample": "requests with M <= 4 of the same stream, each solved alone by the oracle (assembly + "
                      "GMRES(100) to 1e-6 per right-hand side, FP64), until %.0f s were spent" % budget_s}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--solves", type=int, default=5)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--level", type=int, default=6)
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="run one plain device copy of this many MB first (counter calibration)")
    ap.add_argument("--no-breakdown", action="store_true", help="skip the per-kernel assembly breakdown")
    ap.add_argument("--no-batched", action="store_true", help="skip the 64-RHS Helmholtz workload")
    ap.add_argument("--no-headline", action="store_true", help="skip the Helmholtz 24,576-element assembly + GMRES")
    ap.add_argument("--serving", type=int, default=128, help="requests in the serving stream (0: skip)")
    ap.add_argument("--no-maxwell", action="store_true", help="skip the Maxwell EFIE 1,536-element workload")
    ap.add_argument("--batched-profile", action="store_true",
                    help="only the Helmholtz 6,144-element 64-RHS products and GMRES steps (profiling workload)")
    ap.add_argument("--helmholtz-profile", action="store_true",
                    help="only the Helmholtz 24,576-element products and GMRES steps (profiling workload)")
    ap.add_argument("--assembly-profile", action="store_true",
                    help="only one H2 assembly each of the Laplace (np 4) and Helmholtz (kappa 4, np 5) operators "
                         "on 24,576 elements, P 2 (profiling workload of the near-field quadrature)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: the H2 path runs on a single MI355X "
              "(sharding over several GPUs is designed, not implemented)", file=sys.stderr)
        sys.exit(2)

    import torch
    restore_affinity()
    import bembel_amd as B
    torch.cuda.set_device(0)
    torch.cuda.init()
    # no cyclic garbage collection inside timed regions (as timeit does)
    gc.collect()
    gc.disable()
    if a.helmholtz_profile:
        run_helmholtz_profile(B, torch)
        return
    if a.batched_profile:
        run_batched_profile(B, torch)
        return
    if a.assembly_profile:
        run_assembly_profile(B, torch)
        return
    M, P, npts = a.level, 2, 4
    geo = B.Geometry()
    op = B.LaplaceSingleLayerOperator()
    t0 = time.perf_counter()
    space = B.AnsatzSpace(geo, M, P, op)
    t_space = time.perf_counter() - t0
    n = space.number_of_dofs
    log("space: %d elements, %d dofs (%.2f s)" % (space.number_of_elements, n, t_space))

    if a.copy_mb:
        src = torch.ones(a.copy_mb * 1024 * 1024 // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    # warm-up: one assembly and product on a 96-element sphere loads the library's code objects
    # and sizes the runtime's pools (a one-time cost of the process, about 0.1 s), untimed
    wsp = B.AnsatzSpace(B.Geometry(), 2, P, op)
    wh = B.H2Matrix(wsp, op, n_cheb=npts)
    wh.matvec(np.ones(wsp.number_of_dofs))
    del wh, wsp
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    H = B.H2Matrix(space, op, n_cheb=npts)
    torch.cuda.synchronize()
    t_asm = time.perf_counter() - t0
    info = H.info
    log("H2 assembly %.3f s: %d dense leaves, %d low-rank leaves, %d near-field pairs, classes %s, %.2f GB" % (
        t_asm, info["dense"], info["lowrank"], info["pairs"], info["pair_classes"], info["device_bytes"] / 1e9))

    # per-kernel breakdown of the assembly (second assembly, events around every launch)
    duffy_ms = reg_ms = cpl_ms = 0.0
    duffy_n = reg_n = cpl_n = 0
    cls_ms, host_ms = {}, {}
    if not a.no_breakdown:
        B.timing.reset()
        B.timing.enable(True)
        H2b = B.H2Matrix(space, op, n_cheb=npts)
        B.timing.enable(False)
        cls_ms = {k: B.timing.get("nf_" + k)[0] for k in ("identical", "edge", "vertex")}
        duffy_ms = sum(cls_ms.values())
        duffy_n = sum(B.timing.get("nf_" + k)[1] for k in ("identical", "edge", "vertex"))
        reg_ms, reg_n = B.timing.get("nf_regular")
        cpl_ms, cpl_n = B.timing.get("coupling_assemble")
        host_ms = {k: B.timing.get("assembly_host_" + k)[0] for k in ("block_tree", "pairs", "pair_sort")}
        del H2b
    log("assembly kernels: nf_duffy %.1f ms (%d launches: %s), nf_regular %.1f ms (%d), coupling %.1f ms; host: %s" % (
        duffy_ms, duffy_n, ", ".join("%s %.1f" % kv for kv in cls_ms.items()), reg_ms, reg_n, cpl_ms,
        ", ".join("%s %.1f ms" % kv for kv in host_ms.items())))

    # ---------------------------------------------------------------- matvec (the step)
    g = torch.Generator(device="cuda")
    g.manual_seed(1234)
    x = torch.rand(n, dtype=torch.float64, device="cuda", generator=g) - 0.5
    # steps are enqueued back to back on the library's stream; the clock stops when the last
    # product has finished on the device
    torch.cuda.synchronize()  # x is written by torch (its own HIP runtime)
    for _ in range(a.warmup):
        y = H.matvec(x, sync=False)
    B.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        y = H.matvec(x, sync=False)
    B.synchronize()
    t_mv = (time.perf_counter() - t0) / a.steps
    value = n / t_mv
    log("matvec %.3f ms -> %.3e DOF/s" % (t_mv * 1e3, value))

    # mixed precision product (FP32 coupling blocks, FP64 near field): speed and error vs FP64
    Hm = B.H2Matrix(space, op, n_cheb=npts, precision=B.MIXED)
    for _ in range(a.warmup):
        ym = Hm.matvec(x, sync=False)
    B.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        ym = Hm.matvec(x, sync=False)
    B.synchronize()
    t_mv_mixed = (time.perf_counter() - t0) / a.steps
    mixed_err = float(torch.linalg.norm(ym - y) / torch.linalg.norm(y))
    log("mixed matvec %.3f ms -> %.3e DOF/s, rel err vs FP64 %.2e" % (t_mv_mixed * 1e3, n / t_mv_mixed, mixed_err))
    del Hm

    # kernel timing pass for the roofline (HIP events on the launch stream, inside the library)
    B.timing.reset()
    B.timing.enable(True)
    keep = [H.matvec(x, sync=False) for _ in range(max(3, a.steps // 4))]  # queued back to back
    B.synchronize()
    B.timing.enable(False)
    del keep
    # the upper-block sweep has its own scope (its transposed partials are added by k_sym_gather)
    su_ms, su_n = B.timing.get("matvec_sweep_upper")
    sym_sweep = su_n > 0
    sw_ms, sw_n = (su_ms, su_n) if sym_sweep else B.timing.get("matvec_sweep")
    sg_ms, sg_n = B.timing.get("matvec_sym_gather")
    fw_ms, fw_n = B.timing.get("matvec_forward")
    bw_ms, bw_n = B.timing.get("matvec_backward")
    sw_avg = sw_ms / max(sw_n, 1)
    dd = info["dense_dim"]
    nsuper = space.number_of_super_dofs
    rows = info["coupling_dim"]
    L = M - 1
    lvl_total = 6 * (4 ** (L + 1) - 1) // 3 * rows
    # algorithmic bytes of the fused sweep: every near-field leaf once, every coupling block once
    # (or, when the autotuner chose the symmetric sweep, the upper half of the coupling blocks plus
    # one transposed partial vector written per upper block), x read and y written once, the
    # multipole vectors read and the local expansions written once
    ncpl = info["lowrank"] // 2 if sym_sweep else info["lowrank"]
    sw_bytes = ((info["dense"] * dd * dd + ncpl * rows * rows) * 8 + 2 * nsuper * 8 + 2 * lvl_total * 8 +
                (ncpl * rows * 8 if sym_sweep else 0))
    achieved = sw_bytes / (sw_avg * 1e-3) / 1e9 if sw_avg > 0 else 0.0
    read_gbs = streaming_read_gbs(torch)
    log("streaming read (torch.sum over 2 GiB) %.0f GB/s; sweep at %.0f %% of it" % (read_gbs,
                                                                                      100 * achieved / read_gbs))
    log("sweep (near field + coupling%s) %.3f ms/launch, %.1f MB -> %.0f GB/s; forward %.3f ms, backward %.3f ms" % (
        ", upper coupling blocks" if sym_sweep else "", sw_avg, sw_bytes / 1e6, achieved, fw_ms / max(fw_n, 1),
        bw_ms / max(bw_n, 1)))

    # near-field quadrature roofline (FP64)
    pairs_by_class = info["pair_classes"]
    deg_touch = int(0.5 * ((4 + 2 * P) * M * math.log(2.0)) / (2 * math.log(2.0)))
    ntouch = deg_touch + 1
    fl = nearfield_flops([(2, ntouch, pairs_by_class[2]), (3, ntouch, pairs_by_class[3]),
                          (4, ntouch, pairs_by_class[4])], P, 5)
    duffy_tflops = fl / (duffy_ms * 1e-3) / 1e12 if duffy_ms > 0 else 0.0
    cls_tflops = {}
    for c, name in ((2, "identical"), (3, "edge"), (4, "vertex")):
        t = cls_ms.get(name, 0.0)
        if t > 0:
            cls_tflops[name] = nearfield_flops([(c, ntouch, pairs_by_class[c])], P, 5) / (t * 1e-3) / 1e12
    log("Duffy FP64: %.1f TFLOP/s overall (%s)" % (
        duffy_tflops, ", ".join("%s %.1f" % kv for kv in cls_tflops.items())))

    # ---------------------------------------------------------------- solves
    rng = np.random.default_rng(2024)
    d = rng.normal(size=3)
    x0 = 0.3 * rng.uniform() ** (1 / 3) * d / np.linalg.norm(d)
    b = B.DiscreteLinearForm(B.DirichletTrace(x0), space).compute(
        device=True).get_discrete_linear_form()
    lat64, latm, it64, rel_err_mixed = [float("nan")], [float("nan")], 0, float("nan")
    if a.solves > 0:
        lat64, latm, it64, rel_err_mixed = run_solves(B, torch, H, space, op, npts, b, a.solves)

    batched, Hb = (None, None) if a.no_batched else run_batched(B, torch, rng, a.steps)
    headline, Hh = (None, None) if a.no_headline else run_headline(B, torch, rng)
    maxwell, Hx = (None, None) if a.no_maxwell else run_maxwell(B, torch, rng)
    serving, reqs, infos = (run_serving(B, torch, np.random.default_rng(55), a.serving) if a.serving > 0
                            else (None, None, None))

    # ---------------------------------------------------------------- CPU baselines (after the GPU legs)
    cpu = None
    host = cpu_info()
    threads = a.cpu_threads or host["omp_threads"]
    host["omp_threads"] = threads
    if not a.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        O.lib()
        restore_affinity()
        crng = np.random.default_rng(77)
        calib = None
        try:
            calib = cpu_calibration(O, crng, threads)
        except Exception as e:  # the CPU legs must not hide the GPU results
            log("CPU calibration failed:", repr(e))
        try:
            cpu = cpu_laplace(O, H, M, P, npts, n, crng, threads, calib)
            cpu["host"] = host
            cpu["calibration"] = calib
        except Exception as e:
            log("CPU baseline failed:", repr(e))
        for name, fn, rec, handle in (("batched", cpu_batched, batched, Hb), ("headline", cpu_headline, headline, Hh),
                                      ("maxwell", cpu_maxwell, maxwell, Hx)):
            if rec is None:
                continue
            try:
                rec["cpu_baseline"] = fn(O, handle, rec, crng, threads, calib)
            except Exception as e:
                log("CPU baseline (%s) failed:" % name, repr(e))
        if serving is not None:
            try:
                serving["cpu_baseline"] = cpu_serving(O, reqs, infos, crng, threads)
            except Exception as e:
                log("CPU baseline (serving) failed:", repr(e))
        restore_affinity()
    del Hb, Hh, Hx

    out = {
        "metric": "H2 assembly s + matvec DOF/s at 1e-6 rel err vs FP64; p50/p99 solve latency",
        "value": value,
        "unit": "DOF/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": t_mv * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded point-source Dirichlet data, random input vectors), unit-sphere geometry file",
        "config": {"workload": "laplace_single_layer_h2_sphere_M%d_P%d_np%d" % (M, P, npts),
                   "elements": space.number_of_elements, "dofs": n, "polynomial_degree": P, "n_cheb": npts,
                   "eta": 1.6, "min_cluster_level": 1, "nrhs": 1},
        "mixed_matvec": {"ms_per_step": t_mv_mixed * 1e3, "dof_per_s": n / t_mv_mixed, "rel_err_vs_fp64": mixed_err,
                         "storage": "near-field blocks FP32, coupling blocks FP32 (recomputed from the Chebyshev "
                                    "surface points or read, whichever measured faster), sums in FP64"},
        "h2_assembly_s": t_asm,
        "space_setup_s": t_space,
        "assembly_kernels_ms": {"nf_duffy": duffy_ms, "nf_duffy_by_class": cls_ms, "nf_regular": reg_ms,
                                "coupling_assemble": cpl_ms, "host_ms": host_ms},
        "h2": {k: info[k] for k in ("dense", "lowrank", "pairs", "pair_classes", "device_bytes")},
        "solve": {"solver": "CG", "tol": 1e-10, "fp64_iterations": it64, "fp64_p50_s": pct(lat64, 50),
                  "fp64_p99_s": pct(lat64, 99), "mixed_p50_s": pct(latm, 50), "mixed_p99_s": pct(latm, 99),
                  "mixed_rel_err_vs_fp64": rel_err_mixed, "solves": a.solves},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(M, P, npts, sym_sweep),
                     "traffic_source": "recorded, not measured in this run: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE "
                                       "of this kernel on this workload, scaled by a plain 1 GiB copy "
                                       "(profiles/k_sweep_fused_laplace_M6_P2_np4_traffic.json)",
                     "kernel": "k_sweep_fused (near-field leaves + coupling blocks of the matvec)",
                     "bytes_per_launch": sw_bytes, "avg_launch_ms": sw_avg, "symmetric_coupling": sym_sweep,
                     "sym_gather_ms": sg_ms / max(sg_n, 1) if sym_sweep else 0.0,
                     "streaming_read_gbs": read_gbs,
                     "forward_ms": fw_ms / max(fw_n, 1), "backward_ms": bw_ms / max(bw_n, 1)},
        "roofline_assembly": {"bound": "fp64", "achieved": duffy_tflops, "peak": FP64_PEAK_TFLOPS,
                              "unit": "TFLOP/s", "frac": duffy_tflops / FP64_PEAK_TFLOPS,
                              "kernel": "k_nearfield identical/edge/vertex (Duffy)", "flops": fl,
                              "by_class_tflops": cls_tflops},
        "cpu_baseline": cpu,
        "batched_64rhs": batched,
        "helmholtz_assembly_gmres": headline,
        "maxwell_efie": maxwell,
        "serving": serving,
        "build": B.build_info(),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
