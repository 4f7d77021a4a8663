"""GPU H2 assembly and matrix-vector product versus the oracle.

Small sizes: identical leaf sets, every leaf matrix and the matvec equal the oracle to 1e-12.
Full size (24,576 elements): sampled blocks equal the oracle; size-independent properties
(symmetry, linearity, the constant-function closed form, multi-RHS consistency) hold.
"""
import math

import numpy as np
import pytest

import bembel_amd as B
import oracle_lib as O

pytestmark = pytest.mark.gpu
TOL = 1e-12


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def make(kind, kappa, M, P, np_, precision=B.FP64):
    os_ = O.Space(O.Geometry(), M, P)
    ps = B.AnsatzSpace(B.Geometry(), M, P)
    op = B.LaplaceSingleLayerOperator() if kind == O.LAPLACE else B.HelmholtzSingleLayerOperator()
    op.set_wavenumber(kappa)
    H = B.H2Matrix(ps, op, n_cheb=np_, precision=precision)
    return os_, ps, op, H


CASES = [(O.LAPLACE, 0j, 3, 2, 4), (O.HELMHOLTZ, 4 + 0j, 3, 2, 5), (O.LAPLACE, 0j, 4, 1, 3),
         (O.HELMHOLTZ, 0j, 2, 1, 9), (O.HELMHOLTZ, 2 + 0.5j, 3, 2, 4)]  # last: damped (complex) wavenumber


@pytest.mark.parametrize("kind,kappa,M,P,np_", CASES)
def test_gpu_h2_leaves_match_oracle(gpu, kind, kappa, M, P, np_):
    os_, ps, op, H = make(kind, kappa, M, P, np_)
    oh = O.H2(os_, kind, kappa, np_)
    ol = oh.leaves()
    gl = H.leaves()
    okeys = {tuple(r): i for i, r in enumerate(ol.tolist())}
    gkeys = {tuple(r): i for i, r in enumerate(gl.tolist())}
    assert set(okeys) == set(gkeys)
    worst = 0.0
    for key, gi in gkeys.items():
        F = H.leaf_matrix(gi)
        R = oh.leaf_matrix(okeys[key])
        worst = max(worst, rel(F, R))
    print("worst leaf rel", worst)
    assert worst <= TOL


@pytest.mark.parametrize("kind,kappa,M,P,np_", CASES)
def test_gpu_h2_matvec_matches_oracle(gpu, kind, kappa, M, P, np_):
    os_, ps, op, H = make(kind, kappa, M, P, np_)
    oh = O.H2(os_, kind, kappa, np_)
    rng = np.random.default_rng(7)
    n = ps.number_of_dofs
    # 67 columns: one full and one partial tile of 64 right-hand sides
    x = rng.uniform(-1, 1, (n, 67))
    if kind != O.LAPLACE:
        x = x + 1j * rng.uniform(-1, 1, (n, 67))
    y = H @ x
    yo = oh.matvec(x)
    r = rel(y, yo)
    print("matvec rel", r)
    assert r <= TOL
    # columns are independent
    assert rel(H @ x[:, 1], yo[:, 1]) <= TOL


@pytest.mark.parametrize("kind,kappa,M,P,np_", [CASES[0], CASES[1]])
def test_gpu_h2_upper_block_sweep(gpu, kind, kappa, M, P, np_):
    """Single-RHS products that read only the upper coupling blocks, and also only the upper
    near-field leaves (the transposed products formed from them, partials gathered per target),
    equal the oracle like the sweep over every block, for the FP64 and the mixed operator; each
    form is forced in turn, with both sweeps in one launch and with the near-field leaves running
    beside the far chain on a side stream."""
    os_, ps, op, H = make(kind, kappa, M, P, np_)
    oh = O.H2(os_, kind, kappa, np_)
    rng = np.random.default_rng(11)
    n = ps.number_of_dofs
    x = rng.uniform(-1, 1, (n, 1))
    if kind != O.LAPLACE:
        x = x + 1j * rng.uniform(-1, 1, (n, 1))
    yo = oh.matvec(x)
    ys = {}
    for mode in ("full", "upper", "upper_near"):
        H.set_sweep(mode)
        for ov in ("fused", "overlap"):
            # near-field leaves in the same launch as the coupling blocks, or beside the far chain
            H.set_overlap(ov)
            ys[mode, ov] = H @ x
            print(mode, ov, "rel", rel(ys[mode, ov], yo))
            assert rel(ys[mode, ov], yo) <= TOL
        # the two forms run the same block code on disjoint output vectors
        print(mode, "overlap == fused bit for bit:", np.array_equal(ys[mode, "overlap"], ys[mode, "fused"]))
        assert rel(ys[mode, "overlap"], ys[mode, "fused"]) <= 1e-15
        H.set_overlap(None)
    assert rel(ys["upper", "fused"], ys["full", "fused"]) <= 1e-14
    assert rel(ys["upper_near", "fused"], ys["full", "fused"]) <= 1e-14
    H.set_sweep(None)
    assert rel(H @ x, yo) <= TOL
    Hm = B.H2Matrix(ps, op, n_cheb=np_, precision=B.MIXED)
    for mode in ("full", "upper"):
        Hm.set_sweep(mode)
        e = rel(Hm @ x, yo)
        print("mixed", mode, "rel", e)
        assert e <= 1e-6


def test_gpu_h2_vs_dense_lazy_eigen_sum(gpu):
    """tests/test_LazyEigenSum.cpp: |D a - H2 a| < 1e-8 (Helmholtz kappa 0, P 1, M 2, np 9)"""
    os_, ps, op, H = make(O.HELMHOLTZ, 0j, 2, 1, 9)
    D = os_.dense_galerkin(O.HELMHOLTZ, 0j)
    a = np.random.default_rng(3).uniform(-1, 1, ps.number_of_dofs).astype(complex)
    assert np.linalg.norm(D @ a - H @ a) < 1e-8
    assert np.linalg.norm(3 * (D @ a) + D @ a - (3 * (H @ a) + H @ a)) < 1e-8


def test_gpu_h2_device_tensors(gpu):
    import torch
    os_, ps, op, H = make(O.HELMHOLTZ, 4 + 0j, 3, 2, 5)
    n = ps.number_of_dofs
    x = torch.randn(n, 4, dtype=torch.complex128, device="cuda")
    y = H @ x
    assert y.is_cuda and y.shape == (n, 4)
    yh = H @ x.cpu().numpy()
    assert rel(y.cpu().numpy(), yh) <= 1e-14


@pytest.mark.slow
def test_gpu_h2_full_size_laplace(gpu):
    """Laplace single layer on 24,576 elements, P = 2, 4x4 Chebyshev nodes (the C2 workload)."""
    M, P, np_ = 6, 2, 4
    ps = B.AnsatzSpace(B.Geometry(), M, P)
    H = B.H2Matrix(ps, B.LaplaceSingleLayerOperator(), n_cheb=np_)
    n = ps.number_of_dofs
    assert n == 26136
    rng = np.random.default_rng(11)
    x, z = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    Ax, Az = H @ x, H @ z
    # symmetric operator; the H2 compression keeps symmetry to its own accuracy
    assert abs(z @ Ax - x @ Az) <= 1e-6 * abs(z @ Ax)
    # linearity and batched columns
    Y = H @ np.stack([x, z, x + 2 * z], axis=1)
    assert rel(Y[:, 0], Ax) <= 1e-14 and rel(Y[:, 2], Ax + 2 * Az) <= 1e-13
    # constant function: c^T A c = 4 pi with c = h * 1
    c = np.ones(n) * 2.0 ** -M
    assert abs(c @ (H @ c) - 4 * math.pi) / (4 * math.pi) < 1e-4
    # sampled near-field pairs against the oracle
    os_ = O.Space(O.Geometry(), M, P)
    el = rng.choice(ps.number_of_elements, 6, replace=False)
    pairs = []
    for e in el:
        for f in range(ps.number_of_elements):
            info, _ = os_.pair_info(int(e), f)
            if info[0] in (2, 3, 4) or (info[0] == 0 and info[3] > 4 and rng.uniform() < 0.05):
                pairs.append((int(e), f))
    pairs = np.array(pairs, dtype=np.int32)
    got = ps.pair_blocks(B.LaplaceSingleLayerOperator(), pairs)
    for k, (i, j) in enumerate(pairs):
        assert rel(got[k], os_.pair_block(O.LAPLACE, int(i), int(j))) <= TOL


def test_gpu_h2_graph_replays_match_eager(gpu):
    """The captured-graph product equals the kernel-by-kernel product on every replay, with the
    program's own device work and synchronisations between replays, and after the workspace
    grows (which re-captures)."""
    import torch
    os_, ps, op, H = make(O.HELMHOLTZ, 4 + 0j, 3, 2, 5)
    n = ps.number_of_dofs
    g = torch.Generator(device="cuda").manual_seed(1)
    # 12 columns take the right-hand-side-fastest sweeps
    xs = [torch.randn(n, k, dtype=torch.complex128, device="cuda", generator=g) for k in (1, 1, 3, 12, 1, 3, 12)]
    H.use_graphs(False)
    eager = [(H @ x).clone() for x in xs]
    torch.cuda.synchronize()
    H.use_graphs(True)
    for rep in range(3):
        for x, ye in zip(xs, eager):
            y = H @ x
            junk = torch.randn(1 << 16, device="cuda") * 2.0  # unrelated device work in between
            torch.cuda.synchronize()
            assert torch.equal(y, ye), rep
            del junk


@pytest.mark.parametrize("kind,kappa,M,P,np_", [CASES[0], CASES[1]])
def test_gpu_h2_super_space_product(gpu, kind, kappa, M, P, np_):
    """A_super xs without T (bmi_h2_matvec_super) equals the oracle's super-space product."""
    import torch
    os_, ps, op, H = make(kind, kappa, M, P, np_)
    oh = O.H2(os_, kind, kappa, np_)
    N = ps.number_of_super_dofs
    rng = np.random.default_rng(4)
    xs = rng.uniform(-1, 1, (N, 2))
    if kind != O.LAPLACE:
        xs = xs + 1j * rng.uniform(-1, 1, (N, 2))
    dt = torch.complex128 if kind != O.LAPLACE else torch.float64
    ys = H.matvec_super(torch.tensor(xs, dtype=dt, device="cuda")).cpu().numpy()
    assert rel(ys, oh.matvec_super(xs)) <= TOL


def test_gpu_mixed_precision_matvec_and_refinement(gpu):
    """Mixed handles: FP32 coupling blocks keep the product within 1e-6 of FP64; BF16 coarse
    blocks are coarser but bounded; iterative refinement with the FP64 residual brings CG and GMRES
    solutions to the FP64 solve."""
    M, P, np_ = 4, 2, 4
    ps = B.AnsatzSpace(B.Geometry(), M, P)
    op = B.LaplaceSingleLayerOperator()
    H64 = B.H2Matrix(ps, op, n_cheb=np_)
    H32 = B.H2Matrix(ps, op, n_cheb=np_, precision=B.MIXED)
    H16 = B.H2Matrix(ps, op, n_cheb=np_, precision=B.MIXED_BF16)
    x = np.random.default_rng(2).uniform(-1, 1, ps.number_of_dofs)
    y64 = H64 @ x
    e32, e16 = rel(H32 @ x, y64), rel(H16 @ x, y64)
    print("mixed matvec rel err fp32 %.2e bf16 %.2e" % (e32, e16))
    assert e32 <= 1e-6 and e16 <= 1e-2 and e16 >= e32
    # 16 columns: the FP32 blocks on the FP32 matrix cores (per-block FP32 sums, FP64 across blocks)
    X = np.random.default_rng(3).uniform(-1, 1, (ps.number_of_dofs, 16))
    e32m = rel(H32 @ X, H64 @ X)
    print("mixed matvec rel err, 16 columns: %.2e" % e32m)
    assert e32m <= 1e-6
    b = B.DiscreteLinearForm(B.DirichletTrace([0.1, -0.2, 0.05]), ps).compute().get_discrete_linear_form()
    sols = {}
    for name, H in (("fp64", H64), ("mixed", H32), ("bf16", H16)):
        cg = B.ConjugateGradient()
        cg.set_tolerance(1e-10)
        cg.set_max_iterations(3000)
        cg.compute(H)
        sols[name] = cg.solve(b)
        assert cg.info() == B.OK, name
        assert rel(H64 @ sols[name], b) <= 2e-10, name
    # both residuals are at 1e-10 of the FP64 operator; solutions of a first-kind equation with
    # condition number ~1e3 may still differ by ~1e-7: the stated bound is 1e-6
    print("refined vs fp64 solve:", rel(sols["mixed"], sols["fp64"]), rel(sols["bf16"], sols["fp64"]))
    assert rel(sols["mixed"], sols["fp64"]) <= 1e-6 and rel(sols["bf16"], sols["fp64"]) <= 1e-6
    # Helmholtz GMRES with refinement
    oph = B.HelmholtzSingleLayerOperator()
    oph.set_wavenumber(2.0)
    ps3 = B.AnsatzSpace(B.Geometry(), 3, 2)
    Hh64 = B.H2Matrix(ps3, oph, n_cheb=5)
    Hh32 = B.H2Matrix(ps3, oph, n_cheb=5, precision=B.MIXED)
    # one column: FP32 near-field and coupling copies with FP64 sums; 3 columns: FP32 coupling
    # blocks with the FP64 near field; 12 columns: FP32 blocks on the FP32 matrix cores
    rng = np.random.default_rng(5)
    nh = ps3.number_of_dofs
    xh = rng.uniform(-1, 1, (nh, 12)) + 1j * rng.uniform(-1, 1, (nh, 12))
    e1, e3, e12 = (rel(Hh32 @ xh[:, 0], Hh64 @ xh[:, 0]), rel(Hh32 @ xh[:, :3], Hh64 @ xh[:, :3]),
                   rel(Hh32 @ xh, Hh64 @ xh))
    print("helmholtz mixed matvec rel err: 1 column %.2e, 3 columns %.2e, 12 columns %.2e" % (e1, e3, e12))
    assert e1 <= 1e-6 and e3 <= 1e-6 and e12 <= 1e-6
    bh = B.DiscreteLinearForm(B.PlaneWaves([[0.0, 0.0, 1.0]], 2.0), ps3).compute().get_discrete_linear_form()
    out = []
    for H in (Hh64, Hh32):
        gm = B.GMRES()
        gm.set_tolerance(1e-10)
        gm.set_restart(60)
        gm.set_max_iterations(3000)
        gm.compute(H)
        out.append(gm.solve(bh))
        assert gm.info() == B.OK
        assert rel(Hh64 @ out[-1], bh) <= 2e-10
    assert rel(out[1], out[0]) <= 1e-6


def test_gpu_h2_torch_work_is_ordered_before_the_product(gpu):
    """torch runs on its own HIP runtime: pending torch work that produces x finishes before the
    library reads x, and the product is complete when torch reads the result."""
    import torch
    os_, ps, op, H = make(O.LAPLACE, 0j, 3, 2, 4)
    n = ps.number_of_dofs
    big = torch.randn(1 << 24, device="cuda", dtype=torch.float64)
    x = torch.zeros(n, device="cuda", dtype=torch.float64)
    for _ in range(20):  # a queue of torch kernels in front of the one that writes x
        big = big * 1.0000001
    x += big[:n] * 0 + 1.0
    y = H @ x
    z = y * 2.0  # torch reads the product right away
    ref = H @ np.ones(n)
    assert rel(z.cpu().numpy() / 2.0, ref) <= 1e-14


def test_gpu_gmres_tasks_interleaved_equal_blocking(gpu):
    """Two GMRES tasks (bmi_gmres_start/poll/finish) on two handles and two streams, interleaved by
    polling from one thread, give the same bits and iteration counts as bmi_solve_gmres alone."""
    import ctypes as C
    import time
    import torch
    lib = B.library()
    ps = B.AnsatzSpace(B.Geometry(), 3, 2)
    cases = []
    for kappa, prec in ((2.0, B.FP64), (3.0, B.MIXED)):
        op = B.HelmholtzSingleLayerOperator()
        op.set_wavenumber(kappa)
        H = B.H2Matrix(ps, op, n_cheb=5, precision=prec)
        b = B.DiscreteLinearForm(B.PlaneWaves([[0.0, 0.6, 0.8], [1.0, 0.0, 0.0]], kappa), ps).compute(
            device=True).get_discrete_linear_form()
        cases.append((H, b.t().contiguous()))  # column-major: (columns, rows) contiguous
    torch.cuda.synchronize()
    tol, restart, maxit = 1e-8, 30, 3000
    n = ps.number_of_dofs
    ref = []
    for H, bt in cases:
        X = torch.zeros_like(bt)
        it = np.zeros(2, dtype=np.int32)
        rs = np.zeros(2)
        dp = C.POINTER(C.c_double)
        st = lib.bmi_solve_gmres(H._h, 2, C.cast(C.c_void_p(bt.data_ptr()), dp), C.cast(C.c_void_p(X.data_ptr()), dp),
                                 tol, restart, maxit, 1, it.ctypes.data_as(C.POINTER(C.c_int)),
                                 rs.ctypes.data_as(C.POINTER(C.c_double)), B.MEM_DEVICE, None)
        assert st == B.OK, (st, B.library().bmi_last_error())
        ref.append((X.clone(), it.copy()))
    streams, tasks, outs = [], [], []
    for H, bt in cases:
        s = C.c_void_p()
        assert lib.bmi_stream_create(C.byref(s)) == B.OK
        X = torch.zeros_like(bt)
        torch.cuda.synchronize()
        t = C.c_void_p()
        assert lib.bmi_gmres_start(H._h, 2, C.c_void_p(bt.data_ptr()), C.c_void_p(X.data_ptr()), tol, restart,
                                   maxit, 1, s, C.byref(t)) == B.OK, lib.bmi_last_error()
        streams.append(s)
        tasks.append(t)
        outs.append(X)
    done = [0, 0]
    turns = [0, 0]
    for _ in range(2_000_000):
        progress = False
        for k in range(2):
            if done[k] == 1:
                continue
            d = C.c_int()
            assert lib.bmi_gmres_poll(tasks[k], C.byref(d)) == B.OK, lib.bmi_last_error()
            done[k] = d.value
            turns[k] += d.value == 2
            progress = progress or d.value != 0
        if done == [1, 1]:
            break
        if not progress:
            time.sleep(50e-6)
    assert done == [1, 1]
    for k in range(2):
        it = np.zeros(2, dtype=np.int32)
        rs = np.zeros(2)
        st = lib.bmi_gmres_finish(tasks[k], it.ctypes.data_as(C.POINTER(C.c_int)),
                                  rs.ctypes.data_as(C.POINTER(C.c_double)))
        assert st == B.OK
        assert lib.bmi_stream_destroy(streams[k]) == B.OK
        torch.cuda.synchronize()
        assert torch.equal(outs[k], ref[k][0]), k
        assert list(it) == list(ref[k][1]), (it, ref[k][1])
        assert turns[k] > 1
    print("turns with work enqueued:", turns, "iterations:", [list(r[1]) for r in ref])
