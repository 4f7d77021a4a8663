"""Pins the CPU oracle against the reference's golden data, its known-answer unit tests and
closed forms.  CPU only.

Reference tests restated here: tests/test_DuffyTrick.cpp (+ tests/DuffyTrick/*),
tests/test_FMMTransferMatrices.cpp, tests/test_FMMForwardTransformation.cpp,
tests/test_FMMBackwardTransformation.cpp, tests/test_LazyEigenSum.cpp,
tests/test_GeometryImportAndEval.cpp, tests/test_Bernstein.cpp, tests/test_Projector.cpp.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle_lib as O

HERE = os.path.dirname(os.path.abspath(__file__))
SCREEN = os.path.join(HERE, "golden", "screen.dat")


# ------------------------------------------------------------------ quadrature
def _oracle_table():
    import re
    txt = open(os.path.join(os.path.dirname(HERE), "oracle", "gauss_table.inc")).read()
    offs = list(map(int, re.search(r"or_gl_offset\[52\] = \{(.*?)\}", txt).group(1).split(",")))
    xs = [float.fromhex(t.strip()) for t in re.search(r"or_gl_x\[\d+\] = \{(.*?)\};", txt, re.S).group(1).split(",") if t.strip()]
    ws = [float.fromhex(t.strip()) for t in re.search(r"or_gl_w\[\d+\] = \{(.*?)\};", txt, re.S).group(1).split(",") if t.strip()]
    return offs, xs, ws


def test_gauss_rules_match_reference_literals():
    """GaussLegendreRule<1..20> literals (Quadrature/GaussLegendreRule.hpp:57-688), bit-exact."""
    ref = json.load(open(os.path.join(HERE, "golden", "gauss_legendre_ref.json")))
    offs, xs, ws = _oracle_table()
    for n in range(1, 21):
        r = ref[str(n)]
        o = offs[n - 1]
        assert [float.fromhex(v) for v in r["x"]] == xs[o:o + n]
        assert [float.fromhex(v) for v in r["w"]] == ws[o:o + n]


def test_gauss_rules_exact_on_polynomials():
    offs, xs, ws = _oracle_table()
    for n in range(1, 52):
        x = np.array(xs[offs[n - 1]:offs[n - 1] + n])
        w = np.array(ws[offs[n - 1]:offs[n - 1] + n])
        assert np.all(np.diff(x) > 0) and abs(w.sum() - 1) < 1e-14
        for k in range(0, 2 * n, max(1, n // 3)):
            assert abs(np.dot(w, x ** k) - 1.0 / (k + 1)) < 2e-14 * (k + 1)


# ------------------------------------------------------------------ basis and geometry
@pytest.mark.parametrize("p", [0, 1, 2, 3, 4, 7])
def test_bernstein_closed_form(p):
    for x in [0.0, 0.13, 0.5, 0.77, 1.0]:
        b = O.bernstein(p, x)
        ex = [math.comb(p, k) * x ** k * (1 - x) ** (p - k) for k in range(p + 1)]
        assert np.allclose(b, ex, rtol=1e-14, atol=1e-15)
        assert abs(b.sum() - 1) < 1e-14
        d = O.bernstein(p, x, der=True)
        eps = 1e-6
        fd = (O.bernstein(p, min(x + eps, 1.0)) - O.bernstein(p, max(x - eps, 0.0))) / (min(x + eps, 1.0) - max(x - eps, 0.0))
        assert np.allclose(d, fd, atol=1e-6 * (p + 1) ** 2)


def test_sphere_geometry():
    """tests/test_GeometryImportAndEval.cpp: |x| = 1 and normal parallel to x (1e-9)."""
    g = O.Geometry()
    rng = np.random.default_rng(0)
    for p in range(6):
        for s, t in rng.uniform(0, 1, (20, 2)):
            sp = g.surface_point(p, (s, t))
            assert abs(np.linalg.norm(sp[3:6]) - 1) < 1e-9
            n = np.cross(sp[6:9], sp[9:12])
            assert abs(abs(np.dot(n / np.linalg.norm(n), sp[3:6])) - 1) < 1e-9
            eps = 1e-6
            fd = (g.surface_point(p, (min(s + eps, 1), t))[3:6] - g.surface_point(p, (max(s - eps, 0), t))[3:6]) / (
                min(s + eps, 1) - max(s - eps, 0))
            assert np.allclose(fd, sp[6:9], atol=1e-6)
            # eval and updateSurfacePoint agree
            assert np.allclose(g.eval(p, s, t), sp[3:6], atol=1e-15)


@pytest.mark.parametrize("M", [0, 1, 2, 3, 4])
def test_element_tree_counts(M):
    S = O.Space(O.Geometry(), M, 1)
    assert S.nel == 6 * 4 ** M
    assert S.npoints == S.nel + 2  # closed quadrilateral surface: V - E + F = 2
    n = 2 ** M
    assert S.ndof == 6 * (n + 1) ** 2  # P=1 discontinuous across patches (test_Projector.cpp)
    el = S.elements()
    # enclosing balls contain the element corners
    assert np.all(el["radius"] > 0)
    if M >= 1:
        mids, rads = S.clusters(M - 1)
        assert len(rads) == 6 * 4 ** (M - 1)
        # parent balls contain the child balls
        for c in range(len(rads)):
            for k in range(4):
                e = 4 * c + k
                assert np.linalg.norm(el["mid"][e] - mids[c]) + el["radius"][e] <= rads[c] * (1 + 1e-12)


def test_neighbour_structure():
    """Every element of the sphere has 4 edge neighbours and 3 or 4 vertex neighbours."""
    S = O.Space(O.Geometry(), 2, 1)
    cls = np.zeros((S.nel, 5), dtype=int)
    for i in range(S.nel):
        for j in range(S.nel):
            c, _ = S.pair_info(i, j)
            cls[i, c[0]] += 1
    assert np.all(cls[:, 2] == 1)
    assert np.all(cls[:, 3] == 4)
    assert set(cls[:, 4].tolist()) <= {3, 4}
    assert np.sum(cls[:, 4] == 3) == 8 * 3  # 3 elements at each of the 8 cube corners


def test_projector_partition_of_unity():
    """T * ones = ones (the B-splines restricted to elements sum to one), test_Projector.cpp."""
    for M, P in [(1, 1), (2, 2), (3, 2)]:
        S = O.Space(O.Geometry(), M, P)
        T = S.T_dense()
        assert np.allclose(T @ np.ones(S.ndof), 1.0, atol=1e-12)
        assert S.ndof == 6 * (2 ** M + P) ** 2


def test_divconforming_dof_count():
    """DivConforming space: 2 (n+P)(n+P-1) per patch minus one identification per edge DOF."""
    M, P = 2, 2
    S = O.Space(O.Geometry(), M, P, form=1)
    n = 2 ** M
    pre = 6 * 2 * (n + P) * (n + P - 1)
    assert S.ndof == pre - 12 * (n + P - 1)


# ------------------------------------------------------------------ Duffy known answers
def _anti(x0, x1):
    sig = 1e-3
    return 0.5 * math.log(sig + (x0 - x1) ** 2) + (x1 - x0) / math.sqrt(sig) * math.atan((x0 - x1) / math.sqrt(sig))


def _int2d(a):
    return _anti(a[1][0], a[1][1]) - _anti(a[1][0], a[0][1]) - _anti(a[0][0], a[1][1]) + _anti(a[0][0], a[0][1])


def near_singular_integral(ax):
    """testFunction4DIntegral (tests/DuffyTrick/test_testFunctions.hpp)"""
    return (math.exp(-ax[0][0]) - math.exp(-ax[1][0])) * _int2d([[ax[0][1], ax[0][2]], [ax[1][1], ax[1][2]]]) * (
        0.5 * math.exp(-2 * ax[0][3]) - 0.5 * math.exp(-2 * ax[1][3]))


def analytic_integral(ax):
    pi = math.pi
    return ((math.cos(pi * ax[0][0]) - math.cos(pi * ax[1][0])) / pi * 0.5 *
            (math.cos(2 * pi * ax[0][2]) - math.cos(2 * pi * ax[1][2])) / pi *
            (math.exp(-ax[0][1]) - math.exp(-ax[1][1])) * (math.exp(ax[1][3]) - math.exp(ax[0][3])))


def duffy_cases(space):
    """(which, e1, rot1, e2, rot2, exact) for the reference Duffy tests on the level-2 screen"""
    el = space.elements()
    h = 0.25
    out = []
    for e1 in range(space.nel):
        for e2 in range(space.nel):
            info, _ = space.pair_info(e1, e2)
            l1, l2 = el["llc"][e1], el["llc"][e2]
            ax = [[l1[0], l1[1], l2[0], l2[1]], [l1[0] + h, l1[1] + h, l2[0] + h, l2[1] + h]]
            out.append((1, e1, 0, e2, 0, analytic_integral(ax) / h ** 2, O.DUMMY_ANALYTIC))
            if e1 == e2:
                out.append((2, e1, 0, e2, 0, near_singular_integral(ax) / h ** 2, O.DUMMY_NEARSINGULAR))
            if info[0] in (3, 4):
                out.append((info[0], e1, info[1], e2, info[2], near_singular_integral(ax) / h ** 2,
                            O.DUMMY_NEARSINGULAR))
    return out


DUFFY_TOL = {1: 1e-12, 2: 1e-6, 3: 1e-4, 4: 1e-4}  # tests/Test.hpp:44-50


def test_duffy_known_answers():
    S = O.Space(O.Geometry(SCREEN), 2, 0)
    worst = {k: 0.0 for k in DUFFY_TOL}
    for which, e1, r1, e2, r2, ex, fn in duffy_cases(S):
        v = S.integrate(O.DUMMY, which, e1, r1, e2, r2, 16, dummy_fn=fn)[0, 0]
        worst[which] = max(worst[which], abs(ex - v) / abs(ex))
    for k, tol in DUFFY_TOL.items():
        assert worst[k] < tol, (k, worst[k])


def test_duffy_rotation_plants_are_caught():
    """The smooth test functions above do not see the Duffy rotations (any parametrisation of
    the squares integrates them); the singular Laplace kernel does.  With the rotations of
    compareElements the edge and vertex rules hit the independent brute-force value to
    quadrature accuracy; any other rotation pair misses it by orders of magnitude."""
    S = O.Space(O.Geometry(SCREEN), 2, 0)
    h = 0.25
    el = S.elements()
    for which, off in [(3, (1, 0)), (4, (1, 1))]:
        ex = _square_pair_integral(off) * h ** 3 / (4 * math.pi) / h ** 2
        e2 = [e for e in range(S.nel) if np.allclose(el["llc"][e] - el["llc"][0], [off[0] * h, off[1] * h])][0]
        info, _ = S.pair_info(0, e2)
        assert info[0] == which
        for r1 in range(4):
            for r2 in range(4):
                v = S.integrate(O.LAPLACE, which, 0, r1, e2, r2, 12)[0, 0]
                err = abs(v - ex) / ex
                if (r1, r2) == (info[1], info[2]):
                    assert err < 1e-10, (which, r1, r2, err)
                else:
                    assert err > 1e-8, (which, r1, r2, err)


# ------------------------------------------------------------------ Laplace closed forms
def _square_pair_integral(offset):
    """int_Q int_Q' 1/|x-y| for unit squares Q and Q' = Q + offset (overlap-correlation form)."""
    from scipy import integrate

    def g(u, c):
        return max(0.0, 1.0 - abs(u - c))

    ox, oy = offset

    def f(v, u):
        r = math.hypot(u, v)
        return g(u, ox) * g(v, oy) / r if r > 0 else 0.0

    tot = 0.0
    xs = sorted(set([ox - 1, ox, ox + 1, 0.0]))
    ys = sorted(set([oy - 1, oy, oy + 1, 0.0]))
    for a, b in zip(xs[:-1], xs[1:]):
        for c, d in zip(ys[:-1], ys[1:]):
            val, _ = integrate.dblquad(f, a, b, c, d, epsabs=1e-13, epsrel=1e-13)
            tot += val
    return tot


@pytest.mark.parametrize("which,offset", [(2, (0, 0)), (3, (1, 0)), (4, (1, 1))])
def test_laplace_flat_pairs_brute_force(which, offset):
    """Laplace single layer (P = 0) on identical, edge- and vertex-adjacent flat squares versus
    an independent adaptive quadrature of the 4D integral reduced to 2D."""
    if which == 2:
        exact_unit = 4 * math.log(1 + math.sqrt(2)) - 4.0 / 3.0 * (math.sqrt(2) - 1)
        assert abs(exact_unit - _square_pair_integral((0, 0))) < 1e-9
    else:
        exact_unit = _square_pair_integral(offset)
    S = O.Space(O.Geometry(SCREEN), 2, 0)
    h = 0.25
    el = S.elements()
    target = np.array([offset[0] * h, offset[1] * h])
    e1 = 0
    e2 = [e for e in range(S.nel) if np.allclose(el["llc"][e] - el["llc"][e1], target)][0]
    info, _ = S.pair_info(e1, e2)
    assert info[0] == which
    v = S.integrate(O.LAPLACE, which, e1, info[1], e2, info[2], 20)[0, 0]
    # Galerkin entry = (1/h^2) int_E int_E' K  with the L2-scaled weights
    ex = exact_unit * h ** 3 / (4 * math.pi) / h ** 2
    assert abs(v - ex) / ex < 1e-10


# ------------------------------------------------------------------ FMM identities
def test_fmm_transfer_identity():
    """tests/test_FMMTransferMatrices.cpp: T_k * moment(level 1) = block k of moment(level 0)."""
    S = O.Space(O.Geometry(SCREEN), 3, 0)
    n, n2 = 9, 81
    T = O.transfer_matrix(n)
    m1 = S.moment(1, 2, n)
    m2 = S.moment(0, 3, n)
    err = sum(np.linalg.norm(m2[:, k * m1.shape[1]:(k + 1) * m1.shape[1]] - T[:, k * n2:(k + 1) * n2] @ m1)
              for k in range(4))
    assert err < 1e-12


def test_fmm_forward_identity():
    """tests/test_FMMForwardTransformation.cpp"""
    S = O.Space(O.Geometry(SCREEN), 6, 0)
    n, cl = 9, 2
    T = O.transfer_matrix(n)
    ms = [np.vstack([m, m]) for m in (S.moment(cl, 6 - cl, n), S.moment(cl - 1, 7 - cl, n), S.moment(cl - 2, 8 - cl, n))]
    rv = np.random.default_rng(0).uniform(-1, 1, ms[2].shape[1])
    Xs = [rv.reshape(-1, m.shape[1]).T for m in ms]
    fw = O.forward_transform(ms[0], T, 4, Xs[0])
    err = sum(np.linalg.norm(ms[k] @ Xs[k] - fw[k]) for k in range(3))
    assert err < 1e-12


def test_fmm_backward_identity():
    """tests/test_FMMBackwardTransformation.cpp: the backward sweep is the transpose of the
    forward sweep, so <fw(x), y> = <x, bw(y)>; checked through the moment identity
    moment(level l)^T y = (T-chain)^T applied to the leaf moments."""
    S = O.Space(O.Geometry(SCREEN), 5, 0)
    n = 9
    T = O.transfer_matrix(n)
    m0 = S.moment(2, 3, n)
    m1 = S.moment(1, 4, n)
    rng = np.random.default_rng(1)
    y = rng.uniform(-1, 1, m1.shape[0])
    # backward of one coarse column: children receive T_k^T y, then moment^T
    lhs = m1.T @ y
    rhs = np.concatenate([m0.T @ (T[:, k * 81:(k + 1) * 81].T @ y) for k in range(4)])
    assert np.linalg.norm(lhs - rhs) < 1e-12


def test_chebyshev_lagrange():
    x = O.cheb_points(7)
    assert np.all(np.diff(x) > 0) and x[0] > 0 and x[-1] < 1
    T = O.transfer_matrix(4)
    # interpolation reproduces constants: each child block's rows sum to one
    for k in range(4):
        assert np.allclose(T[:, k * 16:(k + 1) * 16].sum(axis=0), 1.0, atol=1e-13)


# ------------------------------------------------------------------ assembled operators
def test_h2_vs_dense_lazy_eigen_sum():
    """tests/test_LazyEigenSum.cpp: |D a - H2 a| < 1e-8 (Helmholtz, kappa 0, P 1, M 2, np 9)."""
    S = O.Space(O.Geometry(), 2, 1)
    D = S.dense_galerkin(O.HELMHOLTZ, 0j)
    H = O.H2(S, O.HELMHOLTZ, 0j, 9)
    a = np.random.default_rng(3).uniform(-1, 1, S.ndof)
    assert np.linalg.norm(D @ a - H.matvec(a.astype(complex))) < 1e-8
    assert np.linalg.norm(3 * (D @ a) - 3 * H.matvec(a.astype(complex))) < 1e-8


def test_laplace_constant_function_closed_form():
    """V[1] = 1 on the unit sphere, so c^T A c = |S| = 4 pi for the Laplace single layer, where
    c = h * ones are the coefficients of the constant function in the L2-scaled basis
    (quadrature weights h*w, DuffyTrick/farFieldQuadratureNodes.hpp:36-41)."""
    S = O.Space(O.Geometry(), 2, 1)
    A = S.dense_galerkin(O.LAPLACE)
    one = np.ones(S.ndof) * 2.0 ** -2
    assert abs(one @ A @ one - 4 * math.pi) / (4 * math.pi) < 1e-6
    assert np.abs(A - A.T).max() < 1e-12 * np.abs(A).max()
    assert np.linalg.eigvalsh(A).min() > 0


def test_helmholtz_constant_function_closed_form():
    """V_k[1] = sin(k) e^{-ik}/k on the unit sphere (kernel e^{-ikr}/(4 pi r))."""
    k = 2.0
    S = O.Space(O.Geometry(), 2, 1)
    A = S.dense_galerkin(O.HELMHOLTZ, k + 0j)
    one = np.ones(S.ndof) * 2.0 ** -2
    ex = 4 * math.pi * math.sin(k) * complex(math.cos(k), -math.sin(k)) / k
    assert abs(one @ A @ one - ex) / abs(ex) < 1e-6
    assert np.abs(A - A.T).max() < 1e-12 * np.abs(A).max()


def test_rhs_and_solve_closed_form():
    """Laplace SL with g = 1/|x - x0|, x0 outside the unit sphere: the exterior density
    reproduces g; the H2 solve agrees with the dense solve."""
    S = O.Space(O.Geometry(), 2, 1)
    A = S.dense_galerkin(O.LAPLACE)
    x0 = np.array([2.0, 0.5, -0.3])
    b = S.rhs(0, x0)
    x = np.zeros(S.ndof)
    it = O.lib().or_cg(S.ndof, O._dp(np.asfortranarray(A)), O._dp(b), O._dp(x), 1e-12, 2000)
    assert it < 2000
    assert np.linalg.norm(A @ x - b) < 1e-10 * np.linalg.norm(b)
    assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-8, atol=1e-10)


# ------------------------------------------------------------------ Maxwell (EFIE)
ELLIPSOID = (1.0, 0.7, 0.5)


def _bernstein_np(P, x):
    """Bernstein values and derivatives of degree P from the closed form, vectorised over x."""
    x = np.asarray(x, dtype=np.float64)
    B = np.stack([math.comb(P, k) * x ** k * (1 - x) ** (P - k) for k in range(P + 1)])
    low = [math.comb(P - 1, k) * x ** k * (1 - x) ** (P - 1 - k) for k in range(P)]
    D = np.stack([P * ((low[k - 1] if k >= 1 else 0.0) - (low[k] if k <= P - 1 else 0.0)) for k in range(P + 1)])
    return B, D


def _maxwell_block_numpy(S, geo, e1, e2, kappa, n, div=1.0):
    """EFIE element-pair block from its definition with an n-point tensor Gauss rule per element:
    K = exp(-i k r) / (4 pi r), vector part K w w (t_a . t_b) phi_i phi_j, divergence part
    -K w w / (k^2 h^2) dphi_i dphi_j, rows/cols ordered (component, iy*(P+1)+ix)."""
    P, I = S.P, S.P + 1
    I2 = I * I
    h = 2.0 ** -S.M
    t, w = np.polynomial.legendre.leggauss(n)
    t, w = (t + 1) / 2, w / 2
    el = S.elements()

    def element_points(e):
        xi0 = np.repeat(t, n)
        xi1 = np.tile(t, n)
        wt = h * np.repeat(w, n) * np.tile(w, n)
        X, T0, T1 = [], [], []
        for a, b in zip(xi0, xi1):
            sp = geo.surface_point(int(el["patch"][e]), el["llc"][e] + h * np.array([a, b]))
            X.append(sp[3:6]), T0.append(sp[6:9]), T1.append(sp[9:12])
        Bx, Dx = _bernstein_np(P, xi0)
        By, Dy = _bernstein_np(P, xi1)
        phi = np.stack([Bx[ix] * By[iy] for iy in range(I) for ix in range(I)])
        dphi = [np.stack([Dx[ix] * By[iy] for iy in range(I) for ix in range(I)]),
                np.stack([Bx[ix] * Dy[iy] for iy in range(I) for ix in range(I)])]
        return np.array(X), [np.array(T0), np.array(T1)], wt, phi, dphi

    X1, T1, w1, phi1, d1 = element_points(e1)
    X2, T2, w2, phi2, d2 = element_points(e2)
    r = np.linalg.norm(X1[:, None, :] - X2[None, :, :], axis=2)
    K = np.exp(-1j * kappa * r) / (4 * math.pi * r) * w1[:, None] * w2[None, :]
    blk = np.zeros((2 * I2, 2 * I2), dtype=complex)
    for rc in range(2):
        for cc in range(2):
            tt = T1[rc] @ T2[cc].T
            blk[rc * I2:(rc + 1) * I2, cc * I2:(cc + 1) * I2] = (
                phi1 @ (K * tt) @ phi2.T - div * d1[rc] @ K @ d2[cc].T / (kappa ** 2 * h ** 2))
    return blk


@pytest.mark.parametrize("P", [1, 2])
def test_maxwell_integrand_vs_definition(P):
    """The oracle's EFIE integrand (vector and divergence parts, component layout, kappa and h
    scaling) equals the definition evaluated independently, on a regular pair with a 16-point
    rule where both sides are converged."""
    geo = O.Geometry(scale=ELLIPSOID)
    S = O.Space(geo, 1, P, form=1)
    kappa = 2.0 + 0.3j
    pairs = []
    for e1 in (0, 5):
        far = [e2 for e2 in range(S.nel) if S.pair_info(e1, e2)[0][0] == 0 and S.pair_info(e1, e2)[1] > 0]
        pairs.append((e1, far[len(far) // 2]))
    for e1, e2 in pairs:
        got = S.integrate(O.MAXWELL, 1, e1, 0, e2, 0, 15, kappa)
        ref = _maxwell_block_numpy(S, geo, e1, e2, kappa, 16)
        assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref)
        # both parts matter: a dropped or sign-flipped divergence part misses by far more than
        # the tolerance, and so do swapped vector components
        for div in (0.0, -1.0):
            bad = _maxwell_block_numpy(S, geo, e1, e2, kappa, 16, div)
            assert np.linalg.norm(got - bad) > 1e-3 * np.linalg.norm(ref)
        I2 = S.I2
        sw = ref.copy()
        sw[:I2, I2:], sw[I2:, :I2] = ref[I2:, :I2], ref[:I2, I2:]
        assert np.linalg.norm(got - sw) > 1e-3 * np.linalg.norm(ref)


def test_maxwell_h2_vs_dense():
    """The Maxwell interpolation FMM (derivative moments, 4x4 interpolant per vector component
    pair) reproduces the dense EFIE matrix (tests/test_LazyEigenSum.cpp analogue)."""
    geo = O.Geometry(scale=ELLIPSOID)
    S = O.Space(geo, 2, 1, form=1)
    kappa = 2.0 + 0j
    D = S.dense_galerkin(O.MAXWELL, kappa)
    assert np.abs(D - D.T).max() <= 1e-12 * np.abs(D).max()
    H = O.H2(S, O.MAXWELL, kappa, 9)
    a = np.random.default_rng(5).uniform(-1, 1, S.ndof) + 0j
    Da, Ha = D @ a, H.matvec(a)
    assert np.linalg.norm(Da - Ha) <= 1e-7 * np.linalg.norm(Da)
    # a coarse interpolation is visibly worse, so the comparison has power
    H3 = O.H2(S, O.MAXWELL, kappa, 3)
    assert np.linalg.norm(Da - H3.matvec(a)) > 10 * np.linalg.norm(Da - Ha)


def _dipole_np(x, kappa, pos, length):
    """Hertz dipole field (Jackson, Classical Electrodynamics, 3rd ed., p. 411) with the
    exp(-i kappa r) convention of the reference (examples/Data.hpp:76-90)."""
    c = np.asarray(x, dtype=float) - pos
    r = np.linalg.norm(c)
    n = c / r
    e = np.exp(-1j * kappa * r)
    far = kappa ** 2 * e / r * (length - (n @ length) * n)
    near = (1 / r ** 3 + 1j * kappa / r ** 2) * e * (3 * (n @ length) * n - length)
    return far + near


def test_dipole_field_solves_maxwell():
    """curl curl E = kappa^2 E and div E = 0 away from the dipole (fourth-order differences)."""
    kappa, pos, length = 2.0, np.array([0.1, -0.05, 0.08]), np.array([0.02, 0.05, -0.07])
    x0, hstep = np.array([0.6, 0.3, -0.4]), 1e-3

    def jac(f, x):
        J = np.zeros((3, 3), dtype=complex)
        for k in range(3):
            e = np.zeros(3)
            e[k] = hstep
            J[:, k] = (-f(x + 2 * e) + 8 * f(x + e) - 8 * f(x - e) + f(x - 2 * e)) / (12 * hstep)
        return J

    E = lambda x: _dipole_np(x, kappa, pos, length)
    curl = lambda f: (lambda x: (lambda J: np.array([J[2, 1] - J[1, 2], J[0, 2] - J[2, 0], J[1, 0] - J[0, 1]]))(jac(f, x)))
    cc = curl(curl(E))(x0)
    assert np.linalg.norm(cc - kappa ** 2 * E(x0)) < 1e-5 * np.linalg.norm(kappa ** 2 * E(x0))
    assert abs(np.trace(jac(E, x0))) < 1e-6 * np.linalg.norm(E(x0))


def test_rotated_tangential_trace_vs_definition():
    """RHS of the EFIE: Galerkin projection of the rotated tangential trace of the dipole field
    onto the DivConforming space, evaluated independently with the same (P+2)-point rule."""
    geo = O.Geometry(scale=ELLIPSOID)
    for P in (1, 2):
        S = O.Space(geo, 1, P, form=1)
        kappa, pos, length = 2.0, np.array([0.1, -0.05, 0.08]), np.array([0.02, 0.05, -0.07])
        got = S.rhs(2, np.concatenate([pos, length]), kappa)
        I, I2, h, n = P + 1, S.I2, 2.0 ** -S.M, P + 2
        t, w = np.polynomial.legendre.leggauss(n)
        t, w = (t + 1) / 2, w / 2
        el = S.elements()
        sup = np.zeros(S.nsuper, dtype=complex)
        for e in range(S.nel):
            for a in range(n):
                for b in range(n):
                    sp = geo.surface_point(int(el["patch"][e]), el["llc"][e] + h * np.array([t[a], t[b]]))
                    ds, dt = sp[6:9], sp[9:12]
                    nrm = np.cross(ds, dt)
                    nrm /= np.linalg.norm(nrm)
                    f = _dipole_np(sp[3:6], kappa, pos, length)
                    tang = (f - (f @ nrm) * nrm) * (h * w[a] * w[b])
                    Bx, _ = _bernstein_np(P, np.array([t[a]]))
                    By, _ = _bernstein_np(P, np.array([t[b]]))
                    phi = np.array([Bx[ix, 0] * By[iy, 0] for iy in range(I) for ix in range(I)])
                    sup[e * I2:(e + 1) * I2] += (ds @ tang) * phi
                    sup[S.nel * I2 + e * I2:S.nel * I2 + (e + 1) * I2] += (dt @ tang) * phi
        ref = S.T_dense().T @ sup
        assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref)
