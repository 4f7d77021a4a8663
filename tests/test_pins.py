"""Pins of the oracle functions that no reference golden vector covers directly: the near-field
quadrature degree, the DivConforming glue, the Dirichlet-trace right-hand sides and the
oracle's GMRES.  CPU only.

* Degree rule (LinearOperator/LinearOperatorBase.hpp:69-89, DuffyTrick/evaluateBilinearForm.hpp
  :31-38): tests/golden/nearfield_degree.json holds the reference expression evaluated in IEEE
  doubles, cross-checked against 50-digit arithmetic wherever the value is not on an integer.
* DivConforming glue (AnsatzSpace/Glue.hpp:335,568; restates tests/test_Glue.cpp as an
  invariant): the normal flux of the Piola-mapped field is continuous across every patch
  interface of the sphere and the ellipsoid, and a planted sign flip breaks it.
* Right-hand sides (LinearForm/DiscreteLinearForm.hpp:44-83, LinearForm/DirichletTrace.hpp
  :31-70): with c the coefficients of the constant function, c^T b is the surface integral of
  the data, known in closed form on the unit sphere; the error falls at the order of the
  (P+2)-point Gauss rule.
* GMRES (consumer of examples/HelmholtzSingleLayerH2.cpp:82-85): agrees with a dense LU solve.
"""
import json
import math
import os
import sys

import numpy as np
import pytest

import bembel_amd as B
import oracle_lib as O

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
from make_degree_fixture import degree_ieee  # noqa: E402

FIXTURE = json.load(open(os.path.join(HERE, "golden", "nearfield_degree.json")))["entries"]


# ------------------------------------------------------------------ near-field degree
def test_degree_rule_matches_fixture():
    for e in FIXTURE:
        got = O.lib().or_nfield_degree(e["P"], float.fromhex(e["dist"]), e["L"])
        assert got == e["degree"], e


def test_degree_rule_known_values():
    """Touching pairs: the real value (2+P)L/2 is an integer for even P L and the truncated
    IEEE result falls below it for some (P, L).  A rounded instead of truncated result, or an
    alpha off by one, changes these."""
    want = {(1, 3): 4, (1, 4): 6, (1, 5): 7, (1, 6): 8, (1, 7): 10, (2, 3): 6, (2, 4): 8, (2, 5): 10,
            (2, 6): 12, (2, 7): 13, (3, 2): 4, (3, 6): 14}
    for (P, L), d in want.items():
        assert O.lib().or_nfield_degree(P, 0.0, L) == d, (P, L)
        e = [x for x in FIXTURE if x["P"] == P and x["L"] == L and float.fromhex(x["dist"]) == 0.0][0]
        assert e["degree"] == d
    # the exact values away from integers truncate as 50-digit arithmetic says
    off = [e for e in FIXTURE if not e["boundary"]]
    assert len(off) > 300
    for e in off:
        ex = float(e["exact"])
        assert e["degree"] == (math.floor(ex) if ex >= 0 else math.ceil(ex))


@pytest.mark.parametrize("M,P", [(3, 1), (4, 2), (6, 2)])
def test_degree_of_mesh_pairs(M, P):
    """Degrees of real element pairs from the oracle and from the product's host code equal the
    reference expression (max with the far-field degree P+2) at each pair's distance."""
    og = O.Geometry()
    os_ = O.Space(og, M, P)
    ps = B.AnsatzSpace(B.Geometry(), M, P)
    n = os_.nel
    rng = np.random.default_rng(10 + M)
    e1 = rng.integers(0, n, 300)
    pairs = []
    for a in e1:
        pairs.append((a, a))
        pairs.extend((a, b) for b in rng.integers(0, n, 40))
    el = os_.elements()
    # add near pairs: elements whose balls are within a few element sizes
    mid = el["mid"]
    for a in e1[:60]:
        d = np.linalg.norm(mid - mid[a], axis=1)
        pairs.extend((a, b) for b in np.argsort(d)[:40])
    pairs = np.array(pairs, dtype=np.int32)
    prod = ps.pair_info(pairs)
    seen = set()
    for k, (a, b) in enumerate(pairs):
        info, dist = os_.pair_info(int(a), int(b))
        want = max(degree_ieee(P, dist, M)[1], P + 2)
        assert info[3] == want, (a, b, dist, info, want)
        assert prod[k][3] == want
        seen.add(want)
    assert len(seen) >= (2 if M < 4 else 3)  # the sample spans several degrees


# ------------------------------------------------------------------ DivConforming glue
def _bern(P, x):
    return np.array([math.comb(P, k) * x ** k * (1 - x) ** (P - k) for k in range(P + 1)])


def _interface_fluxes(geo, S, T, c):
    """Normal flux u.nu of the Piola-mapped field u = (ds x c_s + dt x c_t)/|ds x dt| at the
    midpoints of the element edges on patch boundaries, from both sides of each interface
    (nu is the unit conormal of the first side).  Returns an (n, 2) array."""
    P, I, I2, M = S.P, S.P + 1, S.I2, S.M
    h, n = 2.0 ** -M, 2 ** M
    el = S.elements()
    sup = T @ c
    idx = {}
    for e in range(S.nel):
        idx[(int(el["patch"][e]), int(round(el["llc"][e][0] / h)), int(round(el["llc"][e][1] / h)))] = e
    pts = {}
    for p in range(6):
        for i in range(n):
            a = (i + 0.5) * h
            for s, t in ((a, 0.0), (1.0, a), (a, 1.0), (0.0, a)):
                e = idx[(p, min(int(s / h), n - 1), min(int(t / h), n - 1))]
                xi = ((s - el["llc"][e][0]) / h, (t - el["llc"][e][1]) / h)
                sp = geo.surface_point(p, (s, t))
                b0, b1 = _bern(P, xi[0]), _bern(P, xi[1])
                phi = np.array([b0[jx] * b1[jy] for jy in range(I) for jx in range(I)])
                cs = sup[e * I2:(e + 1) * I2] @ phi
                ct = sup[S.nel * I2 + e * I2:S.nel * I2 + (e + 1) * I2] @ phi
                ds, dt = sp[6:9], sp[9:12]
                nrm = np.cross(ds, dt)
                det = np.linalg.norm(nrm)
                tau = dt if s in (0.0, 1.0) else ds
                pts.setdefault(tuple(np.round(sp[3:6], 9)), []).append(
                    ((ds * cs + dt * ct) / det, nrm / det, tau / np.linalg.norm(tau)))
    out = []
    for v in pts.values():
        assert len(v) == 2
        (u1, n1, t1), (u2, _, _) = v
        nu = np.cross(n1, t1)
        out.append((u1 @ nu, u2 @ nu))
    return np.array(out)


@pytest.mark.parametrize("scale", [(1.0, 1.0, 1.0), (1.0, 0.7, 0.5)])
@pytest.mark.parametrize("M,P", [(1, 1), (1, 2), (2, 2), (2, 3)])
def test_divconforming_normal_flux_is_continuous(scale, M, P):
    geo = O.Geometry(scale=scale)
    S = O.Space(geo, M, P, form=1)
    T = S.T_dense()
    c = np.random.default_rng(M * 10 + P).uniform(-1, 1, S.ndof)
    f = _interface_fluxes(geo, S, T, c)
    assert len(f) == 12 * 2 ** M
    scale_f = np.abs(f).max()
    assert np.abs(f[:, 0] - f[:, 1]).max() <= 1e-12 * scale_f
    # the interface DOFs are the ones shared by two patches: one per edge function of each of
    # the 12 cube edges
    el = S.elements()
    patch_of_row = np.concatenate([np.repeat(el["patch"], S.I2)] * 2)
    shared = [j for j in range(S.ndof) if len(set(patch_of_row[np.nonzero(T[:, j])[0]])) == 2]
    assert len(shared) == 12 * (2 ** M + P - 1)
    # planted mistake: the orientation sign of one glued DOF flipped on one side
    j = shared[len(shared) // 3]
    rows = np.nonzero(T[:, j])[0]
    p0 = patch_of_row[rows[0]]
    Tb = T.copy()
    Tb[rows[patch_of_row[rows] == p0], j] *= -1
    cb = np.zeros(S.ndof)
    cb[j] = 1.0
    fb = _interface_fluxes(geo, S, Tb, cb)
    assert np.abs(fb[:, 0] - fb[:, 1]).max() > 1e-3 * np.abs(fb).max()
    fg = _interface_fluxes(geo, S, T, cb)
    assert np.abs(fg[:, 0] - fg[:, 1]).max() <= 1e-12 * np.abs(fg).max()


# ------------------------------------------------------------------ Dirichlet-trace RHS
def _rhs_errors(P, Ms):
    """relative errors of c^T b against the closed forms for point sources inside and outside
    the unit sphere and for a plane wave, c = 2^-M (the constant function in the L2-scaled
    basis, quadrature weights h w: DuffyTrick/farFieldQuadratureNodes.hpp:36-41)"""
    g = O.Geometry()
    x_in = np.array([0.1, -0.2, 0.15])
    x_out = np.array([2.0, 0.5, -0.3])
    k, d = 4.0, np.array([1.0, 2.0, -2.0]) / 3.0
    exact = (4 * math.pi, 4 * math.pi / np.linalg.norm(x_out), 4 * math.pi * math.sin(k) / k)
    errs = []
    for M in Ms:
        S = O.Space(g, M, P)
        c = np.full(S.ndof, 2.0 ** -M)
        vals = (c @ S.rhs(0, x_in), c @ S.rhs(0, x_out), c @ S.rhs(1, d, k))
        errs.append([abs(v - e) / abs(e) for v, e in zip(vals, exact)])
    return np.array(errs)


@pytest.mark.parametrize("P", [1, 2])
def test_rhs_surface_integrals_closed_form(P):
    """int_S 1/|x - x0| dS = 4 pi (|x0| < 1) and 4 pi/|x0| (|x0| > 1); int_S exp(i k d.x) dS =
    4 pi sin(k)/k.  The (P+2)-point rule per direction converges like h^(2P+4)."""
    e = _rhs_errors(P, [1, 2, 3, 4])
    assert np.all(np.diff(e, axis=0) < 0)
    assert e[-1].max() < (1e-8 if P == 1 else 1e-11)
    rate = 2.0 ** (2 * P + 4)
    # outside source and plane wave: the asymptotic rate over the last refinements
    for col in (1, 2):
        assert e[2, col] / e[3, col] > rate / 4, (col, e[:, col])


def test_rhs_plants_are_caught():
    """A dropped surface measure or a real-valued plane wave gives errors far above the
    discretisation error: c^T b computed without the measure is not 4 pi."""
    S = O.Space(O.Geometry(), 3, 2)
    c = np.full(S.ndof, 2.0 ** -3)
    b = S.rhs(0, np.array([0.1, -0.2, 0.15]))
    assert abs(c @ b - 4 * math.pi) < 1e-10
    assert abs(0.5 * (c @ b) - 4 * math.pi) > 1  # a halved weight would be seen
    bw = S.rhs(1, np.array([0.0, 0.0, 1.0]), 4.0)
    assert abs((c @ bw).imag) < 1e-10  # symmetric sphere: the sine part integrates to zero
    assert abs((c @ bw).real - 4 * math.pi * math.sin(4.0) / 4.0) < 1e-9
    # a wrong wavenumber inside the plane wave (4.1 for 4) or a moved source point is seen
    bk = S.rhs(1, np.array([0.0, 0.0, 1.0]), 4.1)
    assert abs((c @ bk).real - 4 * math.pi * math.sin(4.0) / 4.0) > 1e-2
    bo = S.rhs(0, np.array([2.0, 0.5, -0.3]))
    assert abs(c @ bo - 4 * math.pi / np.linalg.norm([2.1, 0.5, -0.3])) > 1e-2


# ------------------------------------------------------------------ oracle GMRES
def test_oracle_gmres_dense_and_h2_agree_with_lu():
    S = O.Space(O.Geometry(), 2, 1)
    kappa = 3.0 + 0j
    A = S.dense_galerkin(O.HELMHOLTZ, kappa)
    b = S.rhs(1, np.array([0.0, 0.6, 0.8]), kappa.real)
    ref = np.linalg.solve(A, b)
    x = np.zeros(S.ndof, dtype=np.complex128)
    Af = np.ascontiguousarray(A.reshape(-1, order="F")).view(np.float64)  # column-major, interleaved
    it = O.lib().or_gmres_dense(S.ndof, O._dp(Af), O._dp(b.view(np.float64)), O._dp(x.view(np.float64)),
                                1e-12, 30, 5000)
    assert 0 < it < 5000
    assert np.linalg.norm(x - ref) <= 1e-9 * np.linalg.norm(ref)
    H = O.H2(S, O.HELMHOLTZ, kappa, 9)
    xh = np.zeros(S.ndof, dtype=np.complex128)
    it2 = O.lib().or_gmres_h2(H.h, O._dp(b.view(np.float64)), O._dp(xh.view(np.float64)), 1e-12, 30, 5000)
    assert 0 < it2 < 5000
    # the H2 operator differs from the dense one by the interpolation error (< 1e-8 here)
    assert np.linalg.norm(xh - ref) <= 1e-6 * np.linalg.norm(ref)
    assert np.linalg.norm(H.matvec(xh) - b) <= 1e-10 * np.linalg.norm(b)
