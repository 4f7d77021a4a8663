"""Pins of the oracle's single-layer potentials (Potential/DiscretePotential.hpp:49-103) and of
the example data functions used as Dirichlet traces (examples/Data.hpp).  CPU only.

* closed forms on the unit sphere: the potential of the constant density is 1 inside and 1/|x|
  outside (Laplace), e^{-ik} sin(k|x|)/(k|x|) inside and sin(k) e^{-ik|x|}/(k|x|) outside
  (Helmholtz); the surface integrals of the data functions;
* the reference's own end-to-end asserts, restated on dense Galerkin matrices and LU solves:
  examples/LaplaceSingleLayerH2.cpp:102 (interior harmonic data, rate >= 0.9 (2p+3) over the
  last two levels), examples/HelmholtzSingleLayerH2.cpp:108 (exterior fundamental solution,
  last three levels), examples/MaxwellSingleLayerH2.cpp:106 (exterior dipole, rate >= 0.9 (2p+1)).
"""
import math

import numpy as np
import pytest

import oracle_lib as O


def sphere_grid(r, n):
    """Util::makeSphereGrid (examples/Grids.hpp:36-49): golden-spiral points on a sphere"""
    gr = (1 + math.sqrt(5.0)) / 2.0
    out = []
    for i in range(n):
        phi = math.acos(1 - 2.0 * (i + 0.5) / n)
        th = 2 * math.pi * (i + 0.5) / gr
        out.append([r * math.cos(th) * math.sin(phi), r * math.sin(th) * math.sin(phi), r * math.cos(phi)])
    return np.array(out)


def cube_grid():
    """Util::makeTensorProductGrid of 10 x 10 x 10 points in [-0.25, 0.25]^3
    (examples/LaplaceSingleLayerH2.cpp:30-33)"""
    lin = np.linspace(-0.25, 0.25, 10)
    return np.array([[a, b, c] for a in lin for b in lin for c in lin])


def rate(errors):
    """estimateRateOfConvergence (examples/Error.hpp:102-109): least-squares slope of log2 errors"""
    e = np.log(np.abs(errors)) / math.log(2)
    A = np.vstack([np.ones(len(e)), np.arange(len(e))]).T
    return -np.linalg.lstsq(A, e, rcond=None)[0][1]


def harmonic(x):
    return 4 * x[:, 0] ** 2 - 3 * x[:, 1] ** 2 - x[:, 2] ** 2


def dipole(x, kappa, pos, length):
    """Data::Dipole (examples/Data.hpp:76-90), Hertz dipole with the exp(-i kappa r) convention"""
    c = x - pos
    r = np.linalg.norm(c)
    n = c / r
    e = np.exp(-1j * kappa * r)
    E = kappa ** 2 * e / r * np.cross(np.cross(n, length), n)
    return E + (1 / r ** 3 + 1j * kappa / r ** 2) * e * (3 * (n @ length) * n - length)


PTS_IN = np.array([[0.1, 0.2, -0.1], [0.0, 0.0, 0.0], [-0.3, 0.2, 0.4]])
PTS_OUT = np.array([[2.0, 0.5, -0.3], [0.0, 0.0, 1.5]])


def test_potential_of_constant_density_closed_forms():
    g = O.Geometry()
    k = 2.0
    r_in, r_out = np.linalg.norm(PTS_IN, axis=1), np.linalg.norm(PTS_OUT, axis=1)
    ex_in = np.exp(-1j * k) * np.where(r_in > 0, np.sin(k * r_in) / (k * np.maximum(r_in, 1e-300)), 1.0)
    ex_out = math.sin(k) * np.exp(-1j * k * r_out) / (k * r_out)
    prev = None
    for M in (1, 2, 3):
        S = O.Space(g, M, 2)
        rho = np.full(S.ndof, 2.0 ** -M)  # the constant density in the L2-scaled basis
        errs = [np.abs(S.potential(O.LAPLACE, rho, PTS_IN) - 1).max(),
                np.abs(S.potential(O.LAPLACE, rho, PTS_OUT) - 1 / r_out).max(),
                np.abs(S.potential(O.HELMHOLTZ, rho.astype(complex), PTS_IN, k + 0j) - ex_in).max(),
                np.abs(S.potential(O.HELMHOLTZ, rho.astype(complex), PTS_OUT, k + 0j) - ex_out).max()]
        if prev is not None:
            assert all(e < p / 50 for e, p in zip(errs, prev)), (errs, prev)
        prev = errs
    assert max(prev) < 1e-9
    # a planted mistake (density without the 1/h of the L2-scaled basis) is far off
    S = O.Space(g, 2, 2)
    assert abs(S.potential(O.LAPLACE, np.ones(S.ndof), PTS_IN)[0] - 1) > 1


def test_data_functions_surface_integrals():
    """int_S exp(-ik|x-p|)/|x-p| dS = 4 pi e^{-ik} sin(k|p|)/(k|p|) for |p| < 1;
    int_S (4x^2 - 3y^2 - z^2) dS = 0 on the unit sphere"""
    g = O.Geometry()
    S = O.Space(g, 3, 2)
    c = np.full(S.ndof, 2.0 ** -3)
    k, p = 3.0, np.array([0.2, -0.1, 0.3])
    rp = np.linalg.norm(p)
    ex = 4 * math.pi * np.exp(-1j * k) * math.sin(k * rp) / (k * rp)
    assert abs(c @ S.rhs(3, p, k) - ex) < 1e-9 * abs(ex)
    assert abs(c @ S.rhs(4, np.zeros(3))) < 1e-10
    # and 4 pi x^2 weights: int_S x^2 dS = 4 pi / 3 (checks the harmonic data is not identically 0)
    assert abs(c @ np.abs(S.rhs(4, np.zeros(3)))) > 1e-2


def test_laplace_example_convergence_rate():
    """examples/LaplaceSingleLayerH2.cpp: interior harmonic data, 10^3 grid in [-.25,.25]^3,
    rate of the last two levels >= 0.9 (2p+3)"""
    g = O.Geometry()
    grid = cube_grid()
    ref = harmonic(grid)
    for P in (1, 2):
        errs = []
        for M in (2, 3):
            S = O.Space(g, M, P)
            rho = np.linalg.solve(S.dense_galerkin(O.LAPLACE), S.rhs(4, np.zeros(3)))
            errs.append(np.abs(S.potential(O.LAPLACE, rho, grid) - ref).max())
        assert rate(errs) > 0.9 * (2 * P + 3), (P, errs, rate(errs))


def test_helmholtz_example_convergence_rate():
    """examples/HelmholtzSingleLayerH2.cpp: kappa 2, exterior data exp(-ikr)/r centred at 0,
    100 points on the sphere of radius 2, rate of the last three levels >= 0.9 (2p+3)"""
    g = O.Geometry()
    grid = sphere_grid(2.0, 100)
    k = 2.0
    ref = np.exp(-1j * k * np.linalg.norm(grid, axis=1)) / np.linalg.norm(grid, axis=1)
    for P in (1, 2):
        errs = []
        for M in (1, 2, 3):
            S = O.Space(g, M, P)
            rho = np.linalg.solve(S.dense_galerkin(O.HELMHOLTZ, k + 0j), S.rhs(3, np.zeros(3), k))
            errs.append(np.abs(S.potential(O.HELMHOLTZ, rho, grid, k + 0j) - ref).max())
        assert rate(errs) > 0.9 * (2 * P + 3), (P, errs, rate(errs))


@pytest.mark.parametrize("P", [1, 2])
def test_maxwell_example_convergence_rate(P):
    """examples/MaxwellSingleLayerH2.cpp: kappa 2, dipole at (.2,.2,.2) with moment (0,.1,.1),
    100 points on the sphere of radius 2, rate of the last three levels >= 0.9 (2p+1)"""
    g = O.Geometry()
    grid = sphere_grid(2.0, 100)
    k = 2.0
    pos, mom = np.array([0.2, 0.2, 0.2]), np.array([0.0, 0.1, 0.1])
    ref = np.array([dipole(x, k, pos, mom) for x in grid])
    errs = []
    for M in (1, 2, 3):
        S = O.Space(g, M, P, form=1)
        rho = np.linalg.solve(S.dense_galerkin(O.MAXWELL, k + 0j), S.rhs(2, np.concatenate([pos, mom]), k))
        errs.append(np.linalg.norm(S.potential(O.MAXWELL, rho, grid, k + 0j) - ref, axis=1).max())
    assert rate(errs) > 0.9 * (2 * P + 1), (P, errs, rate(errs))
