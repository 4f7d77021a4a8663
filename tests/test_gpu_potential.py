"""Single-layer potentials and the example data on the GPU.

* bmi_potential_evaluate (DiscretePotential, Potential/DiscretePotential.hpp:49-103) equals the
  oracle to 1e-12 for the Laplace, Helmholtz (real and damped wavenumbers) and Maxwell potentials;
* the Dirichlet traces of the example data (fundamental solution, harmonic polynomial) equal
  the oracle to 1e-12;
* the reference's end-to-end asserts run through the device path (H2 assembly with the reference
  defaults: 9 Chebyshev nodes, eta 1.6; device CG/GMRES; device potential):
  examples/LaplaceSingleLayerH2.cpp:102, HelmholtzSingleLayerH2.cpp:108, MaxwellSingleLayerH2.cpp:106;
* at 24,576 elements the potential of the constant density is 1 inside the sphere and 1/|x|
  outside.
"""
import numpy as np
import pytest

import bembel_amd as B
import oracle_lib as O
from test_potential import PTS_IN, PTS_OUT, cube_grid, dipole, harmonic, rate, sphere_grid

pytestmark = pytest.mark.gpu
TOL = 1e-12
ELLIPSOID = (1.0, 0.7, 0.5)


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _points(rng):
    return np.vstack([PTS_IN, PTS_OUT, 0.4 * rng.uniform(-1, 1, (40, 3)), sphere_grid(1.7, 60)])


@pytest.mark.parametrize("M,P", [(2, 1), (3, 2)])
@pytest.mark.parametrize("kind,kappa", [(O.LAPLACE, 0j), (O.HELMHOLTZ, 2 + 0j), (O.HELMHOLTZ, 2 + 0.5j)])
def test_gpu_potential_matches_oracle(gpu, kind, kappa, M, P):
    rng = np.random.default_rng(10 * M + P)
    os_ = O.Space(O.Geometry(), M, P)
    ps = B.AnsatzSpace(B.Geometry(), M, P)
    n = ps.number_of_dofs
    rho = rng.uniform(-1, 1, n) if kind == O.LAPLACE else rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    pts = _points(rng)
    pot = B.LaplaceSingleLayerPotential() if kind == O.LAPLACE else B.HelmholtzSingleLayerPotential()
    pot.set_wavenumber(kappa)
    dp = B.DiscretePotential(pot, ps)
    dp.set_cauchy_data(rho)
    got = dp.evaluate(pts)
    ref = os_.potential(kind, rho, pts, kappa)
    print("rel", rel(got, ref))
    assert rel(got, ref) <= TOL


@pytest.mark.parametrize("M,P", [(2, 1), (2, 2)])
def test_gpu_maxwell_potential_matches_oracle(gpu, M, P):
    rng = np.random.default_rng(7 + M + P)
    os_ = O.Space(O.Geometry(scale=ELLIPSOID), M, P, form=1)
    ps = B.AnsatzSpace(B.Geometry(scale=ELLIPSOID), M, P, B.MaxwellSingleLayerOperator())
    n = ps.number_of_dofs
    rho = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    pts = _points(rng)
    pot = B.MaxwellSingleLayerPotential()
    pot.set_wavenumber(2.0)
    dp = B.DiscretePotential(pot, ps)
    dp.set_cauchy_data(rho)
    got = dp.evaluate(pts)
    assert got.shape == (len(pts), 3)
    ref = os_.potential(O.MAXWELL, rho, pts, 2 + 0j)
    print("rel", rel(got, ref))
    assert rel(got, ref) <= TOL


@pytest.mark.parametrize("M,P", [(2, 1), (3, 2)])
def test_gpu_example_data_traces_match_oracle(gpu, M, P):
    os_ = O.Space(O.Geometry(), M, P)
    ps = B.AnsatzSpace(B.Geometry(), M, P)
    c = np.array([[0.1, -0.2, 0.15], [0.0, 0.0, 0.0]])
    f = B.DiscreteLinearForm(B.HelmholtzFundamentalSolution(c, 2.0), ps).compute().get_discrete_linear_form()
    for k in range(2):
        assert rel(f[:, k], os_.rhs(3, c[k], 2.0)) <= TOL
    g = B.DiscreteLinearForm(B.HarmonicFunction(), ps).compute().get_discrete_linear_form()
    assert rel(g, os_.rhs(4, np.zeros(3))) <= TOL


def _solve(H, b, kind):
    if kind == O.LAPLACE:
        s = B.ConjugateGradient()
    else:
        s = B.GMRES()
        s.set_restart(2000)
    s.set_tolerance(1e-12)
    s.set_max_iterations(5000)
    s.compute(H)
    x = s.solve(b)
    assert s.info() == B.OK
    return x


def test_gpu_laplace_example_convergence(gpu):
    """examples/LaplaceSingleLayerH2.cpp:102 through the device path"""
    grid = cube_grid()
    ref = harmonic(grid)
    for P in (1, 2):
        errs = []
        for M in (2, 3):
            ps = B.AnsatzSpace(B.Geometry(), M, P)
            op = B.LaplaceSingleLayerOperator()
            H = B.DiscreteOperator(B.H2Matrix, ps, op).compute().get_discrete_operator()
            b = B.DiscreteLinearForm(B.HarmonicFunction(), ps).compute().get_discrete_linear_form()
            rho = _solve(H, b, O.LAPLACE)
            dp = B.DiscretePotential(B.LaplaceSingleLayerPotential(), ps)
            dp.set_cauchy_data(rho)
            errs.append(np.abs(dp.evaluate(grid) - ref).max())
        print("laplace P", P, errs, rate(errs))
        assert rate(errs) > 0.9 * (2 * P + 3)


def test_gpu_helmholtz_example_convergence(gpu):
    """examples/HelmholtzSingleLayerH2.cpp:108 through the device path"""
    grid = sphere_grid(2.0, 100)
    k = 2.0
    r = np.linalg.norm(grid, axis=1)
    ref = np.exp(-1j * k * r) / r
    for P in (1, 2):
        errs = []
        for M in (1, 2, 3):
            ps = B.AnsatzSpace(B.Geometry(), M, P)
            op = B.HelmholtzSingleLayerOperator()
            op.set_wavenumber(k)
            H = B.DiscreteOperator(B.H2Matrix, ps, op).compute().get_discrete_operator()
            b = B.DiscreteLinearForm(B.HelmholtzFundamentalSolution([0.0, 0.0, 0.0], k), ps).compute(
            ).get_discrete_linear_form()
            rho = _solve(H, b, O.HELMHOLTZ)
            pot = B.HelmholtzSingleLayerPotential()
            pot.set_wavenumber(k)
            dp = B.DiscretePotential(pot, ps)
            dp.set_cauchy_data(rho)
            errs.append(np.abs(dp.evaluate(grid) - ref).max())
        print("helmholtz P", P, errs, rate(errs))
        assert rate(errs) > 0.9 * (2 * P + 3)


def test_gpu_maxwell_example_convergence(gpu):
    """examples/MaxwellSingleLayerH2.cpp:106 through the device path"""
    grid = sphere_grid(2.0, 100)
    k = 2.0
    pos, mom = np.array([0.2, 0.2, 0.2]), np.array([0.0, 0.1, 0.1])
    ref = np.array([dipole(x, k, pos, mom) for x in grid])
    for P in (1, 2):
        errs = []
        for M in (1, 2, 3):
            op = B.MaxwellSingleLayerOperator()
            op.set_wavenumber(k)
            ps = B.AnsatzSpace(B.Geometry(), M, P, op)
            H = B.DiscreteOperator(B.H2Matrix, ps, op).compute().get_discrete_operator()
            b = B.DiscreteLinearForm(B.RotatedTangentialTrace(pos, mom, k), ps).compute().get_discrete_linear_form()
            rho = _solve(H, b, O.MAXWELL)
            pot = B.MaxwellSingleLayerPotential()
            pot.set_wavenumber(k)
            dp = B.DiscretePotential(pot, ps)
            dp.set_cauchy_data(rho)
            errs.append(np.linalg.norm(dp.evaluate(grid) - ref, axis=1).max())
        print("maxwell P", P, errs, rate(errs))
        assert rate(errs) > 0.9 * (2 * P + 1)


def test_gpu_potential_full_size_closed_form(gpu):
    """M = 6 (24,576 elements, 393,216 sources), P = 2: constant density"""
    ps = B.AnsatzSpace(B.Geometry(), 6, 2)
    dp = B.DiscretePotential(B.LaplaceSingleLayerPotential(), ps)
    dp.set_cauchy_data(np.full(ps.number_of_dofs, 2.0 ** -6))
    u = dp.evaluate(np.vstack([PTS_IN, PTS_OUT]))
    ex = np.concatenate([np.ones(len(PTS_IN)), 1 / np.linalg.norm(PTS_OUT, axis=1)])
    assert np.abs(u - ex).max() <= 1e-12
