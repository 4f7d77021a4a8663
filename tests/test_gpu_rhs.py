"""Right-hand sides on the GPU (bmi_rhs_assemble) versus the oracle.

DiscreteLinearForm<DirichletTrace> (LinearForm/DiscreteLinearForm.hpp:44-83,
LinearForm/DirichletTrace.hpp:31-70) with point sources 1/|x - x0| and plane waves
exp(i kappa d.x), several columns per call, on the sphere and the ellipsoid: equal to the oracle
to 1e-12 (relative, per column).  At 24,576 elements the surface-integral closed forms hold.
"""
import math

import numpy as np
import pytest

import bembel_amd as B
import oracle_lib as O

pytestmark = pytest.mark.gpu
TOL = 1e-12
ELLIPSOID = (1.0, 0.7, 0.5)


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _unit(rng, n):
    d = rng.normal(size=(n, 3))
    return d / np.linalg.norm(d, axis=1, keepdims=True)


@pytest.mark.parametrize("scale", [None, ELLIPSOID])
@pytest.mark.parametrize("M,P", [(2, 1), (3, 2), (4, 2), (3, 1)])
def test_gpu_rhs_point_sources_and_plane_waves(gpu, scale, M, P):
    rng = np.random.default_rng(100 * M + P)
    og = O.Geometry(scale=scale or (1.0, 1.0, 1.0))
    os_ = O.Space(og, M, P)
    ps = B.AnsatzSpace(B.Geometry(scale=scale), M, P)
    # point sources inside and outside the surface
    x0 = np.vstack([0.3 * rng.uniform(-1, 1, (3, 3)), 2.0 + rng.uniform(0, 1, (2, 3))])
    f = B.DiscreteLinearForm(B.DirichletTrace(x0), ps).compute().get_discrete_linear_form()
    assert f.shape == (ps.number_of_dofs, len(x0))
    for k in range(len(x0)):
        assert rel(f[:, k], os_.rhs(0, x0[k])) <= TOL, ("point source", k)
    d = _unit(rng, 5)
    for kappa in (1.0, 4.0):
        g = B.DiscreteLinearForm(B.PlaneWaves(d, kappa), ps).compute().get_discrete_linear_form()
        for k in range(len(d)):
            assert rel(g[:, k], os_.rhs(1, d[k], kappa)) <= TOL, ("plane wave", kappa, k)
    # device output equals host output bit for bit
    gd = B.DiscreteLinearForm(B.PlaneWaves(d, 4.0), ps).compute(device=True).get_discrete_linear_form()
    assert np.array_equal(gd.cpu().numpy(), g)


def test_gpu_rhs_full_size_closed_forms(gpu):
    """M = 6 (24,576 elements), P = 2: c^T b equals the surface integral of the data on the unit
    sphere (4 pi inside, 4 pi/|x0| outside, 4 pi sin(k)/k for a plane wave); c = 2^-6 is the
    constant function in the L2-scaled basis."""
    ps = B.AnsatzSpace(B.Geometry(), 6, 2)
    c = np.full(ps.number_of_dofs, 2.0 ** -6)
    x_in, x_out = np.array([0.1, -0.2, 0.15]), np.array([2.0, 0.5, -0.3])
    f = B.DiscreteLinearForm(B.DirichletTrace(np.vstack([x_in, x_out])), ps).compute().get_discrete_linear_form()
    assert abs(c @ f[:, 0] - 4 * math.pi) <= 1e-12 * 4 * math.pi
    assert abs(c @ f[:, 1] - 4 * math.pi / np.linalg.norm(x_out)) <= 1e-12 * 4 * math.pi
    k = 4.0
    g = B.DiscreteLinearForm(B.PlaneWaves(_unit(np.random.default_rng(1), 64), k), ps).compute().get_discrete_linear_form()
    ex = 4 * math.pi * math.sin(k) / k
    assert np.abs(c @ g - ex).max() <= 1e-11 * abs(ex)


@pytest.mark.parametrize("M,P", [(2, 1), (3, 2)])
def test_gpu_rhs_from_caller_values(gpu, M, P):
    """DirichletTrace / RotatedTangentialTrace with an arbitrary function (evaluated by the
    caller at bmi_linear_form_points) equal the built-in device data functions"""
    ps = B.AnsatzSpace(B.Geometry(), M, P)
    x0 = np.array([0.2, -0.1, 0.05])
    f = B.DiscreteLinearForm(B.FunctionTrace(lambda X: 1.0 / np.linalg.norm(X - x0, axis=1)), ps).compute()
    g = B.DiscreteLinearForm(B.DirichletTrace(x0), ps).compute()
    assert rel(f.get_discrete_linear_form(), g.get_discrete_linear_form()) <= TOL
    d, k = np.array([0.0, 0.6, 0.8]), 3.0
    f = B.DiscreteLinearForm(B.FunctionTrace(lambda X: np.exp(1j * k * (X @ d)), np.complex128), ps).compute()
    g = B.DiscreteLinearForm(B.PlaneWaves(d, k), ps).compute()
    assert rel(f.get_discrete_linear_form(), g.get_discrete_linear_form()) <= TOL
    op = B.MaxwellSingleLayerOperator()
    pm = B.AnsatzSpace(B.Geometry(scale=ELLIPSOID), M, P, op)
    pos, mom = np.array([0.1, -0.05, 0.08]), np.array([0.02, 0.05, -0.07])
    from test_oracle import _dipole_np
    f = B.DiscreteLinearForm(B.FunctionTrace(lambda X: np.array([_dipole_np(x, 2.0, pos, mom) for x in X]),
                                             vector=True), pm).compute()
    g = B.DiscreteLinearForm(B.RotatedTangentialTrace(pos, mom, 2.0), pm).compute()
    assert rel(f.get_discrete_linear_form(), g.get_discrete_linear_form()) <= TOL
