"""GPU near-field quadrature versus the oracle (parity bar: 1e-12 relative, FP64).

Covers every pair class of evaluateBilinearForm (DuffyTrick/evaluateBilinearForm.hpp:21-71):
regular pairs on the far-field table (integrate0) and on the tensor rule (integrate1), identical
(integrate2), edge-adjacent (integrate3) and vertex-adjacent (integrate4) pairs, for the Laplace
and Helmholtz single layers, plus the reference's own Duffy known-answer tests on the GPU.
"""
import math

import numpy as np
import pytest

import bembel_amd as B
import oracle_lib as O
from test_oracle import DUFFY_TOL, SCREEN, duffy_cases

pytestmark = pytest.mark.gpu
TOL = 1e-12


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_gpu_duffy_known_answers(gpu):
    """tests/test_DuffyTrick.cpp on the GPU: integrate1..4 with the reference's test functions,
    rule with 17 points, level-2 screen; GPU equals the oracle to 1e-12 and both meet the
    reference tolerances against the closed forms."""
    os_ = O.Space(O.Geometry(SCREEN), 2, 0)
    ps = B.AnsatzSpace(B.Geometry(SCREEN), 2, 0, B.DummyOperator())
    cases = duffy_cases(os_)
    for fn in (O.DUMMY_ANALYTIC, O.DUMMY_NEARSINGULAR):
        sel = [c for c in cases if c[6] == fn]
        jobs = np.array([c[:5] for c in sel], dtype=np.int32)
        got = ps.integrate_jobs(B.DummyOperator(fn), jobs, 16)
        worst = {}
        for k, (which, e1, r1, e2, r2, ex, _) in enumerate(sel):
            want = os_.integrate(O.DUMMY, which, e1, r1, e2, r2, 16, dummy_fn=fn)[0, 0]
            assert abs(got[k] - want) <= TOL * abs(want), (which, e1, e2, got[k], want)
            worst[which] = max(worst.get(which, 0.0), abs(got[k] - ex) / abs(ex))
        for which, err in worst.items():
            assert err < DUFFY_TOL[which], (which, err)


def sample_pairs(ps, n_per_class=12, seed=0):
    """element pairs covering every quadrature class of the space"""
    rng = np.random.default_rng(seed)
    n = ps.number_of_elements
    cand = np.array([(i, j) for i in rng.choice(n, min(n, 48), replace=False) for j in range(n)], dtype=np.int32)
    info = ps.pair_info(cand)
    ff = ps.polynomial_degree + 2
    keys = np.where(info[:, 0] == 0, np.where(info[:, 3] == ff, 0, 1), info[:, 0])
    out = []
    for c in range(5):
        idx = np.nonzero(keys == c)[0]
        if len(idx):
            out.append(cand[rng.choice(idx, min(n_per_class, len(idx)), replace=False)])
    return np.vstack(out), keys


@pytest.mark.parametrize("kind,kappa,M,P", [(O.LAPLACE, 0j, 3, 1), (O.LAPLACE, 0j, 3, 2), (O.HELMHOLTZ, 4 + 0j, 3, 2),
                                            (O.HELMHOLTZ, 2 + 0.5j, 2, 1)])
def test_gpu_pair_blocks_match_oracle(gpu, kind, kappa, M, P):
    og = O.Geometry()
    os_ = O.Space(og, M, P)
    ps = B.AnsatzSpace(B.Geometry(), M, P)
    op = B.LaplaceSingleLayerOperator() if kind == O.LAPLACE else B.HelmholtzSingleLayerOperator()
    op.set_wavenumber(kappa)
    pairs, _ = sample_pairs(ps)
    got = ps.pair_blocks(op, pairs)
    classes = set()
    for k, (i, j) in enumerate(pairs):
        want = os_.pair_block(kind, int(i), int(j), kappa=kappa)
        info, _ = os_.pair_info(int(i), int(j))
        classes.add(info[0])
        r = rel(got[k], want)
        print("pair", i, j, "class", info[0], "deg", info[3], "rel", r)
        assert r <= TOL, (i, j, info, r)
    assert {0, 2, 3, 4} <= classes


def test_gpu_dense_c1(gpu):
    """Dense Laplace Galerkin matrix of the 384-element sphere (P = 1): T^T A T equals the oracle
    to 1e-12, is symmetric and positive definite."""
    og = O.Geometry()
    os_ = O.Space(og, 3, 1)
    ps = B.AnsatzSpace(B.Geometry(), 3, 1)
    A_gpu = B.DiscreteOperator(B.DenseMatrix, ps, B.LaplaceSingleLayerOperator()).compute().get_discrete_operator().matrix
    A_ref = os_.dense_galerkin(O.LAPLACE)
    assert A_gpu.shape == (486, 486)
    assert rel(A_gpu, A_ref) <= TOL
    assert np.abs(A_gpu - A_ref).max() <= TOL * np.abs(A_ref).max()
    assert np.abs(A_gpu - A_gpu.T).max() <= 1e-12 * np.abs(A_gpu).max()
    c = np.ones(486) * 2.0 ** -3
    assert abs(c @ A_gpu @ c - 4 * math.pi) / (4 * math.pi) < 1e-6
