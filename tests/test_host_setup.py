"""Host side of the product versus the oracle, and the C ABI surface.  CPU only: none of these
entry points touches the GPU.

The mesh, enclosing balls, transformation matrix, pair classification (compareElements
rotations and quadrature degrees) and block-cluster tree decide every later result, so they are
required to agree with the oracle exactly.
"""
import ctypes
import os
import re

import numpy as np
import pytest

import bembel_amd as B
import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SCREEN = os.path.join(ROOT, "tests", "golden", "screen.dat")


def canon(v):
    m = {}
    out = np.empty_like(v)
    flat = out.reshape(-1)
    for i, x in enumerate(v.reshape(-1)):
        flat[i] = m.setdefault(int(x), len(m))
    return out


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "bembel_amd.h")).read()
    names = sorted(set(re.findall(r"\b(bmi_[a-z0-9_]+)\s*\(", hdr)))
    assert len(names) >= 25
    lib = ctypes.CDLL(B.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_library_built_from_these_sources():
    """the loaded library reports the digest of the sources it was built from; it must be the
    digest of the sources in this tree (a stale build fails here, not silently on the GPU)"""
    info = B.build_info()
    assert info["match"], info


def test_error_reporting_without_gpu():
    with pytest.raises(B.BembelError) as e:
        B.Geometry("/nonexistent/file.dat")
    assert e.value.status == B.IO_ERROR
    g = B.Geometry()
    with pytest.raises(B.BembelError) as e:
        B.AnsatzSpace(g, -1, 1)
    assert e.value.status == B.INVALID_ARG


def test_invalid_arguments_return_status_without_gpu():
    """Null handles and out-of-range arguments come back as BMI_INVALID_ARG before any device
    call (no abort, no crash), for the product, sweep control and task entry points."""
    lib = B.library()
    dp = ctypes.POINTER(ctypes.c_double)
    one = (ctypes.c_double * 2)(1.0, 0.0)
    x = (ctypes.c_double * 4)()
    assert lib.bmi_h2_matvec(None, 1, x, 1, x, 1, one, None, B.MEM_HOST, None) == B.INVALID_ARG
    assert lib.bmi_h2_set_sweep(None, 0) == B.INVALID_ARG
    assert lib.bmi_h2_set_overlap(None, 0) == B.INVALID_ARG
    assert lib.bmi_h2_use_graphs(None, 1) == B.INVALID_ARG
    t = ctypes.c_void_p()
    assert lib.bmi_gmres_start(None, 1, x, x, 1e-8, 30, 10, 1, None, ctypes.byref(t)) == B.INVALID_ARG
    assert lib.bmi_gmres_poll(None, None) == B.INVALID_ARG
    assert lib.bmi_gmres_finish(None, None, None) == B.INVALID_ARG
    assert lib.bmi_stream_destroy(None) == B.OK
    assert lib.bmi_stream_create(None) == B.INVALID_ARG
    assert lib.bmi_solve_gmres(None, 1, ctypes.cast(x, dp), ctypes.cast(x, dp), 1e-8, 30, 10, 1, None, None,
                               B.MEM_HOST, None) == B.INVALID_ARG
    assert lib.bmi_last_error()


CASES = [(None, 0, 1, 0), (None, 2, 2, 0), (None, 4, 2, 0), ((1.0, 0.7, 0.5), 3, 2, 0), ((1.0, 0.7, 0.5), 3, 2, 1),
         (None, 3, 1, 1), ("screen", 2, 0, 0), (None, 6, 2, 0)]


@pytest.mark.parametrize("scale,M,P,form", CASES)
def test_space_matches_oracle(scale, M, P, form):
    path = SCREEN if scale == "screen" else B.SPHERE_DAT
    sc = None if scale in (None, "screen") else scale
    og = O.Geometry(path, sc or (1.0, 1.0, 1.0))
    os_ = O.Space(og, M, P, form)
    pg = B.Geometry(path, sc)
    op = B.MaxwellSingleLayerOperator() if form else B.LaplaceSingleLayerOperator()
    ps = B.AnsatzSpace(pg, M, P, op)
    assert (ps.number_of_elements, ps.number_of_dofs, ps.number_of_super_dofs) == (os_.nel, os_.ndof, os_.nsuper)
    oe, pe = os_.elements(), ps.elements()
    assert np.array_equal(canon(oe["vert"]), canon(pe["vert"]))
    assert np.array_equal(oe["patch"], pe["patch"])
    assert np.array_equal(oe["llc"], pe["llc"])
    assert np.array_equal(oe["mid"], pe["mid"]) and np.array_equal(oe["radius"], pe["radius"])
    for lvl in range(M + 1):
        om, orad = os_.clusters(lvl)
        pm, prad = ps.clusters(lvl)
        assert np.array_equal(om, pm) and np.array_equal(orad, prad)
    orp, oc, ov = os_.T()
    prp, pc, pv = ps.get_transformation_matrix()
    assert np.array_equal(orp, prp) and np.array_equal(oc, pc)
    assert np.abs(ov - pv).max() <= 1e-14


@pytest.mark.parametrize("M,P", [(2, 1), (3, 2)])
def test_pair_classification_matches_oracle(M, P):
    og = O.Geometry()
    os_ = O.Space(og, M, P)
    ps = B.AnsatzSpace(B.Geometry(), M, P)
    n = ps.number_of_elements
    rng = np.random.default_rng(M)
    pairs = np.array([(i, j) for i in range(n) for j in range(n)], dtype=np.int32)
    if len(pairs) > 40000:
        pairs = pairs[rng.choice(len(pairs), 40000, replace=False)]
        # keep every touching pair of element 0..63
        extra = np.array([(i, j) for i in range(64) for j in range(n)], dtype=np.int32)
        pairs = np.vstack([pairs, extra])
    got = ps.pair_info(pairs)
    for k, (i, j) in enumerate(pairs):
        want, _ = os_.pair_info(int(i), int(j))
        assert list(got[k]) == list(want), (i, j, got[k], want)


def test_linear_form_points_on_the_surface():
    """bmi_linear_form_points (host code): (P+2)^2 points per element on the unit sphere, in the
    element order and local order of the oracle's right-hand-side rule"""
    for M, P in [(1, 1), (2, 2)]:
        ps = B.AnsatzSpace(B.Geometry(), M, P)
        n = ctypes.c_int64()
        assert B.library().bmi_linear_form_points(ps._h, ctypes.byref(n), None) == B.OK
        assert n.value == ps.number_of_elements * (P + 2) ** 2
        pts = np.zeros((n.value, 3))
        B.library().bmi_linear_form_points(ps._h, ctypes.byref(n), B._dp(pts))
        assert np.abs(np.linalg.norm(pts, axis=1) - 1).max() < 1e-14
        # first element, first point: the oracle's map2surface at the first Gauss node pair
        og = O.Geometry()
        osp = O.Space(og, M, P)
        el = osp.elements()
        t, _ = np.polynomial.legendre.leggauss(P + 2)
        t = (t + 1) / 2
        h = 2.0 ** -M
        for e in (0, ps.number_of_elements - 1):
            for q in (0, (P + 2) ** 2 - 1):
                st = el["llc"][e] + h * np.array([t[q // (P + 2)], t[q % (P + 2)]])
                ref = og.surface_point(int(el["patch"][e]), st)[3:6]
                assert np.allclose(pts[e * (P + 2) ** 2 + q], ref, atol=1e-14)
