"""The allocation-free product bmi_h2_matvec_noalloc (workspace in the library's per-GPU scratch
array) against bmi_h2_matvec through the C ABI: two calls in a row on each handle, the second
with fewer columns, then a smaller problem on another handle; FP64 handles within 1e-12
(relative, per column), the reduced operator of a mixed handle against bmi_h2_matvec of the same
handle within 1e-12; a product above the scratch size is refused before any device work."""
import ctypes as C

import numpy as np
import pytest

import bembel_amd as B

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _dp(t):
    return C.cast(C.c_void_p(t.data_ptr()), C.POINTER(C.c_double))


def _products(H, X, Y0, alpha, beta):
    """(reference, allocation-free) results of Y = alpha H X + beta Y0, column-major torch tensors
    holding (n, nrhs) with complex entries as interleaved doubles"""
    import torch
    lib = B.library()
    n = H.rows()
    nrhs = X.shape[0]
    al = np.array([alpha.real, alpha.imag])
    be = np.array([beta.real, beta.imag])
    Yr, Yn = Y0.clone(), Y0.clone()
    torch.cuda.synchronize()
    st = lib.bmi_h2_matvec(H._h, nrhs, _dp(X), n, _dp(Yr), n, B._dp(al), B._dp(be), B.MEM_DEVICE, None)
    assert st == B.OK, lib.bmi_last_error()
    B.synchronize()
    st = lib.bmi_h2_matvec_noalloc(H._h, nrhs, _dp(X), n, _dp(Yn), n, B._dp(al), B._dp(be), None)
    assert st == B.OK, lib.bmi_last_error()
    B.synchronize()
    return Yr.cpu().numpy(), Yn.cpu().numpy()


def _check(H, nrhs, seed, noalloc_first=False):
    import torch
    n = H.rows()
    S = 2 if H.scalar == np.complex128 else 1
    g = torch.Generator(device="cpu").manual_seed(seed)
    X = torch.rand(nrhs, S * n, generator=g, dtype=torch.float64).cuda() - 0.5
    Y0 = torch.rand(nrhs, S * n, generator=g, dtype=torch.float64).cuda() - 0.5
    alpha, beta = complex(0.75, 0.25 if S == 2 else 0.0), complex(-1.5, 0.0)
    if noalloc_first:
        # a handle that has not run bmi_h2_matvec yet: the untimed default variants
        lib = B.library()
        al, be = np.array([alpha.real, alpha.imag]), np.array([beta.real, beta.imag])
        Yn = Y0.clone()
        torch.cuda.synchronize()
        assert lib.bmi_h2_matvec_noalloc(H._h, nrhs, _dp(X), n, _dp(Yn), n, B._dp(al), B._dp(be), None) == B.OK
        B.synchronize()
        first = Yn.cpu().numpy()
    ref, got = _products(H, X, Y0, alpha, beta)
    for k in range(nrhs):
        err = np.linalg.norm(got[k] - ref[k]) / np.linalg.norm(ref[k])
        assert err <= TOL, (k, err)
        if noalloc_first:
            # before the first timed product a mixed handle may use another reduced-operator form
            # (FP32 coupling recomputed on the fly instead of the FP32 copy): both are within 1e-8
            # of the FP64 operator
            err0 = np.linalg.norm(first[k] - ref[k]) / np.linalg.norm(ref[k])
            assert err0 <= (TOL if H.precision == B.FP64 else 1e-7), (k, err0)
    print("nrhs", nrhs, "max rel", max(np.linalg.norm(got[k] - ref[k]) / np.linalg.norm(ref[k]) for k in range(nrhs)))


def _handle(kind, M, P, np_, precision=B.FP64, kappa=4.0):
    ps = B.AnsatzSpace(B.Geometry(), M, P)
    op = B.LaplaceSingleLayerOperator() if kind == 0 else B.HelmholtzSingleLayerOperator()
    if kind == 1:
        op.set_wavenumber(kappa)
    return ps, B.H2Matrix(ps, op, n_cheb=np_, precision=precision)


@pytest.mark.parametrize("kind,M,P,np_,precision", [(0, 4, 2, 4, B.FP64), (1, 4, 2, 5, B.FP64),
                                                     (0, 4, 2, 4, B.MIXED), (1, 3, 2, 5, B.MIXED)])
def test_gpu_noalloc_matches_matvec(gpu, kind, M, P, np_, precision):
    ps, H = _handle(kind, M, P, np_, precision)
    _check(H, 3, 1, noalloc_first=True)  # untimed variants, then after bmi_h2_matvec
    _check(H, 1, 2)  # second call, fewer columns
    _check(H, 12, 3)  # many-RHS layout
    _check(H, 2, 4)
    # a smaller problem on another handle, right after: the scratch keeps the last call's data
    ps2, H2 = _handle(kind, M - 1, P, np_, precision)
    _check(H2, 1, 5)


def test_gpu_noalloc_refuses_products_above_the_scratch(gpu):
    import torch
    ps, H = _handle(0, 4, 2, 4)
    n = H.rows()
    lib = B.library()
    nrhs = 400  # 3 x 13,824 x 8 B x 400 > 64 MiB
    X = torch.zeros(nrhs, n, dtype=torch.float64, device="cuda")
    one, zero = np.array([1.0, 0.0]), np.zeros(2)
    st = lib.bmi_h2_matvec_noalloc(H._h, nrhs, _dp(X), n, _dp(X), n, B._dp(one), B._dp(zero), None)
    assert st == B.INVALID_ARG
    assert b"scratch" in lib.bmi_last_error()
