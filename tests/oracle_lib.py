"""ctypes access to the CPU oracle (oracle/liboracle_bembel.so) for the tests.

The oracle is test infrastructure: it restates the reference algorithm on the CPU and is used
only as the checker.  Nothing in the product package imports this module.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle_bembel.so")
SPHERE = os.path.join(ROOT, "bembel_amd", "geo", "sphere.dat")

LAPLACE, HELMHOLTZ, MAXWELL, DUMMY = 0, 1, 2, 3
DUMMY_ANALYTIC, DUMMY_NEARSINGULAR, DUMMY_ONE = 0, 1, 2

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library missing: run __graft_entry__.build() (make -C oracle)")
        L = C.CDLL(LIB_PATH)
        vp, ip, dp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)
        L.or_geometry_load.restype = vp
        L.or_geometry_load.argtypes = [C.c_char_p, dp]
        L.or_geometry_free.argtypes = [vp]
        L.or_patch_eval.argtypes = [vp, C.c_int, C.c_double, C.c_double, dp]
        L.or_patch_surface_point.argtypes = [vp, C.c_int, dp, C.c_double, dp, dp]
        L.or_bernstein.argtypes = [C.c_int, C.c_double, dp]
        L.or_bernstein_der.argtypes = [C.c_int, C.c_double, dp]
        L.or_space_create.restype = vp
        L.or_space_create.argtypes = [vp, C.c_int, C.c_int, C.c_int]
        L.or_space_free.argtypes = [vp]
        L.or_space_info.argtypes = [vp, ip]
        L.or_space_elements.argtypes = [vp, ip, ip, dp, dp, dp]
        L.or_space_clusters.restype = C.c_int
        L.or_space_clusters.argtypes = [vp, C.c_int, dp, dp]
        L.or_space_T.argtypes = [vp, ip, ip, dp]
        L.or_pair_info.argtypes = [vp, C.c_int, C.c_int, ip, dp]
        L.or_nfield_degree.restype = C.c_int
        L.or_nfield_degree.argtypes = [C.c_int, C.c_double, C.c_int]
        L.or_pair_block.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, dp]
        L.or_integrate.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                   C.c_int, C.c_int, C.c_int, C.c_int, dp]
        L.or_dense_super.argtypes = [vp, C.c_int, C.c_double, C.c_double, dp]
        L.or_dense_galerkin.argtypes = [vp, C.c_int, C.c_double, C.c_double, dp]
        L.or_cheb_points.argtypes = [C.c_int, dp]
        L.or_transfer_matrix.argtypes = [C.c_int, dp]
        L.or_moment_matrix.restype = C.c_int
        L.or_moment_matrix.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, dp]
        L.or_forward_transform.argtypes = [C.c_int, C.c_int, dp, C.c_int, dp, C.c_int, dp, C.c_int,
                                           C.POINTER(dp)]
        L.or_h2_build.restype = vp
        L.or_h2_build.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_int, C.c_double, C.c_int,
                                  C.c_int]
        L.or_h2_free.argtypes = [vp]
        L.or_h2_info.argtypes = [vp, ip]
        L.or_h2_leaves.argtypes = [vp, ip]
        L.or_h2_leaf_matrix.argtypes = [vp, C.c_int, C.c_int, C.c_int, dp]
        L.or_h2_set_leaf_matrix.argtypes = [vp, C.c_int, C.c_int, C.c_int, dp]
        L.or_h2_assemble_leaves.argtypes = [vp, C.c_int, ip]
        L.or_h2_import_leaves.restype = C.c_int
        L.or_h2_import_leaves.argtypes = [vp, C.c_int, ip, dp]
        L.or_h2_matvec.argtypes = [vp, C.c_int, dp, dp]
        L.or_h2_matvec_super.argtypes = [vp, C.c_int, dp, dp]
        L.or_rhs.argtypes = [vp, C.c_int, dp, C.c_double, dp]
        L.or_potential.argtypes = [vp, C.c_int, C.c_double, C.c_double, dp, C.c_int, dp, dp]
        L.or_cg.restype = C.c_int
        L.or_cg.argtypes = [C.c_int, dp, dp, dp, C.c_double, C.c_int]
        L.or_gmres_dense.restype = C.c_int
        L.or_gmres_dense.argtypes = [C.c_int, dp, dp, dp, C.c_double, C.c_int, C.c_int]
        L.or_gmres_h2.restype = C.c_int
        L.or_gmres_h2.argtypes = [vp, dp, dp, C.c_double, C.c_int, C.c_int]
        L.or_time_assembly_sample.restype = C.c_double
        L.or_time_assembly_sample.argtypes = [vp, C.c_int, ip, C.c_int, ip, C.c_int]
        L.or_set_num_threads.restype = None
        L.or_set_num_threads.argtypes = [C.c_int]
        L.or_time_matvec.restype = C.c_double
        L.or_time_matvec.argtypes = [vp, C.c_int, C.c_int, C.c_int]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class Geometry:
    def __init__(self, path=SPHERE, scale=(1.0, 1.0, 1.0)):
        sc = np.asarray(scale, dtype=np.float64)
        self.h = lib().or_geometry_load(path.encode(), _dp(sc))
        if not self.h:
            raise RuntimeError("oracle: cannot load geometry " + path)
        self._keep = sc

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_geometry_free(self.h)
            self.h = None

    def eval(self, patch, s, t):
        out = np.zeros(3)
        lib().or_patch_eval(self.h, patch, s, t, _dp(out))
        return out

    def surface_point(self, patch, st, w=1.0, xi=(0.0, 0.0)):
        out = np.zeros(12)
        st = np.asarray(st, dtype=np.float64)
        xi = np.asarray(xi, dtype=np.float64)
        lib().or_patch_surface_point(self.h, patch, _dp(st), w, _dp(xi), _dp(out))
        return out


def bernstein(p, x, der=False):
    out = np.zeros(p + 1)
    (lib().or_bernstein_der if der else lib().or_bernstein)(p, x, _dp(out))
    return out


class Space:
    def __init__(self, geo, M, P, form=0):
        self.geo = geo
        self.h = lib().or_space_create(geo.h, M, P, form)
        if not self.h:
            raise RuntimeError("oracle: space creation failed")
        info = np.zeros(8, dtype=np.int32)
        lib().or_space_info(self.h, _ip(info))
        (self.nel, self.npoints, self.I2, self.vd, self.nsuper, self.ndof, self.nnz,
         self.M) = [int(v) for v in info]
        self.P = P

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_space_free(self.h)
            self.h = None

    def elements(self):
        n = self.nel
        vert = np.zeros((n, 4), dtype=np.int32)
        patch = np.zeros(n, dtype=np.int32)
        llc = np.zeros((n, 2))
        mid = np.zeros((n, 3))
        rad = np.zeros(n)
        lib().or_space_elements(self.h, _ip(vert), _ip(patch), _dp(llc), _dp(mid), _dp(rad))
        return dict(vert=vert, patch=patch, llc=llc, mid=mid, radius=rad)

    def clusters(self, level):
        n = 6 * 4 ** level if level >= 0 else 0
        mid = np.zeros((max(n, 1) * 64, 3))
        rad = np.zeros(max(n, 1) * 64)
        k = lib().or_space_clusters(self.h, level, _dp(mid), _dp(rad))
        return mid[:k].copy(), rad[:k].copy()

    def T(self):
        rp = np.zeros(self.nsuper + 1, dtype=np.int32)
        col = np.zeros(self.nnz, dtype=np.int32)
        val = np.zeros(self.nnz)
        lib().or_space_T(self.h, _ip(rp), _ip(col), _dp(val))
        return rp, col, val

    def T_dense(self):
        rp, col, val = self.T()
        Tm = np.zeros((self.nsuper, self.ndof))
        for r in range(self.nsuper):
            for k in range(rp[r], rp[r + 1]):
                Tm[r, col[k]] += val[k]
        return Tm

    def pair_info(self, e1, e2):
        out = np.zeros(4, dtype=np.int32)
        d = C.c_double(0)
        lib().or_pair_info(self.h, e1, e2, _ip(out), C.byref(d))
        return [int(v) for v in out], d.value

    def pair_block(self, kind, e1, e2, kappa=0j, dummy_fn=0):
        n = self.vd * self.I2
        cplx = kind in (HELMHOLTZ, MAXWELL)
        out = np.zeros(n * n * (2 if cplx else 1))
        lib().or_pair_block(self.h, kind, kappa.real, kappa.imag, dummy_fn, e1, e2, _dp(out))
        if cplx:
            return out.view(np.complex128).reshape(n, n)
        return out.reshape(n, n)

    def integrate(self, kind, which, e1, rot1, e2, rot2, qdeg, kappa=0j, dummy_fn=0):
        n = self.vd * self.I2
        cplx = kind in (HELMHOLTZ, MAXWELL)
        out = np.zeros(n * n * (2 if cplx else 1))
        lib().or_integrate(self.h, kind, kappa.real, kappa.imag, dummy_fn, which, e1, rot1, e2, rot2,
                           qdeg, _dp(out))
        if cplx:
            return out.view(np.complex128).reshape(n, n)
        return out.reshape(n, n)

    def dense_super(self, kind, kappa=0j):
        cplx = kind in (HELMHOLTZ, MAXWELL)
        N = self.nsuper
        A = np.zeros(N * N * (2 if cplx else 1))
        lib().or_dense_super(self.h, kind, kappa.real, kappa.imag, _dp(A))
        if cplx:
            return A.view(np.complex128).reshape(N, N).T.copy()
        return A.reshape(N, N).T.copy()

    def dense_galerkin(self, kind, kappa=0j):
        cplx = kind in (HELMHOLTZ, MAXWELL)
        D = self.ndof
        A = np.zeros(D * D * (2 if cplx else 1))
        lib().or_dense_galerkin(self.h, kind, kappa.real, kappa.imag, _dp(A))
        if cplx:
            return A.view(np.complex128).reshape(D, D).T.copy()
        return A.reshape(D, D).T.copy()

    def potential(self, kind, rho, points, kappa=0j):
        """DiscretePotential of the Cauchy data rho at points (n, 3): (n,) or (n, 3) for Maxwell"""
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        n = len(pts)
        dim = 3 if kind == MAXWELL else 1
        cplx = kind != LAPLACE
        r = np.ascontiguousarray(rho, dtype=np.complex128 if cplx else np.float64)
        rin = r.view(np.float64) if cplx else r
        out = np.zeros(n * dim * (2 if cplx else 1))
        lib().or_potential(self.h, kind, kappa.real, kappa.imag, _dp(rin), n, _dp(pts), _dp(out))
        v = out.view(np.complex128) if cplx else out
        v = v.reshape(dim, n).T
        return v[:, 0].copy() if dim == 1 else v.copy()

    def rhs(self, fn, p, kappa=0.0):
        cplx = fn in (1, 2, 3)
        out = np.zeros(self.ndof * (2 if cplx else 1))
        pa = np.asarray(p, dtype=np.float64)
        lib().or_rhs(self.h, fn, _dp(pa), kappa, _dp(out))
        return out.view(np.complex128) if cplx else out

    def moment(self, cl, cr, n, comp=0):
        rows = n * n * (2 if comp else 1)
        cols = 4 ** cr * self.I2
        out = np.zeros(rows * cols)
        lib().or_moment_matrix(self.h, cl, cr, n, comp, _dp(out))
        return out.reshape(cols, rows).T.copy()


def cheb_points(n):
    x = np.zeros(n)
    lib().or_cheb_points(n, _dp(x))
    return x


def transfer_matrix(n):
    T = np.zeros(n * n * 4 * n * n)
    lib().or_transfer_matrix(n, _dp(T))
    return T.reshape(4 * n * n, n * n).T.copy()


def forward_transform(mom, T, steps, X):
    n2 = T.shape[0]
    nc = mom.shape[0] // n2
    momF = np.asfortranarray(mom)
    TF = np.asfortranarray(T)
    XF = np.asfortranarray(X)
    ncols = X.shape[1]
    levels = []
    ptrs = (C.POINTER(C.c_double) * (steps + 1))()
    n = ncols
    for k in range(steps + 1):
        arr = np.zeros((mom.shape[0], max(n, 1)), order="F")
        levels.append(arr)
        ptrs[k] = _dp(arr)
        n //= 4
    lib().or_forward_transform(n2, nc, _dp(momF), mom.shape[1], _dp(TF), steps, _dp(XF), ncols, ptrs)
    return levels


class H2:
    def __init__(self, space, kind, kappa=0j, np_=9, eta=1.6, mcl=1, assemble=True):
        self.space = space
        self.kind = kind
        self.cplx = kind in (HELMHOLTZ, MAXWELL)
        self.h = lib().or_h2_build(space.h, kind, kappa.real, kappa.imag, np_, eta, mcl, int(assemble))
        info = np.zeros(6, dtype=np.int32)
        lib().or_h2_info(self.h, _ip(info))
        self.nleaves, self.ndense, self.nlowrank, self.dense_dim, self.lr_dim, self.levels = [int(v) for v in info]

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_h2_free(self.h)
            self.h = None

    def leaves(self):
        out = np.zeros((self.nleaves, 4), dtype=np.int32)
        lib().or_h2_leaves(self.h, _ip(out))
        return out

    def leaf_matrix(self, i, r=0, c=0, dim=None):
        cc = int(self.leaves()[i, 0]) if dim is None else None
        d = dim if dim is not None else (self.dense_dim if cc == 2 else self.lr_dim)
        out = np.zeros(d * d * (2 if self.cplx else 1))
        lib().or_h2_leaf_matrix(self.h, i, r, c, _dp(out))
        if self.cplx:
            return out.view(np.complex128).reshape(d, d).T.copy()
        return out.reshape(d, d).T.copy()

    def set_leaf_matrix(self, i, F, r=0, c=0):
        a = np.asfortranarray(F)
        if self.cplx:
            a = np.asfortranarray(a.astype(np.complex128)).view(np.float64)
            a = np.ascontiguousarray(a.T.reshape(-1))
            a = np.ascontiguousarray(np.asfortranarray(F.astype(np.complex128)).T.reshape(-1).view(np.float64))
        else:
            a = np.ascontiguousarray(F.T.reshape(-1))
        lib().or_h2_set_leaf_matrix(self.h, i, r, c, _dp(a))

    def assemble_leaves(self, idx):
        a = np.ascontiguousarray(idx, dtype=np.int32)
        lib().or_h2_assemble_leaves(self.h, len(a), _ip(a))

    def import_leaves(self, keys, data):
        k = np.ascontiguousarray(keys, dtype=np.int32)
        d = np.ascontiguousarray(data, dtype=np.float64)
        return lib().or_h2_import_leaves(self.h, len(k), _ip(k), _dp(d))

    def time_matvec(self, nrhs=1, reps=3, threads=0):
        return lib().or_time_matvec(self.h, nrhs, reps, threads)

    def time_assembly_sample(self, pairs, blocks, threads=0):
        p = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1)
        b = np.ascontiguousarray(blocks, dtype=np.int32).reshape(-1)
        return lib().or_time_assembly_sample(self.h, len(p) // 2, _ip(p), len(b), _ip(b), threads)

    def matvec_super(self, xs):
        """super-space product (no T, T^T), xs of shape (nsuper, nrhs)"""
        X = np.asarray(xs).reshape(xs.shape[0], -1)
        nrhs, N = X.shape[1], self.space.nsuper
        if self.cplx:
            xin = np.ascontiguousarray(X.T.astype(np.complex128)).view(np.float64)
            out = np.zeros(2 * N * nrhs)
            lib().or_h2_matvec_super(self.h, nrhs, _dp(xin), _dp(out))
            return out.view(np.complex128).reshape(nrhs, N).T
        xin = np.ascontiguousarray(X.T.astype(np.float64))
        out = np.zeros(N * nrhs)
        lib().or_h2_matvec_super(self.h, nrhs, _dp(xin), _dp(out))
        return out.reshape(nrhs, N).T

    def matvec(self, x):
        x = np.asarray(x)
        X = x.reshape(x.shape[0], -1)
        nrhs = X.shape[1]
        D = self.space.ndof
        if self.cplx:
            xin = np.ascontiguousarray(X.T.astype(np.complex128)).view(np.float64)
            out = np.zeros(2 * D * nrhs)
            lib().or_h2_matvec(self.h, nrhs, _dp(xin), _dp(out))
            Y = out.view(np.complex128).reshape(nrhs, D).T
        else:
            xin = np.ascontiguousarray(X.T.astype(np.float64))
            out = np.zeros(D * nrhs)
            lib().or_h2_matvec(self.h, nrhs, _dp(xin), _dp(out))
            Y = out.reshape(nrhs, D).T
        return Y.reshape(x.shape)
