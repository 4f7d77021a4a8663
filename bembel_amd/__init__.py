"""bembel_amd: MI355X implementation of Bembel's Galerkin assembly and H2 matrix-vector product.

Python mirror of the reference API over the C ABI in include/bembel_amd.h:

    geometry = Geometry("sphere.dat")                              # Geometry/Geometry.hpp:37
    op = HelmholtzSingleLayerOperator(); op.set_wavenumber(4.0)    # Helmholtz/SingleLayerOperator.hpp:124
    space = AnsatzSpace(geometry, 5, 2, op)                        # AnsatzSpace/AnsatzSpace.hpp:80
    disc = DiscreteOperator(H2Matrix, space, op); disc.compute()   # LinearOperator/DiscreteOperator.hpp:107
    A = disc.get_discrete_operator()
    y = A @ x                                                      # H2Matrix/EigenHelper/H2DenseProduct.hpp

Host numpy arrays and device torch tensors are both accepted; device tensors stay on the GPU.
The compute path is the HIP library libbembel_amd.so; there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

__all__ = [
    "BembelError", "Geometry", "AnsatzSpace", "LaplaceSingleLayerOperator", "HelmholtzSingleLayerOperator",
    "MaxwellSingleLayerOperator", "DummyOperator", "DiscreteOperator", "H2Matrix", "DenseMatrix",
    "DiscreteLinearForm", "DirichletTrace", "PlaneWaves", "RotatedTangentialTrace", "HelmholtzFundamentalSolution",
    "HarmonicFunction", "FunctionTrace", "DiscretePotential", "LaplaceSingleLayerPotential", "HelmholtzSingleLayerPotential",
    "MaxwellSingleLayerPotential", "GMRES", "ConjugateGradient", "library", "timing",
    "SPHERE_DAT", "FP64", "MIXED", "MIXED_BF16", "Server", "synchronize",
]

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbembel_amd.so")
SPHERE_DAT = os.path.join(HERE, "geo", "sphere.dat")

OK, INVALID_ARG, OUT_OF_MEMORY, DEVICE_ERROR, NOT_CONVERGED, ADMISSION_REJECTED, IO_ERROR, UNSUPPORTED, PENDING = range(9)
LAPLACE, HELMHOLTZ, MAXWELL, DUMMY = 0, 1, 2, 3
FP64, MIXED, MIXED_BF16 = 0, 1, 2
MEM_HOST, MEM_DEVICE = 0, 1


class BembelError(RuntimeError):
    def __init__(self, status, message):
        super().__init__("bembel_amd status %d: %s" % (status, message))
        self.status = status


class OpDesc(C.Structure):
    _fields_ = [("kind", C.c_int), ("kappa_re", C.c_double), ("kappa_im", C.c_double), ("n_cheb", C.c_int),
                ("eta", C.c_double), ("min_cluster_level", C.c_int), ("precision", C.c_int),
                ("dummy_fn", C.c_int)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_double), ("max_active", C.c_int), ("max_batch_rhs", C.c_int),
                ("slice_iterations", C.c_int)]


class Request(C.Structure):
    _fields_ = [("geometry_path", C.c_char_p), ("scale", C.c_double * 3), ("refinement_level", C.c_int),
                ("polynomial_degree", C.c_int), ("op", OpDesc), ("data_fn", C.c_int), ("nrhs", C.c_int),
                ("data_params", C.POINTER(C.c_double)), ("data_kappa", C.c_double), ("tol", C.c_double),
                ("restart", C.c_int), ("maxit", C.c_int)]


class ResultInfo(C.Structure):
    _fields_ = [("status", C.c_int), ("ndof", C.c_int64), ("nrhs", C.c_int), ("submit_s", C.c_double),
                ("start_s", C.c_double), ("end_s", C.c_double), ("iterations", C.c_int), ("residual", C.c_double),
                ("batch_requests", C.c_int), ("batch_rhs", C.c_int), ("cache_hit", C.c_int), ("slices", C.c_int)]


_lib = None


def library():
    """Loads libbembel_amd.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is not None:
        return _lib
    # torch bundles its own HIP runtime on the same HSA runtime as /opt/rocm; when torch is
    # loaded first, this library binds to that HSA runtime instead of starting a second one
    try:
        import torch  # noqa: F401
    except Exception:
        pass
    path = os.environ.get("BEMBEL_AMD_LIBRARY", LIB_PATH)  # another build of the same library
    if not os.path.exists(path):
        raise RuntimeError("libbembel_amd.so is missing: build it with `python -m bembel_amd.build`")
    L = C.CDLL(path)
    vp, ip, dp, i64p = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_int64)
    pp = C.POINTER(C.c_void_p)
    sig = {
        "bmi_last_error": (C.c_char_p, []),
        "bmi_version": (C.c_char_p, []),
        "bmi_op_desc_default": (C.c_int, [C.c_int, C.POINTER(OpDesc)]),
        "bmi_geometry_load": (C.c_int, [C.c_char_p, dp, pp]),
        "bmi_geometry_destroy": (C.c_int, [vp]),
        "bmi_geometry_num_patches": (C.c_int, [vp, ip]),
        "bmi_geometry_surface_point": (C.c_int, [vp, C.c_int, C.c_double, C.c_double, dp]),
        "bmi_space_create": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, pp]),
        "bmi_space_destroy": (C.c_int, [vp]),
        "bmi_space_info": (C.c_int, [vp, i64p]),
        "bmi_space_elements": (C.c_int, [vp, ip, ip, dp, dp, dp]),
        "bmi_space_clusters": (C.c_int, [vp, C.c_int, dp, dp, ip]),
        "bmi_space_transformation": (C.c_int, [vp, ip, ip, dp]),
        "bmi_pair_info": (C.c_int, [vp, C.c_int, ip, ip]),
        "bmi_pair_blocks": (C.c_int, [vp, C.POINTER(OpDesc), C.c_int, ip, dp, vp]),
        "bmi_integrate_jobs": (C.c_int, [vp, C.POINTER(OpDesc), C.c_int, ip, C.c_int, dp, vp]),
        "bmi_dense_assemble": (C.c_int, [vp, C.POINTER(OpDesc), C.c_int, dp, C.c_int, vp]),
        "bmi_h2_assemble": (C.c_int, [vp, C.POINTER(OpDesc), vp, pp]),
        "bmi_h2_destroy": (C.c_int, [vp]),
        "bmi_h2_info": (C.c_int, [vp, i64p]),
        "bmi_h2_leaves": (C.c_int, [vp, ip]),
        "bmi_h2_leaf_matrix": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, dp]),
        "bmi_h2_export_leaves": (C.c_int, [vp, dp]),
        "bmi_h2_matvec": (C.c_int, [vp, C.c_int, dp, C.c_int64, dp, C.c_int64, dp, dp, C.c_int, vp]),
        "bmi_h2_matvec_noalloc": (C.c_int, [vp, C.c_int, dp, C.c_int64, dp, C.c_int64, dp, dp, vp]),
        "bmi_h2_matvec_super": (C.c_int, [vp, C.c_int, dp, dp, vp]),
        "bmi_h2_use_graphs": (C.c_int, [vp, C.c_int]),
        "bmi_h2_set_sweep": (C.c_int, [vp, C.c_int]),
        "bmi_h2_set_overlap": (C.c_int, [vp, C.c_int]),
        "bmi_gmres_start": (C.c_int, [vp, C.c_int, vp, vp, C.c_double, C.c_int, C.c_int, C.c_int, vp,
                                       C.POINTER(vp)]),
        "bmi_gmres_poll": (C.c_int, [vp, C.POINTER(C.c_int)]),
        "bmi_gmres_finish": (C.c_int, [vp, ip, dp]),
        "bmi_stream_create": (C.c_int, [C.POINTER(vp)]),
        "bmi_stream_destroy": (C.c_int, [vp]),
        "bmi_synchronize": (C.c_int, [vp]),
        "bmi_rhs_assemble": (C.c_int, [vp, C.c_int, dp, C.c_int, C.c_double, dp, C.c_int, vp]),
        "bmi_potential_evaluate": (C.c_int, [vp, C.POINTER(OpDesc), dp, C.c_int, dp, dp, C.c_int, vp]),
        "bmi_linear_form_points": (C.c_int, [vp, i64p, dp]),
        "bmi_rhs_from_values": (C.c_int, [vp, C.c_int, dp, C.c_int, dp, C.c_int, vp]),
        "bmi_solve_gmres": (C.c_int, [vp, C.c_int, dp, dp, C.c_double, C.c_int, C.c_int, C.c_int, ip, dp,
                                      C.c_int, vp]),
        "bmi_solve_cg": (C.c_int, [vp, C.c_int, dp, dp, C.c_double, C.c_int, C.c_int, ip, dp, C.c_int, vp]),
        "bmi_timing_enable": (C.c_int, [C.c_int]),
        "bmi_timing_reset": (C.c_int, []),
        "bmi_timing_get": (C.c_int, [C.c_char_p, dp, i64p]),
        "bmi_h2_estimate_bytes": (C.c_int, [vp, vp, C.c_int, i64p]),
        "bmi_server_config_default": (C.c_int, [vp]),
        "bmi_server_create": (C.c_int, [vp, pp]),
        "bmi_server_submit": (C.c_int, [vp, vp, i64p]),
        "bmi_server_wait": (C.c_int, [vp, C.c_int64, C.c_double, vp]),
        "bmi_server_fetch": (C.c_int, [vp, C.c_int64, dp]),
        "bmi_server_stats": (C.c_int, [vp, i64p]),
        "bmi_server_destroy": (C.c_int, [vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(status, allow=()):
    if status != OK and status not in allow:
        raise BembelError(status, library().bmi_last_error().decode())
    return status


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _stream_ptr(stream):
    """A stream of this library's HIP runtime (an integer handle) or None for its default
    stream.  torch streams belong to torch's own HIP runtime and are not passed through: work on
    them is synchronised at the boundary instead (_torch_sync)."""
    if stream is None:
        return None
    if isinstance(stream, int):
        return C.c_void_p(stream)
    raise BembelError(INVALID_ARG, "pass None or an integer stream handle of this library's HIP runtime")


def _torch_sync():
    """Finish pending torch work before the library reads or writes torch tensors."""
    try:
        import torch
        if torch.cuda.is_available() and torch.cuda.is_initialized():
            torch.cuda.current_stream().synchronize()
    except ImportError:
        pass


# temporaries (column-major copies of torch inputs) of products enqueued with sync=False; the
# library may still read them, so they are kept until the next synchronize()
_pending = []


def build_info():
    """Provenance of the loaded library: the source digest it was built from (bmi_version), the
    digest of the sources in this tree, whether they match, and the library file's own sha256."""
    import hashlib
    from . import build as _build
    version = library().bmi_version().decode()
    built = version.rsplit(" ", 1)[-1] if " sources " in version else "unknown"
    try:
        tree = _build.source_digest()
    except OSError:
        tree = "unavailable"
    path = os.environ.get("BEMBEL_AMD_LIBRARY", LIB_PATH)
    with open(path, "rb") as f:
        lib_sha = hashlib.sha256(f.read()).hexdigest()[:16]
    return {"version": version, "library_sources": built, "tree_sources": tree, "match": built == tree,
            "library_sha256": lib_sha}


def synchronize(stream=None):
    """Wait for the library's work on `stream` (default stream if None)."""
    _check(library().bmi_synchronize(_stream_ptr(stream)))
    del _pending[:]


def _is_torch_cuda(x):
    return type(x).__module__.startswith("torch") and getattr(x, "is_cuda", False)


class timing:
    """Per-kernel HIP-event timing inside the library (bmi_timing_*)."""

    @staticmethod
    def enable(on=True):
        _check(library().bmi_timing_enable(1 if on else 0))

    @staticmethod
    def reset():
        _check(library().bmi_timing_reset())

    @staticmethod
    def get(name):
        ms, n = C.c_double(0), C.c_int64(0)
        _check(library().bmi_timing_get(name.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value


# ---------------------------------------------------------------------------- geometry
class Geometry:
    """Geometry(filename) of rational Bezier patches (Geometry/Geometry.hpp:37).

    scale multiplies the homogeneous x, y, z control coordinates (the sphere file scaled by
    (a, b, c) is an exact ellipsoid)."""

    def __init__(self, filename=SPHERE_DAT, scale=None):
        h = C.c_void_p()
        sc = None if scale is None else np.ascontiguousarray(scale, dtype=np.float64)
        _check(library().bmi_geometry_load(filename.encode(), None if sc is None else _dp(sc), C.byref(h)))
        self._destroy = library().bmi_geometry_destroy
        self._h = h
        self.filename = filename
        self.scale = None if scale is None else tuple(scale)

    def __del__(self):
        # the destroy function is bound when the handle is made: at interpreter exit the module's
        # globals may already be gone
        if getattr(self, "_h", None):
            self._destroy(self._h)
            self._h = None

    @property
    def number_of_patches(self):
        n = C.c_int(0)
        _check(library().bmi_geometry_num_patches(self._h, C.byref(n)))
        return n.value

    def surface_point(self, patch, s, t):
        """Patch::updateSurfacePoint: (s, t, 1, x, dx/ds, dx/dt)."""
        out = np.zeros(12)
        _check(library().bmi_geometry_surface_point(self._h, patch, s, t, _dp(out)))
        return out


# ---------------------------------------------------------------------------- operators
class _Operator:
    kind = None
    form = 0
    scalar = np.float64

    def __init__(self):
        self.wavenumber = 0j

    def set_wavenumber(self, k):
        self.wavenumber = complex(k)

    def get_wavenumber(self):
        return self.wavenumber

    def desc(self, n_cheb=9, eta=1.6, min_cluster_level=1, precision=FP64, dummy_fn=0):
        d = OpDesc()
        _check(library().bmi_op_desc_default(self.kind, C.byref(d)))
        d.kappa_re = self.wavenumber.real
        d.kappa_im = self.wavenumber.imag
        d.n_cheb = n_cheb
        d.eta = eta
        d.min_cluster_level = min_cluster_level
        d.precision = precision
        d.dummy_fn = dummy_fn
        return d


class LaplaceSingleLayerOperator(_Operator):
    """kernel 1/(4 pi |x-y|) (Laplace/SingleLayerOperator.hpp:115-118)"""
    kind = LAPLACE


class HelmholtzSingleLayerOperator(_Operator):
    """kernel exp(-i kappa |x-y|)/(4 pi |x-y|) (Helmholtz/SingleLayerOperator.hpp:115-120)"""
    kind = HELMHOLTZ
    scalar = np.complex128


class MaxwellSingleLayerOperator(_Operator):
    """EFIE single layer (Maxwell/SingleLayerOperator.hpp), DivConforming space"""
    kind = MAXWELL
    form = 1
    scalar = np.complex128


class DummyOperator(_Operator):
    """Test operator integrating a fixed function f(x, y) (LinearOperator/Dummy/DummyOperator.hpp)."""
    kind = DUMMY

    def __init__(self, test_function=0):
        super().__init__()
        self.test_function = test_function

    def desc(self, **kw):
        kw.setdefault("dummy_fn", self.test_function)
        return super().desc(**kw)


# ---------------------------------------------------------------------------- space
class AnsatzSpace:
    """AnsatzSpace<Op>(geometry, refinement_level, polynomial_degree) (AnsatzSpace.hpp:80-111)."""

    def __init__(self, geometry, refinement_level, polynomial_degree, operator=None, knot_repetition=1):
        if knot_repetition != 1:
            raise BembelError(UNSUPPORTED, "only knot_repetition = 1 is supported")
        op = operator if operator is not None else LaplaceSingleLayerOperator()
        form = op.form
        h = C.c_void_p()
        _check(library().bmi_space_create(geometry._h, refinement_level, polynomial_degree, form, C.byref(h)))
        self._destroy = library().bmi_space_destroy
        self._h = h
        self.geometry = geometry  # keeps the geometry alive
        info = np.zeros(8, dtype=np.int64)
        _check(library().bmi_space_info(self._h, info.ctypes.data_as(C.POINTER(C.c_int64))))
        (self.number_of_elements, self.number_of_vertices, self.I2, self.vector_dimension, self.number_of_super_dofs,
         self.number_of_dofs, self.nnz, self.refinement_level) = [int(v) for v in info]
        self.polynomial_degree = polynomial_degree
        self.form = form

    def __del__(self):
        # the destroy function is bound when the handle is made: at interpreter exit the module's
        # globals may already be gone
        if getattr(self, "_h", None):
            self._destroy(self._h)
            self._h = None

    def get_number_of_dofs(self):
        return self.number_of_dofs

    def get_polynomial_degree(self):
        return self.polynomial_degree

    def get_refinement_level(self):
        return self.refinement_level

    def get_number_of_elements(self):
        return self.number_of_elements

    def elements(self):
        n = self.number_of_elements
        vert = np.zeros((n, 4), dtype=np.int32)
        patch = np.zeros(n, dtype=np.int32)
        llc = np.zeros((n, 2))
        mid = np.zeros((n, 3))
        rad = np.zeros(n)
        _check(library().bmi_space_elements(self._h, _ip(vert), _ip(patch), _dp(llc), _dp(mid), _dp(rad)))
        return dict(vert=vert, patch=patch, llc=llc, mid=mid, radius=rad)

    def clusters(self, level):
        cnt = C.c_int(0)
        _check(library().bmi_space_clusters(self._h, level, None, None, C.byref(cnt)))
        mid = np.zeros((cnt.value, 3))
        rad = np.zeros(cnt.value)
        _check(library().bmi_space_clusters(self._h, level, _dp(mid), _dp(rad), C.byref(cnt)))
        return mid, rad

    def get_transformation_matrix(self):
        """T as (rowptr, col, val) CSR arrays (AnsatzSpace::get_transformation_matrix)."""
        rp = np.zeros(self.number_of_super_dofs + 1, dtype=np.int32)
        col = np.zeros(self.nnz, dtype=np.int32)
        val = np.zeros(self.nnz)
        _check(library().bmi_space_transformation(self._h, _ip(rp), _ip(col), _dp(val)))
        return rp, col, val

    def pair_info(self, pairs):
        p = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        out = np.zeros((len(p), 4), dtype=np.int32)
        _check(library().bmi_pair_info(self._h, len(p), _ip(p), _ip(out)))
        return out

    def pair_blocks(self, operator, pairs, stream=None):
        """evaluateBilinearForm on the GPU for element pairs; returns (n, vd*I2, vd*I2)."""
        p = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        d = operator.desc()
        nb = self.vector_dimension * self.I2
        cplx = operator.scalar == np.complex128
        out = np.zeros(len(p) * nb * nb * (2 if cplx else 1))
        _check(library().bmi_pair_blocks(self._h, C.byref(d), len(p), _ip(p), _dp(out), _stream_ptr(stream)))
        if cplx:
            return out.view(np.complex128).reshape(len(p), nb, nb)
        return out.reshape(len(p), nb, nb)

    def integrate_jobs(self, operator, jobs, qdeg, stream=None):
        """integrate0..4 with explicit rotations on the GPU (reference Duffy unit tests)."""
        j = np.ascontiguousarray(jobs, dtype=np.int32).reshape(-1, 5)
        d = operator.desc()
        out = np.zeros(len(j))
        _check(library().bmi_integrate_jobs(self._h, C.byref(d), len(j), _ip(j), qdeg, _dp(out),
                                            _stream_ptr(stream)))
        return out


# ---------------------------------------------------------------------------- matrices
def _columns(x, n, scalar):
    """(array, nrhs, ld, mem, squeeze, backend) for a vector or column-major matrix."""
    if _is_torch_cuda(x):
        import torch
        squeeze = x.dim() == 1
        X = x.reshape(-1, 1) if squeeze else x
        if X.shape[0] != n:
            raise BembelError(INVALID_ARG, "dimension mismatch")
        want = torch.complex128 if scalar == np.complex128 else torch.float64
        if X.dtype != want:
            X = X.to(want)
        Xc = X.t().contiguous()  # columns contiguous
        # a copy made here runs on torch's current stream: the caller must order it before the
        # library reads it (_torch_sync) and keep it alive until the library is done with it
        copied = Xc.data_ptr() != x.data_ptr()
        return Xc, X.shape[1], n, MEM_DEVICE, squeeze, "torch" if not copied else "torch-copy"
    a = np.asarray(x)
    squeeze = a.ndim == 1
    A = a.reshape(-1, 1) if squeeze else a
    if A.shape[0] != n:
        raise BembelError(INVALID_ARG, "dimension mismatch")
    Ac = np.ascontiguousarray(A.T.astype(scalar))
    return Ac, A.shape[1], n, MEM_HOST, squeeze, "numpy"


def _dptr(a, backend):
    if backend.startswith("torch"):
        return C.cast(C.c_void_p(a.data_ptr()), C.POINTER(C.c_double))
    return a.ctypes.data_as(C.POINTER(C.c_double))


class H2Matrix:
    """Device-resident H2 matrix (Eigen::H2Matrix, H2Matrix/H2Matrix.hpp:59-296)."""

    def __init__(self, space, operator, n_cheb=9, eta=1.6, min_cluster_level=1, precision=FP64, stream=None):
        self.space = space
        self.operator = operator
        self.scalar = operator.scalar
        self.precision = precision
        self.n_cheb = n_cheb
        d = operator.desc(n_cheb=n_cheb, eta=eta, min_cluster_level=min_cluster_level, precision=precision)
        h = C.c_void_p()
        _check(library().bmi_h2_assemble(space._h, C.byref(d), _stream_ptr(stream), C.byref(h)))
        self._destroy = library().bmi_h2_destroy
        self._h = h
        info = np.zeros(16, dtype=np.int64)
        _check(library().bmi_h2_info(self._h, info.ctypes.data_as(C.POINTER(C.c_int64))))
        self.info = dict(leaves=int(info[0]), dense=int(info[1]), lowrank=int(info[2]), pairs=int(info[3]),
                         dense_dim=int(info[4]), coupling_dim=int(info[5]), levels=int(info[6]),
                         dofs=int(info[7]), device_bytes=int(info[8]), complex=bool(info[9]),
                         pair_classes=[int(v) for v in info[11:16]])

    def __del__(self):
        # the destroy function is bound when the handle is made: at interpreter exit the module's
        # globals may already be gone
        if getattr(self, "_h", None):
            self._destroy(self._h)
            self._h = None

    def rows(self):
        return self.space.number_of_dofs

    def cols(self):
        return self.space.number_of_dofs

    @property
    def shape(self):
        return (self.rows(), self.cols())

    def leaves(self):
        out = np.zeros((self.info["leaves"], 4), dtype=np.int32)
        _check(library().bmi_h2_leaves(self._h, _ip(out)))
        return out

    def leaf_matrix(self, leaf, r=0, c=0):
        lv = self.leaves()[leaf]
        d = self.info["dense_dim"] if lv[0] == 2 else self.info["coupling_dim"]
        cplx = self.scalar == np.complex128
        out = np.zeros(d * d * (2 if cplx else 1))
        _check(library().bmi_h2_leaf_matrix(self._h, leaf, r, c, _dp(out)))
        if cplx:
            return out.view(np.complex128).reshape(d, d).T.copy()
        return out.reshape(d, d).T.copy()

    def export_leaf_data(self, indices):
        """(keys, data) of the leaves `indices` only, in the layout of export_leaves"""
        idx = np.asarray(indices, dtype=np.int64)
        keys = self.leaves()[idx]
        vd = self.space.vector_dimension
        S = 2 if self.scalar == np.complex128 else 1
        dims = np.where(keys[:, 0] == 2, self.info["dense_dim"], self.info["coupling_dim"])
        data = np.empty(int(np.sum(dims.astype(np.int64) ** 2)) * vd * vd * S)
        o = 0
        for k, d in zip(idx.tolist(), dims.tolist()):
            for r in range(vd):
                for c in range(vd):
                    _check(library().bmi_h2_leaf_matrix(self._h, int(k), r, c, _dp(data[o:])))
                    o += d * d * S
        return keys, data

    def export_leaves(self):
        """(keys, data): keys as in leaves(), data = all leaf matrices back to back (float64,
        complex entries interleaved)."""
        vd = self.space.vector_dimension
        S = 2 if self.scalar == np.complex128 else 1
        n = (self.info["dense"] * self.info["dense_dim"] ** 2 +
             self.info["lowrank"] * self.info["coupling_dim"] ** 2) * vd * vd * S
        data = np.empty(n)
        _check(library().bmi_h2_export_leaves(self._h, _dp(data)))
        return self.leaves(), data

    def matvec(self, x, alpha=1.0, stream=None, sync=True):
        """alpha * H x for a vector or a (n x nrhs) matrix (numpy on host or torch on device).
        With device tensors the call waits for pending torch work before and for the product
        after (torch runs on its own HIP runtime).  sync=False leaves that ordering to the caller
        (torch.cuda.synchronize() before, bembel_amd.synchronize() after): the product is only
        enqueued, so back-to-back products overlap their host overhead with the device work."""
        n = self.rows()
        X, nrhs, ld, mem, squeeze, backend = _columns(x, n, self.scalar)
        if backend == "torch-copy" or (backend == "torch" and sync):
            _torch_sync()
        if backend == "torch-copy" and not sync:
            _pending.append(X)  # released by synchronize(), after the library's reads
        # beta = 0: Y is written, never read
        if backend.startswith("torch"):
            import torch
            Y = torch.empty_like(X)
        else:
            Y = np.empty_like(X)
        al = np.array([complex(alpha).real, complex(alpha).imag])
        be = np.zeros(2)
        _check(library().bmi_h2_matvec(self._h, nrhs, _dptr(X, backend), ld, _dptr(Y, backend), ld, _dp(al), _dp(be),
                                       mem, _stream_ptr(stream)))
        if backend.startswith("torch") and sync:
            synchronize(stream)
        Y = Y.t() if backend.startswith("torch") else Y.T
        return Y.reshape(-1) if squeeze else Y

    def __matmul__(self, x):
        return self.matvec(x)

    def matvec_super(self, xs, stream=None):
        """Super-space product A_super xs (no T, T^T) for a torch device tensor of shape
        (number_of_super_dofs, nrhs) or (number_of_super_dofs,)."""
        import torch
        if not _is_torch_cuda(xs):
            raise BembelError(INVALID_ARG, "matvec_super takes a device tensor")
        N = self.space.number_of_super_dofs
        X, nrhs, ld, mem, squeeze, backend = _columns(xs, N, self.scalar)
        Y = torch.empty_like(X)
        _torch_sync()
        _check(library().bmi_h2_matvec_super(self._h, nrhs, _dptr(X, backend), _dptr(Y, backend),
                                             _stream_ptr(stream)))
        Y = Y.t()
        return Y.reshape(-1) if squeeze else Y

    def use_graphs(self, enable=True):
        """Replay products of two or more columns from captured HIP graphs (default) or launch
        kernel by kernel (single-column products always launch kernel by kernel)."""
        _check(library().bmi_h2_use_graphs(self._h, 1 if enable else 0))

    def set_sweep(self, mode):
        """Sweep of single-RHS products: None (timed choice, default), "full" (every block),
        "upper" (upper coupling blocks of a symmetric kernel with the transposed products formed
        from them) or "upper_near" (the same for the near-field leaves too)."""
        m = {None: -1, "full": 0, "upper": 1, "upper_near": 2}[mode]
        _check(library().bmi_h2_set_sweep(self._h, m))

    def set_overlap(self, mode):
        """Single-RHS products: None (timed choice, default), "fused" (near-field leaves and
        coupling blocks in one launch) or "overlap" (near-field leaves beside the coupling
        blocks and the backward transformation on a side stream)."""
        m = {None: -1, "fused": 0, "overlap": 1}[mode]
        _check(library().bmi_h2_set_overlap(self._h, m))

    def get_dense(self):
        """Dense matrix by applying H to the identity (H2Matrix::get_dense, H2Matrix.hpp:254-264)."""
        n = self.rows()
        return self.matvec(np.eye(n, dtype=self.scalar))


class DenseMatrix:
    """Dense Galerkin matrix T^T A T (DiscreteOperatorComputer<Eigen::MatrixX>)."""

    def __init__(self, space, operator, super_only=False, stream=None):
        d = operator.desc()
        n = space.number_of_super_dofs if super_only else space.number_of_dofs
        cplx = operator.scalar == np.complex128
        out = np.zeros(n * n * (2 if cplx else 1))
        _check(library().bmi_dense_assemble(space._h, C.byref(d), 1 if super_only else 0, _dp(out), MEM_HOST,
                                            _stream_ptr(stream)))
        A = out.view(np.complex128) if cplx else out
        self.matrix = A.reshape(n, n).T.copy()

    def __matmul__(self, x):
        return self.matrix @ x


class DiscreteOperator:
    """DiscreteOperator<Format, Op>(ansatz_space) with compute() (LinearOperator/DiscreteOperator.hpp:107-145)."""

    def __init__(self, fmt, space, operator=None, **h2_options):
        self.fmt = fmt
        self.space = space
        self.operator = operator if operator is not None else LaplaceSingleLayerOperator()
        self.h2_options = h2_options
        self._disc = None

    def get_linear_operator(self):
        return self.operator

    def compute(self, stream=None):
        if self.fmt is H2Matrix:
            self._disc = H2Matrix(self.space, self.operator, stream=stream, **self.h2_options)
        else:
            self._disc = DenseMatrix(self.space, self.operator, stream=stream)
        return self

    def get_discrete_operator(self):
        if self._disc is None:
            raise BembelError(INVALID_ARG, "compute() has not been called")
        return self._disc


# ---------------------------------------------------------------------------- linear forms
class DirichletTrace:
    """g_k(x) = 1/|x - x0_k| (point sources; one column per source)."""

    def __init__(self, x0):
        p = np.asarray(x0, dtype=np.float64).reshape(-1, 3)
        self.fn = 0
        self.params = np.ascontiguousarray(p)
        self.nrhs = p.shape[0]
        self.kappa = 0.0


class PlaneWaves:
    """g_k(x) = exp(i kappa d_k . x) for unit directions d_k (one column per direction)."""

    def __init__(self, directions, kappa):
        d = np.asarray(directions, dtype=np.float64).reshape(-1, 3)
        self.fn = 1
        self.params = np.ascontiguousarray(d)
        self.nrhs = d.shape[0]
        self.kappa = float(kappa)


class HelmholtzFundamentalSolution:
    """g_k(x) = exp(-i kappa |x - c_k|) / |x - c_k| (Data::HelmholtzFundamentalSolution,
    examples/Data.hpp:37-44; one column per center)."""

    def __init__(self, centers, kappa):
        c = np.asarray(centers, dtype=np.float64).reshape(-1, 3)
        self.fn = 3
        self.params = np.ascontiguousarray(c)
        self.nrhs = c.shape[0]
        self.kappa = float(kappa)


class HarmonicFunction:
    """g(x) = 4 x^2 - 3 y^2 - z^2 (Data::HarmonicFunction, examples/Data.hpp:14-16)."""

    def __init__(self):
        self.fn = 4
        self.params = np.zeros((1, 3))
        self.nrhs = 1
        self.kappa = 0.0


class FunctionTrace:
    """Data given as a Python function of the surface points, as DirichletTrace<Scalar>::
    set_function / RotatedTangentialTrace<Scalar>::set_function take an arbitrary function
    (LinearForm/DirichletTrace.hpp:31-70, LinearForm/RotatedTangentialTrace.hpp:30-87).
    f maps an (n, 3) array of points to (n,) scalars (real or complex) or, with vector=True, to
    (n, 3) complex field values whose rotated tangential trace is taken (DivConforming spaces).
    One column; the function is called once with every point of the (P+2)^2-point rule."""

    def __init__(self, f, scalar=np.float64, vector=False):
        self.f = f
        self.vector = vector
        self.fn = 7 if vector else (5 if scalar == np.float64 else 6)
        self.nrhs = 1
        self.kappa = 0.0


class RotatedTangentialTrace:
    """Rotated tangential trace of Hertz dipole fields for the EFIE right-hand side
    (LinearForm/RotatedTangentialTrace.hpp with the examples/Data.hpp:76-90 Dipole); one column per
    (position, moment) pair; DivConforming spaces."""

    def __init__(self, positions, moments, kappa):
        p = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
        m = np.asarray(moments, dtype=np.float64).reshape(-1, 3)
        if p.shape != m.shape:
            raise BembelError(INVALID_ARG, "one moment per dipole position")
        self.fn = 2
        self.params = np.ascontiguousarray(np.concatenate([p, m], axis=1))
        self.nrhs = p.shape[0]
        self.kappa = float(kappa)


class DiscreteLinearForm:
    """DiscreteLinearForm<DirichletTrace | RotatedTangentialTrace>(ansatz_space)
    (LinearForm/DiscreteLinearForm.hpp:44-83)."""

    def __init__(self, data, space):
        self.data = data
        self.space = space
        self._value = None

    def _values(self):
        """data of a FunctionTrace at the linear-form points, flattened for bmi_rhs_from_values"""
        npts = C.c_int64()
        _check(library().bmi_linear_form_points(self.space._h, C.byref(npts), None))
        pts = np.zeros((npts.value, 3))
        _check(library().bmi_linear_form_points(self.space._h, C.byref(npts), _dp(pts)))
        v = np.asarray(self.data.f(pts))
        if self.data.fn == 5:
            return np.ascontiguousarray(v.reshape(-1), dtype=np.float64)
        if self.data.fn == 6:
            return np.ascontiguousarray(v.reshape(-1), dtype=np.complex128).view(np.float64)
        return np.ascontiguousarray(v.reshape(-1, 3), dtype=np.complex128).reshape(-1).view(np.float64)

    def compute(self, stream=None, device=False):
        n = self.space.number_of_dofs
        cplx = self.data.fn not in (0, 4, 5)
        S = 2 if cplx else 1
        nrhs = self.data.nrhs
        if self.data.fn >= 5:
            vals = self._values()
            out = np.zeros(S * n)
            _check(library().bmi_rhs_from_values(self.space._h, self.data.fn - 5, _dp(vals), 1, _dp(out), MEM_HOST,
                                                 _stream_ptr(stream)))
            self._value = out.view(np.complex128) if cplx else out
            return self
        if device:
            import torch
            out = torch.zeros(S * n * nrhs, dtype=torch.float64, device="cuda")
            _torch_sync()
            _check(library().bmi_rhs_assemble(self.space._h, self.data.fn, _dp(self.data.params), nrhs,
                                              self.data.kappa, _dptr(out, "torch"), MEM_DEVICE, _stream_ptr(stream)))
            v = torch.view_as_complex(out.view(-1, 2)) if cplx else out
            self._value = v.view(nrhs, n).t()
        else:
            out = np.zeros(S * n * nrhs)
            _check(library().bmi_rhs_assemble(self.space._h, self.data.fn, _dp(self.data.params), nrhs,
                                              self.data.kappa, _dp(out), MEM_HOST, _stream_ptr(stream)))
            v = out.view(np.complex128) if cplx else out
            self._value = v.reshape(nrhs, n).T
        if nrhs == 1:
            self._value = self._value.reshape(-1)
        return self

    def get_discrete_linear_form(self):
        return self._value


# ---------------------------------------------------------------------------- potentials
class _Potential:
    kind = None

    def __init__(self):
        self._kappa = 0j

    def set_wavenumber(self, k):
        self._kappa = complex(k)

    def get_wavenumber(self):
        return self._kappa


class LaplaceSingleLayerPotential(_Potential):
    """Laplace/SingleLayerPotential.hpp: 1/(4 pi |x - y|)."""
    kind = LAPLACE


class HelmholtzSingleLayerPotential(_Potential):
    """Helmholtz/SingleLayerPotential.hpp: exp(-i kappa |x - y|)/(4 pi |x - y|)."""
    kind = HELMHOLTZ


class MaxwellSingleLayerPotential(_Potential):
    """Maxwell/SingleLayerPotential.hpp: kernel times the vector density plus the kernel gradient
    times its surface divergence over kappa^2."""
    kind = MAXWELL


class DiscretePotential:
    """DiscretePotential<Potential, Op>(ansatz_space) (Potential/DiscretePotential.hpp:20-120):
    set_cauchy_data(rho), then evaluate(points) -> (n,) for Laplace/Helmholtz, (n, 3) for Maxwell."""

    def __init__(self, potential, space):
        self._pot = potential
        self.space = space
        self._rho = None

    def get_potential(self):
        return self._pot

    def set_cauchy_data(self, rho):
        cplx = self._pot.kind != LAPLACE
        r = rho.detach().cpu().numpy() if _is_torch_cuda(rho) else np.asarray(rho)
        self._rho = np.ascontiguousarray(r.reshape(-1), dtype=np.complex128 if cplx else np.float64)
        if len(self._rho) != self.space.number_of_dofs:
            raise BembelError(INVALID_ARG, "Cauchy data must have one entry per DOF")

    def evaluate(self, points, stream=None):
        if self._rho is None:
            raise BembelError(INVALID_ARG, "set_cauchy_data() has not been called")
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        n = len(pts)
        kind = self._pot.kind
        cplx = kind != LAPLACE
        dim = 3 if kind == MAXWELL else 1
        d = OpDesc()
        _check(library().bmi_op_desc_default(kind, C.byref(d)))
        k = self._pot.get_wavenumber()
        d.kappa_re, d.kappa_im = k.real, k.imag
        out = np.zeros(n * dim * (2 if cplx else 1))
        rin = self._rho.view(np.float64) if cplx else self._rho
        _check(library().bmi_potential_evaluate(self.space._h, C.byref(d), _dp(rin), n, _dp(pts), _dp(out), MEM_HOST,
                                                _stream_ptr(stream)))
        v = (out.view(np.complex128) if cplx else out).reshape(dim, n).T
        return v[:, 0].copy() if dim == 1 else v.copy()


# ---------------------------------------------------------------------------- solvers
class _Solver:
    def __init__(self):
        self.tol = 1e-10
        self.maxit = 1000
        self._A = None
        self._iters = None
        self._resid = None
        self._info = OK

    def set_tolerance(self, tol):
        self.tol = float(tol)

    def set_max_iterations(self, n):
        self.maxit = int(n)

    def compute(self, A):
        self._A = A
        return self

    def iterations(self):
        return self._iters

    def error(self):
        return self._resid

    def info(self):
        return self._info

    def _solve(self, b, x0, fn, stream):
        A = self._A
        n = A.rows()
        B, nrhs, ld, mem, squeeze, backend = _columns(b, n, A.scalar)
        if x0 is None:
            if backend.startswith("torch"):
                import torch
                X = torch.zeros_like(B)
            else:
                X = np.zeros_like(B)
            zero = 1
        else:
            X = _columns(x0, n, A.scalar)[0]
            X = X.clone() if backend.startswith("torch") else X.copy()
            zero = 0
        it = np.zeros(nrhs, dtype=np.int32)
        rs = np.zeros(nrhs)
        if backend.startswith("torch"):
            _torch_sync()
        st = fn(A._h, nrhs, _dptr(B, backend), _dptr(X, backend), zero, _ip(it), _dp(rs), mem, _stream_ptr(stream))
        self._info = _check(st, allow=(NOT_CONVERGED,))
        self._iters = it
        self._resid = rs
        X = X.t() if backend.startswith("torch") else X.T
        return X.reshape(-1) if squeeze else X


class GMRES(_Solver):
    """Restarted GMRES with FP64 Arnoldi (consumer: Eigen::GMRES)."""

    def __init__(self):
        super().__init__()
        self.restart = 30

    def set_restart(self, r):
        self.restart = int(r)

    def solve(self, b, x0=None, stream=None):
        L = library()

        def fn(h, nrhs, B, X, zero, it, rs, mem, s):
            return L.bmi_solve_gmres(h, nrhs, B, X, self.tol, self.restart, self.maxit, zero, it, rs, mem, s)

        return self._solve(b, x0, fn, stream)


class ConjugateGradient(_Solver):
    """CG for the real symmetric Laplace operator (consumer: Eigen::ConjugateGradient)."""

    def solve(self, b, x0=None, stream=None):
        L = library()

        def fn(h, nrhs, B, X, zero, it, rs, mem, s):
            return L.bmi_solve_cg(h, nrhs, B, X, self.tol, self.maxit, zero, it, rs, mem, s)

        return self._solve(b, x0, fn, stream)


# ---------------------------------------------------------------------------- serving
class Server:
    """Request loop on the GPU with continuous batching (bmi_server_*): submit() returns at once,
    result() waits and returns (solution, info).  Up to max_active batches are admitted; their
    GMRES solves run as tasks on separate streams driven by one device thread, so their work
    overlaps on the device (slice_iterations=0: one batch at a time, each to the end)."""

    def __init__(self, hbm_budget_bytes=None, max_active=None, max_batch_rhs=64, slice_iterations=None):
        cfg = ServerConfig()
        _check(library().bmi_server_config_default(C.byref(cfg)))
        if hbm_budget_bytes is not None:
            cfg.hbm_budget_bytes = float(hbm_budget_bytes)
        if max_active is not None:
            cfg.max_active = int(max_active)
        cfg.max_batch_rhs = int(max_batch_rhs)
        if slice_iterations is not None:
            cfg.slice_iterations = int(slice_iterations)
        h = C.c_void_p()
        _check(library().bmi_server_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self._lib = library()
        self._meta = {}

    def submit(self, operator, refinement_level, polynomial_degree, data, scale=(1.0, 1.0, 1.0),
               geometry=SPHERE_DAT, n_cheb=9, eta=1.6, min_cluster_level=1, precision=FP64, tol=1e-10,
               restart=30, maxit=5000):
        """data: DirichletTrace (Laplace) or PlaneWaves (Helmholtz)"""
        r = Request()
        r.geometry_path = geometry.encode()
        for k in range(3):
            r.scale[k] = float(scale[k])
        r.refinement_level = int(refinement_level)
        r.polynomial_degree = int(polynomial_degree)
        r.op = operator.desc(n_cheb=n_cheb, eta=eta, min_cluster_level=min_cluster_level, precision=precision)
        params = np.ascontiguousarray(data.params, dtype=np.float64).reshape(-1)
        r.data_fn = data.fn
        r.nrhs = data.nrhs
        r.data_params = _dp(params)
        r.data_kappa = data.kappa
        r.tol = float(tol)
        r.restart = int(restart)
        r.maxit = int(maxit)
        rid = C.c_int64()
        _check(library().bmi_server_submit(self._h, C.byref(r), C.byref(rid)))
        self._meta[rid.value] = (data.nrhs, operator.scalar)
        return rid.value

    def result(self, rid, timeout=-1.0):
        """(solution ndof x nrhs or None on failure, info dict); raises TimeoutError if pending"""
        info = ResultInfo()
        st = library().bmi_server_wait(self._h, rid, float(timeout), C.byref(info))
        if st == PENDING:
            raise TimeoutError("request %d still running" % rid)
        _check(st)
        nrhs, scalar = self._meta.pop(rid)
        out = dict(status=info.status, ndof=info.ndof, submit_s=info.submit_s, start_s=info.start_s,
                   end_s=info.end_s, latency_s=info.end_s - info.submit_s, iterations=info.iterations,
                   residual=info.residual, batch_requests=info.batch_requests, batch_rhs=info.batch_rhs,
                   cache_hit=bool(info.cache_hit), slices=info.slices)
        if info.status != OK and info.status != NOT_CONVERGED:
            out["error"] = library().bmi_last_error().decode()
            _check(library().bmi_server_fetch(self._h, rid, None))
            return None, out
        S = 2 if scalar == np.complex128 else 1
        x = np.zeros(S * info.ndof * nrhs)
        _check(library().bmi_server_fetch(self._h, rid, _dp(x)))
        X = (x.view(np.complex128) if S == 2 else x).reshape(nrhs, info.ndof).T
        return (X[:, 0] if nrhs == 1 else X), out

    def stats(self):
        v = np.zeros(6, dtype=np.int64)
        _check(library().bmi_server_stats(self._h, v.ctypes.data_as(C.POINTER(C.c_int64))))
        return dict(submitted=int(v[0]), finished=int(v[1]), queued=int(v[2]), running=int(v[3]),
                    resident_operators=int(v[4]), reserved_bytes=int(v[5]))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.bmi_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
