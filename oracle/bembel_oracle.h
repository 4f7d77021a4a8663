/*
 * bembel_oracle.h -- CPU restatement of Bembel's Galerkin assembly and H2 matvec.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity oracle and the timed CPU
 * baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it; the product library (bembel_amd/) never links or calls it.
 *
 * Every function restates the reference algorithm it names (file:line relative to
 * the Bembel source tree).  Floating-point expressions keep the reference's operand
 * order; the library is compiled with -ffp-contract=off so that the host-side
 * decisions (vertex dedup, enclosing balls, admissibility, quadrature degrees) are
 * deterministic.
 *
 * Parity status per function (what pins it):
 *   Gauss rules 1..20 points .... bit-exact vs the reference literals (golden fixture)
 *   Gauss rules 21..51 points ... parity unpinned bit-wise (reference uses an
 *                                 eigensolver); pinned only by exactness on polynomials
 *   Bernstein / patch eval ...... closed forms, sphere |x|=1, normal || x, FD Jacobian
 *   element tree / balls ........ counts, ids, neighbour counts, closed-form radii
 *   Duffy integrate0..4 ......... reference known-answer tests (closed-form integrals)
 *   Laplace/Helmholtz blocks .... brute-force adaptive quadrature on flat pairs,
 *                                 symmetry A = A^T, H2-vs-dense agreement
 *   Maxwell blocks .............. symmetry, kappa->scaling identities (partial)
 *   Chebyshev/transfer/moments .. the reference FMM identity tests (1e-12)
 *   H2 matvec ................... vs the dense Galerkin matrix (reference LazyEigenSum test)
 *   projector T ................. partition of unity, DOF counts, de Boor agreement
 *   solvers ..................... dense LU on small problems; closed-form solution
 */
#ifndef BEMBEL_ORACLE_H_
#define BEMBEL_ORACLE_H_
#ifdef __cplusplus
extern "C" {
#endif

/* operator kinds */
enum { OR_LAPLACE = 0, OR_HELMHOLTZ = 1, OR_MAXWELL = 2, OR_DUMMY = 3 };
/* test functions of the dummy operator (tests/DuffyTrick/test_testFunctions.hpp) */
enum { OR_DUMMY_ANALYTIC = 0, OR_DUMMY_NEARSINGULAR = 1, OR_DUMMY_ONE = 2 };

typedef struct or_geometry or_geometry;
typedef struct or_space or_space;
typedef struct or_h2 or_h2;

/* geometry (Geometry/GeometryIO.hpp:25-122, Geometry/Patch.hpp:63-104) */
or_geometry* or_geometry_load(const char* path, const double scale[3]);
void or_geometry_free(or_geometry* g);
int or_geometry_npatch(const or_geometry* g);
/* Patch::updateSurfacePoint (Patch.hpp:294-365): out[12] = xi(2), w, x(3), dsx(3), dtx(3) */
void or_patch_surface_point(const or_geometry* g, int patch, const double st[2],
                            double w, const double xi[2], double out[12]);
/* Patch::eval (Patch.hpp:115-155) */
void or_patch_eval(const or_geometry* g, int patch, double s, double t, double out[3]);
/* Bernstein basis of degree p and its derivative (Spline/Bernstein.hpp:197-244) */
void or_bernstein(int p, double x, double* out);
void or_bernstein_der(int p, double x, double* out);

/* discrete space: element tree + super space + transformation matrix.
   form: 0 = Discontinuous (Laplace/Helmholtz), 1 = DivConforming (Maxwell).
   (AnsatzSpace/AnsatzSpace.hpp:102-111, ClusterTree/ElementTree.hpp:105-161) */
or_space* or_space_create(or_geometry* g, int M, int P, int form);
void or_space_free(or_space* s);
/* info[0..7] = nel, npoints, I2, vd, nsuper, ndof, nnz(T), M */
void or_space_info(const or_space* s, int info[8]);
/* per element in Morton order: vertices[4], patch, level, llc[2], midpoint[3], radius */
void or_space_elements(const or_space* s, int* vert, int* patch, double* llc,
                       double* mid, double* radius);
/* cluster midpoints/radii of level l (6*4^l clusters for 6 patches) */
int or_space_clusters(const or_space* s, int level, double* mid, double* radius);
/* T in CSR (rows = super DOFs, cols = DOFs) */
void or_space_T(const or_space* s, int* rowptr, int* col, double* val);

/* compareElements + near-field degree (DuffyTrick/compareElements.hpp:23-69,
   LinearOperator/LinearOperatorBase.hpp:56-89, evaluateBilinearForm.hpp:21-40).
   out[4] = class, rot1, rot2, degree (max with far-field degree) */
void or_pair_info(const or_space* s, int e1, int e2, int out[4], double* dist);
/* getNearfieldQuadratureDegree of a single-layer operator (LinearOperatorBase.hpp:69-89),
   before the max with the far-field degree */
int or_nfield_degree(int P, double dist, int level);

/* evaluateBilinearForm for one element pair.  block: (vd*I2) x (vd*I2) row-major,
   rows = local functions of e1; complex kinds write interleaved re/im. */
void or_pair_block(const or_space* s, int kind, double kre, double kim, int dummy_fn,
                   int e1, int e2, double* block);
/* integrate0..4 driven directly with a chosen rule and rotations (DuffyTrick tests) */
void or_integrate(const or_space* s, int kind, double kre, double kim, int dummy_fn,
                  int which, int e1, int rot1, int e2, int rot2, int qdeg, double* block);

/* dense Galerkin matrices (LinearOperator/DiscreteOperator.hpp:27-87):
   super matrix (nsuper x nsuper) and T^T A T (ndof x ndof), column-major. */
void or_dense_super(const or_space* s, int kind, double kre, double kim, double* A);
void or_dense_galerkin(const or_space* s, int kind, double kre, double kim, double* A);

/* FMM pieces (H2Matrix/H2Multipole.hpp) */
void or_cheb_points(int n, double* x);
void or_transfer_matrix(int n, double* T /* n^2 x 4n^2, column-major */);
/* moment matrix of one component: (nc*n^2) x (4^cr*I2) column-major; comp selects the
   Maxwell derivative variant (0: plain, 1: [M;M_dx], 2: [M;M_dy]) */
int or_moment_matrix(const or_space* s, int cluster_level, int cluster_ref, int n, int comp,
                     double* out);
/* forward/backward transformation of an input matrix (H2Multipole.hpp:380-469) */
void or_forward_transform(int n2, int nc, const double* mom, int mom_cols, const double* T,
                          int steps, const double* X, int ncols, double** levels);

/* H2 matrix (H2Matrix/H2Matrix.hpp:93-250, BlockClusterTree.hpp:113-297) */
or_h2* or_h2_build(or_space* s, int kind, double kre, double kim, int np, double eta,
                   int min_cluster_level, int assemble);
void or_h2_free(or_h2* h);
/* info: nleaves, ndense, nlowrank, block_dim_dense, block_dim_lowrank, fmm_levels */
void or_h2_info(const or_h2* h, int info[6]);
/* leaves in reference DFS order: cc (1 LowRank, 2 Dense), level, id1, id2 */
void or_h2_leaves(const or_h2* h, int* leaves);
/* leaf matrix (component block rc for Maxwell), column-major, complex interleaved */
void or_h2_leaf_matrix(const or_h2* h, int leaf, int comp_r, int comp_c, double* F);
void or_h2_set_leaf_matrix(or_h2* h, int leaf, int comp_r, int comp_c, const double* F);
/* assemble only the listed leaves of an H2 built with assemble = 0 (sampled full-size checks) */
void or_h2_assemble_leaves(or_h2* h, int n, const int* leaves);
int or_h2_import_leaves(or_h2* h, int n, const int* keys, const double* data);
/* res = A * x for nrhs columns (H2DenseProduct.hpp:39-164), DOF space; complex interleaved */
void or_h2_matvec(const or_h2* h, int nrhs, const double* x, double* y);
/* same in the super space (skips T and T^T) */
void or_h2_matvec_super(const or_h2* h, int nrhs, const double* x, double* y);

/* right-hand sides (LinearForm/DiscreteLinearForm.hpp:44-83, DirichletTrace.hpp:31-70).
   fn: 0 point source 1/|x-x0| (p = x0), 1 plane wave exp(i kappa d.x) (p = d, kre),
   2 rotated tangential trace of a Hertz dipole (p = position[3], moment[3], kre;
   DivConforming spaces, LinearForm/RotatedTangentialTrace.hpp, examples/Data.hpp:76-90),
   out: ndof entries (real) or 2*ndof (complex) */
void or_rhs(const or_space* s, int fn, const double* p, double kre, double* out);
/* (continued) fn 3: Helmholtz fundamental solution exp(-i kappa |x-p|)/|x-p| (complex),
   fn 4: the harmonic polynomial 4x^2 - 3y^2 - z^2 (real; p unused) (examples/Data.hpp) */

/* DiscretePotential<*SingleLayerPotential, Op>::evaluate (Potential/DiscretePotential.hpp:49-103)
   of the Cauchy data rho (ndof, complex interleaved for Helmholtz/Maxwell) at npts points;
   out: npts x dim column-major (dim 3 for Maxwell), complex interleaved for Helmholtz/Maxwell */
void or_potential(const or_space* s, int kind, double kre, double kim, const double* rho, int npts,
                  const double* pts, double* out);

/* solvers (reference consumers are Eigen CG/GMRES; restated textbook algorithms) */
int or_cg(int n, const double* A, const double* b, double* x, double tol, int maxit);
int or_gmres_dense(int n, const double* A, const double* b, double* x, double tol,
                   int restart, int maxit);
int or_gmres_h2(const or_h2* h, const double* b, double* x, double tol, int restart,
                int maxit);

/* timing helper for the CPU baseline: assemble only the listed element pairs and
   coupling blocks (no storage) and return the wall time in seconds */
double or_time_assembly_sample(const or_h2* h, int npairs, const int* pairs, int nblocks,
                               const int* blocks, int nthreads);
double or_time_matvec(const or_h2* h, int nrhs, int reps, int nthreads);
/* OpenMP threads of the oracle's later parallel regions (or_h2_build, matvecs): CPU baseline */
void or_set_num_threads(int nthreads);

#ifdef __cplusplus
}
#endif
#endif
