/*
 * bembel_amd.h -- C ABI of the MI355X implementation of Bembel's Galerkin assembly and H2
 * matrix-vector product for the Laplace, Helmholtz and Maxwell single-layer operators.
 *
 * Every entry point takes plain pointers and sizes.  Complex data is interleaved (re, im)
 * double pairs.  Matrices are column-major.  Vectors are in the B-spline DOF order of the
 * reference (columns of the transformation matrix T), so they are interchangeable with
 * Bembel's Eigen vectors.  Functions return a bmi_status; bmi_last_error() gives the message
 * of the last failure on the calling thread.  No entry point aborts the process.
 *
 * Threading: calls on different handles may run concurrently; one handle is not reentrant.
 * Streams: a NULL stream is the default stream of the current device.
 */
#ifndef BEMBEL_AMD_H_
#define BEMBEL_AMD_H_
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  BMI_OK = 0,
  BMI_INVALID_ARG = 1,
  BMI_OUT_OF_MEMORY = 2,
  BMI_DEVICE_ERROR = 3,
  BMI_NOT_CONVERGED = 4,
  BMI_ADMISSION_REJECTED = 5,
  BMI_IO_ERROR = 6,
  BMI_UNSUPPORTED = 7,
  BMI_PENDING = 8 /* a served request did not finish within the wait time */
} bmi_status;

/* operator kinds: the single-layer operators of the reference
   (Laplace/SingleLayerOperator.hpp, Helmholtz/SingleLayerOperator.hpp,
   Maxwell/SingleLayerOperator.hpp) and the test operator of
   LinearOperator/Dummy/DummyOperator.hpp */
typedef enum {
  BMI_LAPLACE_SL = 0,
  BMI_HELMHOLTZ_SL = 1,
  BMI_MAXWELL_SL = 2,
  BMI_DUMMY = 3
} bmi_operator_kind;

/* storage / arithmetic of the far field (coupling and transfer) */
typedef enum {
  BMI_PRECISION_FP64 = 0,       /* everything FP64 */
  BMI_PRECISION_MIXED = 1,      /* reduced operator for products and inner solves: coupling blocks in FP32
                                   (one RHS: an FP32 copy of the near-field blocks with the coupling
                                   blocks read from their FP32 copy or, for the real Laplace kernel,
                                   recomputed in FP32 from the Chebyshev surface points, whichever the
                                   first product measures faster; from 8 RHS on: the FP32 near-field and
                                   coupling blocks widened to FP64 on the matrix cores; 2-7 RHS: FP32
                                   coupling blocks with the FP64 near field), sums in FP64; the FP64
                                   operator is kept and solvers refine with it */
  BMI_PRECISION_MIXED_BF16 = 2  /* as MIXED, and the coupling blocks of the two coarsest levels in BF16 */
} bmi_precision;

/* spaces: DifferentialForm of the operator (LinearOperator/DifferentialFormEnum.hpp) */
typedef enum { BMI_FORM_DISCONTINUOUS = 0, BMI_FORM_DIVCONFORMING = 1 } bmi_form;

typedef enum { BMI_MEM_HOST = 0, BMI_MEM_DEVICE = 1 } bmi_memory;

/* Operator descriptor.  Replaces the LinearOperatorTraits plug-in of the reference
   (LinearOperator/LinearOperatorTraits.hpp) and H2Matrix::init_H2Matrix's number_of_points
   (H2Matrix/H2Matrix.hpp:93-96) plus BlockClusterTree::set_parameters
   (H2Matrix/BlockClusterTree.hpp:113-118).  Reference defaults: n_cheb 9, eta 1.6,
   min_cluster_level 1. */
typedef struct {
  int kind;              /* bmi_operator_kind */
  double kappa_re;       /* wavenumber (Helmholtz/Maxwell set_wavenumber) */
  double kappa_im;
  int n_cheb;            /* Chebyshev nodes per direction of a cluster */
  double eta;            /* admissibility parameter */
  int min_cluster_level; /* leaves hold 4^min_cluster_level elements */
  int precision;         /* bmi_precision */
  int dummy_fn;          /* test function of BMI_DUMMY (0 analytic, 1 near-singular, 2 one) */
} bmi_op_desc;

typedef struct bmi_geometry bmi_geometry;
typedef struct bmi_space bmi_space;
typedef struct bmi_h2 bmi_h2;

/* ------------------------------------------------------------------ errors, device */
const char* bmi_last_error(void);
const char* bmi_version(void);
/* reference default descriptor for an operator kind */
bmi_status bmi_op_desc_default(int kind, bmi_op_desc* out);

/* ------------------------------------------------------------------ geometry */
/* Geometry(filename) for .dat files (Geometry/Geometry.hpp:37,89-102,
   Geometry/GeometryIO.hpp:25-122).  scale (may be NULL) multiplies the homogeneous
   x, y, z control coordinates, which maps the sphere file to an exact ellipsoid. */
bmi_status bmi_geometry_load(const char* path, const double* scale, bmi_geometry** out);
bmi_status bmi_geometry_destroy(bmi_geometry* g);
bmi_status bmi_geometry_num_patches(const bmi_geometry* g, int* n);
/* Patch::updateSurfacePoint (Geometry/Patch.hpp:294-365): out = xi(2) w x(3) dsx(3) dtx(3) */
bmi_status bmi_geometry_surface_point(const bmi_geometry* g, int patch, double s, double t,
                                      double out[12]);

/* ------------------------------------------------------------------ discrete space */
/* AnsatzSpace<Op>(geometry, refinement_level, polynomial_degree)
   (AnsatzSpace/AnsatzSpace.hpp:80-111).  form selects Discontinuous (Laplace, Helmholtz)
   or DivConforming (Maxwell). The geometry must outlive the space. */
bmi_status bmi_space_create(const bmi_geometry* g, int refinement_level, int polynomial_degree,
                            int form, bmi_space** out);
bmi_status bmi_space_destroy(bmi_space* s);
/* info: [0] elements [1] vertices [2] (P+1)^2 [3] vector dimension [4] super DOFs
   [5] DOFs (AnsatzSpace::get_number_of_dofs) [6] nnz(T) [7] refinement level */
bmi_status bmi_space_info(const bmi_space* s, int64_t info[8]);
/* element data in Morton order (ClusterTree/ElementTree.hpp): vertex ids (4 per element,
   numbered by first appearance), patch, lower-left corner, enclosing-ball midpoint/radius */
bmi_status bmi_space_elements(const bmi_space* s, int* vertices, int* patch, double* llc,
                              double* midpoint, double* radius);
/* cluster balls of one level (6*4^level clusters for a 6-patch geometry) */
bmi_status bmi_space_clusters(const bmi_space* s, int level, double* midpoint, double* radius,
                              int* count);
/* transformation matrix T (AnsatzSpace::get_transformation_matrix) in CSR */
bmi_status bmi_space_transformation(const bmi_space* s, int* rowptr, int* col, double* val);

/* ------------------------------------------------------------------ near field */
/* compareElements + quadrature degree of element pairs (DuffyTrick/compareElements.hpp:23-69,
   LinearOperator/LinearOperatorBase.hpp:56-89): out[4*k] = class, rot1, rot2, degree */
bmi_status bmi_pair_info(const bmi_space* s, int npairs, const int* pairs, int* out);
/* evaluateBilinearForm on the GPU for a list of element pairs (DuffyTrick/
   evaluateBilinearForm.hpp:21-71); blocks are (vd*I2)^2, row-major, rows = functions of
   the first element; complex kinds interleaved.  out is a host pointer. */
bmi_status bmi_pair_blocks(const bmi_space* s, const bmi_op_desc* op, int npairs,
                           const int* pairs, double* out, void* stream);
/* integrate0..4 driven directly with a given rule and rotations (the reference Duffy
   unit tests, tests/DuffyTrick/test_integrate*.hpp); job[5k..5k+4] = which, e1, rot1, e2,
   rot2.  Only the (0,0) entry of each block is returned (P = 0 test spaces). */
bmi_status bmi_integrate_jobs(const bmi_space* s, const bmi_op_desc* op, int njobs,
                              const int* jobs, int qdeg, double* out, void* stream);

/* ------------------------------------------------------------------ dense assembly */
/* DiscreteOperator<Eigen::MatrixX, Op>::compute (LinearOperator/DiscreteOperator.hpp:27-87):
   A = T^T A_super T (ndof x ndof), or the super matrix if super_only; column-major, out is
   a host or device pointer (mem). */
bmi_status bmi_dense_assemble(const bmi_space* s, const bmi_op_desc* op, int super_only,
                              double* out, int mem, void* stream);

/* ------------------------------------------------------------------ H2 */
/* DiscreteOperator<Eigen::H2Matrix<S>, Op>::compute -> H2Matrix::init_H2Matrix
   (LinearOperator/DiscreteOperator.hpp:92-102, H2Matrix/H2Matrix.hpp:93-250).
   All leaves are assembled on the device and stay resident. The space must outlive it. */
bmi_status bmi_h2_assemble(const bmi_space* s, const bmi_op_desc* op, void* stream,
                           bmi_h2** out);
/* device bytes an H2 assembly of `op` on `s` will hold, including the workspace of an
   nrhs-column product (block tree built on the host, nothing allocated).  No counterpart in the
   reference; used for admission by device-memory budget. */
bmi_status bmi_h2_estimate_bytes(const bmi_space* s, const bmi_op_desc* op, int nrhs, int64_t* bytes);
bmi_status bmi_h2_destroy(bmi_h2* h);
/* info: [0] leaves [1] dense leaves [2] low-rank leaves [3] near-field pairs
   [4] dense leaf dimension [5] coupling dimension [6] FMM levels [7] rows=cols (DOFs)
   [8] device bytes held [9] complex (0/1) [10] precision */
bmi_status bmi_h2_info(const bmi_h2* h, int64_t info[16]);
/* leaves (cc 1 low rank, 2 dense; level; row cluster id; column cluster id) in the
   storage order of this implementation */
bmi_status bmi_h2_leaves(const bmi_h2* h, int* leaves);
/* copy one leaf matrix (component block r,c) to the host, column-major */
bmi_status bmi_h2_leaf_matrix(const bmi_h2* h, int leaf, int comp_r, int comp_c, double* F);
/* all leaf matrices to the host in the order of bmi_h2_leaves: per leaf vd*vd component
   blocks (row component major), each column-major; dense leaves first.  data holds
   (dense * dd^2 + lowrank * cd^2) * vd^2 scalars (doubles, or interleaved complex). */
bmi_status bmi_h2_export_leaves(const bmi_h2* h, double* data);
/* Y = alpha * H * X + beta * Y for nrhs columns (the product of
   H2Matrix/EigenHelper/H2DenseProduct.hpp:33-165 and the Eigen scaleAndAddTo glue
   :268-290).  X, Y column-major with leading dimensions ldx, ldy (in scalars); alpha, beta
   point to (re, im) pairs; mem says whether X and Y are host or device pointers.  With
   device operands the product is enqueued on `stream` and the call returns without waiting;
   with host operands it returns when Y holds the result. */
bmi_status bmi_h2_matvec(const bmi_h2* h, int nrhs, const double* X, int64_t ldx, double* Y,
                         int64_t ldy, const double* alpha, const double* beta, int mem,
                         void* stream);
/* Replay the super-space part of products with two or more columns from a captured HIP graph
   (one per RHS count, default on) or launch it kernel by kernel; single-column products always
   launch kernel by kernel (a graph's completion delays the next kernel on the stream by more
   than its few launches save).  No counterpart in the reference (its product is CPU code);
   exposed so callers can compare both modes. */
bmi_status bmi_h2_use_graphs(bmi_h2* h, int enable);
/* Sweep of single-RHS products: -1 (default) lets the first product time the forms and keep
   the fastest, 0 reads every block, 1 reads only the upper coupling blocks of a symmetric kernel
   (Laplace, Helmholtz) and forms the transposed products from them, 2 does the same for the
   near-field leaves as well (BMI_UNSUPPORTED where a form is not available).  No counterpart in
   the reference (its product is CPU code); exposed so callers can compare the forms. */
bmi_status bmi_h2_set_sweep(bmi_h2* h, int mode);
/* Single-RHS products on the handle's workspace: -1 (default) lets the first product time both
   forms and keep the faster, 0 runs the near-field leaves and coupling blocks in one launch after
   the forward transformation, 1 runs the near-field leaves on the caller's stream while the
   coupling blocks and the backward transformation run on a side stream (both forms give the same
   bits).  No counterpart in the reference; exposed so callers can
   compare the forms. */
bmi_status bmi_h2_set_overlap(bmi_h2* h, int mode);
/* wait for the library's work on `stream` (NULL: the default stream) */
bmi_status bmi_synchronize(void* stream);
/* Allocation-free form of bmi_h2_matvec for callers that cannot allocate device memory (code
   under HIP graph capture, tight serving loops); no counterpart in the reference.  Device X and Y
   only; the product is enqueued on `stream` and the call returns without waiting.  The product's
   workspace (super vectors and multipole coefficients: 8 bytes x (3 x super DOFs + (vd + 1) x
   multipole rows) x nrhs, doubled for complex operators) lives in one fixed scratch array of
   BMI_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; larger products are refused
   with BMI_INVALID_ARG.  Every call on a GPU shares that GPU's array: calls that overlap on one
   GPU (several streams or threads) race and must be serialised by the caller.  Every kernel runs
   on `stream` (no side stream, no graph, no in-library timing); results equal bmi_h2_matvec's to
   rounding (the same kernels, on the variants bmi_h2_matvec timed if it has run). */
#define BMI_NOALLOC_SCRATCH_BYTES (64ull << 20)
bmi_status bmi_h2_matvec_noalloc(const bmi_h2* h, int nrhs, const double* X, int64_t ldx, double* Y,
                                 int64_t ldy, const double* alpha, const double* beta, void* stream);
/* same operator applied in the super space (no T, no T^T); device pointers only */
bmi_status bmi_h2_matvec_super(const bmi_h2* h, int nrhs, const double* Xs, double* Ys,
                               void* stream);

/* ------------------------------------------------------------------ right-hand sides */
/* DiscreteLinearForm<DirichletTrace>::compute (LinearForm/DiscreteLinearForm.hpp:44-83,
   LinearForm/DirichletTrace.hpp:31-70) for built-in data functions:
   fn 0: g(x) = 1/|x - p_k|          (p_k = params[3k..3k+2]; real, one column per k)
   fn 1: g(x) = exp(i kappa d_k . x) (d_k = params[3k..3k+2]; complex, one column per k)
   fn 2: rotated tangential trace of a Hertz dipole field for the EFIE
         (LinearForm/RotatedTangentialTrace.hpp:39-82 with examples/Data.hpp:76-90 Dipole):
         params[6k..6k+5] = position, moment; DivConforming spaces; complex
   fn 3: g(x) = exp(-i kappa |x - p_k|) / |x - p_k| (Data::HelmholtzFundamentalSolution,
         examples/Data.hpp:37-44; p_k = params[3k..3k+2]; complex)
   fn 4: g(x) = 4 x^2 - 3 y^2 - z^2 (Data::HarmonicFunction, examples/Data.hpp:14-16; real;
         params: 3 ignored doubles per column)
   out: ndof x nrhs, host or device (mem) */
bmi_status bmi_rhs_assemble(const bmi_space* s, int fn, const double* params, int nrhs,
                            double kappa, double* out, int mem, void* stream);

/* The same linear forms for data the caller evaluates (a DirichletTrace or RotatedTangentialTrace
   with an arbitrary function, LinearForm/DirichletTrace.hpp:31-70,
   LinearForm/RotatedTangentialTrace.hpp:30-87).  bmi_linear_form_points gives the points of the
   (P+2)^2-point rule on every element (point k = element * (P+2)^2 + q, 3 doubles each; host
   code, no device use; out may be NULL to query *npoints).  bmi_rhs_from_values takes the data
   at those points, column by column: kind 0 real scalars, 1 complex scalars (interleaved),
   2 complex 3-vectors (6 doubles per point, DivConforming spaces: the rotated tangential trace
   is formed on the device).  values are host or device pointers (mem), like out. */
bmi_status bmi_linear_form_points(const bmi_space* s, int64_t* npoints, double* out);
bmi_status bmi_rhs_from_values(const bmi_space* s, int kind, const double* values, int nrhs,
                               double* out, int mem, void* stream);

/* ------------------------------------------------------------------ potentials */
/* DiscretePotential<{Laplace,Helmholtz,Maxwell}SingleLayerPotential, Op>::evaluate
   (Potential/DiscretePotential.hpp:49-103; Laplace/SingleLayerPotential.hpp,
   Helmholtz/SingleLayerPotential.hpp, Maxwell/SingleLayerPotential.hpp): the single-layer
   potential of the Cauchy data rho (ndof entries, complex interleaved for Helmholtz and Maxwell)
   at npoints points (3 doubles each), with the (P+2)^2-point rule per element.  op gives the
   kind and the wavenumber.  out: npoints x dim column-major (dim 1, or 3 for Maxwell), complex
   interleaved for Helmholtz and Maxwell.  rho, points and out are host or device pointers (mem). */
bmi_status bmi_potential_evaluate(const bmi_space* s, const bmi_op_desc* op, const double* rho,
                                  int npoints, const double* points, double* out, int mem,
                                  void* stream);

/* ------------------------------------------------------------------ solvers */
/* restarted GMRES (consumer: Eigen::GMRES in examples/HelmholtzSingleLayerH2.cpp:82-85)
   with FP64 Arnoldi, one solve per column; X holds the initial guess (zeroed if
   zero_guess).  iters/resid receive per-column values (may be NULL). */
bmi_status bmi_solve_gmres(const bmi_h2* h, int nrhs, const double* B, double* X, double tol,
                           int restart, int maxit, int zero_guess, int* iters, double* resid,
                           int mem, void* stream);
/* conjugate gradients for the real symmetric Laplace operator
   (consumer: Eigen::ConjugateGradient in examples/LaplaceSingleLayerH2.cpp:78-81) */
bmi_status bmi_solve_cg(const bmi_h2* h, int nrhs, const double* B, double* X, double tol,
                        int maxit, int zero_guess, int* iters, double* resid, int mem,
                        void* stream);

/* The GMRES of bmi_solve_gmres as a task that one thread can interleave with others (no
   counterpart in the reference).  start enqueues the first work on `stream` (device B and X);
   poll enqueues the next piece once the convergence flags of the previous one have arrived and
   returns without waiting on the device (*done: 0 flags not there yet, nothing enqueued;
   1 the task is finished; 2 work enqueued); finish waits,
   writes iters/resid (may be NULL) and frees the task, returning BMI_OK or BMI_NOT_CONVERGED.
   Tasks on different streams overlap on the device.  A handle's product workspace is shared:
   run one task per H2 handle at a time.  The device work, and so the result, equals
   bmi_solve_gmres with the same arguments. */
typedef struct bmi_gmres_task bmi_gmres_task;
bmi_status bmi_gmres_start(const bmi_h2* h, int nrhs, const double* B, double* X, double tol,
                           int restart, int maxit, int zero_guess, void* stream,
                           bmi_gmres_task** out);
bmi_status bmi_gmres_poll(bmi_gmres_task* task, int* done);
bmi_status bmi_gmres_finish(bmi_gmres_task* task, int* iters, double* resid);
/* a non-blocking stream of this library's HIP runtime (default priority); NULL destroys nothing */
bmi_status bmi_stream_create(void** out);
bmi_status bmi_stream_destroy(void* stream);

/* ------------------------------------------------------------------ serving */
/* A request loop on one GPU (no counterpart in the reference, which solves one problem per
   program).  Submitted solves are queued; a scheduler admits them (smallest first, anything
   older than 5 s before all else) while fewer than max_active batches are admitted and the
   device-memory budget holds the batch's workspace plus, for an operator that is not resident
   yet, its estimated H2 state (idle resident operators are evicted least recently used first).
   Queued requests with equal operator and solver settings are merged into one batched solve of
   up to max_batch_rhs right-hand sides.  One device thread drives every admitted batch: a
   GMRES batch runs as a task on its own stream (bmi_gmres_start/poll), enqueueing its next work
   whenever its convergence flags arrive, so the kernels of several batches are in flight on the
   device together and a short solve does not wait for a long one to finish; CG batches run to
   the end in their first turn.  Assembled operators stay resident between requests. */
typedef struct bmi_server bmi_server;
typedef struct {
  double hbm_budget_bytes; /* device memory the server may hold */
  int max_active;          /* batches admitted at once (their work overlaps on the device) */
  int max_batch_rhs;       /* right-hand sides merged into one solve */
  int slice_iterations;    /* > 0: GMRES batches run as overlapping tasks; 0: one batch at a time,
                              each to the end */
} bmi_server_config;
typedef struct {
  const char* geometry_path; /* .dat file (GeometryIO.hpp:25-122) */
  double scale[3];           /* per-axis scaling of the control points (0 = 1) */
  int refinement_level, polynomial_degree;
  bmi_op_desc op;            /* Laplace or Helmholtz single layer */
  int data_fn;               /* 0: point sources (Laplace), 1: plane waves (Helmholtz) */
  int nrhs;
  const double* data_params; /* 3 per right-hand side: source point or direction */
  double data_kappa;         /* plane-wave wavenumber */
  double tol;                /* relative residual */
  int restart, maxit;        /* GMRES restart (Helmholtz), iteration limit */
} bmi_request;
typedef struct {
  bmi_status status;         /* of the solve itself */
  int64_t ndof;
  int nrhs;
  double submit_s, start_s, end_s; /* seconds since the server started */
  int iterations;            /* largest over the request's columns */
  double residual;           /* largest relative residual over the request's columns */
  int batch_requests, batch_rhs; /* size of the batch the request was solved in */
  int cache_hit;             /* the operator was resident when the batch started */
  int slices;                /* device turns in which the batch enqueued work */
} bmi_result_info;
bmi_status bmi_server_config_default(bmi_server_config* cfg);
bmi_status bmi_server_create(const bmi_server_config* cfg, bmi_server** out);
/* queue a request (returns at once); *id names it in wait/fetch */
bmi_status bmi_server_submit(bmi_server* srv, const bmi_request* req, int64_t* id);
/* wait up to timeout_s seconds (< 0: no limit) for the request; BMI_PENDING if still running */
bmi_status bmi_server_wait(bmi_server* srv, int64_t id, double timeout_s, bmi_result_info* info);
/* copy the solution (ndof x nrhs, column-major, complex interleaved for Helmholtz) and release
   the request */
bmi_status bmi_server_fetch(bmi_server* srv, int64_t id, double* x);
/* submitted, finished, queued, running batches, resident operators, reserved device bytes */
bmi_status bmi_server_stats(bmi_server* srv, int64_t out[6]);
/* waits for running batches, then frees everything */
bmi_status bmi_server_destroy(bmi_server* srv);

/* ------------------------------------------------------------------ measurement */
/* per-kernel timing with HIP events recorded on the stream each kernel is launched on; timed
   launches do not wait for their events (they queue as untimed ones do), bmi_timing_get waits for
   and reads the events recorded so far */
bmi_status bmi_timing_enable(int enable);
bmi_status bmi_timing_reset(void);
/* total milliseconds and launch count of a kernel family ("nf_duffy", "nf_regular",
   "coupling", "matvec_nearfield", ...) since the last reset */
bmi_status bmi_timing_get(const char* name, double* total_ms, int64_t* launches);

#ifdef __cplusplus
}
#endif
#endif
