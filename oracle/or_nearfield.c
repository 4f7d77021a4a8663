/* Oracle near field: operators, compareElements, quadrature degrees, integrate0..4,
 * evaluateBilinearForm and dense assembly.  Test infrastructure only. */
#include <omp.h>
#include <stdio.h>

#include "or_internal.h"

static int op_order(int kind) { return kind == OR_DUMMY ? 0 : -1; }
int or_op_ncomp(const or_op* op) { return op->kind == OR_MAXWELL ? 2 : 1; }
int or_op_is_complex(const or_op* op) {
  return op->kind == OR_HELMHOLTZ || op->kind == OR_MAXWELL;
}

/* LinearOperatorBase::get_FarfieldQuadratureDegree (LinearOperatorBase.hpp:56-58), SL */
int or_ffield_degree(int P) { return P - (-1) + 1; }
static int ffield_degree_kind(int kind, int P) { return P - op_order(kind) + 1; }

/* LinearOperatorBase::getNearfieldQuadratureDegree (LinearOperatorBase.hpp:69-89) */
static int nfield_degree_oo(int P, double dist, int level, int OO) {
  double dlog = (dist * (1 << level) < 1) ? -(level * log(2.)) : log(dist);
  int alpha = 2 - OO + 2 * P;
  double num = (alpha + P) * level * log(2.) - (2 - P + OO) * dlog;
  double den = (level + 2) * log(2.) + dlog;
  return (int)(0.5 * num / den);
}
int or_nfield_degree(int P, double dist, int level) { return nfield_degree_oo(P, dist, level, -1); }

/* DuffyTrick::compareElements (DuffyTrick/compareElements.hpp:23-69) */
void or_compare_elements(const or_space* s, int n1, int n2, int out[3], double* dist) {
  const or_node* e1 = &s->nodes[n1];
  const or_node* e2 = &s->nodes[n2];
  out[0] = out[1] = out[2] = 0;
  if (n1 == n2) {
    out[0] = 4, out[1] = 4, out[2] = 2;
    *dist = 0;
    return;
  }
  double d[3] = {e1->mid[0] - e2->mid[0], e1->mid[1] - e2->mid[1], e1->mid[2] - e2->mid[2]};
  double r = sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]) - e1->radius - e2->radius;
  *dist = r >= 0 ? r : 0;
  if (*dist > .5 / (1 << e1->level)) {
    out[0] = 4, out[1] = 4, out[2] = 0;
    return;
  }
  for (int r1 = 0; r1 < 4; ++r1)
    for (int r2 = 0; r2 < 4; ++r2)
      if (e1->vert[r1] == e2->vert[r2]) {
        if (e1->vert[3] == e2->vert[(r2 + 1) % 4]) {
          out[0] = 3, out[1] = r2, out[2] = 3;
        } else if (e1->vert[(r1 + 1) % 4] == e2->vert[(r2 + 3) % 4]) {
          out[0] = r1, out[1] = (r2 + 3) % 4, out[2] = 3;
        } else {
          out[0] = r1, out[1] = r2, out[2] = 4;
        }
        return;
      }
  out[0] = 4, out[1] = 4, out[2] = 0;
}

/* ------------------------------------------------------------ integrands */
static double norm_cross(const double* a, const double* b) {
  double c0 = a[1] * b[2] - a[2] * b[1];
  double c1 = a[2] * b[0] - a[0] * b[2];
  double c2 = a[0] * b[1] - a[1] * b[0];
  return sqrt((c0 * c0 + c1 * c1) + c2 * c2);
}
static double dot3(const double* a, const double* b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
static double dist3(const double* x, const double* y) {
  double d0 = x[0] - y[0], d1 = x[1] - y[1], d2 = x[2] - y[2];
  return sqrt((d0 * d0 + d1 * d1) + d2 * d2);
}
/* LaplaceSingleLayerOperator::evaluateKernel (Laplace/SingleLayerOperator.hpp:115-118) */
static double laplace_kernel(const double* x, const double* y) {
  return 1. / 4. / OR_PI / dist3(x, y);
}
/* HelmholtzSingleLayerOperator::evaluateKernel (Helmholtz/SingleLayerOperator.hpp:115-120),
   same in Maxwell/SingleLayerOperator.hpp:125-130: exp(-i kappa r)/(4 pi r) */
static double complex helmholtz_kernel(double complex kappa, const double* x, const double* y) {
  double r = dist3(x, y);
  double complex e = cexp(-OR_I * kappa * r);
  return e / 4. / OR_PI / r;
}

/* dummy test functions (tests/DuffyTrick/test_testFunctions.hpp:18-183) */
static double dummy_fn(int fn, const double* x, const double* y) {
  if (fn == OR_DUMMY_ANALYTIC)
    return sin(OR_PI * x[0]) * sin(2 * OR_PI * y[0]) * exp(-x[1]) * exp(y[1]);
  if (fn == OR_DUMMY_NEARSINGULAR) {
    double c = x[1] - y[0];
    return exp(-x[0] - 2 * y[1]) * (1. / (1e-3 + c * c));
  }
  return 1.;
}

/* tensor Bernstein values at xi, index iy*I+ix (Spline/Basis.hpp:57-164) */
static void tensor_basis(int P, const double* xi, double* out) {
  int I = P + 1;
  double X[32], Y[32];
  or_bernstein(P, xi[0], X);
  or_bernstein(P, xi[1], Y);
  for (int iy = 0; iy < I; ++iy)
    for (int ix = 0; ix < I; ++ix) out[iy * I + ix] = X[ix] * Y[iy];
}

/* evaluateIntegrand_impl of the operators, accumulating into a complex block of
   dimension (vd*I2)^2, row index = local function of p1. */
void or_eval_integrand(const or_op* op, const double* p1, const double* p2, double complex* blk) {
  int P = op->P, I = P + 1, I2 = op->I2;
  double X[32], Y[32];
  if (op->kind == OR_DUMMY) {
    blk[0] += dummy_fn(op->dummy_fn, p1 + 3, p2 + 3) * p1[2] * p2[2];
    return;
  }
  if (op->kind == OR_LAPLACE || op->kind == OR_HELMHOLTZ) {
    double xk = norm_cross(p1 + 6, p1 + 9);
    double yk = norm_cross(p2 + 6, p2 + 9);
    double complex integrand;
    if (op->kind == OR_LAPLACE)
      integrand = laplace_kernel(p1 + 3, p2 + 3) * xk * yk * p1[2] * p2[2];
    else
      integrand = helmholtz_kernel(op->kappa, p1 + 3, p2 + 3) * xk * yk * p1[2] * p2[2];
    /* Phi_times_Phi_ (Basis.hpp:141-164) */
    double complex a[1024];
    double b[1024];
    or_bernstein(P, p1[0], X);
    or_bernstein(P, p1[1], Y);
    for (int iy = 0; iy < I; ++iy)
      for (int ix = 0; ix < I; ++ix) a[iy * I + ix] = integrand * X[ix] * Y[iy];
    or_bernstein(P, p2[0], X);
    or_bernstein(P, p2[1], Y);
    for (int iy = 0; iy < I; ++iy)
      for (int ix = 0; ix < I; ++ix) b[iy * I + ix] = X[ix] * Y[iy];
    for (int i = 0; i < I2; ++i)
      for (int j = 0; j < I2; ++j) blk[i * I2 + j] += a[i] * b[j];
    return;
  }
  /* Maxwell (Maxwell/SingleLayerOperator.hpp:44-88) */
  int n = 2 * I2;
  double h = 1. / (1 << op->M);
  double complex kev = helmholtz_kernel(op->kappa, p1 + 3, p2 + 3) * p1[2] * p2[2];
  double complex ivec = kev;
  double complex idiv = -kev / op->kappa2 / h / h;
  /* addScaledVectorBasisInteraction (SuperSpace.hpp:236-258) */
  double bi[1024], ta[1024];
  tensor_basis(P, p1, ta);
  double tb[1024];
  tensor_basis(P, p2, tb);
  for (int i = 0; i < I2; ++i)
    for (int j = 0; j < I2; ++j) bi[i * I2 + j] = ta[i] * tb[j];
  const double *xdx = p1 + 6, *xdy = p1 + 9, *ydx = p2 + 6, *ydy = p2 + 9;
  double complex c00 = ivec * dot3(xdx, ydx), c01 = ivec * dot3(xdx, ydy);
  double complex c10 = ivec * dot3(xdy, ydx), c11 = ivec * dot3(xdy, ydy);
  for (int i = 0; i < I2; ++i)
    for (int j = 0; j < I2; ++j) {
      blk[i * n + j] += c00 * bi[i * I2 + j];
      blk[i * n + j + I2] += c01 * bi[i * I2 + j];
      blk[(i + I2) * n + j] += c10 * bi[i * I2 + j];
      blk[(i + I2) * n + j + I2] += c11 * bi[i * I2 + j];
    }
  /* Div_Phi_times_Div_Phi_ (Basis.hpp:170-200) */
  double dX[32], dY[32];
  double complex adx[1024], ady[1024];
  double bdx[1024], bdy[1024];
  or_bernstein(P, p1[0], X);
  or_bernstein(P, p1[1], Y);
  or_bernstein_der(P, p1[0], dX);
  or_bernstein_der(P, p1[1], dY);
  for (int iy = 0; iy < I; ++iy)
    for (int ix = 0; ix < I; ++ix) {
      adx[iy * I + ix] = 0. + idiv * dX[ix] * Y[iy];
      ady[iy * I + ix] = 0. + idiv * X[ix] * dY[iy];
    }
  or_bernstein(P, p2[0], X);
  or_bernstein(P, p2[1], Y);
  or_bernstein_der(P, p2[0], dX);
  or_bernstein_der(P, p2[1], dY);
  for (int iy = 0; iy < I; ++iy)
    for (int ix = 0; ix < I; ++ix) {
      bdx[iy * I + ix] = 0. + 1. * dX[ix] * Y[iy];
      bdy[iy * I + ix] = 0. + 1. * X[ix] * dY[iy];
    }
  for (int i = 0; i < I2; ++i)
    for (int j = 0; j < I2; ++j) blk[i * n + j] += adx[i] * bdx[j];
  for (int i = 0; i < I2; ++i)
    for (int j = 0; j < I2; ++j) blk[i * n + j + I2] += adx[i] * bdy[j];
  for (int i = 0; i < I2; ++i)
    for (int j = 0; j < I2; ++j) blk[(i + I2) * n + j] += ady[i] * bdx[j];
  for (int i = 0; i < I2; ++i)
    for (int j = 0; j < I2; ++j) blk[(i + I2) * n + j + I2] += ady[i] * bdy[j];
}

/* evaluateFMMInterpolation_impl (Laplace :87-110, Helmholtz :87-110, Maxwell :90-120);
   out is (vd*nc) x (vd*nc) row-major */
void or_fmm_interp(const or_op* op, const double* p1, const double* p2, double complex* out) {
  if (op->kind == OR_LAPLACE || op->kind == OR_DUMMY) {
    double xk = norm_cross(p1 + 6, p1 + 9), yk = norm_cross(p2 + 6, p2 + 9);
    out[0] = laplace_kernel(p1 + 3, p2 + 3) * xk * yk;
    return;
  }
  if (op->kind == OR_HELMHOLTZ) {
    double xk = norm_cross(p1 + 6, p1 + 9), yk = norm_cross(p2 + 6, p2 + 9);
    out[0] = helmholtz_kernel(op->kappa, p1 + 3, p2 + 3) * xk * yk;
    return;
  }
  double complex k = helmholtz_kernel(op->kappa, p1 + 3, p2 + 3);
  for (int i = 0; i < 16; ++i) out[i] = 0;
  out[0 * 4 + 0] = k * dot3(p1 + 6, p2 + 6);
  out[0 * 4 + 2] = k * dot3(p1 + 6, p2 + 9);
  out[2 * 4 + 0] = k * dot3(p1 + 9, p2 + 6);
  out[2 * 4 + 2] = k * dot3(p1 + 9, p2 + 9);
  double complex d = -k / op->kappa2;
  out[1 * 4 + 1] = d;
  out[1 * 4 + 3] = d;
  out[3 * 4 + 1] = d;
  out[3 * 4 + 3] = d;
}

/* ------------------------------------------------------------ Duffy rules */
/* DuffyTrick::tau (DuffyTrick/tau.hpp:26-42) */
static void tau(double x, double y, int c, double* out) {
  switch (c) {
    case 1:
      out[0] = 1 - y, out[1] = x;
      return;
    case 2:
      out[0] = 1 - x, out[1] = 1 - y;
      return;
    case 3:
      out[0] = y, out[1] = 1 - x;
      return;
    default:
      out[0] = x, out[1] = y;
  }
}

typedef struct {
  int n; /* points per direction */
  const double* x;
  const double* w;
} rule1d;

static rule1d get_rule(int deg) {
  rule1d r;
  r.n = deg + 1;
  r.x = or_gl_nodes(r.n);
  r.w = or_gl_weights(r.n);
  return r;
}
/* tensor point k of GaussSquare: (x[k/n], x[k%n]), weight w[k/n]*w[k%n] */
#define QX0(r, k) ((r).x[(k) / (r).n])
#define QX1(r, k) ((r).x[(k) % (r).n])
#define QW(r, k) ((r).w[(k) / (r).n] * (r).w[(k) % (r).n])

static void map(const or_space* s, int node, double a, double b, double w, double* out) {
  double xi[2] = {a, b};
  or_map2surface(s, node, xi, w, out);
}

/* integrate0 (DuffyTrick/integrate0.hpp:22-38) */
static void integrate0(const or_space* s, const or_op* op, int e1, int e2, double complex* blk) {
  int nq = s->nff;
  const double* f1 = s->ffield + (size_t)12 * nq * s->nodes[e1].id;
  const double* f2 = s->ffield + (size_t)12 * nq * s->nodes[e2].id;
  for (int i = 0; i < nq; ++i)
    for (int j = 0; j < nq; ++j) or_eval_integrand(op, f1 + 12 * i, f2 + 12 * j, blk);
}
/* integrate1 (DuffyTrick/integrate1.hpp:25-48) */
static void integrate1(const or_space* s, const or_op* op, int e1, int e2, rule1d r,
                       double complex* blk) {
  double h = 1. / (1 << s->nodes[e1].level);
  int nq = r.n * r.n;
  double q1[12], q2[12];
  for (int i = 0; i < nq; ++i) {
    map(s, e1, QX0(r, i), QX1(r, i), h * QW(r, i), q1);
    for (int j = 0; j < nq; ++j) {
      map(s, e2, QX0(r, j), QX1(r, j), h * QW(r, j), q2);
      or_eval_integrand(op, q1, q2, blk);
    }
  }
}
/* integrate2 (DuffyTrick/integrate2.hpp:27-67) */
static void integrate2(const or_space* s, const or_op* op, int e1, rule1d r, double complex* blk) {
  double h = 1. / (1 << s->nodes[e1].level);
  int nq = r.n * r.n;
  double q1[12], q2[12];
  for (int i = 0; i < nq; ++i) {
    double xi0 = QX0(r, i), xi1 = QX1(r, i);
    double w = h * h * QW(r, i) * xi0 * (1 - xi0) * (1 - xi0 * xi1);
    for (int j = 0; j < nq; ++j) {
      double xj0 = QX0(r, j), xj1 = QX1(r, j), wj = QW(r, j);
      double t1 = xj0 * (1 - xi0);
      double t2 = xj1 * (1 - xi0 * xi1);
      double t3 = xj0 * (1 - xi0) + xi0;
      double t4 = xj1 * (1 - xi0 * xi1) + xi0 * xi1;
      /* pts columns (row-major comma initialiser): (t1,t2),(t3,t4),(t1,t4),(t3,t2),
         (t2,t1),(t4,t3),(t2,t3),(t4,t1) */
      double pts[8][2] = {{t1, t2}, {t3, t4}, {t1, t4}, {t3, t2},
                          {t2, t1}, {t4, t3}, {t2, t3}, {t4, t1}};
      for (int k = 0; k < 4; ++k) {
        map(s, e1, pts[2 * k][0], pts[2 * k][1], w, q1);
        map(s, e1, pts[2 * k + 1][0], pts[2 * k + 1][1], wj, q2);
        or_eval_integrand(op, q1, q2, blk);
        or_eval_integrand(op, q2, q1, blk);
      }
    }
  }
}
/* integrate3 (DuffyTrick/integrate3.hpp:27-80) */
static void integrate3(const or_space* s, const or_op* op, int e1, int rot1, int e2, int rot2,
                       rule1d r, double complex* blk) {
  double h = 1. / (1 << s->nodes[e1].level);
  int nq = r.n * r.n;
  double q1[12], q2[12], a[2], b[2];
  for (int i = 0; i < nq; ++i) {
    double xi0 = QX0(r, i), xi1 = QX1(r, i);
    double w = h * h * QW(r, i) * xi1 * xi1;
    double t1 = xi0 * (1 - xi1);
    double t2 = (1 - xi0) * (1 - xi1);
    for (int j = 0; j < nq; ++j) {
      double eta0 = xi1 * QX0(r, j), eta1 = xi1 * QX1(r, j), wj = QW(r, j);
      double t3 = xi0 * (1 - eta0);
      double t4 = (1 - xi0) * (1 - eta0);
      double wa = wj * (1 - xi1), wb = wj * (1 - eta0);
      tau(t1, eta0, rot1, a), tau(t2, eta1, rot2, b);
      or_map2surface(s, e1, a, w, q1), or_map2surface(s, e2, b, wa, q2);
      or_eval_integrand(op, q1, q2, blk);
      tau(1 - t1, eta0, rot1, a), tau(1 - t2, eta1, rot2, b);
      or_map2surface(s, e1, a, w, q1), or_map2surface(s, e2, b, wa, q2);
      or_eval_integrand(op, q1, q2, blk);
      tau(t3, xi1, rot1, a), tau(t4, eta1, rot2, b);
      or_map2surface(s, e1, a, w, q1), or_map2surface(s, e2, b, wb, q2);
      or_eval_integrand(op, q1, q2, blk);
      tau(1 - t3, xi1, rot1, a), tau(1 - t4, eta1, rot2, b);
      or_map2surface(s, e1, a, w, q1), or_map2surface(s, e2, b, wb, q2);
      or_eval_integrand(op, q1, q2, blk);
      tau(t4, eta1, rot1, a), tau(t3, xi1, rot2, b);
      or_map2surface(s, e1, a, w, q1), or_map2surface(s, e2, b, wb, q2);
      or_eval_integrand(op, q1, q2, blk);
      tau(1 - t4, eta1, rot1, a), tau(1 - t3, xi1, rot2, b);
      or_map2surface(s, e1, a, w, q1), or_map2surface(s, e2, b, wb, q2);
      or_eval_integrand(op, q1, q2, blk);
    }
  }
}
/* integrate4 (DuffyTrick/integrate4.hpp:27-64) */
static void integrate4(const or_space* s, const or_op* op, int e1, int rot1, int e2, int rot2,
                       rule1d r, double complex* blk) {
  double h = 1. / (1 << s->nodes[e1].level);
  int nq = r.n * r.n;
  double q1[12], q2[12], q3[12], q4[12], q5[12], q6[12], a[2];
  for (int i = 0; i < nq; ++i) {
    double xi0 = QX0(r, i), xi1 = QX1(r, i);
    double w = h * h * QW(r, i) * pow(xi0, 3.);
    xi1 *= xi0;
    tau(xi0, xi1, rot1, a), or_map2surface(s, e1, a, w, q1);
    tau(xi1, xi0, rot1, a), or_map2surface(s, e1, a, w, q2);
    tau(xi0, xi1, rot2, a), or_map2surface(s, e2, a, w, q3);
    tau(xi1, xi0, rot2, a), or_map2surface(s, e2, a, w, q4);
    for (int j = 0; j < nq; ++j) {
      double eta0 = xi0 * QX0(r, j), eta1 = xi0 * QX1(r, j), wj = QW(r, j);
      tau(eta0, eta1, rot2, a), or_map2surface(s, e2, a, wj, q5);
      tau(eta0, eta1, rot1, a), or_map2surface(s, e1, a, wj, q6);
      or_eval_integrand(op, q1, q5, blk);
      or_eval_integrand(op, q2, q5, blk);
      or_eval_integrand(op, q6, q3, blk);
      or_eval_integrand(op, q6, q4, blk);
    }
  }
}

void or_integrate_n(const or_space* s, const or_op* op, int which, int n1, int rot1, int n2,
                    int rot2, int qdeg, double complex* blk) {
  rule1d r = get_rule(qdeg);
  switch (which) {
    case 0:
      integrate0(s, op, n1, n2, blk);
      break;
    case 1:
      integrate1(s, op, n1, n2, r, blk);
      break;
    case 2:
      integrate2(s, op, n1, r, blk);
      break;
    case 3:
      integrate3(s, op, n1, rot1, n2, rot2, r, blk);
      break;
    case 4:
      integrate4(s, op, n1, rot1, n2, rot2, r, blk);
      break;
  }
}

/* DuffyTrick::computeFfieldQnodes (DuffyTrick/farFieldQuadratureNodes.hpp:22-46) for the
   single-layer far-field degree P+2 */
void or_build_ffield(or_space* s) {
  int deg = or_ffield_degree(s->P);
  rule1d r = get_rule(deg);
  s->ffdeg = deg;
  s->nff = r.n * r.n;
  s->ffield = (double*)malloc(sizeof(double) * 12 * (size_t)s->nff * s->nel);
  double h = 1. / (1 << s->nodes[s->elem[0]].level);
  for (int e = 0; e < s->nel; ++e) {
    int node = s->elem[e];
    int id = s->nodes[node].id;
    for (int k = 0; k < s->nff; ++k)
      map(s, node, QX0(r, k), QX1(r, k), h * QW(r, k), s->ffield + 12 * ((size_t)id * s->nff + k));
  }
}

/* DuffyTrick::evaluateBilinearForm (DuffyTrick/evaluateBilinearForm.hpp:21-71) */
void or_eval_bilinear(const or_space* s, const or_op* op, int n1, int n2, double complex* blk) {
  int n = op->vd * op->I2;
  memset(blk, 0, sizeof(double complex) * n * n);
  double dist = 0;
  int ff = ffield_degree_kind(op->kind, op->P);
  int cp[3];
  or_compare_elements(s, n1, n2, cp, &dist);
  int nf = nfield_degree_oo(op->P, dist, s->nodes[n1].level, op_order(op->kind));
  nf = nf >= ff ? nf : ff;
  rule1d r = get_rule(nf);
  switch (cp[2]) {
    case 0:
      if (nf == ff && s->ffdeg == ff)
        integrate0(s, op, n1, n2, blk);
      else
        integrate1(s, op, n1, n2, r, blk);
      return;
    case 2:
      integrate2(s, op, n1, r, blk);
      return;
    case 3:
      integrate3(s, op, n1, cp[0], n2, cp[1], r, blk);
      return;
    case 4:
      integrate4(s, op, n1, cp[0], n2, cp[1], r, blk);
      return;
  }
}

static or_op make_op(const or_space* s, int kind, double kre, double kim, int dfn) {
  or_op op;
  op.kind = kind;
  op.kappa = kre + OR_I * kim;
  op.kappa2 = op.kappa * op.kappa;
  op.dummy_fn = dfn;
  op.P = s->P;
  op.I2 = s->I2;
  op.vd = s->vd;
  op.M = s->M;
  return op;
}

void or_pair_info(const or_space* s, int e1, int e2, int out[4], double* dist) {
  int n1 = s->elem[e1], n2 = s->elem[e2];
  int cp[3];
  or_compare_elements(s, n1, n2, cp, dist);
  int ff = or_ffield_degree(s->P);
  int nf = or_nfield_degree(s->P, *dist, s->nodes[n1].level);
  out[0] = cp[2];
  out[1] = cp[0];
  out[2] = cp[1];
  out[3] = nf >= ff ? nf : ff;
}

static void store_block(const or_op* op, const double complex* blk, int n, double* out) {
  if (or_op_is_complex(op)) {
    for (int i = 0; i < n * n; ++i) out[2 * i] = creal(blk[i]), out[2 * i + 1] = cimag(blk[i]);
  } else {
    for (int i = 0; i < n * n; ++i) out[i] = creal(blk[i]);
  }
}

void or_pair_block(const or_space* s, int kind, double kre, double kim, int dfn, int e1, int e2,
                   double* block) {
  or_op op = make_op(s, kind, kre, kim, dfn);
  int n = op.vd * op.I2;
  double complex* blk = (double complex*)malloc(sizeof(double complex) * n * n);
  or_eval_bilinear(s, &op, s->elem[e1], s->elem[e2], blk);
  store_block(&op, blk, n, block);
  free(blk);
}

void or_integrate(const or_space* s, int kind, double kre, double kim, int dfn, int which, int e1,
                  int rot1, int e2, int rot2, int qdeg, double* block) {
  or_op op = make_op(s, kind, kre, kim, dfn);
  int n = op.vd * op.I2;
  double complex* blk = (double complex*)calloc(n * n, sizeof(double complex));
  or_integrate_n(s, &op, which, s->elem[e1], rot1, s->elem[e2], rot2, qdeg, blk);
  store_block(&op, blk, n, block);
  free(blk);
}

/* DiscreteOperatorComputer<Matrix>::compute (LinearOperator/DiscreteOperator.hpp:31-86):
   super matrix, column-major, complex interleaved for complex kinds */
void or_dense_super(const or_space* s, int kind, double kre, double kim, double* A) {
  or_op op = make_op(s, kind, kre, kim, 0);
  int I2 = s->I2, vd = s->vd, n = vd * I2, N = s->nsuper, nel = s->nel;
  int cplx = or_op_is_complex(&op);
#pragma omp parallel for schedule(dynamic, 1)
  for (long pr = 0; pr < (long)nel * nel; ++pr) {
    int e1 = (int)(pr / nel), e2 = (int)(pr % nel);
    double complex blk[4 * 1024];
    or_eval_bilinear(s, &op, s->elem[e1], s->elem[e2], blk);
    for (int i = 0; i < vd; ++i)
      for (int j = 0; j < vd; ++j)
        for (int a = 0; a < I2; ++a)
          for (int b = 0; b < I2; ++b) {
            long row = (long)I2 * (i * nel + e1) + a, col = (long)I2 * (j * nel + e2) + b;
            double complex v = blk[(i * I2 + a) * n + j * I2 + b];
            if (cplx) {
              A[2 * (col * N + row)] = creal(v);
              A[2 * (col * N + row) + 1] = cimag(v);
            } else {
              A[col * N + row] = creal(v);
            }
          }
  }
}

/* projector^T * (A * projector) (DiscreteOperator.hpp:83-84) */
void or_dense_galerkin(const or_space* s, int kind, double kre, double kim, double* out) {
  int N = s->nsuper, D = s->ndof;
  int cplx = kind == OR_HELMHOLTZ || kind == OR_MAXWELL;
  int c = cplx ? 2 : 1;
  double* A = (double*)malloc(sizeof(double) * c * (size_t)N * N);
  or_dense_super(s, kind, kre, kim, A);
  /* AT = A * T  (N x D) */
  double* AT = (double*)calloc((size_t)c * N * D, sizeof(double));
  for (int r = 0; r < N; ++r)
    for (int k = s->rowptr[r]; k < s->rowptr[r + 1]; ++k) {
      int col = s->colidx[k];
      double v = s->vals[k];
      /* column col of AT += v * column r of A */
      for (int i = 0; i < N; ++i)
        for (int q = 0; q < c; ++q)
          AT[c * ((size_t)col * N + i) + q] += v * A[c * ((size_t)r * N + i) + q];
    }
  memset(out, 0, sizeof(double) * c * (size_t)D * D);
  for (int r = 0; r < N; ++r)
    for (int k = s->rowptr[r]; k < s->rowptr[r + 1]; ++k) {
      int row = s->colidx[k];
      double v = s->vals[k];
      for (int j = 0; j < D; ++j)
        for (int q = 0; q < c; ++q)
          out[c * ((size_t)j * D + row) + q] += v * AT[c * ((size_t)j * N + r) + q];
    }
  free(A);
  free(AT);
}
