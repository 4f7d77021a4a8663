/* Oracle H2: block-cluster tree, interpolation FMM, matvec, RHS, solvers, timing.
 * Test infrastructure only. */
#include <omp.h>
#include <stdio.h>

#include "or_internal.h"

/* ------------------------------------------------------------ Chebyshev / Lagrange */
/* ChebychevRoots (H2Matrix/H2Multipole.hpp:20-32): ascending nodes in [0,1] */
void or_cheb_points(int n, double* x) {
  double alpha = OR_PI / (2. * (double)n);
  for (int i = 0; i < n; ++i) {
    double g = (double)(n - i);
    x[i] = 0.5 * cos(alpha * (2 * g - 1)) + 0.5;
  }
}
/* computeLagrangePolynomials (H2Multipole.hpp:39-49): Newton coefficients, L[k*n+i] = L(k,i) */
static void lagrange(int n, const double* x, double* L) {
  for (int k = 0; k < n; ++k)
    for (int i = 0; i < n; ++i) L[k * n + i] = k == i ? 1. : 0.;
  for (int i = 0; i < n; ++i)
    for (int j = 1; j < n; ++j)
      for (int k = n - 1; k >= j; --k)
        L[k * n + i] = (L[k * n + i] - L[(k - 1) * n + i]) / (x[k] - x[k - j]);
}
/* evaluatePolynomial (H2Multipole.hpp:55-67) for column i of L */
static double eval_poly(int n, const double* x, const double* L, int i, double xi) {
  double r = L[(n - 1) * n + i];
  for (int k = n - 2; k >= 0; --k) r = r * (xi - x[k]) + L[k * n + i];
  return r;
}

/* computeTransferMatrices (H2Multipole.hpp:76-115), column-major n^2 x 4n^2 */
void or_transfer_matrix(int n, double* T) {
  int n2 = n * n;
  double x[64], L[4096], E[64 * 128];
  or_cheb_points(n, x);
  lagrange(n, x, L);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      E[i * 2 * n + j] = eval_poly(n, x, L, j, 0.5 * x[i]);
      E[i * 2 * n + j + n] = eval_poly(n, x, L, j, 0.5 * x[i] + 0.5);
    }
  static const int perm[4] = {0, 3, 1, 2};
  for (int k = 0; k < 4; ++k)
    for (int i = 0; i < n; ++i)
      for (int ii = 0; ii < n; ++ii)
        for (int j = 0; j < n; ++j)
          for (int jj = 0; jj < n; ++jj) {
            int row = j * n + jj, col = i * n + ii + n2 * perm[k];
            T[(size_t)col * n2 + row] =
                E[i * 2 * n + j + (k / 2) * n] * E[ii * 2 * n + jj + (k % 2) * n];
          }
}

/* Moment1D / Moment1DDerivative (H2Multipole.hpp:120-218): n x (2^cr*(P+1)), row-major */
static void moment1d(int P, int cl, int cr, int np, int deriv, double* out) {
  int n = 1 << cr;
  double h = 1. / n;
  int N = 1 << cl;
  double H = 1. / N;
  int pp1 = P + 1;
  int qi = (int)ceil(0.5 * (np + P - 2));
  int nq = qi + 1;
  const double* qx = or_gl_nodes(nq);
  const double* qw = or_gl_weights(nq);
  double x[64], L[4096];
  or_cheb_points(np, x);
  lagrange(np, x, L);
  int cols = n * pp1;
  for (int i = 0; i < np; ++i)
    for (int j = 0; j < n; ++j) {
      double acc[32] = {0};
      for (int k = 0; k < nq; ++k) {
        double base[32];
        double w;
        if (!deriv) {
          w = qw[k] * sqrt(h * H) * eval_poly(np, x, L, i, h * (j + qx[k]));
          or_bernstein(P, qx[k], base);
        } else {
          w = qw[k] / sqrt(h * H) * eval_poly(np, x, L, i, h * (j + qx[k]));
          or_bernstein_der(P, qx[k], base);
        }
        for (int m = 0; m < pp1; ++m) acc[m] += w * base[m];
      }
      for (int m = 0; m < pp1; ++m) out[i * cols + j * pp1 + m] = acc[m];
    }
}

/* moment2DComputer (H2Multipole.hpp:224-293): np^2 x (4^cr*I2), column-major with leading
   dimension ld (rows offset by roff) */
static void moment2d(int P, int cl, int cr, int np, int d1, int d2, double* out, int ld,
                     int roff) {
  int n = 1 << cr, n2 = n * n, pp1 = P + 1, I2 = pp1 * pp1;
  double* m1 = (double*)malloc(sizeof(double) * np * n * pp1);
  double* m2 = (double*)malloc(sizeof(double) * np * n * pp1);
  moment1d(P, cl, cr, np, d1, m1);
  moment1d(P, cl, cr, np, d2, m2);
  int* is = (int*)calloc(n2, sizeof(int));
  int* it = (int*)calloc(n2, sizeof(int));
  for (int j = 0; j < cr; ++j) {
    int inc = 1 << (2 * j);
    for (int i = 0; i < inc; ++i) {
      int s = is[i], t = it[i];
      is[4 * i] = 2 * s, is[4 * i + 1] = 2 * s + 1, is[4 * i + 2] = 2 * s + 1, is[4 * i + 3] = 2 * s;
      it[4 * i] = 2 * t, it[4 * i + 1] = 2 * t, it[4 * i + 2] = 2 * t + 1, it[4 * i + 3] = 2 * t + 1;
    }
  }
  int cols = n * pp1;
  for (int i = 0; i < np; ++i)
    for (int j = 0; j < np; ++j)
      for (int k = 0; k < n2; ++k)
        for (int a = 0; a < pp1; ++a)
          for (int b = 0; b < pp1; ++b) {
            int row = i * np + j + roff, col = I2 * k + a * pp1 + b;
            out[(size_t)col * ld + row] = m1[i * cols + is[k] * pp1 + b] * m2[j * cols + it[k] * pp1 + a];
          }
  free(m1), free(m2), free(is), free(it);
}

int or_moment_matrix(const or_space* s, int cl, int cr, int np, int comp, double* out) {
  int np2 = np * np;
  if (comp == 0) {
    moment2d(s->P, cl, cr, np, 0, 0, out, np2, 0);
    return np2;
  }
  /* Maxwell: [M; M_dx] (comp 1) or [M; M_dy] (comp 2) (Maxwell/SingleLayerOperator.hpp:153-183) */
  moment2d(s->P, cl, cr, np, 0, 0, out, 2 * np2, 0);
  if (comp == 1)
    moment2d(s->P, cl, cr, np, 1, 0, out, 2 * np2, np2);
  else
    moment2d(s->P, cl, cr, np, 0, 1, out, 2 * np2, np2);
  return 2 * np2;
}

/* ------------------------------------------------------------ H2 structure */
typedef struct {
  int cc, level, n1, n2, id1, id2;
} or_leaf;

struct or_h2 {
  or_space* s;
  or_op op;
  int cplx, np, np2, nc, vd, mcl, steps, max_level;
  double eta;
  int nleaves, cap;
  or_leaf* leaves;
  double** F; /* F[leaf*vd*vd + r*vd + c]: column-major, complex interleaved if cplx */
  int* dim;   /* rows = cols of F per leaf */
  double* T;  /* transfer np2 x 4np2 */
  double* mom[2];
  int mom_rows, mom_cols;
  double* nodes2d; /* np2 x 2 */
};

/* BlockClusterTree::compareCluster (H2Matrix/BlockClusterTree.hpp:255-278) */
static int compare_cluster(const or_h2* h, const or_node* c1, const or_node* c2) {
  double d[3] = {c1->mid[0] - c2->mid[0], c1->mid[1] - c2->mid[1], c1->mid[2] - c2->mid[2]};
  double dist = sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]) - c1->radius - c2->radius;
  dist = dist > 0 ? dist : 0;
  double mr = c1->radius > c2->radius ? c1->radius : c2->radius;
  int r = (dist < 0 || mr >= h->eta * dist) ? 0 : 1;
  if (r == 0 && h->max_level - c1->level <= h->mcl) r = 2;
  return r;
}

/* appendSubtree + updateLeafPointers_recursion (BlockClusterTree.hpp:154-182, 279-297):
   leaves in DFS order with the column son index outermost */
static void append_subtree(or_h2* h, int n1, int n2) {
  const or_space* s = h->s;
  const or_node* c1 = &s->nodes[n1];
  const or_node* c2 = &s->nodes[n2];
  int cc = compare_cluster(h, c1, c2);
  if (cc == 0) {
    for (int j = 0; j < c2->nsons; ++j)
      for (int i = 0; i < c1->nsons; ++i) append_subtree(h, s->nodes[n1].sons + i, s->nodes[n2].sons + j);
    return;
  }
  if (h->nleaves == h->cap) {
    h->cap = h->cap ? 2 * h->cap : 1024;
    h->leaves = (or_leaf*)realloc(h->leaves, sizeof(or_leaf) * h->cap);
  }
  or_leaf* l = &h->leaves[h->nleaves++];
  l->cc = cc;
  l->level = c1->level;
  l->n1 = n1;
  l->n2 = n2;
  l->id1 = c1->id;
  l->id2 = c2->id;
}

static int cluster_nel(const or_space* s, int level) { return 1 << (2 * (s->M - level)); }

/* interpolateKernel (H2Multipole.hpp:336-375): F column-major (vd*nc*np2)^2 */
static void interpolate_kernel(const or_h2* h, int n1, int n2, double complex* F) {
  int np2 = h->np2, vd = h->vd, nc = h->nc, dim = vd * nc * np2;
  double q1[12], q2[12];
  double complex out[16];
  for (int i = 0; i < np2; ++i) {
    or_map2surface(h->s, n1, h->nodes2d + 2 * i, 1., q1);
    for (int j = 0; j < np2; ++j) {
      or_map2surface(h->s, n2, h->nodes2d + 2 * j, 1., q2);
      or_fmm_interp(&h->op, q1, q2, out);
      for (int ii = 0; ii < vd; ++ii)
        for (int jj = 0; jj < vd; ++jj)
          for (int a = 0; a < nc; ++a)
            for (int b = 0; b < nc; ++b) {
              int row = i + (ii * nc + a) * np2, col = j + (jj * nc + b) * np2;
              F[(size_t)col * dim + row] = out[(a + ii * nc) * (vd * nc) + (b + jj * nc)];
            }
    }
  }
}

static void assemble_leaf(or_h2* h, int li) {
  const or_space* s = h->s;
  or_leaf* l = &h->leaves[li];
  int vd = h->vd, c = h->cplx ? 2 : 1;
  if (l->cc == 2) {
    int I2 = s->I2, ne = cluster_nel(s, l->level), dim = ne * I2, nb = vd * I2;
    for (int q = 0; q < vd * vd; ++q) h->F[li * vd * vd + q] = (double*)malloc(sizeof(double) * c * dim * dim);
    h->dim[li] = dim;
    double complex* blk = (double complex*)malloc(sizeof(double complex) * nb * nb);
    int f1 = l->id1 * ne, f2 = l->id2 * ne;
    for (int a = 0; a < ne; ++a)
      for (int b = 0; b < ne; ++b) {
        or_eval_bilinear(s, &h->op, s->elem[f1 + a], s->elem[f2 + b], blk);
        for (int r = 0; r < vd; ++r)
          for (int cc = 0; cc < vd; ++cc) {
            double* F = h->F[li * vd * vd + r * vd + cc];
            for (int i = 0; i < I2; ++i)
              for (int j = 0; j < I2; ++j) {
                size_t idx = (size_t)(b * I2 + j) * dim + a * I2 + i;
                double complex v = blk[(r * I2 + i) * nb + cc * I2 + j];
                if (c == 2)
                  F[2 * idx] = creal(v), F[2 * idx + 1] = cimag(v);
                else
                  F[idx] = creal(v);
              }
          }
      }
    free(blk);
  } else {
    int sub = h->nc * h->np2, dim = vd * sub;
    double complex* F = (double complex*)malloc(sizeof(double complex) * dim * dim);
    interpolate_kernel(h, l->n1, l->n2, F);
    h->dim[li] = sub;
    for (int r = 0; r < vd; ++r)
      for (int cc = 0; cc < vd; ++cc) {
        double* G = (double*)malloc(sizeof(double) * c * sub * sub);
        for (int j = 0; j < sub; ++j)
          for (int i = 0; i < sub; ++i) {
            double complex v = F[(size_t)(cc * sub + j) * dim + r * sub + i];
            if (c == 2)
              G[2 * (j * sub + i)] = creal(v), G[2 * (j * sub + i) + 1] = cimag(v);
            else
              G[j * sub + i] = creal(v);
          }
        h->F[li * vd * vd + r * vd + cc] = G;
      }
    free(F);
  }
}

or_h2* or_h2_build(or_space* s, int kind, double kre, double kim, int np, double eta, int mcl,
                   int assemble) {
  or_h2* h = (or_h2*)calloc(1, sizeof(or_h2));
  h->s = s;
  h->op.kind = kind;
  h->op.kappa = kre + OR_I * kim;
  h->op.kappa2 = h->op.kappa * h->op.kappa;
  h->op.dummy_fn = 0;
  h->op.P = s->P;
  h->op.I2 = s->I2;
  h->op.vd = s->vd;
  h->op.M = s->M;
  h->cplx = or_op_is_complex(&h->op);
  h->np = np;
  h->np2 = np * np;
  h->nc = or_op_ncomp(&h->op);
  h->vd = s->vd;
  h->mcl = mcl;
  h->eta = eta;
  h->max_level = s->max_level;
  append_subtree(h, s->root, s->root);
  /* H2Matrix::init_H2Matrix (H2Matrix.hpp:93-137) */
  int cl = h->max_level - mcl;
  if (cl < 0) cl = 0;
  int cr = mcl;
  if (cr > h->max_level) cr = h->max_level;
  h->steps = h->max_level - mcl;
  if (h->steps < 0) h->steps = 0;
  h->T = (double*)malloc(sizeof(double) * h->np2 * 4 * h->np2);
  or_transfer_matrix(np, h->T);
  h->mom_cols = (1 << (2 * cr)) * s->I2;
  h->mom_rows = h->nc * h->np2;
  for (int comp = 0; comp < h->vd; ++comp) {
    h->mom[comp] = (double*)malloc(sizeof(double) * h->mom_rows * h->mom_cols);
    or_moment_matrix(s, cl, cr, np, h->vd == 2 ? comp + 1 : 0, h->mom[comp]);
  }
  double x[64];
  or_cheb_points(np, x);
  h->nodes2d = (double*)malloc(sizeof(double) * 2 * h->np2);
  for (int i = 0; i < np; ++i)
    for (int j = 0; j < np; ++j) h->nodes2d[2 * (i * np + j)] = x[i], h->nodes2d[2 * (i * np + j) + 1] = x[j];
  h->F = (double**)calloc((size_t)h->nleaves * h->vd * h->vd, sizeof(double*));
  h->dim = (int*)calloc(h->nleaves, sizeof(int));
  for (int li = 0; li < h->nleaves; ++li)
    h->dim[li] = h->leaves[li].cc == 2 ? cluster_nel(s, h->leaves[li].level) * s->I2 : h->nc * h->np2;
  if (assemble) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int li = 0; li < h->nleaves; ++li) assemble_leaf(h, li);
  }
  return h;
}

void or_h2_free(or_h2* h) {
  if (!h) return;
  for (long i = 0; i < (long)h->nleaves * h->vd * h->vd; ++i) free(h->F[i]);
  free(h->F);
  free(h->dim);
  free(h->leaves);
  free(h->T);
  free(h->mom[0]);
  free(h->mom[1]);
  free(h->nodes2d);
  free(h);
}

void or_h2_info(const or_h2* h, int info[6]) {
  int nd = 0, nl = 0;
  for (int i = 0; i < h->nleaves; ++i) h->leaves[i].cc == 2 ? ++nd : ++nl;
  info[0] = h->nleaves;
  info[1] = nd;
  info[2] = nl;
  info[3] = cluster_nel(h->s, h->max_level - h->mcl < 0 ? 0 : h->max_level - h->mcl) * h->s->I2;
  info[4] = h->nc * h->np2;
  info[5] = h->steps + 1;
}

void or_h2_leaves(const or_h2* h, int* out) {
  for (int i = 0; i < h->nleaves; ++i) {
    out[4 * i] = h->leaves[i].cc;
    out[4 * i + 1] = h->leaves[i].level;
    out[4 * i + 2] = h->leaves[i].id1;
    out[4 * i + 3] = h->leaves[i].id2;
  }
}

void or_h2_leaf_matrix(const or_h2* h, int leaf, int r, int c, double* F) {
  int d = h->dim[leaf], m = h->cplx ? 2 : 1;
  const double* src = h->F[leaf * h->vd * h->vd + r * h->vd + c];
  if (src)
    memcpy(F, src, sizeof(double) * m * d * d);
  else
    memset(F, 0, sizeof(double) * m * d * d);
}

void or_h2_assemble_leaves(or_h2* h, int n, const int* leaves) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int k = 0; k < n; ++k) {
    int li = leaves[k];
    if (li >= 0 && li < h->nleaves && !h->F[(size_t)li * h->vd * h->vd]) assemble_leaf(h, li);
  }
}

void or_h2_set_leaf_matrix(or_h2* h, int leaf, int r, int c, const double* F) {
  int d = h->dim[leaf], m = h->cplx ? 2 : 1;
  double** dst = &h->F[leaf * h->vd * h->vd + r * h->vd + c];
  if (!*dst) *dst = (double*)malloc(sizeof(double) * m * d * d);
  memcpy(*dst, F, sizeof(double) * m * d * d);
}

/* ------------------------------------------------------------ matvec */
/* y(rows) += A(rows x cols, column-major) * x, real or complex interleaved */
static void gemv(int cplx, int rows, int cols, const double* A, const double* x, double* y) {
  if (!cplx) {
    for (int j = 0; j < cols; ++j) {
      double xj = x[j];
      const double* a = A + (size_t)j * rows;
      for (int i = 0; i < rows; ++i) y[i] += a[i] * xj;
    }
  } else {
    for (int j = 0; j < cols; ++j) {
      double xr = x[2 * j], xi = x[2 * j + 1];
      const double* a = A + (size_t)2 * j * rows;
      for (int i = 0; i < rows; ++i) {
        y[2 * i] += a[2 * i] * xr - a[2 * i + 1] * xi;
        y[2 * i + 1] += a[2 * i] * xi + a[2 * i + 1] * xr;
      }
    }
  }
}
/* y(rows) += A^T x with real A (rows_A x cols_A), x complex or real */
static void gemv_real_t(int cplx, int ra, int ca, const double* A, const double* x, double* y) {
  int c = cplx ? 2 : 1;
  for (int j = 0; j < ca; ++j)
    for (int q = 0; q < c; ++q) {
      double acc = 0;
      for (int i = 0; i < ra; ++i) acc += A[(size_t)j * ra + i] * x[c * i + q];
      y[c * j + q] += acc;
    }
}
static void gemv_real(int cplx, int ra, int ca, const double* A, const double* x, double* y) {
  int c = cplx ? 2 : 1;
  for (int j = 0; j < ca; ++j)
    for (int q = 0; q < c; ++q) {
      double xj = x[c * j + q];
      for (int i = 0; i < ra; ++i) y[c * i + q] += A[(size_t)j * ra + i] * xj;
    }
}

/* forwardTransformation (H2Multipole.hpp:380-422); levels[k] has (nc*np2) x ncl_k entries */
static void forward(const or_h2* h, const double* mom, const double* xs, double** lv, int nleaf) {
  int c = h->cplx ? 2 : 1, rows = h->mom_rows, np2 = h->np2, nc = h->nc;
  memset(lv[0], 0, sizeof(double) * c * rows * nleaf);
  for (int col = 0; col < nleaf; ++col)
    gemv_real(h->cplx, rows, h->mom_cols, mom, xs + (size_t)c * col * h->mom_cols, lv[0] + (size_t)c * col * rows);
  int ncl = nleaf;
  for (int st = 0; st < h->steps; ++st) {
    int nn = ncl / 4;
    memset(lv[st + 1], 0, sizeof(double) * c * rows * nn);
    for (int col = 0; col < nn; ++col)
      for (int comp = 0; comp < nc; ++comp) {
        /* to_transfer = child blocks (j + child*nc)*np2 of the reshaped column */
        double buf[2 * 4 * 1024];
        for (int ch = 0; ch < 4; ++ch)
          for (int r = 0; r < np2; ++r)
            for (int q = 0; q < c; ++q)
              buf[c * (ch * np2 + r) + q] = lv[st][(size_t)c * ((4 * col + ch) * rows + comp * np2 + r) + q];
        gemv_real(h->cplx, np2, 4 * np2, h->T, buf, lv[st + 1] + (size_t)c * (col * rows + comp * np2));
      }
    ncl = nn;
  }
}

/* backwardTransformation (H2Multipole.hpp:428-469): consumes bw, adds to y (super) */
static void backward(const or_h2* h, const double* mom, double** bw, int nleaf, double* y) {
  int c = h->cplx ? 2 : 1, rows = h->mom_rows, np2 = h->np2, nc = h->nc;
  int* ncl = (int*)malloc(sizeof(int) * (h->steps + 1));
  ncl[0] = nleaf;
  for (int k = 1; k <= h->steps; ++k) ncl[k] = ncl[k - 1] / 4;
  for (int st = h->steps; st > 0; --st) {
    for (int col = 0; col < ncl[st]; ++col)
      for (int comp = 0; comp < nc; ++comp) {
        double prod[2 * 4 * 1024];
        memset(prod, 0, sizeof(double) * c * 4 * np2);
        gemv_real_t(h->cplx, np2, 4 * np2, h->T, bw[st] + (size_t)c * (col * rows + comp * np2), prod);
        for (int ch = 0; ch < 4; ++ch)
          for (int r = 0; r < np2; ++r)
            for (int q = 0; q < c; ++q)
              bw[st - 1][(size_t)c * ((4 * col + ch) * rows + comp * np2 + r) + q] += prod[c * (ch * np2 + r) + q];
      }
  }
  for (int col = 0; col < nleaf; ++col)
    gemv_real_t(h->cplx, rows, h->mom_cols, mom, bw[0] + (size_t)c * col * rows, y + (size_t)c * col * h->mom_cols);
  free(ncl);
}

/* leaf sweep + transforms for one (row, col) component block in the super space
   (H2DenseProduct.hpp:60-158) */
static void matvec_component(const or_h2* h, int rc, int cc, const double* xs, double* ys) {
  const or_space* s = h->s;
  int c = h->cplx ? 2 : 1, rows = h->mom_rows;
  int vcs = s->nel * s->I2;
  int nleaf = vcs / h->mom_cols;
  double** fw = (double**)malloc(sizeof(double*) * (h->steps + 1));
  int n = nleaf;
  for (int k = 0; k <= h->steps; ++k, n /= 4) fw[k] = (double*)malloc(sizeof(double) * c * rows * (n > 0 ? n : 1));
  forward(h, h->mom[cc], xs, fw, nleaf);
#pragma omp parallel
  {
    double* my = (double*)calloc((size_t)c * vcs, sizeof(double));
    double** bw = (double**)malloc(sizeof(double*) * (h->steps + 1));
    int nn = nleaf;
    for (int k = 0; k <= h->steps; ++k, nn /= 4) bw[k] = (double*)calloc((size_t)c * rows * (nn > 0 ? nn : 1), sizeof(double));
#pragma omp for schedule(dynamic, 16)
    for (int li = 0; li < h->nleaves; ++li) {
      const or_leaf* l = &h->leaves[li];
      const double* F = h->F[li * h->vd * h->vd + rc * h->vd + cc];
      if (!F) continue;
      if (l->cc == 2) {
        int ne = cluster_nel(s, l->level), d = ne * s->I2;
        gemv(h->cplx, d, d, F, xs + (size_t)c * l->id2 * d, my + (size_t)c * l->id1 * d);
      } else {
        int lvl = h->max_level - h->mcl - l->level;
        gemv(h->cplx, rows, rows, F, fw[lvl] + (size_t)c * l->id2 * rows, bw[lvl] + (size_t)c * l->id1 * rows);
      }
    }
    backward(h, h->mom[rc], bw, nleaf, my);
#pragma omp critical
    for (size_t i = 0; i < (size_t)c * vcs; ++i) ys[i] += my[i];
    for (int k = 0; k <= h->steps; ++k) free(bw[k]);
    free(bw);
    free(my);
  }
  for (int k = 0; k <= h->steps; ++k) free(fw[k]);
  free(fw);
}

static void super_apply(const or_h2* h, const double* xs, double* ys) {
  int c = h->cplx ? 2 : 1;
  size_t vcs = (size_t)h->s->nel * h->s->I2;
  memset(ys, 0, sizeof(double) * c * vcs * h->vd);
  for (int cc = 0; cc < h->vd; ++cc)
    for (int rc = 0; rc < h->vd; ++rc) matvec_component(h, rc, cc, xs + c * vcs * cc, ys + c * vcs * rc);
}

void or_h2_matvec_super(const or_h2* h, int nrhs, const double* x, double* y) {
  int c = h->cplx ? 2 : 1;
  size_t N = h->s->nsuper;
  for (int k = 0; k < nrhs; ++k) super_apply(h, x + c * N * k, y + c * N * k);
}

/* res += T^T (A_super (T x)) (H2DenseProduct.hpp:50-163) with res initialised to zero */
void or_h2_matvec(const or_h2* h, int nrhs, const double* x, double* y) {
  const or_space* s = h->s;
  int c = h->cplx ? 2 : 1;
  size_t N = s->nsuper, D = s->ndof;
  double* xs = (double*)malloc(sizeof(double) * c * N);
  double* ys = (double*)malloc(sizeof(double) * c * N);
  for (int k = 0; k < nrhs; ++k) {
    const double* xk = x + c * D * k;
    double* yk = y + c * D * k;
    for (size_t r = 0; r < N; ++r)
      for (int q = 0; q < c; ++q) {
        double acc = 0;
        for (int p = s->rowptr[r]; p < s->rowptr[r + 1]; ++p) acc += s->vals[p] * xk[c * s->colidx[p] + q];
        xs[c * r + q] = acc;
      }
    super_apply(h, xs, ys);
    memset(yk, 0, sizeof(double) * c * D);
    for (size_t r = 0; r < N; ++r)
      for (int p = s->rowptr[r]; p < s->rowptr[r + 1]; ++p)
        for (int q = 0; q < c; ++q) yk[c * s->colidx[p] + q] += s->vals[p] * ys[c * r + q];
  }
  free(xs);
  free(ys);
}

/* ------------------------------------------------------------ RHS */
/* DiscreteLinearForm::compute with DirichletTrace (LinearForm/DiscreteLinearForm.hpp:44-83,
   DirichletTrace.hpp:38-66).  fn 0: 1/|x-p|, fn 1: exp(i kre p.x). */
/* Data::Dipole (examples/Data.hpp:76-90): Hertz dipole at pos with moment len */
static void dipole_field(const double* x, double kappa, const double* pos, const double* len,
                         double complex* E) {
  double c[3] = {x[0] - pos[0], x[1] - pos[1], x[2] - pos[2]};
  double r = sqrt((c[0] * c[0] + c[1] * c[1]) + c[2] * c[2]);
  double n[3] = {c[0] / r, c[1] / r, c[2] / r};
  double complex expc = cexp(-OR_I * kappa * r);
  /* (n x len) x n and 3 (n . len) n - len */
  double nxl[3] = {n[1] * len[2] - n[2] * len[1], n[2] * len[0] - n[0] * len[2], n[0] * len[1] - n[1] * len[0]};
  double tr[3] = {nxl[1] * n[2] - nxl[2] * n[1], nxl[2] * n[0] - nxl[0] * n[2], nxl[0] * n[1] - nxl[1] * n[0]};
  double nl = (n[0] * len[0] + n[1] * len[1]) + n[2] * len[2];
  double complex a = kappa * kappa * expc / r;
  double complex ec = ((1 / (r * r * r)) - (OR_I * (-kappa) / (r * r))) * expc;
  for (int k = 0; k < 3; ++k) E[k] = a * tr[k] + ec * (3 * nl * n[k] - len[k]);
}

/* DiscreteLinearForm<RotatedTangentialTrace> (LinearForm/RotatedTangentialTrace.hpp:39-82) of
   the dipole field: per element (I2 x 2) block, column c = t_c . (f - (f.n) n) w phi */
static void rhs_rotated_tangential(const or_space* s, const double* p, double kappa, double* out) {
  int P = s->P, I = P + 1, I2 = s->I2, nq1 = P + 2, nq = nq1 * nq1;
  const double* qx = or_gl_nodes(nq1);
  const double* qw = or_gl_weights(nq1);
  size_t N = s->nsuper, vcs = (size_t)s->nel * I2;
  double* sup = (double*)calloc(2 * N, sizeof(double));
  for (int e = 0; e < s->nel; ++e) {
    int node = s->elem[e];
    double h = 1. / (1 << s->nodes[node].level);
    double complex acc[2][1024];
    for (int m = 0; m < I2; ++m) acc[0][m] = acc[1][m] = 0;
    for (int i = 0; i < nq; ++i) {
      double xi[2] = {qx[i / nq1], qx[i % nq1]}, q[12];
      or_map2surface(s, node, xi, h * (qw[i / nq1] * qw[i % nq1]), q);
      double c0 = q[7] * q[11] - q[8] * q[10], c1 = q[8] * q[9] - q[6] * q[11], c2 = q[6] * q[10] - q[7] * q[9];
      double nn = sqrt((c0 * c0 + c1 * c1) + c2 * c2);
      double n[3] = {c0 / nn, c1 / nn, c2 / nn};
      double complex E[3];
      dipole_field(q + 3, kappa, p, p + 3, E);
      double complex fn_ = (E[0] * n[0] + E[1] * n[1]) + E[2] * n[2];
      double complex t[3];
      for (int k = 0; k < 3; ++k) t[k] = (E[k] - fn_ * n[k]) * q[2];
      double complex cx = (q[6] * t[0] + q[7] * t[1]) + q[8] * t[2];
      double complex cy = (q[9] * t[0] + q[10] * t[1]) + q[11] * t[2];
      double X[32], Y[32];
      or_bernstein(P, xi[0], X);
      or_bernstein(P, xi[1], Y);
      for (int iy = 0; iy < I; ++iy)
        for (int ix = 0; ix < I; ++ix) {
          acc[0][iy * I + ix] += cx * (X[ix] * Y[iy]);
          acc[1][iy * I + ix] += cy * (X[ix] * Y[iy]);
        }
    }
    int id = s->nodes[node].id;
    for (int c = 0; c < 2; ++c)
      for (int m = 0; m < I2; ++m) {
        size_t r = c * vcs + (size_t)id * I2 + m;
        sup[2 * r] = creal(acc[c][m]), sup[2 * r + 1] = cimag(acc[c][m]);
      }
  }
  memset(out, 0, sizeof(double) * 2 * s->ndof);
  for (size_t r = 0; r < N; ++r)
    for (int k = s->rowptr[r]; k < s->rowptr[r + 1]; ++k)
      for (int q = 0; q < 2; ++q) out[2 * s->colidx[k] + q] += s->vals[k] * sup[2 * r + q];
  free(sup);
}

void or_rhs(const or_space* s, int fn, const double* p, double kre, double* out) {
  if (fn == 2) {
    rhs_rotated_tangential(s, p, kre, out);
    return;
  }
  int P = s->P, I = P + 1, I2 = s->I2, deg = P + 1, nq1 = deg + 1, nq = nq1 * nq1;
  int cplx = fn == 1 || fn == 3;
  int c = cplx ? 2 : 1;
  const double* qx = or_gl_nodes(nq1);
  const double* qw = or_gl_weights(nq1);
  size_t N = s->nsuper;
  double* sup = (double*)calloc(c * N, sizeof(double));
  for (int e = 0; e < s->nel; ++e) {
    int node = s->elem[e];
    double h = 1. / (1 << s->nodes[node].level);
    double complex acc[1024];
    for (int m = 0; m < I2; ++m) acc[m] = 0;
    for (int i = 0; i < nq; ++i) {
      double xi[2] = {qx[i / nq1], qx[i % nq1]}, q[12];
      or_map2surface(s, node, xi, h * (qw[i / nq1] * qw[i % nq1]), q);
      double c0 = q[7] * q[11] - q[8] * q[10], c1 = q[8] * q[9] - q[6] * q[11], c2 = q[6] * q[10] - q[7] * q[9];
      double kap = sqrt((c0 * c0 + c1 * c1) + c2 * c2);
      double complex f;
      if (fn == 0) {
        double d0 = q[3] - p[0], d1 = q[4] - p[1], d2 = q[5] - p[2];
        f = 1. / sqrt((d0 * d0 + d1 * d1) + d2 * d2);
      } else if (fn == 1) {
        double ph = kre * ((p[0] * q[3] + p[1] * q[4]) + p[2] * q[5]);
        f = cos(ph) + OR_I * sin(ph);
      } else if (fn == 3) {
        /* Data::HelmholtzFundamentalSolution (examples/Data.hpp:37-44): exp(-i k r)/r */
        double d0 = q[3] - p[0], d1 = q[4] - p[1], d2 = q[5] - p[2];
        double r = sqrt((d0 * d0 + d1 * d1) + d2 * d2);
        f = cexp(-OR_I * kre * r) / r;
      } else {
        /* Data::HarmonicFunction (examples/Data.hpp:14-16) */
        f = 4 * q[3] * q[3] - 3 * q[4] * q[4] - q[5] * q[5];
      }
      double complex integrand = f * kap * q[2];
      double X[32], Y[32];
      or_bernstein(P, xi[0], X);
      or_bernstein(P, xi[1], Y);
      for (int iy = 0; iy < I; ++iy)
        for (int ix = 0; ix < I; ++ix) acc[iy * I + ix] += integrand * X[ix] * Y[iy];
    }
    int id = s->nodes[node].id;
    for (int m = 0; m < I2; ++m) {
      if (cplx)
        sup[2 * ((size_t)id * I2 + m)] = creal(acc[m]), sup[2 * ((size_t)id * I2 + m) + 1] = cimag(acc[m]);
      else
        sup[(size_t)id * I2 + m] = creal(acc[m]);
    }
  }
  memset(out, 0, sizeof(double) * c * s->ndof);
  for (size_t r = 0; r < (size_t)s->nel * I2; ++r)
    for (int k = s->rowptr[r]; k < s->rowptr[r + 1]; ++k)
      for (int q = 0; q < c; ++q) out[c * s->colidx[k] + q] += s->vals[k] * sup[c * r + q];
  free(sup);
}

/* ------------------------------------------------------------ potentials */
/* DiscretePotential<*SingleLayerPotential, Op>::evaluate (Potential/DiscretePotential.hpp:49-103):
   rule GS[P+1] on every element, weight h^2 w (map2surface), Cauchy data through
   FunctionEvaluator (AnsatzSpace/FunctionEvaluatorEval.hpp: value c^T phi / h, vector value
   (dsx c_s + dtx c_t) / h, divergence (phi_dx c_s + phi_dy c_t) / h^2).  Integrands:
   Laplace/Helmholtz K(x,y) rho(y) |dsx x dtx| w (Laplace/SingleLayerPotential.hpp:44-78,
   Helmholtz/SingleLayerPotential.hpp:44-78); Maxwell (K j(y) + grad_x K div(y) / kappa^2) w
   without the surface measure (Maxwell/SingleLayerPotential.hpp:44-104).
   out: npts x dim column-major (dim 1 or 3), complex interleaved for Helmholtz/Maxwell. */
void or_potential(const or_space* s, int kind, double kre, double kim, const double* rho, int npts,
                  const double* pts, double* out) {
  int P = s->P, I = P + 1, I2 = s->I2, nq1 = P + 2, nq = nq1 * nq1;
  int cplx = kind != OR_LAPLACE, vd = s->vd, dim = kind == OR_MAXWELL ? 3 : 1;
  double complex kappa = kre + OR_I * kim, kappa2 = kappa * kappa;
  const double* qx = or_gl_nodes(nq1);
  const double* qw = or_gl_weights(nq1);
  size_t N = s->nsuper;
  /* super coefficients T rho */
  double complex* sup = (double complex*)calloc(N, sizeof(double complex));
  for (size_t r = 0; r < N; ++r)
    for (int k = s->rowptr[r]; k < s->rowptr[r + 1]; ++k) {
      int j = s->colidx[k];
      sup[r] += s->vals[k] * (cplx ? rho[2 * j] + OR_I * rho[2 * j + 1] : rho[j]);
    }
  size_t vcs = (size_t)s->nel * I2;
  double complex* acc = (double complex*)calloc((size_t)npts * dim, sizeof(double complex));
  for (int e = 0; e < s->nel; ++e) {
    int node = s->elem[e];
    double h = 1. / (1 << s->nodes[node].level);
    int id = s->nodes[node].id;
    for (int i = 0; i < nq; ++i) {
      double xi[2] = {qx[i / nq1], qx[i % nq1]}, q[12];
      or_map2surface(s, node, xi, h * h * (qw[i / nq1] * qw[i % nq1]), q);
      double X[32], Y[32], DX[32], DY[32];
      or_bernstein(P, xi[0], X);
      or_bernstein(P, xi[1], Y);
      or_bernstein_der(P, xi[0], DX);
      or_bernstein_der(P, xi[1], DY);
      double complex v0 = 0, v1 = 0, dv = 0;
      for (int iy = 0; iy < I; ++iy)
        for (int ix = 0; ix < I; ++ix) {
          int m = iy * I + ix;
          double complex c0 = sup[(size_t)id * I2 + m];
          double complex c1 = vd > 1 ? sup[vcs + (size_t)id * I2 + m] : 0;
          v0 += c0 * (X[ix] * Y[iy]);
          v1 += c1 * (X[ix] * Y[iy]);
          dv += c0 * (DX[ix] * Y[iy]) + c1 * (X[ix] * DY[iy]);
        }
      double ws = q[2];
      for (int t = 0; t < npts; ++t) {
        const double* x = pts + 3 * t;
        double d0 = x[0] - q[3], d1 = x[1] - q[4], d2 = x[2] - q[5];
        double r = sqrt((d0 * d0 + d1 * d1) + d2 * d2);
        if (kind == OR_LAPLACE) {
          double c0 = q[7] * q[11] - q[8] * q[10], c1 = q[8] * q[9] - q[6] * q[11], c2 = q[6] * q[10] - q[7] * q[9];
          double kap = sqrt((c0 * c0 + c1 * c1) + c2 * c2);
          acc[t] += 1. / 4. / OR_PI / r * (v0 / h) * kap * ws;
        } else if (kind == OR_HELMHOLTZ) {
          double c0 = q[7] * q[11] - q[8] * q[10], c1 = q[8] * q[9] - q[6] * q[11], c2 = q[6] * q[10] - q[7] * q[9];
          double kap = sqrt((c0 * c0 + c1 * c1) + c2 * c2);
          acc[t] += cexp(-OR_I * kappa * r) / 4. / OR_PI / r * (v0 / h) * kap * ws;
        } else {
          double complex K = cexp(-OR_I * kappa * r) / 4. / OR_PI / r;
          double complex G = cexp(-OR_I * kappa * r) * (-1. - OR_I * kappa * r) / 4. / OR_PI / (r * r * r);
          double complex div = dv / h / h;
          double dd[3] = {d0, d1, d2};
          for (int k = 0; k < 3; ++k) {
            double complex val = (q[6 + k] * v0 + q[9 + k] * v1) / h;
            acc[(size_t)k * npts + t] += (K * val + 1. / kappa2 * (G * dd[k]) * div) * ws;
          }
        }
      }
    }
  }
  for (size_t k = 0; k < (size_t)npts * dim; ++k) {
    if (cplx)
      out[2 * k] = creal(acc[k]), out[2 * k + 1] = cimag(acc[k]);
    else
      out[k] = creal(acc[k]);
  }
  free(acc), free(sup);
}

/* ------------------------------------------------------------ solvers */
int or_cg(int n, const double* A, const double* b, double* x, double tol, int maxit) {
  double* r = (double*)malloc(sizeof(double) * n);
  double* p = (double*)malloc(sizeof(double) * n);
  double* Ap = (double*)malloc(sizeof(double) * n);
  double bn = 0;
  for (int i = 0; i < n; ++i) x[i] = 0, r[i] = b[i], p[i] = b[i], bn += b[i] * b[i];
  bn = sqrt(bn);
  double rr = bn * bn;
  int it = 0;
  for (; it < maxit && sqrt(rr) > tol * bn; ++it) {
    for (int i = 0; i < n; ++i) {
      double acc = 0;
      for (int j = 0; j < n; ++j) acc += A[(size_t)j * n + i] * p[j];
      Ap[i] = acc;
    }
    double pAp = 0;
    for (int i = 0; i < n; ++i) pAp += p[i] * Ap[i];
    double alpha = rr / pAp, rr2 = 0;
    for (int i = 0; i < n; ++i) x[i] += alpha * p[i], r[i] -= alpha * Ap[i], rr2 += r[i] * r[i];
    for (int i = 0; i < n; ++i) p[i] = r[i] + (rr2 / rr) * p[i];
    rr = rr2;
  }
  free(r), free(p), free(Ap);
  return it;
}

typedef void (*apply_fn)(const void* ctx, const double complex* x, double complex* y);

/* restarted GMRES with modified Gram-Schmidt and Givens rotations */
static int gmres(int n, apply_fn ap, const void* ctx, const double complex* b, double complex* x,
                 double tol, int restart, int maxit) {
  double complex* V = (double complex*)malloc(sizeof(double complex) * (size_t)n * (restart + 1));
  double complex* H = (double complex*)calloc((size_t)(restart + 1) * restart, sizeof(double complex));
  double complex *cs = malloc(sizeof(double complex) * restart), *sn = malloc(sizeof(double complex) * restart);
  double complex* g = (double complex*)malloc(sizeof(double complex) * (restart + 1));
  double complex* w = (double complex*)malloc(sizeof(double complex) * n);
  double bn = 0;
  for (int i = 0; i < n; ++i) bn += creal(b[i] * conj(b[i]));
  bn = sqrt(bn);
  if (bn == 0) bn = 1;
  int total = 0;
  for (int i = 0; i < n; ++i) x[i] = 0;
  while (total < maxit) {
    ap(ctx, x, w);
    double beta = 0;
    for (int i = 0; i < n; ++i) w[i] = b[i] - w[i], beta += creal(w[i] * conj(w[i]));
    beta = sqrt(beta);
    if (beta <= tol * bn) break;
    for (int i = 0; i < n; ++i) V[i] = w[i] / beta;
    for (int i = 0; i <= restart; ++i) g[i] = 0;
    g[0] = beta;
    int k = 0;
    for (; k < restart && total < maxit; ++k, ++total) {
      ap(ctx, V + (size_t)k * n, w);
      for (int j = 0; j <= k; ++j) {
        double complex hj = 0;
        for (int i = 0; i < n; ++i) hj += conj(V[(size_t)j * n + i]) * w[i];
        H[k * (restart + 1) + j] = hj;
        for (int i = 0; i < n; ++i) w[i] -= hj * V[(size_t)j * n + i];
      }
      double hn = 0;
      for (int i = 0; i < n; ++i) hn += creal(w[i] * conj(w[i]));
      hn = sqrt(hn);
      H[k * (restart + 1) + k + 1] = hn;
      if (hn != 0)
        for (int i = 0; i < n; ++i) V[(size_t)(k + 1) * n + i] = w[i] / hn;
      for (int j = 0; j < k; ++j) {
        double complex a = H[k * (restart + 1) + j], bb = H[k * (restart + 1) + j + 1];
        H[k * (restart + 1) + j] = conj(cs[j]) * a + conj(sn[j]) * bb;
        H[k * (restart + 1) + j + 1] = -sn[j] * a + cs[j] * bb;
      }
      double complex a = H[k * (restart + 1) + k], bb = H[k * (restart + 1) + k + 1];
      double den = sqrt(creal(a * conj(a)) + creal(bb * conj(bb)));
      cs[k] = a / den;
      sn[k] = bb / den;
      H[k * (restart + 1) + k] = den;
      H[k * (restart + 1) + k + 1] = 0;
      g[k + 1] = -sn[k] * g[k];
      g[k] = conj(cs[k]) * g[k];
      if (cabs(g[k + 1]) <= tol * bn) {
        ++k;
        ++total;
        break;
      }
    }
    /* solve upper triangular */
    double complex* yv = (double complex*)malloc(sizeof(double complex) * (k > 0 ? k : 1));
    for (int i = k - 1; i >= 0; --i) {
      double complex acc = g[i];
      for (int j = i + 1; j < k; ++j) acc -= H[j * (restart + 1) + i] * yv[j];
      yv[i] = acc / H[i * (restart + 1) + i];
    }
    for (int j = 0; j < k; ++j)
      for (int i = 0; i < n; ++i) x[i] += yv[j] * V[(size_t)j * n + i];
    free(yv);
    if (k == 0) break;
  }
  free(V), free(H), free(cs), free(sn), free(g), free(w);
  return total;
}

typedef struct {
  int n;
  const double* A;
} dense_ctx;
static void dense_apply(const void* c, const double complex* x, double complex* y) {
  const dense_ctx* d = (const dense_ctx*)c;
  for (int i = 0; i < d->n; ++i) y[i] = 0;
  for (int j = 0; j < d->n; ++j)
    for (int i = 0; i < d->n; ++i)
      y[i] += (d->A[2 * ((size_t)j * d->n + i)] + OR_I * d->A[2 * ((size_t)j * d->n + i) + 1]) * x[j];
}
int or_gmres_dense(int n, const double* A, const double* b, double* x, double tol, int restart,
                   int maxit) {
  dense_ctx d = {n, A};
  return gmres(n, dense_apply, &d, (const double complex*)b, (double complex*)x, tol, restart, maxit);
}
static void h2_apply(const void* c, const double complex* x, double complex* y) {
  or_h2_matvec((const or_h2*)c, 1, (const double*)x, (double*)y);
}
int or_gmres_h2(const or_h2* h, const double* b, double* x, double tol, int restart, int maxit) {
  return gmres(h->s->ndof, h2_apply, h, (const double complex*)b, (double complex*)x, tol, restart, maxit);
}

/* ------------------------------------------------------------ timing (CPU baseline) */
double or_time_assembly_sample(const or_h2* h, int npairs, const int* pairs, int nblocks,
                               const int* blocks, int nthreads) {
  const or_space* s = h->s;
  int nb = h->vd * s->I2;
  int dim = h->vd * h->nc * h->np2;
  if (nthreads > 0) omp_set_num_threads(nthreads);
  double t0 = omp_get_wtime();
  double sink = 0;
#pragma omp parallel reduction(+ : sink)
  {
    double complex* blk = (double complex*)malloc(sizeof(double complex) * nb * nb);
    double complex* F = (double complex*)malloc(sizeof(double complex) * dim * dim);
#pragma omp for schedule(dynamic, 1) nowait
    for (int k = 0; k < npairs; ++k) {
      or_eval_bilinear(s, &h->op, s->elem[pairs[2 * k]], s->elem[pairs[2 * k + 1]], blk);
      sink += creal(blk[0]);
    }
#pragma omp for schedule(dynamic, 1)
    for (int k = 0; k < nblocks; ++k) {
      const or_leaf* l = &h->leaves[blocks[k]];
      interpolate_kernel(h, l->n1, l->n2, F);
      sink += creal(F[0]);
    }
    free(blk), free(F);
  }
  double t1 = omp_get_wtime();
  if (sink == 12345.6789) fprintf(stderr, "%g\n", sink);
  return t1 - t0;
}

void or_set_num_threads(int nthreads) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
}

double or_time_matvec(const or_h2* h, int nrhs, int reps, int nthreads) {
  int c = h->cplx ? 2 : 1;
  size_t D = h->s->ndof;
  double* x = (double*)malloc(sizeof(double) * c * D * nrhs);
  double* y = (double*)malloc(sizeof(double) * c * D * nrhs);
  for (size_t i = 0; i < c * D * nrhs; ++i) x[i] = 1.0 / (1.0 + (double)(i % 97));
  if (nthreads > 0) omp_set_num_threads(nthreads);
  double t0 = omp_get_wtime();
  for (int r = 0; r < reps; ++r) or_h2_matvec(h, nrhs, x, y);
  double t1 = omp_get_wtime();
  free(x), free(y);
  return (t1 - t0) / reps;
}

/* forwardTransformation on a real input matrix X (mom_cols x ncols), for the FMM identity
   tests (tests/test_FMMForwardTransformation.cpp): levels[k] receives (nc*n2) x (ncols/4^k) */
void or_forward_transform(int n2, int nc, const double* mom, int mom_cols, const double* T,
                          int steps, const double* X, int ncols, double** levels) {
  int rows = nc * n2;
  for (int col = 0; col < ncols; ++col) {
    double* out = levels[0] + (size_t)col * rows;
    memset(out, 0, sizeof(double) * rows);
    gemv_real(0, rows, mom_cols, mom, X + (size_t)col * mom_cols, out);
  }
  int ncl = ncols;
  for (int st = 0; st < steps; ++st) {
    int nn = ncl / 4;
    for (int col = 0; col < nn; ++col)
      for (int comp = 0; comp < nc; ++comp) {
        double* buf = (double*)malloc(sizeof(double) * 4 * n2);
        for (int ch = 0; ch < 4; ++ch)
          for (int r = 0; r < n2; ++r) buf[ch * n2 + r] = levels[st][(size_t)(4 * col + ch) * rows + comp * n2 + r];
        double* out = levels[st + 1] + (size_t)col * rows + comp * n2;
        memset(out, 0, sizeof(double) * n2);
        gemv_real(0, n2, 4 * n2, T, buf, out);
        free(buf);
      }
    ncl = nn;
  }
}

/* bulk import of leaf matrices in the order of `keys` (cc, level, id1, id2 per leaf); the
   data holds vd*vd component blocks per leaf, column-major, back to back.  Used by the CPU
   baseline to time the matvec on leaves assembled elsewhere. Returns leaves matched. */
int or_h2_import_leaves(or_h2* h, int n, const int* keys, const double* data) {
  int matched = 0;
  int m = h->cplx ? 2 : 1;
  /* index of oracle leaves by (level, id1, id2, cc) through a sort-free hash */
  long cap = 1;
  while (cap < 4L * h->nleaves) cap <<= 1;
  long* tab = (long*)malloc(sizeof(long) * cap);
  for (long i = 0; i < cap; ++i) tab[i] = -1;
  for (int i = 0; i < h->nleaves; ++i) {
    const or_leaf* l = &h->leaves[i];
    unsigned long k = ((unsigned long)l->level * 1000003UL + (unsigned long)l->id1) * 1000033UL + (unsigned long)l->id2;
    long s = (long)(k & (cap - 1));
    while (tab[s] >= 0) s = (s + 1) & (cap - 1);
    tab[s] = i;
  }
  size_t off = 0;
  for (int q = 0; q < n; ++q) {
    int cc = keys[4 * q], lv = keys[4 * q + 1], a = keys[4 * q + 2], b = keys[4 * q + 3];
    unsigned long k = ((unsigned long)lv * 1000003UL + (unsigned long)a) * 1000033UL + (unsigned long)b;
    long s = (long)(k & (cap - 1));
    int found = -1;
    while (tab[s] >= 0) {
      const or_leaf* l = &h->leaves[tab[s]];
      if (l->level == lv && l->id1 == a && l->id2 == b && l->cc == cc) {
        found = (int)tab[s];
        break;
      }
      s = (s + 1) & (cap - 1);
    }
    int d = cc == 2 ? cluster_nel(h->s, lv) * h->s->I2 : h->nc * h->np2;
    size_t blk = (size_t)m * d * d;
    if (found >= 0) {
      for (int r = 0; r < h->vd; ++r)
        for (int c = 0; c < h->vd; ++c) {
          double** dst = &h->F[found * h->vd * h->vd + r * h->vd + c];
          if (!*dst) *dst = (double*)malloc(sizeof(double) * blk);
          memcpy(*dst, data + off + (size_t)(r * h->vd + c) * blk, sizeof(double) * blk);
        }
      ++matched;
    }
    off += (size_t)h->vd * h->vd * blk;
  }
  free(tab);
  return matched;
}
