/* Internal structures of the oracle (test infrastructure only). */
#ifndef OR_INTERNAL_H_
#define OR_INTERNAL_H_
#include <complex.h>
#undef I
#define OR_I _Complex_I
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "bembel_oracle.h"

#define OR_PI M_PI /* BEMBEL_PI is M_PI where math.h defines it (util/Macros.hpp:15-19) */

typedef struct {
  int pdx, pdy;  /* number of Bernstein functions per direction (degree + 1) */
  double* data;  /* 4*pdx*pdy homogeneous control points, index 4*(i*pdy+j)+k */
} or_patch;

struct or_geometry {
  int npatch;
  or_patch* patches;
};

/* ElementTreeNode (ClusterTree/ElementTreeNode.hpp:21-278) with indices for pointers */
typedef struct {
  int sons;  /* index of first son, -1 for a leaf */
  int nsons;
  int adj[4];
  int vert[4];
  double mid[3];
  double llc[2];
  double radius;
  int id, level, patch;
  int prev, next;
} or_node;

typedef struct {
  int n;  /* number of entries */
  int* row;
  int* col;
  double* val;
} or_triplets;

struct or_space {
  or_geometry* geo;
  int M, P, I2, vd, form;
  /* element tree */
  or_node* nodes;
  int nnodes, cap;
  int root;
  int npoints, nel, max_level;
  int first, last;
  int* elem;      /* Morton element id -> node index */
  int** cluster;  /* cluster[l][id] -> node index, l = 0..M */
  double* points; /* 3*npoints */
  /* transformation matrix, CSR (nsuper x ndof) */
  int nsuper, ndof, nnz;
  int* rowptr;
  int* colidx;
  double* vals;
  /* far-field quadrature table: nel x (ffdeg+1)^2 surface points */
  int ffdeg, nff;
  double* ffield;
};

/* quadrature */
int or_gl_n(int deg); /* number of points of GaussLegendre[deg] = deg+1 */
const double* or_gl_nodes(int npts);
const double* or_gl_weights(int npts);

/* shared helpers */
void or_map2surface(const or_space* s, int node, const double xi[2], double w,
                    double out[12]);
int or_ffield_degree(int P);
int or_nfield_degree(int P, double dist, int level);
void or_compare_elements(const or_space* s, int n1, int n2, int out[3], double* dist);
void or_build_ffield(or_space* s);

/* integrand accumulation for one operator into a block (complex storage) */
typedef struct {
  int kind;
  double complex kappa;
  double complex kappa2;
  int dummy_fn;
  int P, I2, vd, M;
} or_op;

void or_eval_integrand(const or_op* op, const double* p1, const double* p2,
                       double complex* block);
void or_eval_bilinear(const or_space* s, const or_op* op, int n1, int n2,
                      double complex* block);
void or_integrate_n(const or_space* s, const or_op* op, int which, int n1, int rot1,
                    int n2, int rot2, int qdeg, double complex* block);
void or_fmm_interp(const or_op* op, const double* p1, const double* p2,
                   double complex* out /* (vd*nc)^2 row-major */);
int or_op_ncomp(const or_op* op);
int or_op_is_complex(const or_op* op);

#endif
