// Host-side setup of the discrete space and the H2 structure.
//
// The mesh is built directly from the uniform-refinement structure: element ids are Morton
// codes (ClusterTree/ElementTree.hpp:706), vertex ids come from a union-find over the grid
// points of all patches glued along the patch edges found by corner matching
// (ElementTree.hpp:105-161).  This yields the same vertex identification as the reference's
// recursive refinement, which is all compareElements needs.  Enclosing balls are computed
// level by level with the reference's arithmetic (util/GeometryHelper.hpp:32-52,
// ElementTree.hpp:354-396).  This file must be compiled with -ffp-contract=off so that the
// admissibility decisions and quadrature degrees are reproducible.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <numeric>
#include <sstream>

#include "bmi_internal.h"

namespace bmi {

// ------------------------------------------------------------------ Bernstein basis
// Bernstein<N,P>(x) = C(P,N) x^N (1-x)^(P-N) with the powers built as x*(x*(...))
// (Spline/Bernstein.hpp:23-48); derivative P*(B_{N-1,P-1} - B_{N,P-1}) (:239-244).
static double ipow_left(int n, double x) {
  if (n <= 0) return n == 0 ? 1.0 : 0.0;
  double r = x;
  for (int k = 2; k <= n; ++k) r = x * r;
  return r;
}
static int binomial(int k, int n) {
  if (k < 0 || k > n) return 0;
  long r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return static_cast<int>(r);
}
static double bern1(int N, int P, double x) {
  if (N < 0 || N > P) return 0.0;
  return static_cast<double>(binomial(N, P)) * ipow_left(N, x) * ipow_left(P - N, 1. - x);
}
void bernstein(int p, double x, double* out) {
  for (int k = 0; k <= p; ++k) out[k] = bern1(k, p, x);
}
void bernstein_der(int p, double x, double* out) {
  if (p == 0) {
    out[0] = 0.0;
    return;
  }
  out[p] = p * bern1(p - 1, p - 1, x);
  for (int k = p - 1; k >= 1; --k) out[k] = p * (bern1(k - 1, p - 1, x) - bern1(k, p - 1, x));
  out[0] = (-p) * bern1(0, p - 1, x);
}

// Patch::eval (Geometry/Patch.hpp:115-155) for a Bezier patch
void patch_eval(const Patch& p, double s, double t, double out[3]) {
  double xb[32], yb[32];
  bernstein(p.pdx - 1, s, xb);
  bernstein(p.pdy - 1, t, yb);
  double acc[4] = {0., 0., 0., 0.};
  for (int i = 0; i < p.pdx; ++i)
    for (int j = 0; j < p.pdy; ++j) {
      const double w = xb[i] * yb[j];
      const double* c = &p.cp[4 * (p.pdy * i + j)];
      for (int k = 0; k < 4; ++k) acc[k] += c[k] * w;
    }
  for (int k = 0; k < 3; ++k) out[k] = acc[k] / acc[3];
}

// Patch::updateSurfacePoint (Patch.hpp:294-365): out = (s, t, 1, x, dsx, dtx)
void patch_surface_point(const Patch& p, double s, double t, double out[12]) {
  double xb[32], yb[32], xd[32], yd[32];
  bernstein(p.pdx - 1, s, xb);
  bernstein(p.pdy - 1, t, yb);
  bernstein_der(p.pdx - 1, s, xd);
  bernstein_der(p.pdy - 1, t, yd);
  double v[4] = {0, 0, 0, 0}, vs[4] = {0, 0, 0, 0}, vt[4] = {0, 0, 0, 0};
  for (int i = 0; i < p.pdx; ++i)
    for (int j = 0; j < p.pdy; ++j) {
      const double* c = &p.cp[4 * (p.pdy * i + j)];
      const double w = xb[i] * yb[j], ws = xd[i] * yb[j], wt = xb[i] * yd[j];
      for (int k = 0; k < 4; ++k) {
        v[k] += c[k] * w;
        vs[k] += c[k] * ws;
        vt[k] += c[k] * wt;
      }
    }
  const double bot = 1. / v[3], b2 = bot * bot;
  out[0] = s;
  out[1] = t;
  out[2] = 1.0;
  for (int k = 0; k < 3; ++k) {
    out[3 + k] = v[k] * bot;
    out[6 + k] = (vs[k] * v[3] - v[k] * vs[3]) * b2;
    out[9 + k] = (vt[k] * v[3] - v[k] * vt[3]) * b2;
  }
}

static double norm3(double a, double b, double c) { return std::sqrt((a * a + b * b) + c * c); }

// util::computeEnclosingBall (util/GeometryHelper.hpp:32-52)
static void enclosing_ball(double* mid, double* r, const double* m1, double r1, const double* m2,
                           double r2) {
  const double z[3] = {m1[0] - m2[0], m1[1] - m2[1], m1[2] - m2[2]};
  const double nz = norm3(z[0], z[1], z[2]);
  if (nz + r2 <= r1) {
    std::copy(m1, m1 + 3, mid);
    *r = r1;
  } else if (nz + r1 <= r2) {
    std::copy(m2, m2 + 3, mid);
    *r = r2;
  } else {
    const double f = (r1 - r2) / nz;
    for (int k = 0; k < 3; ++k) mid[k] = 0.5 * (m1[k] + m2[k] + f * z[k]);
    *r = 0.5 * (r1 + r2 + nz);
  }
}

struct UnionFind {
  std::vector<int> parent;
  explicit UnionFind(int n) : parent(n) { std::iota(parent.begin(), parent.end(), 0); }
  int find(int a) {
    while (parent[a] != a) a = parent[a] = parent[parent[a]];
    return a;
  }
  void unite(int a, int b) {
    a = find(a), b = find(b);
    if (a != b) parent[std::max(a, b)] = std::min(a, b);
  }
};

// ------------------------------------------------------------------ mesh
void build_mesh(bmi_space* s) {
  const auto& P = s->geo->patches;
  const int np = static_cast<int>(P.size());
  const int M = s->M, n = 1 << M;
  s->npatch = np;
  s->n = n;
  s->nel = np << (2 * M);
  // patch corners, deduplicated with tolerance 1e-9 in patch order (ElementTree.hpp:128-154)
  static const double corner_st[4][2] = {{0., 0.}, {1., 0.}, {1., 1.}, {0., 1.}};
  std::vector<std::array<double, 3>> uniq;
  std::vector<std::array<int, 4>> corner(np);
  for (int p = 0; p < np; ++p)
    for (int j = 0; j < 4; ++j) {
      double v[3];
      patch_eval(P[p], corner_st[j][0], corner_st[j][1], v);
      size_t idx = 0;
      for (; idx < uniq.size(); ++idx)
        if (norm3(uniq[idx][0] - v[0], uniq[idx][1] - v[1], uniq[idx][2] - v[2]) < 1e-9) break;
      if (idx == uniq.size()) uniq.push_back({v[0], v[1], v[2]});
      corner[p][j] = static_cast<int>(idx);
    }
  // grid points of all patches and their identification along shared patch edges
  const int g = n + 1;
  auto key = [&](int p, int i, int j) { return (p * g + j) * g + i; };
  UnionFind uf(np * g * g);
  // corners shared between patches
  const int cgrid[4][2] = {{0, 0}, {n, 0}, {n, n}, {0, n}};
  std::vector<int> first_corner(uniq.size(), -1);
  for (int p = 0; p < np; ++p)
    for (int j = 0; j < 4; ++j) {
      const int k = key(p, cgrid[j][0], cgrid[j][1]);
      int& f = first_corner[corner[p][j]];
      if (f < 0)
        f = k;
      else
        uf.unite(f, k);
    }
  // grid point t along edge j, walking from corner j to corner j+1
  auto edge_point = [&](int j, int t, int* i, int* jj) {
    switch (j) {
      case 0: *i = t, *jj = 0; break;
      case 1: *i = n, *jj = t; break;
      case 2: *i = n - t, *jj = n; break;
      default: *i = 0, *jj = n - t; break;
    }
  };
  struct EdgeRef {
    int a, b, p, j;
  };
  std::vector<EdgeRef> edges;
  for (int p = 0; p < np; ++p)
    for (int j = 0; j < 4; ++j) {
      int a = corner[p][j], b = corner[p][(j + 1) % 4];
      edges.push_back({std::min(a, b), std::max(a, b), p, j});
    }
  std::stable_sort(edges.begin(), edges.end(), [](const EdgeRef& x, const EdgeRef& y) {
    return x.a != y.a ? x.a < y.a : x.b < y.b;
  });
  for (size_t q = 0; q + 1 < edges.size(); ++q) {
    const EdgeRef& A = edges[q];
    const EdgeRef& B = edges[q + 1];
    if (A.a != B.a || A.b != B.b) continue;
    const bool same_dir = corner[A.p][A.j] == corner[B.p][B.j];
    for (int t = 0; t <= n; ++t) {
      int ia, ja, ib, jb;
      edge_point(A.j, t, &ia, &ja);
      edge_point(B.j, same_dir ? t : n - t, &ib, &jb);
      uf.unite(key(A.p, ia, ja), key(B.p, ib, jb));
    }
    ++q;
  }
  // elements in Morton order: son i of a node sits at (i==1||i==2, i>=2) (Constants.hpp:28)
  s->vert.assign(4 * static_cast<size_t>(s->nel), 0);
  s->elem_patch.assign(s->nel, 0);
  s->elem_llc.assign(2 * static_cast<size_t>(s->nel), 0.0);
  std::vector<int> canon_id(np * g * g, -1);
  int nv = 0;
  const double h = 1.0 / static_cast<double>(n);
  for (int e = 0; e < s->nel; ++e) {
    const int p = e >> (2 * M);
    const int m = e & ((1 << (2 * M)) - 1);
    int ix = 0, iy = 0;
    for (int l = M - 1; l >= 0; --l) {
      const int d = (m >> (2 * l)) & 3;
      ix = (ix << 1) | ((d == 1 || d == 2) ? 1 : 0);
      iy = (iy << 1) | (d >= 2 ? 1 : 0);
    }
    s->elem_patch[e] = p;
    s->elem_llc[2 * e] = ix * h;
    s->elem_llc[2 * e + 1] = iy * h;
    const int cx[4] = {ix, ix + 1, ix + 1, ix}, cy[4] = {iy, iy, iy + 1, iy + 1};
    for (int k = 0; k < 4; ++k) {
      const int root = uf.find(key(p, cx[k], cy[k]));
      if (canon_id[root] < 0) canon_id[root] = nv++;
      s->vert[4 * e + k] = canon_id[root];
    }
  }
  s->nvert = nv;
  // vertex coordinates: every element writes its corners, the last writer wins
  // (ElementTree::generatePointList, ElementTree.hpp:232-256)
  s->points.assign(3 * static_cast<size_t>(nv), 0.0);
  for (int e = 0; e < s->nel; ++e) {
    const double x = s->elem_llc[2 * e], y = s->elem_llc[2 * e + 1];
    const double cs[4][2] = {{x, y}, {x + h, y}, {x + h, y + h}, {x, y + h}};
    for (int k = 0; k < 4; ++k)
      patch_eval(P[s->elem_patch[e]], cs[k][0], cs[k][1], &s->points[3 * s->vert[4 * e + k]]);
  }
  // enclosing balls, leaves then parents (ElementTree.hpp:366-396)
  s->cmid.assign(M + 1, {});
  s->crad.assign(M + 1, {});
  s->cmid[M].assign(3 * static_cast<size_t>(s->nel), 0.0);
  s->crad[M].assign(s->nel, 0.0);
  for (int e = 0; e < s->nel; ++e) {
    const double* pt = s->points.data();
    const int* v = &s->vert[4 * e];
    double m1[3], m2[3], r1, r2;
    enclosing_ball(m1, &r1, pt + 3 * v[0], 0, pt + 3 * v[2], 0);
    enclosing_ball(m2, &r2, pt + 3 * v[1], 0, pt + 3 * v[3], 0);
    enclosing_ball(&s->cmid[M][3 * e], &s->crad[M][e], m1, r1, m2, r2);
  }
  for (int l = M - 1; l >= 0; --l) {
    const int nc = np << (2 * l);
    s->cmid[l].assign(3 * static_cast<size_t>(nc), 0.0);
    s->crad[l].assign(nc, 0.0);
    const auto& cm = s->cmid[l + 1];
    const auto& cr = s->crad[l + 1];
    for (int c = 0; c < nc; ++c) {
      const int s0 = 4 * c;
      double m1[3], m2[3], r1, r2;
      enclosing_ball(m1, &r1, &cm[3 * s0], cr[s0], &cm[3 * (s0 + 2)], cr[s0 + 2]);
      enclosing_ball(m2, &r2, &cm[3 * (s0 + 1)], cr[s0 + 1], &cm[3 * (s0 + 3)], cr[s0 + 3]);
      enclosing_ball(&s->cmid[l][3 * c], &s->crad[l][c], m1, r1, m2, r2);
    }
  }
  // patch topology for the DivConforming glue is rebuilt from `corner` when needed
}

// ------------------------------------------------------------------ near-field rules
static int operator_order(int kind) { return kind == BMI_DUMMY ? 0 : -1; }

// LinearOperatorBase::get_FarfieldQuadratureDegree (LinearOperatorBase.hpp:56-58)
int farfield_degree(int kind, int P) { return P - operator_order(kind) + 1; }

// LinearOperatorBase::getNearfieldQuadratureDegree (LinearOperatorBase.hpp:69-89); the
// double expression is kept in the reference's order and truncated toward zero
int nearfield_degree(int kind, int P, double dist, int level) {
  const int OO = operator_order(kind);
  const double dlog = (dist * (1 << level) < 1) ? -(level * std::log(2.)) : std::log(dist);
  const int alpha = 2 - OO + 2 * P;
  const double num = (alpha + P) * level * std::log(2.) - (2 - P + OO) * dlog;
  const double den = (level + 2) * std::log(2.) + dlog;
  return static_cast<int>(0.5 * num / den);
}

// DuffyTrick::compareElements (DuffyTrick/compareElements.hpp:23-69): out = rot1, rot2, class
void compare_elements(const bmi_space* s, int e1, int e2, int out[3], double* dist) {
  if (e1 == e2) {
    out[0] = 4, out[1] = 4, out[2] = 2;
    *dist = 0;
    return;
  }
  const double* m1 = &s->cmid[s->M][3 * e1];
  const double* m2 = &s->cmid[s->M][3 * e2];
  double d = norm3(m1[0] - m2[0], m1[1] - m2[1], m1[2] - m2[2]) - s->crad[s->M][e1] -
             s->crad[s->M][e2];
  d = d >= 0 ? d : 0;
  *dist = d;
  if (d > .5 / (1 << s->M)) {
    out[0] = 4, out[1] = 4, out[2] = 0;
    return;
  }
  const int* v1 = &s->vert[4 * e1];
  const int* v2 = &s->vert[4 * e2];
  for (int r1 = 0; r1 < 4; ++r1)
    for (int r2 = 0; r2 < 4; ++r2)
      if (v1[r1] == v2[r2]) {
        if (v1[3] == v2[(r2 + 1) % 4]) {
          out[0] = 3, out[1] = r2, out[2] = 3;
        } else if (v1[(r1 + 1) % 4] == v2[(r2 + 3) % 4]) {
          out[0] = r1, out[1] = (r2 + 3) % 4, out[2] = 3;
        } else {
          out[0] = r1, out[1] = r2, out[2] = 4;
        }
        return;
      }
  out[0] = 4, out[1] = 4, out[2] = 0;
}

// ------------------------------------------------------------------ block cluster tree
// BlockClusterTree::compareCluster (H2Matrix/BlockClusterTree.hpp:255-278)
static int compare_cluster(double eta, int mcl, int M, const double* m1, double r1, const double* m2,
                           double r2, int level) {
  double d = norm3(m1[0] - m2[0], m1[1] - m2[1], m1[2] - m2[2]) - r1 - r2;
  d = d > 0 ? d : 0;
  const double mr = r1 > r2 ? r1 : r2;
  int r = (d < 0 || mr >= eta * d) ? 0 : 1;
  if (r == 0 && M - level <= mcl) r = 2;
  return r;
}

// Level-synchronous construction of the same leaf set as the reference's recursion
// (BlockClusterTree.hpp:121-182).  The root (radius infinity) always refines into all patch
// pairs unless the whole matrix is one dense leaf.
void build_block_tree(const bmi_space* s, double eta, int mcl, std::vector<Leaf>& dense,
                      std::vector<Leaf>& lowrank) {
  dense.clear();
  lowrank.clear();
  const int M = s->M;
  BMI_REQUIRE(M - (-1) > mcl, BMI_UNSUPPORTED,
              "H2 matrix needs refinement_level >= min_cluster_level");
  std::vector<std::pair<int, int>> cur, next;
  for (int j = 0; j < s->npatch; ++j)
    for (int i = 0; i < s->npatch; ++i) cur.push_back({i, j});
  for (int l = 0; l <= M && !cur.empty(); ++l) {
    next.clear();
    const auto& cm = s->cmid[l];
    const auto& cr = s->crad[l];
    for (auto& pr : cur) {
      const int c1 = pr.first, c2 = pr.second;
      const int cc = compare_cluster(eta, mcl, M, &cm[3 * c1], cr[c1], &cm[3 * c2], cr[c2], l);
      if (cc == 1) {
        lowrank.push_back({1, l, c1, c2});
      } else if (cc == 2) {
        dense.push_back({2, l, c1, c2});
      } else {
        for (int j = 0; j < 4; ++j)
          for (int i = 0; i < 4; ++i) next.push_back({4 * c1 + i, 4 * c2 + j});
      }
    }
    std::swap(cur, next);
  }
  // storage order: dense by (row cluster, column cluster), low rank by (level, row, column)
  std::sort(dense.begin(), dense.end(), [](const Leaf& a, const Leaf& b) {
    return a.c1 != b.c1 ? a.c1 < b.c1 : a.c2 < b.c2;
  });
  std::sort(lowrank.begin(), lowrank.end(), [](const Leaf& a, const Leaf& b) {
    if (a.level != b.level) return a.level < b.level;
    return a.c1 != b.c1 ? a.c1 < b.c1 : a.c2 < b.c2;
  });
}

// ------------------------------------------------------------------ FMM constants
// ChebychevRoots (H2Matrix/H2Multipole.hpp:20-32)
void cheb_points(int n, double* x) {
  const double alpha = M_PI / (2. * static_cast<double>(n));
  for (int i = 0; i < n; ++i) {
    const double gp = static_cast<double>(n - i);
    x[i] = 0.5 * std::cos(alpha * (2 * gp - 1)) + 0.5;
  }
}
// Newton-form Lagrange basis (H2Multipole.hpp:39-67)
static std::vector<double> lagrange_newton(int n, const double* x) {
  std::vector<double> L(n * n, 0.0);  // L[k*n+i]
  for (int i = 0; i < n; ++i) L[i * n + i] = 1.0;
  for (int i = 0; i < n; ++i)
    for (int j = 1; j < n; ++j)
      for (int k = n - 1; k >= j; --k) L[k * n + i] = (L[k * n + i] - L[(k - 1) * n + i]) / (x[k] - x[k - j]);
  return L;
}
static double newton_eval(int n, const double* x, const std::vector<double>& L, int i, double t) {
  double r = L[(n - 1) * n + i];
  for (int k = n - 2; k >= 0; --k) r = r * (t - x[k]) + L[k * n + i];
  return r;
}
// computeTransferMatrices (H2Multipole.hpp:76-115): column-major n^2 x 4n^2, child blocks in
// son order (0,0),(1/2,0),(1/2,1/2),(0,1/2)
void transfer_matrix(int n, std::vector<double>& T) {
  const int n2 = n * n;
  std::vector<double> x(n);
  cheb_points(n, x.data());
  const auto L = lagrange_newton(n, x.data());
  std::vector<double> E(n * 2 * n);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      E[i * 2 * n + j] = newton_eval(n, x.data(), L, j, 0.5 * x[i]);
      E[i * 2 * n + j + n] = newton_eval(n, x.data(), L, j, 0.5 * x[i] + 0.5);
    }
  static const int son_of_quadrant[4] = {0, 3, 1, 2};
  T.assign(static_cast<size_t>(n2) * 4 * n2, 0.0);
  for (int k = 0; k < 4; ++k)
    for (int i = 0; i < n; ++i)
      for (int ii = 0; ii < n; ++ii)
        for (int j = 0; j < n; ++j)
          for (int jj = 0; jj < n; ++jj)
            T[static_cast<size_t>(i * n + ii + n2 * son_of_quadrant[k]) * n2 + j * n + jj] =
                E[i * 2 * n + j + (k / 2) * n] * E[ii * 2 * n + jj + (k % 2) * n];
}

// Moment1D / Moment1DDerivative (H2Multipole.hpp:120-218), row-major np x 2^cr (P+1)
static void moment1d(int P, int cl, int cr, int np, bool deriv, std::vector<double>& out);

}  // namespace bmi

#include "gauss_legendre_table.inc"
static const int bmi_gl_offset_host[52] = BMI_GL_OFFSET_INIT;
static const double bmi_gl_nodes_host[BMI_GL_TABLE_SIZE] = BMI_GL_NODES_INIT;
static const double bmi_gl_weights_host[BMI_GL_TABLE_SIZE] = BMI_GL_WEIGHTS_INIT;

namespace bmi {

// points of the (P+2)^2-point rule of the linear forms on every element (Morton order), point
// k = element * (P+2)^2 + a * (P+2) + b at local coordinates (x_a, x_b)
void linear_form_points(const bmi_space* s, double* out) {
  const int n = s->P + 2;
  const double* qx = bmi_gl_nodes_host + bmi_gl_offset_host[n - 1];
  const double h = 1.0 / s->n;
  for (int e = 0; e < s->nel; ++e) {
    const Patch& pt = s->geo->patches[s->elem_patch[e]];
    for (int q = 0; q < n * n; ++q) {
      double sp[12];
      patch_surface_point(pt, s->elem_llc[2 * e] + h * qx[q / n], s->elem_llc[2 * e + 1] + h * qx[q % n], sp);
      double* o = out + 3 * (static_cast<int64_t>(e) * n * n + q);
      o[0] = sp[3], o[1] = sp[4], o[2] = sp[5];
    }
  }
}

static void moment1d(int P, int cl, int cr, int np, bool deriv, std::vector<double>& out) {
  const int n = 1 << cr, N = 1 << cl, pp1 = P + 1;
  const double h = 1. / n, H = 1. / N;
  const int nq = static_cast<int>(std::ceil(0.5 * (np + P - 2))) + 1;
  const double* qx = bmi_gl_nodes_host + bmi_gl_offset_host[nq - 1];
  const double* qw = bmi_gl_weights_host + bmi_gl_offset_host[nq - 1];
  std::vector<double> x(np);
  cheb_points(np, x.data());
  const auto L = lagrange_newton(np, x.data());
  const int cols = n * pp1;
  out.assign(static_cast<size_t>(np) * cols, 0.0);
  for (int i = 0; i < np; ++i)
    for (int j = 0; j < n; ++j) {
      double acc[32] = {0};
      for (int k = 0; k < nq; ++k) {
        double b[32];
        const double lv = newton_eval(np, x.data(), L, i, h * (j + qx[k]));
        double w;
        if (!deriv) {
          w = qw[k] * std::sqrt(h * H) * lv;
          bernstein(P, qx[k], b);
        } else {
          w = qw[k] / std::sqrt(h * H) * lv;
          bernstein_der(P, qx[k], b);
        }
        for (int m = 0; m < pp1; ++m) acc[m] += w * b[m];
      }
      for (int m = 0; m < pp1; ++m) out[i * cols + j * pp1 + m] = acc[m];
    }
}

// moment2DComputer (H2Multipole.hpp:224-293) into a column-major matrix with leading
// dimension ld, rows offset by roff
static void moment2d(int P, int cl, int cr, int np, bool d1, bool d2, double* out, int ld, int roff) {
  const int n = 1 << cr, n2 = n * n, pp1 = P + 1, I2 = pp1 * pp1;
  std::vector<double> m1, m2;
  moment1d(P, cl, cr, np, d1, m1);
  moment1d(P, cl, cr, np, d2, m2);
  std::vector<int> is(n2, 0), it(n2, 0);
  for (int j = 0; j < cr; ++j) {
    const int inc = 1 << (2 * j);
    for (int i = 0; i < inc; ++i) {
      const int a = is[i], b = it[i];
      is[4 * i] = 2 * a, is[4 * i + 1] = 2 * a + 1, is[4 * i + 2] = 2 * a + 1, is[4 * i + 3] = 2 * a;
      it[4 * i] = 2 * b, it[4 * i + 1] = 2 * b, it[4 * i + 2] = 2 * b + 1, it[4 * i + 3] = 2 * b + 1;
    }
  }
  const int cols = n * pp1;
  for (int i = 0; i < np; ++i)
    for (int j = 0; j < np; ++j)
      for (int k = 0; k < n2; ++k)
        for (int a = 0; a < pp1; ++a)
          for (int b = 0; b < pp1; ++b)
            out[static_cast<size_t>(I2 * k + a * pp1 + b) * ld + i * np + j + roff] =
                m1[i * cols + is[k] * pp1 + b] * m2[j * cols + it[k] * pp1 + a];
}

void moment_matrix(int P, int cl, int cr, int np, int comp, std::vector<double>& out, int* rows,
                   int* cols) {
  const int np2 = np * np, I2 = (P + 1) * (P + 1);
  *cols = (1 << (2 * cr)) * I2;
  if (comp == 0) {
    *rows = np2;
    out.assign(static_cast<size_t>(np2) * *cols, 0.0);
    moment2d(P, cl, cr, np, false, false, out.data(), np2, 0);
    return;
  }
  // Maxwell: [M; M_dx] for the first vector component, [M; M_dy] for the second
  // (Maxwell/SingleLayerOperator.hpp:153-183)
  *rows = 2 * np2;
  out.assign(static_cast<size_t>(2 * np2) * *cols, 0.0);
  moment2d(P, cl, cr, np, false, false, out.data(), 2 * np2, 0);
  moment2d(P, cl, cr, np, comp == 1, comp == 2, out.data(), 2 * np2, np2);
}

// ------------------------------------------------------------------ transformation matrix
// Spl::DeBoor of one coefficient vector (Spline/DeBoor.hpp:67-110)
static double de_boor(const std::vector<double>& coef, const std::vector<double>& knots, double x) {
  const int nc = static_cast<int>(coef.size());
  const int p = static_cast<int>(knots.size()) - nc - 1;
  int l = 0;
  while (knots[l] <= x && l != nc) l++;
  l = l - 1;
  double temp[32], ws[32];
  for (int i = 0; i <= p; ++i) temp[i] = coef[l - p + i];
  for (int k = p; k != 0; k--) {
    for (int i = 0; i < k; i++) ws[i] = (x - knots[l - k + 1 + i]) / (knots[l + 1 + i] - knots[l - k + 1 + i]);
    for (int i = 0; i < k; i++) temp[i] = ((ws[i] * temp[1 + i]) + ((1 - ws[i]) * temp[i]));
  }
  return temp[0];
}

struct Triplet {
  int r, c;
  double v;
};

// Local projection of the B-spline space onto the element Bernstein bases by interpolation
// at equispaced points (ProjectorRoutines::makeLocalProjectionTriplets,
// AnsatzSpace/Projector.hpp:165-337); coefficients below 1e-4 are dropped.
static void local_projection(const bmi_space* s, int pp1x, int pp1y, std::vector<Triplet>& out,
                             int* nc_dof, int* ndc_dof) {
  const int mpd = std::max(pp1x, pp1y), minp = std::min(pp1x, pp1y);
  const int rep = (mpd <= 1) ? minp : 1;
  const int n = s->n;
  auto knots = [&](int pp1) {
    std::vector<double> k;
    const double h = 1. / n;
    for (int i = 0; i < pp1; ++i) k.push_back(0);
    for (int i = 1; i < n; ++i)
      for (int r = 0; r < rep; ++r) k.push_back(i * h);
    for (int i = 0; i < pp1; ++i) k.push_back(1);
    return k;
  };
  const auto kx = knots(pp1x), ky = knots(pp1y);
  const int dimx = static_cast<int>(kx.size()) - pp1x, dimy = static_cast<int>(ky.size()) - pp1y;
  const int dim = dimx * dimy;
  std::vector<double> mask(mpd);
  for (int i = 0; i < mpd; ++i) mask[i] = (i + 1) * (1. / (mpd + 1));
  const int sz = mpd * mpd;
  // interpolation system, LU with partial pivoting
  std::vector<double> A(sz * sz);
  for (int iy = 0; iy < mpd; ++iy) {
    double vy[32], vx[32];
    bernstein(mpd - 1, mask[iy], vy);
    for (int ix = 0; ix < mpd; ++ix) {
      bernstein(mpd - 1, mask[ix], vx);
      for (int jy = 0; jy < mpd; ++jy)
        for (int jx = 0; jx < mpd; ++jx) A[(iy * mpd + ix) * sz + jy * mpd + jx] = vx[jx] * vy[jy];
    }
  }
  std::vector<int> piv(sz);
  for (int k = 0; k < sz; ++k) {
    int p = k;
    for (int i = k + 1; i < sz; ++i)
      if (std::fabs(A[i * sz + k]) > std::fabs(A[p * sz + k])) p = i;
    piv[k] = p;
    if (p != k)
      for (int j = 0; j < sz; ++j) std::swap(A[k * sz + j], A[p * sz + j]);
    for (int i = k + 1; i < sz; ++i) {
      A[i * sz + k] /= A[k * sz + k];
      for (int j = k + 1; j < sz; ++j) A[i * sz + j] -= A[i * sz + k] * A[k * sz + j];
    }
  }
  std::vector<double> cx(dimx, 0.0), cy(dimy, 0.0), rhs(sz);
  for (int e = 0; e < s->nel; ++e) {
    const double px = s->elem_llc[2 * e], py = s->elem_llc[2 * e + 1];
    const double h = 1.0 / n;
    const double mx = px + .5 * h, my = py + .5 * h;
    for (int dy = 0; dy < dimy; ++dy) {
      if (!(ky[dy] <= my && ky[dy + pp1y] >= my)) continue;
      for (int dx = 0; dx < dimx; ++dx) {
        if (!(kx[dx] <= mx && kx[dx + pp1x] >= mx)) continue;
        double vx[32], vy[32];
        cx[dx] = 1.0;
        cy[dy] = 1.0;
        for (int i = 0; i < mpd; ++i) {
          vx[i] = de_boor(cx, kx, px + (1.0 / n) * mask[i]);
          vy[i] = de_boor(cy, ky, py + (1.0 / n) * mask[i]);
        }
        cx[dx] = 0.0;
        cy[dy] = 0.0;
        for (int iy = 0; iy < mpd; ++iy)
          for (int ix = 0; ix < mpd; ++ix) rhs[iy * mpd + ix] = vx[ix] * vy[iy];
        for (int k = 0; k < sz; ++k) std::swap(rhs[k], rhs[piv[k]]);
        for (int i = 0; i < sz; ++i)
          for (int j = 0; j < i; ++j) rhs[i] -= A[i * sz + j] * rhs[j];
        for (int i = sz - 1; i >= 0; --i) {
          for (int j = i + 1; j < sz; ++j) rhs[i] -= A[i * sz + j] * rhs[j];
          rhs[i] /= A[i * sz + i];
        }
        for (int i = 0; i < sz; ++i)
          if (std::fabs(rhs[i]) > 1e-4)
            out.push_back({e * sz + i, s->elem_patch[e] * dim + dimx * dy + dx, rhs[i]});
      }
    }
  }
  *nc_dof = dim * s->npatch;
  *ndc_dof = sz * n * n * s->npatch;
}

// DivConforming identification of normal-component DOFs across patch edges
// (AnsatzSpace/Glue.hpp:153-232, 249-331, 528-598).
static void glue_divconforming(const bmi_space* s, std::vector<Triplet>& t, int* ncols) {
  const auto& P = s->geo->patches;
  const int np = s->npatch, n = s->n, Pd = s->P;
  const int small_dim = Pd + (n - 1), large_dim = Pd + 1 + (n - 1);
  const int dpp = small_dim * large_dim;
  const int pre = *ncols;
  // patch topology (ElementTree::patchTopologyInfo, ElementTree.hpp:510-533)
  static const double corner_st[4][2] = {{0., 0.}, {1., 0.}, {1., 1.}, {0., 1.}};
  std::vector<std::array<double, 3>> uniq;
  std::vector<std::array<int, 4>> corner(np);
  for (int p = 0; p < np; ++p)
    for (int j = 0; j < 4; ++j) {
      double v[3];
      patch_eval(P[p], corner_st[j][0], corner_st[j][1], v);
      size_t idx = 0;
      for (; idx < uniq.size(); ++idx)
        if (norm3(uniq[idx][0] - v[0], uniq[idx][1] - v[1], uniq[idx][2] - v[2]) < 1e-9) break;
      if (idx == uniq.size()) uniq.push_back({v[0], v[1], v[2]});
      corner[p][j] = static_cast<int>(idx);
    }
  auto nb = [&](int p, int j, int* q, int* k) {
    const int a = std::min(corner[p][j], corner[p][(j + 1) % 4]);
    const int b = std::max(corner[p][j], corner[p][(j + 1) % 4]);
    for (int pp = 0; pp < np; ++pp)
      for (int jj = 0; jj < 4; ++jj) {
        if (pp == p && jj == j) continue;
        const int c = std::min(corner[pp][jj], corner[pp][(jj + 1) % 4]);
        const int d = std::max(corner[pp][jj], corner[pp][(jj + 1) % 4]);
        if (a == c && b == d) {
          *q = pp, *k = jj;
          return true;
        }
      }
    return false;
  };
  auto edge_dofs = [](int ec, int dx, int dy, int shift) {
    std::vector<int> o;
    if (ec == 0)
      for (int i = 0; i < dx; ++i) o.push_back(i + shift);
    else if (ec == 1)
      for (int i = 0; i < dy; ++i) o.push_back(dx * (i + 1) - 1 + shift);
    else if (ec == 2)
      for (int i = 0; i < dx; ++i) o.push_back(dx * (dy - 1) + i + shift);
    else
      for (int i = 0; i < dy; ++i) o.push_back(dx * i + shift);
    return o;
  };
  auto first_comp = [](int e) { return !(e == 0 || e == 2); };
  auto fwd = [](int e) { return e == 0 || e == 1; };
  auto inward = [](int e) { return e == 0 || e == 3; };
  std::vector<int> partner(pre, -1), coef(pre, 1);
  std::vector<char> slave(pre, 0), master(pre, 0), boundary(pre, 0);
  for (int p = 0; p < np; ++p)
    for (int j = 0; j < 4; ++j) {
      int q = -1, k = -1;
      const bool has = nb(p, j, &q, &k);
      if (has && q < p) continue;  // each interface once, from the lower patch
      const int sh1 = p * dpp + (first_comp(j) ? 0 : dpp * np);
      const auto d1 = edge_dofs(j, first_comp(j) ? large_dim : small_dim,
                                first_comp(j) ? small_dim : large_dim, sh1);
      if (!has) {
        for (int v : d1) master[v] = 1, boundary[v] = 1;
        continue;
      }
      const int sh2 = q * dpp + (first_comp(k) ? 0 : dpp * np);
      const auto d2 = edge_dofs(k, first_comp(k) ? large_dim : small_dim,
                                first_comp(k) ? small_dim : large_dim, sh2);
      const bool rev = fwd(j) == fwd(k);
      const int cf = (inward(j) == inward(k)) ? -1 : 1;
      const int m = static_cast<int>(d1.size());
      for (int i = 0; i < m; ++i) {
        const int jj = rev ? m - 1 - i : i;
        const int a = std::min(d1[i], d2[jj]), b = std::max(d1[i], d2[jj]);
        master[a] = 1, slave[b] = 1, partner[a] = b, coef[a] = cf;
      }
    }
  std::vector<int> post(pre, -1);
  std::vector<double> pc(pre, 0.0);
  int idx = 0;
  for (int i = 0; i < pre; ++i) {
    if (slave[i] || (master[i] && boundary[i])) continue;
    post[i] = idx;
    pc[i] = 1.0;
    if (partner[i] >= 0) post[partner[i]] = idx, pc[partner[i]] = coef[i];
    ++idx;
  }
  std::vector<Triplet> o;
  o.reserve(t.size());
  for (const auto& tr : t)
    if (post[tr.c] >= 0) o.push_back({tr.r, post[tr.c], tr.v * pc[tr.c]});
  t.swap(o);
  *ncols = idx;
}

static void to_csr(int nrows, int ncols, std::vector<Triplet>& t, std::vector<int>& rp, std::vector<int>& ci,
                   std::vector<double>& v) {
  std::stable_sort(t.begin(), t.end(), [](const Triplet& a, const Triplet& b) {
    return a.r != b.r ? a.r < b.r : a.c < b.c;
  });
  rp.assign(nrows + 1, 0);
  ci.clear();
  v.clear();
  for (size_t k = 0; k < t.size(); ++k) {
    if (!ci.empty() && k > 0 && t[k].r == t[k - 1].r && t[k].c == t[k - 1].c) {
      v.back() += t[k].v;
      continue;
    }
    ci.push_back(t[k].c);
    v.push_back(t[k].v);
    rp[t[k].r + 1]++;
  }
  for (int r = 0; r < nrows; ++r) rp[r + 1] += rp[r];
  (void)ncols;
}

// T = Projector * Glue (AnsatzSpace/AnsatzSpace.hpp:102-111)
void build_transformation(bmi_space* s) {
  std::vector<Triplet> t;
  int nc = 0, nd = 0;
  if (s->form == BMI_FORM_DISCONTINUOUS) {
    local_projection(s, s->P + 1, s->P + 1, t, &nc, &nd);
  } else {
    std::vector<Triplet> t2;
    int nc2 = 0, nd2 = 0;
    local_projection(s, s->P + 1, s->P, t, &nc, &nd);
    local_projection(s, s->P, s->P + 1, t2, &nc2, &nd2);
    for (auto& x : t2) t.push_back({x.r + nd, x.c + nc, x.v});
    nc += nc2;
    nd += nd2;
    glue_divconforming(s, t, &nc);
  }
  s->nsuper = nd;
  s->ndof = nc;
  to_csr(nd, nc, t, s->rowptr, s->col, s->val);
  // transpose
  std::vector<Triplet> tt;
  tt.reserve(s->val.size());
  for (int r = 0; r < nd; ++r)
    for (int k = s->rowptr[r]; k < s->rowptr[r + 1]; ++k) tt.push_back({s->col[k], r, s->val[k]});
  to_csr(nc, nd, tt, s->trowptr, s->tcol, s->tval);
}

}  // namespace bmi
