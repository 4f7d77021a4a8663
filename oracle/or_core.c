/* Oracle core: geometry, Bernstein basis, element tree, enclosing balls, projector.
 * Test infrastructure only (see bembel_oracle.h). */
#include <stdio.h>

#include "or_internal.h"
#include "gauss_table.inc"

/* ---------------------------------------------------------------- quadrature */
/* GaussSquare<N>[deg] is the tensor rule of GaussLegendreRule<deg+1>
   (Quadrature/TensorProductQuadratureVector.hpp:23-58) */
int or_gl_n(int deg) { return deg + 1; }
const double* or_gl_nodes(int npts) { return or_gl_x + or_gl_offset[npts - 1]; }
const double* or_gl_weights(int npts) { return or_gl_w + or_gl_offset[npts - 1]; }

/* ---------------------------------------------------------------- Bernstein */
/* x^N as the template recursion BernsteinX<N> = x * BernsteinX<N-1>
   (Spline/Bernstein.hpp:23-42) */
static double xpow(int n, double x) {
  if (n <= 0) return n == 0 ? 1. : 0.;
  double r = x;
  for (int k = 2; k <= n; ++k) r = x * r;
  return r;
}
static int binom(int k, int n) {
  if (k < 0 || k > n) return 0;
  long r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return (int)r;
}
/* Bernstein<N,P>(x) = Binomial<N,P> * x^N * (1-x)^(P-N) (Bernstein.hpp:44-48) */
static double bern(int N, int P, double x) {
  if (N < 0 || N > P) {
    /* Bernstein<-1,-1>: Binomial<-1,-1> matches Binomial<N,N> = 1, BernsteinX<-1> = 0 */
    return 0.;
  }
  return (double)binom(N, P) * xpow(N, x) * xpow(P - N, 1. - x);
}
/* EvalBernsteinBasis (Bernstein.hpp:197-201) */
void or_bernstein(int p, double x, double* out) {
  for (int n = 0; n <= p; ++n) out[n] = bern(n, p, x);
}
/* EvalBernsteinDerBasis (Bernstein.hpp:239-244 with HiddenBernsteinClass::EvalDerBasis) */
void or_bernstein_der(int p, double x, double* out) {
  if (p == 0) {
    out[0] = 0. * 0.;
    return;
  }
  out[p] = p * bern(p - 1, p - 1, x);
  for (int n = p - 1; n >= 1; --n) out[n] = p * (bern(n - 1, p - 1, x) - bern(n, p - 1, x));
  out[0] = (-p) * bern(0, p - 1, x);
}

/* ---------------------------------------------------------------- geometry */
static char* next_line(FILE* f, char* buf, int n) { return fgets(buf, n, f); }

/* LoadGeometryFileDAT (Geometry/GeometryIO.hpp:25-122) + Patch::init_Patch for Bezier
   patches (Patch.hpp:63-87; Unroll Spline/Unroll.hpp). Patches with interior knots are
   rejected (the shredder path is not part of this restatement). */
or_geometry* or_geometry_load(const char* path, const double scale[3]) {
  FILE* f = fopen(path, "r");
  if (!f) return NULL;
  static char buf[1 << 20];
  for (int i = 0; i < 5; ++i)
    if (!next_line(f, buf, sizeof(buf))) {
      fclose(f);
      return NULL;
    }
  int info5[5];
  if (sscanf(buf, "%d %d %d %d %d", &info5[0], &info5[1], &info5[2], &info5[3], &info5[4]) != 5) {
    fclose(f);
    return NULL;
  }
  or_geometry* g = (or_geometry*)calloc(1, sizeof(or_geometry));
  g->npatch = info5[2];
  g->patches = (or_patch*)calloc(g->npatch, sizeof(or_patch));
  for (int p = 0; p < g->npatch; ++p) {
    int info[4];
    next_line(f, buf, sizeof(buf)); /* PATCH k */
    next_line(f, buf, sizeof(buf));
    sscanf(buf, "%d %d", &info[0], &info[1]);
    next_line(f, buf, sizeof(buf));
    sscanf(buf, "%d %d", &info[2], &info[3]);
    int nk1 = info[0] + info[2] + 1, nk2 = info[1] + info[3] + 1;
    double* k1 = (double*)malloc(sizeof(double) * nk1);
    double* k2 = (double*)malloc(sizeof(double) * nk2);
    for (int line = 0; line < 2; ++line) {
      next_line(f, buf, sizeof(buf));
      char* c = buf;
      int nk = line ? nk2 : nk1;
      double* k = line ? k2 : k1;
      for (int i = 0; i < nk; ++i) k[i] = strtod(c, &c);
    }
    int N = info[2], Mr = info[3];
    double* mats[4];
    for (int m = 0; m < 4; ++m) {
      mats[m] = (double*)malloc(sizeof(double) * N * Mr);
      next_line(f, buf, sizeof(buf));
      char* c = buf;
      for (int i = 0; i < Mr; ++i)
        for (int j = 0; j < N; ++j) mats[m][i * N + j] = strtod(c, &c); /* mat(i,j) */
    }
    /* unique knots with tolerance 1e-6 (Spline/Knots.hpp ExtractUniqueKnotVector) */
    int u1 = 1, u2 = 1;
    for (int i = 1; i < nk1; ++i)
      if (k1[i] > k1[i - 1] + 1e-6) ++u1;
    for (int i = 1; i < nk2; ++i)
      if (k2[i] > k2[i - 1] + 1e-6) ++u2;
    if (u1 != 2 || u2 != 2) {
      fprintf(stderr, "oracle: patch %d has interior knots (unsupported)\n", p);
      fclose(f);
      return NULL;
    }
    or_patch* pt = &g->patches[p];
    pt->pdx = nk1 - N;
    pt->pdy = nk2 - Mr;
    pt->data = (double*)malloc(sizeof(double) * 4 * N * Mr);
    for (int m = 0; m < 4; ++m)
      for (int i = 0; i < N; ++i)
        for (int j = 0; j < Mr; ++j) pt->data[(i * Mr + j) * 4 + m] = mats[m][j * N + i];
    if (scale)
      for (int q = 0; q < N * Mr; ++q)
        for (int m = 0; m < 3; ++m) pt->data[4 * q + m] *= scale[m];
    for (int m = 0; m < 4; ++m) free(mats[m]);
    free(k1);
    free(k2);
  }
  fclose(f);
  return g;
}

void or_geometry_free(or_geometry* g) {
  if (!g) return;
  for (int p = 0; p < g->npatch; ++p) free(g->patches[p].data);
  free(g->patches);
  free(g);
}
int or_geometry_npatch(const or_geometry* g) { return g->npatch; }

/* Patch::eval (Patch.hpp:115-155) for a Bezier patch */
void or_patch_eval(const or_geometry* g, int patch, double s, double t, double out[3]) {
  const or_patch* p = &g->patches[patch];
  double xb[32], yb[32];
  or_bernstein(p->pdx - 1, s, xb);
  or_bernstein(p->pdy - 1, t, yb);
  double tmp[4] = {0., 0., 0., 0.};
  for (int i = 0; i < p->pdx; ++i)
    for (int j = 0; j < p->pdy; ++j) {
      double tp = xb[i] * yb[j];
      const double* d = p->data + 4 * (p->pdy * i + j);
      for (int k = 0; k < 4; ++k) tmp[k] += d[k] * tp;
    }
  for (int k = 0; k < 3; ++k) out[k] = tmp[k] / tmp[3];
}

/* Patch::updateSurfacePoint (Patch.hpp:294-365) */
void or_patch_surface_point(const or_geometry* g, int patch, const double st[2], double w,
                            const double xi[2], double out[12]) {
  const or_patch* p = &g->patches[patch];
  double xb[32], yb[32], xd[32], yd[32];
  or_bernstein(p->pdx - 1, st[0], xb);
  or_bernstein(p->pdy - 1, st[1], yb);
  or_bernstein_der(p->pdx - 1, st[0], xd);
  or_bernstein_der(p->pdy - 1, st[1], yd);
  double tmp[4] = {0, 0, 0, 0}, tdx[4] = {0, 0, 0, 0}, tdy[4] = {0, 0, 0, 0};
  for (int i = 0; i < p->pdx; ++i)
    for (int j = 0; j < p->pdy; ++j) {
      double tp = xb[i] * yb[j];
      double tpx = xd[i] * yb[j];
      double tpy = xb[i] * yd[j];
      const double* d = p->data + 4 * (p->pdy * i + j);
      for (int k = 0; k < 4; ++k) {
        tmp[k] += d[k] * tp;
        tdx[k] += d[k] * tpx;
        tdy[k] += d[k] * tpy;
      }
    }
  double bot = 1. / tmp[3];
  double bs = bot * bot;
  out[0] = xi[0];
  out[1] = xi[1];
  out[2] = w;
  for (int k = 0; k < 3; ++k) {
    out[3 + k] = tmp[k] * bot;
    out[6 + k] = (tdx[k] * tmp[3] - tmp[k] * tdx[3]) * bs;
    out[9 + k] = (tdy[k] * tmp[3] - tmp[k] * tdy[3]) * bs;
  }
}

/* ---------------------------------------------------------------- element tree */
static int new_nodes(or_space* s, int n) {
  if (s->nnodes + n > s->cap) {
    s->cap = (s->nnodes + n) * 2;
    s->nodes = (or_node*)realloc(s->nodes, sizeof(or_node) * s->cap);
  }
  int first = s->nnodes;
  for (int i = 0; i < n; ++i) {
    or_node* e = &s->nodes[first + i];
    memset(e, 0, sizeof(*e));
    e->sons = -1;
    e->nsons = 0;
    for (int k = 0; k < 4; ++k) e->adj[k] = -1, e->vert[k] = -1;
    e->radius = INFINITY;
    e->id = -1;
    e->level = -1;
    e->patch = -1;
    e->prev = e->next = -1;
  }
  s->nnodes += n;
  return first;
}

static double norm3(const double* a) { return sqrt((a[0] * a[0] + a[1] * a[1]) + a[2] * a[2]); }

/* ElementTree::updateTopology (ElementTree.hpp:550-582): edge map local to the call */
static void update_topology(or_space* s, const int* el, int n) {
  int keys[64][2], owner[64], nk = 0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < 4; ++j) {
      or_node* e = &s->nodes[el[i]];
      int v1 = e->vert[j], v2 = e->vert[(j + 1) % 4];
      int a = v1 < v2 ? v1 : v2, b = v1 < v2 ? v2 : v1;
      int f = -1;
      for (int q = 0; q < nk; ++q)
        if (keys[q][0] == a && keys[q][1] == b) {
          f = q;
          break;
        }
      if (f >= 0) {
        e->adj[j] = owner[f];
        or_node* o = &s->nodes[owner[f]];
        for (int k = 0; k < 4; ++k) {
          int v3 = o->vert[k], v4 = o->vert[(k + 1) % 4];
          int c = v3 < v4 ? v3 : v4, d = v3 < v4 ? v4 : v3;
          if (a == c && b == d) {
            o->adj[k] = el[i];
            break;
          }
        }
      } else {
        keys[nk][0] = a;
        keys[nk][1] = b;
        owner[nk] = el[i];
        ++nk;
      }
    }
}

static const double llcs[2][4] = {{0., .5, .5, .0}, {0., 0., .5, .5}};

/* ElementTree::refineLeaf (ElementTree.hpp:584-739) */
static void refine_leaf(or_space* s, int cur) {
  if (s->nodes[cur].nsons) return;
  int ptIds[5], refN[4] = {0, 0, 0, 0};
  int elements[16], ne = 0;
  for (int i = 0; i < 4; ++i) {
    int a = s->nodes[cur].adj[i];
    if (a >= 0) {
      for (int j = 0; j < 4; ++j)
        if (s->nodes[a].adj[j] == cur) {
          refN[i] = j;
          break;
        }
    } else {
      refN[i] = -1;
    }
  }
  for (int i = 0; i < 4; ++i) {
    int a = s->nodes[cur].adj[i];
    if (a >= 0 && s->nodes[a].nsons) {
      int r = refN[i];
      int s0 = s->nodes[a].sons;
      ptIds[i] = s->nodes[s0 + r].vert[(r + 1) % 4];
      elements[ne++] = s0 + r;
      elements[ne++] = s0 + (r + 1) % 4;
    } else {
      ptIds[i] = s->npoints++;
    }
  }
  ptIds[4] = s->npoints++;
  int sons = new_nodes(s, 4);
  or_node* c = &s->nodes[cur];
  c->sons = sons;
  c->nsons = 4;
  s->nel += 3;
  if (s->max_level < c->level + 1) s->max_level = c->level + 1;
  for (int i = 0; i < 4; ++i) {
    or_node* son = &s->nodes[sons + i];
    c = &s->nodes[cur];
    if (i == 0) {
      son->prev = c->prev;
      if (c->prev >= 0) s->nodes[c->prev].next = sons;
      if (s->first == cur) s->first = sons;
    } else {
      son->prev = sons + i - 1;
    }
    if (i == 3) {
      son->next = c->next;
      if (c->next >= 0) s->nodes[c->next].prev = sons + 3;
      if (s->last == cur) s->last = sons + 3;
    } else {
      son->next = sons + i + 1;
    }
    son->patch = c->patch;
    son->level = c->level + 1;
    son->id = 4 * c->id + i;
    son->llc[0] = c->llc[0] + llcs[0][i] / (double)(1 << c->level);
    son->llc[1] = c->llc[1] + llcs[1][i] / (double)(1 << c->level);
    elements[ne++] = sons + i;
  }
  c = &s->nodes[cur];
  c->prev = -1;
  c->next = -1;
  const int* v = c->vert;
  int sv[4][4] = {{v[0], ptIds[0], ptIds[4], ptIds[3]},
                  {ptIds[0], v[1], ptIds[1], ptIds[4]},
                  {ptIds[4], ptIds[1], v[2], ptIds[2]},
                  {ptIds[3], ptIds[4], ptIds[2], v[3]}};
  for (int i = 0; i < 4; ++i)
    for (int k = 0; k < 4; ++k) s->nodes[sons + i].vert[k] = sv[i][k];
  update_topology(s, elements, ne);
}

static void refine_recursion(or_space* s, int n) {
  int ns = s->nodes[n].nsons;
  if (ns) {
    int s0 = s->nodes[n].sons;
    for (int i = 0; i < ns; ++i) refine_recursion(s, s0 + i);
  } else {
    refine_leaf(s, n);
  }
}

/* util::computeEnclosingBall (util/GeometryHelper.hpp:32-52) */
static void enclosing_ball(double* mid, double* r, const double* m1, double r1, const double* m2,
                           double r2) {
  double z[3] = {m1[0] - m2[0], m1[1] - m2[1], m1[2] - m2[2]};
  double nrm = norm3(z);
  if (nrm + r2 <= r1) {
    for (int k = 0; k < 3; ++k) mid[k] = m1[k];
    *r = r1;
  } else if (nrm + r1 <= r2) {
    for (int k = 0; k < 3; ++k) mid[k] = m2[k];
    *r = r2;
  } else {
    double f = (r1 - r2) / nrm;
    for (int k = 0; k < 3; ++k) mid[k] = 0.5 * (m1[k] + m2[k] + f * z[k]);
    *r = 0.5 * (r1 + r2 + nrm);
  }
}

/* ElementTree::computeElementEnclosings_recursion (ElementTree.hpp:366-396) */
static void enclosings(or_space* s, int n) {
  double mp1[3], mp2[3], r1, r2;
  or_node* e = &s->nodes[n];
  if (!e->nsons) {
    const double* P = s->points;
    enclosing_ball(mp1, &r1, P + 3 * e->vert[0], 0, P + 3 * e->vert[2], 0);
    enclosing_ball(mp2, &r2, P + 3 * e->vert[1], 0, P + 3 * e->vert[3], 0);
    enclosing_ball(e->mid, &e->radius, mp1, r1, mp2, r2);
  } else {
    int s0 = e->sons;
    for (int i = 0; i < 4; ++i) enclosings(s, s0 + i);
    e = &s->nodes[n];
    or_node* so = &s->nodes[s0];
    enclosing_ball(mp1, &r1, so[0].mid, so[0].radius, so[2].mid, so[2].radius);
    enclosing_ball(mp2, &r2, so[1].mid, so[1].radius, so[3].mid, so[3].radius);
    enclosing_ball(e->mid, &e->radius, mp1, r1, mp2, r2);
  }
}

static void collect_clusters(or_space* s, int n) {
  or_node* e = &s->nodes[n];
  if (e->level >= 0 && e->level <= s->M) s->cluster[e->level][e->id] = n;
  for (int i = 0; i < e->nsons; ++i) collect_clusters(s, e->sons + i);
}

/* ElementTree::init_ElementTree (ElementTree.hpp:105-161) + computeElementEnclosings */
static void build_tree(or_space* s) {
  or_geometry* g = s->geo;
  int np = g->npatch;
  s->cap = 64;
  s->nodes = (or_node*)malloc(sizeof(or_node) * s->cap);
  s->nnodes = 0;
  s->root = new_nodes(s, 1);
  int sons = new_nodes(s, np);
  s->nodes[s->root].sons = sons;
  s->nodes[s->root].nsons = np;
  s->npoints = 0;
  s->nel = np;
  s->max_level = s->M;
  double* uniq = (double*)malloc(sizeof(double) * 3 * 4 * np);
  int nuniq = 0;
  static const double corners[2][4] = {{0., 1., 1., 0}, {0., 0., 1., 1.}};
  for (int i = 0; i < np; ++i) {
    or_node* e = &s->nodes[sons + i];
    e->id = i;
    e->level = 0;
    e->patch = i;
    e->llc[0] = e->llc[1] = 0.;
    e->prev = i == 0 ? -1 : sons + i - 1;
    e->next = i == np - 1 ? -1 : sons + i + 1;
    for (int j = 0; j < 4; ++j) {
      double v[3];
      or_patch_eval(g, i, corners[0][j], corners[1][j], v);
      int idx = 0;
      for (; idx < nuniq; ++idx) {
        double d[3] = {uniq[3 * idx] - v[0], uniq[3 * idx + 1] - v[1], uniq[3 * idx + 2] - v[2]};
        if (norm3(d) < 1e-9) break;
      }
      if (idx != nuniq) {
        e->vert[j] = idx;
      } else {
        memcpy(uniq + 3 * nuniq, v, sizeof(v));
        ++nuniq;
        e->vert[j] = s->npoints++;
      }
    }
  }
  s->first = sons;
  s->last = sons + np - 1;
  free(uniq);
  int* els = (int*)malloc(sizeof(int) * np);
  for (int i = 0; i < np; ++i) els[i] = sons + i;
  /* patch-level topology uses the same edge matching with a larger table */
  {
    int (*keys)[2] = malloc(sizeof(int) * 2 * 4 * np);
    int* owner = malloc(sizeof(int) * 4 * np);
    int nk = 0;
    for (int i = 0; i < np; ++i)
      for (int j = 0; j < 4; ++j) {
        or_node* e = &s->nodes[els[i]];
        int v1 = e->vert[j], v2 = e->vert[(j + 1) % 4];
        int a = v1 < v2 ? v1 : v2, b = v1 < v2 ? v2 : v1;
        int f = -1;
        for (int q = 0; q < nk; ++q)
          if (keys[q][0] == a && keys[q][1] == b) {
            f = q;
            break;
          }
        if (f >= 0) {
          e->adj[j] = owner[f];
          or_node* o = &s->nodes[owner[f]];
          for (int k = 0; k < 4; ++k) {
            int v3 = o->vert[k], v4 = o->vert[(k + 1) % 4];
            int c = v3 < v4 ? v3 : v4, d = v3 < v4 ? v4 : v3;
            if (a == c && b == d) {
              o->adj[k] = els[i];
              break;
            }
          }
        } else {
          keys[nk][0] = a;
          keys[nk][1] = b;
          owner[nk] = els[i];
          ++nk;
        }
      }
    free(keys);
    free(owner);
  }
  free(els);
  for (int l = 0; l < s->M; ++l) refine_recursion(s, s->root);
  /* Morton leaf list */
  s->elem = (int*)malloc(sizeof(int) * s->nel);
  int cnt = 0;
  for (int n = s->first; n >= 0; n = s->nodes[n].next) s->elem[cnt++] = n;
  /* generatePointList (ElementTree.hpp:232-256): last writer wins */
  s->points = (double*)calloc(3 * s->npoints, sizeof(double));
  for (int e = 0; e < s->nel; ++e) {
    or_node* n = &s->nodes[s->elem[e]];
    double h = 1. / (double)(1 << n->level);
    or_patch_eval(g, n->patch, n->llc[0], n->llc[1], s->points + 3 * n->vert[0]);
    or_patch_eval(g, n->patch, n->llc[0] + h, n->llc[1], s->points + 3 * n->vert[1]);
    or_patch_eval(g, n->patch, n->llc[0] + h, n->llc[1] + h, s->points + 3 * n->vert[2]);
    or_patch_eval(g, n->patch, n->llc[0], n->llc[1] + h, s->points + 3 * n->vert[3]);
  }
  for (int i = 0; i < np; ++i) enclosings(s, s->nodes[s->root].sons + i);
  s->cluster = (int**)malloc(sizeof(int*) * (s->M + 1));
  for (int l = 0; l <= s->M; ++l) s->cluster[l] = (int*)malloc(sizeof(int) * np * (1 << (2 * l)));
  collect_clusters(s, s->root);
}

/* SuperSpace::map2surface (AnsatzSpace/SuperSpace.hpp:160-165) */
void or_map2surface(const or_space* s, int node, const double xi[2], double w, double out[12]) {
  const or_node* e = &s->nodes[node];
  double h = 1. / (double)(1 << e->level);
  double st[2] = {e->llc[0] + h * xi[0], e->llc[1] + h * xi[1]};
  or_patch_surface_point(s->geo, e->patch, st, w, xi, out);
}

/* ---------------------------------------------------------------- projector */
/* Spl::DeBoor for a scalar coefficient vector (Spline/DeBoor.hpp:67-110) */
static double deboor(const double* coef, int ncoef, const double* knots, int nknots, double x) {
  int p = nknots - ncoef - 1;
  int l = 0;
  while (knots[l] <= x && l != ncoef) l++;
  l = l - 1;
  double temp[32], ws[32];
  for (int i = 0; i <= p; ++i) temp[i] = coef[l - p + i];
  for (int k = p; k != 0; k--) {
    for (int i = 0; i < k; i++)
      ws[i] = (x - knots[l - k + 1 + i]) / (knots[l + 1 + i] - knots[l - k + 1 + i]);
    for (int i = 0; i < k; i++) temp[i] = ((ws[i] * temp[1 + i]) + ((1 - ws[i]) * temp[i]));
  }
  return temp[0];
}

/* Spl::MakeUniformKnotVector (Spline/Knots.hpp) */
static int uniform_knots(int pp1, int interior, int rep, double* out) {
  int n = 0;
  double h = 1. / (interior + 1);
  for (int i = 0; i < pp1; ++i) out[n++] = 0;
  for (int i = 1; i < interior + 1; ++i)
    for (int k = 0; k < rep; ++k) out[n++] = i * h;
  for (int i = 0; i < pp1; ++i) out[n++] = 1;
  return n;
}

/* dense LU with partial pivoting (stands in for Eigen::PartialPivLU, Projector.hpp:238) */
static void lu_factor(int n, double* A, int* piv) {
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = fabs(A[k * n + k]);
    for (int i = k + 1; i < n; ++i)
      if (fabs(A[i * n + k]) > best) best = fabs(A[i * n + k]), p = i;
    piv[k] = p;
    if (p != k)
      for (int j = 0; j < n; ++j) {
        double t = A[k * n + j];
        A[k * n + j] = A[p * n + j];
        A[p * n + j] = t;
      }
    for (int i = k + 1; i < n; ++i) {
      A[i * n + k] /= A[k * n + k];
      for (int j = k + 1; j < n; ++j) A[i * n + j] -= A[i * n + k] * A[k * n + j];
    }
  }
}
static void lu_solve(int n, const double* A, const int* piv, double* b) {
  for (int k = 0; k < n; ++k)
    if (piv[k] != k) {
      double t = b[k];
      b[k] = b[piv[k]];
      b[piv[k]] = t;
    }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < i; ++j) b[i] -= A[i * n + j] * b[j];
  for (int i = n - 1; i >= 0; --i) {
    for (int j = i + 1; j < n; ++j) b[i] -= A[i * n + j] * b[j];
    b[i] /= A[i * n + i];
  }
}

/* ProjectorRoutines::makeLocalProjectionTriplets (AnsatzSpace/Projector.hpp:165-337) */
static void local_projection(const or_space* s, int pp1x, int pp1y, or_triplets* out, int* ncdof,
                             int* ndcdof) {
  int M = s->M, mpd = pp1x > pp1y ? pp1x : pp1y, minp = pp1x < pp1y ? pp1x : pp1y;
  int rep = (mpd <= 1) ? minp : 1;
  int n = 1 << M, np = s->geo->npatch;
  double* kx = (double*)malloc(sizeof(double) * (2 * pp1x + n + 2));
  double* ky = (double*)malloc(sizeof(double) * (2 * pp1y + n + 2));
  int nkx = uniform_knots(pp1x, n - 1, rep, kx);
  int nky = uniform_knots(pp1y, n - 1, rep, ky);
  int dimx = nkx - pp1x, dimy = nky - pp1y, dim = dimx * dimy;
  double mask[32];
  for (int i = 0; i < mpd; ++i) mask[i] = (i + 1) * (1. / (mpd + 1));
  int ms = mpd, sz = ms * ms;
  double* sys = (double*)malloc(sizeof(double) * sz * sz);
  for (int iy = 0; iy < mpd; ++iy) {
    double vy[32], vx[32];
    or_bernstein(mpd - 1, mask[iy], vy);
    for (int ix = 0; ix < mpd; ++ix) {
      or_bernstein(mpd - 1, mask[ix], vx);
      for (int jy = 0; jy < mpd; ++jy)
        for (int jx = 0; jx < mpd; ++jx) sys[(iy * ms + ix) * sz + jy * mpd + jx] = vx[jx] * vy[jy];
    }
  }
  int* piv = (int*)malloc(sizeof(int) * sz);
  lu_factor(sz, sys, piv);
  int cap = s->nel * sz * sz + 16;
  out->row = (int*)malloc(sizeof(int) * cap);
  out->col = (int*)malloc(sizeof(int) * cap);
  out->val = (double*)malloc(sizeof(double) * cap);
  out->n = 0;
  double* cx = (double*)calloc(dimx, sizeof(double));
  double* cy = (double*)calloc(dimy, sizeof(double));
  for (int e = 0; e < s->nel; ++e) {
    const or_node* el = &s->nodes[s->elem[e]];
    double px = el->llc[0], py = el->llc[1];
    double h = 1. / (double)(1 << el->level);
    double mx = px + (.5 * h), my = py + (.5 * h);
    int nzx[32], nzy[32], nnx = 0, nny = 0;
    for (int d = 0; d < dimy; ++d)
      if (ky[d] <= my && ky[d + pp1y] >= my) nzy[nny++] = d;
    for (int d = 0; d < dimx; ++d)
      if (kx[d] <= mx && kx[d + pp1x] >= mx) nzx[nnx++] = d;
    for (int a = 0; a < nny; ++a)
      for (int b = 0; b < nnx; ++b) {
        int dy = nzy[a], dx = nzx[b];
        double lmx[32], lmy[32], vx[32], vy[32], rhs[1024];
        for (int i = 0; i < ms; ++i) {
          lmx[i] = px + (1.0 / n) * mask[i];
          lmy[i] = py + (1.0 / n) * mask[i];
        }
        cx[dx] = 1.;
        cy[dy] = 1.;
        for (int i = 0; i < ms; ++i) {
          vx[i] = deboor(cx, dimx, kx, nkx, lmx[i]);
          vy[i] = deboor(cy, dimy, ky, nky, lmy[i]);
        }
        cx[dx] = 0.;
        cy[dy] = 0.;
        for (int iy = 0; iy < ms; ++iy)
          for (int ix = 0; ix < ms; ++ix) rhs[iy * ms + ix] = vx[ix] * vy[iy];
        lu_solve(sz, sys, piv, rhs);
        for (int i = 0; i < sz; ++i)
          if (fabs(rhs[i]) > 1e-4) {
            out->col[out->n] = el->patch * dim + dimx * dy + dx;
            out->row[out->n] = e * sz + i;
            out->val[out->n] = rhs[i];
            ++out->n;
          }
      }
  }
  *ncdof = dim * np;
  *ndcdof = sz * n * n * np;
  free(cx);
  free(cy);
  free(piv);
  free(sys);
  free(kx);
  free(ky);
}

static int cmp_trip(const void* a, const void* b) {
  const long* x = (const long*)a;
  const long* y = (const long*)b;
  return (*x > *y) - (*x < *y);
}

/* triplets -> CSR, summing duplicates (Eigen::SparseMatrix::setFromTriplets) */
static void to_csr(or_space* s, const or_triplets* t, int nrows, int ncols) {
  long* key = (long*)malloc(sizeof(long) * 2 * t->n);
  for (int i = 0; i < t->n; ++i) {
    key[2 * i] = (long)t->row[i] * ncols + t->col[i];
    key[2 * i + 1] = i;
  }
  qsort(key, t->n, 2 * sizeof(long), cmp_trip);
  s->rowptr = (int*)calloc(nrows + 1, sizeof(int));
  s->colidx = (int*)malloc(sizeof(int) * (t->n + 1));
  s->vals = (double*)malloc(sizeof(double) * (t->n + 1));
  int nnz = 0;
  long last = -1;
  for (int i = 0; i < t->n; ++i) {
    long k = key[2 * i];
    int src = (int)key[2 * i + 1];
    if (k == last) {
      s->vals[nnz - 1] += t->val[src];
    } else {
      s->colidx[nnz] = (int)(k % ncols);
      s->vals[nnz] = t->val[src];
      s->rowptr[k / ncols + 1]++;
      ++nnz;
      last = k;
    }
  }
  for (int r = 0; r < nrows; ++r) s->rowptr[r + 1] += s->rowptr[r];
  s->nnz = nnz;
  s->nsuper = nrows;
  s->ndof = ncols;
  free(key);
}

void or_glue_divconforming(or_space* s, or_triplets* t, int* ncols);

or_space* or_space_create(or_geometry* g, int M, int P, int form) {
  if (!g || M < 0 || P < 0 || P > 18) return NULL;
  or_space* s = (or_space*)calloc(1, sizeof(or_space));
  s->geo = g;
  s->M = M;
  s->P = P;
  s->I2 = (P + 1) * (P + 1);
  s->form = form;
  s->vd = form == 1 ? 2 : 1;
  build_tree(s);
  or_triplets t;
  int ncdof, ndcdof;
  if (form == 0) {
    local_projection(s, P + 1, P + 1, &t, &ncdof, &ndcdof);
    to_csr(s, &t, ndcdof, ncdof);
  } else {
    or_triplets t1, t2;
    int nc1, nd1, nc2, nd2;
    local_projection(s, P + 1, P, &t1, &nc1, &nd1);
    local_projection(s, P, P + 1, &t2, &nc2, &nd2);
    t.n = t1.n + t2.n;
    t.row = (int*)malloc(sizeof(int) * t.n);
    t.col = (int*)malloc(sizeof(int) * t.n);
    t.val = (double*)malloc(sizeof(double) * t.n);
    for (int k = 0; k < t1.n; ++k) t.row[k] = t1.row[k], t.col[k] = t1.col[k], t.val[k] = t1.val[k];
    for (int k = 0; k < t2.n; ++k) {
      t.row[t1.n + k] = t2.row[k] + nd1;
      t.col[t1.n + k] = t2.col[k] + nc1;
      t.val[t1.n + k] = t2.val[k];
    }
    free(t1.row), free(t1.col), free(t1.val), free(t2.row), free(t2.col), free(t2.val);
    int ncols = nc1 + nc2;
    or_glue_divconforming(s, &t, &ncols);
    to_csr(s, &t, nd1 + nd2, ncols);
  }
  free(t.row);
  free(t.col);
  free(t.val);
  or_build_ffield(s);
  return s;
}

void or_space_free(or_space* s) {
  if (!s) return;
  free(s->nodes);
  free(s->elem);
  if (s->cluster) {
    for (int l = 0; l <= s->M; ++l) free(s->cluster[l]);
    free(s->cluster);
  }
  free(s->points);
  free(s->rowptr);
  free(s->colidx);
  free(s->vals);
  free(s->ffield);
  free(s);
}

void or_space_info(const or_space* s, int info[8]) {
  info[0] = s->nel;
  info[1] = s->npoints;
  info[2] = s->I2;
  info[3] = s->vd;
  info[4] = s->nsuper;
  info[5] = s->ndof;
  info[6] = s->nnz;
  info[7] = s->M;
}

void or_space_elements(const or_space* s, int* vert, int* patch, double* llc, double* mid,
                       double* radius) {
  for (int e = 0; e < s->nel; ++e) {
    const or_node* n = &s->nodes[s->elem[e]];
    for (int k = 0; k < 4; ++k) vert[4 * e + k] = n->vert[k];
    patch[e] = n->patch;
    llc[2 * e] = n->llc[0];
    llc[2 * e + 1] = n->llc[1];
    for (int k = 0; k < 3; ++k) mid[3 * e + k] = n->mid[k];
    radius[e] = n->radius;
  }
}

int or_space_clusters(const or_space* s, int level, double* mid, double* radius) {
  if (level < 0 || level > s->M) return 0;
  int nc = s->geo->npatch * (1 << (2 * level));
  for (int c = 0; c < nc; ++c) {
    const or_node* n = &s->nodes[s->cluster[level][c]];
    for (int k = 0; k < 3; ++k) mid[3 * c + k] = n->mid[k];
    radius[c] = n->radius;
  }
  return nc;
}

void or_space_T(const or_space* s, int* rowptr, int* col, double* val) {
  memcpy(rowptr, s->rowptr, sizeof(int) * (s->nsuper + 1));
  memcpy(col, s->colidx, sizeof(int) * s->nnz);
  memcpy(val, s->vals, sizeof(double) * s->nnz);
}

/* ---------------------------------------------------------------- glue */
/* ElementTree::patchTopologyInfo (ClusterTree/ElementTree.hpp:510-533):
   rows of {patch1, patch2, edge1, edge2}, patch2 = -1 on a boundary edge */
static int patch_topology(const or_space* s, int (*out)[4]) {
  int n = 0, np = s->geo->npatch, s0 = s->nodes[s->root].sons;
  for (int i = 0; i < np; ++i) {
    const or_node* it = &s->nodes[s0 + i];
    for (int j = 0; j < 4; ++j) {
      if (it->adj[j] >= 0) {
        const or_node* nb = &s->nodes[it->adj[j]];
        if (it->id < nb->id) {
          int k = 0;
          for (; k < 4; ++k)
            if (nb->adj[k] >= 0 && s->nodes[nb->adj[k]].id == it->id) break;
          out[n][0] = it->id, out[n][1] = nb->id, out[n][2] = j, out[n][3] = k;
          ++n;
        }
      } else {
        out[n][0] = it->id, out[n][1] = -1, out[n][2] = j, out[n][3] = -1;
        ++n;
      }
    }
  }
  return n;
}

/* GlueRoutines::getEdgeDofIndices (AnsatzSpace/Glue.hpp:249-296) */
static int edge_dofs(int ec, int dx, int dy, int shift, int* out) {
  int n = 0;
  switch (ec) {
    case 0:
      for (int i = 0; i < dx; ++i) out[n++] = i + shift;
      break;
    case 1:
      for (int i = 0; i < dy; ++i) out[n++] = dx * (i + 1) - 1 + shift;
      break;
    case 2:
      for (int i = 0; i < dx; ++i) out[n++] = dx * (dy - 1) + i + shift;
      break;
    case 3:
      for (int i = 0; i < dy; ++i) out[n++] = dx * i + shift;
      break;
  }
  return n;
}
static int fwd(int e) { return e == 0 || e == 1; }
static int inward(int e) { return e == 0 || e == 3; }
static int first_comp(int e) { return !(e == 0 || e == 2); }

/* DivConforming identification (Glue.hpp:528-598) and assembleGlueMatrix (:153-232),
   applied to the projector triplets: column c -> glued column with coefficient */
void or_glue_divconforming(or_space* s, or_triplets* t, int* ncols) {
  int np = s->geo->npatch, n = 1 << s->M, P = s->P;
  int small_dim = P + (n - 1), large_dim = P + 1 + (n - 1);
  int dpp = small_dim * large_dim;
  int pre = *ncols;
  int (*edges)[4] = malloc(sizeof(int) * 4 * 4 * np);
  int ne = patch_topology(s, edges);
  /* identification sets: master (smaller), partner (or -1), coef */
  int* partner = (int*)malloc(sizeof(int) * pre);
  int* coefs = (int*)malloc(sizeof(int) * pre);
  char* is_slave = (char*)calloc(pre, 1);
  char* is_master = (char*)calloc(pre, 1);
  char* boundary = (char*)calloc(pre, 1);
  for (int i = 0; i < pre; ++i) partner[i] = -1, coefs[i] = 1;
  int* d1 = (int*)malloc(sizeof(int) * (large_dim + 2));
  int* d2 = (int*)malloc(sizeof(int) * (large_dim + 2));
  for (int q = 0; q < ne; ++q) {
    int* e = edges[q];
    int sh1 = e[0] * dpp + (first_comp(e[2]) ? 0 : dpp * np);
    int xd1 = first_comp(e[2]) ? large_dim : small_dim;
    int yd1 = first_comp(e[2]) ? small_dim : large_dim;
    int m = edge_dofs(e[2], xd1, yd1, sh1, d1);
    if (e[1] > -1 && e[0] > -1) {
      int sh2 = e[1] * dpp + (first_comp(e[3]) ? 0 : dpp * np);
      int xd2 = first_comp(e[3]) ? large_dim : small_dim;
      int yd2 = first_comp(e[3]) ? small_dim : large_dim;
      edge_dofs(e[3], xd2, yd2, sh2, d2);
      int rev = (fwd(e[2]) && fwd(e[3])) || (!fwd(e[2]) && !fwd(e[3]));
      int coef = ((inward(e[2]) && inward(e[3])) || (!inward(e[2]) && !inward(e[3]))) ? -1 : 1;
      for (int i = 0; i < m; ++i) {
        int j = rev ? m - 1 - i : i;
        int a = d1[i] < d2[j] ? d1[i] : d2[j], b = d1[i] < d2[j] ? d2[j] : d1[i];
        is_master[a] = 1;
        is_slave[b] = 1;
        partner[a] = b;
        coefs[a] = coef;
      }
    } else {
      for (int i = 0; i < m; ++i) is_master[d1[i]] = 1, boundary[d1[i]] = 1;
    }
  }
  /* post index per pre dof, and coefficient */
  int* post = (int*)malloc(sizeof(int) * pre);
  double* pc = (double*)malloc(sizeof(double) * pre);
  for (int i = 0; i < pre; ++i) post[i] = -1, pc[i] = 0;
  int pi = 0;
  for (int i = 0; i < pre; ++i) {
    if (is_slave[i]) continue;
    if (is_master[i] && boundary[i]) continue;
    post[i] = pi;
    pc[i] = 1.;
    if (is_master[i] && partner[i] >= 0) {
      post[partner[i]] = pi;
      pc[partner[i]] = (double)coefs[i];
    }
    ++pi;
  }
  int w = 0;
  for (int k = 0; k < t->n; ++k) {
    int c = t->col[k];
    if (post[c] < 0) continue;
    t->row[w] = t->row[k];
    t->col[w] = post[c];
    t->val[w] = t->val[k] * pc[c];
    ++w;
  }
  t->n = w;
  *ncols = pi;
  free(edges), free(partner), free(coefs), free(is_slave), free(is_master), free(boundary);
  free(d1), free(d2), free(post), free(pc);
}
