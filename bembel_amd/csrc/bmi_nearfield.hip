// Near-field Galerkin quadrature on the GPU: one workgroup per element pair.
//
// Restates DuffyTrick::integrate0..4 (DuffyTrick/integrate0.hpp:22-38, integrate1.hpp:25-48,
// integrate2.hpp:27-67, integrate3.hpp:27-80, integrate4.hpp:27-64) for the Laplace and
// Helmholtz single-layer integrands (Laplace/SingleLayerOperator.hpp:44-85,
// Helmholtz/SingleLayerOperator.hpp:44-85), the Maxwell single layer (EFIE,
// Maxwell/SingleLayerOperator.hpp:44-88) and the test operator
// (LinearOperator/Dummy/DummyOperator.hpp:53-59).
//
// Design: all lanes of a workgroup work on the same element pair, so the patch control
// points are wave-uniform and are read through the scalar cache.  Each lane evaluates a
// strided share of the quadrature index space (i, j), computes its surface points in
// registers and accumulates the (P+1)^2 x (P+1)^2 block in registers; the block is reduced
// across the workgroup at the end.  The work is FP64-VALU bound (rational Bezier surface
// points dominate), which is why FP64 MFMA (same peak as FP64 VALU on this chip, 9x9 blocks
// padded to 16x16) is not used for the block accumulation.  Maxwell blocks have four
// (P+1)^2 x (P+1)^2 complex component blocks; one workgroup computes one component block
// (blockIdx.y = 2*row component + column component) so the accumulators stay in registers.
#include <hip/hip_runtime.h>

#include "bmi_internal.h"
#include "bmi_device.h"

namespace bmi {

struct NFParams {
  const double* cp;  // control points of all patches
  int cp_stride;     // doubles per patch
  const int* elem_patch;
  const double* elem_llc;
  double h;        // element width in the parameter domain
  double kre, kim;  // wavenumber
  int dummy_fn;
  const double* ffield;  // far-field table, FF_STRIDE doubles per point
  int nff;               // points per element in the table
  int ffs;               // doubles per point in the table
  double* out;
  int64_t rs, cs;  // row / column stride of a block in the output (in scalars)
  int64_t comp_rs, comp_cs;  // offsets of the vector components (Maxwell)
  double qre, qim;           // -1 / (kappa^2 h^2): divergence-part factor (Maxwell)
};

// far-field table record: x0 x1 x2 kappa w s0 s1 pad [ds0 ds1 ds2 dt0 dt1 dt2 pad pad]

// 1 / sqrt(s), s > 0: the hardware estimate (v_rsq_f64) refined by two steps of Newton's
// iteration for 1/sqrt, y <- y + y (1 - s y^2) / 2, each of which squares the relative error
// (the residual 1 - s y^2 formed with one rounding by an FMA); within a few ulps of the rounded
// quotient 1 / sqrt(s), for a third of the instructions of a square root and a division
__device__ __forceinline__ double rsqrt_newton(double s) {
  double y = __builtin_amdgcn_rsq(s);
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    const double e = fma(-(s * y), y, 1.0);
    y = fma(0.5 * y, e, y);
  }
  return y;
}

// 1 / x, x != 0: the hardware estimate (v_rcp_f64) refined by two steps of Newton's iteration
// for the reciprocal, y <- y + y (1 - x y) (residual by an FMA); within an ulp or two of the
// rounded quotient, for half the instructions of a division
__device__ __forceinline__ double rcp_newton(double x) {
  double y = __builtin_amdgcn_rcp(x);
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    const double e = fma(-x, y, 1.0);
    y = fma(y, e, y);
  }
  return y;
}

// evaluation point of the integrand: physical point, weight (kappa*w or w), local coords;
// Maxwell also keeps the two tangent vectors
template <int KIND>
struct Pt {
  double x0, x1, x2, W, s0, s1;
};
template <>
struct Pt<BMI_MAXWELL_SL> {
  double x0, x1, x2, W, s0, s1;
  double t[6];  // d/ds x, d/dt x
};

template <int PD, int KIND>
__device__ __forceinline__ Pt<KIND> make_point(const double* __restrict__ cp, double lx, double ly,
                                               double h, double u, double v, double w) {
  double X[3], Ds[3], Dt[3];
  surface_point<PD>(cp, lx + h * u, ly + h * v, X, Ds, Dt);
  Pt<KIND> p;
  p.x0 = X[0], p.x1 = X[1], p.x2 = X[2];
  p.s0 = u, p.s1 = v;
  if constexpr (KIND == BMI_MAXWELL_SL) {
    // the vector basis carries no surface measure: the Piola factor cancels
    // (Maxwell/SingleLayerOperator.hpp:71-85)
    p.W = w;
#pragma unroll
    for (int k = 0; k < 3; ++k) p.t[k] = Ds[k], p.t[3 + k] = Dt[k];
  } else if (KIND == BMI_DUMMY) {
    p.W = w;
  } else {
    const double c0 = Ds[1] * Dt[2] - Ds[2] * Dt[1];
    const double c1 = Ds[2] * Dt[0] - Ds[0] * Dt[2];
    const double c2 = Ds[0] * Dt[1] - Ds[1] * Dt[0];
    p.W = sqrt(c0 * c0 + c1 * c1 + c2 * c2) * w;
  }
  return p;
}

// test functions of the reference Duffy unit tests (tests/DuffyTrick/test_testFunctions.hpp)
template <class PT>
__device__ __forceinline__ double dummy_value(int fn, const PT& a, const PT& b) {
  if (fn == 0)
    return sin(M_PI * a.x0) * sin(2 * M_PI * b.x0) * exp(-a.x1) * exp(b.x1);
  if (fn == 1) {
    const double c = a.x1 - b.x0;
    return exp(-a.x0 - 2 * b.x1) * (1. / (1e-3 + c * c));
  }
  return 1.0;
}

// Block accumulation split over groups of G lanes.  Each lane evaluates its own integrand point
// (surface points, kernel, Bernstein values); the G lanes of a group exchange those values and
// each adds all G evaluations to its part of the block: for complex operators the lane role
// (lane mod G) picks real or imaginary parts (role & 1) and, with G = 4, a half of the rows
// (role >> 1).  Per lane this divides the accumulator registers by G at the same number of
// multiply-adds; real blocks (G = 1) stay whole in every lane.  Complex blocks use G = 2 (a
// four-way split halves the accumulators again but doubles the exchanges, which measured
// slower for P = 2).
template <int P, int KIND>
struct BlockAcc {
  static constexpr int I = P + 1;
  static constexpr int I2 = I * I;
  static constexpr bool CPLX = KIND == BMI_HELMHOLTZ_SL || KIND == BMI_MAXWELL_SL;
  static constexpr bool MAXW = KIND == BMI_MAXWELL_SL;
  // lanes per group: two lanes hold a block, real blocks split into two halves of their rows,
  // complex ones into their real and imaginary parts (splitting complex blocks into four lanes
  // as well measured slower: Helmholtz edge rule 182 -> 303 ms, more sums exchanged per pass)
  static constexpr int G = 2;
  static constexpr int RH = CPLX ? G / 2 : G;         // row halves
  static constexpr int NR = (I2 + RH - 1) / RH;       // rows held per lane
  static constexpr int ID = MAXW ? I : 1;
  static constexpr int TOT = CPLX ? 2 * I2 * I2 : I2 * I2;
  double acc[NR * I2];

  // what the block update needs from one evaluation
  struct Eval {
    double vr, vi;  // kernel x weights (x tangent product for Maxwell)
    double qr, qi;  // Maxwell divergence factor
    double X[I], Y[I], U[I], V[I];
    double DX[ID], DY[ID], DU[ID], DV[ID];
  };

  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int k = 0; k < NR * I2; ++k) acc[k] = 0.0;
  }

  // one evaluation of the integrand; a lane without work (valid = false) contributes zero
  __device__ __forceinline__ Eval eval(const Pt<KIND>& a, const Pt<KIND>& b, const NFParams& prm, int rc, int cc,
                                       bool valid) const {
    Eval e;
    const double d0 = a.x0 - b.x0, d1 = a.x1 - b.x1, d2 = a.x2 - b.x2;
    const double s = d0 * d0 + d1 * d1 + d2 * d2;
    const double ri = rsqrt_newton(s);  // 1 / r
    const double r = s * ri;
    e.vi = 0.0, e.qr = 0.0, e.qi = 0.0;
    if constexpr (KIND == BMI_LAPLACE_SL) {
      e.vr = (a.W * b.W) * (0.25 * M_1_PI) * ri;
    } else if constexpr (CPLX) {
      // exp(-i kappa r) / (4 pi r), kappa = kre + i kim (Helmholtz/SingleLayerOperator.hpp:115-120)
      double sn, cs;
      sincos(prm.kre * r, &sn, &cs);
      const double damp = prm.kim != 0.0 ? exp(prm.kim * r) : 1.0;
      const double mag = damp * (a.W * b.W) * (0.25 * M_1_PI) * ri;
      const double kr = mag * cs, ki = -mag * sn;
      if constexpr (MAXW) {
        // vector part K w w (t_rc(a) . t_cc(b)), divergence part -K w w / (kappa^2 h^2)
        // (Maxwell/SingleLayerOperator.hpp:44-88)
        const double* ta = a.t + 3 * rc;
        const double* tb = b.t + 3 * cc;
        const double dt = ta[0] * tb[0] + ta[1] * tb[1] + ta[2] * tb[2];
        e.vr = kr * dt, e.vi = ki * dt;
        e.qr = kr * prm.qre - ki * prm.qim, e.qi = kr * prm.qim + ki * prm.qre;
      } else {
        e.vr = kr, e.vi = ki;
      }
    } else {
      e.vr = dummy_value(prm.dummy_fn, a, b) * a.W * b.W;
    }
    if (!valid) e.vr = 0.0, e.vi = 0.0, e.qr = 0.0, e.qi = 0.0;
    if constexpr (MAXW) {
      bernstein_basis_der<P>(a.s0, e.X, e.DX);
      bernstein_basis_der<P>(a.s1, e.Y, e.DY);
      bernstein_basis_der<P>(b.s0, e.U, e.DU);
      bernstein_basis_der<P>(b.s1, e.V, e.DV);
    } else {
      bernstein_basis<P>(a.s0, e.X);
      bernstein_basis<P>(a.s1, e.Y);
      bernstein_basis<P>(b.s0, e.U);
      bernstein_basis<P>(b.s1, e.V);
    }
    return e;
  }

  // the evaluation held by lane `src` of the wave
  __device__ __forceinline__ static Eval fetch(const Eval& e, int src) {
    Eval o;
    o.vr = __shfl(e.vr, src, 64);
    if constexpr (CPLX) o.vi = __shfl(e.vi, src, 64);
    if constexpr (MAXW) o.qr = __shfl(e.qr, src, 64), o.qi = __shfl(e.qi, src, 64);
#pragma unroll
    for (int k = 0; k < I; ++k) {
      o.X[k] = __shfl(e.X[k], src, 64), o.Y[k] = __shfl(e.Y[k], src, 64);
      o.U[k] = __shfl(e.U[k], src, 64), o.V[k] = __shfl(e.V[k], src, 64);
      if constexpr (MAXW) {
        o.DX[k] = __shfl(e.DX[k], src, 64), o.DY[k] = __shfl(e.DY[k], src, 64);
        o.DU[k] = __shfl(e.DU[k], src, 64), o.DV[k] = __shfl(e.DV[k], src, 64);
      }
    }
    return o;
  }

  // add one evaluation to this lane's part of the block
  __device__ __forceinline__ void accumulate(const Eval& e, int role, int rc, int cc) {
    const int part = CPLX ? (role & 1) : 0;
    const int half = CPLX ? (role >> 1) : role;
    double bv[I2], bd[MAXW ? I2 : 1], av[I2], ad[MAXW ? I2 : 1];
#pragma unroll
    for (int jy = 0; jy < I; ++jy)
#pragma unroll
      for (int jx = 0; jx < I; ++jx) {
        bv[jy * I + jx] = e.U[jx] * e.V[jy];
        if constexpr (MAXW) bd[jy * I + jx] = cc == 0 ? e.DU[jx] * e.V[jy] : e.U[jx] * e.DV[jy];
      }
#pragma unroll
    for (int iy = 0; iy < I; ++iy)
#pragma unroll
      for (int ix = 0; ix < I; ++ix) {
        av[iy * I + ix] = e.X[ix] * e.Y[iy];
        if constexpr (MAXW) ad[iy * I + ix] = rc == 0 ? e.DX[ix] * e.Y[iy] : e.X[ix] * e.DY[iy];
      }
    const double s = part ? e.vi : e.vr;
    const double q = part ? e.qi : e.qr;
#pragma unroll
    for (int ii = 0; ii < NR; ++ii) {
      // row ii of this lane's half (rows past the end hold zero)
      constexpr int dummy = 0;
      (void)dummy;
      double ar = av[ii], dr = MAXW ? ad[ii < (MAXW ? I2 : 1) ? ii : 0] : 0.0;
      if constexpr (RH == 2) {
        const bool in = NR + ii < I2;
        const double ah = in ? av[in ? NR + ii : 0] : 0.0;
        ar = half ? ah : ar;
        if constexpr (MAXW) {
          const double dh = in ? ad[in ? NR + ii : 0] : 0.0;
          dr = half ? dh : dr;
        }
      }
      const double as = s * ar;
      const double aq = MAXW ? q * dr : 0.0;
#pragma unroll
      for (int j = 0; j < I2; ++j) {
        if constexpr (MAXW)
          acc[ii * I2 + j] += as * bv[j] + aq * bd[j];
        else
          acc[ii * I2 + j] += as * bv[j];
      }
    }
  }

  // the lane's own evaluation and those of the other lanes of its group, in role order (all
  // lanes of the wave must call this together)
  __device__ __forceinline__ void add(const Pt<KIND>& a, const Pt<KIND>& b, const NFParams& prm, int rc, int cc,
                                      bool valid, int lane) {
    const Eval e = eval(a, b, prm, rc, cc, valid);
    if constexpr (G == 1) {
      accumulate(e, 0, rc, cc);
    } else {
      const int role = lane & (G - 1), base = lane & ~(G - 1);
#pragma unroll 1
      for (int r = 0; r < G; ++r) accumulate(fetch(e, base + r), role, rc, cc);
    }
  }

  // output slot of accumulator k of a lane with this role (complex blocks store all real parts,
  // then all imaginary parts); -1 if unused
  __device__ __forceinline__ static int slot(int role, int k) {
    const int part = CPLX ? (role & 1) : 0;
    const int half = CPLX ? (role >> 1) : role;
    const int row = half * NR + k / I2;
    return row < I2 ? part * I2 * I2 + row * I2 + k % I2 : -1;
  }
};

// Duffy rules reuse surface points: some points of the edge rule depend on the outer index i and
// one coordinate of the inner index j only, and the inner points of the vertex rule on the outer
// first coordinate and j only.  Those points are tabulated in LDS for a chunk of outer indices and
// the (i, j) work of the chunk is spread over the workgroup.  The chunk is the smallest number of
// outer indices that keeps at least 90 % of the lanes busy (or the best one that fits).
__host__ __device__ constexpr int pick_chunk(int items_per_unit, int nt, int max_units) {
  int best = 1;
  int best_num = 0, best_den = 1;
  for (int c = 1; c <= max_units; ++c) {
    const int items = c * items_per_unit;
    const int slots = ((items + nt - 1) / nt) * nt;
    if (10 * items >= 9 * slots) return c;
    if (items * best_den > best_num * slots) best = c, best_num = items, best_den = slots;
  }
  return best;
}
// LDS for the point tables (bytes), shared by the launch and the kernel
constexpr int DUFFY_TABLE_BYTES = 48 * 1024;
template <class PtK>
__host__ __device__ constexpr int edge_chunk(int n, int nt) {
  const int cap = DUFFY_TABLE_BYTES / (8 * n * static_cast<int>(sizeof(PtK)));
  const int mx = cap < n * n ? cap : n * n;
  return pick_chunk(n * n, nt, mx > 0 ? mx : 1);
}
template <class PtK>
__host__ __device__ constexpr int vertex_chunk(int n, int nt) {
  const int cap = DUFFY_TABLE_BYTES / (2 * n * n * static_cast<int>(sizeof(PtK)));
  const int mx = cap < n ? cap : n;
  return pick_chunk(n * n * n, nt, mx > 0 ? mx : 1);
}

// Edge rule, scalar operators: the computed points of sub-regions g = 1, 2 are tau(t4 or 1 - t4,
// e1; r) with t4 fixed by (i, j0), so for one outer point they lie on a coordinate line of the
// patch.  The patch is contracted once per line in the fixed direction (a degree-(PD-1) curve of
// homogeneous points and cross derivatives, 2 PD x 4 values held in registers) and each point on
// the line costs a one-dimensional evaluation.
template <int PD>
__device__ __forceinline__ void line_contract(const double* __restrict__ cp, bool fix_s, double f, double* L) {
  double B[PD], E[PD];
  bernstein_basis_der<PD - 1>(f, B, E);
#pragma unroll
  for (int a = 0; a < PD; ++a) {
    double r[4] = {0, 0, 0, 0}, rs[4] = {0, 0, 0, 0};
#pragma unroll
    for (int b = 0; b < PD; ++b) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const double c = fix_s ? cp[4 * (b * PD + a) + k] : cp[4 * (a * PD + b) + k];
        r[k] += B[b] * c;
        rs[k] += E[b] * c;
      }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) L[4 * a + k] = r[k], L[4 * PD + 4 * a + k] = rs[k];
  }
}

// a point of a contracted line at local coordinates (u, v) with weight w, as make_point
template <int PD, int KIND>
__device__ __forceinline__ Pt<KIND> line_point(const double* L, bool fix_s, double lx, double ly, double h, double u,
                                               double v, double w) {
  double B[PD], E[PD];
  bernstein_basis_der<PD - 1>(fix_s ? ly + h * v : lx + h * u, B, E);
  double p[4] = {0, 0, 0, 0}, d1[4] = {0, 0, 0, 0}, d2[4] = {0, 0, 0, 0};
#pragma unroll
  for (int a = 0; a < PD; ++a)
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      p[k] += B[a] * L[4 * a + k];
      d1[k] += B[a] * L[4 * PD + 4 * a + k];  // across the line (the fixed direction)
      d2[k] += E[a] * L[4 * a + k];           // along the line
    }
  const double bot = rcp_newton(p[3]);
  const double b2 = bot * bot;
  double X[3], Ds[3], Dt[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    X[k] = p[k] * bot;
    const double ds = fix_s ? d1[k] : d2[k], dt = fix_s ? d2[k] : d1[k];
    const double dsw = fix_s ? d1[3] : d2[3], dtw = fix_s ? d2[3] : d1[3];
    Ds[k] = (ds * p[3] - p[k] * dsw) * b2;
    Dt[k] = (dt * p[3] - p[k] * dtw) * b2;
  }
  Pt<KIND> q;
  q.x0 = X[0], q.x1 = X[1], q.x2 = X[2];
  q.s0 = u, q.s1 = v;
  if constexpr (KIND == BMI_DUMMY) {
    q.W = w;
  } else {
    const double c0 = Ds[1] * Dt[2] - Ds[2] * Dt[1];
    const double c1 = Ds[2] * Dt[0] - Ds[0] * Dt[2];
    const double c2 = Ds[0] * Dt[1] - Ds[1] * Dt[0];
    q.W = sqrt(c0 * c0 + c1 * c1 + c2 * c2) * w;
  }
  return q;
}

// Factorised vertex rule (scalar operators): a lane fixes one outer point and sums
// kernel x basis vector of the varying point over (its share of) the inner index, then adds the
// outer product with the fixed point's basis vector to its block: (P+1)^2 multiply-adds per
// evaluation instead of (P+1)^4.  A shape is the chunk of outer indices staged per LDS table
// and the number of lanes that split one outer point's inner range (even for complex
// operators, whose lane pairs hold the real and imaginary parts of the block).  The score
// trades busy lanes against the outer products, which every lane of an outer point repeats.
struct FactorShape {
  int chunk, lanes;
};
__host__ __device__ constexpr FactorShape factor_shape(int outer_per_chunk, int max_chunk, int nt, int g,
                                                       int eval_cost) {
  FactorShape best{1, g};
  long best_score = -1;
  for (int c = 1; c <= max_chunk; ++c) {
    const int units = c * outer_per_chunk;
    int S = nt / units;
    S -= S % g;  // whole lane groups per unit
    if (S < g) break;
    const long score = static_cast<long>(units) * S * 1000 / (eval_cost + S);
    if (score > best_score) best_score = score, best = FactorShape{c, S};
  }
  return best;
}
// vertex rule: chunks of first outer coordinates (n outer points each), inner range n^2
__host__ __device__ constexpr FactorShape vertex_shape(int n, int nt, int ptsize, int g) {
  const int cap = DUFFY_TABLE_BYTES / (2 * n * n * ptsize);
  return factor_shape(n, cap < n ? (cap > 0 ? cap : 1) : n, nt, g, 90);
}
// edge rule, factorised: chunks of outer indices i (n outer points (i, j0) each), inner range n
__host__ __device__ constexpr FactorShape edge_shape(int n, int nt, int ptsize, int g) {
  const int cap = DUFFY_TABLE_BYTES / (8 * n * ptsize);
  const int mx = cap < n * n ? cap : n * n;
  return factor_shape(n, mx > 0 ? mx : 1, nt, g, 48);
}

// kernel times both point weights, (kr, ki): the scalar integrands of eval() without the basis
template <int KIND>
__device__ __forceinline__ void kernel_ww(const Pt<KIND>& a, const Pt<KIND>& b, const NFParams& prm, double& kr,
                                          double& ki) {
  const double d0 = a.x0 - b.x0, d1 = a.x1 - b.x1, d2 = a.x2 - b.x2;
  const double s = d0 * d0 + d1 * d1 + d2 * d2;
  const double ri = rsqrt_newton(s);  // 1 / r
  if constexpr (KIND == BMI_LAPLACE_SL) {
    kr = (a.W * b.W) * (0.25 * M_1_PI) * ri, ki = 0.0;
  } else if constexpr (KIND == BMI_HELMHOLTZ_SL) {
    const double r = s * ri;
    double sn, cs;
    sincos(prm.kre * r, &sn, &cs);
    const double damp = prm.kim != 0.0 ? exp(prm.kim * r) : 1.0;
    const double mag = damp * (a.W * b.W) * (0.25 * M_1_PI) * ri;
    kr = mag * cs, ki = -mag * sn;
  } else {
    kr = dummy_value(prm.dummy_fn, a, b) * a.W * b.W, ki = 0.0;
  }
}

// lane-private sum of (kernel x basis vector) over inner points
template <int P, bool CPLX>
struct VecSum {
  static constexpr int I = P + 1, I2 = I * I;
  double r[I2], i[CPLX ? I2 : 1];
  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int k = 0; k < I2; ++k) {
      r[k] = 0.0;
      if constexpr (CPLX) i[k] = 0.0;
    }
  }
  // += (kr + i ki) * basis(s0, s1)
  __device__ __forceinline__ void add(double kr, double ki, double s0, double s1) {
    double X[I], Y[I];
    bernstein_basis<P>(s0, X);
    bernstein_basis<P>(s1, Y);
#pragma unroll
    for (int iy = 0; iy < I; ++iy)
#pragma unroll
      for (int ix = 0; ix < I; ++ix) {
        const double b = X[ix] * Y[iy];
        r[iy * I + ix] += kr * b;
        if constexpr (CPLX) i[iy * I + ix] += ki * b;
      }
  }
  // the vector summed over the G lanes of this lane's group (each holds a share of the inner
  // range): real parts, or for complex vectors the part of this lane's role (role & 1: 0 real,
  // 1 imaginary); all lanes must call this together
  template <int G>
  __device__ __forceinline__ void total(int role, double* w) const {
#pragma unroll
    for (int k = 0; k < I2; ++k) {
      double sr = r[k], si = CPLX ? i[k] : 0.0;
#pragma unroll
      for (int m = 1; m < G; m <<= 1) {
        sr += __shfl_xor(sr, m, 64);
        if constexpr (CPLX) si += __shfl_xor(si, m, 64);
      }
      w[k] = (CPLX && (role & 1)) ? si : sr;
    }
  }
};

// One factorised sub-region: the fixed point F against the inner points inner(j) for j = sub,
// sub + S, ... < count.  COLS: F supplies the rows (element 1) and the inner points the columns;
// otherwise the inner points supply the rows.  Adds the outer product of F's basis with the
// summed vector to this lane's part of the block.  All lanes of the wave must call this
// together (the lane pairs of complex operators exchange their sums).
template <int P, int KIND, bool COLS, class Acc, class PtK, class Inner>
__device__ __forceinline__ void factor_pass(Acc& acc, const PtK& F, int role, bool act, int sub, int S, int count,
                                            const NFParams& prm, Inner inner) {
  constexpr int I = P + 1, I2 = I * I;
  VecSum<P, Acc::CPLX> u;
  u.zero();
  if (act) {
#pragma unroll 1
    for (int j = sub; j < count; j += S) {
      const PtK b = inner(j);
      double kr, ki;
      if constexpr (COLS)
        kernel_ww<KIND>(F, b, prm, kr, ki);
      else
        kernel_ww<KIND>(b, F, prm, kr, ki);
      u.add(kr, ki, b.s0, b.s1);
    }
  }
  double w[I2];
  u.template total<Acc::G>(role, w);
  if (!act) return;
  double X[I], Y[I];
  bernstein_basis<P>(F.s0, X);
  bernstein_basis<P>(F.s1, Y);
  double f[I2];
#pragma unroll
  for (int iy = 0; iy < I; ++iy)
#pragma unroll
    for (int ix = 0; ix < I; ++ix) f[iy * I + ix] = X[ix] * Y[iy];
  // this lane's rows: half (role bit above the real/imaginary bit) x NR .. + NR
  const int half = Acc::RH == 2 ? (Acc::CPLX ? (role >> 1) : role) : 0;
#pragma unroll
  for (int ii = 0; ii < Acc::NR; ++ii) {
    const int row = half * Acc::NR + ii;
    const bool in = Acc::RH == 1 || row < I2;
    const int rr = in ? row : 0;
    // COLS: F's basis gives the rows; otherwise the summed vector does
    const double a = in ? (COLS ? f[rr] : w[rr]) : 0.0;
#pragma unroll
    for (int k = 0; k < I2; ++k) acc.acc[ii * I2 + k] += a * (COLS ? w[k] : f[k]);
  }
}

// f(std::integral_constant<int, r>()) for a rotation r that is uniform in the workgroup: the
// callee's tau and line evaluations then fold the rotation at compile time
template <class F>
__device__ __forceinline__ void with_rotation(int r, F&& f) {
  switch (r) {
    case 1: f(std::integral_constant<int, 1>()); break;
    case 2: f(std::integral_constant<int, 2>()); break;
    case 3: f(std::integral_constant<int, 3>()); break;
    default: f(std::integral_constant<int, 0>()); break;
  }
}

// DuffyTrick::tau (DuffyTrick/tau.hpp:26-42)
__device__ __forceinline__ void tau(double x, double y, int c, double& u, double& v) {
  switch (c) {
    case 1: u = 1 - y, v = x; return;
    case 2: u = 1 - x, v = 1 - y; return;
    case 3: u = y, v = 1 - x; return;
    default: u = x, v = y; return;
  }
}

// sum the per-lane partial blocks over the workgroup: lanes of equal role within each wave by
// shuffles, then the waves through LDS; the result is the full block in `sum` (Acc::TOT values)
template <int NT, class Acc>
__device__ __forceinline__ void reduce_block(Acc& acc, double* lds, int tid, double* sum) {
  constexpr int N = Acc::NR * Acc::I2;
  constexpr int TOT = Acc::TOT;
  constexpr int NW = NT / 64;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    double v = acc.acc[k];
#pragma unroll
    for (int off = 32; off >= Acc::G; off >>= 1) v += __shfl_xor(v, off, 64);
    acc.acc[k] = v;
  }
  const int wave = tid >> 6, lane = tid & 63;
  if (lane < Acc::G) {
#pragma unroll
    for (int k = 0; k < N; ++k) {
      const int sl = Acc::slot(lane, k);
      if (sl >= 0) lds[wave * TOT + sl] = acc.acc[k];
    }
  }
  __syncthreads();
  for (int k = tid; k < TOT; k += NT) {
    double v = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) v += lds[w * TOT + k];
    sum[k] = v;
  }
  __syncthreads();
}

template <int P, int KIND, int PD, int CLS, int NT>
__global__ __launch_bounds__(NT) void k_nearfield(const NFPair* __restrict__ pairs, NFParams prm) {
  extern __shared__ double lds[];
  using Acc = BlockAcc<P, KIND>;
  constexpr int I2 = Acc::I2;
  const NFPair pr = pairs[blockIdx.x];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const double* cp1 = prm.cp + static_cast<int64_t>(prm.elem_patch[pr.e1]) * prm.cp_stride;
  const double* cp2 = prm.cp + static_cast<int64_t>(prm.elem_patch[pr.e2]) * prm.cp_stride;
  const double l1x = prm.elem_llc[2 * pr.e1], l1y = prm.elem_llc[2 * pr.e1 + 1];
  const double l2x = prm.elem_llc[2 * pr.e2], l2y = prm.elem_llc[2 * pr.e2 + 1];
  const double h = prm.h;
  const int n = pr.deg + 1;
  const int nq = n * n;
  const double* gx = c_gl_nodes + c_gl_offset[n - 1];
  const double* gw = c_gl_weights + c_gl_offset[n - 1];
  using PtK = Pt<KIND>;
  // Maxwell component block of this workgroup (0 for the scalar operators)
  const int rc = KIND == BMI_MAXWELL_SL ? static_cast<int>(blockIdx.y) >> 1 : 0;
  const int cc = KIND == BMI_MAXWELL_SL ? static_cast<int>(blockIdx.y) & 1 : 0;
  Acc acc;
  acc.zero();
  double aval = 0.0;  // the block entry a thread owns (factorised identical rule)
  // every lane runs the same number of iterations (lane groups swap values); lanes past the end
  // evaluate a valid point with zero weight

  if constexpr (CLS == 0) {
    // integrate0: far-field tables of both elements
    const int nf = prm.nff, fs = prm.ffs;
    const double* f1 = prm.ffield + static_cast<int64_t>(pr.e1) * nf * fs;
    const double* f2 = prm.ffield + static_cast<int64_t>(pr.e2) * nf * fs;
    const int total = nf * nf;
    for (int base = 0; base < total; base += NT) {
      const bool valid = base + tid < total;
      const int idx = valid ? base + tid : total - 1;
      const int i = idx / nf, j = idx - (idx / nf) * nf;
      const double* a = f1 + i * fs;
      const double* b = f2 + j * fs;
      constexpr bool PLAIN_W = KIND == BMI_DUMMY || KIND == BMI_MAXWELL_SL;
      PtK pa, pb;
      pa.x0 = a[0], pa.x1 = a[1], pa.x2 = a[2], pa.W = PLAIN_W ? a[4] : a[3] * a[4], pa.s0 = a[5], pa.s1 = a[6];
      pb.x0 = b[0], pb.x1 = b[1], pb.x2 = b[2], pb.W = PLAIN_W ? b[4] : b[3] * b[4], pb.s0 = b[5], pb.s1 = b[6];
      if constexpr (KIND == BMI_MAXWELL_SL) {
#pragma unroll
        for (int k = 0; k < 6; ++k) pa.t[k] = a[8 + k], pb.t[k] = b[8 + k];
      }
      acc.add(pa, pb, prm, rc, cc, valid, lane);
    }
  } else if constexpr (CLS == 1 && !Acc::MAXW) {
    // integrate1, factorised: surface points of both elements staged in LDS; S lanes share a
    // point of element 1 and each sums kernel x basis vector over its share of the points of
    // element 2.  The lanes' vectors and the outer points' basis values go to LDS, and every
    // thread adds the outer products into the one block entry it owns (no per-lane block, no
    // workgroup reduction).
    constexpr bool CPLX = Acc::CPLX;
    constexpr int I = P + 1;
    constexpr int VR = CPLX ? 2 * I2 : I2;  // doubles of one lane's vector
    constexpr int NE = Acc::TOT;
    static_assert(NE <= NT, "at most one block entry per thread");
    PtK* q1 = reinterpret_cast<PtK*>(lds);
    PtK* q2 = q1 + nq;
    double* rec = reinterpret_cast<double*>(q2 + nq);  // NT vectors
    double* bas = rec + NT * VR;                       // basis products of the round's outer points
    for (int k = tid; k < nq; k += NT) {
      const double u = gx[k / n], v = gx[k % n], w = h * (gw[k / n] * gw[k % n]);
      q1[k] = make_point<PD, KIND>(cp1, l1x, l1y, h, u, v, w);
      q2[k] = make_point<PD, KIND>(cp2, l2x, l2y, h, u, v, w);
    }
    __syncthreads();
    const int S = NT / nq > 1 ? NT / nq : 1;
    const int per_round = NT / S;
    const int epart = tid / (I2 * I2), erem = tid - epart * (I2 * I2);
    const int erow = erem / I2, ecol = erem - erow * I2;
    for (int u0 = 0; u0 < nq; u0 += per_round) {
      const int unit = u0 + tid / S, sub = tid - (tid / S) * S;
      const bool act = tid / S < per_round && unit < nq;
      VecSum<P, CPLX> vs;
      vs.zero();
      if (act) {
        const PtK F = q1[unit];
#pragma unroll 1
        for (int j = sub; j < nq; j += S) {
          const PtK b = q2[j];
          double kr, ki;
          kernel_ww<KIND>(F, b, prm, kr, ki);
          vs.add(kr, ki, b.s0, b.s1);
        }
        if (sub == 0) {
          double X[I], Y[I];
          bernstein_basis<P>(F.s0, X);
          bernstein_basis<P>(F.s1, Y);
#pragma unroll
          for (int iy = 0; iy < I; ++iy)
#pragma unroll
            for (int ix = 0; ix < I; ++ix) bas[(tid / S) * I2 + iy * I + ix] = X[ix] * Y[iy];
        }
      }
      double* r = rec + tid * VR;
#pragma unroll
      for (int k = 0; k < I2; ++k) {
        r[k] = vs.r[k];
        if constexpr (CPLX) r[I2 + k] = vs.i[k];
      }
      __syncthreads();
      if (tid < NE) {
        const int nu = nq - u0 < per_round ? nq - u0 : per_round;
        const int o = epart * I2 + ecol;
#pragma unroll 1
        for (int m = 0; m < nu; ++m) {
          double v = 0.0;
          for (int s2 = 0; s2 < S; ++s2) v += rec[(m * S + s2) * VR + o];
          aval += bas[m * I2 + erow] * v;
        }
      }
      __syncthreads();
    }
  } else if constexpr (CLS == 1) {
    // integrate1: tensor Gauss rule; surface points of both elements staged in LDS
    PtK* q1 = reinterpret_cast<PtK*>(lds);
    PtK* q2 = q1 + nq;
    for (int k = tid; k < nq; k += NT) {
      const double u = gx[k / n], v = gx[k % n], w = h * (gw[k / n] * gw[k % n]);
      q1[k] = make_point<PD, KIND>(cp1, l1x, l1y, h, u, v, w);
      q2[k] = make_point<PD, KIND>(cp2, l2x, l2y, h, u, v, w);
    }
    __syncthreads();
    const int total = nq * nq;
    for (int base = 0; base < total; base += NT) {
      const bool valid = base + tid < total;
      const int idx = valid ? base + tid : total - 1;
      const int i = idx / nq, j = idx - (idx / nq) * nq;
      acc.add(q1[i], q2[j], prm, rc, cc, valid, lane);
    }
    __syncthreads();
  } else if constexpr (CLS == 2 && (KIND == BMI_LAPLACE_SL || KIND == BMI_HELMHOLTZ_SL)) {
    // integrate2, factorised (symmetric kernels).  The sub-simplex pairs of the unfactorised form
    // below are, with t1 = xj0 (1 - xi0), t3 = t1 + xi0 (outer unit (i, j0)) and t2 = xj1 (1 - xi0 xi1),
    // t4 = t2 + xi0 xi1 (inner index j1): (t1,t2)-(t3,t4), (t1,t4)-(t3,t2), (t2,t1)-(t4,t3),
    // (t2,t3)-(t4,t1).  For one unit the eight points lie on four coordinate lines of the patch
    // (s = t1, s = t3, t = t1, t = t3), contracted once per round into LDS (one line per thread),
    // and each point is a one-dimensional evaluation.  The tensor basis splits the block: with
    // B(x) the Bernstein values, A((ix,iy),(jx,jy)) += B_ix(t1) B_jx(t3) U01(iy,jy) + B_iy(t1)
    // B_jy(t3) U23(ix,jx), U01 = sum over j1 of K0 B(t2) B(t4)^T + K1 B(t4) B(t2)^T and U23 likewise
    // with K2, K3 (the fourth pair taken in the other order, which the symmetric kernel allows).
    // The rule adds every pair in both orders, so the block is A + A^T, formed when it is written.
    // S = 4 lanes share a unit (a lane pair of an even wave for the s-lines, one of an odd wave for
    // the t-lines, each pair taking every second j1) and add their sums; the unit's sums, basis
    // values and outer weight go to LDS, and after each round every thread adds the units'
    // contributions to the one entry of A it owns (no per-lane block in registers).
    constexpr bool CPLX = Acc::CPLX;
    constexpr int I = P + 1;
    constexpr int S = 4;
    constexpr int UPR = NT / S;                // units per round
    constexpr int LS = 8 * PD;                 // doubles of one contracted line
    constexpr int US = 4 * LS + 2;             // doubles per unit: the 16-byte pad puts the 16 units
                                               // a wave reads at once on different LDS banks
    constexpr int UR = 7 + (CPLX ? 4 : 2) * I2;  // record of a unit: sc, B(t1), B(t3), U01, U23
    constexpr int NE = Acc::TOT;               // entries of A (real parts, then imaginary parts)
    static_assert(UPR * 4 == NT && NE <= NT, "one line and at most one block entry per thread");
    static_assert(NT == 256, "four waves: two lane-pair halves of 32 units each");
    double* ua = lds + UPR * US;
    const int units = nq * n;
    // evaluation lanes: even waves sum U01 over the s-lines, odd waves U23 over the t-lines (the
    // line direction is uniform in a wave and fixed at compile time below); a lane pair holds a
    // unit and takes every second j1
    const int wv = tid >> 6, ln = tid & 63;
    const bool tl = (wv & 1) != 0;
    const int myu = (wv >> 1) * 32 + (ln >> 1), par = ln & 1;
    // the block entry this thread owns
    const int epart = tid / (I2 * I2), erem = tid - epart * (I2 * I2);
    const int erow = erem / I2, ecol = erem - erow * I2;
    const int eiy = erow / I, eix = erow - eiy * I, ejy = ecol / I, ejx = ecol - ejy * I;
    for (int u0 = 0; u0 < units; u0 += UPR) {
      __syncthreads();
      {
        const int uu = u0 + (tid >> 2), q = tid & 3;
        if (uu < units) {
          const int i = uu / n, j0 = uu - (uu / n) * n;
          const double xi0 = gx[i / n];
          const double t1 = gx[j0] * (1 - xi0), tf = (q & 1) ? t1 + xi0 : t1;
          const bool fs = q < 2;
          line_contract<PD>(cp1, fs, fs ? l1x + h * tf : l1y + h * tf, lds + (tid >> 2) * US + q * LS);
        }
      }
      __syncthreads();
      const int uu = u0 + myu;
      const bool act = uu < units;
      const int uc = act ? uu : 0;
      const int i = uc / n, j0 = uc - (uc / n) * n;
      const double xi0 = gx[i / n], xi1 = gx[i % n], wi = gw[i / n] * gw[i % n];
      const double t1 = gx[j0] * (1 - xi0), t3 = t1 + xi0;
      double ur[I2], ui[CPLX ? I2 : 1];
#pragma unroll
      for (int k = 0; k < I2; ++k) {
        ur[k] = 0.0;
        if constexpr (CPLX) ui[k] = 0.0;
      }
      auto sums = [&](auto TLC) {
        constexpr bool TL = decltype(TLC)::value;
#pragma unroll 1
        for (int j1 = par; j1 < n; j1 += 2) {
          const double t2 = gx[j1] * (1 - xi0 * xi1), t4 = t2 + xi0 * xi1;
          const double wj = gw[j1];
          // the lines are read again in every iteration, one at a time (an opaque offset keeps
          // the compiler from holding them in registers across the loop)
          int off = myu * US + (TL ? 2 * LS : 0);
          asm volatile("" : "+v"(off));
          const double* Lq = lds + off;
          // s-lines: (t1,t2)-(t3,t4) and (t1,t4)-(t3,t2); t-lines: (t2,t1)-(t4,t3) and
          // (t4,t1)-(t2,t3) (the second pair in the other order); either way the first line
          // holds the points at (t2, t4) of the first member, the second line those of the second
          const PtK a0 = line_point<PD, KIND>(Lq, !TL, l1x, l1y, h, TL ? t2 : t1, TL ? t1 : t2, 1.0);
          const PtK a1 = line_point<PD, KIND>(Lq, !TL, l1x, l1y, h, TL ? t4 : t1, TL ? t1 : t4, TL ? wj : 1.0);
          __builtin_amdgcn_sched_barrier(0);
          const PtK b0 = line_point<PD, KIND>(Lq + LS, !TL, l1x, l1y, h, TL ? t4 : t3, TL ? t3 : t4, wj);
          const PtK b1 = line_point<PD, KIND>(Lq + LS, !TL, l1x, l1y, h, TL ? t2 : t3, TL ? t3 : t2, TL ? 1.0 : wj);
          double kar, kai, kbr, kbi;
          kernel_ww<KIND>(a0, b0, prm, kar, kai);
          kernel_ww<KIND>(a1, b1, prm, kbr, kbi);
          double B2[I], B4[I];
          bernstein_basis<P>(t2, B2);
          bernstein_basis<P>(t4, B4);
#pragma unroll
          for (int p = 0; p < I; ++p)
#pragma unroll
            for (int q = 0; q < I; ++q) {
              const double f = B2[p] * B4[q], g = B4[p] * B2[q];
              ur[p * I + q] += kar * f + kbr * g;
              if constexpr (CPLX) ui[p * I + q] += kai * f + kbi * g;
            }
        }
      };
      if (act) {
        if (tl)
          sums(std::integral_constant<bool, true>());
        else
          sums(std::integral_constant<bool, false>());
      }
      // the two lanes of a unit add their sums; the first lane of the s-line pair writes the
      // unit's record with U01, the first lane of the t-line pair adds U23
#pragma unroll
      for (int k = 0; k < I2; ++k) {
        ur[k] += __shfl_xor(ur[k], 1, 64);
        if constexpr (CPLX) ui[k] += __shfl_xor(ui[k], 1, 64);
      }
      if (par == 0) {
        double* r = ua + myu * UR;
        if (!tl) {
          double B1[I], B3[I];
          bernstein_basis<P>(t1, B1);
          bernstein_basis<P>(t3, B3);
          r[0] = act ? h * h * wi * xi0 * (1 - xi0) * (1 - xi0 * xi1) * gw[j0] : 0.0;
#pragma unroll
          for (int k = 0; k < I; ++k) r[1 + k] = B1[k], r[1 + I + k] = B3[k];
        }
        const int o = tl ? I2 : 0;
#pragma unroll
        for (int k = 0; k < I2; ++k) {
          r[7 + o + k] = ur[k];
          if constexpr (CPLX) r[7 + 2 * I2 + o + k] = ui[k];
        }
      }
      __syncthreads();
      if (tid < NE) {
        const int o01 = 7 + (epart ? 2 * I2 : 0), o23 = o01 + I2;
#pragma unroll 4
        for (int m = 0; m < UPR; ++m) {
          const double* r = ua + m * UR;
          aval += r[0] * (r[1 + eix] * r[1 + I + ejx] * r[o01 + eiy * I + ejy] +
                          r[1 + eiy] * r[1 + I + ejy] * r[o23 + eix * I + ejx]);
        }
      }
    }
    __syncthreads();
  } else if constexpr (CLS == 2) {
    // integrate2: identical elements, 8 Duffy sub-simplices, symmetrised
    const int total = nq * nq;
    for (int base = 0; base < total; base += NT) {
      const bool valid = base + tid < total;
      const int idx = valid ? base + tid : total - 1;
      const int i = idx / nq, j = idx - (idx / nq) * nq;
      const double xi0 = gx[i / n], xi1 = gx[i % n], wi = gw[i / n] * gw[i % n];
      const double xj0 = gx[j / n], xj1 = gx[j % n], wj = gw[j / n] * gw[j % n];
      const double w = h * h * wi * xi0 * (1 - xi0) * (1 - xi0 * xi1);
      const double t1 = xj0 * (1 - xi0);
      const double t2 = xj1 * (1 - xi0 * xi1);
      const double t3 = t1 + xi0;
      const double t4 = t2 + xi0 * xi1;
      // point pairs (t1,t2)-(t3,t4), (t1,t4)-(t3,t2), (t2,t1)-(t4,t3), (t2,t3)-(t4,t1), written as
      // selects; the loop is unrolled only when the block accumulator leaves registers to spare
      constexpr int UNR2 = Acc::G == 1 ? 4 : 1;
#pragma unroll UNR2
      for (int k = 0; k < 4; ++k) {
        const bool sw = k >= 2, odd = k & 1;
        const double pu = sw ? t2 : t1, qu = sw ? t4 : t3;
        const double r = sw ? t1 : t2, sv = sw ? t3 : t4;
        const PtK a = make_point<PD, KIND>(cp1, l1x, l1y, h, pu, odd ? sv : r, w);
        const PtK b = make_point<PD, KIND>(cp1, l1x, l1y, h, qu, odd ? r : sv, wj);
        acc.add(a, b, prm, rc, cc, valid, lane);
        acc.add(b, a, prm, rc, cc, valid, lane);
      }
    }
  } else if constexpr (CLS == 3 && !Acc::MAXW) {
    // integrate3, factorised: the sub-regions and points of the unfactorised form below.  Outer
    // point (i, j0), inner index j1.  g = 0: rows (t1, e0) and columns (t2, e1) from the tables;
    // g = 1: rows (t3, xi1) from the tables, columns on element 2 at tau(t4 or 1 - t4, e1; r2);
    // g = 2: rows on element 1 at tau(t4 or 1 - t4, e1; r1), columns (t3, xi1) from the tables.
    // The computed points come from the contracted line of their (i, j0).  The weight products
    // are split into the outer factor (w and the j0 part of the inner weights) and gw[j1].
    const int r1 = pr.rot1, r2 = pr.rot2;
    constexpr bool CPLX = Acc::CPLX;
    const FactorShape shp = edge_shape(n, NT, static_cast<int>(sizeof(PtK)), Acc::G);
    const int CI = shp.chunk, S = shp.lanes;
    const int role = lane & (Acc::G - 1);
    // tau(x, y; r) fixes u for r = 0, 2: the line of a computed point runs along v there
    PtK* tab = reinterpret_cast<PtK*>(lds);
    for (int i0 = 0; i0 < nq; i0 += CI) {
      const int cn = nq - i0 < CI ? nq - i0 : CI;
      const int per = cn * n;
      __syncthreads();
      for (int e = tid; e < 8 * per; e += NT) {
        const int kind = e / per, rem = e - kind * per;
        const int ci = rem / n, jx = rem - ci * n;
        const int i = i0 + ci;
        const double xi0 = gx[i / n], xi1 = gx[i % n], xj = gx[jx];
        const bool odd = kind & 1;
        double u, v;
        if (kind < 2) {
          const double t1 = xi0 * (1 - xi1);
          tau(odd ? 1 - t1 : t1, xi1 * xj, r1, u, v);
          tab[e] = make_point<PD, KIND>(cp1, l1x, l1y, h, u, v, 1.0);
        } else if (kind < 4) {
          const double t3 = xi0 * (1 - xi1 * xj);
          tau(odd ? 1 - t3 : t3, xi1, r1, u, v);
          tab[e] = make_point<PD, KIND>(cp1, l1x, l1y, h, u, v, 1.0);
        } else if (kind < 6) {
          const double t3 = xi0 * (1 - xi1 * xj);
          tau(odd ? 1 - t3 : t3, xi1, r2, u, v);
          tab[e] = make_point<PD, KIND>(cp2, l2x, l2y, h, u, v, 1.0);
        } else {
          const double t2 = (1 - xi0) * (1 - xi1);
          tau(odd ? 1 - t2 : t2, xi1 * xj, r2, u, v);
          tab[e] = make_point<PD, KIND>(cp2, l2x, l2y, h, u, v, 1.0);
        }
      }
      __syncthreads();
      const int units = cn * n;
      const int unit = tid / S, sub = tid - (tid / S) * S;
      const bool act = unit < units;
      const int uu = act ? unit : 0;
      const int ci = uu / n, j0 = uu - (uu / n) * n;
      const int i = i0 + ci;
      const double xi0 = gx[i / n], xi1 = gx[i % n], wi = gw[i / n] * gw[i % n];
      const double w = h * h * wi * xi1 * xi1;
      const double e0 = xi1 * gx[j0];
      const double t4 = (1 - xi0) * (1 - e0);
      const double fa = w * (gw[j0] * (1 - xi1)), fb = w * (gw[j0] * (1 - e0));
      const int ta = ci * n + j0;
      const PtK* colt = tab + ci * n;
#pragma unroll 1
      for (int odd = 0; odd < 2; ++odd) {
        const double x = odd ? 1 - t4 : t4;
        // g = 0: table rows and columns
        PtK F = tab[odd * per + ta];
        F.W *= fa;
        factor_pass<P, KIND, true>(acc, F, role, act, sub, S, n, prm, [&](int j1) {
          PtK b = colt[(6 + odd) * per + j1];
          b.W *= gw[j1];
          return b;
        });
        // g = 1: table rows, columns on the line of element 2 (the rotation, uniform in the
        // workgroup, fixed at compile time: tau and the line evaluation without selects)
        F = tab[(2 + odd) * per + ta];
        F.W *= fb;
        with_rotation(r2, [&](auto R) {
          constexpr int RC = decltype(R)::value;
          constexpr bool FS = (RC & 1) == 0;
          double L[8 * PD];
          line_contract<PD>(cp2, FS, FS ? l2x + h * (RC < 2 ? x : 1 - x) : l2y + h * (RC < 2 ? x : 1 - x), L);
          factor_pass<P, KIND, true>(acc, F, role, act, sub, S, n, prm, [&](int j1) {
            double u, v;
            tau(x, xi1 * gx[j1], RC, u, v);
            return line_point<PD, KIND>(L, FS, l2x, l2y, h, u, v, gw[j1]);
          });
        });
        // g = 2: rows on the line of element 1, table columns
        F = tab[(4 + odd) * per + ta];
        F.W *= fb;
        with_rotation(r1, [&](auto R) {
          constexpr int RC = decltype(R)::value;
          constexpr bool FS = (RC & 1) == 0;
          double L[8 * PD];
          line_contract<PD>(cp1, FS, FS ? l1x + h * (RC < 2 ? x : 1 - x) : l1y + h * (RC < 2 ? x : 1 - x), L);
          factor_pass<P, KIND, false>(acc, F, role, act, sub, S, n, prm, [&](int j1) {
            double u, v;
            tau(x, xi1 * gx[j1], RC, u, v);
            return line_point<PD, KIND>(L, FS, l1x, l1y, h, u, v, gw[j1]);
          });
        });
      }
    }
    __syncthreads();
  } else if constexpr (CLS == 4 && !Acc::MAXW) {
    // integrate4, factorised: the four outer points per i as below; per chunk of first outer
    // coordinates the inner points p5 (element 2) and p6 (element 1) are tabulated, and a lane
    // fixed to one outer index i sums over its share of j: p5 gives columns for the rows of
    // q0, q1, p6 gives rows for the columns of q2, q3.
    const int r1 = pr.rot1, r2 = pr.rot2;
    constexpr bool CPLX = Acc::CPLX;
    PtK* q = reinterpret_cast<PtK*>(lds);  // 4 * nq
    PtK* tin = q + 4 * nq;
    for (int i = tid; i < nq; i += NT) {
      const double xi0 = gx[i / n], wi = gw[i / n] * gw[i % n];
      const double xi1 = gx[i % n] * xi0;
      const double w = h * h * wi * (xi0 * xi0 * xi0);
      double u, v;
      tau(xi0, xi1, r1, u, v);
      q[4 * i + 0] = make_point<PD, KIND>(cp1, l1x, l1y, h, u, v, w);
      tau(xi1, xi0, r1, u, v);
      q[4 * i + 1] = make_point<PD, KIND>(cp1, l1x, l1y, h, u, v, w);
      tau(xi0, xi1, r2, u, v);
      q[4 * i + 2] = make_point<PD, KIND>(cp2, l2x, l2y, h, u, v, w);
      tau(xi1, xi0, r2, u, v);
      q[4 * i + 3] = make_point<PD, KIND>(cp2, l2x, l2y, h, u, v, w);
    }
    const FactorShape shp = vertex_shape(n, NT, static_cast<int>(sizeof(PtK)), Acc::G);
    const int C0 = shp.chunk, S = shp.lanes;
    const int role = lane & (Acc::G - 1);
    for (int a0 = 0; a0 < n; a0 += C0) {
      const int cn = n - a0 < C0 ? n - a0 : C0;
      __syncthreads();
      for (int e = tid; e < 2 * cn * nq; e += NT) {
        const int which = e / (cn * nq), rem = e - which * (cn * nq);
        const int c0 = rem / nq, j = rem - c0 * nq;
        const double xi0 = gx[a0 + c0];
        const double e0 = xi0 * gx[j / n], e1 = xi0 * gx[j % n];
        const double wj = gw[j / n] * gw[j % n];
        double u, v;
        if (which == 0) {
          tau(e0, e1, r2, u, v);
          tin[e] = make_point<PD, KIND>(cp2, l2x, l2y, h, u, v, wj);
        } else {
          tau(e0, e1, r1, u, v);
          tin[e] = make_point<PD, KIND>(cp1, l1x, l1y, h, u, v, wj);
        }
      }
      __syncthreads();
      const int units = cn * n;
      const int unit = tid / S, sub = tid - (tid / S) * S;
      const bool act = unit < units;
      const int uu = act ? unit : 0;
      const int c0 = uu / n, i1 = uu - (uu / n) * n;
      const int i = (a0 + c0) * n + i1;
      const PtK* p5 = tin + c0 * nq;
      const PtK* p6 = tin + (cn + c0) * nq;
#pragma unroll 1
      for (int k = 0; k < 2; ++k)  // rows q0, q1 (element 1), columns p5 (element 2)
        factor_pass<P, KIND, true>(acc, q[4 * i + k], role, act, sub, S, nq, prm, [&](int j) { return p5[j]; });
#pragma unroll 1
      for (int k = 2; k < 4; ++k)  // rows p6 (element 1), columns q2, q3 (element 2)
        factor_pass<P, KIND, false>(acc, q[4 * i + k], role, act, sub, S, nq, prm, [&](int j) { return p6[j]; });
    }
    __syncthreads();
  } else if constexpr (CLS == 3) {
    // integrate3: common edge, 6 Duffy sub-regions.  With t1 = xi0 (1 - xi1), t2 = (1 - xi0)(1 - xi1),
    // e0 = xi1 xj0, e1 = xi1 xj1, t3 = xi0 (1 - e0), t4 = (1 - xi0)(1 - e0), sub-region k = 2g + odd
    // pairs a = tau(a0 or 1 - a0, a1; r1) on element 1 (weight w) with b = tau(b0 or 1 - b0, b1; r2)
    // on element 2 (weight wa for g = 0, else wb), where (a0, a1) = (t1, e0), (t3, xi1), (t4, e1)
    // and (b0, b1) = (t2, e1), (t4, e1), (t3, xi1).  The points of g = 0 (a: i, xj0; b: i, xj1), the
    // a points of g = 1 and the b points of g = 2 (i, xj0) come from per-chunk tables (W holds the
    // surface measure; the weight is applied where the point is used, as make_point does).
    const int r1 = pr.rot1, r2 = pr.rot2;
    const int CI = edge_chunk<PtK>(n, NT);
    PtK* tab = reinterpret_cast<PtK*>(lds);
    for (int i0 = 0; i0 < nq; i0 += CI) {
      const int cn = nq - i0 < CI ? nq - i0 : CI;
      const int per = cn * n;  // entries of one table kind
      __syncthreads();
      for (int e = tid; e < 8 * per; e += NT) {
        const int kind = e / per, rem = e - kind * per;
        const int ci = rem / n, jx = rem - ci * n;
        const int i = i0 + ci;
        const double xi0 = gx[i / n], xi1 = gx[i % n], xj = gx[jx];
        const bool odd = kind & 1;
        double u, v;
        if (kind < 2) {  // g = 0, a: (t1, e0) on element 1
          const double t1 = xi0 * (1 - xi1);
          tau(odd ? 1 - t1 : t1, xi1 * xj, r1, u, v);
          tab[e] = make_point<PD, KIND>(cp1, l1x, l1y, h, u, v, 1.0);
        } else if (kind < 4) {  // g = 1, a: (t3, xi1) on element 1
          const double t3 = xi0 * (1 - xi1 * xj);
          tau(odd ? 1 - t3 : t3, xi1, r1, u, v);
          tab[e] = make_point<PD, KIND>(cp1, l1x, l1y, h, u, v, 1.0);
        } else if (kind < 6) {  // g = 2, b: (t3, xi1) on element 2
          const double t3 = xi0 * (1 - xi1 * xj);
          tau(odd ? 1 - t3 : t3, xi1, r2, u, v);
          tab[e] = make_point<PD, KIND>(cp2, l2x, l2y, h, u, v, 1.0);
        } else {  // g = 0, b: (t2, e1) on element 2 (xj = xj1)
          const double t2 = (1 - xi0) * (1 - xi1);
          tau(odd ? 1 - t2 : t2, xi1 * xj, r2, u, v);
          tab[e] = make_point<PD, KIND>(cp2, l2x, l2y, h, u, v, 1.0);
        }
      }
      __syncthreads();
      const int total = cn * nq;
      for (int base = 0; base < total; base += NT) {
        const bool valid = base + tid < total;
        const int idx = valid ? base + tid : total - 1;
        const int ci = idx / nq, j = idx - (idx / nq) * nq;
        const int i = i0 + ci;
        const int j0 = j / n, j1 = j - (j / n) * n;
        const double xi0 = gx[i / n], xi1 = gx[i % n], wi = gw[i / n] * gw[i % n];
        const double xj0 = gx[j0], xj1 = gx[j1], wj = gw[j0] * gw[j1];
        const double w = h * h * wi * xi1 * xi1;
        const double e0 = xi1 * xj0, e1 = xi1 * xj1;
        const double t4 = (1 - xi0) * (1 - e0);
        const double wa = wj * (1 - xi1), wb = wj * (1 - e0);
        const int ta = ci * n + j0, tb = ci * n + j1;
#pragma unroll 1
        for (int k = 0; k < 6; ++k) {
          const int g = k >> 1;
          const bool odd = k & 1;
          PtK a, b;
          if (g == 0) {
            a = tab[(0 + odd) * per + ta];
            b = tab[(6 + odd) * per + tb];
          } else if (g == 1) {
            double u, v;
            a = tab[(2 + odd) * per + ta];
            tau(odd ? 1 - t4 : t4, e1, r2, u, v);
            b = make_point<PD, KIND>(cp2, l2x, l2y, h, u, v, 1.0);
          } else {
            double u, v;
            tau(odd ? 1 - t4 : t4, e1, r1, u, v);
            a = make_point<PD, KIND>(cp1, l1x, l1y, h, u, v, 1.0);
            b = tab[(4 + odd) * per + ta];
          }
          a.W *= w;
          b.W *= g == 0 ? wa : wb;
          acc.add(a, b, prm, rc, cc, valid, lane);
        }
      }
    }
    __syncthreads();
  } else {
    // integrate4: common vertex.  The four outer points per i are staged in LDS; the two inner
    // points p5 = tau(e0, e1; r2) on element 2 and p6 = tau(e0, e1; r1) on element 1, e = xi0 xj,
    // depend on the first outer coordinate and j only and are tabulated per chunk of xi0 values
    const int r1 = pr.rot1, r2 = pr.rot2;
    PtK* q = reinterpret_cast<PtK*>(lds);  // 4 * nq
    PtK* tin = q + 4 * nq;                 // 2 * nq per xi0 of the chunk
    for (int i = tid; i < nq; i += NT) {
      const double xi0 = gx[i / n], wi = gw[i / n] * gw[i % n];
      const double xi1 = gx[i % n] * xi0;
      const double w = h * h * wi * (xi0 * xi0 * xi0);
      double u, v;
      tau(xi0, xi1, r1, u, v);
      q[4 * i + 0] = make_point<PD, KIND>(cp1, l1x, l1y, h, u, v, w);
      tau(xi1, xi0, r1, u, v);
      q[4 * i + 1] = make_point<PD, KIND>(cp1, l1x, l1y, h, u, v, w);
      tau(xi0, xi1, r2, u, v);
      q[4 * i + 2] = make_point<PD, KIND>(cp2, l2x, l2y, h, u, v, w);
      tau(xi1, xi0, r2, u, v);
      q[4 * i + 3] = make_point<PD, KIND>(cp2, l2x, l2y, h, u, v, w);
    }
    const int C0 = vertex_chunk<PtK>(n, NT);
    for (int a0 = 0; a0 < n; a0 += C0) {
      const int cn = n - a0 < C0 ? n - a0 : C0;
      __syncthreads();
      for (int e = tid; e < 2 * cn * nq; e += NT) {
        const int which = e / (cn * nq), rem = e - which * (cn * nq);
        const int c0 = rem / nq, j = rem - c0 * nq;
        const double xi0 = gx[a0 + c0];
        const double e0 = xi0 * gx[j / n], e1 = xi0 * gx[j % n];
        double u, v;
        if (which == 0) {
          tau(e0, e1, r2, u, v);
          tin[e] = make_point<PD, KIND>(cp2, l2x, l2y, h, u, v, 1.0);
        } else {
          tau(e0, e1, r1, u, v);
          tin[e] = make_point<PD, KIND>(cp1, l1x, l1y, h, u, v, 1.0);
        }
      }
      __syncthreads();
      const int total = cn * n * nq;  // (xi0 of the chunk, xi1, j)
      for (int base = 0; base < total; base += NT) {
        const bool valid = base + tid < total;
        const int idx = valid ? base + tid : total - 1;
        const int c0 = idx / (n * nq), rem = idx - c0 * (n * nq);
        const int i1 = rem / nq, j = rem - i1 * nq;
        const int i = (a0 + c0) * n + i1;
        const double wj = gw[j / n] * gw[j % n];
        PtK p5 = tin[c0 * nq + j], p6 = tin[(cn + c0) * nq + j];
        p5.W *= wj;
        p6.W *= wj;
        acc.add(q[4 * i + 0], p5, prm, rc, cc, valid, lane);
        acc.add(q[4 * i + 1], p5, prm, rc, cc, valid, lane);
        acc.add(p6, q[4 * i + 2], prm, rc, cc, valid, lane);
        acc.add(p6, q[4 * i + 3], prm, rc, cc, valid, lane);
      }
    }
    __syncthreads();
  }

  // reduce the per-lane partial blocks and write the result
  double* red = lds;
  double* sum = lds + (NT / 64) * Acc::TOT;
  if constexpr ((CLS == 2 && (KIND == BMI_LAPLACE_SL || KIND == BMI_HELMHOLTZ_SL)) || (CLS == 1 && !Acc::MAXW)) {
    if (tid < Acc::TOT) sum[tid] = aval;  // each entry summed by its owner
    __syncthreads();
  } else {
    reduce_block<NT>(acc, red, tid, sum);
  }
  double* o = prm.out;
  constexpr bool CPLX = Acc::CPLX;
  const int64_t base = pr.off + rc * prm.comp_rs + cc * prm.comp_cs;
  // the factorised identical rule accumulated A; its block is A + A^T
  constexpr bool SYMW = CLS == 2 && (KIND == BMI_LAPLACE_SL || KIND == BMI_HELMHOLTZ_SL);
  for (int k = tid; k < I2 * I2; k += NT) {
    const int i = k / I2, j = k - (k / I2) * I2;
    const int64_t idx = base + i * prm.rs + j * prm.cs;
    const int kt = j * I2 + i;
    if (CPLX) {
      o[2 * idx] = SYMW ? sum[k] + sum[kt] : sum[k];
      o[2 * idx + 1] = SYMW ? sum[I2 * I2 + k] + sum[I2 * I2 + kt] : sum[I2 * I2 + k];
    } else {
      o[idx] = SYMW ? sum[k] + sum[kt] : sum[k];
    }
  }
}

// far-field surface point table of all elements (DuffyTrick/farFieldQuadratureNodes.hpp:22-46)
template <int PD>
__global__ void k_ffield(const double* __restrict__ cp, int cp_stride, const int* __restrict__ elem_patch,
                         const double* __restrict__ elem_llc, int nel, int n, double h, int fs, double* out) {
  const int nq = n * n;
  const int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= static_cast<int64_t>(nel) * nq) return;
  const int e = static_cast<int>(gid / nq), k = static_cast<int>(gid % nq);
  const double* gx = c_gl_nodes + c_gl_offset[n - 1];
  const double* gw = c_gl_weights + c_gl_offset[n - 1];
  const double u = gx[k / n], v = gx[k % n], w = h * (gw[k / n] * gw[k % n]);
  double X[3], Ds[3], Dt[3];
  const double* c = cp + static_cast<int64_t>(elem_patch[e]) * cp_stride;
  surface_point<PD>(c, elem_llc[2 * e] + h * u, elem_llc[2 * e + 1] + h * v, X, Ds, Dt);
  const double c0 = Ds[1] * Dt[2] - Ds[2] * Dt[1];
  const double c1 = Ds[2] * Dt[0] - Ds[0] * Dt[2];
  const double c2 = Ds[0] * Dt[1] - Ds[1] * Dt[0];
  double* o = out + gid * fs;
  o[0] = X[0], o[1] = X[1], o[2] = X[2];
  o[3] = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
  o[4] = w, o[5] = u, o[6] = v, o[7] = 0.0;
  if (fs == FF_STRIDE_TAN) {
    o[8] = Ds[0], o[9] = Ds[1], o[10] = Ds[2];
    o[11] = Dt[0], o[12] = Dt[1], o[13] = Dt[2];
    o[14] = 0.0, o[15] = 0.0;
  }
}

void launch_ffield_table(const bmi_space* s, int n, int fs, double* out, hipStream_t st) {
  const int64_t total = static_cast<int64_t>(s->nel) * n * n;
  const int blocks = static_cast<int>((total + 255) / 256);
  const double h = 1.0 / s->n;
  const int stride = 4 * s->pd * s->pd;
  if (s->pd == 2)
    hipLaunchKernelGGL(k_ffield<2>, dim3(blocks), dim3(256), 0, st, s->d_cp.ptr, stride,
                       s->d_elem_patch.ptr, s->d_elem_llc.ptr, s->nel, n, h, fs, out);
  else if (s->pd == 5)
    hipLaunchKernelGGL(k_ffield<5>, dim3(blocks), dim3(256), 0, st, s->d_cp.ptr, stride,
                       s->d_elem_patch.ptr, s->d_elem_llc.ptr, s->nel, n, h, fs, out);
  else
    throw Error(BMI_UNSUPPORTED, "patch degree not supported on the device");
  BMI_CHECK_HIP(hipGetLastError());
}

// dynamic LDS of one workgroup of class CLS at quadrature degree deg (each workgroup sizes its
// point tables from its own pair's degree)
template <int P, int KIND, int PD, int CLS, int NT>
static size_t class_lds_bytes(int deg) {
  const int n = deg + 1, nq = n * n;
  size_t lds = sizeof(double) * ((NT / 64) + 1) * BlockAcc<P, KIND>::TOT;
  if (CLS == 1) lds = std::max(lds, 2 * nq * sizeof(Pt<KIND>));
  if (CLS == 1 && KIND != BMI_MAXWELL_SL)  // points, the lanes' vectors, the outer basis values
    lds = std::max(lds, 2 * nq * sizeof(Pt<KIND>) + static_cast<size_t>(NT) *
                                                        ((BlockAcc<P, KIND>::CPLX ? 3 : 2) * (P + 1) * (P + 1)) *
                                                        sizeof(double));
  constexpr bool FACT = KIND != BMI_MAXWELL_SL;  // factorised vertex rule
  constexpr int PS = static_cast<int>(sizeof(Pt<KIND>));
  const int cv = FACT ? vertex_shape(n, NT, PS, BlockAcc<P, KIND>::G).chunk : vertex_chunk<Pt<KIND>>(n, NT);
  const int ce = FACT ? edge_shape(n, NT, PS, BlockAcc<P, KIND>::G).chunk : edge_chunk<Pt<KIND>>(n, NT);
  if (CLS == 3) lds = std::max(lds, static_cast<size_t>(8 * n * ce) * sizeof(Pt<KIND>));
  if (CLS == 2 && (KIND == BMI_LAPLACE_SL || KIND == BMI_HELMHOLTZ_SL))
    lds = std::max(lds, static_cast<size_t>(NT / 4) * (32 * PD + 2 + 7 + (BlockAcc<P, KIND>::CPLX ? 4 : 2) *
                                                     (P + 1) * (P + 1)) * sizeof(double));  // lines, unit records
  if (CLS == 4) lds = std::max(lds, static_cast<size_t>(4 * nq + 2 * nq * cv) * sizeof(Pt<KIND>));
  return lds;
}

// one launch for the pairs [d_pairs, d_pairs + count) of class CLS; degs lists the distinct
// quadrature degrees among them, and the launch gets the largest LDS any of them needs (the
// table size is not monotone in the degree: the chunk sizes change with it)
template <int P, int KIND, int PD, int CLS>
static void launch_class(const NFPair* d_pairs, int count, const std::vector<int>& degs, NFParams prm,
                         hipStream_t st) {
  if (count == 0) return;
  constexpr int NT = (CLS == 0) ? 64 : 256;
  size_t lds = 0;
  for (int d : degs) lds = std::max(lds, class_lds_bytes<P, KIND, PD, CLS, NT>(d));
  BMI_REQUIRE(lds <= 160 * 1024, BMI_UNSUPPORTED, "quadrature degree too large for LDS staging");
  const char* name = CLS == 2 ? "nf_identical" : CLS == 3 ? "nf_edge" : CLS == 4 ? "nf_vertex" : "nf_regular";
  TimedScope ts(name, st);
  const int ncomp = KIND == BMI_MAXWELL_SL ? 4 : 1;
  hipLaunchKernelGGL((k_nearfield<P, KIND, PD, CLS, NT>), dim3(count, ncomp), dim3(NT), lds, st, d_pairs,
                     prm);
  BMI_CHECK_HIP(hipGetLastError());
}

template <int P, int KIND, int PD>
static void launch_all(const std::vector<NFPair>& pairs, const NFPair* d_pairs, NFParams prm,
                       hipStream_t st) {
  // pairs are sorted by class, then by falling degree: one launch per class holds every
  // quadrature degree of the class (each workgroup reads its pair's degree; LDS is sized for the
  // largest need among the degrees present), the most expensive pairs first so the tail of the
  // launch is short
  size_t a = 0;
  while (a < pairs.size()) {
    size_t b = a;
    std::vector<int> degs;
    while (b < pairs.size() && pairs[b].cls == pairs[a].cls) {
      if (degs.empty() || degs.back() != pairs[b].deg) degs.push_back(pairs[b].deg);
      ++b;
    }
    const int cnt = static_cast<int>(b - a);
    switch (pairs[a].cls) {
      case 0: launch_class<P, KIND, PD, 0>(d_pairs + a, cnt, degs, prm, st); break;
      case 1: launch_class<P, KIND, PD, 1>(d_pairs + a, cnt, degs, prm, st); break;
      case 2: launch_class<P, KIND, PD, 2>(d_pairs + a, cnt, degs, prm, st); break;
      case 3: launch_class<P, KIND, PD, 3>(d_pairs + a, cnt, degs, prm, st); break;
      default: launch_class<P, KIND, PD, 4>(d_pairs + a, cnt, degs, prm, st); break;
    }
    a = b;
  }
}

static void dispatch(const bmi_space* s, int kind, int P, const std::vector<NFPair>& pairs,
                     const NFPair* d_pairs, NFParams prm, hipStream_t st) {
  const int pd = s->pd;
#define BMI_NF_CASE(PP, KK, DD) \
  if (P == PP && kind == KK && pd == DD) return launch_all<PP, KK, DD>(pairs, d_pairs, prm, st);
  BMI_NF_CASE(1, BMI_LAPLACE_SL, 5)
  BMI_NF_CASE(2, BMI_LAPLACE_SL, 5)
  BMI_NF_CASE(1, BMI_HELMHOLTZ_SL, 5)
  BMI_NF_CASE(2, BMI_HELMHOLTZ_SL, 5)
  BMI_NF_CASE(1, BMI_LAPLACE_SL, 2)
  BMI_NF_CASE(2, BMI_LAPLACE_SL, 2)
  BMI_NF_CASE(1, BMI_HELMHOLTZ_SL, 2)
  BMI_NF_CASE(2, BMI_HELMHOLTZ_SL, 2)
  BMI_NF_CASE(1, BMI_MAXWELL_SL, 5)
  BMI_NF_CASE(2, BMI_MAXWELL_SL, 5)
  BMI_NF_CASE(1, BMI_MAXWELL_SL, 2)
  BMI_NF_CASE(2, BMI_MAXWELL_SL, 2)
  BMI_NF_CASE(0, BMI_DUMMY, 2)
#undef BMI_NF_CASE
  throw Error(BMI_UNSUPPORTED, "near-field kernel not instantiated for this operator/degree/patch order");
}

void launch_nearfield(const bmi_space* s, const bmi_op_desc& op, std::vector<NFPair>& pairs, double* out,
                      int64_t rs, int64_t cs, int64_t comp_rs, int64_t comp_cs, hipStream_t st) {
  ensure_device_space(s, st);
  const int ff = farfield_degree(op.kind, s->P);
  const FFTable* tab = ensure_ffield(s, ff, st, op.kind == BMI_MAXWELL_SL);
  DevBuf<NFPair> d_pairs;
  {
    // stable counting sort by class, then by decreasing degree (degrees < 50): the launches take
    // one class each, its pairs with the largest rules first
    HostTimedScope hs("assembly_host_pair_sort");
    constexpr int ND = 50;
    std::vector<int64_t> start(5 * ND + 1, 0);
    for (const NFPair& p : pairs) {
      BMI_REQUIRE(p.cls >= 0 && p.cls < 5 && p.deg >= 0 && p.deg < ND, BMI_UNSUPPORTED,
                  "near-field pair class or quadrature degree out of range");
      start[p.cls * ND + (ND - 1 - p.deg) + 1]++;
    }
    for (int b = 0; b < 5 * ND; ++b) start[b + 1] += start[b];
    std::vector<NFPair> sorted(pairs.size());
    for (const NFPair& p : pairs) sorted[start[p.cls * ND + (ND - 1 - p.deg)]++] = p;
    pairs.swap(sorted);
    d_pairs.upload(pairs, st);
  }
  NFParams prm;
  prm.cp = s->d_cp.ptr;
  prm.cp_stride = 4 * s->pd * s->pd;
  prm.elem_patch = s->d_elem_patch.ptr;
  prm.elem_llc = s->d_elem_llc.ptr;
  prm.h = 1.0 / s->n;
  prm.kre = op.kappa_re;
  prm.kim = op.kappa_im;
  prm.dummy_fn = op.dummy_fn;
  prm.ffield = tab->buf.ptr;
  prm.nff = (ff + 1) * (ff + 1);
  prm.ffs = tab->fs;
  prm.out = out;
  prm.rs = rs;
  prm.cs = cs;
  prm.comp_rs = comp_rs;
  prm.comp_cs = comp_cs;
  prm.qre = prm.qim = 0.0;
  if (op.kind == BMI_MAXWELL_SL) {
    // -1 / (kappa^2 h^2) with h = 2^-M (Maxwell/SingleLayerOperator.hpp:71-76)
    const double k2r = op.kappa_re * op.kappa_re - op.kappa_im * op.kappa_im;
    const double k2i = 2 * op.kappa_re * op.kappa_im;
    const double h2 = prm.h * prm.h;
    const double den = (k2r * k2r + k2i * k2i) * h2;
    prm.qre = -k2r / den;
    prm.qim = k2i / den;
  }
  dispatch(s, op.kind, s->P, pairs, d_pairs.ptr, prm, st);
  BMI_CHECK_HIP(hipStreamSynchronize(st));
}

void launch_integrate_jobs(const bmi_space* s, const bmi_op_desc& op, const std::vector<int>& jobs,
                           int qdeg, double* out, hipStream_t st) {
  ensure_device_space(s, st);
  const FFTable* tab = ensure_ffield(s, qdeg, st, false);
  std::vector<NFPair> pairs;
  const int nj = static_cast<int>(jobs.size() / 5);
  for (int k = 0; k < nj; ++k) {
    NFPair p;
    p.cls = jobs[5 * k];
    p.e1 = jobs[5 * k + 1];
    p.rot1 = jobs[5 * k + 2];
    p.e2 = jobs[5 * k + 3];
    p.rot2 = jobs[5 * k + 4];
    p.deg = qdeg;
    p.off = k;
    pairs.push_back(p);
  }
  DevBuf<NFPair> d_pairs;
  d_pairs.upload(pairs, st);
  NFParams prm;
  prm.cp = s->d_cp.ptr;
  prm.cp_stride = 4 * s->pd * s->pd;
  prm.elem_patch = s->d_elem_patch.ptr;
  prm.elem_llc = s->d_elem_llc.ptr;
  prm.h = 1.0 / s->n;
  prm.kre = op.kappa_re;
  prm.kim = op.kappa_im;
  prm.dummy_fn = op.dummy_fn;
  prm.ffield = tab->buf.ptr;
  prm.nff = (qdeg + 1) * (qdeg + 1);
  prm.ffs = tab->fs;
  prm.out = out;
  prm.rs = 0;
  prm.cs = 0;
  prm.comp_rs = prm.comp_cs = 0;
  prm.qre = prm.qim = 0.0;
  // jobs are launched one group per class, keeping their output slots
  std::vector<NFPair> sorted = pairs;
  std::stable_sort(sorted.begin(), sorted.end(), [](const NFPair& a, const NFPair& b) { return a.cls < b.cls; });
  DevBuf<NFPair> d_sorted;
  d_sorted.upload(sorted, st);
  dispatch(s, op.kind, s->P, sorted, d_sorted.ptr, prm, st);
  BMI_CHECK_HIP(hipStreamSynchronize(st));
}

}  // namespace bmi
