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
constexpr int FF_STRIDE = 8;
constexpr int FF_STRIDE_TAN = 16;

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

// per-lane block accumulator
template <int P, int KIND>
struct BlockAcc {
  static constexpr int I = P + 1;
  static constexpr int I2 = I * I;
  static constexpr bool CPLX = KIND == BMI_HELMHOLTZ_SL || KIND == BMI_MAXWELL_SL;
  double re[I2 * I2];
  double im[CPLX ? I2 * I2 : 1];

  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int k = 0; k < I2 * I2; ++k) re[k] = 0.0;
    if (CPLX) {
#pragma unroll
      for (int k = 0; k < I2 * I2; ++k) im[k] = 0.0;
    }
  }

  // Maxwell: component block (rc, cc) of the EFIE integrand, vector part
  // K w w (t_rc(a) . t_cc(b)) phi_i(a) phi_j(b) plus divergence part
  // -K w w / (kappa^2 h^2) dphi_i^rc(a) dphi_j^cc(b) (Maxwell/SingleLayerOperator.hpp:44-88,
  // SuperSpace.hpp:236-283, Basis.hpp:170-200)
  __device__ __forceinline__ void add_maxwell(const Pt<KIND>& a, const Pt<KIND>& b, const NFParams& prm, int rc,
                                              int cc) {
    const double d0 = a.x0 - b.x0, d1 = a.x1 - b.x1, d2 = a.x2 - b.x2;
    const double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    double sn, cs;
    sincos(prm.kre * r, &sn, &cs);
    const double mag = exp(prm.kim * r) * (a.W * b.W) * (0.25 * M_1_PI) / r;
    const double kr = mag * cs, ki = -mag * sn;
    const double* ta = a.t + 3 * rc;
    const double* tb = b.t + 3 * cc;
    const double dt = ta[0] * tb[0] + ta[1] * tb[1] + ta[2] * tb[2];
    const double vr = kr * dt, vi = ki * dt;
    const double qr = kr * prm.qre - ki * prm.qim, qi = kr * prm.qim + ki * prm.qre;
    double X[I], Y[I], U[I], V[I], DX[I], DY[I], DU[I], DV[I];
    bernstein_basis_der<P>(a.s0, X, DX);
    bernstein_basis_der<P>(a.s1, Y, DY);
    bernstein_basis_der<P>(b.s0, U, DU);
    bernstein_basis_der<P>(b.s1, V, DV);
    double bv[I2], bd[I2];
#pragma unroll
    for (int jy = 0; jy < I; ++jy)
#pragma unroll
      for (int jx = 0; jx < I; ++jx) {
        bv[jy * I + jx] = U[jx] * V[jy];
        bd[jy * I + jx] = cc == 0 ? DU[jx] * V[jy] : U[jx] * DV[jy];
      }
#pragma unroll
    for (int iy = 0; iy < I; ++iy)
#pragma unroll
      for (int ix = 0; ix < I; ++ix) {
        const double av = X[ix] * Y[iy];
        const double ad = rc == 0 ? DX[ix] * Y[iy] : X[ix] * DY[iy];
        const double ar = vr * av, ai = vi * av, dr = qr * ad, di = qi * ad;
        const int row = (iy * I + ix) * I2;
#pragma unroll
        for (int j = 0; j < I2; ++j) {
          re[row + j] += ar * bv[j] + dr * bd[j];
          im[row + j] += ai * bv[j] + di * bd[j];
        }
      }
  }

  // integrand of the pair (a, b) times the basis product, added to the block
  __device__ __forceinline__ void add(const Pt<KIND>& a, const Pt<KIND>& b, const NFParams& prm, int rc = 0,
                                      int cc = 0) {
    if constexpr (KIND == BMI_MAXWELL_SL) {
      add_maxwell(a, b, prm, rc, cc);
      return;
    } else {
    double vr, vi = 0.0;
    const double d0 = a.x0 - b.x0, d1 = a.x1 - b.x1, d2 = a.x2 - b.x2;
    if (KIND == BMI_LAPLACE_SL) {
      const double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
      vr = (a.W * b.W) * (0.25 * M_1_PI) / r;
    } else if (KIND == BMI_HELMHOLTZ_SL) {
      // exp(-i kappa r) / (4 pi r), kappa = kre + i kim (Helmholtz/SingleLayerOperator.hpp:115-120)
      const double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
      double sn, cs;
      sincos(prm.kre * r, &sn, &cs);
      const double mag = exp(prm.kim * r) * (a.W * b.W) * (0.25 * M_1_PI) / r;
      vr = mag * cs;
      vi = -mag * sn;
    } else {
      vr = dummy_value(prm.dummy_fn, a, b) * a.W * b.W;
    }
    double X[I], Y[I], U[I], V[I];
    bernstein_basis<P>(a.s0, X);
    bernstein_basis<P>(a.s1, Y);
    bernstein_basis<P>(b.s0, U);
    bernstein_basis<P>(b.s1, V);
    double bv[I2];
#pragma unroll
    for (int jy = 0; jy < I; ++jy)
#pragma unroll
      for (int jx = 0; jx < I; ++jx) bv[jy * I + jx] = U[jx] * V[jy];
#pragma unroll
    for (int iy = 0; iy < I; ++iy)
#pragma unroll
      for (int ix = 0; ix < I; ++ix) {
        const double av = X[ix] * Y[iy];
        const double ar = vr * av;
        const int row = (iy * I + ix) * I2;
#pragma unroll
        for (int j = 0; j < I2; ++j) re[row + j] += ar * bv[j];
        if (CPLX) {
          const double ai = vi * av;
#pragma unroll
          for (int j = 0; j < I2; ++j) im[row + j] += ai * bv[j];
        }
      }
    }
  }
};

// DuffyTrick::tau (DuffyTrick/tau.hpp:26-42)
__device__ __forceinline__ void tau(double x, double y, int c, double& u, double& v) {
  switch (c) {
    case 1: u = 1 - y, v = x; return;
    case 2: u = 1 - x, v = 1 - y; return;
    case 3: u = y, v = 1 - x; return;
    default: u = x, v = y; return;
  }
}

template <int NT, int N>
__device__ __forceinline__ void block_reduce_store(double (&acc)[N], double* lds, int tid,
                                                   double* out_sum) {
  // lane reduction inside each wave, then across waves through LDS
#pragma unroll
  for (int k = 0; k < N; ++k) {
    double v = acc[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    acc[k] = v;
  }
  constexpr int NW = NT / 64;
  const int wave = tid >> 6, lane = tid & 63;
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < N; ++k) lds[wave * N + k] = acc[k];
  }
  __syncthreads();
  for (int k = tid; k < N; k += NT) {
    double s = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) s += lds[w * N + k];
    out_sum[k] = s;
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

  if constexpr (CLS == 0) {
    // integrate0: far-field tables of both elements
    const int nf = prm.nff, fs = prm.ffs;
    const double* f1 = prm.ffield + static_cast<int64_t>(pr.e1) * nf * fs;
    const double* f2 = prm.ffield + static_cast<int64_t>(pr.e2) * nf * fs;
    for (int idx = tid; idx < nf * nf; idx += NT) {
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
      acc.add(pa, pb, prm, rc, cc);
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
    for (int idx = tid; idx < nq * nq; idx += NT) {
      const int i = idx / nq, j = idx - (idx / nq) * nq;
      acc.add(q1[i], q2[j], prm, rc, cc);
    }
    __syncthreads();
  } else if constexpr (CLS == 2) {
    // integrate2: identical elements, 8 Duffy sub-simplices, symmetrised
    for (int idx = tid; idx < nq * nq; idx += NT) {
      const int i = idx / nq, j = idx - (idx / nq) * nq;
      const double xi0 = gx[i / n], xi1 = gx[i % n], wi = gw[i / n] * gw[i % n];
      const double xj0 = gx[j / n], xj1 = gx[j % n], wj = gw[j / n] * gw[j % n];
      const double w = h * h * wi * xi0 * (1 - xi0) * (1 - xi0 * xi1);
      const double t1 = xj0 * (1 - xi0);
      const double t2 = xj1 * (1 - xi0 * xi1);
      const double t3 = t1 + xi0;
      const double t4 = t2 + xi0 * xi1;
      const double pu[8] = {t1, t3, t1, t3, t2, t4, t2, t4};
      const double pv[8] = {t2, t4, t4, t2, t1, t3, t3, t1};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const PtK a = make_point<PD, KIND>(cp1, l1x, l1y, h, pu[2 * k], pv[2 * k], w);
        const PtK b = make_point<PD, KIND>(cp1, l1x, l1y, h, pu[2 * k + 1], pv[2 * k + 1], wj);
        acc.add(a, b, prm, rc, cc);
        acc.add(b, a, prm, rc, cc);
      }
    }
  } else if constexpr (CLS == 3) {
    // integrate3: common edge, 6 Duffy sub-regions
    const int r1 = pr.rot1, r2 = pr.rot2;
    for (int idx = tid; idx < nq * nq; idx += NT) {
      const int i = idx / nq, j = idx - (idx / nq) * nq;
      const double xi0 = gx[i / n], xi1 = gx[i % n], wi = gw[i / n] * gw[i % n];
      const double xj0 = gx[j / n], xj1 = gx[j % n], wj = gw[j / n] * gw[j % n];
      const double w = h * h * wi * xi1 * xi1;
      const double t1 = xi0 * (1 - xi1);
      const double t2 = (1 - xi0) * (1 - xi1);
      const double e0 = xi1 * xj0, e1 = xi1 * xj1;
      const double t3 = xi0 * (1 - e0);
      const double t4 = (1 - xi0) * (1 - e0);
      const double wa = wj * (1 - xi1), wb = wj * (1 - e0);
      const double A0[6] = {t1, 1 - t1, t3, 1 - t3, t4, 1 - t4};
      const double A1[6] = {e0, e0, xi1, xi1, e1, e1};
      const double B0[6] = {t2, 1 - t2, t4, 1 - t4, t3, 1 - t3};
      const double B1[6] = {e1, e1, e1, e1, xi1, xi1};
      const double BW[6] = {wa, wa, wb, wb, wb, wb};
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        double u, v;
        tau(A0[k], A1[k], r1, u, v);
        const PtK a = make_point<PD, KIND>(cp1, l1x, l1y, h, u, v, w);
        tau(B0[k], B1[k], r2, u, v);
        const PtK b = make_point<PD, KIND>(cp2, l2x, l2y, h, u, v, BW[k]);
        acc.add(a, b, prm, rc, cc);
      }
    }
  } else {
    // integrate4: common vertex; the four outer points per i are staged in LDS
    const int r1 = pr.rot1, r2 = pr.rot2;
    PtK* q = reinterpret_cast<PtK*>(lds);  // 4 * nq
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
    __syncthreads();
    for (int idx = tid; idx < nq * nq; idx += NT) {
      const int i = idx / nq, j = idx - (idx / nq) * nq;
      const double xi0 = gx[i / n];
      const double e0 = xi0 * gx[j / n], e1 = xi0 * gx[j % n], wj = gw[j / n] * gw[j % n];
      double u, v;
      tau(e0, e1, r2, u, v);
      const PtK p5 = make_point<PD, KIND>(cp2, l2x, l2y, h, u, v, wj);
      tau(e0, e1, r1, u, v);
      const PtK p6 = make_point<PD, KIND>(cp1, l1x, l1y, h, u, v, wj);
      acc.add(q[4 * i + 0], p5, prm, rc, cc);
      acc.add(q[4 * i + 1], p5, prm, rc, cc);
      acc.add(p6, q[4 * i + 2], prm, rc, cc);
      acc.add(p6, q[4 * i + 3], prm, rc, cc);
    }
    __syncthreads();
  }

  // reduce the per-lane blocks and write the result
  double* red = lds;
  double* sum = lds + (NT / 64) * I2 * I2;
  block_reduce_store<NT>(acc.re, red, tid, sum);
  double* o = prm.out;
  constexpr bool CPLX = Acc::CPLX;
  const int64_t base = pr.off + rc * prm.comp_rs + cc * prm.comp_cs;
  for (int k = tid; k < I2 * I2; k += NT) {
    const int i = k / I2, j = k - (k / I2) * I2;
    const int64_t idx = base + i * prm.rs + j * prm.cs;
    if (CPLX)
      o[2 * idx] = sum[k];
    else
      o[idx] = sum[k];
  }
  if constexpr (CPLX) {
    __syncthreads();
    block_reduce_store<NT>(acc.im, red, tid, sum);
    for (int k = tid; k < I2 * I2; k += NT) {
      const int i = k / I2, j = k - (k / I2) * I2;
      const int64_t idx = base + i * prm.rs + j * prm.cs;
      o[2 * idx + 1] = sum[k];
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

// the table always carries the tangents when they are asked for once (Maxwell), so one
// table serves every operator of the space
void ensure_ffield(const bmi_space* s, int deg, hipStream_t st, bool tangents) {
  ensure_device_space(s, st);
  const int fs = (tangents || s->ffield_stride == FF_STRIDE_TAN) ? FF_STRIDE_TAN : FF_STRIDE;
  if (s->ffield_deg == deg && s->ffield_stride == fs) return;
  const int n = deg + 1;
  s->d_ffield.alloc(static_cast<size_t>(s->nel) * n * n * fs);
  const int64_t total = static_cast<int64_t>(s->nel) * n * n;
  const int blocks = static_cast<int>((total + 255) / 256);
  const double h = 1.0 / s->n;
  const int stride = 4 * s->pd * s->pd;
  if (s->pd == 2)
    hipLaunchKernelGGL(k_ffield<2>, dim3(blocks), dim3(256), 0, st, s->d_cp.ptr, stride,
                       s->d_elem_patch.ptr, s->d_elem_llc.ptr, s->nel, n, h, fs, s->d_ffield.ptr);
  else if (s->pd == 5)
    hipLaunchKernelGGL(k_ffield<5>, dim3(blocks), dim3(256), 0, st, s->d_cp.ptr, stride,
                       s->d_elem_patch.ptr, s->d_elem_llc.ptr, s->nel, n, h, fs, s->d_ffield.ptr);
  else
    throw Error(BMI_UNSUPPORTED, "patch degree not supported on the device");
  BMI_CHECK_HIP(hipGetLastError());
  s->ffield_deg = deg;
  s->ffield_stride = fs;
}

template <int P, int KIND, int PD, int CLS>
static void launch_class(const NFPair* d_pairs, int count, int deg, NFParams prm, hipStream_t st) {
  if (count == 0) return;
  constexpr int NT = (CLS == 0) ? 64 : 256;
  constexpr int I2 = (P + 1) * (P + 1);
  const int nq = (deg + 1) * (deg + 1);
  size_t lds = sizeof(double) * ((NT / 64) * I2 * I2 + I2 * I2);
  if (CLS == 1) lds = std::max(lds, 2 * nq * sizeof(Pt<KIND>));
  if (CLS == 4) lds = std::max(lds, 4 * nq * sizeof(Pt<KIND>));
  BMI_REQUIRE(lds <= 160 * 1024, BMI_UNSUPPORTED, "quadrature degree too large for LDS staging");
  const char* name = CLS >= 2 ? "nf_duffy" : "nf_regular";
  TimedScope ts(name, st);
  const int ncomp = KIND == BMI_MAXWELL_SL ? 4 : 1;
  hipLaunchKernelGGL((k_nearfield<P, KIND, PD, CLS, NT>), dim3(count, ncomp), dim3(NT), lds, st, d_pairs,
                     prm);
  BMI_CHECK_HIP(hipGetLastError());
}

template <int P, int KIND, int PD>
static void launch_all(const std::vector<NFPair>& pairs, const NFPair* d_pairs, NFParams prm,
                       hipStream_t st) {
  // pairs are sorted by (class, degree); one launch per group
  size_t a = 0;
  while (a < pairs.size()) {
    size_t b = a;
    while (b < pairs.size() && pairs[b].cls == pairs[a].cls && pairs[b].deg == pairs[a].deg) ++b;
    const int cnt = static_cast<int>(b - a), deg = pairs[a].deg;
    switch (pairs[a].cls) {
      case 0: launch_class<P, KIND, PD, 0>(d_pairs + a, cnt, deg, prm, st); break;
      case 1: launch_class<P, KIND, PD, 1>(d_pairs + a, cnt, deg, prm, st); break;
      case 2: launch_class<P, KIND, PD, 2>(d_pairs + a, cnt, deg, prm, st); break;
      case 3: launch_class<P, KIND, PD, 3>(d_pairs + a, cnt, deg, prm, st); break;
      default: launch_class<P, KIND, PD, 4>(d_pairs + a, cnt, deg, prm, st); break;
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
  ensure_ffield(s, ff, st, op.kind == BMI_MAXWELL_SL);
  std::stable_sort(pairs.begin(), pairs.end(), [](const NFPair& a, const NFPair& b) {
    if (a.cls != b.cls) return a.cls < b.cls;
    return a.deg > b.deg;
  });
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
  prm.ffield = s->d_ffield.ptr;
  prm.nff = (ff + 1) * (ff + 1);
  prm.ffs = s->ffield_stride;
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
  ensure_ffield(s, qdeg, st, false);
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
  prm.ffield = s->d_ffield.ptr;
  prm.nff = (qdeg + 1) * (qdeg + 1);
  prm.ffs = s->ffield_stride;
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
