// Device helpers shared by the kernels: Gauss-Legendre tables in constant memory, Bernstein
// bases and rational Bezier surface points.
#pragma once
#include <hip/hip_runtime.h>

#include "gauss_legendre_table.inc"

namespace bmi {

static __constant__ int c_gl_offset[52] = BMI_GL_OFFSET_INIT;
static __constant__ double c_gl_nodes[BMI_GL_TABLE_SIZE] = BMI_GL_NODES_INIT;
static __constant__ double c_gl_weights[BMI_GL_TABLE_SIZE] = BMI_GL_WEIGHTS_INIT;

template <int P>
struct Binom {
  static constexpr double c(int k) {
    double r = 1.0;
    for (int i = 1; i <= k; ++i) r = r * (P - k + i) / i;
    return r;
  }
};

// Bernstein basis of degree P at x (Spline/Bernstein.hpp:44-48)
template <int P>
__device__ __forceinline__ void bernstein_basis(double x, double* B) {
  if constexpr (P == 0) {
    B[0] = 1.0;
  } else {
    const double y = 1.0 - x;
    double xp[P + 1], yp[P + 1];
    xp[0] = 1.0, yp[0] = 1.0;
#pragma unroll
    for (int k = 1; k <= P; ++k) xp[k] = xp[k - 1] * x, yp[k] = yp[k - 1] * y;
#pragma unroll
    for (int k = 0; k <= P; ++k) B[k] = Binom<P>::c(k) * xp[k] * yp[P - k];
  }
}

// Bernstein basis and derivative of degree P (Bernstein.hpp:239-244)
template <int P>
__device__ __forceinline__ void bernstein_basis_der(double x, double* B, double* D) {
  if constexpr (P == 0) {
    B[0] = 1.0;
    D[0] = 0.0;
  } else {
    const double y = 1.0 - x;
    double xp[P + 1], yp[P + 1];
    xp[0] = 1.0, yp[0] = 1.0;
#pragma unroll
    for (int k = 1; k <= P; ++k) xp[k] = xp[k - 1] * x, yp[k] = yp[k - 1] * y;
#pragma unroll
    for (int k = 0; k <= P; ++k) B[k] = Binom<P>::c(k) * xp[k] * yp[P - k];
    double L[P];  // degree P-1 basis
#pragma unroll
    for (int k = 0; k < P; ++k) L[k] = Binom<P - 1>::c(k) * xp[k] * yp[P - 1 - k];
#pragma unroll
    for (int k = 0; k <= P; ++k) {
      const double lo = k >= 1 ? L[k - 1] : 0.0;
      const double hi = k <= P - 1 ? L[k] : 0.0;
      D[k] = P * (lo - hi);
    }
  }
}

// Rational Bezier patch with PD x PD homogeneous control points (cp[4*(i*PD+j)+k]):
// point and parameter derivatives by the quotient rule (Geometry/Patch.hpp:294-365).
// The contraction is done one direction at a time (x first, then y).
template <int PD>
__device__ __forceinline__ void surface_point(const double* __restrict__ cp, double s, double t, double X[3],
                                              double Ds[3], double Dt[3]) {
  double Bs[PD], Es[PD], Bt[PD], Et[PD];
  bernstein_basis_der<PD - 1>(s, Bs, Es);
  bernstein_basis_der<PD - 1>(t, Bt, Et);
  double v[4] = {0, 0, 0, 0}, vs[4] = {0, 0, 0, 0}, vt[4] = {0, 0, 0, 0};
#pragma unroll
  for (int j = 0; j < PD; ++j) {
    double r[4] = {0, 0, 0, 0}, rs[4] = {0, 0, 0, 0};
#pragma unroll
    for (int i = 0; i < PD; ++i) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const double c = cp[4 * (i * PD + j) + k];
        r[k] += Bs[i] * c;
        rs[k] += Es[i] * c;
      }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      v[k] += Bt[j] * r[k];
      vs[k] += Bt[j] * rs[k];
      vt[k] += Et[j] * r[k];
    }
  }
  const double bot = 1.0 / v[3];
  const double b2 = bot * bot;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    X[k] = v[k] * bot;
    Ds[k] = (vs[k] * v[3] - v[k] * vs[3]) * b2;
    Dt[k] = (vt[k] * v[3] - v[k] * vt[3]) * b2;
  }
}

}  // namespace bmi
