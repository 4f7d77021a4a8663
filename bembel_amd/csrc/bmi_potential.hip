// Single-layer potentials of a discrete density: DiscretePotential<*SingleLayerPotential, Op>::
// evaluate (Potential/DiscretePotential.hpp:49-103) for the Laplace, Helmholtz and Maxwell
// single-layer potentials (Laplace/SingleLayerPotential.hpp, Helmholtz/SingleLayerPotential.hpp,
// Maxwell/SingleLayerPotential.hpp).
//
// Two steps.  k_pot_sources turns the density into point sources, one per (element, quadrature
// point) of the (P+2)^2-point rule: the surface point and the weighted density there (scalar
// kinds: rho |dsx x dtx| h^2 w with rho = c^T phi / h; Maxwell: the vector value
// (dsx c_s + dtx c_t) h w and the divergence (phi_dx c_s + phi_dy c_t) w / kappa^2, without the
// surface measure).  k_pot_sum then sums kernel x source over a chunk of sources for every target
// point (sources staged through LDS in tiles, one target per lane) and k_pot_reduce adds the
// chunk sums in chunk order, so the result does not depend on scheduling.  The sum is
// compute-bound (about 40 flops per target and source for the scalar kinds).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "bmi_device.h"
#include "bmi_fmm.h"
#include "bmi_internal.h"

namespace bmi {

constexpr int POT_TILE = 256;  // sources per LDS tile = threads per workgroup
constexpr double POT_4PI_INV = 0.25 / M_PI;

// doubles per source: scalar kinds x, y, z, rho_re, rho_im; Maxwell x, y, z, J_re(3), J_im(3),
// d_re, d_im, (pad)
template <int KIND>
struct PotSrc {
  static constexpr int N = KIND == BMI_MAXWELL_SL ? 12 : 5;
};

template <int PD, int P, int KIND>
__global__ __launch_bounds__(256) void k_pot_sources(const double* __restrict__ cp, int cp_stride,
                                                     const int* __restrict__ elem_patch,
                                                     const double* __restrict__ elem_llc, int nel, double h,
                                                     const double* __restrict__ sup, int64_t vcs, double q2r,
                                                     double q2i, double* __restrict__ src) {
  constexpr int I = P + 1, I2 = I * I, n = P + 2, NQ = n * n;
  constexpr bool C = KIND != BMI_LAPLACE_SL;
  const int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= static_cast<int64_t>(nel) * NQ) return;
  const int e = static_cast<int>(gid / NQ), q = static_cast<int>(gid % NQ);
  const double* gx = c_gl_nodes + c_gl_offset[n - 1];
  const double* gw = c_gl_weights + c_gl_offset[n - 1];
  const double u = gx[q / n], v = gx[q % n], w = gw[q / n] * gw[q % n];
  double X[3], Ds[3], Dt[3];
  surface_point<PD>(cp + static_cast<int64_t>(elem_patch[e]) * cp_stride, elem_llc[2 * e] + h * u,
                    elem_llc[2 * e + 1] + h * v, X, Ds, Dt);
  double Bu[I], Bv[I], Du[I], Dv[I];
  bernstein_basis_der<P>(u, Bu, Du);
  bernstein_basis_der<P>(v, Bv, Dv);
  // density values at the point: component 0, component 1 (Maxwell) and the divergence
  double v0r = 0, v0i = 0, v1r = 0, v1i = 0, dr = 0, di = 0;
#pragma unroll
  for (int iy = 0; iy < I; ++iy)
#pragma unroll
    for (int ix = 0; ix < I; ++ix) {
      const int64_t m = static_cast<int64_t>(e) * I2 + iy * I + ix;
      const double b = Bu[ix] * Bv[iy];
      const double c0r = C ? sup[2 * m] : sup[m], c0i = C ? sup[2 * m + 1] : 0.0;
      v0r += c0r * b, v0i += c0i * b;
      if constexpr (KIND == BMI_MAXWELL_SL) {
        const double c1r = sup[2 * (m + vcs)], c1i = sup[2 * (m + vcs) + 1];
        v1r += c1r * b, v1i += c1i * b;
        const double bx = Du[ix] * Bv[iy], by = Bu[ix] * Dv[iy];
        dr += c0r * bx + c1r * by;
        di += c0i * bx + c1i * by;
      }
    }
  double* o = src + gid * PotSrc<KIND>::N;
  o[0] = X[0], o[1] = X[1], o[2] = X[2];
  if constexpr (KIND == BMI_MAXWELL_SL) {
    // (dsx c_s + dtx c_t) / h * (h^2 w) and (phi_dx c_s + phi_dy c_t) / h^2 * (h^2 w) / kappa^2
    const double hw = h * w;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      o[3 + k] = (Ds[k] * v0r + Dt[k] * v1r) * hw;
      o[6 + k] = (Ds[k] * v0i + Dt[k] * v1i) * hw;
    }
    const double ar = dr * w, ai = di * w;
    o[9] = ar * q2r - ai * q2i;  // times 1 / kappa^2 = q2r + i q2i
    o[10] = ar * q2i + ai * q2r;
    o[11] = 0.0;
  } else {
    const double c0 = Ds[1] * Dt[2] - Ds[2] * Dt[1], c1 = Ds[2] * Dt[0] - Ds[0] * Dt[2];
    const double c2 = Ds[0] * Dt[1] - Ds[1] * Dt[0];
    const double kap = sqrt(c0 * c0 + c1 * c1 + c2 * c2) * h * w;  // (c^T phi / h) |dsx x dtx| h^2 w
    o[3] = v0r * kap;
    o[4] = v0i * kap;
  }
}

// partial[chunk][comp][point] (complex interleaved for C) = sum over the chunk's sources
template <int KIND>
__global__ __launch_bounds__(POT_TILE) void k_pot_sum(const double* __restrict__ src, int64_t nsrc, int64_t chunk,
                                                      const double* __restrict__ pts, int npts, double kre,
                                                      double kim, double* __restrict__ partial) {
  constexpr int NS = PotSrc<KIND>::N;
  constexpr bool C = KIND != BMI_LAPLACE_SL;
  constexpr int DIM = KIND == BMI_MAXWELL_SL ? 3 : 1;
  __shared__ double tile[POT_TILE * NS];
  const int t = blockIdx.x * POT_TILE + threadIdx.x;
  const bool act = t < npts;
  const double x0 = act ? pts[3 * t] : 0.0, x1 = act ? pts[3 * t + 1] : 0.0, x2 = act ? pts[3 * t + 2] : 0.0;
  const int64_t s0 = static_cast<int64_t>(blockIdx.y) * chunk;
  const int64_t s1 = std::min(nsrc, s0 + chunk);
  double ar[DIM], ai[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) ar[d] = 0.0, ai[d] = 0.0;
  for (int64_t b = s0; b < s1; b += POT_TILE) {
    const int cnt = static_cast<int>(std::min<int64_t>(POT_TILE, s1 - b));
    __syncthreads();
    for (int k = threadIdx.x; k < cnt * NS; k += POT_TILE) tile[k] = src[b * NS + k];
    __syncthreads();
    for (int j = 0; j < cnt; ++j) {
      const double* sj = tile + j * NS;
      const double d0 = x0 - sj[0], d1 = x1 - sj[1], d2 = x2 - sj[2];
      const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
      const double r = sqrt(r2);
      const double ir = 1.0 / r;
      if constexpr (KIND == BMI_LAPLACE_SL) {
        ar[0] += POT_4PI_INV * ir * sj[3];
      } else {
        // exp(-i kappa r) / (4 pi r) = exp(kim r) (cos(kre r) - i sin(kre r)) / (4 pi r)
        double sn, cs;
        sincos(kre * r, &sn, &cs);
        const double mag = exp(kim * r) * POT_4PI_INV * ir;
        const double Kr = mag * cs, Ki = -mag * sn;
        if constexpr (KIND == BMI_HELMHOLTZ_SL) {
          ar[0] += Kr * sj[3] - Ki * sj[4];
          ai[0] += Kr * sj[4] + Ki * sj[3];
        } else {
          // grad_x K = K (-1 - i kappa r) / r^2 (x - y)
          const double fr = -1.0 + kim * r, fi = -kre * r;
          const double ir2 = ir * ir;
          const double Gr = (Kr * fr - Ki * fi) * ir2, Gi = (Kr * fi + Ki * fr) * ir2;
          const double dd[3] = {d0, d1, d2};
          // G * div (complex)
          const double gdr = Gr * sj[9] - Gi * sj[10], gdi = Gr * sj[10] + Gi * sj[9];
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            ar[k] += Kr * sj[3 + k] - Ki * sj[6 + k] + gdr * dd[k];
            ai[k] += Kr * sj[6 + k] + Ki * sj[3 + k] + gdi * dd[k];
          }
        }
      }
    }
  }
  if (!act) return;
  const int64_t base = static_cast<int64_t>(blockIdx.y) * DIM * npts;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const int64_t o = base + static_cast<int64_t>(d) * npts + t;
    if constexpr (C) {
      partial[2 * o] = ar[d], partial[2 * o + 1] = ai[d];
    } else {
      partial[o] = ar[d];
    }
  }
}

// out[comp * npts + point] = sum of the chunk partials in chunk order
__global__ void k_pot_reduce(const double* __restrict__ partial, int nchunk, int64_t per_chunk,
                             double* __restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= per_chunk) return;
  double acc = 0.0;
  for (int c = 0; c < nchunk; ++c) acc += partial[static_cast<int64_t>(c) * per_chunk + i];
  out[i] = acc;
}

}  // namespace bmi

using namespace bmi;

extern "C" bmi_status bmi_potential_evaluate(const bmi_space* s, const bmi_op_desc* op, const double* rho,
                                             int npoints, const double* points, double* out, int mem,
                                             void* stream) {
  try {
    BMI_REQUIRE(s && op && rho && points && out && npoints >= 0, BMI_INVALID_ARG, "bad argument");
    const int kind = op->kind;
    BMI_REQUIRE(kind == BMI_LAPLACE_SL || kind == BMI_HELMHOLTZ_SL || kind == BMI_MAXWELL_SL, BMI_INVALID_ARG,
                "potentials exist for the Laplace, Helmholtz and Maxwell single layers");
    BMI_REQUIRE((kind == BMI_MAXWELL_SL) == (s->form == BMI_FORM_DIVCONFORMING), BMI_INVALID_ARG,
                "the Maxwell potential needs a DivConforming space, the scalar ones a discontinuous space");
    if (npoints == 0) return BMI_OK;
    hipStream_t st = static_cast<hipStream_t>(stream);
    ensure_device_space(s, st);
    const bool C = kind != BMI_LAPLACE_SL;
    const int S = C ? 2 : 1;
    const int dim = kind == BMI_MAXWELL_SL ? 3 : 1;
    const int64_t ndof = s->ndof;
    // operands on the device
    DevBuf<double> d_rho, d_pts, d_out, sup;
    const double* r = rho;
    const double* p = points;
    double* o = out;
    if (mem != BMI_MEM_DEVICE) {
      d_rho.upload(std::vector<double>(rho, rho + S * ndof), st);
      d_pts.upload(std::vector<double>(points, points + 3 * static_cast<int64_t>(npoints)), st);
      d_out.alloc(static_cast<size_t>(S) * dim * npoints);
      r = d_rho.ptr, p = d_pts.ptr, o = d_out.ptr;
    }
    // super coefficients T rho
    sup.alloc(static_cast<size_t>(S) * s->nsuper);
    launch_spmv(C, s->nsuper, 1, s->d_rowptr.ptr, s->d_col.ptr, s->d_val.ptr, r, ndof, sup.ptr, s->nsuper, nullptr,
                nullptr, st);
    // sources
    const int NQ = (s->P + 2) * (s->P + 2);
    const int64_t nsrc = static_cast<int64_t>(s->nel) * NQ;
    const int ns = kind == BMI_MAXWELL_SL ? PotSrc<BMI_MAXWELL_SL>::N : PotSrc<BMI_LAPLACE_SL>::N;
    DevBuf<double> src;
    src.alloc(static_cast<size_t>(nsrc) * ns);
    // 1 / kappa^2 for the Maxwell divergence term
    const double k2r = op->kappa_re * op->kappa_re - op->kappa_im * op->kappa_im;
    const double k2i = 2 * op->kappa_re * op->kappa_im;
    const double den = k2r * k2r + k2i * k2i;
    BMI_REQUIRE(kind != BMI_MAXWELL_SL || den > 0, BMI_INVALID_ARG, "the Maxwell potential needs kappa != 0");
    const double q2r = kind == BMI_MAXWELL_SL ? k2r / den : 0.0, q2i = kind == BMI_MAXWELL_SL ? -k2i / den : 0.0;
    const unsigned sb = static_cast<unsigned>((nsrc + 255) / 256);
    const int stride = 4 * s->pd * s->pd;
    const double h = 1.0 / s->n;
    const int64_t vcs = static_cast<int64_t>(s->nel) * s->I2;
#define BMI_POT_SRC(PD, PP, KK)                                                                                  \
  if (s->pd == PD && s->P == PP && kind == KK) {                                                                 \
    hipLaunchKernelGGL((k_pot_sources<PD, PP, KK>), dim3(sb), dim3(256), 0, st, s->d_cp.ptr, stride,           \
                       s->d_elem_patch.ptr, s->d_elem_llc.ptr, s->nel, h, sup.ptr, vcs, q2r, q2i, src.ptr);    \
  } else
    BMI_POT_SRC(5, 1, BMI_LAPLACE_SL) BMI_POT_SRC(5, 2, BMI_LAPLACE_SL) BMI_POT_SRC(2, 1, BMI_LAPLACE_SL)
    BMI_POT_SRC(2, 2, BMI_LAPLACE_SL) BMI_POT_SRC(5, 1, BMI_HELMHOLTZ_SL) BMI_POT_SRC(5, 2, BMI_HELMHOLTZ_SL)
    BMI_POT_SRC(2, 1, BMI_HELMHOLTZ_SL) BMI_POT_SRC(2, 2, BMI_HELMHOLTZ_SL) BMI_POT_SRC(5, 1, BMI_MAXWELL_SL)
    BMI_POT_SRC(5, 2, BMI_MAXWELL_SL) BMI_POT_SRC(2, 1, BMI_MAXWELL_SL) BMI_POT_SRC(2, 2, BMI_MAXWELL_SL) {
      throw Error(BMI_UNSUPPORTED, "potential kernel not instantiated for this degree / patch order");
    }
#undef BMI_POT_SRC
    BMI_CHECK_HIP(hipGetLastError());
    // sums: enough chunks that the grid holds about 1024 workgroups, each chunk whole tiles
    const int ptiles = (npoints + POT_TILE - 1) / POT_TILE;
    const int64_t want = std::max<int64_t>(1, 1024 / ptiles);
    int64_t chunk = (nsrc + want - 1) / want;
    chunk = std::max<int64_t>(POT_TILE, (chunk + POT_TILE - 1) / POT_TILE * POT_TILE);
    const int nchunk = static_cast<int>((nsrc + chunk - 1) / chunk);
    const int64_t per_chunk = static_cast<int64_t>(S) * dim * npoints;
    DevBuf<double> part;
    part.alloc(static_cast<size_t>(nchunk) * per_chunk);
    const dim3 grid(static_cast<unsigned>(ptiles), static_cast<unsigned>(nchunk));
    {
      TimedScope ts("potential", st);
      if (kind == BMI_LAPLACE_SL)
        hipLaunchKernelGGL(k_pot_sum<BMI_LAPLACE_SL>, grid, dim3(POT_TILE), 0, st, src.ptr, nsrc, chunk, p, npoints,
                           op->kappa_re, op->kappa_im, part.ptr);
      else if (kind == BMI_HELMHOLTZ_SL)
        hipLaunchKernelGGL(k_pot_sum<BMI_HELMHOLTZ_SL>, grid, dim3(POT_TILE), 0, st, src.ptr, nsrc, chunk, p,
                           npoints, op->kappa_re, op->kappa_im, part.ptr);
      else
        hipLaunchKernelGGL(k_pot_sum<BMI_MAXWELL_SL>, grid, dim3(POT_TILE), 0, st, src.ptr, nsrc, chunk, p, npoints,
                           op->kappa_re, op->kappa_im, part.ptr);
      BMI_CHECK_HIP(hipGetLastError());
      hipLaunchKernelGGL(k_pot_reduce, dim3(static_cast<unsigned>((per_chunk + 255) / 256)), dim3(256), 0, st,
                         part.ptr, nchunk, per_chunk, o);
      BMI_CHECK_HIP(hipGetLastError());
    }
    if (mem != BMI_MEM_DEVICE)
      BMI_CHECK_HIP(hipMemcpyAsync(out, d_out.ptr, d_out.bytes(), hipMemcpyDeviceToHost, st));
    BMI_CHECK_HIP(hipStreamSynchronize(st));
    return BMI_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.status;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return BMI_DEVICE_ERROR;
  }
}
