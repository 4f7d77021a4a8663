// Device solvers (batched restarted GMRES with CGS2 Arnoldi in FP64, batched CG) and the
// Dirichlet-trace right-hand sides.
//
// The consumers of the reference are Eigen::GMRES / Eigen::ConjugateGradient over the H2
// product (examples/HelmholtzSingleLayerH2.cpp:82-85, examples/LaplaceSingleLayerH2.cpp:78-81).
// Here all right-hand sides advance together: one batched H2 product per Arnoldi step, all
// inner products of a step in one launch (classical Gram-Schmidt applied twice), and the tiny
// Hessenberg/Givens work per column on the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <cmath>
#include <complex>
#include <vector>

#include "bmi_device.h"
#include "bmi_fmm.h"
#include "bmi_internal.h"
#include "bmi_solver.h"

namespace bmi {

__global__ void k_to_fp32(const double* src, float* dst, size_t n) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = static_cast<float>(src[i]);
}
void convert_to_fp32(const double* src, float* dst, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_to_fp32, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, st, src, dst, n);
  BMI_CHECK_HIP(hipGetLastError());
}

// Inner products of many vectors with one in two passes: block (chunk, column c, group g of four
// vectors) sums its rows of conj(V_j) W for j = 4g .. 4g + 3 into
// part[((j * nrhs + c) * nchunk + chunk)] (the groups run side by side, so a Gram-Schmidt pass
// over k vectors costs one round of loads, not k / 4), then one wave per (j, c) adds the chunk
// sums in a fixed order (deterministic: the chunk count depends on the sizes only).
constexpr int DOT_CHUNKS = 64;
// Hessenberg column and rotations the Givens kernel stages in LDS (48 KB: 1,000 complex steps);
// later steps of longer cycles read them from memory
constexpr int GIVENS_STAGE_DOUBLES = 6144;
// Gram-Schmidt coefficients the subtraction kernel holds in LDS (48 KB: 3,072 complex vectors)
constexpr int GS_LDS_DOUBLES = 6144;
constexpr int DOT_JB = 4;
// chunks per column: 128 for one or two columns; with more columns about 512 workgroups per group
// of four vectors, each chunk at least 256 rows (a workgroup's reduction then serves several rows
// per thread instead of one)
static inline int dot_chunks(int64_t n, int nrhs) {
  if (nrhs <= 2) return 2 * DOT_CHUNKS;
  const int64_t want = (512 + nrhs - 1) / nrhs, cap = std::max<int64_t>(1, (n + 255) / 256);
  return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(std::min<int64_t>(want, cap), 2 * DOT_CHUNKS)));
}
template <bool C>
__global__ __launch_bounds__(256) void k_dots_part(const double* __restrict__ V, int64_t vstride, int nv,
                                                    const double* __restrict__ W, int64_t n, int nrhs, int nchunk,
                                                    double* __restrict__ part) {
  constexpr int S = C ? 2 : 1;
  constexpr int JB = DOT_JB;
  __shared__ double red[2][JB][4];
  const int chunk = blockIdx.x, c = blockIdx.y;
  const int64_t per = (n + nchunk - 1) / nchunk;
  const int64_t r0 = chunk * per, r1 = r0 + per < n ? r0 + per : n;
  const double* w = W + static_cast<int64_t>(S) * n * c;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  {
    const int j0 = static_cast<int>(blockIdx.z) * JB;
    const int nj = nv - j0 < JB ? nv - j0 : JB;
    double ar[JB], ai[JB];
#pragma unroll
    for (int q = 0; q < JB; ++q) ar[q] = 0.0, ai[q] = 0.0;
    for (int64_t i = r0 + threadIdx.x; i < r1; i += blockDim.x) {
      const double wr = w[S * i], wi = C ? w[S * i + 1] : 0.0;
#pragma unroll
      for (int q = 0; q < JB; ++q) {
        if (q < nj) {
          const double* v = V + (j0 + q) * vstride + static_cast<int64_t>(S) * n * c;
          if (C) {
            const double vr = v[2 * i], vi = v[2 * i + 1];
            ar[q] += vr * wr + vi * wi;
            ai[q] += vr * wi - vi * wr;
          } else {
            ar[q] += v[i] * wr;
          }
        }
      }
    }
#pragma unroll
    for (int q = 0; q < JB; ++q) {
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        ar[q] += __shfl_xor(ar[q], off, 64);
        if (C) ai[q] += __shfl_xor(ai[q], off, 64);
      }
      if (lane == 0) red[0][q][wave] = ar[q], red[1][q][wave] = ai[q];
    }
    __syncthreads();
    if (threadIdx.x < nj) {
      const int q = threadIdx.x;
      const int64_t o = (static_cast<int64_t>(j0 + q) * nrhs + c) * nchunk + chunk;
      part[S * o] = (red[0][q][0] + red[0][q][1]) + (red[0][q][2] + red[0][q][3]);
      if (C) part[S * o + 1] = (red[1][q][0] + red[1][q][1]) + (red[1][q][2] + red[1][q][3]);
    }
  }
}
// one wave per inner product: lane l adds chunks l, l + 64, ... in order, then a fixed shuffle tree
template <bool C>
__global__ __launch_bounds__(256) void k_dots_final(const double* __restrict__ part, int count, int nchunk,
                                                     double* out) {
  constexpr int S = C ? 2 : 1;
  const int q = static_cast<int>(blockIdx.x) * 4 + static_cast<int>(threadIdx.x >> 6);
  if (q >= count) return;
  const int lane = threadIdx.x & 63;
  double ar = 0.0, ai = 0.0;
  for (int k = lane; k < nchunk; k += 64) {
    ar += part[S * (static_cast<int64_t>(q) * nchunk + k)];
    if (C) ai += part[S * (static_cast<int64_t>(q) * nchunk + k) + 1];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    ar += __shfl_xor(ar, off, 64);
    if (C) ai += __shfl_xor(ai, off, 64);
  }
  if (lane == 0) {
    out[S * q] = ar;
    if (C) out[S * q + 1] = ai;
  }
}

// W[i, c] -= sum_{j < nv} V_j[i, c] * h[j, c]
template <bool C>
__global__ void k_subtract(const double* __restrict__ V, int64_t vstride, int nv, const double* __restrict__ h,
                           double* W, int64_t n, int nrhs) {
  const int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= n * nrhs) return;
  const int c = static_cast<int>(gid / n);
  const int S = C ? 2 : 1;
  double wr = W[S * gid], wi = C ? W[S * gid + 1] : 0.0;
#pragma unroll 4
  for (int j = 0; j < nv; ++j) {
    const double* v = V + j * vstride + S * gid;
    const double hr = h[S * (j * nrhs + c)];
    if (C) {
      const double hi = h[S * (j * nrhs + c) + 1];
      wr -= v[0] * hr - v[1] * hi;
      wi -= v[0] * hi + v[1] * hr;
    } else {
      wr -= v[0] * hr;
    }
  }
  W[S * gid] = wr;
  if (C) W[S * gid + 1] = wi;
}

// One Gram-Schmidt subtraction of an Arnoldi step for the rows of one chunk of column c:
// h[j, c] = the sum of the chunk sums of <V_j, W> in `hpart` (lane l adds chunks l, l + 64, ... in
// order, then a fixed shuffle tree, as k_dots_final; the chunk-0 workgroup writes h), then
// W -= sum_{j < nv} V_j h[j, c] (as k_subtract) and, with NORM, the chunk sums of ||W[:, c]||^2
// (as k_dots_part of W with itself) into `npart`: one launch instead of two or three, and W read
// once.  `hpart` and `npart` must not overlap (other workgroups still read hpart).
template <bool C, bool NORM>
__global__ __launch_bounds__(256) void k_gs_subtract(const double* __restrict__ V, int64_t vstride, int nv,
                                                      const double* __restrict__ hpart, double* __restrict__ h,
                                                      double* W, int64_t n, int nrhs, int nchunk,
                                                      double* __restrict__ npart) {
  constexpr int S = C ? 2 : 1;
  extern __shared__ double hs[];  // nv coefficients (S doubles each)
  __shared__ double red[4];
  const int chunk = blockIdx.x, c = blockIdx.y;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int j = wave; j < nv; j += 4) {
    const int64_t q = static_cast<int64_t>(j) * nrhs + c;
    double ar = 0.0, ai = 0.0;
    for (int k = lane; k < nchunk; k += 64) {
      ar += hpart[S * (q * nchunk + k)];
      if (C) ai += hpart[S * (q * nchunk + k) + 1];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      ar += __shfl_xor(ar, off, 64);
      if (C) ai += __shfl_xor(ai, off, 64);
    }
    if (lane == 0) {
      hs[S * j] = ar;
      if (C) hs[S * j + 1] = ai;
      if (chunk == 0) {
        h[S * q] = ar;
        if (C) h[S * q + 1] = ai;
      }
    }
  }
  __syncthreads();
  const int64_t per = (n + nchunk - 1) / nchunk;
  const int64_t r0 = chunk * per, r1 = r0 + per < n ? r0 + per : n;
  double nr = 0.0;
  for (int64_t i = r0 + threadIdx.x; i < r1; i += blockDim.x) {
    const int64_t gid = static_cast<int64_t>(c) * n + i;
    double wr = W[S * gid], wi = C ? W[S * gid + 1] : 0.0;
#pragma unroll 4
    for (int j = 0; j < nv; ++j) {
      const double* v = V + j * vstride + S * gid;
      const double hr = hs[S * j];
      if (C) {
        const double hi = hs[S * j + 1];
        wr -= v[0] * hr - v[1] * hi;
        wi -= v[0] * hi + v[1] * hr;
      } else {
        wr -= v[0] * hr;
      }
    }
    W[S * gid] = wr;
    if (C) W[S * gid + 1] = wi;
    if (NORM) {
      if (C)
        nr += wr * wr + wi * wi;
      else
        nr += wr * wr;
    }
  }
  if constexpr (NORM) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) nr += __shfl_xor(nr, off, 64);
    if (lane == 0) red[wave] = nr;
    __syncthreads();
    if (threadIdx.x == 0) {
      const int64_t o = static_cast<int64_t>(c) * nchunk + chunk;
      npart[S * o] = (red[0] + red[1]) + (red[2] + red[3]);
      if (C) npart[S * o + 1] = 0.0;
    }
  }
}

// Y[i, c] = a_c * X[i, c] + b_c * Y[i, c] with per-column complex coefficients (or real)
template <bool C>
__global__ void k_axpby(const double* __restrict__ a, const double* __restrict__ X, const double* __restrict__ b,
                        double* Y, int64_t n, int nrhs) {
  const int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= n * nrhs) return;
  const int c = static_cast<int>(gid / n);
  // b = 0 never reads Y (a new Krylov basis vector lands in memory nothing has written yet)
  if (C) {
    const double ar = a[2 * c], ai = a[2 * c + 1], br = b[2 * c], bi = b[2 * c + 1];
    const double xr = X[2 * gid], xi = X[2 * gid + 1];
    const bool rd = br != 0.0 || bi != 0.0;
    const double yr = rd ? Y[2 * gid] : 0.0, yi = rd ? Y[2 * gid + 1] : 0.0;
    Y[2 * gid] = ar * xr - ai * xi + br * yr - bi * yi;
    Y[2 * gid + 1] = ar * xi + ai * xr + br * yi + bi * yr;
  } else {
    const double bc = b[c];
    Y[gid] = a[c] * X[gid] + (bc != 0.0 ? bc * Y[gid] : 0.0);
  }
}

// X[i, c] += sum_j V_j[i, c] y[j, c]
template <bool C>
__global__ void k_combine(const double* __restrict__ V, int64_t vstride, int nv, const double* __restrict__ y,
                          double* X, int64_t n, int nrhs) {
  const int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= n * nrhs) return;
  const int c = static_cast<int>(gid / n);
  const int S = C ? 2 : 1;
  double xr = X[S * gid], xi = C ? X[S * gid + 1] : 0.0;
  for (int j = 0; j < nv; ++j) {
    const double* v = V + j * vstride + S * gid;
    const double yr = y[S * (j * nrhs + c)];
    if (C) {
      const double yi = y[S * (j * nrhs + c) + 1];
      xr += v[0] * yr - v[1] * yi;
      xi += v[0] * yi + v[1] * yr;
    } else {
      xr += v[0] * yr;
    }
  }
  X[S * gid] = xr;
  if (C) X[S * gid + 1] = xi;
}

static inline unsigned nb(int64_t n) { return static_cast<unsigned>((n + 255) / 256); }

struct Vecs {
  bool C;
  int64_t n;
  int nrhs;
  hipStream_t st;
  DevBuf<double> part;   // chunk sums of the inner products
  DevBuf<double> npart;  // chunk sums of ||w||^2 of an Arnoldi step (read by the Givens kernel)
  int S() const { return C ? 2 : 1; }
  int64_t block() const { return static_cast<int64_t>(S()) * n * nrhs; }
  void dots(const double* V, int nv, const double* W, double* d_out, std::vector<double>& host) {
    dots_async(V, nv, W, d_out);
    host.resize(static_cast<size_t>(S()) * nv * nrhs);
    BMI_CHECK_HIP(hipMemcpyAsync(host.data(), d_out, host.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    BMI_CHECK_HIP(hipStreamSynchronize(st));
  }
  void dots_async(const double* V, int nv, const double* W, double* d_out) {
    const int nchunk = dot_chunks(n, nrhs);
    const size_t need = static_cast<size_t>(S()) * nv * nrhs * nchunk;
    if (part.n < need) {
      BMI_CHECK_HIP(hipStreamSynchronize(st));  // earlier launches may still use the old buffer
      part.alloc(std::max(need, std::max(2 * part.n, static_cast<size_t>(S()) * 32 * nrhs * nchunk)));
    }
    const dim3 grid(static_cast<unsigned>(nchunk), static_cast<unsigned>(nrhs),
                    static_cast<unsigned>((nv + DOT_JB - 1) / DOT_JB));
    const int count = nv * nrhs;
    if (C) {
      hipLaunchKernelGGL(k_dots_part<true>, grid, dim3(256), 0, st, V, block(), nv, W, n, nrhs, nchunk, part.ptr);
      hipLaunchKernelGGL(k_dots_final<true>, dim3((count + 3) / 4), dim3(256), 0, st, part.ptr, count, nchunk,
                         d_out);
    } else {
      hipLaunchKernelGGL(k_dots_part<false>, grid, dim3(256), 0, st, V, block(), nv, W, n, nrhs, nchunk, part.ptr);
      hipLaunchKernelGGL(k_dots_final<false>, dim3((count + 3) / 4), dim3(256), 0, st, part.ptr, count, nchunk,
                         d_out);
    }
    BMI_CHECK_HIP(hipGetLastError());
  }
  void subtract(const double* V, int nv, const double* d_h, double* W) const {
    if (C)
      hipLaunchKernelGGL(k_subtract<true>, dim3(nb(n * nrhs)), dim3(256), 0, st, V, block(), nv, d_h, W, n, nrhs);
    else
      hipLaunchKernelGGL(k_subtract<false>, dim3(nb(n * nrhs)), dim3(256), 0, st, V, block(), nv, d_h, W, n, nrhs);
    BMI_CHECK_HIP(hipGetLastError());
  }
  // chunk sums of <V_j, W> into `part` (k_dots_part alone); returns the chunk count
  int dots_part_async(const double* V, int nv, const double* W) {
    const int nchunk = dot_chunks(n, nrhs);
    const size_t need = static_cast<size_t>(S()) * nv * nrhs * nchunk;
    if (part.n < need) {
      BMI_CHECK_HIP(hipStreamSynchronize(st));  // earlier launches may still use the old buffer
      part.alloc(std::max(need, std::max(2 * part.n, static_cast<size_t>(S()) * 32 * nrhs * nchunk)));
    }
    const dim3 grid(static_cast<unsigned>(nchunk), static_cast<unsigned>(nrhs),
                    static_cast<unsigned>((nv + DOT_JB - 1) / DOT_JB));
    if (C)
      hipLaunchKernelGGL(k_dots_part<true>, grid, dim3(256), 0, st, V, block(), nv, W, n, nrhs, nchunk, part.ptr);
    else
      hipLaunchKernelGGL(k_dots_part<false>, grid, dim3(256), 0, st, V, block(), nv, W, n, nrhs, nchunk, part.ptr);
    BMI_CHECK_HIP(hipGetLastError());
    return nchunk;
  }
  // h = the inner products whose chunk sums dots_part_async left in `part`, then W -= V h; with
  // norm, the chunk sums of ||W||^2 per column into `npart` (returned chunk count, read by the
  // Givens kernel)
  int gs_subtract(const double* V, int nv, double* d_h, double* W, bool norm) {
    const int nchunk = dot_chunks(n, nrhs);
    const size_t need = static_cast<size_t>(S()) * nrhs * nchunk;
    if (norm && npart.n < need) {
      BMI_CHECK_HIP(hipStreamSynchronize(st));
      npart.alloc(need);
    }
    const dim3 grid(static_cast<unsigned>(nchunk), static_cast<unsigned>(nrhs));
    const size_t lds = static_cast<size_t>(S()) * nv * sizeof(double);
    if (C && norm)
      hipLaunchKernelGGL((k_gs_subtract<true, true>), grid, dim3(256), lds, st, V, block(), nv, part.ptr, d_h, W, n,
                         nrhs, nchunk, npart.ptr);
    else if (C)
      hipLaunchKernelGGL((k_gs_subtract<true, false>), grid, dim3(256), lds, st, V, block(), nv, part.ptr, d_h, W,
                         n, nrhs, nchunk, nullptr);
    else if (norm)
      hipLaunchKernelGGL((k_gs_subtract<false, true>), grid, dim3(256), lds, st, V, block(), nv, part.ptr, d_h, W,
                         n, nrhs, nchunk, npart.ptr);
    else
      hipLaunchKernelGGL((k_gs_subtract<false, false>), grid, dim3(256), lds, st, V, block(), nv, part.ptr, d_h, W,
                         n, nrhs, nchunk, nullptr);
    BMI_CHECK_HIP(hipGetLastError());
    return nchunk;
  }
  void axpby(const double* d_a, const double* X, const double* d_b, double* Y) const {
    if (C)
      hipLaunchKernelGGL(k_axpby<true>, dim3(nb(n * nrhs)), dim3(256), 0, st, d_a, X, d_b, Y, n, nrhs);
    else
      hipLaunchKernelGGL(k_axpby<false>, dim3(nb(n * nrhs)), dim3(256), 0, st, d_a, X, d_b, Y, n, nrhs);
    BMI_CHECK_HIP(hipGetLastError());
  }
  void combine(const double* V, int nv, const double* d_y, double* X) const {
    if (C)
      hipLaunchKernelGGL(k_combine<true>, dim3(nb(n * nrhs)), dim3(256), 0, st, V, block(), nv, d_y, X, n, nrhs);
    else
      hipLaunchKernelGGL(k_combine<false>, dim3(nb(n * nrhs)), dim3(256), 0, st, V, block(), nv, d_y, X, n, nrhs);
    BMI_CHECK_HIP(hipGetLastError());
  }
};

// ------------------------------------------------------------------ GMRES state on the device
// Per column c (complex entries as S doubles): Hessenberg H[c] ((m+1) x m, column-major), the
// rotated right-hand side g[c] (m+1), rotations cs[c], sn[c] (m), the active flag, the number of
// Arnoldi columns kc[c] of this cycle, the step count and the residual estimate.  The host reads
// the flags back only every few steps, so an Arnoldi step costs no host round trip.
struct GmresState {
  double *H, *g, *cs, *sn, *bn, *resid, *coef;
  int *kc, *iters;
  char* active;
};

__device__ __forceinline__ void cmul(double ar, double ai, double br, double bi, double& r, double& i) {
  r = ar * br - ai * bi;
  i = ar * bi + ai * br;
}

// start of a cycle: beta = ||R||, g = (beta, 0, ...), V0 = R / beta (coefficients for axpby)
template <bool C>
__global__ void k_gm_cycle(const double* __restrict__ nrm, GmresState s, int m, int nrhs, double tol, int first) {
  constexpr int S = C ? 2 : 1;
  const int c = threadIdx.x + blockIdx.x * blockDim.x;
  if (c >= nrhs) return;
  const double beta = sqrt(fmax(nrm[S * c], 0.0));
  if (first) {
    s.bn[c] = beta > 0 ? beta : 1.0;
    s.iters[c] = 0;
  }
  s.resid[c] = beta / s.bn[c];
  s.active[c] = (s.resid[c] > tol && beta > 0) ? 1 : 0;
  s.kc[c] = 0;
  double* g = s.g + static_cast<int64_t>(c) * (m + 1) * S;
  for (int j = 0; j <= m; ++j) {
    g[S * j] = j == 0 ? beta : 0.0;
    if (C) g[S * j + 1] = 0.0;
  }
  s.coef[S * c] = beta > 0 ? 1.0 / beta : 0.0;
  s.coef[S * (nrhs + c)] = 0.0;
  if (C) s.coef[S * c + 1] = 0.0, s.coef[S * (nrhs + c) + 1] = 0.0;
}

// Arnoldi step k: H(:, k) = h1 + h2 (two Gram-Schmidt passes), H(k+1, k) = ||w||, previous
// rotations, a new rotation, g and the residual estimate |g(k+1)| / ||b||; coefficients 1/||w||
// for the next basis vector.  One wave per column: ||w||^2 is the sum of the column's chunk sums
// in `part` (lane l adds chunks l, l + 64, ... in order, then a fixed shuffle tree, as
// k_dots_final), the lanes stage h1 + h2 and the earlier rotations in LDS together, and lane 0
// applies the rotations in order (each rotation needs the previous one's result).
template <bool C>
__global__ __launch_bounds__(64) void k_gm_givens(const double* __restrict__ h1, const double* __restrict__ h2,
                                                  const double* __restrict__ part, int nchunk, GmresState s, int k,
                                                  int m, int nrhs, double tol) {
  constexpr int S = C ? 2 : 1;
  extern __shared__ double stage[];  // (k + 1) column entries, k cs, k sn (S doubles each)
  const bool staged = S * (3 * k + 1) <= GIVENS_STAGE_DOUBLES;
  const int c = blockIdx.x;
  const int lane = threadIdx.x;
  double nr2 = 0.0;
  for (int q = lane; q < nchunk; q += 64) nr2 += part[S * (static_cast<int64_t>(c) * nchunk + q)];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) nr2 += __shfl_xor(nr2, off, 64);
  const bool act = s.active[c] != 0;
  double* hst = stage;
  double* cst = stage + S * (k + 1);
  double* snt = cst + S * k;
  const double* __restrict__ cs0 = s.cs + static_cast<int64_t>(c) * m * S;
  const double* __restrict__ sn0 = s.sn + static_cast<int64_t>(c) * m * S;
  if (act && staged) {
    for (int j = lane; j <= k; j += 64) {
      const int64_t q = S * (static_cast<int64_t>(j) * nrhs + c);
      hst[S * j] = h1[q] + h2[q];
      if (C) hst[S * j + 1] = h1[q + 1] + h2[q + 1];
    }
    for (int e = lane; e < S * k; e += 64) cst[e] = cs0[e], snt[e] = sn0[e];
  }
  __syncthreads();
  if (lane != 0) return;
  const double hn = sqrt(fmax(nr2, 0.0));
  s.coef[S * c] = hn > 0 ? 1.0 / hn : 0.0;
  s.coef[S * (nrhs + c)] = 0.0;
  if (C) s.coef[S * c + 1] = 0.0, s.coef[S * (nrhs + c) + 1] = 0.0;
  if (!act) return;
  double* __restrict__ col = s.H + (static_cast<int64_t>(c) * (m + 1) * m + static_cast<int64_t>(k) * (m + 1)) * S;
  // new Hessenberg column h = h1 + h2 with the earlier rotations applied in order; entry j + 1 is
  // carried from rotation j to rotation j + 1
  auto hcol = [&](int j, double& r, double& i) {
    if (staged) {
      r = hst[S * j], i = C ? hst[S * j + 1] : 0.0;
    } else {
      const int64_t q = S * (static_cast<int64_t>(j) * nrhs + c);
      r = h1[q] + h2[q];
      i = C ? h1[q + 1] + h2[q + 1] : 0.0;
    }
  };
  const double* __restrict__ csr = staged ? cst : cs0;
  const double* __restrict__ snr = staged ? snt : sn0;
  double t1r, t1i;
  hcol(0, t1r, t1i);
  for (int j = 0; j < k; ++j) {
    double t2r, t2i;
    hcol(j + 1, t2r, t2i);
    const double cr = csr[S * j], ci = C ? csr[S * j + 1] : 0.0, sr = snr[S * j], si = C ? snr[S * j + 1] : 0.0;
    double ar, ai, br, bi;
    cmul(cr, -ci, t1r, t1i, ar, ai);  // conj(cs) t1 + conj(sn) t2
    cmul(sr, -si, t2r, t2i, br, bi);
    col[S * j] = ar + br;
    if (C) col[S * j + 1] = ai + bi;
    cmul(-sr, -si, t1r, t1i, ar, ai);  // -sn t1 + cs t2
    cmul(cr, ci, t2r, t2i, br, bi);
    t1r = ar + br;
    t1i = ai + bi;
  }
  double* cs = s.cs + static_cast<int64_t>(c) * m * S;
  double* sn = s.sn + static_cast<int64_t>(c) * m * S;
  double* g = s.g + static_cast<int64_t>(c) * (m + 1) * S;
  const double x1r = t1r, x1i = C ? t1i : 0.0;
  const double x2r = hn, x2i = 0.0;
  const double den = sqrt((x1r * x1r + x1i * x1i) + (x2r * x2r + x2i * x2i));
  const double cr = den > 0 ? x1r / den : 1.0, ci = den > 0 ? x1i / den : 0.0;
  const double sr = den > 0 ? x2r / den : 0.0, si = den > 0 ? x2i / den : 0.0;
  cs[S * k] = cr, sn[S * k] = sr;
  if (C) cs[S * k + 1] = ci, sn[S * k + 1] = si;
  col[S * k] = den;
  col[S * (k + 1)] = 0.0;
  if (C) col[S * k + 1] = 0.0, col[S * (k + 1) + 1] = 0.0;
  const double gr = g[S * k], gi = C ? g[S * k + 1] : 0.0;
  double nr, ni, kr, ki;
  cmul(-sr, -si, gr, gi, nr, ni);  // g(k+1) = -sn g(k)
  cmul(cr, -ci, gr, gi, kr, ki);   // g(k) = conj(cs) g(k)
  g[S * (k + 1)] = nr;
  g[S * k] = kr;
  if (C) g[S * (k + 1) + 1] = ni, g[S * k + 1] = ki;
  s.kc[c] = k + 1;
  s.iters[c] += 1;
  s.resid[c] = hypot(nr, ni) / s.bn[c];
  if (s.resid[c] <= tol || hn == 0.0) s.active[c] = 0;
}

// end of a cycle: y from the triangular system R y = g (kc[c] unknowns), zero beyond
template <bool C>
__global__ void k_gm_backsolve(GmresState s, int m, int nv, int nrhs, double* y) {
  constexpr int S = C ? 2 : 1;
  const int c = threadIdx.x + blockIdx.x * blockDim.x;
  if (c >= nrhs) return;
  const int kk = s.kc[c];
  const double* H = s.H + static_cast<int64_t>(c) * (m + 1) * m * S;
  const double* g = s.g + static_cast<int64_t>(c) * (m + 1) * S;
  for (int j = kk; j < nv; ++j) {
    y[S * (static_cast<int64_t>(j) * nrhs + c)] = 0.0;
    if (C) y[S * (static_cast<int64_t>(j) * nrhs + c) + 1] = 0.0;
  }
  for (int i = kk - 1; i >= 0; --i) {
    double ar = g[S * i], ai = C ? g[S * i + 1] : 0.0;
    for (int j = i + 1; j < kk; ++j) {
      const double* hij = H + S * (static_cast<int64_t>(j) * (m + 1) + i);
      const double* yj = y + S * (static_cast<int64_t>(j) * nrhs + c);
      double pr, pi;
      cmul(hij[0], C ? hij[1] : 0.0, yj[0], C ? yj[1] : 0.0, pr, pi);
      ar -= pr;
      ai -= pi;
    }
    const double d = H[S * (static_cast<int64_t>(i) * (m + 1) + i)];  // real after the rotation
    y[S * (static_cast<int64_t>(i) * nrhs + c)] = ar / d;
    if (C) y[S * (static_cast<int64_t>(i) * nrhs + c) + 1] = ai / d;
  }
}

// Batched restarted GMRES (classical Gram-Schmidt applied twice, FP64 Arnoldi) as a resumable
// task: poll() enqueues the next piece of work on the task's stream once the convergence flags
// of the previous piece have arrived (an asynchronous copy into pinned host memory, then an
// event), so the calling thread never waits on the device and tasks on different streams
// overlap there. Pieces: a cycle start (residual R = B - A X -- with res64 the FP64 operator,
// each cycle then being one refinement step whose correction comes from an Arnoldi basis of the
// reduced operator -- its norm and V0 = R / beta), then Arnoldi steps up to the next check
// (every 4 steps), then the cycle's end (back substitution, X += V y). A column that converges
// inside a window takes no further steps (its flag is cleared on the device), the others go on.
// gmres() drives one task to the end; the device work is the same either way.
class GmresRun {
 public:
  GmresRun(bmi_h2* h, int nrhs, const double* B, double* X, double tol, int restart, int maxit, hipStream_t st,
           bool reduced, bool res64)
      : h_(h), nrhs_(nrhs), B_(B), X_(X), tol_(tol), maxit_(maxit), st_(st), reduced_(reduced), res64_(res64),
        C_(h2_complex(h)), n_(h2_rows(h)), vs_{C_, n_, nrhs, st}, S_(vs_.S()), m_(std::max(1, std::min(restart, maxit))) {
    V_.alloc(static_cast<size_t>(vs_.block()) * (m_ + 1));
    W_.alloc(vs_.block());
    R_.alloc(vs_.block());
    d_h1_.alloc(static_cast<size_t>(S_) * (m_ + 1) * nrhs);
    d_h2_.alloc(static_cast<size_t>(S_) * (m_ + 1) * nrhs);
    d_nrm_.alloc(static_cast<size_t>(S_) * nrhs);
    d_H_.alloc(static_cast<size_t>(S_) * (m_ + 1) * m_ * nrhs);
    d_g_.alloc(static_cast<size_t>(S_) * (m_ + 1) * nrhs);
    d_cs_.alloc(static_cast<size_t>(S_) * m_ * nrhs);
    d_sn_.alloc(static_cast<size_t>(S_) * m_ * nrhs);
    d_bn_.alloc(nrhs);
    d_resid_.alloc(nrhs);
    d_coef_.alloc(static_cast<size_t>(2) * S_ * nrhs);
    d_y_.alloc(static_cast<size_t>(S_) * m_ * nrhs);
    d_kc_.alloc(nrhs);
    d_iters_.alloc(nrhs);
    d_active_.alloc(nrhs);
    gs_ = GmresState{d_H_.ptr, d_g_.ptr, d_cs_.ptr, d_sn_.ptr, d_bn_.ptr, d_resid_.ptr, d_coef_.ptr,
                     d_kc_.ptr, d_iters_.ptr, d_active_.ptr};
    sb_ = static_cast<unsigned>((nrhs + 63) / 64);
    BMI_CHECK_HIP(hipHostMalloc(reinterpret_cast<void**>(&act_), nrhs, hipHostMallocDefault));
    const hipError_t e = hipEventCreateWithFlags(&ev_, hipEventDisableTiming);
    if (e != hipSuccess) {
      (void)hipHostFree(act_);
      act_ = nullptr;
      BMI_CHECK_HIP(e);
    }
    try {
      // ||B|| per column
      vs_.dots_async(B_, 1, B_, d_nrm_.ptr);
      cycle_kernel(1);
    } catch (...) {
      release();
      throw;
    }
  }
  ~GmresRun() { release(); }
  void release() {
    if (ev_) {
      (void)hipEventSynchronize(ev_);
      (void)hipEventDestroy(ev_);
    }
    ev_ = nullptr;
    (void)hipStreamSynchronize(st_);  // the buffers may still be in use
    if (act_) (void)hipHostFree(act_);
    act_ = nullptr;
  }
  GmresRun(const GmresRun&) = delete;
  GmresRun& operator=(const GmresRun&) = delete;

  // enqueue the next piece if the flags it depends on have arrived: 0 flags not there yet (nothing
  // enqueued), 1 finished, 2 work enqueued
  int poll() {
    while (true) {
      if (phase_ == DONE) return 1;
      if (pending_) {
        const hipError_t q = hipEventQuery(ev_);
        if (q == hipErrorNotReady) return 0;
        BMI_CHECK_HIP(q);
        pending_ = false;
      }
      switch (phase_) {
        case CYCLE:
          start_cycle();
          return 2;
        case AFTER_CYCLE:
          if (!any_active()) {
            converged_ = true;
            phase_ = DONE;
            break;
          }
          if (total_ >= maxit_) {
            phase_ = DONE;
            break;
          }
          vs_.axpby(d_coef_.ptr, R_.ptr, d_coef_.ptr + S_ * nrhs_, V_.ptr);
          k_ = 0;
          steps();
          if (phase_ == CYCLE) continue;  // the cycle ended without a check: start the next one
          return 2;
        case AFTER_STEPS:
          if (!any_active()) {
            end_cycle();
            continue;
          }
          steps();
          if (phase_ == CYCLE) continue;
          return 2;
        default:
          return 1;
      }
    }
  }
  // block until the flags of the pending piece have arrived
  void wait() {
    if (pending_) BMI_CHECK_HIP(hipEventSynchronize(ev_));
  }
  bool converged() const { return converged_; }
  void results(std::vector<int>& iters, std::vector<double>& resid) {
    iters.assign(nrhs_, 0);
    resid.assign(nrhs_, 0.0);
    BMI_CHECK_HIP(hipMemcpyAsync(iters.data(), d_iters_.ptr, nrhs_ * sizeof(int), hipMemcpyDeviceToHost, st_));
    BMI_CHECK_HIP(hipMemcpyAsync(resid.data(), d_resid_.ptr, nrhs_ * sizeof(double), hipMemcpyDeviceToHost, st_));
    BMI_CHECK_HIP(hipStreamSynchronize(st_));
    // a NaN or infinite residual (the device flags treat it as inactive) never counts as converged
    for (const double r : resid)
      if (!std::isfinite(r)) converged_ = false;
  }

 private:
  enum Phase { CYCLE, AFTER_CYCLE, AFTER_STEPS, DONE };
  static constexpr int CHECK = 4;  // Arnoldi steps between convergence checks

  void cycle_kernel(int first) {
    if (C_)
      hipLaunchKernelGGL(k_gm_cycle<true>, dim3(sb_), dim3(64), 0, st_, d_nrm_.ptr, gs_, m_, nrhs_, tol_, first);
    else
      hipLaunchKernelGGL(k_gm_cycle<false>, dim3(sb_), dim3(64), 0, st_, d_nrm_.ptr, gs_, m_, nrhs_, tol_, first);
    BMI_CHECK_HIP(hipGetLastError());
  }
  void request_flags() {
    BMI_CHECK_HIP(hipMemcpyAsync(act_, d_active_.ptr, nrhs_, hipMemcpyDeviceToHost, st_));
    BMI_CHECK_HIP(hipEventRecord(ev_, st_));
    pending_ = true;
  }
  bool any_active() const {
    for (int c = 0; c < nrhs_; ++c)
      if (act_[c]) return true;
    return false;
  }
  void start_cycle() {
    const double one[2] = {1.0, 0.0}, mone[2] = {-1.0, 0.0};
    BMI_CHECK_HIP(hipMemcpyAsync(R_.ptr, B_, vs_.block() * sizeof(double), hipMemcpyDeviceToDevice, st_));
    h2_apply_device(h_, nrhs_, X_, n_, R_.ptr, n_, mone, one, st_, res64_ ? false : reduced_);
    vs_.dots_async(R_.ptr, 1, R_.ptr, d_nrm_.ptr);
    cycle_kernel(0);
    request_flags();
    phase_ = AFTER_CYCLE;
  }
  // Arnoldi steps up to the next check, then the flags of that check and, queued behind them, the
  // steps up to the check after (so the device keeps working while the host waits for the
  // flags); or the end of the cycle.  A queued batch that turns out to follow convergence changes
  // nothing: the Givens kernel leaves inactive columns as they are (kc, H, g, the step count),
  // and the new basis vectors are unit vectors or zero, weighted by zero at the end of the cycle.
  void steps() {
    if (!ahead_) batch();
    ahead_ = false;
    if (k_ < m_ && total_ < maxit_) {
      request_flags();
      phase_ = AFTER_STEPS;
      if (k_ < m_ && total_ < maxit_) {
        batch();
        ahead_ = true;
      }
    } else {
      end_cycle();
    }
  }
  // Arnoldi steps from k_ up to the next multiple of CHECK (or the end of the cycle)
  void batch() {
    const double one[2] = {1.0, 0.0};
    do {
      double* Vk = V_.ptr + vs_.block() * k_;
      h2_apply_device(h_, nrhs_, Vk, n_, W_.ptr, n_, one, nullptr, st_, reduced_);
      int nchunk;
      const double* npart;
      if (S_ * (k_ + 1) <= GS_LDS_DOUBLES) {
        // the subtractions reduce the inner products themselves (coefficients in LDS)
        vs_.dots_part_async(V_.ptr, k_ + 1, W_.ptr);
        vs_.gs_subtract(V_.ptr, k_ + 1, d_h1_.ptr, W_.ptr, false);
        vs_.dots_part_async(V_.ptr, k_ + 1, W_.ptr);
        nchunk = vs_.gs_subtract(V_.ptr, k_ + 1, d_h2_.ptr, W_.ptr, true);
        npart = vs_.npart.ptr;
      } else {
        // long cycles: separate reductions, coefficients from memory
        vs_.dots_async(V_.ptr, k_ + 1, W_.ptr, d_h1_.ptr);
        vs_.subtract(V_.ptr, k_ + 1, d_h1_.ptr, W_.ptr);
        vs_.dots_async(V_.ptr, k_ + 1, W_.ptr, d_h2_.ptr);
        vs_.subtract(V_.ptr, k_ + 1, d_h2_.ptr, W_.ptr);
        nchunk = vs_.dots_part_async(W_.ptr, 1, W_.ptr);  // chunk sums of ||w||^2, column-major
        npart = vs_.part.ptr;
      }
      const int stage = S_ * (3 * k_ + 1);
      const size_t lds = stage <= GIVENS_STAGE_DOUBLES ? stage * sizeof(double) : 0;
      if (C_)
        hipLaunchKernelGGL(k_gm_givens<true>, dim3(nrhs_), dim3(64), lds, st_, d_h1_.ptr, d_h2_.ptr, npart,
                           nchunk, gs_, k_, m_, nrhs_, tol_);
      else
        hipLaunchKernelGGL(k_gm_givens<false>, dim3(nrhs_), dim3(64), lds, st_, d_h1_.ptr, d_h2_.ptr, npart,
                           nchunk, gs_, k_, m_, nrhs_, tol_);
      BMI_CHECK_HIP(hipGetLastError());
      vs_.axpby(d_coef_.ptr, W_.ptr, d_coef_.ptr + S_ * nrhs_, V_.ptr + vs_.block() * (k_ + 1));
      ++k_;
      ++total_;
    } while (k_ < m_ && total_ < maxit_ && k_ % CHECK != 0);
  }
  // X += V y, then the next cycle
  void end_cycle() {
    ahead_ = false;
    if (k_ == 0) {
      phase_ = DONE;
      return;
    }
    if (C_)
      hipLaunchKernelGGL(k_gm_backsolve<true>, dim3(sb_), dim3(64), 0, st_, gs_, m_, k_, nrhs_, d_y_.ptr);
    else
      hipLaunchKernelGGL(k_gm_backsolve<false>, dim3(sb_), dim3(64), 0, st_, gs_, m_, k_, nrhs_, d_y_.ptr);
    BMI_CHECK_HIP(hipGetLastError());
    vs_.combine(V_.ptr, k_, d_y_.ptr, X_);
    phase_ = CYCLE;
  }

  bmi_h2* h_;
  int nrhs_;
  const double* B_;
  double* X_;
  double tol_;
  int maxit_;
  hipStream_t st_;
  bool reduced_, res64_, C_;
  int64_t n_;
  Vecs vs_;
  int S_, m_;
  DevBuf<double> V_, W_, R_, d_h1_, d_h2_, d_nrm_, d_H_, d_g_, d_cs_, d_sn_, d_bn_, d_resid_, d_coef_, d_y_;
  DevBuf<int> d_kc_, d_iters_;
  DevBuf<char> d_active_;
  GmresState gs_{};
  unsigned sb_ = 1;
  char* act_ = nullptr;
  hipEvent_t ev_ = nullptr;
  Phase phase_ = CYCLE;
  bool pending_ = false, converged_ = false;
  bool ahead_ = false;  // a batch of steps is queued past the pending check
  int k_ = 0, total_ = 0;
};

// returns true when every column converged
static bool gmres(bmi_h2* h, int nrhs, const double* B, double* X, double tol, int restart, int maxit,
                  std::vector<int>& iters, std::vector<double>& resid, hipStream_t st, bool reduced,
                  bool res64 = false) {
  GmresRun run(h, nrhs, B, X, tol, restart, maxit, st, reduced, res64);
  while (run.poll() != 1) run.wait();
  run.results(iters, resid);
  return run.converged();
}

// CG scalars on the device: no host round trip per iteration
// alpha_c = rr_c / pAp_c (active columns); X += alpha P; R -= alpha AP
__global__ void k_cg_xr(const double* __restrict__ rr, const double* __restrict__ pap, const char* __restrict__ active,
                        const double* __restrict__ P, const double* __restrict__ AP, double* X, double* R, int64_t n,
                        int nrhs) {
  const int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= n * nrhs) return;
  const int c = static_cast<int>(gid / n);
  if (!active[c] || pap[c] == 0.0) return;
  const double a = rr[c] / pap[c];
  X[gid] += a * P[gid];
  R[gid] -= a * AP[gid];
}
// per column: count the step, beta = rr_new / rr, deactivate converged columns
__global__ void k_cg_scalars(const double* __restrict__ rr, double* __restrict__ rrn, const double* __restrict__ bn,
                             char* active, int* iters, double* beta, double tol, int nrhs) {
  const int c = threadIdx.x + blockIdx.x * blockDim.x;
  if (c >= nrhs) return;
  if (!active[c]) {
    beta[c] = 0.0;
    rrn[c] = rr[c];
    return;
  }
  iters[c] += 1;
  beta[c] = rr[c] != 0.0 ? rrn[c] / rr[c] : 0.0;
  if (sqrt(fmax(rrn[c], 0.0)) <= tol * bn[c]) active[c] = 0;
}
// P = R + beta P (active columns)
__global__ void k_cg_p(const double* __restrict__ beta, const char* __restrict__ active, const double* __restrict__ R,
                       double* P, int64_t n, int nrhs) {
  const int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= n * nrhs) return;
  const int c = static_cast<int>(gid / n);
  if (!active[c]) return;
  P[gid] = R[gid] + beta[c] * P[gid];
}
__global__ void k_cg_init(const double* __restrict__ rr, const double* __restrict__ bn, char* active, int* iters,
                          double tol, int nrhs) {
  const int c = threadIdx.x + blockIdx.x * blockDim.x;
  if (c >= nrhs) return;
  active[c] = sqrt(fmax(rr[c], 0.0)) > tol * bn[c] ? 1 : 0;
  iters[c] = 0;
}

// batched conjugate gradients (real symmetric operators); convergence is read back every
// `check` iterations
static bool cg(bmi_h2* h, int nrhs, const double* B, double* X, double tol, int maxit, std::vector<int>& iters,
               std::vector<double>& resid, hipStream_t st, bool reduced) {
  const int64_t n = h2_rows(h);
  Vecs vs{false, n, nrhs, st};
  DevBuf<double> R, Pv, AP, d_rr[2], d_pap, d_bn, d_beta;
  DevBuf<char> d_active;
  DevBuf<int> d_iters;
  R.alloc(vs.block());
  Pv.alloc(vs.block());
  AP.alloc(vs.block());
  d_rr[0].alloc(nrhs);
  d_rr[1].alloc(nrhs);
  d_pap.alloc(nrhs);
  d_bn.alloc(nrhs);
  d_beta.alloc(nrhs);
  d_active.alloc(nrhs);
  d_iters.alloc(nrhs);
  const double one[2] = {1.0, 0.0}, mone[2] = {-1.0, 0.0};
  std::vector<double> buf;
  vs.dots(B, 1, B, d_pap.ptr, buf);
  std::vector<double> bn(nrhs);
  for (int c = 0; c < nrhs; ++c) bn[c] = std::sqrt(buf[c]) > 0 ? std::sqrt(buf[c]) : 1.0;
  d_bn.upload(bn, st);
  BMI_CHECK_HIP(hipMemcpyAsync(R.ptr, B, vs.block() * sizeof(double), hipMemcpyDeviceToDevice, st));
  h2_apply_device(h, nrhs, X, n, R.ptr, n, mone, one, st, reduced);
  BMI_CHECK_HIP(hipMemcpyAsync(Pv.ptr, R.ptr, vs.block() * sizeof(double), hipMemcpyDeviceToDevice, st));
  vs.dots_async(R.ptr, 1, R.ptr, d_rr[0].ptr);
  const unsigned sb = static_cast<unsigned>((nrhs + 63) / 64);
  hipLaunchKernelGGL(k_cg_init, dim3(sb), dim3(64), 0, st, d_rr[0].ptr, d_bn.ptr, d_active.ptr, d_iters.ptr, tol, nrhs);
  BMI_CHECK_HIP(hipGetLastError());
  const int check = 8;
  int cur = 0;
  std::vector<char> act(nrhs);
  const unsigned eb = nb(n * nrhs);
  for (int it = 0; it < maxit; ++it) {
    if (it % check == 0) {
      BMI_CHECK_HIP(hipMemcpyAsync(act.data(), d_active.ptr, nrhs, hipMemcpyDeviceToHost, st));
      BMI_CHECK_HIP(hipStreamSynchronize(st));
      if (std::none_of(act.begin(), act.end(), [](char a) { return a != 0; })) break;
    }
    h2_apply_device(h, nrhs, Pv.ptr, n, AP.ptr, n, one, nullptr, st, reduced);
    vs.dots_async(Pv.ptr, 1, AP.ptr, d_pap.ptr);
    hipLaunchKernelGGL(k_cg_xr, dim3(eb), dim3(256), 0, st, d_rr[cur].ptr, d_pap.ptr, d_active.ptr, Pv.ptr, AP.ptr, X,
                       R.ptr, n, nrhs);
    vs.dots_async(R.ptr, 1, R.ptr, d_rr[1 - cur].ptr);
    hipLaunchKernelGGL(k_cg_scalars, dim3(sb), dim3(64), 0, st, d_rr[cur].ptr, d_rr[1 - cur].ptr, d_bn.ptr,
                       d_active.ptr, d_iters.ptr, d_beta.ptr, tol, nrhs);
    hipLaunchKernelGGL(k_cg_p, dim3(eb), dim3(256), 0, st, d_beta.ptr, d_active.ptr, R.ptr, Pv.ptr, n, nrhs);
    BMI_CHECK_HIP(hipGetLastError());
    cur = 1 - cur;
  }
  std::vector<double> rr(nrhs);
  iters.assign(nrhs, 0);
  BMI_CHECK_HIP(hipMemcpyAsync(rr.data(), d_rr[cur].ptr, nrhs * sizeof(double), hipMemcpyDeviceToHost, st));
  BMI_CHECK_HIP(hipMemcpyAsync(iters.data(), d_iters.ptr, nrhs * sizeof(int), hipMemcpyDeviceToHost, st));
  BMI_CHECK_HIP(hipStreamSynchronize(st));
  resid.assign(nrhs, 0.0);
  bool all = true;
  for (int c = 0; c < nrhs; ++c) {
    resid[c] = std::sqrt(std::max(rr[c], 0.0)) / bn[c];
    all = all && resid[c] <= tol;
  }
  return all;
}

// mixed-precision iterative refinement: residuals with the FP64 operator, corrections from the
// inner solver with the reduced-precision operator
static bool refine(bmi_h2* h, int nrhs, const double* B, double* X, double tol, int restart, int maxit,
                   std::vector<int>& iters, std::vector<double>& resid, hipStream_t st, bool use_cg, int* outer_steps) {
  const bool C = h2_complex(h);
  const int64_t n = h2_rows(h);
  Vecs vs{C, n, nrhs, st};
  const int S = vs.S();
  DevBuf<double> R, D, d_small, d_coef;
  R.alloc(vs.block());
  D.alloc(vs.block());
  d_small.alloc(static_cast<size_t>(S) * nrhs);
  d_coef.alloc(static_cast<size_t>(2) * S * nrhs);
  const double one[2] = {1.0, 0.0}, mone[2] = {-1.0, 0.0};
  std::vector<double> buf;
  vs.dots(B, 1, B, d_small.ptr, buf);
  std::vector<double> bn(nrhs);
  for (int c = 0; c < nrhs; ++c) bn[c] = std::sqrt(std::max(buf[S * c], 0.0)) > 0 ? std::sqrt(buf[S * c]) : 1.0;
  std::vector<double> ones(static_cast<size_t>(2) * S * nrhs, 0.0);
  for (int c = 0; c < nrhs; ++c) ones[S * c] = 1.0, ones[S * (nrhs + c)] = 1.0;
  d_coef.upload(ones, st);
  iters.assign(nrhs, 0);
  resid.assign(nrhs, 0.0);
  int used = 0, outer = 0;
  bool all = false;
  // Inner tolerance: an outer step reduces the residual by about max(inner tolerance, operator
  // perturbation x conditioning); solving the inner system further than the second term buys
  // nothing.  Start at 1e-4; when a step's reduction stayed near its inner tolerance, tighten it
  // tenfold (down to 1e-6), when the reduction fell short of it, use the reduction achieved.
  double tin = 1e-4, last = 0.0;
  std::vector<double> prev(nrhs, 0.0);
  while (true) {
    // R = B - A64 X
    BMI_CHECK_HIP(hipMemcpyAsync(R.ptr, B, vs.block() * sizeof(double), hipMemcpyDeviceToDevice, st));
    h2_apply_device(h, nrhs, X, n, R.ptr, n, mone, one, st, false);
    vs.dots(R.ptr, 1, R.ptr, d_small.ptr, buf);
    all = true;
    bool finite = true;  // a NaN or infinite residual ends the refinement unconverged
    double need = 1.0, ratio = 0.0;
    for (int c = 0; c < nrhs; ++c) {
      const double rn = std::sqrt(std::max(buf[S * c], 0.0));
      resid[c] = rn / bn[c];
      if (!std::isfinite(rn)) {
        all = finite = false;
        continue;
      }
      if (outer > 0 && prev[c] > 0) ratio = std::max(ratio, rn / prev[c]);
      prev[c] = rn;
      if (resid[c] > tol) {
        all = false;
        need = std::min(need, tol * bn[c] / rn);
      }
    }
    if (all || !finite || used >= maxit || outer >= 30) break;
    if (outer > 0) tin = ratio > 3 * last ? std::min(std::max(ratio, 1e-6), 1e-2) : std::max(last * 0.1, 1e-6);
    // inner solve A_reduced D = R to the accuracy this step needs
    const double tstep = std::max(need, tin);
    last = tstep;
    device_fill(D.ptr, vs.block(), 0.0, st);
    std::vector<int> it;
    std::vector<double> rs;
    if (use_cg)
      cg(h, nrhs, R.ptr, D.ptr, tstep, maxit - used, it, rs, st, true);
    else
      gmres(h, nrhs, R.ptr, D.ptr, tstep, restart, maxit - used, it, rs, st, true);
    int mx = 0;
    for (int c = 0; c < nrhs; ++c) iters[c] += it[c], mx = std::max(mx, it[c]);
    used += std::max(mx, 1);
    vs.axpby(d_coef.ptr, D.ptr, d_coef.ptr + S * nrhs, X);  // X = D + X
    ++outer;
  }
  if (outer_steps) *outer_steps = outer;
  return all;
}

// ------------------------------------------------------------------ right-hand sides
// Hertz dipole at pos with moment len, exp(-i kappa r) convention (examples/Data.hpp:76-90)
__device__ __forceinline__ void dipole_field(const double* X, double kappa, const double* pos, const double* len,
                                             double* Er, double* Ei) {
  const double c0 = X[0] - pos[0], c1 = X[1] - pos[1], c2 = X[2] - pos[2];
  const double r = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
  const double n[3] = {c0 / r, c1 / r, c2 / r};
  double sn, cs;
  sincos(kappa * r, &sn, &cs);  // exp(-i kappa r) = cs - i sn
  const double nl = n[0] * len[0] + n[1] * len[1] + n[2] * len[2];
  const double ar = kappa * kappa * cs / r, ai = -kappa * kappa * sn / r;
  // (1/r^3 + i kappa/r^2) (cs - i sn)
  const double p = 1.0 / (r * r * r), q = kappa / (r * r);
  const double br = p * cs + q * sn, bi = q * cs - p * sn;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double tr = len[k] - nl * n[k];   // (n x len) x n
    const double hk = 3 * nl * n[k] - len[k];
    Er[k] = ar * tr + br * hk;
    Ei[k] = ai * tr + bi * hk;
  }
}

// DiscreteLinearForm::compute (LinearForm/DiscreteLinearForm.hpp:44-83) with DirichletTrace
// (fn 0: point source, fn 1: plane waves, fn 3: Helmholtz fundamental solution, fn 4: harmonic
// polynomial) or RotatedTangentialTrace of a dipole field (fn 2, RotatedTangentialTrace.hpp:39-82):
// super-space load vector, one thread per (element, column), (P+2)^2-point rule
template <int PD, int P>
__global__ void k_rhs(const double* __restrict__ cp, int cp_stride, const int* __restrict__ elem_patch,
                      const double* __restrict__ elem_llc, int nel, double h, int fn, const double* __restrict__ prm,
                      int nrhs, double kappa, int64_t nsuper, double* out, const double* __restrict__ vals) {
  constexpr int I = P + 1, I2 = I * I, n = P + 2;
  const int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= static_cast<int64_t>(nel) * nrhs) return;
  const int e = static_cast<int>(gid % nel), k = static_cast<int>(gid / nel);
  const double* gx = c_gl_nodes + c_gl_offset[n - 1];
  const double* gw = c_gl_weights + c_gl_offset[n - 1];
  const double* c = cp + static_cast<int64_t>(elem_patch[e]) * cp_stride;
  double ar[I2], ai[I2], br[I2], bi[I2];
#pragma unroll
  for (int m = 0; m < I2; ++m) ar[m] = 0.0, ai[m] = 0.0, br[m] = 0.0, bi[m] = 0.0;
  for (int q = 0; q < n * n; ++q) {
    const double u = gx[q / n], v = gx[q % n], w = h * (gw[q / n] * gw[q % n]);
    double X[3], Ds[3], Dt[3];
    surface_point<PD>(c, elem_llc[2 * e] + h * u, elem_llc[2 * e + 1] + h * v, X, Ds, Dt);
    const double c0 = Ds[1] * Dt[2] - Ds[2] * Dt[1], c1 = Ds[2] * Dt[0] - Ds[0] * Dt[2], c2 = Ds[0] * Dt[1] - Ds[1] * Dt[0];
    const double nn = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
    double fr, fi = 0.0, gr = 0.0, gi = 0.0;
    if (fn == 0) {
      const double* x0 = prm + 3 * k;
      const double d0 = X[0] - x0[0], d1 = X[1] - x0[1], d2 = X[2] - x0[2];
      fr = nn * w / sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    } else if (fn == 1) {
      const double* d = prm + 3 * k;
      sincos(kappa * (d[0] * X[0] + d[1] * X[1] + d[2] * X[2]), &fi, &fr);
      fr *= nn * w, fi *= nn * w;
    } else if (fn == 3) {
      // Helmholtz fundamental solution exp(-i kappa r) / r (examples/Data.hpp:37-44)
      const double* x0 = prm + 3 * k;
      const double d0 = X[0] - x0[0], d1 = X[1] - x0[1], d2 = X[2] - x0[2];
      const double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
      double sn, cs;
      sincos(kappa * r, &sn, &cs);
      fr = cs * nn * w / r, fi = -sn * nn * w / r;
    } else if (fn == 4) {
      // harmonic polynomial 4x^2 - 3y^2 - z^2 (examples/Data.hpp:14-16)
      fr = (4 * X[0] * X[0] - 3 * X[1] * X[1] - X[2] * X[2]) * nn * w;
    } else if (fn == 5 || fn == 6) {
      // caller-evaluated data at the rule's points: real or complex scalars
      const int64_t vi = (static_cast<int64_t>(k) * nel + e) * (n * n) + q;
      fr = (fn == 5 ? vals[vi] : vals[2 * vi]) * nn * w;
      fi = fn == 5 ? 0.0 : vals[2 * vi + 1] * nn * w;
    } else {
      double Er[3], Ei[3];
      if (fn == 7) {
        // caller-evaluated complex field (3 components, interleaved) at the rule's points
        const double* v = vals + 6 * ((static_cast<int64_t>(k) * nel + e) * (n * n) + q);
        Er[0] = v[0], Ei[0] = v[1], Er[1] = v[2], Ei[1] = v[3], Er[2] = v[4], Ei[2] = v[5];
      } else {
        const double* d = prm + 6 * k;
        dipole_field(X, kappa, d, d + 3, Er, Ei);
      }
      const double nv[3] = {c0 / nn, c1 / nn, c2 / nn};
      const double er = Er[0] * nv[0] + Er[1] * nv[1] + Er[2] * nv[2];
      const double ei = Ei[0] * nv[0] + Ei[1] * nv[1] + Ei[2] * nv[2];
      double tr[3], ti[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) tr[j] = (Er[j] - er * nv[j]) * w, ti[j] = (Ei[j] - ei * nv[j]) * w;
      fr = Ds[0] * tr[0] + Ds[1] * tr[1] + Ds[2] * tr[2];
      fi = Ds[0] * ti[0] + Ds[1] * ti[1] + Ds[2] * ti[2];
      gr = Dt[0] * tr[0] + Dt[1] * tr[1] + Dt[2] * tr[2];
      gi = Dt[0] * ti[0] + Dt[1] * ti[1] + Dt[2] * ti[2];
    }
    double Bu[I], Bv[I];
    bernstein_basis<P>(u, Bu);
    bernstein_basis<P>(v, Bv);
#pragma unroll
    for (int iy = 0; iy < I; ++iy)
#pragma unroll
      for (int ix = 0; ix < I; ++ix) {
        const double b = Bu[ix] * Bv[iy];
        ar[iy * I + ix] += fr * b;
        ai[iy * I + ix] += fi * b;
        br[iy * I + ix] += gr * b;
        bi[iy * I + ix] += gi * b;
      }
  }
  const int64_t vcs = static_cast<int64_t>(nel) * I2;  // second vector component (fn 2)
  for (int m = 0; m < I2; ++m) {
    const int64_t o = static_cast<int64_t>(k) * nsuper + static_cast<int64_t>(e) * I2 + m;
    if (fn == 0 || fn == 4 || fn == 5) {
      out[o] = ar[m];
    } else {
      out[2 * o] = ar[m], out[2 * o + 1] = ai[m];
      if (fn == 2 || fn == 7) out[2 * (o + vcs)] = br[m], out[2 * (o + vcs) + 1] = bi[m];
    }
  }
}

}  // namespace bmi

using namespace bmi;

extern "C" {

// fn 0-4: built-in data with params; fn 5-7: values at the linear-form points (vals, host or
// device per mem)
static bmi_status rhs_common(const bmi_space* s, int fn, const double* params, const double* values, int nrhs,
                             double kappa, double* out, int mem, void* stream) {
  try {
    BMI_REQUIRE(s && out && nrhs >= 1, BMI_INVALID_ARG, "bad argument");
    BMI_REQUIRE(fn >= 0 && fn <= 7, BMI_INVALID_ARG, "unknown data function");
    BMI_REQUIRE(fn >= 5 ? values != nullptr : params != nullptr, BMI_INVALID_ARG, "null data");
    const bool vec = fn == 2 || fn == 7;
    BMI_REQUIRE(vec == (s->form == BMI_FORM_DIVCONFORMING), BMI_INVALID_ARG,
                "Dirichlet traces need a scalar space, the rotated tangential trace a DivConforming one");
    hipStream_t st = static_cast<hipStream_t>(stream);
    ensure_device_space(s, st);
    const bool real = fn == 0 || fn == 4 || fn == 5;
    const int S = real ? 1 : 2;
    DevBuf<double> prm, sup, dout, dval;
    const int per = fn == 2 ? 6 : 3;
    if (fn < 5) prm.upload(std::vector<double>(params, params + per * nrhs), st);
    const double* vals = values;
    if (fn >= 5 && mem != BMI_MEM_DEVICE) {
      const int64_t npts = static_cast<int64_t>(s->nel) * (s->P + 2) * (s->P + 2);
      const int per_v = fn == 5 ? 1 : fn == 6 ? 2 : 6;
      dval.upload(std::vector<double>(values, values + per_v * npts * nrhs), st);
      vals = dval.ptr;
    }
    sup.alloc(static_cast<size_t>(S) * s->nsuper * nrhs);
    const int64_t tot = static_cast<int64_t>(s->nel) * nrhs;
    const unsigned blocks = static_cast<unsigned>((tot + 127) / 128);
    const int stride = 4 * s->pd * s->pd;
    const double h = 1.0 / s->n;
#define BMI_RHS(PD, PP)                                                                                       \
  if (s->pd == PD && s->P == PP) {                                                                            \
    hipLaunchKernelGGL((k_rhs<PD, PP>), dim3(blocks), dim3(128), 0, st, s->d_cp.ptr, stride, s->d_elem_patch.ptr, \
                       s->d_elem_llc.ptr, s->nel, h, fn, prm.ptr, nrhs, kappa, static_cast<int64_t>(s->nsuper),   \
                       sup.ptr, vals);                                                                        \
  } else
    BMI_RHS(5, 1) BMI_RHS(5, 2) BMI_RHS(2, 1) BMI_RHS(2, 2) { throw Error(BMI_UNSUPPORTED, "rhs kernel not instantiated"); }
#undef BMI_RHS
    BMI_CHECK_HIP(hipGetLastError());
    double* y = out;
    if (mem != BMI_MEM_DEVICE) {
      dout.alloc(static_cast<size_t>(S) * s->ndof * nrhs);
      y = dout.ptr;
    }
    launch_spmv(!real, s->ndof, nrhs, s->d_trowptr.ptr, s->d_tcol.ptr, s->d_tval.ptr, sup.ptr, s->nsuper, y, s->ndof,
                nullptr, nullptr, st);
    if (mem != BMI_MEM_DEVICE)
      BMI_CHECK_HIP(hipMemcpyAsync(out, dout.ptr, dout.bytes(), hipMemcpyDeviceToHost, st));
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

bmi_status bmi_rhs_assemble(const bmi_space* s, int fn, const double* params, int nrhs, double kappa, double* out,
                            int mem, void* stream) {
  if (fn < 0 || fn > 4) {
    set_last_error("unknown data function");
    return BMI_INVALID_ARG;
  }
  return rhs_common(s, fn, params, nullptr, nrhs, kappa, out, mem, stream);
}

bmi_status bmi_rhs_from_values(const bmi_space* s, int kind, const double* values, int nrhs, double* out, int mem,
                               void* stream) {
  if (kind < 0 || kind > 2) {
    set_last_error("unknown value kind");
    return BMI_INVALID_ARG;
  }
  return rhs_common(s, 5 + kind, nullptr, values, nrhs, 0.0, out, mem, stream);
}

static bmi_status solve_common(const bmi_h2* hc, int nrhs, const double* B, double* X, double tol, int restart, int maxit,
                               int zero_guess, int* iters, double* resid, int mem, void* stream, bool use_cg) {
  try {
    BMI_REQUIRE(hc && B && X && nrhs >= 1 && tol > 0 && maxit >= 1, BMI_INVALID_ARG, "bad argument");
    bmi_h2* h = const_cast<bmi_h2*>(hc);
    if (use_cg) BMI_REQUIRE(h2_is_symmetric_real(h), BMI_INVALID_ARG, "CG needs the real symmetric Laplace operator");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int S = h2_complex(h) ? 2 : 1;
    const int64_t n = h2_rows(h);
    DevBuf<double> dB, dX;
    const double* b = B;
    double* x = X;
    if (mem != BMI_MEM_DEVICE) {
      dB.upload(std::vector<double>(B, B + S * n * nrhs), st);
      dX.upload(std::vector<double>(X, X + S * n * nrhs), st);
      b = dB.ptr;
      x = dX.ptr;
    }
    if (zero_guess) device_fill(x, S * n * nrhs, 0.0, st);
    std::vector<int> it;
    std::vector<double> rs;
    bool ok;
    if (h2_is_mixed(h) && !use_cg)
      // restarted GMRES whose cycles start from the FP64 residual: iterative refinement at the
      // granularity of one restart cycle, so the reduced operator costs no extra restarts
      ok = gmres(h, nrhs, b, x, tol, restart, maxit, it, rs, st, true, true);
    else if (h2_is_mixed(h))
      ok = refine(h, nrhs, b, x, tol, restart, maxit, it, rs, st, use_cg, nullptr);
    else
      ok = use_cg ? cg(h, nrhs, b, x, tol, maxit, it, rs, st, false)
                  : gmres(h, nrhs, b, x, tol, restart, maxit, it, rs, st, false);
    if (mem != BMI_MEM_DEVICE) BMI_CHECK_HIP(hipMemcpyAsync(X, dX.ptr, dX.bytes(), hipMemcpyDeviceToHost, st));
    BMI_CHECK_HIP(hipStreamSynchronize(st));
    if (iters) std::copy(it.begin(), it.end(), iters);
    if (resid) std::copy(rs.begin(), rs.end(), resid);
    return ok ? BMI_OK : BMI_NOT_CONVERGED;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.status;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return BMI_DEVICE_ERROR;
  }
}

struct bmi_gmres_task {
  std::unique_ptr<bmi::GmresRun> run;
  hipStream_t st = nullptr;
};

bmi_status bmi_gmres_start(const bmi_h2* hc, int nrhs, const double* B, double* X, double tol, int restart,
                           int maxit, int zero_guess, void* stream, bmi_gmres_task** out) {
  using namespace bmi;
  try {
    BMI_REQUIRE(hc && B && X && out && nrhs >= 1 && tol > 0 && maxit >= 1 && restart >= 1, BMI_INVALID_ARG,
                "bad argument");
    *out = nullptr;
    bmi_h2* h = const_cast<bmi_h2*>(hc);
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int S = h2_complex(h) ? 2 : 1;
    if (zero_guess) device_fill(X, S * h2_rows(h) * nrhs, 0.0, st);
    // mixed handles: cycles start from the FP64 residual (as bmi_solve_gmres)
    const bool mixed = h2_is_mixed(h);
    auto t = std::make_unique<bmi_gmres_task>();
    t->st = st;
    t->run = std::make_unique<GmresRun>(h, nrhs, B, X, tol, restart, maxit, st, mixed, mixed);
    *out = t.release();
    return BMI_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.status;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return BMI_DEVICE_ERROR;
  }
}

bmi_status bmi_gmres_poll(bmi_gmres_task* t, int* done) {
  using namespace bmi;
  try {
    BMI_REQUIRE(t && t->run && done, BMI_INVALID_ARG, "bad argument");
    *done = t->run->poll();
    return BMI_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.status;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return BMI_DEVICE_ERROR;
  }
}

bmi_status bmi_gmres_finish(bmi_gmres_task* t, int* iters, double* resid) {
  using namespace bmi;
  if (!t) {
    set_last_error("null task");
    return BMI_INVALID_ARG;
  }
  bmi_status status = BMI_OK;
  try {
    BMI_REQUIRE(t->run, BMI_INVALID_ARG, "task already finished");
    while (t->run->poll() != 1) t->run->wait();
    std::vector<int> it;
    std::vector<double> rs;
    t->run->results(it, rs);
    if (iters) std::copy(it.begin(), it.end(), iters);
    if (resid) std::copy(rs.begin(), rs.end(), resid);
    status = t->run->converged() ? BMI_OK : BMI_NOT_CONVERGED;
  } catch (const Error& e) {
    set_last_error(e.what());
    status = e.status;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    status = BMI_DEVICE_ERROR;
  }
  try {
    t->run.reset();
  } catch (...) {
  }
  delete t;
  return status;
}

bmi_status bmi_stream_create(void** out) {
  if (!out) {
    bmi::set_last_error("null argument");
    return BMI_INVALID_ARG;
  }
  hipStream_t s = nullptr;
  const hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e != hipSuccess) {
    bmi::set_last_error(hipGetErrorString(e));
    return BMI_DEVICE_ERROR;
  }
  *out = s;
  return BMI_OK;
}

bmi_status bmi_stream_destroy(void* stream) {
  if (!stream) return BMI_OK;
  const hipError_t e = hipStreamDestroy(static_cast<hipStream_t>(stream));
  if (e != hipSuccess) {
    bmi::set_last_error(hipGetErrorString(e));
    return BMI_DEVICE_ERROR;
  }
  return BMI_OK;
}

bmi_status bmi_solve_gmres(const bmi_h2* h, int nrhs, const double* B, double* X, double tol, int restart, int maxit,
                           int zero_guess, int* iters, double* resid, int mem, void* stream) {
  return solve_common(h, nrhs, B, X, tol, restart, maxit, zero_guess, iters, resid, mem, stream, false);
}

bmi_status bmi_solve_cg(const bmi_h2* h, int nrhs, const double* B, double* X, double tol, int maxit, int zero_guess,
                        int* iters, double* resid, int mem, void* stream) {
  return solve_common(h, nrhs, B, X, tol, 0, maxit, zero_guess, iters, resid, mem, stream, true);
}

}  // extern "C"
