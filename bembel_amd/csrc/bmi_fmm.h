// Launchers of the interpolation-FMM kernels (bmi_fmm.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

struct bmi_space;

namespace bmi {

constexpr int CP_STRIDE = 12;  // doubles per cluster node: x(3) kappa ds(3) dt(3) pad(2)

void launch_cluster_points(const bmi_space* s, int level, int np, const double* d_cheb, double* out,
                           hipStream_t st);
void launch_coupling_assemble(int kind, const int4* d_leaves, int nleaves, int np2,
                              const double* const* d_level_pts, double kre, double kim, double* out,
                              hipStream_t st);
void launch_spmv(bool cplx, int nrows, int nrhs, const int* rp, const int* ci, const double* v, const double* x,
                 int64_t ldx, double* y, int64_t ldy, const double* alpha, const double* beta, hipStream_t st);
void launch_moment(bool cplx, int ncl, int nrhs, int rows, int mc, const double* mom, const double* xs,
                   int64_t ldxs, int64_t seg, double* fw, hipStream_t st);
void launch_m2m(bool cplx, int ncl, int nrhs, int rows, int np2, const double* T, const double* child,
                double* parent, hipStream_t st);
void launch_coupling_apply(bool cplx, bool fp32, int ntgt, int nrhs, int cd, const int2* tgt, const int* rp,
                           const int* src, const void* F, int64_t fstride, int64_t fcomp, const int64_t* lvl_off,
                           const double* fw, double* bw, int accumulate, hipStream_t st);
void launch_l2l(bool cplx, int nchild, int nrhs, int rows, int np2, const double* T, const double* parent,
                double* child, hipStream_t st);
void launch_leaf(bool cplx, int ncl, int nrhs, int dd, const int* rp, const int* src, const double* F,
                 int64_t fstride, int64_t fcomp, const double* xs, int64_t ldxs, int64_t xseg, const double* mom,
                 int rows, const double* bw, double* ys, int64_t ldys, int64_t yseg, int accumulate,
                 hipStream_t st);

}  // namespace bmi
