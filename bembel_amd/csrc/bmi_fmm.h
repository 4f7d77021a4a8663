// Launchers of the interpolation-FMM kernels (bmi_fmm.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

struct bmi_space;

namespace bmi {

// from this many right-hand sides on, block rows run one wave per owner and 64 columns
constexpr int MULTI_MIN_RHS = 8;

constexpr int CP_STRIDE = 12;  // doubles per cluster node: x(3) kappa ds(3) dt(3) pad(2)

void launch_cluster_points(const bmi_space* s, int level, int np, const double* d_cheb, double* out,
                           hipStream_t st);
void launch_coupling_assemble(int kind, const int4* d_leaves, int nleaves, int np2,
                              const double* const* d_level_pts, double kre, double kim, double* out,
                              hipStream_t st);
// y = alpha A (x + x2) + beta y with A in CSR; x2 may be null; long_rows: 16 lanes per row
void launch_spmv(bool cplx, int nrows, int nrhs, const int* rp, const int* ci, const double* v, const double* x,
                 int64_t ldx, double* y, int64_t ldy, const double* alpha, const double* beta, hipStream_t st,
                 const double* x2 = nullptr, bool long_rows = false);
// variant: 0 = one block per lane group per iteration, 1 = two (chosen by the autotuner)
// Fc/ncoarse: BF16 copies of the first ncoarse blocks (fp32 only); ld: column stride of F (0: cd)
void launch_coupling_apply(bool cplx, bool fp32, int variant, int ntgt, int nrhs, int cd, const int2* tgt,
                           const int* rp, const int* src, const void* F, int64_t fstride, int64_t fcomp,
                           const int64_t* lvl_off, const double* fw, double* bw, int accumulate, hipStream_t st,
                           const unsigned short* Fc = nullptr, int ncoarse = 0, int ld = 0);
// many right-hand sides: internal vectors with the RHS index fastest (see bmi_fmm.hip)
void launch_spmv_tin(bool cplx, int nrows, int nrhs, const int* rp, const int* ci, const double* v, const double* X,
                     int64_t ldx, double* Yt, hipStream_t st);
void launch_spmv_tout(bool cplx, int nrows, int nrhs, const int* rp, const int* ci, const double* v, const double* At,
                      const double* Bt, double* Y, int64_t ldy, const double* alpha, const double* beta,
                      hipStream_t st);
// lvl_off: host array of the level offsets
void launch_forward_t(bool cplx, int npatch, int L, int rows, int np2, int dd, int nrhs, const double* mom,
                      const double* T, const double* xs_t, int64_t xseg, double* fw_t, const int64_t* lvl_off,
                      hipStream_t st);
// momt: Mom^T (dd x rows, column-major); Tt: the four T_q transposed
void launch_backward_t(bool cplx, int npatch, int L, int rows, int np2, int dd, int nrhs, const double* momt,
                       const double* Tt, double* bw_t, const int64_t* lvl_off, double* yf_t, int64_t yseg,
                       hipStream_t st);
// block sizes the RHS-fastest sweeps are instantiated for
bool multi_t_supported(bool cplx, int dd, int cd);
bool launch_multi_near_t(bool cplx, bool fp32, int ncl, int nrhs, int dd, const int* rp, const int* src,
                         const void* F, int64_t fstride, int64_t fcomp, const double* xs_t, int64_t xseg,
                         double* ys_t, int64_t yseg, int accumulate, hipStream_t st);
// lvl_off: device array
bool launch_multi_cpl_t(bool cplx, bool fp32, int ntgt, int nrhs, int cd, int ld, const int2* tgt, const int* rp,
                        const int* src, const void* F, int64_t fstride, int64_t fcomp, const int64_t* lvl_off,
                        const double* fw_t, double* bw_t, int accumulate, hipStream_t st);
// column stride of the FP32 coupling copy: cd rounded up to whole 16-byte row vectors
int fp32_coupling_ld(bool cplx, int cd);
void convert_blocks_to_fp32(const double* src, float* dst, int64_t nb, int cd, int ld, bool cplx, hipStream_t st);
// Maxwell far field from the kernel values of each coupling leaf: the four columns (tangent-
// weighted value moments, summed derivative moments) of every cluster entry, then one sweep
// that writes both components' local expansions of every cluster entry
void launch_mx_prep(int64_t ne, int nrhs, int np2, int cd, const double* fw0, const double* fw1, const double* tan,
                    double* Z, hipStream_t st);
void launch_mx_coupling(bool fp32, int ntgt_all, int nrhs, int np2, int cd, const int* order, const int2* tgt_all,
                        const int* rp_all, const int* src, const void* K, const int64_t* lvl_off, const double* tan, const double* Z,
                        double* bw0, double* bw1, double qre, double qim, hipStream_t st);
void convert_to_bf16(const double* src, unsigned short* dst, int64_t n, hipStream_t st);
// forward transformation of all levels (moments at level L, transfers up to level 0)
void launch_forward(bool cplx, int npatch, int L, int rows, int np2, int dd, int nrhs, const double* mom,
                    const double* T, const double* xs, int64_t ldxs, int64_t xseg, double* fw, const int64_t* lvl_off,
                    hipStream_t st, double* bw_zero = nullptr);
// backward transformation of all levels, ending in the leaf rows yf = Mom^T bw_L (written)
void launch_backward(bool cplx, int npatch, int L, int rows, int np2, int dd, int nrhs, const double* mom,
                     const double* T, double* bw, const int64_t* lvl_off, double* yf, int64_t ldy, int64_t yseg,
                     hipStream_t st);
void launch_leaf(bool cplx, int variant, int ncl, int nrhs, int dd, const int* rp, const int* src, const double* F,
                 int64_t fstride, int64_t fcomp, const double* xs, int64_t ldxs, int64_t xseg, double* ys, int64_t ldys,
                 int64_t yseg, int accumulate, hipStream_t st);
// near-field sweep and coupling sweep of a single-RHS, single-component product in one launch
// (ys written, bw written); false when no fused kernel exists for these sizes
// near32: the near-field blocks dF are the FP32 copy (with FP32 coupling blocks)
// block sizes with symmetric single-RHS sweeps
bool sym_sweep_supported(bool cplx, int dd, int cd);
// upper half of the coupling blocks of a symmetric kernel, for single-RHS sweeps
struct SymList {
  int ntgt;  // targets with upper blocks, their CSR (sources) and the block index of each entry
  const int2* tgt;
  const int* rp;
  const int* src;
  const int* blk;
  double* part;      // B^T x_t of every entry (S * cd doubles), stored in gather order
  const int* slot;   // per entry: its slot in part
  int ngather;       // targets receiving partials and the ranges of their slots
  const int2* gtgt;
  const int* grp;
  // near-field lists (row clusters of the leaf level as targets of level 0; the diagonal leaf of a
  // cluster has slot -1): their level offsets are one zero
  const int64_t* lvl_off = nullptr;
};

// timing-scope name of the next single-RHS sweep launches on this thread (null: the sweep's own)
extern thread_local const char* tl_sweep_scope;
bool launch_sweep_fused(bool cplx, bool fp32, int leaf_variant, int cpl_variant, int ncl, int dd, const int* drp,
                        const int* dsrc, const void* dF, const double* xs, double* ys, int ntgt, int cd,
                        const int2* tgt, const int* crp, const int* csrc, const void* cF, const int64_t* lvl_off,
                        const double* fw, double* bw, hipStream_t st, const unsigned short* Fc = nullptr,
                        int ncoarse = 0, bool near32 = false, int ld = 0,
                        const SymList* sym = nullptr, const SymList* nsym = nullptr);
// FP32 node tables (offsets from each cluster's node 0, weights; node 0 in FP64) of one level
void launch_fly_points(const double* pts, int ncl, int np2, float4* out, double4* org, hipStream_t st);
// single-RHS sweep of the reduced operator: FP32 near-field blocks, coupling evaluated on the fly
// in FP32 (ys and bw written); false when no kernel exists for these sizes
// variant: near-field blocks loaded per iteration, 0 = one, 1 = two
bool launch_sweep_fly(bool cplx, int variant, int ncl, int dd, const int* drp, const int* dsrc, const float* dF32,
                      const double* xs, double* ys, int ntgt, int cd, const int2* tgt, const int* crp,
                      const int* csrc, const int64_t* lvl_off, const float4* pts, const double4* org,
                      double kre, double kim, const double* fw, double* bw, hipStream_t st);
void launch_leaf_backward(bool cplx, int ncl, int nrhs, int dd, const double* mom, int rows, const double* bw,
                          double* yf, int64_t ldy, int64_t yseg, hipStream_t st);

}  // namespace bmi
