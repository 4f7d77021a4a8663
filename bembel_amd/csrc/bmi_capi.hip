// C ABI of bembel_amd: handles, device-resident H2 state, assembly and matvec orchestration.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <thread>

#include "bmi_fmm.h"
#include "bmi_internal.h"
#include "bmi_solver.h"

namespace bmi {

static thread_local std::string g_last_error;
static void set_error(const std::string& m) { g_last_error = m; }
void set_last_error(const std::string& m) { g_last_error = m; }

// ------------------------------------------------------------------ timing registry
static std::atomic<bool> g_timing{false};
static std::mutex g_tmu;
static std::map<std::string, std::pair<double, int64_t>> g_times;
// event pairs of finished scopes, read when the times are asked for: a scope does not wait for
// its kernels, so timed launches queue back to back as they do untimed (a wait per scope would
// leave the device idle before each launch and count the host's launch latency as kernel time)
struct PendingTime {
  const char* name;
  hipEvent_t a, b;
};
static std::vector<PendingTime> g_pending;
// the caller holds g_tmu
static void resolve_pending(bool keep) {
  for (const PendingTime& p : g_pending) {
    float ms = 0.f;
    if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess && keep) {
      auto& e = g_times[p.name];
      e.first += ms;
      e.second += 1;
    }
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  g_pending.clear();
}

static thread_local bool tl_timing_off = false;
bool timing_enabled() { return g_timing && !tl_timing_off; }

TimedScope::TimedScope(const char* n, hipStream_t s) : name(n), st(s) {
  if (!timing_enabled()) return;
  (void)hipEventCreate(&a);
  (void)hipEventCreate(&b);
  (void)hipEventRecord(a, st);
}
TimedScope::~TimedScope() {
  if (!a) return;
  (void)hipEventRecord(b, st);
  std::lock_guard<std::mutex> g(g_tmu);
  g_pending.push_back(PendingTime{name, a, b});
  if (g_pending.size() >= 4096) resolve_pending(true);  // bounds the events held
}

static double wall_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
HostTimedScope::HostTimedScope(const char* n) : name(n) {
  if (timing_enabled()) t0 = wall_ms();
}
HostTimedScope::~HostTimedScope() {
  if (t0 < 0) return;
  const double ms = wall_ms() - t0;
  std::lock_guard<std::mutex> g(g_tmu);
  auto& e = g_times[name];
  e.first += ms;
  e.second += 1;
}

// element-pair sub-blocks from their mirrors: component (r, c) of the I2 x I2 sub-block at
// map[2m] = (component (c, r) of the sub-block at map[2m + 1])^T; sub-blocks are column-major with
// leading dimension dd inside their leaf, components dblk apart, S doubles per entry
__global__ void k_mirror_blocks(int64_t nm, const int64_t* __restrict__ map, int I2, int dd, int64_t dblk, int vd,
                                int S, double* dense) {
  const int64_t per = static_cast<int64_t>(vd) * vd * I2 * I2;
  const int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= nm * per) return;
  const int64_t m = gid / per;
  const int q = static_cast<int>(gid - m * per);
  const int comp = q / (I2 * I2), e = q - comp * (I2 * I2);
  const int r = comp / vd, c = comp - r * vd, j = e / I2, i = e - j * I2;  // entry (i, j) of component (r, c)
  const int64_t dst = map[2 * m] + comp * dblk + static_cast<int64_t>(j) * dd + i;
  const int64_t src = map[2 * m + 1] + (c * vd + r) * dblk + static_cast<int64_t>(i) * dd + j;
  for (int t = 0; t < S; ++t) dense[S * dst + t] = dense[S * src + t];
}
void launch_mirror_blocks(int64_t nm, const int64_t* map, int I2, int dd, int64_t dblk, int vd, int S, double* dense,
                          hipStream_t st) {
  const int64_t n = nm * vd * vd * I2 * I2;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_mirror_blocks, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, st, nm, map, I2, dd,
                     dblk, vd, S, dense);
  BMI_CHECK_HIP(hipGetLastError());
}

__global__ void k_fill(double* p, int64_t n, double v) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_add(double* y, const double* x, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) y[i] += x[i];
}
// y += x (n doubles)
static void device_add(double* y, const double* x, int64_t n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_add, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, st, y, x, n);
  BMI_CHECK_HIP(hipGetLastError());
}
void device_fill(double* p, int64_t n, double v, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, st, p, n, v);
  BMI_CHECK_HIP(hipGetLastError());
}

}  // namespace bmi

using namespace bmi;

// ------------------------------------------------------------------ H2 handle
struct bmi_h2 {
  const bmi_space* s = nullptr;
  bmi_op_desc op{};
  bool cplx = false;
  int vd = 1, nc = 1, np = 0, np2 = 0, mcl = 1, L = 0, dd = 0, cd = 0;
  std::vector<Leaf> dense, lowrank;
  int64_t npairs = 0;
  std::vector<int> pair_class_count = std::vector<int>(5, 0);
  // leaf storage
  DevBuf<double> d_dense, d_coup;
  DevBuf<float> d_coup32;
  int ld32 = 0;  // column stride of the FP32 coupling blocks
  DevBuf<unsigned short> d_coup16;  // BF16 copies of the coarsest coupling blocks (MIXED_BF16)
  int ncoarse = 0;
  // MIXED single-component operators (red1): FP32 copy of the near-field blocks for the reduced
  // single-RHS product; for real kernels (fly) also the FP32 node tables from which that product
  // may evaluate the coupling blocks on the fly
  bool red1 = false, fly = false;
  DevBuf<float> d_dense32;
  DevBuf<float4> d_fly_pts;
  DevBuf<double4> d_fly_org;
  // dense CSR over leaf-level clusters
  int ncl_leaf = 0;
  DevBuf<int> d_drp, d_dsrc;
  // coupling CSR over (level, cluster) targets
  int ntgt = 0;
  DevBuf<int2> d_tgt;
  // every cluster of levels 0..L in (level, cluster) order with its coupling blocks (empty rows for
  // clusters without far-field blocks): the many-RHS coupling sweep writes every cluster's sum,
  // so the local expansions need no clearing pass
  int ntgt_all = 0;
  DevBuf<int2> d_tgt_all;
  DevBuf<int> d_crp_all;
  DevBuf<int> d_crp, d_csrc;
  std::vector<int64_t> lvl_off;  // per level start in per-RHS units (clusters * rows)
  DevBuf<int64_t> d_lvl_off;
  int64_t lvl_total = 0;
  // FMM constants
  DevBuf<double> d_T, d_mom[2];
  DevBuf<double> d_Tt, d_momt[2];  // transposed copies for the RHS-fastest backward steps
  // upper half of the coupling blocks (symmetric kernels) for single-RHS sweeps
  bool sym = false;
  int sym_fused = 0;  // chosen by the autotuner: FP64 single-RHS sweeps read only the upper blocks
  int sweep_mode = -1;  // -1: timed choice, 0: every coupling block, 1: upper blocks
  // single-RHS products with the handle's workspace: 1 runs the near-field leaves beside the far
  // chain on the side stream, 0 both sweeps in one launch after the forward transformation
  int overlap_mode = -1;  // -1: timed choice, else forced
  int overlap = 0;
  DevBuf<int2> d_stgt, d_sgtgt;
  // the same for the near field (upper and diagonal leaves, partials added into the rows)
  bool nsym = false;
  bmi::SymList nsl{};
  DevBuf<int2> d_ntgt, d_ngtgt;
  DevBuf<int> d_nrp, d_nsrc, d_nblk, d_nslot, d_ngrp;
  DevBuf<double> d_npart;
  DevBuf<int64_t> d_zero;
  DevBuf<int> d_srp, d_ssrc, d_sblk, d_sgrp, d_sgslot;
  DevBuf<double> d_part;
  bmi::SymList sl{};
  // workspace (grown on demand)
  int ws_rhs = 0;
  DevBuf<double> d_xs, d_ys, d_fw[2], d_bw, d_xin, d_yout;
  // super-space apply captured once per RHS count and replayed (workspace pointers are fixed
  // until the workspace grows, which drops the graphs)
  bool use_graphs = true;
  hipStream_t cap_stream = nullptr;
  std::map<int, hipGraphExec_t> graphs;  // key: 2 * nrhs + reduced
  // the far-field chain runs on a side stream, concurrently with the near-field sweep, and
  // writes its leaf rows into d_yf; the final T^T product adds both
  hipStream_t side_stream = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  DevBuf<double> d_yf;
  // Maxwell far field (mxf): d_coup holds the kernel values K (np2 x np2) of each coupling leaf;
  // the tangents of every cluster entry's nodes (6 doubles per node, device and host copies) and
  // the four-column workspace of the sweep
  bool mxf = false;
  double mx_qre = 0.0, mx_qim = 0.0;  // -1 / kappa^2
  DevBuf<double> d_mx_tan, d_mxz;
  DevBuf<int> d_mx_order;  // cluster entries by decreasing number of coupling leaves
  std::vector<double> h_mx_tan;
  // sweep kernel variants picked by timing on first use (-1: not yet); the fused single-RHS
  // sweep has its own pair
  int leaf_variant = -1, cpl_variant = -1;
  int fused_leaf_variant = -1, fused_cpl_variant = -1;
  int fly_variant = -1;
  void drop_graphs() {
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
    graphs.clear();
  }
  ~bmi_h2() {
    drop_graphs();
    if (cap_stream) (void)hipStreamDestroy(cap_stream);
    if (side_stream) (void)hipStreamDestroy(side_stream);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
  }
  size_t bytes() const {
    return d_dense.bytes() + d_coup.bytes() + d_coup32.bytes() + d_coup16.bytes() + d_dense32.bytes() +
           d_fly_pts.bytes() + d_fly_org.bytes() + d_drp.bytes() + d_dsrc.bytes() + d_tgt.bytes() + d_tgt_all.bytes() + d_crp_all.bytes() +
           d_crp.bytes() + d_csrc.bytes() + d_T.bytes() + d_mom[0].bytes() + d_mom[1].bytes() + d_Tt.bytes() + d_momt[0].bytes() + d_momt[1].bytes() + d_stgt.bytes() + d_sgtgt.bytes() + d_srp.bytes() +
           d_ssrc.bytes() + d_sblk.bytes() + d_sgrp.bytes() + d_sgslot.bytes() + d_part.bytes() + d_xs.bytes() +
           d_ys.bytes() + d_fw[0].bytes() + d_fw[1].bytes() + d_bw.bytes() + d_yf.bytes() + d_mx_tan.bytes() +
           d_mxz.bytes() + d_mx_order.bytes() + d_ntgt.bytes() + d_ngtgt.bytes() + d_nrp.bytes() + d_nsrc.bytes() +
           d_nblk.bytes() + d_nslot.bytes() + d_ngrp.bytes() + d_npart.bytes() + d_zero.bytes();
  }
};

namespace bmi {

static int ncl_at(const bmi_space* s, int l) { return s->npatch << (2 * l); }

// coupling blocks of the FP64 operator, or of the reduced-precision one of a mixed handle
struct CouplingOperand {
  bool fp32;
  const void* F;
  const unsigned short* Fc;
  int ncoarse;
  int ld;       // column stride of a block
  int64_t blk;  // entries per component block
};
static CouplingOperand coupling_operand(const bmi_h2* h, bool reduced) {
  if (h->mxf) {
    const int64_t blk = static_cast<int64_t>(h->np2) * h->np2;
    if (reduced && h->op.precision != BMI_PRECISION_FP64) return {true, h->d_coup32.ptr, nullptr, 0, h->np2, blk};
    return {false, h->d_coup.ptr, nullptr, 0, h->np2, blk};
  }
  if (reduced && h->op.precision != BMI_PRECISION_FP64)
    return {true, h->d_coup32.ptr, h->ncoarse > 0 ? h->d_coup16.ptr : nullptr, h->ncoarse, h->ld32,
            static_cast<int64_t>(h->ld32) * h->cd};
  return {false, h->d_coup.ptr, nullptr, 0, h->cd, static_cast<int64_t>(h->cd) * h->cd};
}

static void ensure_workspace(bmi_h2* h, int nrhs, hipStream_t st) {
  (void)st;
  if (nrhs <= h->ws_rhs) return;
  h->drop_graphs();
  const int S = h->cplx ? 2 : 1;
  const int64_t N = h->s->nsuper;
  h->d_xs.alloc(S * N * nrhs);
  h->d_ys.alloc(S * N * nrhs);
  for (int c = 0; c < h->vd; ++c) h->d_fw[c].alloc(S * h->lvl_total * nrhs);
  h->d_bw.alloc(S * h->lvl_total * nrhs);
  h->d_yf.alloc(S * N * nrhs);
  if (h->mxf) h->d_mxz.alloc(2 * S * h->lvl_total * nrhs);  // 4 np2 columns per entry = 2 cd
  h->ws_rhs = nrhs;
}

static bool has_far(const bmi_h2* h) { return h->L >= 0 && h->lvl_total > 0; }

// Workspace of one product: the handle's own buffers (bmi_h2_matvec, the solvers), or a region of
// the per-device scratch array (bmi_h2_matvec_noalloc).  Only products on the handle's own
// buffers may create side streams, events and graphs; the others run every kernel on the
// caller's stream.
struct Work {
  double* xs;     // super vector of the input
  double* ys;     // near-field rows
  double* yf;     // far-field rows
  double* fw[2];  // multipole coefficients per FMM component
  double* bw;     // local coefficients
  bool own;
  double* z;      // Maxwell far field: four columns per cluster entry
};
static Work handle_work(bmi_h2* h) {
  return Work{h->d_xs.ptr, h->d_ys.ptr, h->d_yf.ptr, {h->d_fw[0].ptr, h->d_fw[1].ptr}, h->d_bw.ptr, true,
              h->d_mxz.ptr};
}

// far-field chain: forward transformation, coupling, backward transformation, leaf rows -> yf
static void apply_far(bmi_h2* h, const Work& w, int nrhs, const double* xs, double* yf, bool reduced,
                      hipStream_t st) {
  const bmi_space* s = h->s;
  const bool C = h->cplx;
  const int S = C ? 2 : 1;
  const int64_t N = s->nsuper;
  const int rows = h->cd;
  const int64_t vcs = static_cast<int64_t>(s->nel) * s->I2;  // super DOFs per component
  const CouplingOperand co = coupling_operand(h, reduced);
  const int64_t cblk = co.blk;
  for (int cc = 0; cc < h->vd; ++cc) {
    TimedScope ts("matvec_forward", st);
    launch_forward(C, s->npatch, h->L, rows, h->np2, h->dd, nrhs, h->d_mom[cc].ptr, h->d_T.ptr, xs, N, cc * vcs,
                   w.fw[cc], h->d_lvl_off.ptr, st);
  }
  if (h->mxf) {
    // four columns per cluster entry from both components' moments, one sweep over the kernel
    // values writing both components' local expansions (into the moment buffers, consumed), then
    // the backward transformation per component
    launch_mx_prep(h->lvl_total / rows, nrhs, h->np2, rows, w.fw[0], w.fw[1], h->d_mx_tan.ptr, w.z, st);
    launch_mx_coupling(co.fp32, h->ntgt_all, nrhs, h->np2, rows, h->d_mx_order.ptr, h->d_tgt_all.ptr,
                       h->d_crp_all.ptr, h->d_csrc.ptr,
                       co.F, h->d_lvl_off.ptr, h->d_mx_tan.ptr, w.z, w.fw[0], w.fw[1], h->mx_qre, h->mx_qim, st);
    for (int rc = 0; rc < h->vd; ++rc) {
      TimedScope ts("matvec_backward", st);
      launch_backward(C, s->npatch, h->L, rows, h->np2, h->dd, nrhs, h->d_mom[rc].ptr, h->d_T.ptr, w.fw[rc],
                      h->d_lvl_off.ptr, yf, N, rc * vcs, st);
    }
    return;
  }
  for (int rc = 0; rc < h->vd; ++rc) {
    device_fill(w.bw, S * h->lvl_total * nrhs, 0.0, st);
    for (int cc = 0; cc < h->vd; ++cc)
      launch_coupling_apply(C, co.fp32, std::max(h->cpl_variant, 0), h->ntgt, nrhs, rows, h->d_tgt.ptr,
                            h->d_crp.ptr, h->d_csrc.ptr, co.F, h->vd * h->vd * cblk, (rc * h->vd + cc) * cblk,
                            h->d_lvl_off.ptr, w.fw[cc], w.bw, 1, st, co.Fc, co.ncoarse, co.ld);
    TimedScope ts("matvec_backward", st);
    launch_backward(C, s->npatch, h->L, rows, h->np2, h->dd, nrhs, h->d_mom[rc].ptr, h->d_T.ptr, w.bw,
                    h->d_lvl_off.ptr, yf, N, rc * vcs, st);
  }
}

// near-field sweep: ys = A_near xs
static void apply_near(bmi_h2* h, int nrhs, const double* xs, double* ys, hipStream_t st) {
  const bmi_space* s = h->s;
  const int64_t N = s->nsuper;
  const int64_t vcs = static_cast<int64_t>(s->nel) * s->I2;
  const int64_t dblk = static_cast<int64_t>(h->dd) * h->dd;
  for (int rc = 0; rc < h->vd; ++rc)
    for (int cc = 0; cc < h->vd; ++cc)
      launch_leaf(h->cplx, std::max(h->leaf_variant, 0), h->ncl_leaf, nrhs, h->dd, h->d_drp.ptr, h->d_dsrc.ptr,
                  h->d_dense.ptr, h->vd * h->vd * dblk, (rc * h->vd + cc) * dblk, xs, N, cc * vcs, ys, N, rc * vcs,
                  cc > 0 ? 1 : 0, st);
}

// time both variants of the two sweeps once (1 RHS) and keep the faster; the variants give
// identical bits, so the choice affects speed only
// names the timing scope of the sweep launches made while it lives
struct SweepScopeName {
  explicit SweepScopeName(const char* n) { tl_sweep_scope = n; }
  ~SweepScopeName() { tl_sweep_scope = nullptr; }
};

// One RHS with the fused sweep kernels: forward transformation (also clears bw), the near-field
// leaves and the coupling blocks, backward transformation.  overlap = false: both sweeps in one
// launch after the forward transformation.  overlap = true: after the forward transformation the
// coupling blocks (with their gather and the backward transformation) run on the side stream and
// the near-field leaves on `st`, so the two sweeps start together and the tail of the far chain
// runs beside the near-field leaves.  (The forward transformation stays ahead of the fork: beside
// the near-field launch its few workgroups wait for free slots, measured 9 -> 167 us.)  The two
// sweeps write disjoint vectors (ys; bw, then yf), so both forms give the same bits.  false if the
// sweep kernel is not instantiated for this operator.
static bool fused_product(bmi_h2* h, const Work& w, const double* xs, double* ys, double* yf, bool reduced,
                          hipStream_t st, bool overlap) {
  const bmi_space* s = h->s;
  const CouplingOperand co = coupling_operand(h, reduced);
  const int64_t N = s->nsuper;
  const bool use_sym = h->sym && h->sym_fused && !co.fp32;
  const SymList* sl = use_sym ? &h->sl : nullptr;
  const SymList* nsl = use_sym && h->sym_fused == 2 ? &h->nsl : nullptr;
  auto sweep = [&](int ncl, int ntgt, const SymList* csl, const SymList* nl, hipStream_t ss) {
    return launch_sweep_fused(h->cplx, co.fp32, h->fused_leaf_variant, h->fused_cpl_variant, ncl, h->dd,
                              h->d_drp.ptr, h->d_dsrc.ptr, h->d_dense.ptr, xs, ys, ntgt, h->cd, h->d_tgt.ptr,
                              h->d_crp.ptr, h->d_csrc.ptr, co.F, h->d_lvl_off.ptr, w.fw[0], w.bw, ss, co.Fc,
                              co.ncoarse, false, co.ld, csl, nl);
  };
  {
    TimedScope ts("matvec_forward", st);
    launch_forward(h->cplx, s->npatch, h->L, h->cd, h->np2, h->dd, 1, h->d_mom[0].ptr, h->d_T.ptr, xs, N, 0,
                   w.fw[0], h->d_lvl_off.ptr, st, w.bw);
  }
  hipStream_t fs = st;
  if (overlap) {
    if (!h->side_stream) BMI_CHECK_HIP(hipStreamCreateWithFlags(&h->side_stream, hipStreamNonBlocking));
    if (!h->ev_fork) BMI_CHECK_HIP(hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming));
    if (!h->ev_join) BMI_CHECK_HIP(hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming));
    BMI_CHECK_HIP(hipEventRecord(h->ev_fork, st));
    BMI_CHECK_HIP(hipStreamWaitEvent(h->side_stream, h->ev_fork, 0));
    fs = h->side_stream;
  }
  bool ok;
  if (!overlap) {
    ok = sweep(h->ncl_leaf, h->ntgt, sl, nsl, st);
  } else {
    const SymList none{};  // no coupling targets in the near-field launch
    bool okc, okn;
    {
      SweepScopeName sn("matvec_sweep_far");
      okc = sweep(0, h->ntgt, sl, nullptr, fs);
    }
    {
      SweepScopeName sn("matvec_sweep_near");
      okn = sweep(h->ncl_leaf, 0, sl ? &none : nullptr, nsl, st);
    }
    ok = okc && okn;
  }
  if (ok) {
    TimedScope ts("matvec_backward", fs);
    launch_backward(h->cplx, s->npatch, h->L, h->cd, h->np2, h->dd, 1, h->d_mom[0].ptr, h->d_T.ptr, w.bw,
                    h->d_lvl_off.ptr, yf, N, 0, fs);
  }
  if (overlap) {
    BMI_CHECK_HIP(hipEventRecord(h->ev_join, fs));
    BMI_CHECK_HIP(hipStreamWaitEvent(st, h->ev_join, 0));
  }
  return ok;
}

static void autotune(bmi_h2* h, const double* xs, bool reduced, hipStream_t st) {
  if (h->leaf_variant >= 0 && h->cpl_variant >= 0) return;
  hipEvent_t a, b;
  BMI_CHECK_HIP(hipEventCreate(&a));
  BMI_CHECK_HIP(hipEventCreate(&b));
  auto time_it = [&](auto&& fn) {
    fn();  // warm-up
    BMI_CHECK_HIP(hipEventRecord(a, st));
    for (int r = 0; r < 3; ++r) fn();
    BMI_CHECK_HIP(hipEventRecord(b, st));
    BMI_CHECK_HIP(hipEventSynchronize(b));
    float ms = 0.f;
    BMI_CHECK_HIP(hipEventElapsedTime(&ms, a, b));
    return ms;
  };
  float best = 0.f;
  int pick = 0;
  for (int v = 0; v < 2; ++v) {
    h->leaf_variant = v;
    const float t = time_it([&] { apply_near(h, 1, xs, h->d_ys.ptr, st); });
    if (v == 0 || t < best) best = t, pick = v;
  }
  h->leaf_variant = pick;
  if (has_far(h) && !h->mxf) {
    const int S = h->cplx ? 2 : 1;
    const CouplingOperand co = coupling_operand(h, reduced);
    const bool fp32 = co.fp32;
    const void* coupF = co.F;
    const int64_t cblk = co.blk;
    device_fill(h->d_fw[0].ptr, S * h->lvl_total, 0.0, st);
    pick = 0;
    for (int v = 0; v < 2; ++v) {
      const float t = time_it([&] {
        launch_coupling_apply(h->cplx, fp32, v, h->ntgt, 1, h->cd, h->d_tgt.ptr, h->d_crp.ptr, h->d_csrc.ptr, coupF,
                              h->vd * h->vd * cblk, 0, h->d_lvl_off.ptr, h->d_fw[0].ptr, h->d_bw.ptr, 0, st, co.Fc,
                              co.ncoarse, co.ld);
      });
      if (v == 0 || t < best) best = t, pick = v;
    }
    h->cpl_variant = pick;
    // the fused sweep: all four combinations (register use, and so occupancy, differs)
    if (h->vd == 1) {
      // full block lists, and (symmetric kernels, FP64 blocks) the upper blocks with transposed
      // partials: the upper form reads half the coupling bytes but adds the column sums
      int pl = -1, pc = -1, ps = 0;
      // forms: 0 every block, 1 upper coupling blocks, 2 upper coupling blocks and upper near-field leaves
      const int nsym = (h->sym && !fp32) ? (h->nsym ? 3 : 2) : 1;
      const int s0 = h->sweep_mode >= 0 ? std::min(h->sweep_mode, nsym - 1) : 0;
      const int s1 = h->sweep_mode >= 0 ? s0 : nsym - 1;
      for (int sy = s0; sy <= s1; ++sy)
        for (int vl = 0; vl < 2; ++vl)
          for (int vc = 0; vc < 2; ++vc) {
            bool ok = true;
            const float t = time_it([&] {
              ok = launch_sweep_fused(h->cplx, fp32, vl, vc, h->ncl_leaf, h->dd, h->d_drp.ptr, h->d_dsrc.ptr,
                                      h->d_dense.ptr, xs, h->d_ys.ptr, h->ntgt, h->cd, h->d_tgt.ptr, h->d_crp.ptr,
                                      h->d_csrc.ptr, coupF, h->d_lvl_off.ptr, h->d_fw[0].ptr, h->d_bw.ptr, st, co.Fc,
                                      co.ncoarse, false, co.ld, sy ? &h->sl : nullptr,
                                      sy == 2 ? &h->nsl : nullptr);
            });
            if (ok && (pl < 0 || t < best)) best = t, pl = vl, pc = vc, ps = sy;
          }
      h->fused_leaf_variant = pl;
      h->fused_cpl_variant = pc;
      h->sym_fused = ps;
      // whole products: both sweeps in one launch, or the near-field leaves beside the far chain
      h->overlap = 0;
      if (pl >= 0 && h->overlap_mode != 0) {
        const Work w = handle_work(h);
        if (h->overlap_mode > 0) {
          h->overlap = 1;
        } else {
          float tb = 0.f;
          for (int o = 0; o < 2; ++o) {
            bool ok = true;
            const float t = time_it([&] { ok = fused_product(h, w, xs, w.ys, w.yf, reduced, st, o == 1); });
            if (ok && (o == 0 || t < tb)) tb = t, h->overlap = o;
          }
        }
      }
    }
    if (h->red1) {
      // reduced single-RHS sweep: coupling recomputed on the fly (variants 0, 1; real kernels:
      // an oscillatory kernel's FP32 phase is less accurate than the FP32-rounded FP64 entries and
      // costs Krylov steps) or read from the FP32 copy (variants 2 + 2 * near blocks per
      // iteration + coupling blocks per iteration), both with the FP32 near-field blocks;
      // whichever is faster (arithmetic vs bytes)
      // (variants 6-9: the FP32 copy's upper blocks with transposed partials)
      int pv = -1;
      const int v0 = (h->sym && h->sweep_mode >= 1) ? 6 : (h->fly ? 0 : 2);
      const int v1 = (h->sym && h->sweep_mode != 0) ? 10 : 6;
      for (int v = v0; v < v1; ++v) {
        bool ok = true;
        const float t = time_it([&] {
          if (v < 2)
            ok = launch_sweep_fly(h->cplx, v, h->ncl_leaf, h->dd, h->d_drp.ptr, h->d_dsrc.ptr, h->d_dense32.ptr, xs,
                                  h->d_ys.ptr, h->ntgt, h->cd, h->d_tgt.ptr, h->d_crp.ptr, h->d_csrc.ptr,
                                  h->d_lvl_off.ptr, h->d_fly_pts.ptr, h->d_fly_org.ptr, h->op.kappa_re,
                                  h->op.kappa_im, h->d_fw[0].ptr, h->d_bw.ptr, st);
          else
            ok = launch_sweep_fused(h->cplx, true, ((v - 2) & 3) >> 1, (v - 2) & 1, h->ncl_leaf, h->dd,
                                    h->d_drp.ptr, h->d_dsrc.ptr, h->d_dense32.ptr, xs, h->d_ys.ptr, h->ntgt, h->cd,
                                    h->d_tgt.ptr, h->d_crp.ptr, h->d_csrc.ptr, h->d_coup32.ptr, h->d_lvl_off.ptr,
                                    h->d_fw[0].ptr, h->d_bw.ptr, st, nullptr, 0, true, h->ld32,
                                    v >= 6 ? &h->sl : nullptr);
        });
        if (ok && (pv < 0 || t < best)) best = t, pv = v;
      }
      h->fly_variant = pv;
    }
  } else {
    h->cpl_variant = 0;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
}

// super-space apply: ys = A_near xs on `st`, yf = A_far xs on the side stream (joined back into
// `st` before returning); the caller adds the two
static void apply_super(bmi_h2* h, const Work& w, int nrhs, const double* xs, double* ys, double* yf,
                        bool reduced, hipStream_t st) {
  const bool far = has_far(h);
  if (far && reduced && h->red1 && nrhs == 1) {
    // reduced operator, one RHS: FP32 near-field blocks and the coupling evaluated on the fly
    const bmi_space* s = h->s;
    const int64_t N = s->nsuper;
    {
      TimedScope ts("matvec_forward", st);
      launch_forward(h->cplx, s->npatch, h->L, h->cd, h->np2, h->dd, 1, h->d_mom[0].ptr, h->d_T.ptr, xs, N, 0,
                     w.fw[0], h->d_lvl_off.ptr, st, w.bw);
    }
    // the variant the first product timed; before that (an allocation-free product on a handle
    // that has not run bmi_h2_matvec yet) the on-the-fly sweep for real kernels, else the FP32 copy
    const int v = h->fly_variant >= 0 ? h->fly_variant : (h->fly ? 0 : 2);
    const bool ok =
        v < 2 ? launch_sweep_fly(h->cplx, v, h->ncl_leaf, h->dd, h->d_drp.ptr, h->d_dsrc.ptr, h->d_dense32.ptr, xs, ys,
                                 h->ntgt, h->cd, h->d_tgt.ptr, h->d_crp.ptr, h->d_csrc.ptr, h->d_lvl_off.ptr,
                                 h->d_fly_pts.ptr, h->d_fly_org.ptr, h->op.kappa_re, h->op.kappa_im, w.fw[0],
                                 w.bw, st)
              : launch_sweep_fused(h->cplx, true, ((v - 2) & 3) >> 1, (v - 2) & 1, h->ncl_leaf, h->dd,
                                   h->d_drp.ptr, h->d_dsrc.ptr, h->d_dense32.ptr, xs, ys, h->ntgt, h->cd, h->d_tgt.ptr,
                                   h->d_crp.ptr, h->d_csrc.ptr, h->d_coup32.ptr, h->d_lvl_off.ptr, w.fw[0],
                                   w.bw, st, nullptr, 0, true, h->ld32, v >= 6 ? &h->sl : nullptr);
    if (ok) {
      TimedScope ts("matvec_backward", st);
      launch_backward(h->cplx, s->npatch, h->L, h->cd, h->np2, h->dd, 1, h->d_mom[0].ptr, h->d_T.ptr, w.bw,
                      h->d_lvl_off.ptr, yf, N, 0, st);
      return;
    }
  }
  if (far && nrhs == 1 && h->vd == 1 && h->fused_leaf_variant >= 0 && h->fused_cpl_variant >= 0 &&
      fused_product(h, w, xs, ys, yf, reduced, st, w.own && h->overlap == 1))
    return;
  if (far && !w.own) {
    // one stream: the two chains write disjoint vectors, so the order gives the same bits
    apply_far(h, w, nrhs, xs, yf, reduced, st);
    apply_near(h, nrhs, xs, ys, st);
    return;
  }
  if (far) {
    if (!h->side_stream) BMI_CHECK_HIP(hipStreamCreateWithFlags(&h->side_stream, hipStreamNonBlocking));
    if (!h->ev_fork) BMI_CHECK_HIP(hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming));
    if (!h->ev_join) BMI_CHECK_HIP(hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming));
    BMI_CHECK_HIP(hipEventRecord(h->ev_fork, st));
    BMI_CHECK_HIP(hipStreamWaitEvent(h->side_stream, h->ev_fork, 0));
    apply_far(h, w, nrhs, xs, yf, reduced, h->side_stream);
    BMI_CHECK_HIP(hipEventRecord(h->ev_join, h->side_stream));
  }
  apply_near(h, nrhs, xs, ys, st);
  if (far) BMI_CHECK_HIP(hipStreamWaitEvent(st, h->ev_join, 0));
}

// many right-hand sides, internal vectors RHS-fastest (xs, ys, yf: [super row][rhs]; multipole
// vectors: [level offset + cluster * rows + row][rhs]): each wave's lanes are 64 consecutive
// right-hand sides, so every vector access of the block sweeps is one contiguous row segment
static bool use_rhs_fastest(const bmi_h2* h, int nrhs, bool reduced) {
  if (h->mxf || nrhs < MULTI_MIN_RHS || !multi_t_supported(h->cplx, h->dd, h->cd)) return false;
  return !(reduced && h->op.precision != BMI_PRECISION_FP64 && h->ncoarse > 0);  // no BF16 levels here
}

static void apply_super_t(bmi_h2* h, const Work& w, int nrhs, bool reduced, hipStream_t st) {
  const bmi_space* s = h->s;
  const bool C = h->cplx;
  const int S = C ? 2 : 1;
  const int vd = h->vd;
  const int64_t vcs = static_cast<int64_t>(s->nel) * s->I2;
  const bool far = has_far(h);
  if (far) {
    // the far chain runs on a side stream beside the near-field sweep (own workspace), or after
    // it on the caller's stream
    hipStream_t fs = st;
    if (w.own) {
      if (!h->side_stream) BMI_CHECK_HIP(hipStreamCreateWithFlags(&h->side_stream, hipStreamNonBlocking));
      if (!h->ev_fork) BMI_CHECK_HIP(hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming));
      if (!h->ev_join) BMI_CHECK_HIP(hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming));
      BMI_CHECK_HIP(hipEventRecord(h->ev_fork, st));
      BMI_CHECK_HIP(hipStreamWaitEvent(h->side_stream, h->ev_fork, 0));
      fs = h->side_stream;
    }
    const CouplingOperand co = coupling_operand(h, reduced);
    for (int cc = 0; cc < vd; ++cc) {
      TimedScope ts("matvec_forward", fs);
      launch_forward_t(C, s->npatch, h->L, h->cd, h->np2, h->dd, nrhs, h->d_mom[cc].ptr, h->d_T.ptr, w.xs,
                       cc * vcs, w.fw[cc], h->lvl_off.data(), fs);
    }
    for (int rc = 0; rc < vd; ++rc) {
      // the first component's pass writes every cluster (zero where there are no blocks)
      for (int cc = 0; cc < vd; ++cc)
        BMI_REQUIRE(launch_multi_cpl_t(C, co.fp32, cc > 0 ? h->ntgt : h->ntgt_all, nrhs, h->cd, co.ld,
                                       cc > 0 ? h->d_tgt.ptr : h->d_tgt_all.ptr,
                                       cc > 0 ? h->d_crp.ptr : h->d_crp_all.ptr, h->d_csrc.ptr, co.F,
                                       vd * vd * co.blk, (rc * vd + cc) * co.blk, h->d_lvl_off.ptr,
                                       w.fw[cc], w.bw, cc > 0 ? 1 : 0, fs),
                    BMI_UNSUPPORTED, "coupling block size without an RHS-fastest sweep");
      TimedScope ts("matvec_backward", fs);
      launch_backward_t(C, s->npatch, h->L, h->cd, h->np2, h->dd, nrhs, h->d_momt[rc].ptr, h->d_Tt.ptr, w.bw,
                        h->lvl_off.data(), w.yf, rc * vcs, fs);
    }
    if (w.own) BMI_CHECK_HIP(hipEventRecord(h->ev_join, fs));
  }
  // FP32 near-field blocks with the reduced operator of a mixed handle
  const bool n32 = reduced && h->red1;
  const int64_t dblk = static_cast<int64_t>(h->dd) * h->dd;
  const void* F = n32 ? static_cast<const void*>(h->d_dense32.ptr) : static_cast<const void*>(h->d_dense.ptr);
  for (int rc = 0; rc < vd; ++rc)
    for (int cc = 0; cc < vd; ++cc)
      BMI_REQUIRE(launch_multi_near_t(C, n32, h->ncl_leaf, nrhs, h->dd, h->d_drp.ptr, h->d_dsrc.ptr, F,
                                      vd * vd * dblk, (rc * vd + cc) * dblk, w.xs, cc * vcs, w.ys,
                                      rc * vcs, cc > 0 ? 1 : 0, st),
                  BMI_UNSUPPORTED, "near-field block size without an RHS-fastest sweep");
  if (far && w.own) BMI_CHECK_HIP(hipStreamWaitEvent(st, h->ev_join, 0));
}

// ys = A_super xs on the workspace, through a captured graph when possible.  The graph holds
// only kernel nodes (buffers are cleared by a fill kernel, not a memset node); in-library
// timing needs host synchronisation per kernel and therefore runs eagerly.  A single-RHS product
// (four launches, the longest a few hundred microseconds) also runs eagerly: measured, the next
// kernel on the stream starts about 8 us after a graph ends, more than the launches save.
static void apply_super_workspace(bmi_h2* h, int nrhs, bool reduced, bool tl, hipStream_t st) {
  const Work w = handle_work(h);
  auto run = [&](hipStream_t rs) {
    if (tl)
      apply_super_t(h, w, nrhs, reduced, rs);
    else
      apply_super(h, w, nrhs, w.xs, w.ys, w.yf, reduced, rs);
  };
  if (!tl) autotune(h, h->d_xs.ptr, reduced, st);
  if (!h->use_graphs || timing_enabled() || (nrhs == 1 && !tl)) {
    run(st);
    return;
  }
  const int key = 4 * nrhs + (tl ? 2 : 0) + (reduced ? 1 : 0);
  auto it = h->graphs.find(key);
  if (it == h->graphs.end()) {
    if (!h->cap_stream) BMI_CHECK_HIP(hipStreamCreateWithFlags(&h->cap_stream, hipStreamNonBlocking));
    hipGraph_t g = nullptr;
    BMI_CHECK_HIP(hipStreamBeginCapture(h->cap_stream, hipStreamCaptureModeThreadLocal));
    try {
      run(h->cap_stream);
    } catch (...) {
      (void)hipStreamEndCapture(h->cap_stream, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    BMI_CHECK_HIP(hipStreamEndCapture(h->cap_stream, &g));
    hipGraphExec_t ex = nullptr;
    const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    BMI_CHECK_HIP(e);
    it = h->graphs.emplace(key, ex).first;
  }
  BMI_CHECK_HIP(hipGraphLaunch(it->second, st));
}

// Y = alpha T^T (ys + yf) + beta Y with ys + yf = A_super T X, on the workspace w
static void apply_dof(bmi_h2* h, const Work& w, int nrhs, const double* X, int64_t ldx, double* Y, int64_t ldy,
                      const double* alpha, const double* beta, hipStream_t st, bool reduced) {
  const bmi_space* s = h->s;
  const int64_t N = s->nsuper;
  const bool tl = use_rhs_fastest(h, nrhs, reduced);
  if (tl)
    launch_spmv_tin(h->cplx, s->nsuper, nrhs, s->d_rowptr.ptr, s->d_col.ptr, s->d_val.ptr, X, ldx, w.xs, st);
  else
    launch_spmv(h->cplx, s->nsuper, nrhs, s->d_rowptr.ptr, s->d_col.ptr, s->d_val.ptr, X, ldx, w.xs, N, nullptr,
                nullptr, st);
  if (w.own)
    apply_super_workspace(h, nrhs, reduced, tl, st);
  else if (tl)
    apply_super_t(h, w, nrhs, reduced, st);
  else
    apply_super(h, w, nrhs, w.xs, w.ys, w.yf, reduced, st);
  if (tl)
    launch_spmv_tout(h->cplx, s->ndof, nrhs, s->d_trowptr.ptr, s->d_tcol.ptr, s->d_tval.ptr, w.ys,
                     has_far(h) ? w.yf : nullptr, Y, ldy, alpha, beta, st);
  else
    launch_spmv(h->cplx, s->ndof, nrhs, s->d_trowptr.ptr, s->d_tcol.ptr, s->d_tval.ptr, w.ys, N, Y, ldy, alpha,
                beta, st, has_far(h) ? w.yf : nullptr, true);
}

void h2_apply_device(bmi_h2* h, int nrhs, const double* X, int64_t ldx, double* Y, int64_t ldy, const double* alpha,
                     const double* beta, hipStream_t st, bool reduced) {
  ensure_workspace(h, nrhs, st);
  apply_dof(h, handle_work(h), nrhs, X, ldx, Y, ldy, alpha, beta, st, reduced);
}

// ------------------------------------------------------------------ allocation-free product
// The workspace of bmi_h2_matvec_noalloc lives in this array: one copy per GPU, held until the
// process ends.  In host code `scratch` names a host-side stand-in, never the device memory: the
// device address of each GPU's copy comes from hipGetSymbolAddress (scratch_ptr below).
static __device__ __align__(256) double scratch[BMI_NOALLOC_SCRATCH_BYTES / sizeof(double)];

static std::mutex g_scratch_mu;
static double* g_scratch_ptr[64] = {};

// device address of the current GPU's scratch array, looked up once per device
static double* scratch_ptr() {
  int dev = 0;
  BMI_CHECK_HIP(hipGetDevice(&dev));
  BMI_REQUIRE(dev >= 0 && dev < 64, BMI_UNSUPPORTED, "device ordinal above 63");
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    BMI_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    BMI_REQUIRE(p != nullptr, BMI_DEVICE_ERROR, "no device address for the scratch array");
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  return g_scratch_ptr[dev];
}

// doubles of a region rounded up to 256 bytes
static int64_t scratch_round(int64_t n) { return (n + 31) / 32 * 32; }

// workspace regions of an nrhs-column product of h in the scratch array; 0 if it does not fit
static int64_t noalloc_doubles(const bmi_h2* h, int nrhs) {
  const int S = h->cplx ? 2 : 1;
  const int64_t vec = scratch_round(static_cast<int64_t>(S) * h->s->nsuper * nrhs);
  const int64_t mp = scratch_round(static_cast<int64_t>(S) * h->lvl_total * nrhs);
  return 3 * vec + (h->vd + 1) * mp + (h->mxf ? 2 * mp : 0);
}

static Work noalloc_work(const bmi_h2* h, int nrhs, double* base) {
  const int S = h->cplx ? 2 : 1;
  const int64_t vec = scratch_round(static_cast<int64_t>(S) * h->s->nsuper * nrhs);
  const int64_t mp = scratch_round(static_cast<int64_t>(S) * h->lvl_total * nrhs);
  Work w;
  w.xs = base;
  w.ys = base + vec;
  w.yf = base + 2 * vec;
  w.fw[0] = base + 3 * vec;
  w.fw[1] = h->vd > 1 ? base + 3 * vec + mp : nullptr;
  w.bw = base + 3 * vec + h->vd * mp;
  w.own = false;
  w.z = h->mxf ? base + 3 * vec + (h->vd + 1) * mp : nullptr;
  return w;
}

// in-library timing creates events; it is off for the calling thread while an allocation-free
// product is enqueued
struct NoTiming {
  NoTiming() { tl_timing_off = true; }
  ~NoTiming() { tl_timing_off = false; }
};

int64_t h2_rows(const bmi_h2* h) { return h->s->ndof; }
bool h2_complex(const bmi_h2* h) { return h->cplx; }
bool h2_is_symmetric_real(const bmi_h2* h) { return h->op.kind == BMI_LAPLACE_SL; }
bool h2_is_mixed(const bmi_h2* h) { return h->op.precision != BMI_PRECISION_FP64; }

// ------------------------------------------------------------------ assembly
// Coupling blocks of a symmetric kernel pair up: (level, c2, c1) is the transpose of (level, c1,
// c2) (same kernel values; the two weights multiply in the other order). For single-RHS sweeps
// list one block per pair under its smaller cluster c1 (the upper block (c1, c2) for real
// kernels, the lower block (c2, c1) for complex ones), and per cluster c2 the entries whose
// partials it receives. Leaves the handle unchanged when a block has no partner.
static void build_symmetric_lists(bmi_h2* h, const std::vector<int2>& tgt, const std::vector<int>& crp,
                                  const std::vector<int>& csrc, hipStream_t st) {
  const int nt = static_cast<int>(tgt.size());
  auto find_target = [&](int level, int c) {
    const auto it = std::lower_bound(tgt.begin(), tgt.end(), make_int2(level, c), [](const int2& a, const int2& b) {
      return a.x != b.x ? a.x < b.x : a.y < b.y;
    });
    return (it != tgt.end() && it->x == level && it->y == c) ? static_cast<int>(it - tgt.begin()) : -1;
  };
  std::vector<int2> utgt;
  std::vector<int> urp(1, 0), usrc, ublk;
  std::vector<int2> ucol;  // (target index of the column cluster, upper entry)
  for (int t = 0; t < nt; ++t) {
    const int level = tgt[t].x, c1 = tgt[t].y;
    bool any = false;
    for (int p = crp[t]; p < crp[t + 1]; ++p) {
      const int c2 = csrc[p];
      const int t2 = find_target(level, c2);
      if (c2 == c1 || t2 < 0 || !std::binary_search(csrc.begin() + crp[t2], csrc.begin() + crp[t2 + 1], c1)) return;
      if (c2 < c1) continue;
      if (!any) utgt.push_back(tgt[t]), urp.push_back(urp.back()), any = true;
      ucol.push_back(make_int2(t2, static_cast<int>(usrc.size())));
      usrc.push_back(c2);
      // complex kernels: the lower block (c2, c1), whose row products are the partials and whose
      // column products sum into the rows of c1
      ublk.push_back(h->cplx ? static_cast<int>(std::lower_bound(csrc.begin() + crp[t2], csrc.begin() + crp[t2 + 1], c1) -
                                                csrc.begin())
                             : p);
      urp.back()++;
    }
  }
  // gather order: column targets in (level, cluster) order, entries in list order; each entry's
  // partial is stored at its position in this order
  std::stable_sort(ucol.begin(), ucol.end(), [](const int2& a, const int2& b) { return a.x < b.x; });
  std::vector<int2> gtgt;
  std::vector<int> grp(1, 0), gslot(usrc.size());
  for (size_t q = 0; q < ucol.size(); ++q) {
    const int2& e = ucol[q];
    if (gtgt.empty() || gtgt.back().x != tgt[e.x].x || gtgt.back().y != tgt[e.x].y) {
      gtgt.push_back(tgt[e.x]);
      grp.push_back(grp.back());
    }
    grp.back()++;
    gslot[e.y] = static_cast<int>(q);
  }
  const int S = h->cplx ? 2 : 1;
  h->d_stgt.upload(utgt, st);
  h->d_srp.upload(urp, st);
  h->d_ssrc.upload(usrc, st);
  h->d_sblk.upload(ublk, st);
  h->d_sgtgt.upload(gtgt, st);
  h->d_sgrp.upload(grp, st);
  h->d_sgslot.upload(gslot, st);
  h->d_part.alloc(static_cast<size_t>(S) * usrc.size() * h->cd);
  h->sl = SymList{static_cast<int>(utgt.size()), h->d_stgt.ptr, h->d_srp.ptr, h->d_ssrc.ptr, h->d_sblk.ptr,
                  h->d_part.ptr, h->d_sgslot.ptr, static_cast<int>(gtgt.size()), h->d_sgtgt.ptr, h->d_sgrp.ptr};
  h->sym = true;
  BMI_CHECK_HIP(hipStreamSynchronize(st));
}

// Near-field leaves of a symmetric kernel pair up as well: leaf (c2, c1) is the transpose of
// (c1, c2) (the assembly copies it so).  For single-RHS sweeps list, per row cluster, its
// diagonal leaf and the leaves (c, c2) with c2 > c; each of the latter also gives the transposed
// product for the rows of c2, gathered per column cluster in list order.  Leaves the handle
// unchanged when a leaf has no partner.
static void build_near_symmetric_lists(bmi_h2* h, const std::vector<int>& drp, const std::vector<int>& dsrc,
                                       hipStream_t st) {
  const int ncl = h->ncl_leaf;
  std::vector<int2> tg(ncl), col;  // col: (column cluster, entry)
  std::vector<int> rp(1, 0), src, blk, slot;
  for (int c = 0; c < ncl; ++c) {
    tg[c] = make_int2(0, c);
    for (int p = drp[c]; p < drp[c + 1]; ++p) {
      const int c2 = dsrc[p];
      if (c2 < c) continue;
      if (c2 > c && !std::binary_search(dsrc.begin() + drp[c2], dsrc.begin() + drp[c2 + 1], c)) return;
      if (c2 > c) col.push_back(make_int2(c2, static_cast<int>(src.size())));
      src.push_back(c2);
      blk.push_back(p);
      slot.push_back(-1);
    }
    rp.push_back(static_cast<int>(src.size()));
  }
  std::stable_sort(col.begin(), col.end(), [](const int2& a, const int2& b) { return a.x < b.x; });
  std::vector<int2> gtgt;
  std::vector<int> grp(1, 0);
  for (size_t q = 0; q < col.size(); ++q) {
    if (gtgt.empty() || gtgt.back().y != col[q].x) gtgt.push_back(make_int2(0, col[q].x)), grp.push_back(grp.back());
    grp.back()++;
    slot[col[q].y] = static_cast<int>(q);
  }
  const int S = h->cplx ? 2 : 1;
  h->d_ntgt.upload(tg, st);
  h->d_nrp.upload(rp, st);
  h->d_nsrc.upload(src, st);
  h->d_nblk.upload(blk, st);
  h->d_nslot.upload(slot, st);
  h->d_ngtgt.upload(gtgt, st);
  h->d_ngrp.upload(grp, st);
  h->d_zero.upload(std::vector<int64_t>(1, 0), st);
  h->d_npart.alloc(static_cast<size_t>(S) * std::max<size_t>(col.size(), 1) * h->dd);
  h->nsl = SymList{ncl, h->d_ntgt.ptr, h->d_nrp.ptr, h->d_nsrc.ptr, h->d_nblk.ptr, h->d_npart.ptr, h->d_nslot.ptr,
                   static_cast<int>(gtgt.size()), h->d_ngtgt.ptr, h->d_ngrp.ptr, h->d_zero.ptr};
  h->nsym = true;
  BMI_CHECK_HIP(hipStreamSynchronize(st));
}

static void assemble_h2(bmi_h2* h, hipStream_t st) {
  const bmi_space* s = h->s;
  const bmi_op_desc& op = h->op;
  ensure_device_space(s, st);
  h->vd = s->vd;
  h->cplx = op.kind == BMI_HELMHOLTZ_SL || op.kind == BMI_MAXWELL_SL;
  h->nc = op.kind == BMI_MAXWELL_SL ? 2 : 1;
  h->np = op.n_cheb;
  h->np2 = h->np * h->np;
  h->mcl = op.min_cluster_level;
  h->L = s->M - h->mcl;
  BMI_REQUIRE(h->L >= 0, BMI_UNSUPPORTED, "refinement level below min_cluster_level");
  h->dd = (1 << (2 * h->mcl)) * s->I2;
  h->cd = h->nc * h->np2;
  {
    HostTimedScope hs("assembly_host_block_tree");
    build_block_tree(s, op.eta, h->mcl, h->dense, h->lowrank);
  }
  const int S = h->cplx ? 2 : 1;
  const int vd = h->vd;
  // ---- near field
  const int ne = 1 << (2 * h->mcl);
  const int64_t dblk = static_cast<int64_t>(h->dd) * h->dd;
  h->d_dense.alloc(static_cast<size_t>(S) * vd * vd * dblk * h->dense.size());
  // The single-layer bilinear forms are symmetric (Laplace, Helmholtz: a(u, v) = a(v, u) with the
  // unconjugated product; Maxwell: component block (r, c) of an element pair (e1, e2) is the
  // transpose of block (c, r) of (e2, e1)), and so is the block tree.  Every element pair (e1, e2)
  // whose mirror (e2, e1) is integrated -- all pairs of a leaf (c1, c2) with c1 > c2 whose partner
  // (c2, c1) exists, and the pairs below the diagonal of a leaf (c, c) -- is copied as a transpose
  // after the quadrature instead of integrated again (equal to integrating it to rounding: the
  // same quadrature points, summed in another order).
  const bool sym_kernel = op.kind != BMI_DUMMY;  // the test integrands are not symmetric
  std::vector<int> quad;       // leaves with integrated pairs
  std::vector<int> partner(h->dense.size(), -1);
  for (int k = 0; k < static_cast<int>(h->dense.size()); ++k) {
    const Leaf& lf = h->dense[k];
    if (lf.c1 > lf.c2 && sym_kernel) {
      const auto it = std::lower_bound(h->dense.begin(), h->dense.end(), lf, [&](const Leaf& x, const Leaf&) {
        return x.c1 != lf.c2 ? x.c1 < lf.c2 : x.c2 < lf.c1;
      });
      if (it != h->dense.end() && it->c1 == lf.c2 && it->c2 == lf.c1) partner[k] = static_cast<int>(it - h->dense.begin());
    }
    if (partner[k] < 0) quad.push_back(k);
  }
  auto diag = [&](int k) { return sym_kernel && h->dense[k].c1 == h->dense[k].c2; };
  // sub-block base of element pair (a, b) (row element a, column element b) of leaf k
  auto sub_off = [&](int64_t k, int a, int b) {
    return k * vd * vd * dblk + static_cast<int64_t>(b * s->I2) * h->dd + a * s->I2;
  };
  std::vector<int64_t> pstart(quad.size() + 1, 0);
  for (size_t q = 0; q < quad.size(); ++q) pstart[q + 1] = pstart[q] + (diag(quad[q]) ? ne * (ne + 1) / 2 : ne * ne);
  // the transposed copies: (destination, source) sub-block bases
  std::vector<int64_t> mirror;
  for (int k = 0; k < static_cast<int>(h->dense.size()); ++k)
    for (int b = 0; b < ne; ++b)
      for (int a = 0; a < ne; ++a) {
        if (partner[k] >= 0)
          mirror.push_back(sub_off(k, a, b)), mirror.push_back(sub_off(partner[k], b, a));
        else if (diag(k) && a > b)
          mirror.push_back(sub_off(k, a, b)), mirror.push_back(sub_off(k, b, a));
      }
  // every integrated element pair: class, Duffy rotations and quadrature degree, in parallel
  // over ranges of leaves (each pair has a fixed slot; the host decisions are pure functions of
  // the space)
  std::vector<NFPair> pairs(pstart.back());
  const int ff = farfield_degree(op.kind, s->P);
  {
    HostTimedScope hs("assembly_host_pairs");
    const int64_t nd = static_cast<int64_t>(quad.size());
    const int nth = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(
        std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency())), nd / 256)));
    std::vector<std::array<int, 5>> counts(nth, std::array<int, 5>{0, 0, 0, 0, 0});
    std::vector<std::string> errors(nth);
    auto work = [&](int t) {
      try {
        for (int64_t kq = nd * t / nth; kq < nd * (t + 1) / nth; ++kq) {
          const int64_t k = quad[kq];
          const Leaf& lf = h->dense[k];
          BMI_REQUIRE(lf.level == h->L, BMI_UNSUPPORTED, "dense leaf above the leaf level");
          const bool dg = diag(static_cast<int>(k));
          int64_t slot = pstart[kq];
          for (int b = 0; b < ne; ++b)
            for (int a = 0; a < ne; ++a) {
              if (dg && a > b) continue;
              NFPair& p = pairs[slot++];
              p.e1 = lf.c1 * ne + a;
              p.e2 = lf.c2 * ne + b;
              int cp[3];
              double dist;
              compare_elements(s, p.e1, p.e2, cp, &dist);
              int nf = nearfield_degree(op.kind, s->P, dist, s->M);
              nf = nf >= ff ? nf : ff;
              BMI_REQUIRE(nf < 50, BMI_UNSUPPORTED, "near-field quadrature degree too large");
              p.cls = (cp[2] == 0) ? (nf == ff ? 0 : 1) : cp[2];
              p.rot1 = cp[0];
              p.rot2 = cp[1];
              p.deg = nf;
              p.off = sub_off(k, a, b);
              counts[t][p.cls]++;
            }
        }
      } catch (const std::exception& e) {
        errors[t] = e.what();
      }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nth; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto& th : pool) th.join();
    for (int t = 0; t < nth; ++t) {
      BMI_REQUIRE(errors[t].empty(), BMI_UNSUPPORTED, errors[t]);
      for (int c = 0; c < 5; ++c) h->pair_class_count[c] += counts[t][c];
    }
  }
  h->npairs = static_cast<int64_t>(pairs.size());
  // component block (rc, cc) of leaf k starts at (k*vd*vd + rc*vd + cc) * dblk
  launch_nearfield(s, op, pairs, h->d_dense.ptr, 1, h->dd, vd * dblk, dblk, st);
  if (!mirror.empty()) {
    DevBuf<int64_t> d_mirror;
    d_mirror.upload(mirror, st);
    launch_mirror_blocks(static_cast<int64_t>(mirror.size() / 2), d_mirror.ptr, s->I2, h->dd, dblk, vd, S,
                         h->d_dense.ptr, st);
    BMI_CHECK_HIP(hipStreamSynchronize(st));
  }
  // dense CSR over leaf clusters (leaves sorted by (c1, c2))
  h->ncl_leaf = ncl_at(s, h->L);
  std::vector<int> drp(h->ncl_leaf + 1, 0), dsrc(h->dense.size());
  for (size_t k = 0; k < h->dense.size(); ++k) drp[h->dense[k].c1 + 1]++, dsrc[k] = h->dense[k].c2;
  for (int c = 0; c < h->ncl_leaf; ++c) drp[c + 1] += drp[c];
  h->d_drp.upload(drp, st);
  h->d_dsrc.upload(dsrc, st);
  if (vd == 1 && (op.kind == BMI_LAPLACE_SL || op.kind == BMI_HELMHOLTZ_SL) && h->dd == 36)
    build_near_symmetric_lists(h, drp, dsrc, st);
  // ---- FMM constants
  std::vector<double> T, mom;
  transfer_matrix(h->np, T);
  h->d_T.upload(T, st);
  {
    // T_q(i, j) = T[(q np2 + j) np2 + i]; Tt holds T_q^T in the same layout
    const int n2 = h->np2;
    std::vector<double> Tt(T.size());
    for (int q = 0; q < 4; ++q)
      for (int j = 0; j < n2; ++j)
        for (int i = 0; i < n2; ++i) Tt[(q * n2 + i) * n2 + j] = T[(q * n2 + j) * n2 + i];
    h->d_Tt.upload(Tt, st);
  }
  for (int c = 0; c < vd; ++c) {
    int rows, cols;
    moment_matrix(s->P, h->L, h->mcl, h->np, vd == 2 ? c + 1 : 0, mom, &rows, &cols);
    BMI_REQUIRE(rows == h->cd && cols == h->dd, BMI_DEVICE_ERROR, "moment matrix shape");
    h->d_mom[c].upload(mom, st);
    // Mom is rows x cols column-major; momt = Mom^T, cols x rows column-major
    std::vector<double> momt(mom.size());
    for (int j = 0; j < cols; ++j)
      for (int i = 0; i < rows; ++i) momt[static_cast<size_t>(i) * cols + j] = mom[static_cast<size_t>(j) * rows + i];
    h->d_momt[c].upload(momt, st);
  }
  // level offsets of the multipole arrays
  h->lvl_off.assign(h->L + 2, 0);
  for (int l = 0; l <= h->L; ++l) h->lvl_off[l + 1] = h->lvl_off[l] + static_cast<int64_t>(ncl_at(s, l)) * h->cd;
  h->lvl_total = h->lvl_off[h->L + 1];
  h->d_lvl_off.upload(h->lvl_off, st);
  // ---- coupling blocks (Maxwell: the kernel values of each leaf, see k_mx_coupling)
  h->mxf = op.kind == BMI_MAXWELL_SL;
  if (h->mxf) {
    const double k2re = op.kappa_re * op.kappa_re - op.kappa_im * op.kappa_im, k2im = 2 * op.kappa_re * op.kappa_im;
    const double den = k2re * k2re + k2im * k2im;
    BMI_REQUIRE(den > 0, BMI_INVALID_ARG, "Maxwell operator with wavenumber zero");
    h->mx_qre = -k2re / den, h->mx_qim = k2im / den;
    BMI_REQUIRE(h->np2 <= 256, BMI_UNSUPPORTED, "Maxwell far field with more than 16x16 Chebyshev nodes");
  }
  const int64_t cblk = h->mxf ? static_cast<int64_t>(h->np2) * h->np2 : static_cast<int64_t>(h->cd) * h->cd;
  const int bpl = h->mxf ? 1 : vd * vd;  // stored blocks per leaf
  const int nl = static_cast<int>(h->lowrank.size());
  h->d_coup.alloc(static_cast<size_t>(S) * bpl * cblk * nl);
  if (nl > 0 || h->mxf) {
    std::vector<double> cheb(h->np);
    cheb_points(h->np, cheb.data());
    DevBuf<double> d_cheb;
    d_cheb.upload(cheb, st);
    std::vector<DevBuf<double>> pts(h->L + 1);
    std::vector<double*> ptrs(h->L + 1, nullptr);
    for (int l = 0; l <= h->L; ++l) {
      pts[l].alloc(static_cast<size_t>(ncl_at(s, l)) * h->np2 * CP_STRIDE);
      launch_cluster_points(s, l, h->np, d_cheb.ptr, pts[l].ptr, st);
      ptrs[l] = pts[l].ptr;
    }
    DevBuf<double*> d_ptrs;
    d_ptrs.upload(ptrs, st);
    std::vector<int4> lv(nl);
    for (int k = 0; k < nl; ++k) lv[k] = make_int4(h->lowrank[k].level, h->lowrank[k].c1, h->lowrank[k].c2, 0);
    DevBuf<int4> d_lv;
    d_lv.upload(lv, st);
    if (nl > 0)
      launch_coupling_assemble(op.kind, d_lv.ptr, nl, h->np2, d_ptrs.ptr, op.kappa_re, op.kappa_im, h->d_coup.ptr,
                               st);
    if (h->mxf) {
      // tangents (d/ds, d/dt) of every cluster entry's nodes, in the entry order of the multipole
      // arrays (level offset / rows + cluster)
      const int64_t ne = h->lvl_total / h->cd;
      h->h_mx_tan.assign(static_cast<size_t>(ne) * h->np2 * 6, 0.0);
      for (int l = 0; l <= h->L; ++l) {
        const int ncl = ncl_at(s, l);
        std::vector<double> hp(static_cast<size_t>(ncl) * h->np2 * CP_STRIDE);
        BMI_CHECK_HIP(hipMemcpyAsync(hp.data(), pts[l].ptr, hp.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        BMI_CHECK_HIP(hipStreamSynchronize(st));
        const int64_t e0 = h->lvl_off[l] / h->cd;
        for (int64_t q = 0; q < static_cast<int64_t>(ncl) * h->np2; ++q)
          for (int x = 0; x < 6; ++x) h->h_mx_tan[(e0 * h->np2 + q) * 6 + x] = hp[q * CP_STRIDE + 4 + x];
      }
      h->d_mx_tan.upload(h->h_mx_tan, st);
    }
    if (op.precision == BMI_PRECISION_MIXED && vd == 1 && !h->cplx) {
      // node tables of every level at the cluster offsets of the multipole arrays
      h->d_fly_pts.alloc(static_cast<size_t>(h->lvl_total));
      h->d_fly_org.alloc(static_cast<size_t>(h->lvl_total / h->cd));
      for (int l = 0; l <= h->L; ++l)
        launch_fly_points(pts[l].ptr, ncl_at(s, l), h->np2, h->d_fly_pts.ptr + h->lvl_off[l],
                          h->d_fly_org.ptr + h->lvl_off[l] / h->cd, st);
      h->fly = true;
    }
    BMI_CHECK_HIP(hipStreamSynchronize(st));
  }
  // targets (level, cluster) and CSR over the low-rank leaves (sorted by level, c1, c2)
  std::vector<int2> tgt;
  std::vector<int> crp(1, 0), csrc(nl);
  for (int k = 0; k < nl; ++k) {
    const Leaf& lf = h->lowrank[k];
    if (tgt.empty() || tgt.back().x != lf.level || tgt.back().y != lf.c1) {
      tgt.push_back(make_int2(lf.level, lf.c1));
      crp.push_back(crp.back());
    }
    crp.back()++;
    csrc[k] = lf.c2;
  }
  h->ntgt = static_cast<int>(tgt.size());
  h->d_tgt.upload(tgt, st);
  h->d_crp.upload(crp, st);
  h->d_csrc.upload(csrc, st);
  {
    std::vector<int2> all;
    std::vector<int> rp_all(1, 0);
    size_t t = 0;
    for (int l = 0; l <= h->L; ++l)
      for (int c = 0; c < (s->npatch << (2 * l)); ++c) {
        all.push_back(make_int2(l, c));
        if (t < tgt.size() && tgt[t].x == l && tgt[t].y == c) {
          rp_all.push_back(crp[t + 1]);
          ++t;
        } else {
          rp_all.push_back(rp_all.back());
        }
      }
    BMI_REQUIRE(t == tgt.size(), BMI_DEVICE_ERROR, "coupling targets out of level order");
    h->ntgt_all = static_cast<int>(all.size());
    h->d_tgt_all.upload(all, st);
    h->d_crp_all.upload(rp_all, st);
    if (h->mxf) {
      std::vector<int> order(all.size());
      for (size_t q = 0; q < order.size(); ++q) order[q] = static_cast<int>(q);
      std::stable_sort(order.begin(), order.end(),
                       [&](int a, int b) { return rp_all[a + 1] - rp_all[a] > rp_all[b + 1] - rp_all[b]; });
      h->d_mx_order.upload(order, st);
    }
  }
  if (vd == 1 && (op.kind == BMI_LAPLACE_SL || op.kind == BMI_HELMHOLTZ_SL) && sym_sweep_supported(h->cplx, h->dd, h->cd))
    build_symmetric_lists(h, tgt, crp, csrc, st);
  // FP32 coupling copy: MIXED pads the columns to whole 16-byte row vectors (the sweeps read rows
  // as vectors); MIXED_BF16 keeps the FP64 layout, shared with its BF16 blocks
  h->ld32 = h->mxf ? h->np2 : op.precision == BMI_PRECISION_MIXED ? fp32_coupling_ld(h->cplx, h->cd) : h->cd;
  const int cdim = h->mxf ? h->np2 : h->cd;  // rows of a stored block
  if (op.precision != BMI_PRECISION_FP64)
    h->d_coup32.alloc(static_cast<size_t>(S) * bpl * h->ld32 * cdim * nl);
  BMI_CHECK_HIP(hipStreamSynchronize(st));
  if (op.precision != BMI_PRECISION_FP64 && h->d_coup.n)
    convert_blocks_to_fp32(h->d_coup.ptr, h->d_coup32.ptr, static_cast<int64_t>(nl) * bpl, cdim, h->ld32, h->cplx,
                           st);
  h->red1 = op.precision == BMI_PRECISION_MIXED && vd == 1;
  if (h->red1) {
    h->d_dense32.alloc(h->d_dense.n);
    if (h->d_dense.n) convert_to_fp32(h->d_dense.ptr, h->d_dense32.ptr, h->d_dense.n, st);
  }
  if (op.precision == BMI_PRECISION_MIXED_BF16 && !h->mxf) {  // Maxwell: FP32 at every level
    // blocks are sorted by level: the coarse ones (levels 0 and 1) come first
    h->ncoarse = 0;
    while (h->ncoarse < nl && h->lowrank[h->ncoarse].level <= 1) ++h->ncoarse;
    const int64_t n16 = static_cast<int64_t>(S) * vd * vd * cblk * h->ncoarse;
    h->d_coup16.alloc(static_cast<size_t>(n16));
    convert_to_bf16(h->d_coup.ptr, h->d_coup16.ptr, n16, st);
  }
  BMI_CHECK_HIP(hipStreamSynchronize(st));
}

}  // namespace bmi

// ====================================================================== C ABI
#define BMI_API_BEGIN try {
#define BMI_API_END                                   \
  return BMI_OK;                                      \
  }                                                   \
  catch (const bmi::Error& e) {                       \
    bmi::set_error(e.what());                         \
    return e.status;                                  \
  }                                                   \
  catch (const std::bad_alloc&) {                     \
    bmi::set_error("host allocation failed");         \
    return BMI_OUT_OF_MEMORY;                         \
  }                                                   \
  catch (const std::exception& e) {                   \
    bmi::set_error(e.what());                         \
    return BMI_DEVICE_ERROR;                          \
  }

extern "C" {

const char* bmi_last_error(void) { return bmi::g_last_error.c_str(); }
#ifndef BMI_SOURCE_DIGEST
#define BMI_SOURCE_DIGEST "unknown"
#endif
const char* bmi_version(void) { return "bembel_amd 0.1 (gfx950) sources " BMI_SOURCE_DIGEST; }

bmi_status bmi_op_desc_default(int kind, bmi_op_desc* out) {
  BMI_API_BEGIN
  BMI_REQUIRE(out, BMI_INVALID_ARG, "null descriptor");
  BMI_REQUIRE(kind >= 0 && kind <= 3, BMI_INVALID_ARG, "unknown operator kind");
  out->kind = kind;
  out->kappa_re = 0.0;
  out->kappa_im = 0.0;
  out->n_cheb = 9;
  out->eta = 1.6;
  out->min_cluster_level = 1;
  out->precision = BMI_PRECISION_FP64;
  out->dummy_fn = 0;
  BMI_API_END
}

bmi_status bmi_geometry_load(const char* path, const double* scale, bmi_geometry** out) {
  BMI_API_BEGIN
  BMI_REQUIRE(path && out, BMI_INVALID_ARG, "null argument");
  *out = nullptr;
  std::ifstream f(path);
  BMI_REQUIRE(f.good(), BMI_IO_ERROR, std::string("cannot open geometry file ") + path);
  // GEOPDE .dat format (Geometry/GeometryIO.hpp:25-122): 4 comment rows, info row, then per
  // patch: name, degrees, control point counts, two knot vectors, four M x N matrices
  std::string line;
  for (int i = 0; i < 5; ++i) BMI_REQUIRE(std::getline(f, line), BMI_IO_ERROR, "truncated geometry header");
  int info[5];
  {
    std::istringstream is(line);
    for (int i = 0; i < 5; ++i) BMI_REQUIRE(static_cast<bool>(is >> info[i]), BMI_IO_ERROR, "bad info row");
  }
  auto g = std::make_unique<bmi_geometry>();
  for (int p = 0; p < info[2]; ++p) {
    int deg[2], ncp[2];
    BMI_REQUIRE(std::getline(f, line), BMI_IO_ERROR, "truncated patch");
    BMI_REQUIRE(std::getline(f, line), BMI_IO_ERROR, "truncated patch");
    {
      std::istringstream is(line);
      is >> deg[0] >> deg[1];
    }
    BMI_REQUIRE(std::getline(f, line), BMI_IO_ERROR, "truncated patch");
    {
      std::istringstream is(line);
      is >> ncp[0] >> ncp[1];
    }
    std::vector<double> k[2];
    for (int kk = 0; kk < 2; ++kk) {
      BMI_REQUIRE(std::getline(f, line), BMI_IO_ERROR, "truncated knots");
      std::istringstream is(line);
      const int nk = deg[kk] + ncp[kk] + 1;
      k[kk].resize(nk);
      for (int i = 0; i < nk; ++i) BMI_REQUIRE(static_cast<bool>(is >> k[kk][i]), BMI_IO_ERROR, "bad knot vector");
      int uniq = 1;
      for (int i = 1; i < nk; ++i)
        if (k[kk][i] > k[kk][i - 1] + 1e-6) ++uniq;
      BMI_REQUIRE(uniq == 2, BMI_UNSUPPORTED, "patches with interior knots are not supported");
    }
    const int N = ncp[0], Mr = ncp[1];
    std::vector<double> mat[4];
    for (int m = 0; m < 4; ++m) {
      BMI_REQUIRE(std::getline(f, line), BMI_IO_ERROR, "truncated control points");
      std::istringstream is(line);
      mat[m].resize(static_cast<size_t>(N) * Mr);
      for (int i = 0; i < Mr * N; ++i) BMI_REQUIRE(static_cast<bool>(is >> mat[m][i]), BMI_IO_ERROR, "bad control points");
    }
    Patch pt;
    pt.pdx = static_cast<int>(k[0].size()) - N;
    pt.pdy = static_cast<int>(k[1].size()) - Mr;
    BMI_REQUIRE(pt.pdx == N && pt.pdy == Mr, BMI_IO_ERROR, "inconsistent Bezier patch");
    pt.cp.resize(4 * static_cast<size_t>(N) * Mr);
    for (int m = 0; m < 4; ++m)
      for (int i = 0; i < N; ++i)
        for (int j = 0; j < Mr; ++j) pt.cp[(i * Mr + j) * 4 + m] = mat[m][j * N + i];
    if (scale)
      for (int q = 0; q < N * Mr; ++q)
        for (int m = 0; m < 3; ++m) pt.cp[4 * q + m] *= scale[m];
    g->patches.push_back(std::move(pt));
  }
  BMI_REQUIRE(!g->patches.empty(), BMI_IO_ERROR, "geometry without patches");
  *out = g.release();
  BMI_API_END
}

bmi_status bmi_geometry_destroy(bmi_geometry* g) {
  delete g;
  return BMI_OK;
}

bmi_status bmi_geometry_num_patches(const bmi_geometry* g, int* n) {
  BMI_API_BEGIN
  BMI_REQUIRE(g && n, BMI_INVALID_ARG, "null argument");
  *n = static_cast<int>(g->patches.size());
  BMI_API_END
}

bmi_status bmi_geometry_surface_point(const bmi_geometry* g, int patch, double s, double t, double out[12]) {
  BMI_API_BEGIN
  BMI_REQUIRE(g && out, BMI_INVALID_ARG, "null argument");
  BMI_REQUIRE(patch >= 0 && patch < static_cast<int>(g->patches.size()), BMI_INVALID_ARG, "patch out of range");
  patch_surface_point(g->patches[patch], s, t, out);
  BMI_API_END
}

bmi_status bmi_linear_form_points(const bmi_space* s, int64_t* npoints, double* out) {
  BMI_API_BEGIN
  BMI_REQUIRE(s && npoints, BMI_INVALID_ARG, "null argument");
  *npoints = static_cast<int64_t>(s->nel) * (s->P + 2) * (s->P + 2);
  if (out) linear_form_points(s, out);
  BMI_API_END
}

bmi_status bmi_space_create(const bmi_geometry* g, int M, int P, int form, bmi_space** out) {
  BMI_API_BEGIN
  BMI_REQUIRE(g && out, BMI_INVALID_ARG, "null argument");
  BMI_REQUIRE(M >= 0 && M <= 10, BMI_INVALID_ARG, "refinement level out of range");
  BMI_REQUIRE(P >= 0 && P <= 8, BMI_INVALID_ARG, "polynomial degree out of range");
  BMI_REQUIRE(form == BMI_FORM_DISCONTINUOUS || form == BMI_FORM_DIVCONFORMING, BMI_INVALID_ARG, "unknown form");
  BMI_REQUIRE(form == BMI_FORM_DISCONTINUOUS || P >= 1, BMI_INVALID_ARG, "DivConforming spaces need P >= 1");
  *out = nullptr;
  auto s = std::make_unique<bmi_space>();
  s->geo = g;
  s->M = M;
  s->P = P;
  s->I2 = (P + 1) * (P + 1);
  s->form = form;
  s->vd = form == BMI_FORM_DIVCONFORMING ? 2 : 1;
  s->pd = g->patches[0].pdx;
  for (const auto& p : g->patches)
    if (p.pdx != s->pd || p.pdy != s->pd) s->pd = -1;
  build_mesh(s.get());
  build_transformation(s.get());
  *out = s.release();
  BMI_API_END
}

bmi_status bmi_space_destroy(bmi_space* s) {
  delete s;
  return BMI_OK;
}

bmi_status bmi_space_info(const bmi_space* s, int64_t info[8]) {
  BMI_API_BEGIN
  BMI_REQUIRE(s && info, BMI_INVALID_ARG, "null argument");
  info[0] = s->nel;
  info[1] = s->nvert;
  info[2] = s->I2;
  info[3] = s->vd;
  info[4] = s->nsuper;
  info[5] = s->ndof;
  info[6] = static_cast<int64_t>(s->val.size());
  info[7] = s->M;
  BMI_API_END
}

bmi_status bmi_space_elements(const bmi_space* s, int* vertices, int* patch, double* llc, double* midpoint,
                              double* radius) {
  BMI_API_BEGIN
  BMI_REQUIRE(s, BMI_INVALID_ARG, "null space");
  if (vertices) std::copy(s->vert.begin(), s->vert.end(), vertices);
  if (patch) std::copy(s->elem_patch.begin(), s->elem_patch.end(), patch);
  if (llc) std::copy(s->elem_llc.begin(), s->elem_llc.end(), llc);
  if (midpoint) std::copy(s->cmid[s->M].begin(), s->cmid[s->M].end(), midpoint);
  if (radius) std::copy(s->crad[s->M].begin(), s->crad[s->M].end(), radius);
  BMI_API_END
}

bmi_status bmi_space_clusters(const bmi_space* s, int level, double* midpoint, double* radius, int* count) {
  BMI_API_BEGIN
  BMI_REQUIRE(s && count, BMI_INVALID_ARG, "null argument");
  BMI_REQUIRE(level >= 0 && level <= s->M, BMI_INVALID_ARG, "level out of range");
  *count = static_cast<int>(s->crad[level].size());
  if (midpoint) std::copy(s->cmid[level].begin(), s->cmid[level].end(), midpoint);
  if (radius) std::copy(s->crad[level].begin(), s->crad[level].end(), radius);
  BMI_API_END
}

bmi_status bmi_space_transformation(const bmi_space* s, int* rowptr, int* col, double* val) {
  BMI_API_BEGIN
  BMI_REQUIRE(s, BMI_INVALID_ARG, "null space");
  if (rowptr) std::copy(s->rowptr.begin(), s->rowptr.end(), rowptr);
  if (col) std::copy(s->col.begin(), s->col.end(), col);
  if (val) std::copy(s->val.begin(), s->val.end(), val);
  BMI_API_END
}

bmi_status bmi_pair_info(const bmi_space* s, int npairs, const int* pairs, int* out) {
  BMI_API_BEGIN
  BMI_REQUIRE(s && (npairs == 0 || (pairs && out)), BMI_INVALID_ARG, "null argument");
  const int ff = farfield_degree(BMI_LAPLACE_SL, s->P);
  for (int k = 0; k < npairs; ++k) {
    const int e1 = pairs[2 * k], e2 = pairs[2 * k + 1];
    BMI_REQUIRE(e1 >= 0 && e1 < s->nel && e2 >= 0 && e2 < s->nel, BMI_INVALID_ARG, "element out of range");
    int cp[3];
    double dist;
    compare_elements(s, e1, e2, cp, &dist);
    int nf = nearfield_degree(BMI_LAPLACE_SL, s->P, dist, s->M);
    out[4 * k] = cp[2];
    out[4 * k + 1] = cp[0];
    out[4 * k + 2] = cp[1];
    out[4 * k + 3] = nf >= ff ? nf : ff;
  }
  BMI_API_END
}

static void check_op(const bmi_space* s, const bmi_op_desc* op) {
  BMI_REQUIRE(s && op, BMI_INVALID_ARG, "null argument");
  BMI_REQUIRE(op->kind >= 0 && op->kind <= 3, BMI_INVALID_ARG, "unknown operator kind");
  BMI_REQUIRE((op->kind == BMI_MAXWELL_SL) == (s->form == BMI_FORM_DIVCONFORMING), BMI_INVALID_ARG,
              "operator and space form do not match");
  BMI_REQUIRE(s->pd > 0, BMI_UNSUPPORTED, "patches of different degrees are not supported on the device");
  if (op->kind == BMI_MAXWELL_SL)
    BMI_REQUIRE(op->kappa_re != 0.0 || op->kappa_im != 0.0, BMI_INVALID_ARG, "Maxwell needs a nonzero wavenumber");
}

bmi_status bmi_pair_blocks(const bmi_space* s, const bmi_op_desc* op, int npairs, const int* pairs, double* out,
                           void* stream) {
  BMI_API_BEGIN
  check_op(s, op);
  BMI_REQUIRE(npairs >= 0 && (npairs == 0 || (pairs && out)), BMI_INVALID_ARG, "null argument");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int n = s->vd * s->I2;
  const bool cplx = op->kind == BMI_HELMHOLTZ_SL || op->kind == BMI_MAXWELL_SL;
  const int ff = farfield_degree(op->kind, s->P);
  std::vector<NFPair> v(npairs);
  for (int k = 0; k < npairs; ++k) {
    NFPair& p = v[k];
    p.e1 = pairs[2 * k];
    p.e2 = pairs[2 * k + 1];
    BMI_REQUIRE(p.e1 >= 0 && p.e1 < s->nel && p.e2 >= 0 && p.e2 < s->nel, BMI_INVALID_ARG, "element out of range");
    int cp[3];
    double dist;
    compare_elements(s, p.e1, p.e2, cp, &dist);
    int nf = nearfield_degree(op->kind, s->P, dist, s->M);
    nf = nf >= ff ? nf : ff;
    BMI_REQUIRE(nf < 50, BMI_UNSUPPORTED, "near-field quadrature degree too large");
    p.cls = (cp[2] == 0) ? (nf == ff ? 0 : 1) : cp[2];
    p.rot1 = cp[0];
    p.rot2 = cp[1];
    p.deg = nf;
    p.off = static_cast<int64_t>(k) * n * n;
  }
  DevBuf<double> d_out;
  d_out.alloc(static_cast<size_t>(npairs) * n * n * (cplx ? 2 : 1));
  if (npairs) {
    launch_nearfield(s, *op, v, d_out.ptr, n, 1, static_cast<int64_t>(s->I2) * n, s->I2, st);
    BMI_CHECK_HIP(hipMemcpyAsync(out, d_out.ptr, d_out.bytes(), hipMemcpyDeviceToHost, st));
    BMI_CHECK_HIP(hipStreamSynchronize(st));
  }
  BMI_API_END
}

bmi_status bmi_integrate_jobs(const bmi_space* s, const bmi_op_desc* op, int njobs, const int* jobs, int qdeg,
                              double* out, void* stream) {
  BMI_API_BEGIN
  BMI_REQUIRE(s && op && (njobs == 0 || (jobs && out)), BMI_INVALID_ARG, "null argument");
  BMI_REQUIRE(op->kind == BMI_DUMMY && s->P == 0, BMI_UNSUPPORTED, "integrate jobs support the test operator on P = 0");
  BMI_REQUIRE(qdeg >= 0 && qdeg < 50, BMI_INVALID_ARG, "quadrature degree out of range");
  hipStream_t st = static_cast<hipStream_t>(stream);
  std::vector<int> j(jobs, jobs + 5 * njobs);
  for (int k = 0; k < njobs; ++k) {
    BMI_REQUIRE(j[5 * k] >= 0 && j[5 * k] <= 4, BMI_INVALID_ARG, "unknown integrate rule");
    BMI_REQUIRE(j[5 * k + 1] >= 0 && j[5 * k + 1] < s->nel && j[5 * k + 3] >= 0 && j[5 * k + 3] < s->nel,
                BMI_INVALID_ARG, "element out of range");
  }
  DevBuf<double> d_out;
  d_out.alloc(njobs);
  if (njobs) {
    launch_integrate_jobs(s, *op, j, qdeg, d_out.ptr, st);
    BMI_CHECK_HIP(hipMemcpyAsync(out, d_out.ptr, d_out.bytes(), hipMemcpyDeviceToHost, st));
    BMI_CHECK_HIP(hipStreamSynchronize(st));
  }
  BMI_API_END
}

bmi_status bmi_dense_assemble(const bmi_space* s, const bmi_op_desc* op, int super_only, double* out, int mem,
                              void* stream) {
  BMI_API_BEGIN
  check_op(s, op);
  BMI_REQUIRE(out, BMI_INVALID_ARG, "null output");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const bool cplx = op->kind == BMI_HELMHOLTZ_SL || op->kind == BMI_MAXWELL_SL;
  const int S = cplx ? 2 : 1;
  const int64_t N = s->nsuper, D = s->ndof;
  const int ff = farfield_degree(op->kind, s->P);
  std::vector<NFPair> v;
  v.reserve(static_cast<size_t>(s->nel) * s->nel);
  for (int e2 = 0; e2 < s->nel; ++e2)
    for (int e1 = 0; e1 < s->nel; ++e1) {
      NFPair p;
      p.e1 = e1;
      p.e2 = e2;
      int cp[3];
      double dist;
      compare_elements(s, e1, e2, cp, &dist);
      int nf = nearfield_degree(op->kind, s->P, dist, s->M);
      nf = nf >= ff ? nf : ff;
      BMI_REQUIRE(nf < 50, BMI_UNSUPPORTED, "near-field quadrature degree too large");
      p.cls = (cp[2] == 0) ? (nf == ff ? 0 : 1) : cp[2];
      p.rot1 = cp[0];
      p.rot2 = cp[1];
      p.deg = nf;
      p.off = static_cast<int64_t>(e2 * s->I2) * N + e1 * s->I2;
      v.push_back(p);
    }
  DevBuf<double> A;
  A.alloc(static_cast<size_t>(S) * N * N);
  // super-space rows and columns are component-major: component c starts at c * nel * I2
  const int64_t vcs = static_cast<int64_t>(s->nel) * s->I2;
  launch_nearfield(s, *op, v, A.ptr, 1, N, vcs, vcs * N, st);
  if (super_only) {
    BMI_CHECK_HIP(hipMemcpyAsync(out, A.ptr, A.bytes(), mem == BMI_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st));
    BMI_CHECK_HIP(hipStreamSynchronize(st));
  } else {
    // A T (N x D) then T^T (A T) (D x D), both with the sparse kernels
    DevBuf<double> AT, ATt, R;
    ensure_device_space(s, st);
    // (A T)^T = T^T A^T: compute column j of A T as A * T(:, j) via spmv on A^T rows is awkward;
    // use the identity T^T A T = (T^T (T^T A^T)^T) with explicit transposes on the host side.
    std::vector<double> hA(static_cast<size_t>(S) * N * N);
    BMI_CHECK_HIP(hipMemcpyAsync(hA.data(), A.ptr, A.bytes(), hipMemcpyDeviceToHost, st));
    BMI_CHECK_HIP(hipStreamSynchronize(st));
    std::vector<double> hAT(static_cast<size_t>(S) * N * D, 0.0);
    for (int64_t r = 0; r < N; ++r)
      for (int k = s->rowptr[r]; k < s->rowptr[r + 1]; ++k) {
        const int64_t c = s->col[k];
        const double v0 = s->val[k];
        const double* src = &hA[S * r * N];
        double* dst = &hAT[S * c * N];
        for (int64_t i = 0; i < S * N; ++i) dst[i] += v0 * src[i];
      }
    std::vector<double> hR(static_cast<size_t>(S) * D * D, 0.0);
    for (int64_t j = 0; j < D; ++j)
      for (int64_t r = 0; r < N; ++r)
        for (int k = s->rowptr[r]; k < s->rowptr[r + 1]; ++k) {
          const int64_t row = s->col[k];
          for (int q = 0; q < S; ++q) hR[S * (j * D + row) + q] += s->val[k] * hAT[S * (j * N + r) + q];
        }
    BMI_CHECK_HIP(hipMemcpyAsync(out, hR.data(), hR.size() * sizeof(double),
                                 mem == BMI_MEM_DEVICE ? hipMemcpyHostToDevice : hipMemcpyHostToHost, st));
    BMI_CHECK_HIP(hipStreamSynchronize(st));
  }
  BMI_API_END
}

bmi_status bmi_h2_assemble(const bmi_space* s, const bmi_op_desc* op, void* stream, bmi_h2** out) {
  BMI_API_BEGIN
  check_op(s, op);
  BMI_REQUIRE(out, BMI_INVALID_ARG, "null output");
  BMI_REQUIRE(op->n_cheb >= 1 && op->n_cheb <= 12, BMI_INVALID_ARG, "n_cheb out of range");
  BMI_REQUIRE(op->eta > 0, BMI_INVALID_ARG, "eta must be positive");
  BMI_REQUIRE(op->min_cluster_level >= 0 && op->min_cluster_level <= 3, BMI_INVALID_ARG, "min_cluster_level out of range");
  BMI_REQUIRE(op->precision >= BMI_PRECISION_FP64 && op->precision <= BMI_PRECISION_MIXED_BF16, BMI_INVALID_ARG,
              "unknown precision");
  *out = nullptr;
  auto h = std::make_unique<bmi_h2>();
  h->s = s;
  h->op = *op;
  assemble_h2(h.get(), static_cast<hipStream_t>(stream));
  *out = h.release();
  BMI_API_END
}

bmi_status bmi_h2_estimate_bytes(const bmi_space* s, const bmi_op_desc* op, int nrhs, int64_t* bytes) {
  BMI_API_BEGIN
  check_op(s, op);
  BMI_REQUIRE(bytes && nrhs >= 0, BMI_INVALID_ARG, "bad argument");
  std::vector<Leaf> dense, lowrank;
  const int mcl = op->min_cluster_level;
  BMI_REQUIRE(s->M - mcl >= 0, BMI_UNSUPPORTED, "refinement level below min_cluster_level");
  build_block_tree(s, op->eta, mcl, dense, lowrank);
  const int64_t S = (op->kind == BMI_HELMHOLTZ_SL || op->kind == BMI_MAXWELL_SL) ? 2 : 1;
  const int64_t vd = s->vd, nc = op->kind == BMI_MAXWELL_SL ? 2 : 1;
  const int64_t dd = (int64_t(1) << (2 * mcl)) * s->I2, cd = nc * op->n_cheb * op->n_cheb;
  const int L = s->M - mcl;
  const int64_t lvl_total = int64_t(s->npatch) * (((int64_t(1) << (2 * (L + 1))) - 1) / 3) * cd;
  // Maxwell stores the kernel values of a coupling leaf once (np2 x np2), plus node tangents
  const bool mxf = op->kind == BMI_MAXWELL_SL;
  const int64_t np2 = int64_t(op->n_cheb) * op->n_cheb;
  const int64_t cpl = mxf ? np2 * np2 : vd * vd * cd * cd;  // stored entries per coupling leaf
  int64_t b = S * (vd * vd * int64_t(dense.size()) * dd * dd + int64_t(lowrank.size()) * cpl) * 8;
  const int64_t ld32 = op->precision == BMI_PRECISION_MIXED ? fp32_coupling_ld(S == 2, static_cast<int>(cd)) : cd;
  if (op->precision != BMI_PRECISION_FP64)
    b += S * int64_t(lowrank.size()) * (mxf ? np2 * np2 : vd * vd * ld32 * cd) * 4;
  if (mxf) b += lvl_total / cd * np2 * 6 * 8 + std::max(nrhs, 1) * S * 2 * lvl_total * 8;
  if (op->precision == BMI_PRECISION_MIXED && vd == 1) {
    b += S * int64_t(dense.size()) * dd * dd * 4;                     // FP32 near field
    if (S == 1) b += lvl_total * 16 + lvl_total / cd * 32;            // node tables
  }
  b += (int64_t(dense.size()) + int64_t(lowrank.size())) * 8;                           // CSR indices
  if (cd > 0) b += lvl_total / cd * 12;                                                  // all-cluster list
  b += std::max(nrhs, 1) * S * (3 * int64_t(s->nsuper) + (vd + 1) * lvl_total) * 8;   // workspace
  b += int64_t(s->nel) * 25 * 16 * 8 + int64_t(s->val.size()) * 24;                    // tables, T
  *bytes = b;
  BMI_API_END
}

bmi_status bmi_h2_destroy(bmi_h2* h) {
  delete h;
  return BMI_OK;
}

bmi_status bmi_h2_info(const bmi_h2* h, int64_t info[16]) {
  BMI_API_BEGIN
  BMI_REQUIRE(h && info, BMI_INVALID_ARG, "null argument");
  std::fill(info, info + 16, 0);
  info[0] = static_cast<int64_t>(h->dense.size() + h->lowrank.size());
  info[1] = static_cast<int64_t>(h->dense.size());
  info[2] = static_cast<int64_t>(h->lowrank.size());
  info[3] = h->npairs;
  info[4] = h->dd;
  info[5] = h->cd;
  info[6] = h->L + 1;
  info[7] = h->s->ndof;
  info[8] = static_cast<int64_t>(h->bytes());
  info[9] = h->cplx ? 1 : 0;
  info[10] = h->op.precision;
  for (int c = 0; c < 5; ++c) info[11 + c] = h->pair_class_count[c];
  BMI_API_END
}

bmi_status bmi_h2_leaves(const bmi_h2* h, int* leaves) {
  BMI_API_BEGIN
  BMI_REQUIRE(h && leaves, BMI_INVALID_ARG, "null argument");
  size_t k = 0;
  for (const auto& l : h->dense) leaves[k++] = 2, leaves[k++] = l.level, leaves[k++] = l.c1, leaves[k++] = l.c2;
  for (const auto& l : h->lowrank) leaves[k++] = 1, leaves[k++] = l.level, leaves[k++] = l.c1, leaves[k++] = l.c2;
  BMI_API_END
}

namespace bmi {
// component block (r, c) of the Maxwell coupling leaf k in the reference layout (cd x cd,
// column-major, complex): [[K o (t_r(x_i) . t_c(y_j)), 0], [0, -K / kappa^2]] from the leaf's
// kernel values Kh (np2 x np2 complex) and the tangents of both clusters' nodes
static void mx_expand_leaf(const bmi_h2* h, int k, int r, int c, const double* Kh, double* F) {
  const int n2 = h->np2, cd = h->cd;
  const Leaf& lf = h->lowrank[k];
  const int64_t e1 = h->lvl_off[lf.level] / cd + lf.c1, e2 = h->lvl_off[lf.level] / cd + lf.c2;
  std::fill(F, F + 2 * static_cast<int64_t>(cd) * cd, 0.0);
  for (int j = 0; j < n2; ++j)
    for (int i = 0; i < n2; ++i) {
      const double* a = &h->h_mx_tan[(e1 * n2 + i) * 6 + 3 * r];
      const double* b = &h->h_mx_tan[(e2 * n2 + j) * 6 + 3 * c];
      const double d = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
      const double kr = Kh[2 * (j * n2 + i)], ki = Kh[2 * (j * n2 + i) + 1];
      const int64_t o = static_cast<int64_t>(j) * cd + i, o1 = static_cast<int64_t>(j + n2) * cd + i + n2;
      F[2 * o] = kr * d, F[2 * o + 1] = ki * d;
      F[2 * o1] = h->mx_qre * kr - h->mx_qim * ki, F[2 * o1 + 1] = h->mx_qre * ki + h->mx_qim * kr;
    }
}
}  // namespace bmi

bmi_status bmi_h2_leaf_matrix(const bmi_h2* h, int leaf, int r, int c, double* F) {
  BMI_API_BEGIN
  BMI_REQUIRE(h && F, BMI_INVALID_ARG, "null argument");
  const int nd = static_cast<int>(h->dense.size()), nl = static_cast<int>(h->lowrank.size());
  BMI_REQUIRE(leaf >= 0 && leaf < nd + nl, BMI_INVALID_ARG, "leaf out of range");
  BMI_REQUIRE(r >= 0 && r < h->vd && c >= 0 && c < h->vd, BMI_INVALID_ARG, "component out of range");
  const int S = h->cplx ? 2 : 1;
  if (leaf < nd) {
    const int64_t blk = static_cast<int64_t>(h->dd) * h->dd;
    const double* src = h->d_dense.ptr + S * ((static_cast<int64_t>(leaf) * h->vd * h->vd + r * h->vd + c) * blk);
    BMI_CHECK_HIP(hipMemcpy(F, src, S * blk * sizeof(double), hipMemcpyDeviceToHost));
  } else if (h->mxf) {
    const int64_t kb = static_cast<int64_t>(h->np2) * h->np2;
    std::vector<double> Kh(2 * kb);
    BMI_CHECK_HIP(hipMemcpy(Kh.data(), h->d_coup.ptr + 2 * (leaf - nd) * kb, 2 * kb * sizeof(double),
                            hipMemcpyDeviceToHost));
    mx_expand_leaf(h, leaf - nd, r, c, Kh.data(), F);
  } else {
    const int64_t blk = static_cast<int64_t>(h->cd) * h->cd;
    const double* src =
        h->d_coup.ptr + S * ((static_cast<int64_t>(leaf - nd) * h->vd * h->vd + r * h->vd + c) * blk);
    BMI_CHECK_HIP(hipMemcpy(F, src, S * blk * sizeof(double), hipMemcpyDeviceToHost));
  }
  BMI_API_END
}

bmi_status bmi_h2_export_leaves(const bmi_h2* h, double* data) {
  BMI_API_BEGIN
  BMI_REQUIRE(h && data, BMI_INVALID_ARG, "null argument");
  if (h->d_dense.n) BMI_CHECK_HIP(hipMemcpy(data, h->d_dense.ptr, h->d_dense.bytes(), hipMemcpyDeviceToHost));
  if (h->mxf && h->d_coup.n) {
    // Maxwell: every leaf's component blocks expanded to the reference layout
    std::vector<double> Kh(h->d_coup.n);
    BMI_CHECK_HIP(hipMemcpy(Kh.data(), h->d_coup.ptr, h->d_coup.bytes(), hipMemcpyDeviceToHost));
    const int64_t kb = static_cast<int64_t>(h->np2) * h->np2, blk = static_cast<int64_t>(h->cd) * h->cd;
    double* out = data + h->d_dense.n;
    for (int k = 0; k < static_cast<int>(h->lowrank.size()); ++k)
      for (int r = 0; r < h->vd; ++r)
        for (int c = 0; c < h->vd; ++c)
          mx_expand_leaf(h, k, r, c, Kh.data() + 2 * k * kb, out + 2 * ((static_cast<int64_t>(k) * h->vd + r) * h->vd + c) * blk);
  } else if (h->d_coup.n) {
    BMI_CHECK_HIP(hipMemcpy(data + h->d_dense.n, h->d_coup.ptr, h->d_coup.bytes(), hipMemcpyDeviceToHost));
  }
  BMI_API_END
}

bmi_status bmi_h2_matvec(const bmi_h2* hc, int nrhs, const double* X, int64_t ldx, double* Y, int64_t ldy,
                         const double* alpha, const double* beta, int mem, void* stream) {
  BMI_API_BEGIN
  BMI_REQUIRE(hc && X && Y, BMI_INVALID_ARG, "null argument");
  BMI_REQUIRE(nrhs >= 0, BMI_INVALID_ARG, "negative nrhs");
  bmi_h2* h = const_cast<bmi_h2*>(hc);
  const int64_t D = h->s->ndof;
  BMI_REQUIRE(ldx >= D && ldy >= D, BMI_INVALID_ARG, "leading dimension smaller than the number of DOFs");
  if (nrhs == 0) return BMI_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int S = h->cplx ? 2 : 1;
  if (mem == BMI_MEM_DEVICE) {
    // device operands: enqueued on the stream, returns without waiting
    h2_apply_device(h, nrhs, X, ldx, Y, ldy, alpha, beta, st, h2_is_mixed(h));
    BMI_CHECK_HIP(hipGetLastError());
    return BMI_OK;
  } else {
    DevBuf<double> dx, dy;
    dx.alloc(static_cast<size_t>(S) * D * nrhs);
    dy.alloc(static_cast<size_t>(S) * D * nrhs);
    for (int k = 0; k < nrhs; ++k) {
      BMI_CHECK_HIP(hipMemcpyAsync(dx.ptr + S * D * k, X + S * ldx * k, S * D * sizeof(double), hipMemcpyHostToDevice, st));
      BMI_CHECK_HIP(hipMemcpyAsync(dy.ptr + S * D * k, Y + S * ldy * k, S * D * sizeof(double), hipMemcpyHostToDevice, st));
    }
    h2_apply_device(h, nrhs, dx.ptr, D, dy.ptr, D, alpha, beta, st, h2_is_mixed(h));
    for (int k = 0; k < nrhs; ++k)
      BMI_CHECK_HIP(hipMemcpyAsync(Y + S * ldy * k, dy.ptr + S * D * k, S * D * sizeof(double), hipMemcpyDeviceToHost, st));
  }
  BMI_CHECK_HIP(hipStreamSynchronize(st));
  BMI_API_END
}

bmi_status bmi_h2_matvec_noalloc(const bmi_h2* hc, int nrhs, const double* X, int64_t ldx, double* Y, int64_t ldy,
                                 const double* alpha, const double* beta, void* stream) {
  BMI_API_BEGIN
  BMI_REQUIRE(hc && X && Y && alpha, BMI_INVALID_ARG, "null argument");
  BMI_REQUIRE(nrhs >= 0, BMI_INVALID_ARG, "negative nrhs");
  bmi_h2* h = const_cast<bmi_h2*>(hc);
  const int64_t D = h->s->ndof;
  BMI_REQUIRE(ldx >= D && ldy >= D, BMI_INVALID_ARG, "leading dimension smaller than the number of DOFs");
  if (nrhs == 0) return BMI_OK;
  const int64_t need = noalloc_doubles(h, nrhs) * static_cast<int64_t>(sizeof(double));
  if (need > static_cast<int64_t>(BMI_NOALLOC_SCRATCH_BYTES))
    throw Error(BMI_INVALID_ARG, "bmi_h2_matvec_noalloc: the product of " + std::to_string(nrhs) +
                                     " columns needs " + std::to_string(need) + " bytes of workspace, above the " +
                                     std::to_string(static_cast<int64_t>(BMI_NOALLOC_SCRATCH_BYTES)) +
                                     "-byte scratch array (use fewer columns or bmi_h2_matvec)");
  NoTiming nt;
  const Work w = noalloc_work(h, nrhs, scratch_ptr());
  apply_dof(h, w, nrhs, X, ldx, Y, ldy, alpha, beta, static_cast<hipStream_t>(stream), h2_is_mixed(h));
  BMI_CHECK_HIP(hipGetLastError());
  BMI_API_END
}

bmi_status bmi_h2_matvec_super(const bmi_h2* hc, int nrhs, const double* Xs, double* Ys, void* stream) {
  BMI_API_BEGIN
  BMI_REQUIRE(hc && Xs && Ys && nrhs >= 0, BMI_INVALID_ARG, "bad argument");
  bmi_h2* h = const_cast<bmi_h2*>(hc);
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (nrhs == 0) return BMI_OK;
  ensure_workspace(h, nrhs, st);
  apply_super(h, handle_work(h), nrhs, Xs, Ys, h->d_yf.ptr, h2_is_mixed(h), st);
  const int S = h->cplx ? 2 : 1;
  if (has_far(h)) device_add(Ys, h->d_yf.ptr, static_cast<int64_t>(S) * h->s->nsuper * nrhs, st);
  BMI_CHECK_HIP(hipStreamSynchronize(st));
  BMI_API_END
}

bmi_status bmi_synchronize(void* stream) {
  BMI_API_BEGIN
  BMI_CHECK_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  BMI_API_END
}

bmi_status bmi_h2_use_graphs(bmi_h2* h, int enable) {
  BMI_API_BEGIN
  BMI_REQUIRE(h, BMI_INVALID_ARG, "null handle");
  h->use_graphs = enable != 0;
  if (!h->use_graphs) h->drop_graphs();
  BMI_API_END
}

bmi_status bmi_h2_set_sweep(bmi_h2* h, int mode) {
  BMI_API_BEGIN
  BMI_REQUIRE(h, BMI_INVALID_ARG, "null handle");
  BMI_REQUIRE(mode >= -1 && mode <= 2, BMI_INVALID_ARG, "sweep mode must be -1, 0, 1 or 2");
  BMI_REQUIRE(mode < 1 || h->sym, BMI_UNSUPPORTED, "no upper-block sweep for this operator or block size");
  BMI_REQUIRE(mode != 2 || h->nsym, BMI_UNSUPPORTED, "no upper near-field sweep for this operator or block size");
  h->sweep_mode = mode;
  // choose again at the next product
  h->leaf_variant = h->cpl_variant = -1;
  h->fused_leaf_variant = h->fused_cpl_variant = -1;
  h->fly_variant = -1;
  h->sym_fused = 0;
  h->overlap = 0;
  h->drop_graphs();
  BMI_API_END
}

bmi_status bmi_h2_set_overlap(bmi_h2* h, int mode) {
  BMI_API_BEGIN
  BMI_REQUIRE(h, BMI_INVALID_ARG, "null handle");
  BMI_REQUIRE(mode >= -1 && mode <= 1, BMI_INVALID_ARG, "overlap mode must be -1, 0 or 1");
  h->overlap_mode = mode;
  // choose again at the next product
  h->leaf_variant = h->cpl_variant = -1;
  h->fused_leaf_variant = h->fused_cpl_variant = -1;
  h->fly_variant = -1;
  h->sym_fused = 0;
  h->overlap = 0;
  h->drop_graphs();
  BMI_API_END
}

bmi_status bmi_timing_enable(int enable) {
  bmi::g_timing = enable != 0;
  return BMI_OK;
}
bmi_status bmi_timing_reset(void) {
  std::lock_guard<std::mutex> g(bmi::g_tmu);
  bmi::resolve_pending(false);
  bmi::g_times.clear();
  return BMI_OK;
}
bmi_status bmi_timing_get(const char* name, double* total_ms, int64_t* launches) {
  BMI_API_BEGIN
  BMI_REQUIRE(name && total_ms && launches, BMI_INVALID_ARG, "null argument");
  std::lock_guard<std::mutex> g(bmi::g_tmu);
  bmi::resolve_pending(true);
  auto it = bmi::g_times.find(name);
  *total_ms = it == bmi::g_times.end() ? 0.0 : it->second.first;
  *launches = it == bmi::g_times.end() ? 0 : it->second.second;
  BMI_API_END
}

}  // extern "C"
