// CPU stand-ins for the HIP runtime and the bembel_amd C ABI used by the serving runtime:
// random delays, real buffers, so ThreadSanitizer / AddressSanitizer see the threading pattern.
#include <hip/hip_runtime.h>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <random>
#include <string>
#include <mutex>
#include <thread>
#include "bembel_amd.h"
static void nap(int max_us) {
  thread_local std::mt19937 g(std::hash<std::thread::id>()(std::this_thread::get_id()));
  std::this_thread::sleep_for(std::chrono::microseconds(g() % (max_us + 1)));
}
// every device call of a running server must come from one thread (its device thread); calls
// during teardown (after the driver set g_teardown) are exempt
std::atomic<bool> g_teardown{false};
std::atomic<int> g_wrong_thread{0};
static std::mutex g_owner_mu;
static std::thread::id g_owner;
void reset_device_owner() {
  std::lock_guard<std::mutex> lk(g_owner_mu);
  g_owner = std::thread::id();
}
// device calls made after the driver stopped waiting on the server (bmi_server_destroy returned):
// none may happen, the device thread is joined before the server is freed
std::atomic<bool> g_destroyed{false};
std::atomic<int> g_after_destroy{0};
static void device_call() {
  if (g_destroyed.load()) g_after_destroy.fetch_add(1);
  if (g_teardown.load()) return;
  std::lock_guard<std::mutex> lk(g_owner_mu);
  if (g_owner == std::thread::id()) g_owner = std::this_thread::get_id();
  else if (g_owner != std::this_thread::get_id()) g_wrong_thread.fetch_add(1);
}
struct ihipStream_t { int x; };
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) { device_call(); nap(50); *s = new ihipStream_t{1}; return 0; }
hipError_t hipStreamDestroy(hipStream_t s) { device_call(); delete s; return 0; }
hipError_t hipStreamSynchronize(hipStream_t) { device_call(); nap(200); return 0; }
hipError_t hipMalloc(void** p, size_t n) { device_call(); *p = std::malloc(n ? n : 1); return 0; }
hipError_t hipFree(void* p) { device_call(); std::free(p); return 0; }
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t) {
  device_call();
  std::memcpy(d, s, n);
  return 0;
}
hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t) {
  device_call();
  std::memset(d, v, n);
  return 0;
}
const char* hipGetErrorString(hipError_t) { return "stub"; }
struct bmi_geometry { int k; };
struct bmi_space { int64_t ndof, nsuper; };
struct bmi_h2 { int64_t ndof; int cplx; };
static thread_local std::string g_err;
namespace bmi { void set_last_error(const std::string& m) { g_err = m; } }
extern "C" {
const char* bmi_last_error(void) { return g_err.c_str(); }
bmi_status bmi_geometry_load(const char*, const double*, bmi_geometry** g) { nap(300); *g = new bmi_geometry{1}; return BMI_OK; }
bmi_status bmi_geometry_destroy(bmi_geometry* g) { delete g; return BMI_OK; }
bmi_status bmi_space_create(const bmi_geometry*, int M, int P, int, bmi_space** s) {
  nap(300);
  *s = new bmi_space{int64_t(6) * (int64_t(1) << M) * (int64_t(1) << M) * (P + 1), int64_t(6) << (2 * M)};
  return BMI_OK;
}
bmi_status bmi_space_destroy(bmi_space* s) { delete s; return BMI_OK; }
bmi_status bmi_space_info(const bmi_space* s, int64_t info[8]) { for (int k = 0; k < 8; ++k) info[k] = 0; info[4] = s->nsuper; info[5] = s->ndof; return BMI_OK; }
bmi_status bmi_h2_estimate_bytes(const bmi_space* s, const bmi_op_desc*, int, int64_t* b) { *b = s->ndof * 1000; return BMI_OK; }
bmi_status bmi_h2_assemble(const bmi_space* s, const bmi_op_desc* op, void*, bmi_h2** h) {
  device_call();
  nap(2000);
  *h = new bmi_h2{s->ndof, op->kind != 0};
  return BMI_OK;
}
bmi_status bmi_h2_destroy(bmi_h2* h) { device_call(); delete h; return BMI_OK; }
bmi_status bmi_h2_use_graphs(bmi_h2*, int) { return BMI_OK; }
bmi_status bmi_rhs_assemble(const bmi_space* s, int fn, const double* p, int nrhs, double, double* out, int, void*) {
  device_call();
  nap(200);
  const int S = fn == 0 ? 1 : 2;
  for (int64_t i = 0; i < S * s->ndof * nrhs; ++i) out[i] = p[(i % nrhs) * 3];
  return BMI_OK;
}
// a solve that needs 45 steps: X = B / 4 after a first call shorter than that (time slice),
// X = B / 2 when it converges; a resumed call (zero_guess 0) starts from X
static bmi_status solve(const bmi_h2* h, int nrhs, const double* B, double* X, int maxit, int zero, int* it,
                        double* rs) {
  device_call();
  nap(3000);
  const int S = h->cplx ? 2 : 1;
  const bool first = zero != 0;
  const bool conv = !first || maxit >= 45;
  for (int64_t i = 0; i < S * h->ndof * nrhs; ++i) X[i] = (conv ? 0.5 : 0.25) * B[i];
  for (int c = 0; c < nrhs; ++c) it[c] = conv ? 10 + c : maxit, rs[c] = conv ? 1e-11 : 1e-3;
  return conv ? BMI_OK : BMI_NOT_CONVERGED;
}
bmi_status bmi_solve_gmres(const bmi_h2* h, int nrhs, const double* B, double* X, double, int, int maxit, int zero,
                           int* it, double* rs, int, void*) { return solve(h, nrhs, B, X, maxit, zero, it, rs); }
bmi_status bmi_solve_cg(const bmi_h2* h, int nrhs, const double* B, double* X, double, int maxit, int zero, int* it,
                        double* rs, int, void*) { return solve(h, nrhs, B, X, maxit, zero, it, rs); }
// a GMRES task that finishes after a random number of polls, some of which find its flags not
// there yet (0) and some enqueue work (2); finish writes X = B / 2
struct bmi_gmres_task {
  const bmi_h2* h;
  int nrhs;
  const double* B;
  double* X;
  int polls;
};
bmi_status bmi_gmres_start(const bmi_h2* h, int nrhs, const double* B, double* X, double, int, int, int, void*,
                           bmi_gmres_task** out) {
  device_call();
  nap(500);
  thread_local std::mt19937 g(std::hash<std::thread::id>()(std::this_thread::get_id()) + 7);
  *out = new bmi_gmres_task{h, nrhs, B, X, 2 + static_cast<int>(g() % 6)};
  return BMI_OK;
}
bmi_status bmi_gmres_poll(bmi_gmres_task* t, int* done) {
  device_call();
  nap(100);
  if (t->polls > 0) {
    --t->polls;
    *done = (t->polls % 2) ? 0 : 2;
  } else {
    *done = 1;
  }
  return BMI_OK;
}
bmi_status bmi_gmres_finish(bmi_gmres_task* t, int* it, double* rs) {
  device_call();
  nap(500);
  const int S = t->h->cplx ? 2 : 1;
  for (int64_t i = 0; i < S * t->h->ndof * t->nrhs; ++i) t->X[i] = 0.5 * t->B[i];
  for (int c = 0; c < t->nrhs; ++c) {
    if (it) it[c] = 10 + c;
    if (rs) rs[c] = 1e-11;
  }
  delete t;
  return BMI_OK;
}
}
extern "C" bmi_status bmi_op_desc_default(int kind, bmi_op_desc* d) {
  d->kind = kind; d->kappa_re = 0; d->kappa_im = 0; d->n_cheb = 9; d->eta = 1.6; d->min_cluster_level = 1;
  d->precision = 0; d->dummy_fn = 0; return BMI_OK;
}
