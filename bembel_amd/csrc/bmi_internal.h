// Internal declarations of the bembel_amd library (host setup, device kernels, C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdlib>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/bembel_amd.h"

namespace bmi {

// Error carrying a bmi_status; converted to a status code at the C ABI.
struct Error : std::runtime_error {
  bmi_status status;
  Error(bmi_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};
#define BMI_CHECK_HIP(expr)                                                              \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      throw ::bmi::Error(e_ == hipErrorOutOfMemory ? BMI_OUT_OF_MEMORY : BMI_DEVICE_ERROR, \
                         std::string(#expr) + ": " + hipGetErrorString(e_));             \
  } while (0)
#define BMI_REQUIRE(cond, status, msg) \
  do {                                 \
    if (!(cond)) throw ::bmi::Error(status, msg); \
  } while (0)

// ------------------------------------------------------------------ host geometry
struct Patch {
  int pdx = 0, pdy = 0;    // Bernstein functions per direction (degree + 1)
  std::vector<double> cp;  // 4*pdx*pdy homogeneous control points, index 4*(i*pdy+j)+k
};

void bernstein(int p, double x, double* out);
void bernstein_der(int p, double x, double* out);
void patch_eval(const Patch& p, double s, double t, double out[3]);
void patch_surface_point(const Patch& p, double s, double t, double out[12]);

// ------------------------------------------------------------------ device buffer
// BMI_POISON_ALLOC=1: every floating-point buffer of the library starts filled with 0xFF bytes (a
// NaN in every double and float), so a kernel that reads values it has not written gives NaN on
// every run instead of only when the allocator hands back memory that held such bit patterns.
// Index buffers are left alone (a poisoned index would turn the same mistake into a fault).
inline bool poison_allocations() {
  static const bool on = [] {
    const char* v = std::getenv("BMI_POISON_ALLOC");
    return v != nullptr && v[0] != '\0' && v[0] != '0';
  }();
  return on;
}
inline void poison_fill(void* p, size_t bytes) {
  if (!poison_allocations() || bytes == 0) return;
  BMI_CHECK_HIP(hipMemsetAsync(p, 0xFF, bytes, nullptr));
  BMI_CHECK_HIP(hipStreamSynchronize(nullptr));
}

template <typename T>
struct DevBuf {
  T* ptr = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    release();
    if (count == 0) return;
    BMI_CHECK_HIP(hipMalloc(&ptr, count * sizeof(T)));
    n = count;
    if constexpr (std::is_floating_point<T>::value) poison_fill(ptr, count * sizeof(T));
  }
  void upload(const std::vector<T>& v, hipStream_t st) {
    alloc(v.size());
    if (n) BMI_CHECK_HIP(hipMemcpyAsync(ptr, v.data(), n * sizeof(T), hipMemcpyHostToDevice, st));
  }
  size_t bytes() const { return n * sizeof(T); }
};

// Far-field quadrature table of a space (surface points of the (deg+1)^2 rule on every
// element, fs doubles per point: FF_STRIDE, or FF_STRIDE_TAN with the two tangents).  Written
// once, then only read until the space is destroyed.
constexpr int FF_STRIDE = 8;
constexpr int FF_STRIDE_TAN = 16;
struct FFTable {
  int deg = -1, fs = 0;
  DevBuf<double> buf;
};

}  // namespace bmi

struct bmi_geometry {
  std::vector<bmi::Patch> patches;
};

// Near-field element pair work item (one Galerkin block).
struct NFPair {
  int e1, e2;
  int cls;   // 0 regular/far rule, 1 regular/tensor rule, 2 identical, 3 edge, 4 vertex
  int rot1, rot2;
  int deg;   // quadrature degree (points per direction = deg + 1)
  int64_t off;  // scalar offset of the (0,0) entry of the block in the output array
};

struct bmi_space {
  const bmi_geometry* geo = nullptr;
  int M = 0, P = 0, I2 = 0, vd = 1, form = 0, npatch = 0, n = 1, nel = 0, pd = 0;
  // mesh (Morton element order)
  std::vector<int> vert;  // 4 per element, canonical vertex ids
  int nvert = 0;
  std::vector<double> points;  // 3 per vertex
  std::vector<int> elem_patch;
  std::vector<double> elem_llc;  // 2 per element
  std::vector<std::vector<double>> cmid, crad;  // per level 0..M (M = elements)
  // transformation matrix T (nsuper x ndof) and T^T, CSR
  int nsuper = 0, ndof = 0;
  std::vector<int> rowptr, col;
  std::vector<double> val;
  std::vector<int> trowptr, tcol;
  std::vector<double> tval;
  // Device copies, made on the first device use under dev_mu and complete (the upload stream
  // synchronised) before dev_ready is set; from then on only read, so calls on different
  // handles that share the space may run concurrently.
  mutable std::mutex dev_mu;
  mutable bool dev_ready = false;
  mutable bmi::DevBuf<double> d_cp;       // all patches, 4*pd*pd each
  mutable bmi::DevBuf<int> d_elem_patch;
  mutable bmi::DevBuf<double> d_elem_llc;
  mutable bmi::DevBuf<int> d_rowptr, d_col, d_trowptr, d_tcol;
  mutable bmi::DevBuf<double> d_val, d_tval;
  // far-field tables, one per (degree, stride), added under dev_mu and never freed or
  // reallocated before the space is destroyed
  mutable std::vector<std::unique_ptr<bmi::FFTable>> ffields;
};

namespace bmi {

// ------------------------------------------------------------------ host setup
void build_mesh(bmi_space* s);
void build_transformation(bmi_space* s);
int farfield_degree(int kind, int P);
int nearfield_degree(int kind, int P, double dist, int level);
void compare_elements(const bmi_space* s, int e1, int e2, int out[3], double* dist);
// points of the (P+2)^2-point rule of the linear forms, 3 doubles per point (host)
void linear_form_points(const bmi_space* s, double* out);

struct Leaf {
  int cc, level, c1, c2;
};
// BlockClusterTree (H2Matrix/BlockClusterTree.hpp): dense leaves and low-rank leaves
void build_block_tree(const bmi_space* s, double eta, int mcl, std::vector<Leaf>& dense,
                      std::vector<Leaf>& lowrank);

// FMM constant matrices (H2Matrix/H2Multipole.hpp), all column-major
void cheb_points(int n, double* x);
void transfer_matrix(int n, std::vector<double>& T);  // n^2 x 4n^2
void moment_matrix(int P, int cl, int cr, int np, int comp, std::vector<double>& out,
                   int* rows, int* cols);

// ------------------------------------------------------------------ device side
// Both return only when the device data are complete; safe to call from several threads.
void ensure_device_space(const bmi_space* s, hipStream_t st);
const FFTable* ensure_ffield(const bmi_space* s, int deg, hipStream_t st, bool tangents);
// enqueues the kernel that fills a far-field table of n^2 points per element (bmi_nearfield.hip)
void launch_ffield_table(const bmi_space* s, int n, int fs, double* out, hipStream_t st);
// Runs the near-field pair kernels; out is a device array of scalars (double or
// interleaved complex).  Entry (i, j) of a block lands at off + i*rs + j*cs (in scalars);
// comp_rs/comp_cs are the offsets between vector components (Maxwell).
void launch_nearfield(const bmi_space* s, const bmi_op_desc& op, std::vector<NFPair>& pairs,
                      double* out, int64_t rs, int64_t cs, int64_t comp_rs, int64_t comp_cs,
                      hipStream_t st);
// integrate0..4 jobs for the Duffy unit tests
void launch_integrate_jobs(const bmi_space* s, const bmi_op_desc& op, const std::vector<int>& jobs,
                           int qdeg, double* out, hipStream_t st);

// kernel timing (HIP events on the launch stream)
struct TimedScope {
  const char* name;
  hipStream_t st;
  hipEvent_t a = nullptr, b = nullptr;
  TimedScope(const char* n, hipStream_t s);
  ~TimedScope();
};
// host phases under the same registry (wall clock, when timing is enabled)
struct HostTimedScope {
  const char* name;
  double t0 = -1.0;
  explicit HostTimedScope(const char* n);
  ~HostTimedScope();
};
bool timing_enabled();
void set_last_error(const std::string& m);

}  // namespace bmi
