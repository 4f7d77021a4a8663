// Device state of a discrete space: the copies of the geometry, the mesh and T that the kernels
// read, and the far-field quadrature tables (one per degree and stride).  Made on first use by
// whichever thread gets there first; afterwards only read.  Host code only: the table itself is
// filled by a kernel in bmi_nearfield.hip.
#include "bmi_internal.h"

namespace bmi {

static void upload_space_locked(const bmi_space* s, hipStream_t st) {
  if (s->dev_ready) return;
  std::vector<double> cp;
  for (const auto& p : s->geo->patches) cp.insert(cp.end(), p.cp.begin(), p.cp.end());
  s->d_cp.upload(cp, st);
  s->d_elem_patch.upload(s->elem_patch, st);
  s->d_elem_llc.upload(s->elem_llc, st);
  s->d_rowptr.upload(s->rowptr, st);
  s->d_col.upload(s->col, st);
  s->d_val.upload(s->val, st);
  s->d_trowptr.upload(s->trowptr, st);
  s->d_tcol.upload(s->tcol, st);
  s->d_tval.upload(s->tval, st);
  // the host vectors are pageable and other threads will read the copies from other streams:
  // both need the copies finished before the flag is set
  BMI_CHECK_HIP(hipStreamSynchronize(st));
  s->dev_ready = true;
}

void ensure_device_space(const bmi_space* s, hipStream_t st) {
  std::lock_guard<std::mutex> g(s->dev_mu);
  upload_space_locked(s, st);
}

// One table per (degree, stride): a table with tangents (Maxwell) also serves operators that
// need none.  The fill kernel has finished before the table is published, so a thread on another
// stream that finds it reads complete data; tables live until the space is destroyed, so a
// pointer handed out here stays valid while the space does.
const FFTable* ensure_ffield(const bmi_space* s, int deg, hipStream_t st, bool tangents) {
  std::lock_guard<std::mutex> g(s->dev_mu);
  upload_space_locked(s, st);
  for (const auto& t : s->ffields)
    if (t->deg == deg && (t->fs == FF_STRIDE_TAN || !tangents)) return t.get();
  const int fs = tangents ? FF_STRIDE_TAN : FF_STRIDE;
  const int n = deg + 1;
  auto tab = std::make_unique<FFTable>();
  tab->deg = deg;
  tab->fs = fs;
  tab->buf.alloc(static_cast<size_t>(s->nel) * n * n * fs);
  launch_ffield_table(s, n, fs, tab->buf.ptr, st);
  BMI_CHECK_HIP(hipStreamSynchronize(st));
  s->ffields.push_back(std::move(tab));
  return s->ffields.back().get();
}

}  // namespace bmi
