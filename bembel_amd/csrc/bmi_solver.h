// Solvers and right-hand sides on the device (bmi_solver.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

struct bmi_h2;

namespace bmi {

void convert_to_fp32(const double* src, float* dst, size_t n, hipStream_t st);
void device_fill(double* p, int64_t n, double v, hipStream_t st);
// Y = alpha * H * X + beta * Y with device pointers; reduced: the reduced-precision operator of a
// mixed handle (the FP64 operator otherwise)
void h2_apply_device(bmi_h2* h, int nrhs, const double* X, int64_t ldx, double* Y, int64_t ldy, const double* alpha,
                     const double* beta, hipStream_t st, bool reduced);
int64_t h2_rows(const bmi_h2* h);
bool h2_complex(const bmi_h2* h);
bool h2_is_symmetric_real(const bmi_h2* h);
bool h2_is_mixed(const bmi_h2* h);

}  // namespace bmi
