// CPU stand-ins for the HIP calls made while a space's device state is set up, and for the
// far-field fill kernel: real buffers and random delays, so ThreadSanitizer and
// AddressSanitizer see the threading pattern of the library's own setup code.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <random>
#include <thread>

#include "bmi_internal.h"

static void nap(int max_us) {
  thread_local std::mt19937 g(std::hash<std::thread::id>()(std::this_thread::get_id()));
  std::this_thread::sleep_for(std::chrono::microseconds(g() % (max_us + 1)));
}
struct ihipStream_t { int x; };
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) { *s = new ihipStream_t{1}; return 0; }
hipError_t hipStreamDestroy(hipStream_t s) { delete s; return 0; }
hipError_t hipStreamSynchronize(hipStream_t) { nap(100); return 0; }
hipError_t hipMalloc(void** p, size_t n) { nap(20); *p = std::malloc(n ? n : 1); return 0; }
hipError_t hipFree(void* p) { std::free(p); return 0; }
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t) {
  std::memcpy(d, s, n);
  return 0;
}
hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t) {
  std::memset(d, v, n);
  return 0;
}
const char* hipGetErrorString(hipError_t) { return "stub"; }

namespace bmi {
// the fill kernel: entry k of the table of n^2 points per element is n * 100000 + fs * 1000 + k % 997
void launch_ffield_table(const bmi_space* s, int n, int fs, double* out, hipStream_t) {
  const int64_t total = static_cast<int64_t>(s->nel) * n * n * fs;
  for (int64_t k = 0; k < total; ++k) out[k] = n * 100000.0 + fs * 1000.0 + static_cast<double>(k % 997);
  nap(200);
}
}  // namespace bmi
