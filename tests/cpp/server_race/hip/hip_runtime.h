// CPU stand-in for the few HIP runtime calls the serving runtime makes (race testing only)
#pragma once
#include <cstddef>
#include <cstdlib>
#include <cstring>
typedef int hipError_t;
enum { hipSuccess = 0, hipErrorOutOfMemory = 2 };
typedef struct ihipStream_t* hipStream_t;
typedef struct ihipEvent_t* hipEvent_t;
enum { hipStreamNonBlocking = 1 };
enum hipMemcpyKind { hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice, hipMemcpyHostToHost };
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipMalloc(void** p, size_t n);
template <class T> hipError_t hipMalloc(T** p, size_t n) { return hipMalloc(reinterpret_cast<void**>(p), n); }
hipError_t hipFree(void* p);
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t);
hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t);
const char* hipGetErrorString(hipError_t);
