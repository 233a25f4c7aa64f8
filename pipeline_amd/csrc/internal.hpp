// internal.hpp -- shared helpers of the C ABI implementation: error reporting
// and launch-shape checks.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/dpcull_hip.h"

namespace dpc {

// Sets the calling thread's last-error text and returns code.
int set_error(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

inline int hip_fail(hipError_t e, const char* what)
{
  return set_error(DPC_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

inline hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

// Error injection for tests (dpc_debug_inject_error): the checked HIP call
// that brings the countdown to zero reports the injected error instead of its
// own result.  Off (a negative countdown) unless a test arms it.
hipError_t injected(hipError_t e);

// Launch check after a kernel launch: reports launch-configuration errors.
inline int check_launch(const char* what)
{
  hipError_t e = injected(hipGetLastError());
  if (e != hipSuccess) return hip_fail(e, what);
  return DPC_OK;
}

inline bool aligned(const void* p, size_t a) { return (reinterpret_cast<uintptr_t>(p) % a) == 0; }

// Number of compute units of the current device (cached per device).
int device_cu_count();

constexpr int kWave = 64;

}  // namespace dpc

#define DPC_HIP_TRY(call, what)                       \
  do {                                                \
    hipError_t e_ = dpc::injected(call);              \
    if (e_ != hipSuccess) return dpc::hip_fail(e_, what); \
  } while (0)
