// host_runtime.cpp -- the host-only stand-in declared in hip/hip_runtime_api.h
// (this directory), plus the few library entry points the serving loop calls,
// on host memory: the matrix uploads and scatters move the matrices (so a
// scene's matrices can be checked after a run), the culls compute nothing.
// Thread-safe: the serving loop calls it from its scheduler, its issuers and
// its callers' threads at once.
#include "hip/hip_runtime_api.h"

#include <atomic>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>

#include "batch_types.hpp"
#include "dpcull_hip.h"

struct HostStream {};
struct HostEvent {
  std::atomic<long long> done_at_ns{0};  // 0: never recorded
};

namespace {

std::mutex g_mu;
std::map<const char*, size_t> g_pinned;  // pinned blocks: base -> bytes

long long now_ns()
{
  return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

}  // namespace

const char* hipGetErrorString(hipError_t e) { return e == hipSuccess ? "no error" : "host runtime error"; }
hipError_t hipGetLastError() { return hipSuccess; }
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetDevice(int* device)
{
  *device = 0;
  return hipSuccess;
}
hipError_t hipMalloc(void** p, size_t bytes)
{
  *p = std::calloc(1, bytes ? bytes : 1);
  return *p ? hipSuccess : hipErrorInvalidValue;
}
hipError_t hipFree(void* p)
{
  std::free(p);
  return hipSuccess;
}
hipError_t hipHostMalloc(void** p, size_t bytes, unsigned)
{
  *p = std::calloc(1, bytes ? bytes : 1);
  if (!*p) return hipErrorInvalidValue;
  std::lock_guard<std::mutex> lock(g_mu);
  g_pinned[static_cast<const char*>(*p)] = bytes;
  return hipSuccess;
}
hipError_t hipHostFree(void* p)
{
  {
    std::lock_guard<std::mutex> lock(g_mu);
    g_pinned.erase(static_cast<const char*>(p));
  }
  std::free(p);
  return hipSuccess;
}
hipError_t hipMemcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind)
{
  if (bytes) std::memmove(dst, src, bytes);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind, hipStream_t)
{
  return hipMemcpy(dst, src, bytes, kind);
}
hipError_t hipStreamCreateWithFlags(hipStream_t* stream, unsigned)
{
  *stream = new HostStream;
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t stream)
{
  delete stream;
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipEventCreateWithFlags(hipEvent_t* event, unsigned)
{
  *event = new HostEvent;
  return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t event)
{
  delete event;
  return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t event, hipStream_t)
{
  event->done_at_ns.store(now_ns() + 20000);  // "completes" 20 us later
  return hipSuccess;
}
hipError_t hipEventQuery(hipEvent_t event)
{
  const long long t = event->done_at_ns.load();
  return t != 0 && now_ns() >= t ? hipSuccess : hipErrorNotReady;
}
hipError_t hipPointerGetAttributes(hipPointerAttribute_t* a, const void* p)
{
  std::lock_guard<std::mutex> lock(g_mu);
  auto it = g_pinned.upper_bound(static_cast<const char*>(p));
  if (it == g_pinned.begin()) return hipErrorInvalidValue;
  --it;
  if (static_cast<const char*>(p) >= it->first + it->second) return hipErrorInvalidValue;
  a->type = hipMemoryTypeHost;
  a->device = 0;
  a->devicePointer = const_cast<void*>(p);
  a->hostPointer = const_cast<void*>(p);
  return hipSuccess;
}
hipError_t hipMemGetAddressRange(hipDeviceptr_t* base, size_t* size, hipDeviceptr_t p)
{
  std::lock_guard<std::mutex> lock(g_mu);
  auto it = g_pinned.upper_bound(static_cast<const char*>(p));
  if (it == g_pinned.begin()) return hipErrorInvalidValue;
  --it;
  *base = const_cast<char*>(it->first);
  *size = it->second;
  return hipSuccess;
}

// ---- the library entry points the serving loop calls ------------------------
namespace {
thread_local char t_error[512] = "";
}

namespace dpc {
int set_error(int code, const char* fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(t_error, sizeof t_error, fmt, ap);
  va_end(ap);
  return code;
}
int launch_serve(const void*, uint32_t, uint64_t, const void*, const void*, const float*, uint64_t*, uint32_t,
                 uint32_t, void*)
{
  return DPC_OK;
}
// row r of matrix s lives at planes[4 * (r * ld + s)]
int launch_scatter_host(float* planes, size_t ld, const uint32_t* idx, const float* src, size_t count, void*)
{
  for (size_t k = 0; k < count; ++k)
    for (int r = 0; r < 4; ++r) std::memcpy(planes + 4 * (r * ld + idx[k]), src + 16 * k + 4 * r, 16);
  return DPC_OK;
}
int launch_scatter_affine(float* planes, size_t ld, const uint32_t* idx, const void* packed, const void* tails,
                          size_t count, void*)
{
  const uint32_t* p = static_cast<const uint32_t*>(packed);
  const uint32_t* t = static_cast<const uint32_t*>(tails);
  const uint32_t affine[4] = {0u, 0u, 0u, 0x3f800000u};
  for (size_t k = 0; k < count; ++k) {
    const uint32_t* w = t ? t + 4 * k : affine;
    const uint32_t* a = p + 12 * k;
    const uint32_t rows[4][4] = {{a[0], a[1], a[2], w[0]},
                                 {a[4], a[5], a[6], w[1]},
                                 {a[8], a[9], a[10], w[2]},
                                 {a[3], a[7], a[11], w[3]}};
    for (int r = 0; r < 4; ++r) std::memcpy(planes + 4 * (r * ld + idx[k]), rows[r], 16);
  }
  return DPC_OK;
}
int launch_copy_out(const void*, uint32_t, const char*, uint64_t, void*) { return DPC_OK; }
}  // namespace dpc

extern "C" {
const char* dpc_last_error(void) { return t_error; }
int dpc_host_alloc(void** h_ptr, size_t bytes) { return hipHostMalloc(h_ptr, bytes, 0) == hipSuccess ? DPC_OK : 2; }
int dpc_host_free(void* h_ptr) { return hipHostFree(h_ptr) == hipSuccess ? DPC_OK : 2; }
int dpc_boxes_from_lower_upper(const float*, size_t, float*, float*, void*) { return DPC_OK; }
int dpc_matrices_from_aos(const void* aos, size_t stride_bytes, size_t count, float* planes, size_t ld, void*)
{
  for (size_t k = 0; k < count; ++k)
    for (int r = 0; r < 4; ++r)
      std::memcpy(planes + 4 * (r * ld + k), static_cast<const char*>(aos) + k * stride_bytes + 16 * r, 16);
  return DPC_OK;
}
int dpc_matrices_scatter(float* planes, size_t ld, const uint32_t* idx, const float* src, size_t count, void* stream)
{
  return dpc::launch_scatter_host(planes, ld, idx, src, count, stream);
}
}
