// runtime.hip -- device memory, streams, events, graphs and small utility
// kernels behind the C ABI.
#include <atomic>
#include <mutex>
#include <vector>

#include "internal.hpp"

namespace dpc {

namespace {
std::atomic<long long> g_inject_countdown{-1};  // < 0: off
std::atomic<int> g_inject_error{hipSuccess};
thread_local std::string g_last_error;
std::mutex g_cu_mutex;
std::vector<int> g_cu_count;  // per device, 0 = unknown
}  // namespace

int set_error(int code, const char* fmt, ...)
{
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

hipError_t injected(hipError_t e)
{
  if (g_inject_countdown.load(std::memory_order_relaxed) < 0) return e;
  if (g_inject_countdown.fetch_sub(1, std::memory_order_relaxed) == 0)
    return static_cast<hipError_t>(g_inject_error.load(std::memory_order_relaxed));
  return e;
}

int device_cu_count()
{
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  std::lock_guard<std::mutex> lock(g_cu_mutex);
  if (static_cast<size_t>(dev) >= g_cu_count.size()) g_cu_count.resize(dev + 1, 0);
  if (g_cu_count[dev] == 0) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    g_cu_count[dev] = cus;
  }
  return g_cu_count[dev];
}

namespace {
// The issue rate of unfused, unpacked FP32 adds (Op 0) or multiplies (Op 1):
// eight independent accumulators per lane, so no operation waits for another.
template <int Op>
__global__ __launch_bounds__(256) void valu_probe_kernel(float* __restrict__ out, float b, uint32_t iters)
{
  float a0 = static_cast<float>(threadIdx.x), a1 = a0 + 1.f, a2 = a0 + 2.f, a3 = a0 + 3.f, a4 = a0 + 4.f,
        a5 = a0 + 5.f, a6 = a0 + 6.f, a7 = a0 + 7.f;
  for (uint32_t i = 0; i < iters; ++i) {
    if (Op == 0) {
      a0 = a0 + b; a1 = a1 + b; a2 = a2 + b; a3 = a3 + b; a4 = a4 + b; a5 = a5 + b; a6 = a6 + b; a7 = a7 + b;
    } else {
      a0 = a0 * b; a1 = a1 * b; a2 = a2 * b; a3 = a3 * b; a4 = a4 * b; a5 = a5 * b; a6 = a6 * b; a7 = a7 * b;
    }
    __builtin_amdgcn_sched_barrier(0);
  }
  out[static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x] = a0 + a1 + a2 + a3 + a4 + a5 + a6 + a7;
}

__global__ void fill_u32_kernel(uint32_t* __restrict__ dst, uint32_t value, size_t n)
{
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = value;
}
}  // namespace

}  // namespace dpc

using namespace dpc;

extern "C" {

int dpc_abi_version(void) { return DPC_ABI_VERSION; }

int dpc_debug_inject_error(int64_t after_calls, int hip_error)
{
  if (after_calls >= 0 && hip_error == 0)
    return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_debug_inject_error: the injected error must be nonzero");
  g_inject_error.store(hip_error, std::memory_order_relaxed);
  g_inject_countdown.store(after_calls < 0 ? -1 : after_calls, std::memory_order_relaxed);
  return DPC_OK;
}

const char* dpc_last_error(void) { return g_last_error.c_str(); }

int dpc_device_info(int device, char* name, size_t name_len, int* cu_count, int* clock_khz, size_t* total_mem,
                    int* l2_bytes)
{
  hipDeviceProp_t prop;
  DPC_HIP_TRY(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
  if (name && name_len) {
    snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
  }
  if (cu_count) *cu_count = prop.multiProcessorCount;
  if (clock_khz) *clock_khz = prop.clockRate;
  if (total_mem) *total_mem = prop.totalGlobalMem;
  if (l2_bytes) *l2_bytes = prop.l2CacheSize;
  return DPC_OK;
}

int dpc_set_device(int device)
{
  DPC_HIP_TRY(hipSetDevice(device), "hipSetDevice");
  return DPC_OK;
}

int dpc_malloc(void** d_ptr, size_t bytes)
{
  if (!d_ptr) return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_malloc: null output pointer");
  *d_ptr = nullptr;
  if (bytes == 0) return DPC_OK;
  hipError_t e = injected(hipMalloc(d_ptr, bytes));
  if (e != hipSuccess && *d_ptr) {  // an injected failure of a call that succeeded
    (void)hipFree(*d_ptr);
    *d_ptr = nullptr;
  }
  if (e == hipErrorOutOfMemory) return set_error(DPC_ERR_OUT_OF_MEMORY, "dpc_malloc: out of memory (%zu bytes)", bytes);
  if (e != hipSuccess) return hip_fail(e, "hipMalloc");
  return DPC_OK;
}

int dpc_free(void* d_ptr)
{
  if (!d_ptr) return DPC_OK;
  DPC_HIP_TRY(hipFree(d_ptr), "hipFree");
  return DPC_OK;
}

int dpc_host_alloc(void** h_ptr, size_t bytes)
{
  if (!h_ptr) return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_host_alloc: null output pointer");
  *h_ptr = nullptr;
  if (bytes == 0) return DPC_OK;
  DPC_HIP_TRY(hipHostMalloc(h_ptr, bytes, hipHostMallocDefault), "hipHostMalloc");
  return DPC_OK;
}

int dpc_host_free(void* h_ptr)
{
  if (!h_ptr) return DPC_OK;
  DPC_HIP_TRY(hipHostFree(h_ptr), "hipHostFree");
  return DPC_OK;
}

int dpc_memcpy_h2d(void* d_dst, const void* src, size_t bytes, void* stream)
{
  if (bytes == 0) return DPC_OK;
  DPC_HIP_TRY(hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, as_stream(stream)), "hipMemcpyAsync H2D");
  return DPC_OK;
}

int dpc_memcpy_d2h(void* dst, const void* d_src, size_t bytes, void* stream)
{
  if (bytes == 0) return DPC_OK;
  DPC_HIP_TRY(hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, as_stream(stream)), "hipMemcpyAsync D2H");
  return DPC_OK;
}

int dpc_memcpy_d2d(void* d_dst, const void* d_src, size_t bytes, void* stream)
{
  if (bytes == 0) return DPC_OK;
  DPC_HIP_TRY(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, as_stream(stream)), "hipMemcpyAsync D2D");
  return DPC_OK;
}

int dpc_probe_valu_rate(int op, uint32_t iters, double* lane_ops_per_s, void* stream)
{
  if (!lane_ops_per_s || (op != 0 && op != 1) || iters == 0)
    return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_probe_valu_rate: op 0 or 1, iters > 0 and an output");
  *lane_ops_per_s = 0.0;
  const unsigned blocks = static_cast<unsigned>(device_cu_count()) * 8;  // 32 waves per CU
  hipStream_t s = as_stream(stream);
  float* d_out = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto launch = [&]() {
    if (op == 0)
      hipLaunchKernelGGL(valu_probe_kernel<0>, dim3(blocks), dim3(256), 0, s, d_out, 1.0000001f, iters);
    else
      hipLaunchKernelGGL(valu_probe_kernel<1>, dim3(blocks), dim3(256), 0, s, d_out, 1.0000001f, iters);
    return check_launch("valu_probe_kernel");
  };
  int rc = DPC_OK;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&d_out), sizeof(float) * 256 * blocks);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e != hipSuccess) rc = set_error(DPC_ERR_HIP, "dpc_probe_valu_rate: %s", hipGetErrorString(e));
  if (rc == DPC_OK) rc = launch();  // (first launch untimed)
  if (rc == DPC_OK && (e = hipEventRecord(e0, s)) != hipSuccess) rc = set_error(DPC_ERR_HIP, "hipEventRecord");
  for (int r = 0; r < 5 && rc == DPC_OK; ++r) rc = launch();
  if (rc == DPC_OK && (e = hipEventRecord(e1, s)) != hipSuccess) rc = set_error(DPC_ERR_HIP, "hipEventRecord");
  float ms = 0.f;
  if (rc == DPC_OK && (e = hipEventSynchronize(e1)) != hipSuccess) rc = set_error(DPC_ERR_HIP, "hipEventSynchronize");
  if (rc == DPC_OK && (e = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess)
    rc = set_error(DPC_ERR_HIP, "hipEventElapsedTime");
  if (rc == DPC_OK && ms > 0.f)
    *lane_ops_per_s = 5.0 * blocks * 256.0 * 8.0 * static_cast<double>(iters) / (ms * 1e-3);
  (void)hipStreamSynchronize(s);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (d_out) (void)hipFree(d_out);
  return rc;
}

int dpc_fill_u32(uint32_t* d_dst, uint32_t value, size_t n, void* stream)
{
  if (n == 0) return DPC_OK;
  if (!d_dst || !aligned(d_dst, 4)) return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_fill_u32: bad destination");
  const unsigned threads = 256;
  size_t blocks = (n + threads - 1) / threads;
  const size_t cap = static_cast<size_t>(device_cu_count()) * 16;
  if (blocks > cap) blocks = cap;
  hipLaunchKernelGGL(fill_u32_kernel, dim3(static_cast<unsigned>(blocks)), dim3(threads), 0, as_stream(stream), d_dst,
                     value, n);
  return check_launch("fill_u32_kernel");
}

int dpc_stream_create(void** stream)
{
  if (!stream) return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_stream_create: null output");
  hipStream_t s = nullptr;
  // Default priority, non-blocking with respect to the null stream.
  DPC_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreateWithFlags");
  *stream = s;
  return DPC_OK;
}

int dpc_stream_destroy(void* stream)
{
  if (!stream) return DPC_OK;
  DPC_HIP_TRY(hipStreamDestroy(as_stream(stream)), "hipStreamDestroy");
  return DPC_OK;
}

int dpc_stream_sync(void* stream)
{
  DPC_HIP_TRY(hipStreamSynchronize(as_stream(stream)), "hipStreamSynchronize");
  return DPC_OK;
}

int dpc_device_sync(void)
{
  DPC_HIP_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
  return DPC_OK;
}

int dpc_event_create(void** event)
{
  if (!event) return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_event_create: null output");
  hipEvent_t e = nullptr;
  DPC_HIP_TRY(hipEventCreate(&e), "hipEventCreate");
  *event = e;
  return DPC_OK;
}

int dpc_event_destroy(void* event)
{
  if (!event) return DPC_OK;
  DPC_HIP_TRY(hipEventDestroy(static_cast<hipEvent_t>(event)), "hipEventDestroy");
  return DPC_OK;
}

int dpc_event_record(void* event, void* stream)
{
  DPC_HIP_TRY(hipEventRecord(static_cast<hipEvent_t>(event), as_stream(stream)), "hipEventRecord");
  return DPC_OK;
}

int dpc_event_sync(void* event)
{
  DPC_HIP_TRY(hipEventSynchronize(static_cast<hipEvent_t>(event)), "hipEventSynchronize");
  return DPC_OK;
}

int dpc_event_elapsed_ms(void* start, void* stop, float* ms)
{
  if (!ms) return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_event_elapsed_ms: null output");
  DPC_HIP_TRY(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)),
              "hipEventElapsedTime");
  return DPC_OK;
}

int dpc_graph_begin(void* stream)
{
  if (!stream) return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_graph_begin: capture needs a created stream");
  DPC_HIP_TRY(hipStreamBeginCapture(as_stream(stream), hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  return DPC_OK;
}

int dpc_graph_end(void* stream, void** graph_exec)
{
  if (!graph_exec) return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_graph_end: null output");
  hipGraph_t g = nullptr;
  DPC_HIP_TRY(hipStreamEndCapture(as_stream(stream), &g), "hipStreamEndCapture");
  hipGraphExec_t ge = nullptr;
  hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
  *graph_exec = ge;
  return DPC_OK;
}

int dpc_graph_launch(void* graph_exec, void* stream)
{
  DPC_HIP_TRY(hipGraphLaunch(static_cast<hipGraphExec_t>(graph_exec), as_stream(stream)), "hipGraphLaunch");
  return DPC_OK;
}

int dpc_graph_destroy(void* graph_exec)
{
  if (!graph_exec) return DPC_OK;
  DPC_HIP_TRY(hipGraphExecDestroy(static_cast<hipGraphExec_t>(graph_exec)), "hipGraphExecDestroy");
  return DPC_OK;
}

}  // extern "C"
