// A host-only stand-in for the part of the HIP runtime API that the serving
// loop (pipeline_amd/csrc/host/Serving.cpp) calls, for its ThreadSanitizer test
// on the CPU (tests/cpp/serve_stress.cpp built with -DDPCULL_HOST_RUNTIME):
// "device" memory is host memory, every asynchronous call completes before it
// returns, and an event reports completion a little after it was recorded, so
// the scheduler's polling and the issuers' calls interleave as on a GPU.  Only
// the threads and locks of the serving loop are under test; no result is
// computed.
#pragma once
#include <cstddef>

typedef enum hipError_t { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorNotReady = 600 } hipError_t;
typedef enum hipMemcpyKind { hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2 } hipMemcpyKind;
typedef enum hipMemoryType { hipMemoryTypeHost = 1, hipMemoryTypeDevice = 2 } hipMemoryType;

typedef struct HostStream* hipStream_t;
typedef struct HostEvent* hipEvent_t;
typedef void* hipDeviceptr_t;

typedef struct hipPointerAttribute_t {
  hipMemoryType type;
  int device;
  void* devicePointer;
  void* hostPointer;
} hipPointerAttribute_t;

#define hipStreamNonBlocking 1u
#define hipEventDisableTiming 2u
#define hipHostMallocDefault 0u

const char* hipGetErrorString(hipError_t e);
hipError_t hipGetLastError();
hipError_t hipSetDevice(int device);
hipError_t hipGetDevice(int* device);
hipError_t hipMalloc(void** p, size_t bytes);
hipError_t hipFree(void* p);
hipError_t hipHostMalloc(void** p, size_t bytes, unsigned flags);
hipError_t hipHostFree(void* p);
hipError_t hipMemcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind kind);
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind, hipStream_t stream);
hipError_t hipStreamCreateWithFlags(hipStream_t* stream, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t stream);
hipError_t hipStreamSynchronize(hipStream_t stream);
hipError_t hipDeviceSynchronize();
hipError_t hipEventCreateWithFlags(hipEvent_t* event, unsigned flags);
hipError_t hipEventDestroy(hipEvent_t event);
hipError_t hipEventRecord(hipEvent_t event, hipStream_t stream);
hipError_t hipEventQuery(hipEvent_t event);
hipError_t hipPointerGetAttributes(hipPointerAttribute_t* attributes, const void* p);
hipError_t hipMemGetAddressRange(hipDeviceptr_t* base, size_t* size, hipDeviceptr_t p);

// (as the HIP headers do, for typed pointers)
template <class T>
inline hipError_t hipMalloc(T** p, size_t bytes)
{
  return hipMalloc(reinterpret_cast<void**>(p), bytes);
}
