// Host-only stand-in for the HIP runtime, used to build damask_amd/csrc/batch.cpp
// on the CPU for the lockstep-batch harness (tests/lockstep_host/harness.cpp).
// Streams are in-order queues executed by one host thread each, events carry
// generation counters, and hipMemcpyAsync copies when its stream reaches it, so
// a missing stream wait or an early reuse of a staging buffer shows up as a
// wrong value read by a later operation.
#pragma once
#include <cstddef>
#include <cstdint>

struct MockStream;
struct MockEvent;
typedef MockStream* hipStream_t;
typedef MockEvent* hipEvent_t;
typedef int hipError_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1 };
enum hipMemcpyKind { hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2, hipMemcpyDeviceToDevice = 3 };
#define hipStreamNonBlocking 1u
#define hipEventDisableTiming 2u
#define hipHostMallocDefault 0u

hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned flags);
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned flags);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind, hipStream_t s);
hipError_t hipHostMalloc(void** p, size_t bytes, unsigned flags);
hipError_t hipHostFree(void* p);
hipError_t hipMalloc(void** p, size_t bytes);
hipError_t hipFree(void* p);
hipError_t hipGetLastError(void);
const char* hipGetErrorString(hipError_t e);
hipError_t hipEventCreate(hipEvent_t* e);
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b);
hipError_t hipMemGetInfo(size_t* free_bytes, size_t* total_bytes);
typedef void (*hipHostFn_t)(void*);
hipError_t hipLaunchHostFunc(hipStream_t s, hipHostFn_t fn, void* user);
struct double2 {
  double x, y;
};
