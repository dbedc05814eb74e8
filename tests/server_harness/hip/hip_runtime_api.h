// Test double of the HIP runtime API for the host-side server harness: the
// calls the build server makes, run on the CPU.  A stream is a worker thread
// that runs its queued operations in order (copies, event completions, the
// fake builds), so the server's host code meets the same asynchrony as on a
// GPU and the sanitizers see every hand-off through real synchronization.
#pragma once

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
#include <functional>
#endif

typedef enum hipError_t {
  hipSuccess = 0,
  hipErrorInvalidValue = 1,
  hipErrorOutOfMemory = 2,
  hipErrorInvalidDevice = 101,
} hipError_t;

typedef enum hipMemcpyKind {
  hipMemcpyHostToHost = 0,
  hipMemcpyHostToDevice = 1,
  hipMemcpyDeviceToHost = 2,
  hipMemcpyDeviceToDevice = 3,
} hipMemcpyKind;

typedef struct FakeStream* hipStream_t;
typedef struct FakeEvent* hipEvent_t;

#define hipEventDisableTiming 0x2u
#define hipHostMallocDefault 0x0u

hipError_t hipSetDevice(int device);
hipError_t hipGetLastError(void);
hipError_t hipMalloc(void** p, size_t bytes);
hipError_t hipFree(void* p);
hipError_t hipHostMalloc(void** p, size_t bytes, unsigned flags);
hipError_t hipHostFree(void* p);
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind,
                          hipStream_t stream);
hipError_t hipStreamSynchronize(hipStream_t stream);
hipError_t hipEventCreateWithFlags(hipEvent_t* ev, unsigned flags);
hipError_t hipEventDestroy(hipEvent_t ev);
hipError_t hipEventRecord(hipEvent_t ev, hipStream_t stream);
hipError_t hipEventSynchronize(hipEvent_t ev);

#ifdef __cplusplus
template <class T>
hipError_t hipMalloc(T** p, size_t bytes) {
  return hipMalloc(reinterpret_cast<void**>(p), bytes);
}
// Harness side: streams and queued host work standing in for kernels.
hipStream_t fake_stream_create();
void fake_stream_destroy(hipStream_t s);
void fake_enqueue(hipStream_t s, std::function<void()> op);
// Device bytes currently allocated by hipMalloc (the server's arena).
size_t fake_device_bytes();
#endif
