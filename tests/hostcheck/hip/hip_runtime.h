// Host-only stand-in for the HIP runtime header, for the sanitizer build of
// the device context (context.cpp) on a machine without a GPU.  Device
// allocations are host heap blocks tracked by a registry, copies are memcpy
// with range checks against that registry, and stream capture records the
// checked kernel launches (mock_kernels.cpp) into graphs that replay them,
// so a replay re-checks every buffer a captured launch names.  Compiled with
// g++ and -fsanitize=address,undefined; nothing here runs on a GPU.
#pragma once

#include <cstddef>
#include <cstdint>
#include <functional>

#define __host__
#define __device__
#define __global__
#define __forceinline__ inline
#define __align__(n) alignas(n)

typedef int hipError_t;
enum {
    hipSuccess = 0,
    hipErrorInvalidValue = 1,
    hipErrorOutOfMemory = 2,
    hipErrorNoDevice = 100,
    hipErrorStreamCaptureUnsupported = 900,
};

struct int4 {
    int x, y, z, w;
};
inline int4 make_int4(int x, int y, int z, int w) { return int4{x, y, z, w}; }

typedef struct mock_stream* hipStream_t;
typedef struct mock_graph* hipGraph_t;
typedef struct mock_graph_exec* hipGraphExec_t;
typedef struct mock_event* hipEvent_t;
typedef struct mock_graph_node* hipGraphNode_t;

enum hipMemcpyKind { hipMemcpyHostToHost, hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice };
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal, hipStreamCaptureModeThreadLocal, hipStreamCaptureModeRelaxed };
enum hipDeviceAttribute_t { hipDeviceAttributeMultiprocessorCount };
#define hipStreamDefault 0u
#define hipStreamNonBlocking 1u

struct hipDeviceProp_t {
    char name[256];
    char gcnArchName[256];
    int multiProcessorCount;
    size_t totalGlobalMem;
    int memoryClockRate;
    int memoryBusWidth;
};

hipError_t hipMalloc(void** p, size_t bytes);
hipError_t hipFree(void* p);
hipError_t hipMemcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind kind);
hipError_t hipMemset(void* p, int value, size_t bytes);
hipError_t hipMemsetAsync(void* p, int value, size_t bytes, hipStream_t s);
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode mode);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g);
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, hipGraphNode_t* err_node, char* log, size_t log_len);
hipError_t hipGraphDestroy(hipGraph_t g);
hipError_t hipGraphExecDestroy(hipGraphExec_t e);
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s);
hipError_t hipEventCreate(hipEvent_t* e);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b);
hipError_t hipGetDeviceCount(int* n);
hipError_t hipGetDevice(int* d);
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t a, int dev);
hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int dev);
hipError_t hipDeviceSynchronize();
hipError_t hipGetLastError();
// A __device__ variable is a host array here; its address is registered as
// a device block (it lives until the process ends).
#define HIP_SYMBOL(x) mock::symbol(&(x), sizeof(x))
hipError_t hipGetSymbolAddress(void** p, const void* symbol);
const char* hipGetErrorString(hipError_t e);

// ---- checks the mock kernels use (mock_runtime.cpp)
namespace mock {
// [p, p + bytes) lies inside one live device allocation; otherwise the
// harness aborts with the launch's name (bytes == 0: p must still be inside
// or at the end of a live allocation, or null when allow_null).
void check_dev(const void* p, size_t bytes, const char* what, bool allow_null = false);
// Record (under capture) or run a launch on stream s.
void launch(hipStream_t s, const char* name, std::function<void()> body);
// Make the next n successful hipMalloc calls succeed and the one after fail
// (n < 0: never fail).
void fail_malloc_after(long long n);
long long live_allocations();
// hipMalloc calls so far (failed ones included).
long long malloc_calls();
const void* symbol(const void* p, size_t bytes);
}  // namespace mock
