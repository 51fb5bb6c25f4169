// CPU emulation of the small part of the HIP runtime and device language that the library uses.
// TEST INFRASTRUCTURE ONLY: it lets the CPU test suite compile the library's own kernel and
// planner sources with g++ and run them (tests/emu/build_emu.py -> libcatabrafft_emu.so).
// GPU threads are fibers of one OS thread; __syncthreads() is a cooperative barrier; "device"
// memory is host memory; every launch runs to completion before the call returns.
#pragma once
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <functional>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __align__(n) alignas(n)
#define __shared__
#define CFFT_DYN_LDS(name) unsigned char* name = emu::smem()
#define CFFT_OPAQUE(x) asm volatile("" : "+r"(x))

struct float2 {
    float x, y;
};
struct double2 {
    double x, y;
};
struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};

namespace emu {
struct Idx {
    unsigned x, y, z;
};
Idx& thread_idx();
Idx& block_idx();
Idx& block_dim();
Idx& grid_dim();
unsigned char* smem();
void barrier();
struct Stream;
void launch(std::function<void()> body, dim3 grid, dim3 block, size_t lds, Stream* s);
}  // namespace emu

#define threadIdx (emu::thread_idx())
#define blockIdx (emu::block_idx())
#define blockDim (emu::block_dim())
#define gridDim (emu::grid_dim())
inline void __syncthreads() { emu::barrier(); }
inline int __ffs(int x) { return __builtin_ffs(x); }

typedef int hipError_t;
enum {
    hipSuccess = 0,
    hipErrorInvalidValue = 1,
    hipErrorOutOfMemory = 2,
    hipErrorNoDevice = 100,
};
typedef emu::Stream* hipStream_t;
struct emu_event;
struct emu_graph;
struct emu_graph_exec;
typedef emu_event* hipEvent_t;
typedef emu_graph* hipGraph_t;
typedef emu_graph_exec* hipGraphExec_t;
typedef void* hipGraphNode_t;
enum hipMemcpyKind {
    hipMemcpyHostToHost = 0,
    hipMemcpyHostToDevice = 1,
    hipMemcpyDeviceToHost = 2,
    hipMemcpyDeviceToDevice = 3,
    hipMemcpyDefault = 4
};
enum hipDeviceAttribute_t {
    hipDeviceAttributeMaxSharedMemoryPerBlock = 1,
    hipDeviceAttributeMultiprocessorCount = 2
};
enum hipFuncAttribute { hipFuncAttributeMaxDynamicSharedMemorySize = 1 };
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal = 0, hipStreamCaptureModeThreadLocal = 1 };
#define hipStreamNonBlocking 1
#define hipEventDisableTiming 2

hipError_t hipGetDeviceCount(int* n);
hipError_t hipSetDevice(int d);
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t a, int d);
hipError_t hipFuncSetAttribute(const void* f, hipFuncAttribute a, int v);
hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int* n, const void* f, int block, size_t lds);
#define hipHostMallocDefault 0
hipError_t hipHostMalloc(void** p, size_t n, unsigned flags);
hipError_t hipHostFree(void* p);
hipError_t hipGetLastError();
const char* hipGetErrorString(hipError_t e);
hipError_t hipDeviceSynchronize();
hipError_t hipMalloc(void** p, size_t n);
hipError_t hipFree(void* p);
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind k);
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t st);
hipError_t hipMemcpy2DAsync(void* d, size_t dp, const void* s, size_t sp, size_t w, size_t h,
                            hipMemcpyKind k, hipStream_t st);
hipError_t hipMemset(void* p, int v, size_t n);
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned f);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipEventCreate(hipEvent_t* e);
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned f);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s);
hipError_t hipEventQuery(hipEvent_t e);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b);
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode m);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g);
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, hipGraphNode_t* n, char* log,
                               size_t len);
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s);
hipError_t hipGraphDestroy(hipGraph_t g);
hipError_t hipGraphExecDestroy(hipGraphExec_t e);

template <class F, class... A>
inline void emu_launch_kernel(F f, dim3 g, dim3 b, size_t lds, hipStream_t s, A... a) {
    emu::launch([=]() { f(a...); }, g, b, lds, s);
}
#define hipLaunchKernelGGL(K, G, B, L, S, ...) emu_launch_kernel(K, dim3(G), dim3(B), L, S, __VA_ARGS__)
