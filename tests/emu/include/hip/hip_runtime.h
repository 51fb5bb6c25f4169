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
// Non-temporal (streaming) global accesses are a cache hint on the device; here they are plain.
#define CFFT_PLAIN_STREAMING 1
// Cross-lane exchanges (the device's DPP / permlane moves within a wavefront): the value of the
// thread whose index differs by xor d, traded through a buffer between two barriers of the block
// (every thread of the block calls it the same number of times).
#define CFFT_EMU_LANES 1
int cfft_emu_lane_xor(int v, int d);

struct float2 {
    float x, y;
};

// IEEE binary16 in software, standing in for the device's _Float16: conversion from float rounds
// to nearest, ties to even (subnormals kept, overflow to infinity), as v_cvt_f16_f32 does.
struct emu_half {
    uint16_t bits;
    emu_half() = default;
    emu_half(float f) : bits(from_float(f)) {}
    operator float() const { return to_float(bits); }
    static uint16_t from_float(float f) {
        uint32_t x;
        std::memcpy(&x, &f, 4);
        const uint32_t sign = (x >> 16) & 0x8000u, ax = x & 0x7fffffffu;
        if (ax >= 0x7f800000u)  // infinity, NaN (quiet)
            return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u | ((ax >> 13) & 0x3ffu) : 0u));
        if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);  // >= 65520 rounds to infinity
        if (ax < 0x38800000u) {                                    // below 2^-14: subnormal
            if (ax < 0x33000000u) return (uint16_t)sign;           // below 2^-25: zero
            const uint32_t e = ax >> 23, m = (ax & 0x7fffffu) | 0x800000u;
            const uint32_t shift = 126u - e;                        // units of 2^-24
            uint32_t q = m >> shift;
            const uint32_t rem = m & ((1u << shift) - 1u), half = 1u << (shift - 1u);
            if (rem > half || (rem == half && (q & 1u))) ++q;
            return (uint16_t)(sign | q);
        }
        uint32_t q = (((ax >> 23) - 112u) << 10) | ((ax & 0x7fffffu) >> 13);
        const uint32_t rem = ax & 0x1fffu;
        if (rem > 0x1000u || (rem == 0x1000u && (q & 1u))) ++q;
        return (uint16_t)(sign | q);
    }
    static float to_float(uint16_t h) {
        const uint32_t sign = (uint32_t)(h & 0x8000u) << 16, e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
        uint32_t x;
        if (e == 0x1f) {
            x = sign | 0x7f800000u | (m << 13);
        } else if (e == 0) {
            const float f = (float)m * 5.9604644775390625e-08f;  // m * 2^-24, exact
            std::memcpy(&x, &f, 4);
            x |= sign;
        } else {
            x = sign | ((e + 112u) << 23) | (m << 13);
        }
        float f;
        std::memcpy(&f, &x, 4);
        return f;
    }
};
#define CFFT_F16 emu_half
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
inline unsigned __umulhi(unsigned a, unsigned b) {
    return (unsigned)(((unsigned long long)a * b) >> 32);
}

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
#define hipHostMallocCoherent 0x40000000
hipError_t hipHostMalloc(void** p, size_t n, unsigned flags);
// host memory is device memory in the emulation
inline hipError_t hipHostGetDevicePointer(void** d, void* h, unsigned) {
    *d = h;
    return hipSuccess;
}
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
hipError_t hipMemsetAsync(void* p, int v, size_t n, hipStream_t st);
// a __device__ variable is an ordinary host object here: its address is its device address
#define HIP_SYMBOL(x) (&(x))
inline hipError_t hipGetSymbolAddress(void** p, const void* sym) {
    *p = const_cast<void*>(sym);
    return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned f);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipEventCreate(hipEvent_t* e);
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned f);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s);
// launches complete in order before they return: nothing to wait for
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
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
