// mfma_dft16 -- a radix-16 DFT stage on the matrix cores against the library's radix-16 codelet.
//
// Both kernels apply y = F16 x / 4 (F16 the 16-point DFT matrix; the 1/4 keeps the norm) REPS times
// to 2^20 independent 16-point vectors held in registers, so that the arithmetic, not memory,
// sets the time:
//   codelet: one thread per vector, cfft::dft16 (the generated straight-line butterfly);
//   mfma:    the stage as a real 32 x 32 matrix [Re F, -Im F; Im F, Re F] times a 32 x 16 block
//            (real parts over imaginary parts of 16 vectors), on v_mfma_f32_16x16x4f32: 16
//            instructions per block.  The input rows are ordered so that the accumulator layout of
//            one stage is the operand layout of the next (no lane exchanges between stages), which
//            leaves the matrix cores nothing but the products: their best case.
// Each result is checked against a double-precision host computation of the same stages.
//
// Build: hipcc --offload-arch=gfx950 -O3 -I <csrc> -x hip mfma_dft16.cpp -o mfma_dft16
// Run:   mfma_dft16   (prints the time of each kernel, its rates and its error)
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "codelets.h"

#define CK(x)                                                                            \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess) {                                                          \
            std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__);  \
            std::exit(1);                                                                \
        }                                                                                \
    } while (0)

constexpr int kVectors = 1 << 20;  // 16-point vectors
constexpr int kReps = 64;          // stages applied in registers
constexpr int kBlocks = 4;         // 16-vector blocks per wave (independent accumulators)
typedef float f4 __attribute__((ext_vector_type(4)));

// Element index of input row k (0..15) of a block: the 4 x 4 transpose, so that lane group q's
// accumulator rows 4q .. 4q + 3 (outputs 4q .. 4q + 3) are the rows it feeds the next stage.
__host__ __device__ constexpr int row_elem(int k) { return 4 * (k % 4) + k / 4; }

__global__ void __launch_bounds__(256) k_codelet(const float2* in, float2* out, float scale) {
    const int v0 = blockIdx.x * blockDim.x + threadIdx.x;
    float2 v[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) v[e] = in[(size_t)v0 * 16 + e];
    for (int r = 0; r < kReps; ++r) {
        cfft::dft16(v);
#pragma unroll
        for (int e = 0; e < 16; ++e) v[e] = float2{v[e].x * scale, v[e].y * scale};
    }
#pragma unroll
    for (int e = 0; e < 16; ++e) out[(size_t)v0 * 16 + e] = v[e];
}

// W: the real 32 x 32 stage matrix, row-major, columns in row_elem order.
__global__ void __launch_bounds__(256) k_mfma(const float2* in, float2* out, const float* W,
                                              float scale) {
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int q = lane >> 4, j = lane & 15;
    // A operand of 16x16x4: lane holds A[i][k], i = lane % 16, k = lane / 16 (row block I,
    // k-step s covers k = 4 s .. 4 s + 3)
    float a[2][8];
#pragma unroll
    for (int I = 0; I < 2; ++I)
#pragma unroll
        for (int s = 0; s < 8; ++s) a[I][s] = W[(16 * I + j) * 32 + 4 * s + q];
    // B operand: lane holds B[k][j], k = 4 s + q: rows 0..15 real, 16..31 imaginary parts of
    // elements row_elem(k), i.e. element 4 q + s of vector j
    float b[kBlocks][8];
#pragma unroll
    for (int blk = 0; blk < kBlocks; ++blk) {
        const size_t vec = (size_t)(wave * kBlocks + blk) * 16 + j;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const float2 x = in[vec * 16 + 4 * q + s];
            b[blk][s] = x.x;
            b[blk][s + 4] = x.y;
        }
    }
    for (int r = 0; r < kReps; ++r) {
        f4 acc[kBlocks][2];
#pragma unroll
        for (int blk = 0; blk < kBlocks; ++blk) acc[blk][0] = acc[blk][1] = f4{0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < 8; ++s)
#pragma unroll
            for (int blk = 0; blk < kBlocks; ++blk)
#pragma unroll
                for (int I = 0; I < 2; ++I)
                    acc[blk][I] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[I][s], b[blk][s],
                                                                       acc[blk][I], 0, 0, 0);
        // accumulator: lane holds D[i][j], i = 4 q + t (t = 0..3): outputs 4 q + t, real parts
        // in row block 0 and imaginary parts in row block 1 -- the next stage's B rows
#pragma unroll
        for (int blk = 0; blk < kBlocks; ++blk)
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                b[blk][t] = acc[blk][0][t] * scale;
                b[blk][t + 4] = acc[blk][1][t] * scale;
            }
    }
#pragma unroll
    for (int blk = 0; blk < kBlocks; ++blk) {
        const size_t vec = (size_t)(wave * kBlocks + blk) * 16 + j;
#pragma unroll
        for (int s = 0; s < 4; ++s) out[vec * 16 + 4 * q + s] = float2{b[blk][s], b[blk][s + 4]};
    }
}

int main() {
    const double pi = 3.14159265358979323846;
    std::vector<float2> h((size_t)kVectors * 16);
    unsigned long long st = 12345;
    auto rnd = [&]() {
        st = st * 6364136223846793005ULL + 1442695040888963407ULL;
        return (double)((st >> 11) & ((1ULL << 53) - 1)) / (double)(1ULL << 53) - 0.5;
    };
    for (auto& x : h) x = float2{(float)rnd(), (float)rnd()};
    std::vector<float> Wh(32 * 32);
    for (int i = 0; i < 16; ++i)
        for (int k = 0; k < 16; ++k) {
            const double ang = -2 * pi * i * row_elem(k) / 16.0;
            const double re = std::cos(ang), im = std::sin(ang);
            Wh[i * 32 + k] = (float)re;
            Wh[i * 32 + 16 + k] = (float)-im;
            Wh[(16 + i) * 32 + k] = (float)im;
            Wh[(16 + i) * 32 + 16 + k] = (float)re;
        }
    const size_t bytes = h.size() * sizeof(float2);
    float2 *din, *dout;
    float* dW;
    CK(hipMalloc(&din, bytes));
    CK(hipMalloc(&dout, bytes));
    CK(hipMalloc(&dW, Wh.size() * sizeof(float)));
    CK(hipMemcpy(din, h.data(), bytes, hipMemcpyHostToDevice));
    CK(hipMemcpy(dW, Wh.data(), Wh.size() * sizeof(float), hipMemcpyHostToDevice));
    // reference: double precision, a sample of vectors
    const int sample = 512;
    std::vector<std::complex<double>> ref((size_t)sample * 16);
    for (int v = 0; v < sample; ++v) {
        const int vec = (int)(((long long)v * 2039) % kVectors);
        std::complex<double> x[16], y[16];
        for (int e = 0; e < 16; ++e) x[e] = {h[(size_t)vec * 16 + e].x, h[(size_t)vec * 16 + e].y};
        for (int r = 0; r < kReps; ++r) {
            for (int m = 0; m < 16; ++m) {
                y[m] = 0;
                for (int n = 0; n < 16; ++n) y[m] += x[n] * std::polar(1.0, -2 * pi * m * n / 16.0);
            }
            for (int e = 0; e < 16; ++e) x[e] = y[e] * 0.25;
        }
        for (int e = 0; e < 16; ++e) ref[(size_t)v * 16 + e] = x[e];
    }
    auto check = [&](const char* name, float ms) {
        std::vector<float2> o(h.size());
        CK(hipMemcpy(o.data(), dout, bytes, hipMemcpyDeviceToHost));
        double num = 0, den = 0;
        for (int v = 0; v < sample; ++v) {
            const int vec = (int)(((long long)v * 2039) % kVectors);
            for (int e = 0; e < 16; ++e) {
                const std::complex<double> g(o[(size_t)vec * 16 + e].x, o[(size_t)vec * 16 + e].y);
                num += std::norm(g - ref[(size_t)v * 16 + e]);
                den += std::norm(ref[(size_t)v * 16 + e]);
            }
        }
        const double stages = (double)kVectors * kReps;
        std::printf("%-8s %9.1f us  %6.3f ns per 16-point stage per CU  %7.1f TFLOP/s nominal "
                    "(5 N log2 N)  %7.1f TFLOP/s executed  rel-L2 %.2e\n",
                    name, ms * 1e3, ms * 1e6 * 256 / stages, stages * 320 / (ms * 1e-3) / 1e12,
                    stages * (name[0] == 'm' ? 2048.0 : 320.0) / (ms * 1e-3) / 1e12,
                    std::sqrt(num / den));
        std::fflush(stdout);
    };
    hipEvent_t a, b;
    CK(hipEventCreate(&a));
    CK(hipEventCreate(&b));
    const float scale = 0.25f;
    for (int pass = 0; pass < 2; ++pass) {
        // codelet: 256-thread workgroups, one vector per thread
        hipLaunchKernelGGL(k_codelet, dim3(kVectors / 256), dim3(256), 0, 0, din, dout, scale);
        CK(hipGetLastError());
        CK(hipEventRecord(a));
        hipLaunchKernelGGL(k_codelet, dim3(kVectors / 256), dim3(256), 0, 0, din, dout, scale);
        CK(hipEventRecord(b));
        CK(hipEventSynchronize(b));
        float ms = 0;
        CK(hipEventElapsedTime(&ms, a, b));
        if (pass) check("codelet", ms);
        // mfma: 4 waves per workgroup, 64 vectors per wave
        const int waves = kVectors / (16 * kBlocks);
        hipLaunchKernelGGL(k_mfma, dim3(waves / 4), dim3(256), 0, 0, din, dout, dW, scale);
        CK(hipGetLastError());
        CK(hipEventRecord(a));
        hipLaunchKernelGGL(k_mfma, dim3(waves / 4), dim3(256), 0, 0, din, dout, dW, scale);
        CK(hipEventRecord(b));
        CK(hipEventSynchronize(b));
        CK(hipEventElapsedTime(&ms, a, b));
        if (pass) check("mfma", ms);
    }
    CK(hipFree(din));
    CK(hipFree(dout));
    CK(hipFree(dW));
    std::printf("done\n");
    return 0;
}
