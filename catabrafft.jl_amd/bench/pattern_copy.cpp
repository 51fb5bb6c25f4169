// pattern_copy -- the HBM access patterns of the four-step passes, with no arithmetic.
//
// Each workgroup copies one tile the way a four-step pass moves it: every thread loads its
// elements of the tile (all loads first), then stores them.  The tile shapes, workgroup sizes,
// LDS allocations (which set the workgroups resident per CU), the XCD remap of block indices (also
// timed without it) and
// the store kinds are those of the passes for 2^20-point transforms:
//   pass A: C adjacent columns of the N1 x N2 row-major input (runs of C elements, row stride
//           N2), each column written as a contiguous line of the middle with streaming stores;
//   pass B: runs of C lines k1 of the middle (stride N1), written in runs of C at stride N1.
// ComplexF32 passes on line pairs load and store 16 bytes per lane (two adjacent columns).
// The time of a pattern bounds what a pass with that tile can reach; comparing tile widths at
// equal occupancy shows what wider runs would give.
//
// Build: hipcc --offload-arch=gfx950 -O3 -o pattern_copy pattern_copy.cpp -x hip
// Run:   pattern_copy   (prints one line per pattern: time and read + write rate)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CK(x)                                                                            \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess) {                                                          \
            std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__);  \
            std::exit(1);                                                                \
        }                                                                                \
    } while (0)

template <typename EL, int VPL>
struct alignas(sizeof(EL) * VPL) Pack {
    EL v[VPL];
};

// Streaming (non-temporal) store of one complex element, as the passes' middle stores.
__device__ __forceinline__ void nt_store(float2 v, float2* p) {
    typedef float W __attribute__((ext_vector_type(2)));
    __builtin_nontemporal_store(W{v.x, v.y}, reinterpret_cast<W*>(p));
}
__device__ __forceinline__ void nt_store(double2 v, double2* p) {
    typedef double W __attribute__((ext_vector_type(2)));
    __builtin_nontemporal_store(W{v.x, v.y}, reinterpret_cast<W*>(p));
}

// Block b runs on XCD b mod 8; remapped, XCD x takes the consecutive tiles [x G/8, (x+1) G/8).
__device__ __forceinline__ int remap_block(int remap) {
    const int b = blockIdx.x, G = gridDim.x;
    if (!remap || (G & 7)) return b;
    return (b & 7) * (G >> 3) + (b >> 3);
}

// PB = false: pass A pattern; true: pass B pattern.  VPL elements per lane (adjacent lines).
template <typename EL, int VPL, int E, bool PB>
__global__ void __launch_bounds__(1024)
k_pattern(const EL* __restrict__ in, EL* __restrict__ out, int N1, int N2, int C, int remap) {
    extern __shared__ unsigned char lds[];
    const int LPR = C / VPL;                 // lanes per matrix row
    const int RPI = (int)blockDim.x / LPR;   // rows per load instruction
    const int lines = PB ? N1 : N2;
    const int tiles = lines / C;
    const int bid = remap_block(remap);
    const long long bt = bid / tiles;
    const int tile = bid - (int)(bt * tiles);
    const int c = (int)threadIdx.x % LPR, t = (int)threadIdx.x / LPR;
    const long long n = (long long)N1 * N2;
    const EL* src = in + bt * n;
    EL* dst = out + bt * n;
    const int col = tile * C + c * VPL;
    const int sstride = PB ? N1 : N2;
    Pack<EL, VPL> r[E];
#pragma unroll
    for (int k = 0; k < E; ++k)
        r[k] = *reinterpret_cast<const Pack<EL, VPL>*>(src + (long long)(t + k * RPI) * sstride + col);
    // The LDS allocation only sets occupancy; this read (never true for the zeroed input) keeps
    // the compiler from dropping it.
    if (threadIdx.x == 0 && lds[0] == 0x5a && r[0].v[0].x == 12345) r[0].v[0].y = 1;
#pragma unroll
    for (int k = 0; k < E; ++k) {
        const int e = t + k * RPI;
        if (PB) {
            *reinterpret_cast<Pack<EL, VPL>*>(dst + (long long)e * N1 + col) = r[k];
        } else {
#pragma unroll
            for (int v = 0; v < VPL; ++v)
                nt_store(r[k].v[v], dst + (long long)(col + v) * N1 + e);
        }
    }
}

template <typename EL, int VPL, int E, bool PB>
static void run(const char* what, void* in, void* out, int N1, int N2, int batch, int C, int T,
                size_t lds, int remap = 1) {
    const int LPR = C / VPL;
    const int lines = PB ? N1 : N2, len = PB ? N2 : N1;
    // every index stays inside one transform: rows t + k RPI < len, columns < lines
    if (C % VPL || T % LPR || (T / LPR) * E != len || lines % C || T > 1024 || lds < 1 ||
        lds > 160 * 1024 || (size_t)N1 * N2 * batch * sizeof(EL) > (1ull << 30)) {
        std::printf("%-46s shape rejected\n", what);
        return;
    }
    auto k = k_pattern<EL, VPL, E, PB>;
    CK(hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                           hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    const unsigned blocks = (unsigned)(lines / C) * batch;
    const EL* i = static_cast<const EL*>(in);
    EL* o = static_cast<EL*>(out);
    hipEvent_t a, b;
    CK(hipEventCreate(&a));
    CK(hipEventCreate(&b));
    for (int w = 0; w < 3; ++w) hipLaunchKernelGGL(k, dim3(blocks), dim3(T), lds, 0, i, o, N1, N2, C, remap);
    CK(hipGetLastError());
    const int reps = 20;
    CK(hipEventRecord(a));
    for (int w = 0; w < reps; ++w) hipLaunchKernelGGL(k, dim3(blocks), dim3(T), lds, 0, i, o, N1, N2, C, remap);
    CK(hipEventRecord(b));
    CK(hipEventSynchronize(b));
    float ms = 0;
    CK(hipEventElapsedTime(&ms, a, b));
    ms /= reps;
    int occ = 0;
    CK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void*>(k), T, lds));
    const double bytes = 2.0 * (double)N1 * N2 * batch * sizeof(EL);
    std::printf("%-46s run %3d B  %4d threads  %d WG/CU  %7.1f us  %.2f TB/s\n", what,
                (int)(C * sizeof(EL)), T, occ, ms * 1e3, bytes / ms / 1e9);
    std::fflush(stdout);
    CK(hipEventDestroy(a));
    CK(hipEventDestroy(b));
}

int main() {
    const size_t bytes = 1ull << 30;
    void *in, *out;
    CK(hipMalloc(&in, bytes));
    CK(hipMalloc(&out, bytes));
    CK(hipMemset(in, 0, bytes));
    CK(hipMemset(out, 0, bytes));
    // 2^20 x 64 ComplexF32: N1 = 2048, N2 = 512, line pairs (16-byte lanes)
    run<float2, 2, 8, false>("F32 A, 2 pairs (bench tile, 2048 x 512)", in, out, 2048, 512, 64, 4, 512, 69 * 1024);
    run<float2, 2, 8, false>("F32 A, 4 pairs, LDS of 4 pairs", in, out, 2048, 512, 64, 8, 1024, 138 * 1024);
    run<float2, 2, 16, false>("F32 A, 4 pairs, LDS of 2 pairs", in, out, 2048, 512, 64, 8, 512, 69 * 1024);
    run<float2, 2, 8, false>("F32 A, 2 pairs, no XCD remap", in, out, 2048, 512, 64, 4, 512, 69 * 1024, 0);
    run<float2, 2, 8, true>("F32 B, 8 pairs (bench tile)", in, out, 2048, 512, 64, 16, 512, 70 * 1024);
    // 2^20 x 64 ComplexF64: N1 = N2 = 1024
    run<double2, 1, 8, false>("F64 A, 4 columns (bench tile)", in, out, 1024, 1024, 64, 4, 512, 70 * 1024);
    run<double2, 1, 8, false>("F64 A, 8 columns, LDS of 8", in, out, 1024, 1024, 64, 8, 1024, 140 * 1024);
    run<double2, 1, 16, false>("F64 A, 8 columns, LDS of 4", in, out, 1024, 1024, 64, 8, 512, 70 * 1024);
    run<double2, 1, 8, false>("F64 A, 4 columns, no XCD remap", in, out, 1024, 1024, 64, 4, 512, 70 * 1024, 0);
    run<double2, 1, 8, true>("F64 B, 4 lines (bench tile)", in, out, 1024, 1024, 64, 4, 512, 72 * 1024);
    run<double2, 1, 8, true>("F64 B, 8 lines, LDS of 8", in, out, 1024, 1024, 64, 8, 1024, 140 * 1024);
    run<double2, 1, 16, true>("F64 B, 8 lines, LDS of 4", in, out, 1024, 1024, 64, 8, 512, 72 * 1024);
    run<double2, 1, 8, true>("F64 B, 4 lines, no XCD remap", in, out, 1024, 1024, 64, 4, 512, 72 * 1024, 0);
    CK(hipFree(in));
    CK(hipFree(out));
    std::printf("done\n");
    return 0;
}
