// kernels_f16.hip -- ComplexF16 storage with float arithmetic.
//
// The rows and four-step kernels of kernels_impl.h instantiated with ch16 global lines: the
// first pass converts while it loads, the last pass rounds while it stores, so a direct plan
// moves 4 bytes in and 4 bytes out per element.  The four-step middle stays float (pass A
// writes it, pass B reads it).  Plan kinds without a fused form (convolutions, N-D) convert
// the whole input to float, run their float kernels and round the result back, with the two
// conversion kernels below.
#include "kernels_impl.h"
#include "kernels_prec.h"

namespace cfft {
using namespace detail;
#define CFFT_E(FN, T, E_, ...) ((E_) <= 8 ? FN<T, 8, ch16>(__VA_ARGS__) : FN<T, 16, ch16>(__VA_ARGS__))

hipError_t rows_f16(const void* in, void* out, const RowsArgs& a, const PassTable& pt,
                    const void* tw, hipStream_t st) {
    return rows_rm_t<float, ch16>(in, out, a, pt, tw, st);
}
hipError_t fsa_f16(const void* in, void* out, const FourStepAArgs& a, const PassTable& pt,
                   const void* tw, hipStream_t st) {
    return CFFT_E(fsa_t, float, a.E, in, out, a, pt, tw, st);
}
hipError_t fsb_f16(const void* in, void* out, const FourStepBArgs& a, const PassTable& pt,
                   const void* tw, hipStream_t st) {
    return CFFT_E(fsb_t, float, a.E, in, out, a, pt, tw, st);
}
#undef CFFT_E

// rows x n elements, row r of the source at r * sdist, of the destination at r * ddist.
__global__ void __launch_bounds__(256)
k_h2f(const ch16* __restrict__ src, float2* __restrict__ dst, int64_t rows, int64_t n,
      int64_t sdist, int64_t ddist) {
    for (int64_t r = blockIdx.y; r < rows; r += gridDim.y)
        for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
             e += (int64_t)gridDim.x * blockDim.x)
            dst[r * ddist + e] = load_elem<float>(src[r * sdist + e]);
}
__global__ void __launch_bounds__(256)
k_f2h(const float2* __restrict__ src, ch16* __restrict__ dst, int64_t rows, int64_t n,
      int64_t sdist, int64_t ddist) {
    for (int64_t r = blockIdx.y; r < rows; r += gridDim.y)
        for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
             e += (int64_t)gridDim.x * blockDim.x)
            store_elem<float>(dst[r * ddist + e], src[r * sdist + e]);
}

static dim3 convert_grid(int64_t rows, int64_t n) {
    const int64_t per = (n + 255) / 256;
    return dim3((unsigned)(per < 4096 ? (per > 0 ? per : 1) : 4096),
                (unsigned)(rows < 65535 ? (rows > 0 ? rows : 1) : 65535));
}
hipError_t convert_h2f(const void* src, void* dst, int64_t rows, int64_t n, int64_t sdist,
                       int64_t ddist, hipStream_t st) {
    if (rows <= 0 || n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_h2f, convert_grid(rows, n), dim3(256), 0, st, (const ch16*)src,
                       (float2*)dst, rows, n, sdist, ddist);
    return hipGetLastError();
}
hipError_t convert_f2h(const void* src, void* dst, int64_t rows, int64_t n, int64_t sdist,
                       int64_t ddist, hipStream_t st) {
    if (rows <= 0 || n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_f2h, convert_grid(rows, n), dim3(256), 0, st, (const float2*)src,
                       (ch16*)dst, rows, n, sdist, ddist);
    return hipGetLastError();
}

void init_kernels_f16() {
    allow_big_lds(k_rows<float, 8, ch16>);
    allow_big_lds(k_rows<float, 16, ch16>);
    allow_big_lds(k_rows<float, 8, ch16, 9>);
    allow_big_lds(k_rows<float, 16, ch16, 9>);
    allow_big_lds(k_fs_a<float, 8, ch16>);
    allow_big_lds(k_fs_a<float, 16, ch16>);
    allow_big_lds(k_fs_b<float, 8, ch16>);
    allow_big_lds(k_fs_b<float, 16, ch16>);
}
}  // namespace cfft
