// kernels_f32.hip -- single-precision instantiation of the kernels (kernels_impl.h).
#include "kernels_impl.h"
#include "kernels_prec.h"

namespace cfft {
using namespace detail;
#define CFFT_E(FN, T, E_, ...) ((E_) <= 8 ? FN<T, 8>(__VA_ARGS__) : FN<T, 16>(__VA_ARGS__))
hipError_t rows_f32(const void* in, void* out, const RowsArgs& a, const PassTable& pt,
                    const void* tw, hipStream_t st) {
    return rows_rm_t<float>(in, out, a, pt, tw, st);
}
hipError_t fsa_f32(const void* in, void* out, const FourStepAArgs& a, const PassTable& pt,
                   const void* tw, hipStream_t st) {
    return CFFT_E(fsa_t, float, a.E, in, out, a, pt, tw, st);
}
hipError_t fsb_f32(const void* in, void* out, const FourStepBArgs& a, const PassTable& pt,
                   const void* tw, hipStream_t st) {
    return CFFT_E(fsb_t, float, a.E, in, out, a, pt, tw, st);
}
hipError_t conv_f32(const void* in, void* out, const ConvSmallArgs& a, const PassTable& pt,
                    const void* tw, hipStream_t st) {
    return CFFT_E(conv_t, float, a.E, in, out, a, pt, tw, st);
}
hipError_t real_rows_f32(const void* in, void* out, const RealArgs& a, const PassTable& pt,
                         const void* tw, int E, hipStream_t st) {
    return CFFT_E(real_rows_t, float, E, in, out, a, pt, tw, st);
}
hipError_t transpose_f32(const void* src, void* dst, int64_t rows, int64_t cols,
                          int64_t batch, const void* tw3, int tbits, double sx, double sy,
                          const void* K, void* dc, hipStream_t st) {
    return transpose_t<float>(src, dst, rows, cols, batch, tw3, tbits, sx, sy, K, dc, st);
}
hipError_t real_aux_f32(const void* in, void* out, const RealAuxArgs& a, hipStream_t st) {
    return real_aux_t<float>(in, out, a, st);
}
// Lets every kernel of this precision use up to 160 KiB of LDS (called once per context).
void init_kernels_f32() {
    allow_big_lds(k_rows<float, 8>);
    allow_big_lds(k_rows<float, 16>);
    allow_big_lds(k_rows<float, 8, float2, 9>);
    allow_big_lds(k_rows<float, 16, float2, 9>);
    allow_big_lds(k_fs_a<float, 8>);
    allow_big_lds(k_fs_a<float, 16>);
    allow_big_lds(k_fs_a<float, 8, V_t<float>, false>);
    allow_big_lds(k_fs_a<float, 16, V_t<float>, false>);
    allow_big_lds(k_fs_b<float, 8>);
    allow_big_lds(k_fs_b<float, 16>);
    allow_big_lds(k_conv_small<float, 8>);
    allow_big_lds(k_conv_small<float, 16>);
    allow_big_lds(k_conv_small<float, 8, true>);
    allow_big_lds(k_conv_small<float, 16, true>);
    allow_big_lds(k_rows<float, 8, float2, 16, true>);
    allow_big_lds(k_rows<float, 16, float2, 16, true>);
    allow_big_lds(k_real_rows<float, 8>);
    allow_big_lds(k_real_rows<float, 16>);
}
}  // namespace cfft
