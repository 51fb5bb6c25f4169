// kernels_f64.hip -- double-precision instantiation of the kernels (kernels_impl.h), plus the
// FP32-I/O / FP64-work gather and scatter of large convolutions.
#include "kernels_impl.h"
#include "kernels_prec.h"

namespace cfft {
using namespace detail;
#define CFFT_E(FN, T, E_, ...) ((E_) <= 8 ? FN<T, 8>(__VA_ARGS__) : FN<T, 16>(__VA_ARGS__))
hipError_t rows_f64(const void* in, void* out, const RowsArgs& a, const PassTable& pt,
                    const void* tw, hipStream_t st) {
    if (pt.wave)  // lines whose last pass runs across wavefront lanes
        return a.E <= 8 ? rows_t<double, 8, double2, 16, true>(in, out, a, pt, tw, st)
                        : rows_t<double, 16, double2, 16, true>(in, out, a, pt, tw, st);
    return CFFT_E(rows_t, double, a.E, in, out, a, pt, tw, st);
}
hipError_t fsa_f64(const void* in, void* out, const FourStepAArgs& a, const PassTable& pt,
                   const void* tw, hipStream_t st) {
    return CFFT_E(fsa_t, double, a.E, in, out, a, pt, tw, st);
}
hipError_t fsb_f64(const void* in, void* out, const FourStepBArgs& a, const PassTable& pt,
                   const void* tw, hipStream_t st) {
    return CFFT_E(fsb_t, double, a.E, in, out, a, pt, tw, st);
}
hipError_t conv_f64(const void* in, void* out, const ConvSmallArgs& a, const PassTable& pt,
                    const void* tw, hipStream_t st) {
    return CFFT_E(conv_t, double, a.E, in, out, a, pt, tw, st);
}
hipError_t real_rows_f64(const void* in, void* out, const RealArgs& a, const PassTable& pt,
                         const void* tw, int E, hipStream_t st) {
    return CFFT_E(real_rows_t, double, E, in, out, a, pt, tw, st);
}
hipError_t transpose_f64(const void* src, void* dst, int64_t rows, int64_t cols,
                          int64_t batch, const void* tw3, int tbits, double sx, double sy,
                          const void* K, void* dc, hipStream_t st) {
    return transpose_t<double>(src, dst, rows, cols, batch, tw3, tbits, sx, sy, K, dc, st);
}
hipError_t real_aux_f64(const void* in, void* out, const RealAuxArgs& a, hipStream_t st) {
    return real_aux_t<double>(in, out, a, st);
}
// Lets every kernel of this precision use up to 160 KiB of LDS (called once per context).
void init_kernels_f64() {
    allow_big_lds(k_rows<double, 8>);
    allow_big_lds(k_rows<double, 16>);
    allow_big_lds(k_fs_a<double, 8>);
    allow_big_lds(k_fs_a<double, 16>);
    allow_big_lds(k_fs_a<double, 8, V_t<double>, false>);
    allow_big_lds(k_fs_a<double, 16, V_t<double>, false>);
    allow_big_lds(k_fs_b<double, 8>);
    allow_big_lds(k_fs_b<double, 16>);
    allow_big_lds(k_conv_small<double, 8>);
    allow_big_lds(k_conv_small<double, 16>);
    allow_big_lds(k_conv_small<double, 8, true>);
    allow_big_lds(k_conv_small<double, 16, true>);
    allow_big_lds(k_rows<double, 8, double2, 16, true>);
    allow_big_lds(k_rows<double, 16, double2, 16, true>);
    allow_big_lds(k_real_rows<double, 8>);
    allow_big_lds(k_real_rows<double, 16>);
}
}  // namespace cfft
