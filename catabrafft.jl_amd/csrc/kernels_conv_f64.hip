// kernels_conv_f64.hip -- double-precision instantiation of the work-list kernels and of the
// large-convolution kernels (gather / scatter, the three fused launches) of kernels_impl.h; a
// translation unit of its own so that the precisions' kernels compile in parallel.
#include "kernels_impl.h"
#include "kernels_prec.h"

namespace cfft {
using namespace detail;
hipError_t gather_f32_f64(const void* x, void* w, const ConvGatherArgs& a, hipStream_t st) {
    return gather_t<float, double>(x, w, a, st);
}
hipError_t gather_f64_f64(const void* x, void* w, const ConvGatherArgs& a, hipStream_t st) {
    return gather_t<double, double>(x, w, a, st);
}
hipError_t scatter_f32_f64(const void* w, void* y, const ConvScatterArgs& a, hipStream_t st) {
    return scatter_t<float, double>(w, y, a, st);
}
hipError_t scatter_f64_f64(const void* w, void* y, const ConvScatterArgs& a, hipStream_t st) {
    return scatter_t<double, double>(w, y, a, st);
}
hipError_t convl_f32_f64(const void* x, void* y, const ConvLArgs& a, const PassTable& ptA,
                         const void* twA, const PassTable& ptB, const void* twB, hipStream_t st) {
    return a.E <= 8 ? convl_t<float, double, 8>(x, y, a, ptA, twA, ptB, twB, st)
                    : convl_t<float, double, 16>(x, y, a, ptA, twA, ptB, twB, st);
}
hipError_t convl_f64_f64(const void* x, void* y, const ConvLArgs& a, const PassTable& ptA,
                         const void* twA, const PassTable& ptB, const void* twB, hipStream_t st) {
    return a.E <= 8 ? convl_t<double, double, 8>(x, y, a, ptA, twA, ptB, twB, st)
                    : convl_t<double, double, 16>(x, y, a, ptA, twA, ptB, twB, st);
}
hipError_t multi_f64(const MultiLaunch& L, const MultiArgs& ma, hipStream_t st) {
    return multi_t<double>(L, ma, st);
}
hipError_t gs_multi_f32_f64(const MultiLaunch& L, const MultiArgs& ma, hipStream_t st) {
    return gs_multi_t<float, double>(L, ma, st);
}
hipError_t gs_multi_f64_f64(const MultiLaunch& L, const MultiArgs& ma, hipStream_t st) {
    return gs_multi_t<double, double>(L, ma, st);
}
// Lets these kernels use up to 160 KiB of LDS (called once per context).
void init_kernels_conv_f64() {
    allow_big_lds(k_convl_a<float, double, 8>);
    allow_big_lds(k_convl_a<float, double, 16>);
    allow_big_lds(k_convl_a<double, double, 8>);
    allow_big_lds(k_convl_a<double, double, 16>);
    allow_big_lds(k_convl_ba<double, 8>);
    allow_big_lds(k_convl_ba<double, 16>);
    allow_big_lds(k_convl_c<float, double, 8>);
    allow_big_lds(k_convl_c<float, double, 16>);
    allow_big_lds(k_convl_c<double, double, 8>);
    allow_big_lds(k_convl_c<double, double, 16>);
    allow_big_lds(k_convl_a_multi<float, double>);
    allow_big_lds(k_convl_c_multi<float, double>);
    allow_big_lds(k_convl_a_multi<double, double>);
    allow_big_lds(k_convl_c_multi<double, double>);
    allow_big_lds(k_convl_ba_multi<double>);
    allow_big_lds(k_rows_multi<double>);
    allow_big_lds(k_fs_a_multi<double>);
    allow_big_lds(k_fs_b_multi<double>);
    allow_big_lds(k_conv_multi<double>);
}
}  // namespace cfft
