// kernels_conv_f32.hip -- single-precision instantiation of the work-list kernels and of the
// large-convolution kernels (gather / scatter, the three fused launches) of kernels_impl.h; a
// translation unit of its own so that the precisions' kernels compile in parallel.
#include "kernels_impl.h"
#include "kernels_prec.h"

namespace cfft {
using namespace detail;
hipError_t gather_f32_f32(const void* x, void* w, const ConvGatherArgs& a, hipStream_t st) {
    return gather_t<float, float>(x, w, a, st);
}
hipError_t scatter_f32_f32(const void* w, void* y, const ConvScatterArgs& a, hipStream_t st) {
    return scatter_t<float, float>(w, y, a, st);
}
hipError_t multi_f32(const MultiLaunch& L, const MultiArgs& ma, hipStream_t st) {
    return multi_t<float>(L, ma, st);
}
hipError_t gs_multi_f32_f32(const MultiLaunch& L, const MultiArgs& ma, hipStream_t st) {
    return gs_multi_t<float, float>(L, ma, st);
}
hipError_t convl_f32_f32(const void* x, void* y, const ConvLArgs& a, const PassTable& ptA,
                         const void* twA, const PassTable& ptB, const void* twB, hipStream_t st) {
    return a.E <= 8 ? convl_t<float, float, 8>(x, y, a, ptA, twA, ptB, twB, st)
                    : convl_t<float, float, 16>(x, y, a, ptA, twA, ptB, twB, st);
}
// Lets these kernels use up to 160 KiB of LDS (called once per context).
void init_kernels_conv_f32() {
    allow_big_lds(k_convl_a<float, float, 8>);
    allow_big_lds(k_convl_a<float, float, 16>);
    allow_big_lds(k_convl_ba<float, 8>);
    allow_big_lds(k_convl_ba<float, 16>);
    allow_big_lds(k_convl_c<float, float, 8>);
    allow_big_lds(k_convl_c<float, float, 16>);
    allow_big_lds(k_convl_a_multi<float, float>);
    allow_big_lds(k_convl_c_multi<float, float>);
    allow_big_lds(k_convl_ba_multi<float>);
    allow_big_lds(k_rows_multi<float>);
    allow_big_lds(k_fs_a_multi<float>);
    allow_big_lds(k_fs_b_multi<float>);
    allow_big_lds(k_conv_multi<float>);
}
}  // namespace cfft
