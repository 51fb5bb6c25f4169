// kernels_prec.h -- per-precision launch entry points (defined in kernels_f32.hip,
// kernels_f64.hip and kernels_f16.hip, dispatched on Prec by kernels_dispatch.cpp).
#pragma once
#include "kernels.h"

namespace cfft {

#define CFFT_DECLARE_LAUNCHERS(SFX)                                                          \
    hipError_t rows_##SFX(const void* in, void* out, const RowsArgs& a, const PassTable& pt, \
                          const void* tw, hipStream_t st);                                  \
    hipError_t fsa_##SFX(const void* in, void* out, const FourStepAArgs& a,                 \
                         const PassTable& pt, const void* tw, hipStream_t st);              \
    hipError_t fsb_##SFX(const void* in, void* out, const FourStepBArgs& a,                 \
                         const PassTable& pt, const void* tw, hipStream_t st);              \
    hipError_t conv_##SFX(const void* in, void* out, const ConvSmallArgs& a,                \
                          const PassTable& pt, const void* tw, hipStream_t st);

CFFT_DECLARE_LAUNCHERS(f32)
CFFT_DECLARE_LAUNCHERS(f64)
#undef CFFT_DECLARE_LAUNCHERS

hipError_t real_rows_f32(const void* in, void* out, const RealArgs& a, const PassTable& pt,
                         const void* tw, int E, hipStream_t st);
hipError_t real_rows_f64(const void* in, void* out, const RealArgs& a, const PassTable& pt,
                         const void* tw, int E, hipStream_t st);
hipError_t real_aux_f32(const void* in, void* out, const RealAuxArgs& a, hipStream_t st);
hipError_t transpose_f32(const void* src, void* dst, int64_t rows, int64_t cols, int64_t batch,
                         const void* tw3, int tbits, double sx, double sy, const void* K, void* dc,
                         hipStream_t st);
hipError_t transpose_f64(const void* src, void* dst, int64_t rows, int64_t cols, int64_t batch,
                         const void* tw3, int tbits, double sx, double sy, const void* K, void* dc,
                         hipStream_t st);
hipError_t real_aux_f64(const void* in, void* out, const RealAuxArgs& a, hipStream_t st);
void init_kernels_f32();
void init_kernels_f64();
hipError_t fsa_pair_f32(const void* in, void* out, const FourStepAArgs& a, const PassTable& pt,
                        const void* tw, hipStream_t st);
hipError_t fsb_pair_f32(const void* in, void* out, const FourStepBArgs& a, const PassTable& pt,
                        const void* tw, hipStream_t st);
void init_kernels_pair_f32();
void init_kernels_conv_f32();
void init_kernels_conv_f64();

hipError_t rows_f16(const void* in, void* out, const RowsArgs& a, const PassTable& pt,
                    const void* tw, hipStream_t st);
hipError_t fsa_f16(const void* in, void* out, const FourStepAArgs& a, const PassTable& pt,
                   const void* tw, hipStream_t st);
hipError_t fsb_f16(const void* in, void* out, const FourStepBArgs& a, const PassTable& pt,
                   const void* tw, hipStream_t st);
hipError_t convert_h2f(const void* src, void* dst, int64_t rows, int64_t n, int64_t sdist,
                       int64_t ddist, hipStream_t st);
hipError_t convert_f2h(const void* src, void* dst, int64_t rows, int64_t n, int64_t sdist,
                       int64_t ddist, hipStream_t st);
void init_kernels_f16();

hipError_t rows_pipe_f32(const void* in, void* out, const RowsArgs& a, const PassTable& pt,
                         const void* tw, size_t twn, hipStream_t st);
hipError_t rows_pipe_f64(const void* in, void* out, const RowsArgs& a, const PassTable& pt,
                         const void* tw, size_t twn, hipStream_t st);
int rows_pipe_occ_f32(int E, int R0, int block, size_t lds);
int rows_pipe_occ_f64(int E, int R0, int block, size_t lds);
hipError_t fsa_pipe_f32(const void* in, void* out, const FourStepAArgs& a, const PassTable& pt,
                        const void* tw, size_t twn, hipStream_t st);
hipError_t fsa_pipe_f64(const void* in, void* out, const FourStepAArgs& a, const PassTable& pt,
                        const void* tw, size_t twn, hipStream_t st);
hipError_t fsb_pipe_f32(const void* in, void* out, const FourStepBArgs& a, const PassTable& pt,
                        const void* tw, size_t twn, hipStream_t st);
hipError_t fsb_pipe_f64(const void* in, void* out, const FourStepBArgs& a, const PassTable& pt,
                        const void* tw, size_t twn, hipStream_t st);
int fs_pipe_occ_f32(int E, int pass, int R0, int block, size_t lds);
int fs_pipe_occ_f64(int E, int pass, int R0, int block, size_t lds);
void init_kernels_pipe_f32();
void init_kernels_pipe_f64();

hipError_t gather_f32_f32(const void* x, void* w, const ConvGatherArgs& a, hipStream_t st);
hipError_t gather_f32_f64(const void* x, void* w, const ConvGatherArgs& a, hipStream_t st);
hipError_t gather_f64_f64(const void* x, void* w, const ConvGatherArgs& a, hipStream_t st);
hipError_t scatter_f32_f32(const void* w, void* y, const ConvScatterArgs& a, hipStream_t st);
hipError_t scatter_f32_f64(const void* w, void* y, const ConvScatterArgs& a, hipStream_t st);
hipError_t scatter_f64_f64(const void* w, void* y, const ConvScatterArgs& a, hipStream_t st);

// Large convolutions over a four-step, three launches (k_convl_a, k_convl_ba, k_convl_c).
hipError_t convl_f32_f32(const void* x, void* y, const ConvLArgs& a, const PassTable& ptA,
                         const void* twA, const PassTable& ptB, const void* twB, hipStream_t st);
hipError_t convl_f32_f64(const void* x, void* y, const ConvLArgs& a, const PassTable& ptA,
                         const void* twA, const PassTable& ptB, const void* twB, hipStream_t st);
hipError_t convl_f64_f64(const void* x, void* y, const ConvLArgs& a, const PassTable& ptA,
                         const void* twA, const PassTable& ptB, const void* twB, hipStream_t st);
}  // namespace cfft
