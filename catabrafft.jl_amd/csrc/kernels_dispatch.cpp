// kernels_dispatch.cpp -- launch_* entry points of kernels.h, dispatched on precision.
#include "kernels.h"
#include "kernels_prec.h"
#include "multi.h"

namespace cfft {

void init_kernels() {
    init_kernels_f32();
    init_kernels_f64();
    init_kernels_pair_f32();
    init_kernels_conv_f32();
    init_kernels_conv_f64();
    init_kernels_f16();
    init_kernels_pipe_f32();
    init_kernels_pipe_f64();
}

hipError_t launch_rows_pipe(Prec p, const void* in, void* out, const RowsArgs& a,
                            const PassTable& pt, const void* tw, size_t twn, hipStream_t st) {
    return p == Prec::F32 ? rows_pipe_f32(in, out, a, pt, tw, twn, st)
                          : rows_pipe_f64(in, out, a, pt, tw, twn, st);
}
size_t rows_pipe_lds(Prec p, int C, int Ls, size_t twn) {
    const size_t el = p == Prec::F32 ? 4 : 8;
    const size_t off = (2 * el * (size_t)C * Ls + 15) & ~(size_t)15;
    return off + twn * 2 * el;
}
hipError_t launch_fourstep_a_pipe(Prec p, const void* in, void* out, const FourStepAArgs& a,
                                  const PassTable& pt, const void* tw, size_t twn,
                                  hipStream_t st) {
    return p == Prec::F32 ? fsa_pipe_f32(in, out, a, pt, tw, twn, st)
                          : fsa_pipe_f64(in, out, a, pt, tw, twn, st);
}
hipError_t launch_fourstep_b_pipe(Prec p, const void* in, void* out, const FourStepBArgs& a,
                                  const PassTable& pt, const void* tw, size_t twn,
                                  hipStream_t st) {
    return p == Prec::F32 ? fsb_pipe_f32(in, out, a, pt, tw, twn, st)
                          : fsb_pipe_f64(in, out, a, pt, tw, twn, st);
}
size_t fs_pipe_lds(Prec p, int C, int Ls, size_t twn, size_t t3n) {
    const size_t v = p == Prec::F32 ? 8 : 16;
    const size_t tw_off = (v * (size_t)C * Ls + 15) & ~(size_t)15;
    const size_t t3_off = (tw_off + twn * v + 15) & ~(size_t)15;
    return t3_off + t3n * v;
}
int fs_pipe_blocks_per_cu(Prec p, int E, int pass, int R0, int block, size_t lds) {
    return p == Prec::F32 ? fs_pipe_occ_f32(E, pass, R0, block, lds)
                          : fs_pipe_occ_f64(E, pass, R0, block, lds);
}
int rows_pipe_blocks_per_cu(Prec p, int E, int R0, int block, size_t lds) {
    return p == Prec::F32 ? rows_pipe_occ_f32(E, R0, block, lds)
                          : rows_pipe_occ_f64(E, R0, block, lds);
}

hipError_t launch_rows(Prec p, const void* in, void* out, const RowsArgs& a, const PassTable& pt,
                       const void* tw, hipStream_t st) {
    return p == Prec::F32 ? rows_f32(in, out, a, pt, tw, st) : rows_f64(in, out, a, pt, tw, st);
}
hipError_t launch_rows_f16(const void* in, void* out, const RowsArgs& a, const PassTable& pt,
                           const void* tw, hipStream_t st) {
    return rows_f16(in, out, a, pt, tw, st);
}
hipError_t launch_fourstep_a_f16(const void* in, void* out, const FourStepAArgs& a,
                                 const PassTable& pt, const void* tw, hipStream_t st) {
    return fsa_f16(in, out, a, pt, tw, st);
}
hipError_t launch_fourstep_b_f16(const void* in, void* out, const FourStepBArgs& a,
                                 const PassTable& pt, const void* tw, hipStream_t st) {
    return fsb_f16(in, out, a, pt, tw, st);
}
hipError_t launch_convert_h2f(const void* src, void* dst, int64_t rows, int64_t n, int64_t sdist,
                              int64_t ddist, hipStream_t st) {
    return convert_h2f(src, dst, rows, n, sdist, ddist, st);
}
hipError_t launch_convert_f2h(const void* src, void* dst, int64_t rows, int64_t n, int64_t sdist,
                              int64_t ddist, hipStream_t st) {
    return convert_f2h(src, dst, rows, n, sdist, ddist, st);
}
hipError_t launch_fourstep_a(Prec p, const void* in, void* out, const FourStepAArgs& a,
                             const PassTable& pt, const void* tw, hipStream_t st) {
    return p == Prec::F32 ? fsa_f32(in, out, a, pt, tw, st) : fsa_f64(in, out, a, pt, tw, st);
}
hipError_t launch_fourstep_b(Prec p, const void* in, void* out, const FourStepBArgs& a,
                             const PassTable& pt, const void* tw, hipStream_t st) {
    return p == Prec::F32 ? fsb_f32(in, out, a, pt, tw, st) : fsb_f64(in, out, a, pt, tw, st);
}
hipError_t launch_fourstep_a_pair(const void* in, void* out, const FourStepAArgs& a,
                                  const PassTable& pt, const void* tw, hipStream_t st) {
    return fsa_pair_f32(in, out, a, pt, tw, st);
}
hipError_t launch_fourstep_b_pair(const void* in, void* out, const FourStepBArgs& a,
                                  const PassTable& pt, const void* tw, hipStream_t st) {
    return fsb_pair_f32(in, out, a, pt, tw, st);
}
hipError_t launch_conv_small(Prec p, const void* in, void* out, const ConvSmallArgs& a,
                             const PassTable& pt, const void* tw, hipStream_t st) {
    return p == Prec::F32 ? conv_f32(in, out, a, pt, tw, st) : conv_f64(in, out, a, pt, tw, st);
}
hipError_t launch_real_rows(Prec p, const void* in, void* out, const RealArgs& a,
                            const PassTable& pt, const void* tw, int E, hipStream_t st) {
    return p == Prec::F32 ? real_rows_f32(in, out, a, pt, tw, E, st)
                          : real_rows_f64(in, out, a, pt, tw, E, st);
}
hipError_t launch_transpose(Prec p, const void* src, void* dst, int64_t rows, int64_t cols,
                            int64_t batch, const void* tw3, int tbits, double sx, double sy,
                            const void* K, void* dc, hipStream_t st) {
    return p == Prec::F32
               ? transpose_f32(src, dst, rows, cols, batch, tw3, tbits, sx, sy, K, dc, st)
               : transpose_f64(src, dst, rows, cols, batch, tw3, tbits, sx, sy, K, dc, st);
}
hipError_t launch_real_aux(Prec p, const void* in, void* out, const RealAuxArgs& a,
                           hipStream_t st) {
    return p == Prec::F32 ? real_aux_f32(in, out, a, st) : real_aux_f64(in, out, a, st);
}
hipError_t launch_conv_gather(Prec pio, Prec pc, const void* x, void* w, const ConvGatherArgs& a,
                              hipStream_t st) {
    if (pio == Prec::F32) return pc == Prec::F32 ? gather_f32_f32(x, w, a, st) : gather_f32_f64(x, w, a, st);
    return gather_f64_f64(x, w, a, st);
}
hipError_t launch_conv_scatter(Prec pio, Prec pc, const void* w, void* y,
                               const ConvScatterArgs& a, hipStream_t st) {
    if (pio == Prec::F32) return pc == Prec::F32 ? scatter_f32_f32(w, y, a, st) : scatter_f32_f64(w, y, a, st);
    return scatter_f64_f64(w, y, a, st);
}

hipError_t launch_convl(Prec pio, Prec pc, const void* x, void* y, const ConvLArgs& a,
                        const PassTable& ptA, const void* twA, const PassTable& ptB,
                        const void* twB, hipStream_t st) {
    if (pio == Prec::F32)
        return pc == Prec::F32 ? convl_f32_f32(x, y, a, ptA, twA, ptB, twB, st)
                               : convl_f32_f64(x, y, a, ptA, twA, ptB, twB, st);
    return convl_f64_f64(x, y, a, ptA, twA, ptB, twB, st);
}

hipError_t launch_multi(const MultiLaunch& L, const MultiArgs& ma, hipStream_t st) {
    if (L.nitems <= 0) return hipSuccess;
    const MultiKernel k = (MultiKernel)L.kernel;
    if (k == MultiKernel::Gather || k == MultiKernel::Scatter || k == MultiKernel::ConvLA ||
        k == MultiKernel::ConvLBA || k == MultiKernel::ConvLC) {
        if (L.prec == (int)Prec::F32)
            return L.cprec == (int)Prec::F32 ? gs_multi_f32_f32(L, ma, st)
                                             : gs_multi_f32_f64(L, ma, st);
        return gs_multi_f64_f64(L, ma, st);
    }
    return L.prec == (int)Prec::F32 ? multi_f32(L, ma, st)
                                    : multi_f64(L, ma, st);
}

}  // namespace cfft
