// kernels_pipe_f32.hip -- single-precision instantiation of the pipelined rows and four-step
// kernels (pipe_impl.h), per first radix (rows) and per-thread capacity.
#include "kernels_prec.h"
#include "pipe_impl.h"

namespace cfft {
using namespace detail;
hipError_t rows_pipe_f32(const void* in, void* out, const RowsArgs& a, const PassTable& pt,
                         const void* tw, size_t twn, hipStream_t st) {
    return a.E <= 8 ? rows_pipe_t<float, 8>(in, out, a, pt, tw, twn, st)
                    : rows_pipe_t<float, 16>(in, out, a, pt, tw, twn, st);
}
int rows_pipe_occ_f32(int E, int R0, int block, size_t lds) {
    return E <= 8 ? rows_pipe_occ_t<float, 8>(R0, block, lds)
                  : rows_pipe_occ_t<float, 16>(R0, block, lds);
}
hipError_t fsa_pipe_f32(const void* in, void* out, const FourStepAArgs& a, const PassTable& pt,
                        const void* tw, size_t twn, hipStream_t st) {
    return a.E <= 8 ? fsa_pipe_t<float, 8>(in, out, a, pt, tw, twn, st)
                    : fsa_pipe_t<float, 16>(in, out, a, pt, tw, twn, st);
}
hipError_t fsb_pipe_f32(const void* in, void* out, const FourStepBArgs& a, const PassTable& pt,
                        const void* tw, size_t twn, hipStream_t st) {
    return a.E <= 8 ? fsb_pipe_t<float, 8>(in, out, a, pt, tw, twn, st)
                    : fsb_pipe_t<float, 16>(in, out, a, pt, tw, twn, st);
}
int fs_pipe_occ_f32(int E, int pass, int R0, int block, size_t lds) {
    return E <= 8 ? fs_pipe_occ_t<float, 8>(pass, R0, block, lds)
                  : fs_pipe_occ_t<float, 16>(pass, R0, block, lds);
}
void init_kernels_pipe_f32() {
    rows_pipe_init_t<float, 8>();
    rows_pipe_init_t<float, 16>();
}
}  // namespace cfft
