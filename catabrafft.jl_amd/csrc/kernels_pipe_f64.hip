// kernels_pipe_f64.hip -- double-precision instantiation of the pipelined rows kernel
// (pipe_impl.h), one kernel per first radix and per-thread capacity.
#include "kernels_prec.h"
#include "pipe_impl.h"

namespace cfft {
using namespace detail;
hipError_t rows_pipe_f64(const void* in, void* out, const RowsArgs& a, const PassTable& pt,
                         const void* tw, size_t twn, hipStream_t st) {
    return a.E <= 8 ? rows_pipe_t<double, 8>(in, out, a, pt, tw, twn, st)
                    : rows_pipe_t<double, 16>(in, out, a, pt, tw, twn, st);
}
int rows_pipe_occ_f64(int E, int R0, int block, size_t lds) {
    return E <= 8 ? rows_pipe_occ_t<double, 8>(R0, block, lds)
                  : rows_pipe_occ_t<double, 16>(R0, block, lds);
}
void init_kernels_pipe_f64() {
    rows_pipe_init_t<double, 8>();
    rows_pipe_init_t<double, 16>();
}
}  // namespace cfft
