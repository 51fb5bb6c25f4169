// pipe_impl.h -- persistent, software-pipelined rows and four-step kernels (gfx950).
//
// Each workgroup loops over groups of C lines.  The first pass's inputs of the NEXT group are
// loaded into registers while the current group runs its LDS passes and stores, so HBM reads are
// in flight during the whole pass sequence instead of only at the start of a workgroup's life.
// vmcnt is in order, so any global load issued after the prefetch would wait for it: the twiddle
// table therefore lives in LDS (copied once per workgroup, k-major per pass), and the only global
// operations in the loop are the prefetch and the last pass's stores.
//
// The arithmetic is the rows kernel's (the same Stockham passes and codelets), so its results are
// bit-identical to k_rows for the same plan.
#pragma once
#include "kernels_impl.h"

namespace cfft {

// LDS offset of the twiddle copy: after the C work lines, 16-byte aligned.
template <typename T>
__host__ __device__ constexpr size_t pipe_tw_offset(int C, int Ls) {
    return (2 * sizeof(T) * (size_t)C * Ls + 15) & ~(size_t)15;
}

// Copies the plan's twiddle table (row-major per pass, [p (R-1) + k - 1]) into LDS, k-major per
// pass ([(k-1) n1 + p]).
template <typename T>
__device__ __forceinline__ void stage_twiddles(const PassTable& pt, int N, const V_t<T>* tw,
                                               V_t<T>* twl) {
    int s = 1;
    for (int ip = 0; ip < pt.npass; ++ip) {
        const int R = pt.radix[ip];
        const int n1 = N / (s * R);
        const int off = pt.twoff[ip];
        batched_for<kStageU>(
            threadIdx.x, blockDim.x, n1 * (R - 1), [&](int i) { return tw[off + i]; },
            [&](int i, V_t<T> v) {
                const int p = i / (R - 1);
                twl[off + (i - p * (R - 1)) * n1 + p] = v;
            });
        s *= R;
    }
}

// First pass (s = 1) of a line whose inputs x[ib][j] = in[b + j N/R] (b = t + ib Tt) are already
// in registers; sg = -1 conjugates them.  The pass is twiddled (the line has more than one pass).
template <typename T, int R, int MB, class DST, class TWA>
__device__ __forceinline__ void first_pass_regs(const int N, const int t, const int Tt,
                                                const bool active, const TWA& tw,
                                                const V_t<T> (&x)[MB][R], const T sg,
                                                const DST& dst) {
    using V = V_t<T>;
    const int nb = N / R;
#pragma unroll
    for (int ib = 0; ib < MB; ++ib) {
        const int b = t + ib * Tt;
        if (active && b < nb) {
            V v[R];
#pragma unroll
            for (int j = 0; j < R; ++j) v[j] = V{x[ib][j].x, x[ib][j].y * sg};
            Dft<R>::run(v);
#pragma unroll
            for (int k = 1; k < R; ++k) v[k] = cmul(v[k], tw.template get<R>(b, k));
#pragma unroll
            for (int k = 0; k < R; ++k) dst.put(R * b + k, v[k]);
        }
    }
}

// Persistent rows kernel for plans whose first radix is R0 and that have at least two passes.
// Grid: at most the number of workgroups resident on the device; group g runs on workgroup
// g mod gridDim.x.  Workgroups of at most kPipeMaxThreads threads: the prefetched inputs stay
// live across the whole pass sequence, which needs more than the 128 registers per lane that
// 1024-thread workgroups allow.

template <typename T, int E, int R0>
__global__ void __launch_bounds__(kPipeMaxThreads)
k_rows_pipe(const V_t<T>* __restrict__ in, V_t<T>* __restrict__ out, RowsArgs a, PassTable pt,
            const V_t<T>* __restrict__ tw) {
    CFFT_DYN_LDS(smem);
    using V = V_t<T>;
    constexpr int MB = bfly_per_thread(E, R0);
    const int Tt = a.Tt, C = a.C, N = a.N;
    const int tid = threadIdx.x;
    const int c = tid / Tt;
    const int t = tid - c * Tt;
    V* twl = reinterpret_cast<V*>(smem + pipe_tw_offset<T>(C, a.Ls));
    stage_twiddles<T>(pt, N, tw, twl);
    const int64_t ngroups = (a.nlines + C - 1) / C;
    const int nb0 = N / R0;
    V x[MB][R0];
    // Loads group g's first-pass inputs of this thread (clamped indices: always valid loads).
    auto load_group = [&](int64_t g) {
        int64_t line = g * C + c;
        if (line >= a.nlines) line = a.nlines - 1;
        const V* src = in + line * a.idist;
#pragma unroll
        for (int ib = 0; ib < MB; ++ib) {
            const int bt = t + ib * Tt;
            const int b = bt < nb0 ? bt : nb0 - 1;
#pragma unroll
            for (int j = 0; j < R0; ++j) x[ib][j] = src[b + j * nb0];
        }
    };
    int64_t g = blockIdx.x;
    if (g < ngroups) load_group(g);
    __syncthreads();  // the twiddle copy is complete
    const LdsLine<T> L = lds_line<T>(smem, C, a.Ls, c);
    const TwLds<V> twt{twl};
    const T sg = (T)a.sgn_in;
    for (; g < ngroups; g += gridDim.x) {
        const int64_t line = g * C + c;
        const bool active = line < a.nlines;
        first_pass_regs<T, R0, MB>(N, t, Tt, active, twt.pass(pt.twoff[0], nb0), x, sg, L);
        __syncthreads();
        if (g + gridDim.x < ngroups) load_group(g + gridDim.x);
        GRowDst<T> gd{out + (active ? line : 0) * a.odist, (T)a.scale, (T)(a.scale * a.sgn_in),
                      nullptr, nullptr};
        run_line_t<T, E>(pt, N, t, Tt, active, twt, L, gd, L, false, true, false, 1, R0);
        __syncthreads();  // the last pass's LDS reads precede the next group's first pass
    }
}

// ------------------------------------------------------------------------------------------
// Pipelined four-step passes.  A workgroup loops over tiles (pass A: C columns of one transform;
// pass B: C rows of the middle) with the plain kernels' column-first lanes (lane = c + C t).  The
// first Stockham pass's inputs of the NEXT tile are loaded straight into registers while the
// current tile runs its remaining passes, so each CU keeps reads in flight through the whole pass
// sequence.  The pass twiddles (and pass B's three-level table) live in LDS: vmcnt counts in
// order, so a global twiddle load issued after the prefetch would wait for it.  Same passes,
// twiddle values and operation order as k_fs_a / k_fs_b, hence the same bits.
//
// Tile order: iteration i of workgroup g runs tile i G + (g mod 8) (G / 8) + g / 8 (G = gridDim.x,
// a multiple of 8; else i G + g), so the tiles in flight on one XCD at a time are consecutive.
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ int64_t pipe_tile(int64_t i, int remap) {
    const int g = blockIdx.x, G = gridDim.x;
    const int gg = (!remap || (G & 7)) ? g : (g & 7) * (G >> 3) + (g >> 3);
    return i * G + gg;
}

// LDS offset of pass B's three-level twiddle table: after the lines and the pass twiddles.
template <typename T>
__host__ __device__ constexpr size_t pipe_t3_offset(int C, int Ls, int twn) {
    return (pipe_tw_offset<T>(C, Ls) + (size_t)twn * sizeof(V_t<T>) + 15) & ~(size_t)15;
}

template <typename T, int E, int R0>
__global__ void __launch_bounds__(kPipeMaxThreads)
k_fs_a_pipe(const V_t<T>* __restrict__ in, V_t<T>* __restrict__ out, FourStepAArgs a,
            PassTable pt, const V_t<T>* __restrict__ tw) {
    CFFT_DYN_LDS(smem);
    using V = V_t<T>;
    constexpr int MB = bfly_per_thread(E, R0);
    const int C = a.C, N1 = a.N1, N2 = a.N2, Tt = a.Tt;
    const int cshift = __ffs(C) - 1;
    const int tid = threadIdx.x;
    const int c = tid & (C - 1);
    const int t = tid >> cshift;
    const int tiles = (N2 + C - 1) / C;
    const int64_t ntiles = (int64_t)tiles * a.batch;
    const int nb0 = N1 / R0;
    V* twl = reinterpret_cast<V*>(smem + pipe_tw_offset<T>(C, a.Ls));
    V x[MB][R0];
    // this thread's first-pass inputs of tile q: x[ib][j] = column j2 element b + j nb0
    auto load_tile = [&](int64_t q) {
        const int64_t bt = q / tiles;
        const int j20 = (int)(q - bt * tiles) * C;
        const int j2 = j20 + c < N2 ? j20 + c : N2 - 1;
        const V* src = in + bt * a.idist + j2;
#pragma unroll
        for (int ib = 0; ib < MB; ++ib) {
            const int bt_ = t + ib * Tt;
            const int b = bt_ < nb0 ? bt_ : nb0 - 1;
#pragma unroll
            for (int j = 0; j < R0; ++j) x[ib][j] = src[(int64_t)(b + j * nb0) * N2];
        }
    };
    int64_t i = 0;
    int64_t q = pipe_tile(0, a.xcd);
    if (q < ntiles) load_tile(q);
    stage_twiddles<T>(pt, N1, tw, twl);
    __syncthreads();  // the twiddle copy is complete
    const LdsLine<T> L = lds_line<T>(smem, C, a.Ls, c);
    const TwLds<V> twt{twl};
    const T sg = (T)a.sgn_in;
    while (q < ntiles) {
        const int64_t bt = q / tiles;
        const int j2 = (int)(q - bt * tiles) * C + c;
        const bool active = j2 < N2;
        first_pass_regs<T, R0, MB>(N1, t, Tt, active, twt.pass(pt.twoff[0], nb0), x, sg, L);
        __syncthreads();
        const int64_t qn = pipe_tile(++i, a.xcd);
        if (qn < ntiles) load_tile(qn);
        GPlainDst<T> d{out + bt * a.odist + (int64_t)(active ? j2 : 0) * N1};
        run_line_t<T, E>(pt, N1, t, Tt, active, twt, L, d, L, false, true, false, 1, R0);
        __syncthreads();  // the last pass's LDS reads precede the next tile's first pass
        q = qn;
    }
}

template <typename T, int E, int R0>
__global__ void __launch_bounds__(kPipeMaxThreads)
k_fs_b_pipe(const V_t<T>* in, V_t<T>* out, FourStepBArgs a, PassTable pt,
            const V_t<T>* __restrict__ tw, int twn) {
    CFFT_DYN_LDS(smem);
    using V = V_t<T>;
    constexpr int MB = bfly_per_thread(E, R0);
    const int C = a.C, N1 = a.N1, N2 = a.N2, Tt = a.Tt;
    const int cshift = __ffs(C) - 1;
    const int tid = threadIdx.x;
    const int c = tid & (C - 1);
    const int t = tid >> cshift;
    const int tiles = (N1 + C - 1) / C;
    const int64_t ntiles = (int64_t)tiles * a.batch;
    const int nb0 = N2 / R0;
    V* twl = reinterpret_cast<V*>(smem + pipe_tw_offset<T>(C, a.Ls));
    V* t3 = reinterpret_cast<V*>(smem + pipe_t3_offset<T>(C, a.Ls, twn));
    const V* tw3 = (const V*)a.tw3;
    V x[MB][R0];
    // this thread's first-pass inputs of tile q: x[ib][j] = middle row b + j nb0, column k1
    auto load_tile = [&](int64_t q) {
        const int64_t bt = q / tiles;
        const int k10 = (int)(q - bt * tiles) * C;
        const int k1 = k10 + c < N1 ? k10 + c : N1 - 1;
        const V* src = in + bt * a.idist + k1;
#pragma unroll
        for (int ib = 0; ib < MB; ++ib) {
            const int bt_ = t + ib * Tt;
            const int b = bt_ < nb0 ? bt_ : nb0 - 1;
#pragma unroll
            for (int j = 0; j < R0; ++j) x[ib][j] = src[(int64_t)(b + j * nb0) * N1];
        }
    };
    int64_t i = 0;
    int64_t q = pipe_tile(0, a.xcd);
    if (q < ntiles) load_tile(q);
    stage_twiddles<T>(pt, N2, tw, twl);
    if (tw3) stage_table<T>(tw3, t3, tw3_entries(a.tbits, a.tnc));
    __syncthreads();  // the tables are in LDS
    const LdsLine<T> L = lds_line<T>(smem, C, a.Ls, c);
    const TwLds<V> twt{twl};
    const T sg = (T)a.sgn_in;
    const T sx = (T)a.scale, sy = (T)(a.scale * a.sgn_out);
    while (q < ntiles) {
        const int64_t bt = q / tiles;
        const int k1 = (int)(q - bt * tiles) * C + c;
        const bool active = k1 < N1;
        if (tw3) {  // the four-step twiddle w_n^{e k1}, before the conjugation in the first pass
#pragma unroll
            for (int ib = 0; ib < MB; ++ib) {
                const int b = t + ib * Tt;
#pragma unroll
                for (int j = 0; j < R0; ++j)
                    x[ib][j] = cmul(x[ib][j], tw3_get<T>(t3, a.tbits, (b + j * nb0) * k1));
            }
        }
        first_pass_regs<T, R0, MB>(N2, t, Tt, active, twt.pass(pt.twoff[0], nb0), x, sg, L);
        __syncthreads();
        const int64_t qn = pipe_tile(++i, a.xcd);
        if (qn < ntiles) load_tile(qn);
        GStrideDst<T> d{out + bt * a.odist + (active ? k1 : 0), N1, sx, sy};
        run_line_t<T, E>(pt, N2, t, Tt, active, twt, L, d, L, false, true, false, 1, R0);
        __syncthreads();
        q = qn;
    }
}

namespace detail {

template <typename T, int E, int R0>
hipError_t rows_pipe_r(const void* in, void* out, const RowsArgs& a, const PassTable& pt,
                       const void* tw, size_t lds, hipStream_t st) {
    hipLaunchKernelGGL((k_rows_pipe<T, E, R0>), dim3((unsigned)a.grid), dim3(a.C * a.Tt), lds, st,
                       (const V_t<T>*)in, (V_t<T>*)out, a, pt, (const V_t<T>*)tw);
    return hipGetLastError();
}

template <typename T, int E, int R0>
int rows_pipe_occ_r(int block, size_t lds) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb, reinterpret_cast<const void*>(k_rows_pipe<T, E, R0>), block, lds) != hipSuccess)
        return 0;
    return nb;
}

template <typename T, int E, int R0>
void rows_pipe_init_r() {
    allow_big_lds(k_rows_pipe<T, E, R0>);
    allow_big_lds(k_fs_a_pipe<T, E, R0>);
    allow_big_lds(k_fs_b_pipe<T, E, R0>);
}

// Dispatch of a pipelined rows launch on its first radix.
#define CFFT_PIPE_R0_SWITCH(R0_, CALL)                                \
    switch (R0_) {                                                    \
        case 16: return CALL(16);                                     \
        case 13: return CALL(13);                                     \
        case 11: return CALL(11);                                     \
        case 9: return CALL(9);                                       \
        case 8: return CALL(8);                                       \
        case 7: return CALL(7);                                       \
        case 5: return CALL(5);                                       \
        case 4: return CALL(4);                                       \
        case 3: return CALL(3);                                       \
        case 2: return CALL(2);                                       \
        default: break;                                               \
    }

template <typename T, int E>
hipError_t rows_pipe_t(const void* in, void* out, const RowsArgs& a, const PassTable& pt,
                       const void* tw, size_t twn, hipStream_t st) {
    const size_t lds = pipe_tw_offset<T>(a.C, a.Ls) + twn * sizeof(V_t<T>);
#define CFFT_PIPE_CALL(RR) rows_pipe_r<T, E, RR>(in, out, a, pt, tw, lds, st)
    CFFT_PIPE_R0_SWITCH(pt.radix[0], CFFT_PIPE_CALL)
#undef CFFT_PIPE_CALL
    return hipErrorInvalidValue;
}

template <typename T, int E>
int rows_pipe_occ_t(int R0, int block, size_t lds) {
#define CFFT_PIPE_CALL(RR) rows_pipe_occ_r<T, E, RR>(block, lds)
    CFFT_PIPE_R0_SWITCH(R0, CFFT_PIPE_CALL)
#undef CFFT_PIPE_CALL
    return 0;
}

template <typename T, int E, int R0>
hipError_t fsa_pipe_r(const void* in, void* out, const FourStepAArgs& a, const PassTable& pt,
                      const void* tw, size_t lds, hipStream_t st) {
    hipLaunchKernelGGL((k_fs_a_pipe<T, E, R0>), dim3((unsigned)a.grid), dim3(a.C * a.Tt), lds, st,
                       (const V_t<T>*)in, (V_t<T>*)out, a, pt, (const V_t<T>*)tw);
    return hipGetLastError();
}
template <typename T, int E, int R0>
hipError_t fsb_pipe_r(const void* in, void* out, const FourStepBArgs& a, const PassTable& pt,
                      const void* tw, size_t twn, size_t lds, hipStream_t st) {
    hipLaunchKernelGGL((k_fs_b_pipe<T, E, R0>), dim3((unsigned)a.grid), dim3(a.C * a.Tt), lds, st,
                       (const V_t<T>*)in, (V_t<T>*)out, a, pt, (const V_t<T>*)tw, (int)twn);
    return hipGetLastError();
}

template <typename T, int E>
hipError_t fsa_pipe_t(const void* in, void* out, const FourStepAArgs& a, const PassTable& pt,
                      const void* tw, size_t twn, hipStream_t st) {
    const size_t lds = pipe_tw_offset<T>(a.C, a.Ls) + twn * sizeof(V_t<T>);
#define CFFT_PIPE_CALL(RR) fsa_pipe_r<T, E, RR>(in, out, a, pt, tw, lds, st)
    CFFT_PIPE_R0_SWITCH(pt.radix[0], CFFT_PIPE_CALL)
#undef CFFT_PIPE_CALL
    return hipErrorInvalidValue;
}

template <typename T, int E>
hipError_t fsb_pipe_t(const void* in, void* out, const FourStepBArgs& a, const PassTable& pt,
                      const void* tw, size_t twn, hipStream_t st) {
    const size_t lds = pipe_t3_offset<T>(a.C, a.Ls, (int)twn) +
                       (a.tw3 ? (size_t)tw3_entries(a.tbits, a.tnc) * sizeof(V_t<T>) : 0);
#define CFFT_PIPE_CALL(RR) fsb_pipe_r<T, E, RR>(in, out, a, pt, tw, twn, lds, st)
    CFFT_PIPE_R0_SWITCH(pt.radix[0], CFFT_PIPE_CALL)
#undef CFFT_PIPE_CALL
    return hipErrorInvalidValue;
}

template <typename T, int E, int R0>
int fs_pipe_occ_r(int pass, int block, size_t lds) {
    int nb = 0;
    const void* k = pass ? reinterpret_cast<const void*>(k_fs_b_pipe<T, E, R0>)
                         : reinterpret_cast<const void*>(k_fs_a_pipe<T, E, R0>);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, block, lds) != hipSuccess) return 0;
    return nb;
}

template <typename T, int E>
int fs_pipe_occ_t(int pass, int R0, int block, size_t lds) {
#define CFFT_PIPE_CALL(RR) fs_pipe_occ_r<T, E, RR>(pass, block, lds)
    CFFT_PIPE_R0_SWITCH(R0, CFFT_PIPE_CALL)
#undef CFFT_PIPE_CALL
    return 0;
}

template <typename T, int E>
void rows_pipe_init_t() {
    rows_pipe_init_r<T, E, 16>();
    rows_pipe_init_r<T, E, 13>();
    rows_pipe_init_r<T, E, 11>();
    rows_pipe_init_r<T, E, 9>();
    rows_pipe_init_r<T, E, 8>();
    rows_pipe_init_r<T, E, 7>();
    rows_pipe_init_r<T, E, 5>();
    rows_pipe_init_r<T, E, 4>();
    rows_pipe_init_r<T, E, 3>();
    rows_pipe_init_r<T, E, 2>();
}

}  // namespace detail
}  // namespace cfft
