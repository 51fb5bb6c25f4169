// kernels_impl.h -- the HIP kernels of the library and their templated launchers (gfx950).
// Instantiated per precision by kernels_f32.hip and kernels_f64.hip.
//
//   rows        batched contiguous line FFTs, one HBM read + one HBM write per element
//               (radix_exec.jl's whole pass sequence in one launch)
//   fs_a/fs_b   the two passes of the four-step n = N1*N2 (mixed_radix.jl:44-78), twiddle
//               w_n^{j2 k1} fused into pass B's load, transpose fused into pass B's store
//   conv        Rader (prime.jl:5-82) or Bluestein fused in LDS
//   gather / scatter
//               the permutation / chirp ends of a large convolution whose FFTs run on the
//               kernels above
// Each kernel is a device-function body called by a single-plan __global__ wrapper (arguments
// in a struct) and by a work-list wrapper (arguments from an MItem / MPlan table), so the two
// paths execute the same arithmetic for the same transform.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "kernels.h"
#include "multi.h"

namespace cfft {

template <typename T> using V_t = typename Vec2<T>::type;

// ComplexF16 storage: (re, im) as two IEEE binary16 values in 4 bytes.  Lines stored this way
// are converted to float at their global loads and rounded (to nearest even) at their global
// stores; all arithmetic in between runs in float.  CFFT_F16 is a macro so that the CPU
// emulation harness of the tests can supply a software binary16 type.
#ifndef CFFT_F16
#define CFFT_F16 _Float16
#endif
struct alignas(4) ch16 {
    CFFT_F16 x, y;
};

// Global element <-> compute vector, for the storage types V_t<T> (no conversion) and ch16.
template <typename T>
__device__ __forceinline__ V_t<T> load_elem(const V_t<T>& g) { return g; }
template <typename T>
__device__ __forceinline__ V_t<T> load_elem(const ch16& g) {
    return V_t<T>{(T)(float)g.x, (T)(float)g.y};
}
template <typename T>
__device__ __forceinline__ void store_elem(V_t<T>& g, V_t<T> v) { g = v; }
template <typename T>
__device__ __forceinline__ void store_elem(ch16& g, V_t<T> v) {
    g = ch16{(CFFT_F16)(float)v.x, (CFFT_F16)(float)v.y};
}

// Streaming (non-temporal) accesses, for data touched once in whole lines: the line inputs and
// outputs of single-plan rows kernels and the four-step middle written by pass A.  They keep
// such streams from evicting what other workgroups still read from the caches.  Column accesses
// (runs of C elements per row, whose 128-byte lines the neighbouring tiles' workgroups read
// through L2) stay ordinary: streaming loads there made the 2^20 F32 four-step 28 % slower.  So do
// work-list launches, whose convolution and four-step intermediates are read again by the next
// launch while still cached, and ComplexF16 lines.
#ifndef CFFT_PLAIN_STREAMING
template <typename T> struct StreamVec;
template <> struct StreamVec<float> { typedef float type __attribute__((ext_vector_type(2))); };
template <> struct StreamVec<double> { typedef double type __attribute__((ext_vector_type(2))); };
template <typename T>
__device__ __forceinline__ V_t<T> load_stream(const V_t<T>& g) {
    typedef typename StreamVec<T>::type W;
    const W w = __builtin_nontemporal_load(reinterpret_cast<const W*>(&g));
    return V_t<T>{w.x, w.y};
}
template <typename T>
__device__ __forceinline__ void store_stream(V_t<T>& g, V_t<T> v) {
    typedef typename StreamVec<T>::type W;
    const W w = {v.x, v.y};
    __builtin_nontemporal_store(w, reinterpret_cast<W*>(&g));
}
#else
template <typename T>
__device__ __forceinline__ V_t<T> load_stream(const V_t<T>& g) { return g; }
template <typename T>
__device__ __forceinline__ void store_stream(V_t<T>& g, V_t<T> v) { g = v; }
#endif
template <typename T>
__device__ __forceinline__ V_t<T> load_stream(const ch16& g) { return load_elem<T>(g); }
template <typename T>
__device__ __forceinline__ void store_stream(ch16& g, V_t<T> v) { store_elem<T>(g, v); }
// S selects streaming (true) or ordinary (false) accesses.
template <typename T, bool S, typename G>
__device__ __forceinline__ V_t<T> load_line_elem(const G& g) {
    if constexpr (S) return load_stream<T>(g);
    else return load_elem<T>(g);
}
template <typename T, bool S, typename G>
__device__ __forceinline__ void store_line_elem(G& g, V_t<T> v) {
    if constexpr (S) store_stream<T>(g, v);
    else store_elem<T>(g, v);
}

// Global-memory line input of storage type G, read once (S: streaming); sg = -1 conjugates.
template <typename T, typename G = V_t<T>, bool S = true>
struct GSrc {
    const G* p;
    T sg;
    __device__ __forceinline__ V_t<T> operator()(int e) const {
        V_t<T> v = load_line_elem<T, S>(p[e]);
        v.y *= sg;
        return v;
    }
};

// Global-memory line output (S: streaming) with scale, conjugation and the optional convolution
// epilogue.
template <typename T, typename G = V_t<T>, bool S = true>
struct GRowDst {
    G* p;
    T sx, sy;
    const V_t<T>* K;
    V_t<T>* dc;
    __device__ __forceinline__ void put(int e, V_t<T> v) const {
        if (K) {
            if (dc && e == 0) *dc = v;
            v = cconj(cmul(v, K[e]));
        }
        store_line_elem<T, S>(p[e], V_t<T>{v.x * sx, v.y * sy});
    }
};

// Plain global-memory line output (S: streaming).
template <typename T, bool S = true>
struct GPlainDst {
    V_t<T>* p;
    __device__ __forceinline__ void put(int e, V_t<T> v) const { store_line_elem<T, S>(p[e], v); }
};

// Global-memory line with elements `stride` apart (a column of a row-major matrix); sg = -1
// conjugates on input, (sx, sy) scale and conjugate on output.
template <typename T, typename G = V_t<T>>
struct GStrideSrc {
    const G* p;
    int64_t stride;
    T sg;
    __device__ __forceinline__ V_t<T> operator()(int e) const {
        V_t<T> v = load_elem<T>(p[(int64_t)e * stride]);
        v.y *= sg;
        return v;
    }
};
template <typename T, typename G = V_t<T>>
struct GStrideDst {
    G* p;
    int64_t stride;
    T sx, sy;
    __device__ __forceinline__ void put(int e, V_t<T> v) const {
        store_elem<T>(p[(int64_t)e * stride], V_t<T>{v.x * sx, v.y * sy});
    }
};

// Staged copy with batched memory latency: for idx = tid, tid + step, ... < total, every thread
// first runs U loads LD(idx) (indices past the end are clamped to total - 1, so every load is
// valid and none is predicated), then the U stores ST(idx, value) of the valid indices.  One wait
// then covers U loads; a plain loop would wait once per element, because each element's store
// may alias the next element's load.
constexpr int kStageU = 8;
// Staged loads of plain data issue 16 per thread per batch: a workgroup then has up to 16 KiB
// (F32) or 32 KiB (F64) per 256 threads in flight, which the memory system needs to cover its
// latency when only one or two workgroups fit a CU.
constexpr int kStageLoadU = 16;

template <int U, class LD, class ST>
__device__ __forceinline__ void batched_for(int tid, int step, int total, const LD& ld,
                                            const ST& st) {
    using R = decltype(ld(0));
    for (int base = tid; base < total; base += U * step) {
        R v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = base + u * step;
            v[u] = ld(i < total ? i : total - 1);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = base + u * step;
            if (i < total) st(i, v[u]);
        }
    }
}

template <typename T>
__device__ __forceinline__ LdsLine<T> lds_line(unsigned char* smem, int C, int Ls, int c,
                                               int sh = PadShift<T>::v) {
    T* re = reinterpret_cast<T*>(smem);
    return LdsLine<T>{re + c * Ls, re + C * Ls + c * Ls, sh};
}

// ------------------------------------------------------------------------------------------
// Line addressing policies: where line c of a workgroup's lines lives.
// ------------------------------------------------------------------------------------------
template <typename T, typename GI = V_t<T>, typename GO = V_t<T>>
struct UniformLines {  // line c at in0 + c*idist / out0 + c*odist, slots dc0 + 2c
    static constexpr bool kStream = true;
    const GI* in0;
    GO* out0;
    int64_t idist, odist;
    V_t<T>* dc0;
    __device__ __forceinline__ const GI* in(int c) const { return in0 + c * idist; }
    __device__ __forceinline__ GO* out(int c) const { return out0 + c * odist; }
    __device__ __forceinline__ V_t<T>* dc(int c) const { return dc0 ? dc0 + 2 * c : nullptr; }
};

template <typename T>
struct ItemLines {  // an MItem's lines: uniform distances or a table of {in, out, dc}
    static constexpr bool kStream = false;
    uint64_t ib, ob;
    const MItem* it;
    __device__ __forceinline__ const uint64_t* entry(int c) const {
        return reinterpret_cast<const uint64_t*>(it->aux) + 3 * c;
    }
    __device__ __forceinline__ const V_t<T>* in(int c) const {
        if (it->flags & kLineTable)
            return reinterpret_cast<const V_t<T>*>(mresolve(entry(c)[0], it->flags & kSelInMask, ib, ob));
        return reinterpret_cast<const V_t<T>*>(mresolve(it->in, it->flags & kSelInMask, ib, ob)) +
               c * it->idist;
    }
    __device__ __forceinline__ V_t<T>* out(int c) const {
        const int so = (it->flags >> kSelOutShift) & 15;
        if (it->flags & kLineTable) return reinterpret_cast<V_t<T>*>(mresolve(entry(c)[1], so, ib, ob));
        return reinterpret_cast<V_t<T>*>(mresolve(it->out, so, ib, ob)) + c * it->odist;
    }
    __device__ __forceinline__ V_t<T>* dc(int c) const {
        if (it->flags & kLineTable) return reinterpret_cast<V_t<T>*>(entry(c)[2]);
        return it->dc ? reinterpret_cast<V_t<T>*>(it->dc) + 2 * c : nullptr;
    }
};

// ------------------------------------------------------------------------------------------
// rows body: up to C lines of length N (`count` valid).  A line with no passes (N = 1) is a
// scaled copy.
// ------------------------------------------------------------------------------------------
template <typename T, int E, int RM = 16, bool WT = false, class LINES>
__device__ __forceinline__ void rows_body(unsigned char* smem, const LINES& lines, int count,
                                          int N, int C, int Tt, int Ls, const PassTable& pt,
                                          const V_t<T>* tw, int sgn, double scale,
                                          const V_t<T>* K, int sh = PadShift<T>::v) {
    const int tid = threadIdx.x;
    const int c = tid / Tt;
    const int t = tid - c * Tt;
    const bool active = (c < C) && (c < count);
    const int li = active ? c : 0;
    using GI = std::remove_cv_t<std::remove_pointer_t<decltype(lines.in(0))>>;
    using GO = std::remove_pointer_t<decltype(lines.out(0))>;
    constexpr bool S = LINES::kStream;
    GSrc<T, GI, S> gs{lines.in(li), (T)sgn};
    GRowDst<T, GO, S> gd{lines.out(li), (T)scale, (T)(scale * sgn), K, K ? lines.dc(li) : nullptr};
    if (pt.npass == 0) {
        if (active && t == 0) gd.put(0, gs(0));
        return;
    }
    const LdsLine<T> L = lds_line<T>(smem, C, Ls, c < C ? c : 0, sh);
    run_line<T, E, RM, WT>(pt, N, t, Tt, active, tw, gs, gd, L, true, true);
}

// RM bounds the radices the kernel holds (lines of radices <= 9 run on a lighter instance); WT
// instances hold the wave-tail path (lines whose last pass runs across wavefront lanes).
template <typename T, int E, typename G = V_t<T>, int RM = 16, bool WT = false>
__global__ void __launch_bounds__(1024)
k_rows(const G* __restrict__ in, G* __restrict__ out, RowsArgs a, PassTable pt,
       const V_t<T>* __restrict__ tw) {
    CFFT_DYN_LDS(smem);
    const int64_t line0 = (int64_t)blockIdx.x * a.C;
    const int64_t left = a.nlines - line0;
    const int count = (int)(left < a.C ? left : a.C);
    UniformLines<T, G, G> lines{in + line0 * a.idist, out + line0 * a.odist, a.idist, a.odist,
                                a.epi.dc ? (V_t<T>*)a.epi.dc + 2 * line0 : nullptr};
    rows_body<T, E, RM, WT>(smem, lines, count, a.N, a.C, a.Tt, a.Ls, pt, tw, a.sgn_in, a.scale,
                            (const V_t<T>*)a.epi.K, a.nopad ? kNoPadShift : PadShift<T>::v);
}

// ------------------------------------------------------------------------------------------
// Four-step pass A body: columns j20 .. j20+C-1 of one transform (N1 x N2 row-major).
//   out[j2*N1 + k1] = DFT_N1(in[j1*N2 + j2])[k1]
// The twiddle w_n^{j2 k1} is applied by pass B while it loads the middle.
// ------------------------------------------------------------------------------------------
// Threads are mapped column first (lane = c + C t): a wave's accesses to a row of the matrix
// cover C adjacent columns, so the first pass reads global memory directly in runs of C
// elements, with no staging through LDS.
template <typename T, int E, typename GI, bool S = true>
__device__ __forceinline__ void fs_a_body(unsigned char* smem, const GI* src, V_t<T>* dst,
                                          int N1, int N2, int C, int Tt, int Ls, int j20,
                                          int sgn, const PassTable& pt, const V_t<T>* tw) {
    const int cshift = __ffs(C) - 1;
    const int tid = threadIdx.x;
    const int c = tid & (C - 1);
    const int t = tid >> cshift;
    const int j2 = j20 + c;
    const bool active = j2 < N2;
    const int j2c = active ? j2 : 0;
    const LdsLine<T> L = lds_line<T>(smem, C, Ls, c);
    GStrideSrc<T, GI> gs{src + j2c, N2, (T)sgn};
    GPlainDst<T, S> d{dst + (int64_t)j2c * N1};
    run_line<T, E>(pt, N1, t, Tt, active, tw, gs, d, L, true, true);
}

// Workgroups are dispatched to the eight XCDs round robin (block b on XCD b mod 8).  With
// remap, XCD x takes the consecutive block range [x G/8, (x+1) G/8) of a grid of G blocks.
__device__ __forceinline__ int xcd_block(int remap) {
    const int b = blockIdx.x, G = gridDim.x;
    if (!remap || (G & 7)) return b;
    return (b & 7) * (G >> 3) + (b >> 3);
}

// Work-list launches order their items by decreasing cost, so one contiguous range per XCD would
// give one XCD the most expensive items.  Instead each window of 64 blocks gives XCD x the eight
// consecutive items [w + 8x, w + 8x + 8) (adjacent tiles of a transform share an L2); blocks past
// the last full window keep their index.
__device__ __forceinline__ int xcd_item() {
    const int b = blockIdx.x;
    if (b >= ((int)gridDim.x & ~63)) return b;
    return (b & ~63) | ((b & 7) << 3) | ((b >> 3) & 7);
}

// S: the middle is written with streaming stores (false: ordinary ones).
template <typename T, int E, typename GI = V_t<T>, bool S = true>
__global__ void __launch_bounds__(1024)
k_fs_a(const GI* __restrict__ in, V_t<T>* __restrict__ out, FourStepAArgs a, PassTable pt,
       const V_t<T>* __restrict__ tw) {
    CFFT_DYN_LDS(smem);
    const int tiles = (a.N2 + a.C - 1) / a.C;
    const int bid = xcd_block(a.xcd);
    const int64_t bt = bid / tiles;
    const int tile = bid - (int)(bt * tiles);
    fs_a_body<T, E, GI, S>(smem, in + bt * a.idist, out + bt * a.odist, a.N1, a.N2, a.C, a.Tt, a.Ls,
                    tile * a.C, a.sgn_in, pt, tw);
}

// ------------------------------------------------------------------------------------------
// Four-step pass B body: lines k10 .. k10+C-1.  The first pass reads the middle from global
// memory and applies the twiddle in registers; the last pass writes the output directly, or
// through a staged store when a convolution epilogue (K) is fused.
//   out[k1 + N1*k2] = DFT_N2(w_n^{j2 k1} in[j2*N1 + k1])[k2];  in == out allowed (same index
// set).  w_n^e = (A[e & m] * B[(e >> tbits) & m]) * C[e >> 2 tbits] from the three-level table
// tw3 (nullptr: no twiddle), which is copied into LDS after the lines.
// ------------------------------------------------------------------------------------------
template <typename T>
struct TwLoad {  // a staged element and a value that goes with it
    V_t<T> v, a;
};

// The four-step twiddle w_n^x from a three-level table in LDS.
template <typename T>
__device__ __forceinline__ V_t<T> tw3_get(const V_t<T>* t3, int tbits, int x) {
    const int m = (1 << tbits) - 1;
    return cmul(cmul(t3[x & m], t3[(1 << tbits) + ((x >> tbits) & m)]),
                t3[(2 << tbits) + (x >> (2 * tbits))]);
}

// Copies n entries of a global table into LDS.
template <typename T>
__device__ __forceinline__ void stage_table(const V_t<T>* g, V_t<T>* l, int n) {
    batched_for<kStageU>(threadIdx.x, blockDim.x, n, [&](int i) { return g[i]; },
                         [&](int i, V_t<T> v) { l[i] = v; });
}

// Pass B's line input straight from the middle: element e of column k1 at p[e * stride] (p =
// middle + k1, stride = N1), then (post, once the pass's loads are issued) times the four-step
// twiddle w_n^{e k1} from the three-level table in LDS (t3 == nullptr: no twiddle); sg = -1
// conjugates.
template <typename T>
struct GTwColSrc {
    const V_t<T>* p;
    int64_t stride;
    const V_t<T>* t3;
    int tbits, k1;
    T sg;
    __device__ __forceinline__ V_t<T> operator()(int e) const { return p[(int64_t)e * stride]; }
    __device__ __forceinline__ V_t<T> post(int e, V_t<T> v) const {
        if (t3) v = cmul(v, tw3_get<T>(t3, tbits, e * k1));
        v.y *= sg;
        return v;
    }
};

template <typename T, int E, typename GO>
__device__ __forceinline__ void fs_b_body(unsigned char* smem, const V_t<T>* src, GO* dst,
                                          int N1, int N2, int C, int Tt, int Ls, int k10,
                                          int sgn_in, int sgn, double scale, const V_t<T>* K,
                                          V_t<T>* dc, const V_t<T>* tw3, int tbits, int tnc,
                                          const PassTable& pt, const V_t<T>* tw) {
    const int cshift = __ffs(C) - 1;
    const int total = C * N2;
    V_t<T>* t3 = reinterpret_cast<V_t<T>*>(smem + fsb_tw_offset(sizeof(T), C, Ls));
    if (tw3) stage_table<T>(tw3, t3, tw3_entries(tbits, tnc));
    __syncthreads();
    // threads column first (lane = c + C t), as in pass A: the first pass reads the middle
    // straight from global memory in runs of C elements (k1) per row j2, applying the four-step
    // twiddle in registers.  Every element of the tile is read by the first pass, before any
    // write (the last pass's writes come after the passes' barriers; a single-pass line gets a
    // barrier between its reads and writes), so in == out is safe.
    const int tid = threadIdx.x;
    const int c = tid & (C - 1);
    const int t = tid >> cshift;
    const bool active = k10 + c < N1;
    const int k1 = active ? k10 + c : 0;
    const LdsLine<T> L = lds_line<T>(smem, C, Ls, c);
    GTwColSrc<T> gs{src + k1, N1, tw3 ? t3 : nullptr, tbits, k1, (T)sgn_in};
    const T sx = (T)scale, sy = (T)(scale * sgn);
    if (!K) {
        // the last pass writes out[k1 + N1 k2] directly, in runs of C (k1) per row k2
        GStrideDst<T, GO> d{dst + k1, N1, sx, sy};
        run_line<T, E>(pt, N2, t, Tt, active, tw, gs, d, L, true, true, pt.npass == 1);
        return;
    }
    run_line<T, E>(pt, N2, t, Tt, active, tw, gs, L, L, true, false);
    // run_line ended with a barrier: the lines' outputs are in LDS.
    // Staged store in runs of C (k1) per output row k2; the epilogue spectrum is loaded in the
    // same batch.  Indices of lines past N1 are skipped.
    batched_for<kStageU>(
        threadIdx.x, blockDim.x, total,
        [&](int idx) {
            const int cc = idx & (C - 1);
            const int e = idx >> cshift;
            const int k1 = k10 + cc < N1 ? k10 + cc : N1 - 1;
            TwLoad<T> r;
            r.v = lds_line<T>(smem, C, Ls, cc)(e);
            if (K) r.a = K[(int64_t)k1 + (int64_t)N1 * e];
            return r;
        },
        [&](int idx, TwLoad<T> r) {
            const int cc = idx & (C - 1);
            const int e = idx >> cshift;
            const int k1 = k10 + cc;
            if (k1 >= N1) return;
            V_t<T> v = r.v;
            const int64_t o = (int64_t)k1 + (int64_t)N1 * e;
            if (K) {
                if (dc && o == 0) *dc = v;
                v = cconj(cmul(v, r.a));
            }
            store_elem<T>(dst[o], V_t<T>{v.x * sx, v.y * sy});
        });
}

template <typename T, int E, typename GO = V_t<T>>
__global__ void __launch_bounds__(1024)
k_fs_b(const V_t<T>* in, GO* out, FourStepBArgs a, PassTable pt,
       const V_t<T>* __restrict__ tw) {
    CFFT_DYN_LDS(smem);
    const int tiles = (a.N1 + a.C - 1) / a.C;
    const int bid = xcd_block(a.xcd);
    const int64_t bt = bid / tiles;
    const int tile = bid - (int)(bt * tiles);
    fs_b_body<T, E>(smem, in + bt * a.idist, out + bt * a.odist, a.N1, a.N2, a.C, a.Tt, a.Ls,
                    tile * a.C, a.sgn_in, a.sgn_out, a.scale, (const V_t<T>*)a.epi.K,
                    a.epi.dc ? (V_t<T>*)a.epi.dc + 2 * bt : nullptr, (const V_t<T>*)a.tw3,
                    a.tbits, a.tnc, pt, tw);
}

// ------------------------------------------------------------------------------------------
// Tiled transpose with an optional four-step twiddle, convolution epilogue and scale (the
// six-step path).  A 32 x 32
// tile goes through LDS (one pad word per row), so both the reads (along a source row) and the
// writes (along a destination row) are coalesced.  Grid: x over column tiles, y over row tiles
// (y wraps when there are more than 65535), z over the matrices of a batch.
// ------------------------------------------------------------------------------------------
constexpr int kTrTile = 32;

template <typename T>
__global__ void __launch_bounds__(256)
k_transpose(const V_t<T>* __restrict__ src, V_t<T>* __restrict__ dst, int64_t rows, int64_t cols,
            const V_t<T>* __restrict__ tw3, int tbits, T sx, T sy, const V_t<T>* __restrict__ K,
            V_t<T>* dc) {
    CFFT_DYN_LDS(smem);
    src += (int64_t)blockIdx.z * rows * cols;  // matrix z of a batch (no twiddle or epilogue)
    dst += (int64_t)blockIdx.z * rows * cols;
    V_t<T>(*tile)[kTrTile + 1] = reinterpret_cast<V_t<T>(*)[kTrTile + 1]>(smem);
    const int tx = threadIdx.x & (kTrTile - 1), ty = threadIdx.x >> 5;  // 32 x 8 threads
    const int64_t c0 = (int64_t)blockIdx.x * kTrTile;
    const int64_t row_tiles = (rows + kTrTile - 1) / kTrTile;
    for (int64_t rt = blockIdx.y; rt < row_tiles; rt += gridDim.y) {
        const int64_t r0 = rt * kTrTile;
#pragma unroll
        for (int k = 0; k < kTrTile; k += 8) {
            const int64_t r = r0 + ty + k, c = c0 + tx;
            if (r < rows && c < cols) {
                V_t<T> v = src[r * cols + c];
                if (tw3) v = cmul(v, tw3_get<T>(tw3, tbits, (int)(r * c)));
                tile[ty + k][tx] = v;
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kTrTile; k += 8) {
            const int64_t c = c0 + ty + k, r = r0 + tx;
            if (r < rows && c < cols) {
                V_t<T> v = tile[tx][ty + k];
                const int64_t o = c * rows + r;
                if (K) {  // a convolution's epilogue, as in GRowDst
                    if (dc && o == 0) *dc = v;
                    v = cconj(cmul(v, K[o]));
                }
                dst[o] = V_t<T>{v.x * sx, v.y * sy};
            }
        }
        __syncthreads();
    }
}

template <typename T>
hipError_t transpose_t(const void* src, void* dst, int64_t rows, int64_t cols, int64_t batch,
                       const void* tw3, int tbits, double sx, double sy, const void* K, void* dc,
                       hipStream_t st) {
    const int64_t ct = (cols + kTrTile - 1) / kTrTile, rt = (rows + kTrTile - 1) / kTrTile;
    if (ct <= 0 || rt <= 0 || batch <= 0) return hipSuccess;
    if (ct > 0x7fffffff || (batch > 1 && (tw3 || K))) return hipErrorInvalidValue;
    const size_t lds = sizeof(V_t<T>) * kTrTile * (kTrTile + 1);
    const int64_t mat = rows * cols;
    for (int64_t z0 = 0; z0 < batch; z0 += 65535) {
        const int64_t nz = batch - z0 < 65535 ? batch - z0 : 65535;
        const dim3 grid((unsigned)ct, (unsigned)(rt < 65535 ? rt : 65535), (unsigned)nz);
        hipLaunchKernelGGL(k_transpose<T>, grid, dim3(256), lds, st,
                           (const V_t<T>*)src + z0 * mat, (V_t<T>*)dst + z0 * mat, rows, cols,
                           (const V_t<T>*)tw3, tbits, (T)sx, (T)sy, (const V_t<T>*)K,
                           (V_t<T>*)dc);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// ------------------------------------------------------------------------------------------
// conv body: Rader / Bluestein, whole transform in LDS.
// LDS: C lines of stride Ls, shared by the input X (then the output) and the convolution work A
// (element e of both at the same word) | dc[C].  x0 is read before the forward FFT_L overwrites
// X; the first pass of that FFT (gather from X, writes to A) and the last pass of the second
// (reads A, scatters into X) separate their reads from their writes with a barrier.
// ------------------------------------------------------------------------------------------
template <typename T>
struct GatherSrc {   // a[e] from the staged input X
    LdsLine<T> X;
    const int* perm;
    const V_t<T>* chirp;
    int n, mode;
    __device__ __forceinline__ V_t<T> operator()(int e) const {
        if (mode == (int)ConvMode::Rader) {
            if (e < n - 1) return X(perm[e]);
            return V_t<T>{0, 0};
        }
        if (e < n) return cmul(X(e), chirp[e]);
        return V_t<T>{0, 0};
    }
};

template <typename T>
struct ConvEpiDst {  // A[e] = conj(v K[e]), dc = v at e == 0
    LdsLine<T> A;
    const V_t<T>* K;
    V_t<T>* dc;
    __device__ __forceinline__ void put(int e, V_t<T> v) const {
        if (e == 0) *dc = v;
        A.put(e, cconj(cmul(v, K[e])));
    }
};

template <typename T>
struct ScatterDst {  // inverse result c = conj(v)/L into the output X
    LdsLine<T> X;
    const int* iperm;
    const V_t<T>* chirp;
    V_t<T> x0;
    T invL;
    int n, mode;
    __device__ __forceinline__ void put(int k, V_t<T> v) const {
        const V_t<T> cv = V_t<T>{v.x * invL, -v.y * invL};
        if (mode == (int)ConvMode::Rader) {
            if (k < n - 1) X.put(iperm[k], V_t<T>{x0.x + cv.x, x0.y + cv.y});
        } else {
            if (k < n) X.put(k, cmul(cv, chirp[k]));
        }
    }
};

template <typename T, int E, bool WT = false, class LINES>
__device__ __forceinline__ void conv_body(unsigned char* smem, const LINES& lines, int nl, int n,
                                          int L, int C, int Tt, int Ls, int mode, int sgn,
                                          double scale, const int* perm, const int* iperm,
                                          const V_t<T>* chirp, const V_t<T>* K,
                                          const PassTable& pt, const V_t<T>* tw) {
    T* xre = reinterpret_cast<T*>(smem);
    T* xim = xre + C * Ls;
    V_t<T>* sdc = reinterpret_cast<V_t<T>*>(xim + C * Ls);
    const T sg = (T)sgn;
    // 1. staged, coalesced load of the nl input lines (flat over lines x elements)
    batched_for<kStageLoadU>(
        threadIdx.x, blockDim.x, nl * n,
        [&](int idx) {
            const int c = idx / n;
            return lines.in(c)[idx - c * n];
        },
        [&](int idx, V_t<T> v) {
            const int c = idx / n;
            v.y *= sg;
            LdsLine<T>{xre + c * Ls, xim + c * Ls}.put(idx - c * n, v);
        });
    __syncthreads();
    const int tid = threadIdx.x;
    const int c = tid / Tt;
    const int t = tid - c * Tt;
    const bool active = (c < C) && (c < nl);
    const int cc = c < C ? c : 0;
    const LdsLine<T> X{xre + cc * Ls, xim + cc * Ls};
    const LdsLine<T> A = X;
    const V_t<T> x0 = X(0);
    // 2. forward FFT_L of the gathered / chirped sequence; pointwise product with K on the way out
    GatherSrc<T> gsrc{X, perm, chirp, n, mode};
    ConvEpiDst<T> edst{A, K, sdc + cc};
    run_line_t<T, E, 16, WT>(pt, L, t, Tt, active, TwGlobal<V_t<T>>{tw}, gsrc, edst, A, true, true, true,
                     0, 1, true);
    __syncthreads();
    // 3. second forward FFT_L of the conjugated product = conj of the inverse; scatter into X
    ScatterDst<T> sdst{X, iperm, chirp, x0, (T)(1.0 / L), n, mode};
    run_line<T, E, 16, WT>(pt, L, t, Tt, active, tw, A, sdst, A, false, true, true);
    __syncthreads();
    if (mode == (int)ConvMode::Rader && active && t == 0) {
        const V_t<T> d = sdc[cc];
        X.put(0, V_t<T>{x0.x + d.x, x0.y + d.y});
    }
    __syncthreads();
    // 4. staged, coalesced store
    const T sx = (T)scale, sy = (T)(scale * sgn);
    batched_for<kStageU>(
        threadIdx.x, blockDim.x, nl * n,
        [&](int idx) {
            const int c2 = idx / n;
            return LdsLine<T>{xre + c2 * Ls, xim + c2 * Ls}(idx - c2 * n);
        },
        [&](int idx, V_t<T> v) {
            const int c2 = idx / n;
            lines.out(c2)[idx - c2 * n] = V_t<T>{v.x * sx, v.y * sy};
        });
}

template <typename T, int E, bool WT = false>
__global__ void __launch_bounds__(1024)
k_conv_small(const V_t<T>* __restrict__ in, V_t<T>* __restrict__ out, ConvSmallArgs a,
             PassTable pt, const V_t<T>* __restrict__ tw) {
    CFFT_DYN_LDS(smem);
    const int64_t line0 = (int64_t)blockIdx.x * a.C;
    const int nl = (int)((a.nlines - line0) < a.C ? (a.nlines - line0) : a.C);
    UniformLines<T> lines{in + line0 * a.idist, out + line0 * a.odist, a.idist, a.odist, nullptr};
    conv_body<T, E, WT>(smem, lines, nl, a.n, a.L, a.C, a.Tt, a.Ls, a.mode, a.sgn_in,
                    a.scale, a.perm, a.iperm, (const V_t<T>*)a.chirp, (const V_t<T>*)a.K, pt, tw);
}

// ------------------------------------------------------------------------------------------
// Large convolution gather / scatter bodies over elements [e0, e1) of one line.
// ------------------------------------------------------------------------------------------
constexpr int kGatherU = 4;  // elements per thread whose loads share one wait

template <typename Ti, typename Tc>
__device__ __forceinline__ void gather_body(const V_t<Ti>* xs, V_t<Tc>* ws, V_t<Tc>* slots,
                                            int64_t e0, int64_t e1, int64_t estep, int64_t n,
                                            int mode, int sgn, const int* perm,
                                            const V_t<Tc>* chirp) {
    const Tc sg = (Tc)sgn;
    const bool rader = mode == (int)ConvMode::Rader;
    const int64_t live_end = rader ? n - 1 : n;  // a[e] = 0 from here on (zero padding)
    const int64_t last = live_end > 0 ? live_end - 1 : 0;
    if (rader && e0 == 0) {  // keep x0 for the scatter (the input may be overwritten in place)
        const V_t<Ti> u0 = xs[0];
        slots[1] = V_t<Tc>{(Tc)u0.x, (Tc)u0.y * sg};
    }
    // Batches of kGatherU elements: permutation indices, then the permuted loads (and chirp),
    // then the stores; every load uses a clamped, valid index.
    for (int64_t e = e0; e < e1; e += kGatherU * estep) {
        int64_t src[kGatherU];
#pragma unroll
        for (int u = 0; u < kGatherU; ++u) {
            const int64_t ee = e + u * estep;
            const int64_t ec = ee < last ? ee : last;
            src[u] = rader ? (int64_t)perm[ec] : ec;
        }
        V_t<Ti> x[kGatherU];
        V_t<Tc> ch[kGatherU];
#pragma unroll
        for (int u = 0; u < kGatherU; ++u) {
            x[u] = xs[src[u]];
            if (!rader) {
                const int64_t ee = e + u * estep;
                ch[u] = chirp[ee < last ? ee : last];
            }
        }
#pragma unroll
        for (int u = 0; u < kGatherU; ++u) {
            const int64_t ee = e + u * estep;
            if (ee >= e1) break;
            V_t<Tc> v = V_t<Tc>{0, 0};
            if (ee < live_end) {
                v = V_t<Tc>{(Tc)x[u].x, (Tc)x[u].y * sg};
                if (!rader) v = cmul(v, ch[u]);
            }
            ws[ee] = v;
        }
    }
}

// slots = {A^[0], x0} of the line (Rader).
template <typename Ti, typename Tc>
__device__ __forceinline__ void scatter_body(const V_t<Tc>* ws, V_t<Ti>* ys,
                                             const V_t<Tc>* slots, int64_t k0, int64_t k1,
                                             int64_t kstep, int64_t n, int mode, int sgn,
                                             double invL_, double scale, const int* iperm,
                                             const V_t<Tc>* chirp) {
    const Tc invL = (Tc)invL_;
    const Tc sx = (Tc)scale, sy = (Tc)(scale * sgn);
    const bool rader = mode == (int)ConvMode::Rader;
    const int64_t kend0 = rader ? n - 1 : n;
    const int64_t kend = k1 < kend0 ? k1 : kend0;  // this thread's range is [k0, kend)
    if (kend <= 0) return;
    const int64_t last = kend0 - 1;
    V_t<Tc> x0 = V_t<Tc>{0, 0};
    if (rader) {
        x0 = slots[1];
        if (k0 == 0) {
            const V_t<Tc> d = slots[0];
            ys[0] = V_t<Ti>{(Ti)((x0.x + d.x) * sx), (Ti)((x0.y + d.y) * sy)};
        }
    }
    for (int64_t k = k0; k < kend; k += kGatherU * kstep) {
        V_t<Tc> v[kGatherU], ch[kGatherU];
        int64_t dst[kGatherU];
#pragma unroll
        for (int u = 0; u < kGatherU; ++u) {
            const int64_t kk = k + u * kstep;
            const int64_t kc = kk < last ? kk : last;
            v[u] = ws[kc];
            if (rader)
                dst[u] = iperm[kc];
            else
                ch[u] = chirp[kc];
        }
#pragma unroll
        for (int u = 0; u < kGatherU; ++u) {
            const int64_t kk = k + u * kstep;
            if (kk >= kend) break;
            if (rader) {
                const V_t<Tc> r = V_t<Tc>{x0.x + v[u].x * invL, x0.y - v[u].y * invL};
                ys[dst[u]] = V_t<Ti>{(Ti)(r.x * sx), (Ti)(r.y * sy)};
            } else {
                const V_t<Tc> r = cmul(V_t<Tc>{v[u].x * invL, -v[u].y * invL}, ch[u]);
                ys[kk] = V_t<Ti>{(Ti)(r.x * sx), (Ti)(r.y * sy)};
            }
        }
    }
}

template <typename Ti, typename Tc>
__global__ void __launch_bounds__(256)
k_conv_gather(const V_t<Ti>* __restrict__ x, V_t<Tc>* __restrict__ w, ConvGatherArgs a) {
    for (int64_t line = blockIdx.y; line < a.nlines; line += gridDim.y)
        gather_body<Ti, Tc>(x + line * a.idist, w + line * a.wdist, (V_t<Tc>*)a.dc + 2 * line,
                            (int64_t)blockIdx.x * blockDim.x + threadIdx.x, a.L,
                            (int64_t)gridDim.x * blockDim.x, a.n, a.mode, a.sgn_in, a.perm,
                            (const V_t<Tc>*)a.chirp);
}

template <typename Ti, typename Tc>
__global__ void __launch_bounds__(256)
k_conv_scatter(const V_t<Tc>* __restrict__ w, V_t<Ti>* __restrict__ y, ConvScatterArgs a) {
    for (int64_t line = blockIdx.y; line < a.nlines; line += gridDim.y)
        scatter_body<Ti, Tc>(w + line * a.wdist, y + line * a.odist, (const V_t<Tc>*)a.dc + 2 * line,
                             (int64_t)blockIdx.x * blockDim.x + threadIdx.x, a.n,
                             (int64_t)gridDim.x * blockDim.x, a.n, a.mode, a.sgn_in, a.invL,
                             a.scale, a.iperm, (const V_t<Tc>*)a.chirp);
}

// ------------------------------------------------------------------------------------------
// Large convolution over a four-step, three launches (ConvLArgs in kernels.h).  Ti: precision
// of x and y; Tc: precision of the convolution.
// ------------------------------------------------------------------------------------------
// Column j2 of the gathered sequence as an N1 x N2 matrix: element e is a[e N2 + j2], with
// Rader a[i] = x[perm[i]] (i < n - 1), Bluestein a[i] = x[i] chirp[i] (i < n), zero beyond (the
// padding); sg = -1 conjugates x.
template <typename Ti, typename Tc>
struct GConvGatherCol {
    const V_t<Ti>* x;
    const int* perm;
    const V_t<Tc>* chirp;
    int64_t live;  // n - 1 (Rader) or n (Bluestein)
    int N2, j2;
    bool rader;
    Tc sg;
    __device__ __forceinline__ V_t<Tc> operator()(int e) const {
        const int64_t i = (int64_t)e * N2 + j2;
        if (i >= live) return V_t<Tc>{0, 0};
        const V_t<Ti> u = x[rader ? (int64_t)perm[i] : i];
        const V_t<Tc> v = V_t<Tc>{(Tc)u.x, (Tc)u.y * sg};
        return rader ? v : cmul(v, chirp[i]);
    }
};

// The bodies below run one tile of one transform: xl / yl its input / output row, m1 / m2 its
// two work lines, slot its {A^[0], x0}.  The single-plan kernels and the work-list kernels
// (one item per workgroup) call the same bodies.

// A: columns j20 .. j20 + C - 1 (j20 = tile C); FFT_N1 of the gathered column -> m1[j2 N1 + k1].
// The first tile keeps x0 for the scatter (Rader).
template <typename Ti, typename Tc, int E>
__device__ __forceinline__ void convl_a_body(unsigned char* smem, const ConvLArgs& a, int tile,
                                             const V_t<Ti>* xl, V_t<Tc>* m1, V_t<Tc>* slot,
                                             const PassTable& pt, const V_t<Tc>* tw) {
    const int C = a.CA;
    const int cshift = __ffs(C) - 1;
    const int tid = threadIdx.x;
    const int c = tid & (C - 1);
    const int t = tid >> cshift;
    const int j2 = tile * C + c;
    const bool active = j2 < a.N2;
    const int j2c = active ? j2 : 0;
    const bool rader = a.mode == (int)ConvMode::Rader;
    const Tc sg = (Tc)a.sgn;
    if (rader && tile == 0 && tid == 0) {
        const V_t<Ti> u0 = xl[0];
        slot[1] = V_t<Tc>{(Tc)u0.x, (Tc)u0.y * sg};
    }
    GConvGatherCol<Ti, Tc> gs{xl, a.perm, (const V_t<Tc>*)a.chirp, rader ? a.n - 1 : a.n, a.N2,
                              j2c, rader, sg};
    GPlainDst<Tc, false> d{m1 + (int64_t)j2c * a.N1};
    run_line<Tc, E>(pt, a.N1, t, a.TtA, active, tw, gs, d, lds_line<Tc>(smem, C, a.LsA, c), true,
                    true);
}

// BA: lines k10 .. k10 + C - 1 (k10 = tile C, k1 < N1): FFT_N2 of w_L^{j2 k1} m1[j2 N1 + k1] (the
// first FFT's pass B) in LDS; conj(X1 K[k1 + N1 k2]) (A^[0] = X1[0] kept in slot[0]); FFT_N2 of
// that (the second FFT's pass A, its column k1) -> m2[k1 N2 + k1'].
template <typename Tc, int E>
__device__ __forceinline__ void convl_ba_body(unsigned char* smem, const ConvLArgs& a, int tile,
                                              const V_t<Tc>* m1, V_t<Tc>* m2, V_t<Tc>* slot,
                                              const PassTable& pt, const V_t<Tc>* tw) {
    const int C = a.CB;
    const int k10 = tile * C;
    const int cshift = __ffs(C) - 1;
    V_t<Tc>* t3 = reinterpret_cast<V_t<Tc>*>(smem + fsb_tw_offset(sizeof(Tc), C, a.LsB));
    stage_table<Tc>((const V_t<Tc>*)a.tw3, t3, tw3_entries(a.tbits, a.tnc));
    __syncthreads();
    const int tid = threadIdx.x;
    const int c = tid & (C - 1);
    const int t = tid >> cshift;
    const bool active = k10 + c < a.N1;
    const int k1 = active ? k10 + c : 0;
    const LdsLine<Tc> Ld = lds_line<Tc>(smem, C, a.LsB, c);
    GTwColSrc<Tc> gs{m1 + k1, a.N1, t3, a.tbits, k1, (Tc)1};
    run_line<Tc, E>(pt, a.N2, t, a.TtB, active, tw, gs, Ld, Ld, true, false);
    // run_line ended with a barrier; the epilogue reads and rewrites each element in place, the
    // K loads of a batch sharing one wait
    const V_t<Tc>* K = (const V_t<Tc>*)a.K;
    batched_for<kStageU>(
        tid, blockDim.x, C * a.N2,
        [&](int idx) {
            const int cc = idx & (C - 1);
            const int e = idx >> cshift;
            const int kk = k10 + cc < a.N1 ? k10 + cc : a.N1 - 1;
            return K[(int64_t)kk + (int64_t)a.N1 * e];
        },
        [&](int idx, V_t<Tc> kv) {
            const int cc = idx & (C - 1);
            const int e = idx >> cshift;
            if (k10 + cc >= a.N1) return;
            const LdsLine<Tc> Lc = lds_line<Tc>(smem, C, a.LsB, cc);
            const V_t<Tc> v = Lc(e);
            if (k10 + cc == 0 && e == 0) slot[0] = v;
            Lc.put(e, cconj(cmul(v, kv)));
        });
    __syncthreads();
    GPlainDst<Tc, false> d{m2 + (int64_t)k1 * a.N2};
    run_line<Tc, E>(pt, a.N2, t, a.TtB, active, tw, Ld, d, Ld, false, true);
}

// C: lines k10 .. k10 + C - 1 (k1' < N2): FFT_N1 of w_L^{j2 k1'} m2[j2 N2 + k1'] (the second
// FFT's pass B) in LDS, then X2[i], i = k1' + N2 k2, scattered into y: Rader y[iperm[i]] =
// x0 + conj(X2[i]) / L (and y[0] = x0 + A^[0]), Bluestein y[i] = chirp[i] conj(X2[i]) / L;
// real part times scale, imaginary part times scale * sgn.
template <typename Tc>
struct ConvLScatterLoad {
    V_t<Tc> v, w;
    int64_t dst;
};

template <typename Ti, typename Tc, int E>
__device__ __forceinline__ void convl_c_body(unsigned char* smem, const ConvLArgs& a, int tile,
                                             const V_t<Tc>* m2, V_t<Ti>* yl,
                                             const V_t<Tc>* slots, const PassTable& pt,
                                             const V_t<Tc>* tw) {
    const int C = a.CC;
    const int k10 = tile * C;
    const int cshift = __ffs(C) - 1;
    const int64_t L = (int64_t)a.N1 * a.N2;
    V_t<Tc>* t3 = reinterpret_cast<V_t<Tc>*>(smem + fsb_tw_offset(sizeof(Tc), C, a.LsC));
    stage_table<Tc>((const V_t<Tc>*)a.tw3, t3, tw3_entries(a.tbits, a.tnc));
    __syncthreads();
    const int tid = threadIdx.x;
    const int c = tid & (C - 1);
    const int t = tid >> cshift;
    const bool active = k10 + c < a.N2;
    const int k1 = active ? k10 + c : 0;
    const LdsLine<Tc> Ld = lds_line<Tc>(smem, C, a.LsC, c);
    GTwColSrc<Tc> gs{m2 + k1, a.N2, t3, a.tbits, k1, (Tc)1};
    run_line<Tc, E>(pt, a.N1, t, a.TtC, active, tw, gs, Ld, Ld, true, false);
    const bool rader = a.mode == (int)ConvMode::Rader;
    const Tc invL = (Tc)(1.0 / (double)L);
    const Tc sx = (Tc)a.scale, sy = (Tc)(a.scale * a.sgn);
    const int64_t live = rader ? a.n - 1 : a.n;
    const V_t<Tc> x0 = rader ? slots[1] : V_t<Tc>{0, 0};
    if (rader && tile == 0 && tid == 0) {
        const V_t<Tc> d0 = slots[0];
        yl[0] = V_t<Ti>{(Ti)((x0.x + d0.x) * sx), (Ti)((x0.y + d0.y) * sy)};
    }
    const int* iperm = a.iperm;
    const V_t<Tc>* chirp = (const V_t<Tc>*)a.chirp;
    batched_for<kStageU>(
        tid, blockDim.x, C * a.N1,
        [&](int idx) {
            const int cc = idx & (C - 1);
            const int e = idx >> cshift;
            const int64_t i0 = (int64_t)(k10 + cc < a.N2 ? k10 + cc : 0) + (int64_t)a.N2 * e;
            const int64_t i = i0 < live ? i0 : live - 1;
            ConvLScatterLoad<Tc> r;
            r.v = lds_line<Tc>(smem, C, a.LsC, cc)(e);
            if (rader) {
                r.dst = iperm[i];
                r.w = V_t<Tc>{0, 0};
            } else {
                r.dst = i;
                r.w = chirp[i];
            }
            return r;
        },
        [&](int idx, ConvLScatterLoad<Tc> r) {
            const int cc = idx & (C - 1);
            const int e = idx >> cshift;
            if (k10 + cc >= a.N2) return;
            const int64_t i = (int64_t)(k10 + cc) + (int64_t)a.N2 * e;
            if (i >= live) return;
            const V_t<Tc> cv = V_t<Tc>{r.v.x * invL, -r.v.y * invL};
            const V_t<Tc> o = rader ? V_t<Tc>{x0.x + cv.x, x0.y + cv.y} : cmul(cv, r.w);
            yl[r.dst] = V_t<Ti>{(Ti)(o.x * sx), (Ti)(o.y * sy)};
        });
}

// Single-plan kernels: grid = tiles x lines, consecutive tiles of a line on one XCD.
__device__ __forceinline__ void convl_tile(int tiles, int64_t& line, int& tile) {
    const int bid = xcd_block(1);
    line = bid / tiles;
    tile = bid - (int)(line * tiles);
}

template <typename Ti, typename Tc, int E>
__global__ void __launch_bounds__(1024)
k_convl_a(const V_t<Ti>* __restrict__ x, ConvLArgs a, PassTable pt,
          const V_t<Tc>* __restrict__ tw) {
    CFFT_DYN_LDS(smem);
    int64_t line;
    int tile;
    convl_tile((a.N2 + a.CA - 1) / a.CA, line, tile);
    const int64_t L = (int64_t)a.N1 * a.N2;
    convl_a_body<Ti, Tc, E>(smem, a, tile, x + line * a.idist,
                            reinterpret_cast<V_t<Tc>*>(a.m1) + line * L,
                            reinterpret_cast<V_t<Tc>*>(a.dc) + 2 * line, pt, tw);
}

template <typename Tc, int E>
__global__ void __launch_bounds__(1024)
k_convl_ba(ConvLArgs a, PassTable pt, const V_t<Tc>* __restrict__ tw) {
    CFFT_DYN_LDS(smem);
    int64_t line;
    int tile;
    convl_tile((a.N1 + a.CB - 1) / a.CB, line, tile);
    const int64_t L = (int64_t)a.N1 * a.N2;
    convl_ba_body<Tc, E>(smem, a, tile, reinterpret_cast<const V_t<Tc>*>(a.m1) + line * L,
                         reinterpret_cast<V_t<Tc>*>(a.m2) + line * L,
                         reinterpret_cast<V_t<Tc>*>(a.dc) + 2 * line, pt, tw);
}

template <typename Ti, typename Tc, int E>
__global__ void __launch_bounds__(1024)
k_convl_c(V_t<Ti>* __restrict__ y, ConvLArgs a, PassTable pt, const V_t<Tc>* __restrict__ tw) {
    CFFT_DYN_LDS(smem);
    int64_t line;
    int tile;
    convl_tile((a.N2 + a.CC - 1) / a.CC, line, tile);
    const int64_t L = (int64_t)a.N1 * a.N2;
    convl_c_body<Ti, Tc, E>(smem, a, tile, reinterpret_cast<const V_t<Tc>*>(a.m2) + line * L,
                            y + line * a.odist,
                            reinterpret_cast<const V_t<Tc>*>(a.dc) + 2 * line, pt, tw);
}

// ------------------------------------------------------------------------------------------
// Real-input transforms, n = 2m even: the reals are read as m complex z[j] = x[2j] + i x[2j+1].
//   forward:  Z = FFT_m(z);  X[k] = (Z[k] + conj Z[m-k])/2 - i w^k (Z[k] - conj Z[m-k])/2
//   backward: Z'[k] = (X[k] + conj X[m-k]) + i conj(w^k) (X[k] - conj X[m-k]);  z = bfft_m(Z')
// with w = e^{-2 pi i / n} (table w[k], k < m) and Z[m] = Z[0].
// ------------------------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void r2c_pair(const V_t<T> zk, const V_t<T> zj, const V_t<T> wk,
                                         V_t<T>& xk) {
    const V_t<T> czj = cconj(zj);
    const V_t<T> a = V_t<T>{(T)0.5 * (zk.x + czj.x), (T)0.5 * (zk.y + czj.y)};
    const V_t<T> d = V_t<T>{(T)0.5 * (zk.x - czj.x), (T)0.5 * (zk.y - czj.y)};
    const V_t<T> wd = cmul(wk, d);
    xk = V_t<T>{a.x + wd.y, a.y - wd.x};  // a - i (w d)
}

// Half spectrum X[0..m] of one line from Z in LDS (or any Z(k) source), threads t of Tt.
template <typename T, class ZSRC>
__device__ __forceinline__ void r2c_post(const ZSRC& Z, V_t<T>* X, const V_t<T>* w, int m, int t,
                                         int Tt, bool active, T scale) {
    if (!active) return;
    constexpr int U = 4;  // pairs per batch: their table loads share one wait
    const int h = m / 2;
    for (int k0 = t; k0 <= h; k0 += U * Tt) {
        V_t<T> wk[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = k0 + u * Tt;
            wk[u] = w[k < h ? k : h];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = k0 + u * Tt;
            if (k > h) break;
            const int j = m - k;
            const V_t<T> zk = Z(k), zj = Z(j == m ? 0 : j);
            V_t<T> xk, xj;
            r2c_pair<T>(zk, zj, wk[u], xk);
            const V_t<T> wj = V_t<T>{-wk[u].x, wk[u].y};  // w^{m-k} = -conj(w^k)
            r2c_pair<T>(zj, zk, wj, xj);
            X[k] = V_t<T>{xk.x * scale, xk.y * scale};
            X[j] = V_t<T>{xj.x * scale, xj.y * scale};
        }
    }
}

// conj(Z'[k]) from the half spectrum in global memory (the engine's first pass reads it).
template <typename T>
struct C2RSrc {
    const V_t<T>* X;
    const V_t<T>* w;
    int m;
    __device__ __forceinline__ V_t<T> operator()(int k) const {
        const V_t<T> xk = X[k], cxj = cconj(X[m - k]);
        const V_t<T> a = V_t<T>{xk.x + cxj.x, xk.y + cxj.y};
        const V_t<T> d = V_t<T>{xk.x - cxj.x, xk.y - cxj.y};
        const V_t<T> wd = cmul(cconj(w[k]), d);
        const V_t<T> z = V_t<T>{a.x - wd.y, a.y + wd.x};  // a + i (conj(w) d)
        return cconj(z);
    }
};

// Fused rows: up to C lines; dir > 0 is r2c (reals -> half spectra), dir < 0 is c2r.
template <typename T, int E>
__global__ void __launch_bounds__(1024)
k_real_rows(const void* __restrict__ in, void* __restrict__ out, RealArgs a, PassTable pt,
            const V_t<T>* __restrict__ tw) {
    CFFT_DYN_LDS(smem);
    const int tid = threadIdx.x;
    const int c = tid / a.Tt;
    const int t = tid - c * a.Tt;
    const int64_t line = (int64_t)blockIdx.x * a.C + c;
    const bool active = (c < a.C) && (line < a.nlines);
    const int64_t li = active ? line : 0;
    const LdsLine<T> L = lds_line<T>(smem, a.C, a.Ls, c < a.C ? c : 0);
    const V_t<T>* w = (const V_t<T>*)a.w;
    if (a.forward) {
        const V_t<T>* z = reinterpret_cast<const V_t<T>*>((const T*)in + li * a.rdist);
        V_t<T>* X = (V_t<T>*)out + li * a.cdist;
        GSrc<T> gs{z, (T)1};
        run_line<T, E>(pt, a.m, t, a.Tt, active, tw, gs, L, L, true, false);
        r2c_post<T>(L, X, w, a.m, t, a.Tt, active, (T)a.scale);
    } else {
        const V_t<T>* X = (const V_t<T>*)in + li * a.cdist;
        V_t<T>* z = reinterpret_cast<V_t<T>*>((T*)out + li * a.rdist);
        C2RSrc<T> src{X, w, a.m};
        GRowDst<T> gd{z, (T)a.scale, (T)(-a.scale), nullptr, nullptr};
        run_line<T, E>(pt, a.m, t, a.Tt, active, tw, src, gd, L, true, true);
    }
}

// Elementwise helpers of the general real path (one thread per element, grid.y over lines).
template <typename T>
__global__ void __launch_bounds__(256)
k_real_aux(const void* __restrict__ in, void* __restrict__ out, RealAuxArgs a) {
    const V_t<T>* w = (const V_t<T>*)a.w;
    for (int64_t line = blockIdx.y; line < a.nlines; line += gridDim.y) {
        for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < a.count;
             k += (int64_t)gridDim.x * blockDim.x) {
            const T s = (T)a.scale;
            switch (a.mode) {
                case 0: {  // r2c post: Z (m per line) -> X (m+1 per line), k <= m/2
                    const V_t<T>* Z = (const V_t<T>*)in + line * a.idist;
                    V_t<T>* X = (V_t<T>*)out + line * a.odist;
                    const int64_t j = a.m - k;
                    V_t<T> xk, xj;
                    r2c_pair<T>(Z[k], Z[j == a.m ? 0 : j], w[k], xk);
                    r2c_pair<T>(Z[j == a.m ? 0 : j], Z[k], V_t<T>{-w[k].x, w[k].y}, xj);
                    X[k] = V_t<T>{xk.x * s, xk.y * s};
                    X[j] = V_t<T>{xj.x * s, xj.y * s};
                    break;
                }
                case 1: {  // c2r pre: X (m+1) -> Z' (m) scaled, k < m
                    const V_t<T>* X = (const V_t<T>*)in + line * a.idist;
                    V_t<T>* Z = (V_t<T>*)out + line * a.odist;
                    const V_t<T> z = cconj(C2RSrc<T>{X, w, (int)a.m}(k));
                    Z[k] = V_t<T>{z.x * s, z.y * s};
                    break;
                }
                case 2: {  // reals -> complex with zero imaginary part, k < n
                    const T* x = (const T*)in + line * a.idist;
                    ((V_t<T>*)out)[line * a.odist + k] = V_t<T>{x[k] * s, (T)0};
                    break;
                }
                case 3: {  // first n/2 + 1 of a full spectrum, k <= n/2
                    ((V_t<T>*)out)[line * a.odist + k] = ((const V_t<T>*)in)[line * a.idist + k];
                    break;
                }
                case 4: {  // half spectrum -> full Hermitian spectrum, k < n
                    const V_t<T>* X = (const V_t<T>*)in + line * a.idist;
                    const int64_t h = k <= a.n / 2 ? k : a.n - k;
                    const V_t<T> v = X[h];
                    ((V_t<T>*)out)[line * a.odist + k] = k <= a.n / 2 ? V_t<T>{v.x * s, v.y * s}
                                                                      : V_t<T>{v.x * s, -v.y * s};
                    break;
                }
                default: {  // 5: real part of a complex line, k < n
                    const V_t<T> v = ((const V_t<T>*)in)[line * a.idist + k];
                    ((T*)out)[line * a.odist + k] = v.x * s;
                    break;
                }
            }
        }
    }
}

// ------------------------------------------------------------------------------------------
// Work-list wrappers (one MItem per workgroup).
// ------------------------------------------------------------------------------------------
constexpr int kMultiE = kMultiEHost;
constexpr int kChunk = 4096;  // elements per gather / scatter item

template <typename V>
__device__ __forceinline__ V* mptr(uint64_t a, int sel, uint64_t ib, uint64_t ob) {
    return reinterpret_cast<V*>(mresolve(a, sel, ib, ob));
}

__device__ __forceinline__ void item_direction(const MItem& it, int n, const MultiArgs& ma,
                                               int& sgn, double& scale) {
    if (it.flags & kLaunchDir) {
        sgn = ma.sgn;
        scale = ma.normalize ? 1.0 / (double)n : 1.0;
    } else {
        sgn = it.sgn;
        scale = it.scale;
    }
}

template <typename T>
__global__ void __launch_bounds__(1024)
k_rows_multi(MultiArgs ma, const MItem* __restrict__ items, const MPlan* __restrict__ plans) {
    CFFT_DYN_LDS(smem);
    const MItem it = items[blockIdx.x];
    const MPlan* P = plans + it.plan;
    int sgn;
    double scale;
    item_direction(it, P->n, ma, sgn, scale);
    ItemLines<T> lines{ma.in_base, ma.out_base, &it};
    rows_body<T, kMultiE>(smem, lines, it.count, P->line.N, P->C, P->line.Tt, P->line.Ls,
                          P->line.pt, (const V_t<T>*)P->line.tw, sgn, scale,
                          (it.flags & kEpilogue) ? (const V_t<T>*)P->K : nullptr);
}

template <typename T>
__global__ void __launch_bounds__(1024)
k_fs_a_multi(MultiArgs ma, const MItem* __restrict__ items, const MPlan* __restrict__ plans) {
    CFFT_DYN_LDS(smem);
    const MItem it = items[xcd_item()];
    const MPlan* P = plans + it.plan;
    int sgn;
    double scale;
    item_direction(it, P->n, ma, sgn, scale);
    fs_a_body<T, kMultiE, V_t<T>, false>(smem, mptr<const V_t<T>>(it.in, it.flags & kSelInMask, ma.in_base, ma.out_base),
                          mptr<V_t<T>>(it.out, (it.flags >> kSelOutShift) & 15, ma.in_base, ma.out_base),
                          P->N1, P->N2, P->CA, P->la.Tt, P->la.Ls, it.count * P->CA, sgn,
                          P->la.pt, (const V_t<T>*)P->la.tw);
}

template <typename T>
__global__ void __launch_bounds__(1024)
k_fs_b_multi(MultiArgs ma, const MItem* __restrict__ items, const MPlan* __restrict__ plans) {
    CFFT_DYN_LDS(smem);
    const MItem it = items[xcd_item()];
    const MPlan* P = plans + it.plan;
    int sgn;
    double scale;
    item_direction(it, P->n, ma, sgn, scale);
    const bool epi = it.flags & kEpilogue;
    fs_b_body<T, kMultiE>(smem, mptr<const V_t<T>>(it.in, it.flags & kSelInMask, ma.in_base, ma.out_base),
                          mptr<V_t<T>>(it.out, (it.flags >> kSelOutShift) & 15, ma.in_base, ma.out_base),
                          P->N1, P->N2, P->CB, P->lb.Tt, P->lb.Ls, it.count * P->CB, 1, sgn, scale,
                          epi ? (const V_t<T>*)P->K : nullptr,
                          epi ? reinterpret_cast<V_t<T>*>(it.dc) : nullptr,
                          (const V_t<T>*)P->tw3, P->tbits, P->tnc, P->lb.pt,
                          (const V_t<T>*)P->lb.tw);
}

template <typename T>
__global__ void __launch_bounds__(1024)
k_conv_multi(MultiArgs ma, const MItem* __restrict__ items, const MPlan* __restrict__ plans) {
    CFFT_DYN_LDS(smem);
    const MItem it = items[blockIdx.x];
    const MPlan* P = plans + it.plan;
    int sgn;
    double scale;
    item_direction(it, P->n, ma, sgn, scale);
    ItemLines<T> lines{ma.in_base, ma.out_base, &it};
    conv_body<T, kMultiE>(smem, lines, it.count, P->n, P->L, P->C, P->line.Tt, P->Ls,
                          P->mode, sgn, scale, P->perm, P->iperm, (const V_t<T>*)P->chirp,
                          (const V_t<T>*)P->K, P->line.pt, (const V_t<T>*)P->line.tw);
}

// gather: in = x line, out = work line; scatter: in = work line, out = y line.  dc = the
// line's slots {A^[0], x0}.  One item covers kChunk elements of one line.
template <typename Ti, typename Tc>
__global__ void __launch_bounds__(256)
k_gather_multi(MultiArgs ma, const MItem* __restrict__ items, const MPlan* __restrict__ plans) {
    const MItem it = items[blockIdx.x];
    const MPlan* P = plans + it.plan;
    int sgn;
    double scale;
    item_direction(it, P->n, ma, sgn, scale);
    const int64_t e0 = (int64_t)it.count * kChunk;
    int64_t e1 = e0 + kChunk;
    if (e1 > P->L) e1 = P->L;
    gather_body<Ti, Tc>(mptr<const V_t<Ti>>(it.in, it.flags & kSelInMask, ma.in_base, ma.out_base),
                        mptr<V_t<Tc>>(it.out, (it.flags >> kSelOutShift) & 15, ma.in_base, ma.out_base),
                        reinterpret_cast<V_t<Tc>*>(it.dc), e0 + threadIdx.x, e1, blockDim.x, P->n,
                        P->mode, sgn, P->perm, (const V_t<Tc>*)P->chirp);
}

template <typename Ti, typename Tc>
__global__ void __launch_bounds__(256)
k_scatter_multi(MultiArgs ma, const MItem* __restrict__ items, const MPlan* __restrict__ plans) {
    const MItem it = items[blockIdx.x];
    const MPlan* P = plans + it.plan;
    int sgn;
    double scale;
    item_direction(it, P->n, ma, sgn, scale);
    const int64_t k0 = (int64_t)it.count * kChunk;
    const int64_t k1 = k0 + kChunk;
    scatter_body<Ti, Tc>(mptr<const V_t<Tc>>(it.in, it.flags & kSelInMask, ma.in_base, ma.out_base),
                         mptr<V_t<Ti>>(it.out, (it.flags >> kSelOutShift) & 15, ma.in_base, ma.out_base),
                         reinterpret_cast<const V_t<Tc>*>(it.dc), k0 + threadIdx.x, k1, blockDim.x,
                         P->n, P->mode, sgn, 1.0 / (double)P->L, scale, P->iperm,
                         (const V_t<Tc>*)P->chirp);
}

// Fused large-convolution launches: one item per tile (count = tile index); in / out are the
// tile's input and output lines (x, m1 / m1, m2 / m2, y), dc its slots; the plan carries the
// convolution tables and its FFT_L four-step (N1, N2, tiles, lines, twiddles).
__device__ __forceinline__ ConvLArgs convl_args_of(const MPlan& P, int sgn, double scale) {
    ConvLArgs a{};
    a.n = P.n;
    a.N1 = P.N1;
    a.N2 = P.N2;
    a.CA = P.CA;
    a.CB = P.CB;
    a.CC = P.CC;
    a.TtA = P.la.Tt;
    a.LsA = P.la.Ls;
    a.TtB = P.lb.Tt;
    a.LsB = P.lb.Ls;
    a.TtC = P.la.Tt;
    a.LsC = P.LsC;
    a.E = kMultiE;
    a.nlines = 1;
    a.mode = P.mode;
    a.sgn = sgn;
    a.scale = scale;
    a.perm = P.perm;
    a.iperm = P.iperm;
    a.chirp = P.chirp;
    a.K = P.K;
    a.tw3 = P.tw3;
    a.tbits = P.tbits;
    a.tnc = P.tnc;
    return a;
}

template <typename Ti, typename Tc>
__global__ void __launch_bounds__(1024)
k_convl_a_multi(MultiArgs ma, const MItem* __restrict__ items, const MPlan* __restrict__ plans) {
    CFFT_DYN_LDS(smem);
    const MItem it = items[xcd_item()];
    const MPlan* P = plans + it.plan;
    int sgn;
    double scale;
    item_direction(it, P->n, ma, sgn, scale);
    const ConvLArgs a = convl_args_of(*P, sgn, scale);
    convl_a_body<Ti, Tc, kMultiE>(
        smem, a, it.count, mptr<const V_t<Ti>>(it.in, it.flags & kSelInMask, ma.in_base, ma.out_base),
        mptr<V_t<Tc>>(it.out, (it.flags >> kSelOutShift) & 15, ma.in_base, ma.out_base),
        reinterpret_cast<V_t<Tc>*>(it.dc), P->la.pt, (const V_t<Tc>*)P->la.tw);
}

template <typename Tc>
__global__ void __launch_bounds__(1024)
k_convl_ba_multi(MultiArgs ma, const MItem* __restrict__ items, const MPlan* __restrict__ plans) {
    CFFT_DYN_LDS(smem);
    const MItem it = items[xcd_item()];
    const MPlan* P = plans + it.plan;
    int sgn;
    double scale;
    item_direction(it, P->n, ma, sgn, scale);
    const ConvLArgs a = convl_args_of(*P, sgn, scale);
    convl_ba_body<Tc, kMultiE>(
        smem, a, it.count, mptr<const V_t<Tc>>(it.in, it.flags & kSelInMask, ma.in_base, ma.out_base),
        mptr<V_t<Tc>>(it.out, (it.flags >> kSelOutShift) & 15, ma.in_base, ma.out_base),
        reinterpret_cast<V_t<Tc>*>(it.dc), P->lb.pt, (const V_t<Tc>*)P->lb.tw);
}

template <typename Ti, typename Tc>
__global__ void __launch_bounds__(1024)
k_convl_c_multi(MultiArgs ma, const MItem* __restrict__ items, const MPlan* __restrict__ plans) {
    CFFT_DYN_LDS(smem);
    const MItem it = items[xcd_item()];
    const MPlan* P = plans + it.plan;
    int sgn;
    double scale;
    item_direction(it, P->n, ma, sgn, scale);
    const ConvLArgs a = convl_args_of(*P, sgn, scale);
    convl_c_body<Ti, Tc, kMultiE>(
        smem, a, it.count, mptr<const V_t<Tc>>(it.in, it.flags & kSelInMask, ma.in_base, ma.out_base),
        mptr<V_t<Ti>>(it.out, (it.flags >> kSelOutShift) & 15, ma.in_base, ma.out_base),
        reinterpret_cast<const V_t<Tc>*>(it.dc), P->la.pt, (const V_t<Tc>*)P->la.tw);
}

// ------------------------------------------------------------------------------------------
// Launchers
// ------------------------------------------------------------------------------------------
namespace detail {

template <typename K>
void allow_big_lds(K kernel) {
    // gfx950 lets one workgroup allocate up to 160 KiB of LDS.
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
}

// Launchers; G is the storage type of the lines' global input / output (V_t<T> or ch16).
template <typename T, int E, typename G = V_t<T>, int RM = 16, bool WT = false>
hipError_t rows_t(const void* in, void* out, const RowsArgs& a, const PassTable& pt,
                  const void* tw, hipStream_t st) {
    const size_t lds = 2 * sizeof(T) * (size_t)a.C * a.Ls;
    const int64_t blocks = (a.nlines + a.C - 1) / a.C;
    hipLaunchKernelGGL((k_rows<T, E, G, RM, WT>), dim3((unsigned)blocks), dim3(a.C * a.Tt), lds,
                       st, (const G*)in, (G*)out, a, pt, (const V_t<T>*)tw);
    return hipGetLastError();
}

// Rows launch on the instance for the plan's radices: lines whose radices are all <= 9 use the
// RM = 9 kernel (float: 60 instead of 92 registers per lane, 8 instead of 5 waves per SIMD).
template <typename T, typename G = V_t<T>>
hipError_t rows_rm_t(const void* in, void* out, const RowsArgs& a, const PassTable& pt,
                     const void* tw, hipStream_t st) {
    if constexpr (std::is_same<G, V_t<T>>::value) {
        if (pt.wave)  // lines with a wave tail (planned for float / double storage only)
            return a.E <= 8 ? rows_t<T, 8, G, 16, true>(in, out, a, pt, tw, st)
                            : rows_t<T, 16, G, 16, true>(in, out, a, pt, tw, st);
    }
    if (pass_table_rmax(pt) <= 9)
        return a.E <= 8 ? rows_t<T, 8, G, 9>(in, out, a, pt, tw, st)
                        : rows_t<T, 16, G, 9>(in, out, a, pt, tw, st);
    return a.E <= 8 ? rows_t<T, 8, G>(in, out, a, pt, tw, st)
                    : rows_t<T, 16, G>(in, out, a, pt, tw, st);
}

template <typename T, int E, typename G = V_t<T>>
hipError_t fsa_t(const void* in, void* out, const FourStepAArgs& a, const PassTable& pt,
                 const void* tw, hipStream_t st) {
    const size_t lds = 2 * sizeof(T) * (size_t)a.C * a.Ls;
    const int64_t blocks = (int64_t)((a.N2 + a.C - 1) / a.C) * a.batch;
    if (a.mid_plain)
        hipLaunchKernelGGL((k_fs_a<T, E, G, false>), dim3((unsigned)blocks), dim3(a.C * a.Tt), lds,
                           st, (const G*)in, (V_t<T>*)out, a, pt, (const V_t<T>*)tw);
    else
        hipLaunchKernelGGL((k_fs_a<T, E, G>), dim3((unsigned)blocks), dim3(a.C * a.Tt), lds, st,
                           (const G*)in, (V_t<T>*)out, a, pt, (const V_t<T>*)tw);
    return hipGetLastError();
}

template <typename T, int E, typename G = V_t<T>>
hipError_t fsb_t(const void* in, void* out, const FourStepBArgs& a, const PassTable& pt,
                 const void* tw, hipStream_t st) {
    const size_t lds = fsb_tw_offset(sizeof(T), a.C, a.Ls) +
                       (a.tw3 ? (size_t)tw3_entries(a.tbits, a.tnc) * sizeof(V_t<T>) : 0);
    const int64_t blocks = (int64_t)((a.N1 + a.C - 1) / a.C) * a.batch;
    hipLaunchKernelGGL((k_fs_b<T, E, G>), dim3((unsigned)blocks), dim3(a.C * a.Tt), lds, st,
                       (const V_t<T>*)in, (G*)out, a, pt, (const V_t<T>*)tw);
    return hipGetLastError();
}

template <typename T, int E>
hipError_t conv_t(const void* in, void* out, const ConvSmallArgs& a, const PassTable& pt,
                  const void* tw, hipStream_t st) {
    const size_t lds = 2 * sizeof(T) * (size_t)a.C * a.Ls + sizeof(V_t<T>) * (size_t)a.C;
    const int64_t blocks = (a.nlines + a.C - 1) / a.C;
    if (pt.wave)
        hipLaunchKernelGGL((k_conv_small<T, E, true>), dim3((unsigned)blocks), dim3(a.C * a.Tt),
                           lds, st, (const V_t<T>*)in, (V_t<T>*)out, a, pt, (const V_t<T>*)tw);
    else
        hipLaunchKernelGGL((k_conv_small<T, E>), dim3((unsigned)blocks), dim3(a.C * a.Tt), lds,
                           st, (const V_t<T>*)in, (V_t<T>*)out, a, pt, (const V_t<T>*)tw);
    return hipGetLastError();
}

template <typename Ti, typename Tc>
hipError_t gather_t(const void* x, void* w, const ConvGatherArgs& a, hipStream_t st) {
    const int64_t per = (a.L + 255) / 256;
    const unsigned gx = (unsigned)(per < 4096 ? per : 4096);
    const unsigned gy = (unsigned)(a.nlines < 65535 ? a.nlines : 65535);
    hipLaunchKernelGGL((k_conv_gather<Ti, Tc>), dim3(gx, gy), dim3(256), 0, st,
                       (const V_t<Ti>*)x, (V_t<Tc>*)w, a);
    return hipGetLastError();
}

template <typename Ti, typename Tc>
hipError_t scatter_t(const void* w, void* y, const ConvScatterArgs& a, hipStream_t st) {
    const int64_t per = (a.n + 255) / 256;
    const unsigned gx = (unsigned)(per < 4096 ? per : 4096);
    const unsigned gy = (unsigned)(a.nlines < 65535 ? a.nlines : 65535);
    hipLaunchKernelGGL((k_conv_scatter<Ti, Tc>), dim3(gx, gy), dim3(256), 0, st,
                       (const V_t<Tc>*)w, (V_t<Ti>*)y, a);
    return hipGetLastError();
}

template <typename Ti, typename Tc, int E>
hipError_t convl_t(const void* x, void* y, const ConvLArgs& a, const PassTable& ptA,
                   const void* twA, const PassTable& ptB, const void* twB, hipStream_t st) {
    const size_t es = sizeof(Tc);
    const size_t t3 = (size_t)tw3_entries(a.tbits, a.tnc) * 2 * es;
    const size_t ldsA = 2 * es * (size_t)a.CA * a.LsA;
    const size_t ldsB = fsb_tw_offset(es, a.CB, a.LsB) + t3;
    const size_t ldsC = fsb_tw_offset(es, a.CC, a.LsC) + t3;
    const int64_t gA = (int64_t)((a.N2 + a.CA - 1) / a.CA) * a.nlines;
    const int64_t gB = (int64_t)((a.N1 + a.CB - 1) / a.CB) * a.nlines;
    const int64_t gC = (int64_t)((a.N2 + a.CC - 1) / a.CC) * a.nlines;
    if (gA > 0x7fffffff || gB > 0x7fffffff || gC > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_convl_a<Ti, Tc, E>), dim3((unsigned)gA), dim3(a.CA * a.TtA), ldsA, st,
                       (const V_t<Ti>*)x, a, ptA, (const V_t<Tc>*)twA);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_convl_ba<Tc, E>), dim3((unsigned)gB), dim3(a.CB * a.TtB), ldsB, st, a,
                       ptB, (const V_t<Tc>*)twB);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_convl_c<Ti, Tc, E>), dim3((unsigned)gC), dim3(a.CC * a.TtC), ldsC, st,
                       (V_t<Ti>*)y, a, ptA, (const V_t<Tc>*)twA);
    return hipGetLastError();
}

template <typename T, int E>
hipError_t real_rows_t(const void* in, void* out, const RealArgs& a, const PassTable& pt,
                       const void* tw, hipStream_t st) {
    const size_t lds = 2 * sizeof(T) * (size_t)a.C * a.Ls;
    const int64_t blocks = (a.nlines + a.C - 1) / a.C;
    hipLaunchKernelGGL((k_real_rows<T, E>), dim3((unsigned)blocks), dim3(a.C * a.Tt), lds, st, in,
                       out, a, pt, (const V_t<T>*)tw);
    return hipGetLastError();
}

template <typename T>
hipError_t real_aux_t(const void* in, void* out, const RealAuxArgs& a, hipStream_t st) {
    const int64_t per = (a.count + 255) / 256;
    const unsigned gx = (unsigned)(per < 4096 ? (per > 0 ? per : 1) : 4096);
    const unsigned gy = (unsigned)(a.nlines < 65535 ? a.nlines : 65535);
    hipLaunchKernelGGL(k_real_aux<T>, dim3(gx, gy), dim3(256), 0, st, in, out, a);
    return hipGetLastError();
}

// Work-list launch for data precision T (rows, four-step, fused convolution).
template <typename T>
hipError_t multi_t(const MultiLaunch& L, const MultiArgs& ma, hipStream_t st) {
    const dim3 grid((unsigned)L.nitems), block(L.block);
    switch ((MultiKernel)L.kernel) {
        case MultiKernel::Rows:
            hipLaunchKernelGGL(k_rows_multi<T>, grid, block, L.lds, st, ma, L.items, L.plans);
            break;
        case MultiKernel::FourStepA:
            hipLaunchKernelGGL(k_fs_a_multi<T>, grid, block, L.lds, st, ma, L.items, L.plans);
            break;
        case MultiKernel::FourStepB:
            hipLaunchKernelGGL(k_fs_b_multi<T>, grid, block, L.lds, st, ma, L.items, L.plans);
            break;
        case MultiKernel::ConvSmall:
            hipLaunchKernelGGL(k_conv_multi<T>, grid, block, L.lds, st, ma, L.items, L.plans);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Work-list gather / scatter of large convolutions (data Ti, work buffer Tc).
template <typename Ti, typename Tc>
hipError_t gs_multi_t(const MultiLaunch& L, const MultiArgs& ma, hipStream_t st) {
    const dim3 grid((unsigned)L.nitems), block(256), lblock(L.block);
    switch ((MultiKernel)L.kernel) {
        case MultiKernel::Gather:
            hipLaunchKernelGGL((k_gather_multi<Ti, Tc>), grid, block, 0, st, ma, L.items, L.plans);
            break;
        case MultiKernel::Scatter:
            hipLaunchKernelGGL((k_scatter_multi<Ti, Tc>), grid, block, 0, st, ma, L.items, L.plans);
            break;
        case MultiKernel::ConvLA:
            hipLaunchKernelGGL((k_convl_a_multi<Ti, Tc>), grid, lblock, L.lds, st, ma, L.items,
                               L.plans);
            break;
        case MultiKernel::ConvLBA:
            hipLaunchKernelGGL((k_convl_ba_multi<Tc>), grid, lblock, L.lds, st, ma, L.items,
                               L.plans);
            break;
        case MultiKernel::ConvLC:
            hipLaunchKernelGGL((k_convl_c_multi<Ti, Tc>), grid, lblock, L.lds, st, ma, L.items,
                               L.plans);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace detail
}  // namespace cfft
