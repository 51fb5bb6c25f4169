// engine.h -- device-side Stockham engine shared by every kernel of the library.
//
// A "line" is one 1-D transform of length N held by Tt threads of a workgroup.  The passes are
// the self-sorting Stockham DIF passes of CatabraFFT.jl's layered codelets
// (radix_2_codelets.jl:402-447 and the radix-3/5/7 twins):
//
//     out[q + s(R p + k)] = w_g^{p k} * sum_j in[q + s(p + j n1)] w_R^{j k}
//     g = N / s, n1 = g / R, q < s, p < n1, butterfly index b = q + s p < N / R
//
// Butterfly b reads in[b + j N/R] (unit stride across lanes) and, in the last pass (n1 = 1),
// writes out[b + k N/R] (unit stride across lanes).  So the first pass can read HBM directly and
// the last pass can write HBM directly, both fully coalesced; passes in between exchange through
// LDS.  LDS holds re and im in separate arrays with one pad word every 2^PS words.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "codelets.h"

// Dynamic LDS of a kernel, and an opaque register copy (an empty asm the compiler cannot see
// through).  Both are macros so that the kernel sources also compile in the CPU emulation
// harness of the tests (tests/emu), which supplies its own definitions.
#ifndef CFFT_DYN_LDS
#define CFFT_DYN_LDS(name) extern __shared__ __align__(16) unsigned char name[]
#endif
#ifndef CFFT_OPAQUE
#define CFFT_OPAQUE(x) asm volatile("" : "+v"(x))
#endif

namespace cfft {

template <typename T> struct Vec2;
template <> struct Vec2<float> { using type = float2; };
template <> struct Vec2<double> { using type = double2; };

// Two ComplexF32 lines computed in lockstep by one thread ("pair" kernels): f2p holds one float
// of each line (a: the first line, b: the second), cp2 one complex element of each.  Every
// operation acts on a and b separately, with the same IEEE operations as the float path, so a
// pair of lines gives the float path's bits.  The engine and the codelets run on them unchanged
// (T = f2p); twiddles, equal for the two lines, stay float2 (cmul(cp2, float2)).
struct f2p {
    float a, b;
    f2p() = default;
    __host__ __device__ constexpr f2p(float x, float y) : a(x), b(y) {}
    __host__ __device__ constexpr explicit f2p(double c) : a((float)c), b((float)c) {}
    __host__ __device__ constexpr explicit f2p(int c) : a((float)c), b((float)c) {}
};
__host__ __device__ __forceinline__ f2p operator+(f2p x, f2p y) { return f2p(x.a + y.a, x.b + y.b); }
__host__ __device__ __forceinline__ f2p operator-(f2p x, f2p y) { return f2p(x.a - y.a, x.b - y.b); }
__host__ __device__ __forceinline__ f2p operator-(f2p x) { return f2p(-x.a, -x.b); }
__host__ __device__ __forceinline__ f2p operator*(f2p x, f2p y) { return f2p(x.a * y.a, x.b * y.b); }
__host__ __device__ __forceinline__ f2p operator*(f2p x, float y) { return f2p(x.a * y, x.b * y); }
__host__ __device__ __forceinline__ f2p& operator*=(f2p& x, f2p y) {
    x.a *= y.a;
    x.b *= y.b;
    return x;
}
__host__ __device__ __forceinline__ f2p fmadd(f2p x, f2p y, f2p z) {
    return f2p(fmadd(x.a, y.a, z.a), fmadd(x.b, y.b, z.b));
}
struct cp2 {
    f2p x, y;
};
template <> struct Vec2<f2p> { using type = cp2; };
// a * w for a pair element and a twiddle shared by both lines
__host__ __device__ __forceinline__ cp2 cmul(cp2 a, float2 w) {
    const f2p wx(w.x, w.x), wy(w.y, w.y);
    return cp2{fmadd(a.x, wx, -(a.y * wy)), fmadd(a.x, wy, a.y * wx)};
}

constexpr int kMaxPasses = 20;

// Radix sequence of one line FFT plus the offsets of each pass's twiddle block
// tw[twoff[i] + p (R-1) + (k-1)] = w_g^{p k}.
struct PassTable {
    int npass;
    // 1: the last pass runs across the lanes of a wavefront (wave_tail_pass); the host sets it
    // only for lines that meet wave_tail_pass's conditions and kernels that map threads row first
    int wave;
    int radix[kMaxPasses];
    int twoff[kMaxPasses];
    // the stride s of pass i (the product of the radices before it) as a multiplier: b / s =
    // umulhi(b, sdiv[i]) with sdiv[i] = ceil(2^32 / s), exact for b, s < 2^16; 0 for s = 1
    unsigned sdiv[kMaxPasses];
};

// b / s from the pass table's multiplier (a multiply-high instead of a runtime division).
__host__ __device__ constexpr unsigned stride_divisor(int s) {
    return s <= 1 ? 0u : (unsigned)(((1ull << 32) + (unsigned long long)s - 1) / (unsigned)s);
}
__device__ __forceinline__ int div_stride(int b, unsigned m) {
    return m ? (int)__umulhi((unsigned)b, m) : b;
}

template <typename V>
__device__ __forceinline__ V cmul(V a, V b) {
    return V{fmadd(a.x, b.x, -(a.y * b.y)), fmadd(a.x, b.y, a.y * b.x)};
}
template <typename V>
__device__ __forceinline__ V cconj(V a) { return V{a.x, -a.y}; }

template <typename T> struct PadShift;
template <> struct PadShift<float> { static constexpr int v = 5; };
template <> struct PadShift<double> { static constexpr int v = 4; };
template <> struct PadShift<f2p> { static constexpr int v = 4; };  // 8-byte words, as double

template <typename T>
__host__ __device__ constexpr int lds_index(int e) { return e + (e >> PadShift<T>::v); }

// Shift of an unpadded LDS line: e >> 31 is 0 for every element index.
constexpr int kNoPadShift = 31;

// One line's slice of LDS (structure of arrays).  Element e sits at word e + (e >> sh): one pad
// word per 2^sh elements (lds_index) spreads power-of-two strides over the banks; sh =
// kNoPadShift leaves the line unpadded, which lets a direct line whose workgroups are bounded
// by LDS fit one more of them per CU.
template <typename T>
struct LdsLine {
    using V = typename Vec2<T>::type;
    T* re;
    T* im;
    int sh = PadShift<T>::v;
    __device__ __forceinline__ V operator()(int e) const {
        const int a = e + (e >> sh);
        return V{re[a], im[a]};
    }
    __device__ __forceinline__ void put(int e, V v) const {
        const int a = e + (e >> sh);
        re[a] = v.x;
        im[a] = v.y;
    }
};

// Registers per lane that a pass holding `data` registers of butterfly inputs may spend on
// twiddles loaded ahead of its butterflies: up to 32, less as the pass's own data grows, so
// that with 1024-thread workgroups (128 registers per lane) nothing spills.
__host__ __device__ constexpr int tw_prefetch_vgprs(int data) {
    return 96 - 2 * data < 0 ? 0 : (96 - 2 * data > 32 ? 32 : 96 - 2 * data);
}

// Per-thread butterfly capacity: MB butterflies of radix R with E complex registers.
__host__ __device__ constexpr int bfly_per_thread(int E, int R) { return E / R > 0 ? E / R : 1; }

// Twiddle tables.  Pass i of a line reads w_g^{p k} (p < n1, 1 <= k < R) from its block at
// twoff[i] through an accessor:
//   TwGlobal: the plan's table in global memory, row-major per pass, [p (R-1) + k - 1];
//   TwLds:    a copy in LDS, k-major per pass, [(k-1) n1 + p]: consecutive butterflies read
//             consecutive words, so a wave's reads are free of bank conflicts.
// Global twiddles are loaded ahead of the butterflies (kPrefetch); LDS ones where they are used.
template <typename V>
struct TwRowAcc {
    const V* p;
    static constexpr bool kPrefetch = true;
    template <int R>
    __device__ __forceinline__ V get(int pp, int k) const { return p[pp * (R - 1) + k - 1]; }
};
template <typename V>
struct TwColAcc {
    const V* p;
    int n1;
    static constexpr bool kPrefetch = false;
    template <int R>
    __device__ __forceinline__ V get(int pp, int k) const { return p[(k - 1) * n1 + pp]; }
};
template <typename V>
struct TwGlobal {
    const V* tw;
    __device__ __forceinline__ TwRowAcc<V> pass(int off, int n1) const { return {tw + off}; }
};
template <typename V>
struct TwLds {
    const V* tw;
    __device__ __forceinline__ TwColAcc<V> pass(int off, int n1) const { return {tw + off, n1}; }
};

// HasPost<S>: S has a member post(e, v), applied to element e after the pass's loads.
template <class S, class = void>
struct HasPost : std::false_type {};
template <class S>
struct HasPost<S, std::void_t<decltype(&S::post)>> : std::true_type {};

// One Stockham pass.  SRC(e) returns element e of the pass input, DST.put(e, v) stores element e
// of its output.  mid_barrier separates the reads from the writes when both go to the same LDS.
template <typename T, int R, int MB, class SRC, class DST, class TWA>
__device__ __forceinline__ void stockham_pass(const int N, const int s, const unsigned sm,
                                              const int t_in, const int Tt,
                                              const bool active, const TWA& tw,
                                              const SRC& src, const DST& dst,
                                              const bool mid_barrier) {
    using V = typename Vec2<T>::type;
    // Keep per-pass address arithmetic inside the pass: an opaque copy of the lane index stops
    // the compiler from hoisting every radix's addresses out of the pass loop.
    int t_ = t_in;
    CFFT_OPAQUE(t_);
    const int t = t_;
    const int nb = N / R;
    // Twiddles first, from a clamped (always valid) butterfly index: their memory latency then
    // overlaps the source reads and the butterflies instead of being paid once per small group
    // of loads.  tw_prefetch_vgprs bounds the registers per lane that hold them (KP twiddles per
    // butterfly); the rest are loaded after the butterflies.  The table has rows for p = 0 (all
    // ones), so every butterfly multiplies.  The last pass (s = N / R) has no twiddles.
    // W: the twiddle type the accessor returns (V, or float2 for pair lines)
    using W = std::remove_cv_t<std::remove_reference_t<decltype(tw.template get<R>(0, 1))>>;
    constexpr int VW = (int)(sizeof(V) / 4), WW = (int)(sizeof(W) / 4);
    constexpr int KP0 = TWA::kPrefetch ? tw_prefetch_vgprs(MB * R * VW) / (MB * WW) : 0;
    constexpr int KP = (R - 1) < KP0 ? (R - 1) : KP0;
    const bool twiddled = s * R < N;
    W w[MB][KP > 0 ? KP : 1];
    if (twiddled) {
#pragma unroll
        for (int ib = 0; ib < MB; ++ib) {
            const int bt = t + ib * Tt;
            const int b = bt < nb ? bt : nb - 1;
#pragma unroll
            for (int k = 0; k < KP; ++k) w[ib][k] = tw.template get<R>(div_stride(b, sm), k + 1);
        }
    }
    V v[MB][R];
#pragma unroll
    for (int ib = 0; ib < MB; ++ib) {
        const int b = t + ib * Tt;
        if (active && b < nb) {
#pragma unroll
            for (int j = 0; j < R; ++j) v[ib][j] = src(b + j * nb);
        }
    }
    // A source with a post step (work on each loaded value, e.g. a twiddle from a table) applies
    // it once all of the pass's loads are issued, so one wait covers them.
    if constexpr (HasPost<SRC>::value) {
#pragma unroll
        for (int ib = 0; ib < MB; ++ib) {
            const int b = t + ib * Tt;
            if (active && b < nb) {
#pragma unroll
                for (int j = 0; j < R; ++j) v[ib][j] = src.post(b + j * nb, v[ib][j]);
            }
        }
    }
    if (mid_barrier) __syncthreads();
#pragma unroll
    for (int ib = 0; ib < MB; ++ib) {
        const int b = t + ib * Tt;
        if (active && b < nb) {
            Dft<R>::run(v[ib]);
            const int p = div_stride(b, sm);
            const int q = b - p * s;
            if (twiddled) {
#pragma unroll
                for (int k = 1; k <= KP; ++k) v[ib][k] = cmul(v[ib][k], w[ib][k - 1]);
#pragma unroll
                for (int k = KP + 1; k < R; ++k) v[ib][k] = cmul(v[ib][k], tw.template get<R>(p, k));
            }
            const int base = q + s * R * p;
#pragma unroll
            for (int k = 0; k < R; ++k) dst.put(base + s * k, v[ib][k]);
        }
    }
}

// ---------------------------------------------------------------------------- lane exchanges
// v of the lane whose index differs from this lane's by xor D, within its wavefront of 64 lanes,
// in VALU instructions only (no LDS): DPP quad_perm for D = 1, 2, row_ror:8 for D = 8,
// row_shl:4 / row_shr:4 and a select for D = 4, and gfx950's v_permlane16_swap / v_permlane32_swap
// and a select for D = 16 / 32.  Every lane of the wavefront executes it (uniform control flow).
// The CPU emulation harness of the tests supplies its own exchange (CFFT_EMU_LANES).
#ifndef CFFT_EMU_LANES
__device__ __forceinline__ int lane_id() {
    return __builtin_amdgcn_mbcnt_hi(-1, __builtin_amdgcn_mbcnt_lo(-1, 0));
}
template <int D>
__device__ __forceinline__ int lane_xor_i32(int v) {
    if constexpr (D == 1) {
        return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
    } else if constexpr (D == 2) {
        return __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
    } else if constexpr (D == 4) {
        const int up = __builtin_amdgcn_update_dpp(0, v, 0x104, 0xF, 0xF, false);  // lane + 4
        const int dn = __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, false);  // lane - 4
        return (lane_id() & 4) ? dn : up;
    } else if constexpr (D == 8) {
        return __builtin_amdgcn_update_dpp(0, v, 0x128, 0xF, 0xF, false);  // row_ror:8
    } else if constexpr (D == 16) {
        // swaps odd rows of the first operand with even rows of the second: with both = v, the
        // first result holds rows {0,0,2,2} and the second rows {1,1,3,3}
        const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
        return (lane_id() & 16) ? (int)r[0] : (int)r[1];
    } else {
        static_assert(D == 32, "lane distance");
        // swaps the upper half of the first operand with the lower half of the second
        const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
        return (lane_id() & 32) ? (int)r[0] : (int)r[1];
    }
}
#else
template <int D>
__device__ __forceinline__ int lane_xor_i32(int v) { return cfft_emu_lane_xor(v, D); }
#endif

template <int D>
__device__ __forceinline__ float lane_xor(float v) {
    return __builtin_bit_cast(float, lane_xor_i32<D>(__builtin_bit_cast(int, v)));
}
template <int D>
__device__ __forceinline__ double lane_xor(double v) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)lane_xor_i32<D>((int)(unsigned)u);
    const unsigned hi = (unsigned)lane_xor_i32<D>((int)(unsigned)(u >> 32));
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
// d is uniform across the wavefront (a scalar branch).
template <typename V>
__device__ __forceinline__ V lane_xor_v(const V v, const int d) {
    switch (d) {
        case 1: return V{lane_xor<1>(v.x), lane_xor<1>(v.y)};
        case 2: return V{lane_xor<2>(v.x), lane_xor<2>(v.y)};
        case 4: return V{lane_xor<4>(v.x), lane_xor<4>(v.y)};
        case 8: return V{lane_xor<8>(v.x), lane_xor<8>(v.y)};
        case 16: return V{lane_xor<16>(v.x), lane_xor<16>(v.y)};
        default: return V{lane_xor<32>(v.x), lane_xor<32>(v.y)};
    }
}

__host__ __device__ constexpr int bit_reverse(int x, int bits) {
    int r = 0;
    for (int i = 0; i < bits; ++i) r |= ((x >> i) & 1) << (bits - 1 - i);
    return r;
}

// The last two passes of a line whose last radix RL (a power of two, <= 16) spans lanes of one
// wavefront.  The penultimate pass (radix R, stride s a power of two, one butterfly per thread,
// b = t = q + s p) computes its butterfly and twiddles as stockham_pass does, but keeps the R
// outputs in registers; the last pass -- for each output k, a DFT_RL over the digit p, whose
// values sit in lanes s apart -- then runs as log2 RL radix-2 decimation-in-frequency stages
// across lanes (lane_xor_v): at the stage of digit bit h the lane with that bit clear keeps
// self + partner, the lane with it set (partner - self) w_{2h}^{p mod h}.  That replaces the
// LDS write, barrier and LDS read between the two passes.  Lane p ends with output
// k' = bitrev(p) of each DFT, which is out[q + s k + (N / RL) k'] of the line.  Every thread of
// the workgroup runs the exchanges (threads without a butterfly carry zeros).  barrier: the
// penultimate pass's reads and the last pass's writes may touch the same LDS.
template <typename T, int R, class SRC, class DST, class TWA>
__device__ __forceinline__ void wave_tail_pass(const int N, const int s, const unsigned sm,
                                               const int RL,
                                               const int t_in, const bool active, const TWA& tw,
                                               const SRC& src, const DST& dst,
                                               const bool barrier) {
    using V = typename Vec2<T>::type;
    int t_ = t_in;
    CFFT_OPAQUE(t_);
    const int b = t_;
    const int nb = N / R;
    const bool live = active && b < nb;
    const int bc = b < nb ? b : nb - 1;
    const int p = div_stride(bc, sm);
    const int q = bc - p * s;
    constexpr int VW = (int)(sizeof(V) / 4);
    constexpr int KP0 = TWA::kPrefetch ? tw_prefetch_vgprs(R * VW) / VW : 0;
    constexpr int KP = (R - 1) < KP0 ? (R - 1) : KP0;
    V w[KP > 0 ? KP : 1];
#pragma unroll
    for (int k = 0; k < KP; ++k) w[k] = tw.template get<R>(p, k + 1);
    V v[R];
#pragma unroll
    for (int j = 0; j < R; ++j) v[j] = V{0, 0};
    if (live) {
#pragma unroll
        for (int j = 0; j < R; ++j) v[j] = src(b + j * nb);
        if constexpr (HasPost<SRC>::value) {
#pragma unroll
            for (int j = 0; j < R; ++j) v[j] = src.post(b + j * nb, v[j]);
        }
        Dft<R>::run(v);
#pragma unroll
        for (int k = 1; k <= KP; ++k) v[k] = cmul(v[k], w[k - 1]);
#pragma unroll
        for (int k = KP + 1; k < R; ++k) v[k] = cmul(v[k], tw.template get<R>(p, k));
    }
    for (int h = RL >> 1; h >= 1; h >>= 1) {
        const bool upper = (p & h) != 0;
        const T sg = upper ? T(-1) : T(1);
        const V wl = lane_w16<V>(((p & (h - 1)) * (8 / h)) & 15);
        const bool twiddled = upper && h > 1;
        const int d = s * h;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const V o = lane_xor_v(v[k], d);
            const V e = V{o.x + sg * v[k].x, o.y + sg * v[k].y};
            const V f = cmul(e, wl);
            v[k] = twiddled ? f : e;
        }
    }
    if (barrier) __syncthreads();
    if (live) {
        int rl_bits = 0;
        while ((1 << rl_bits) < RL) ++rl_bits;
        const int kp = bit_reverse(p, rl_bits);
        const int base = q + (N / RL) * kp;
#pragma unroll
        for (int k = 0; k < R; ++k) dst.put(base + s * k, v[k]);
    }
}

// Dispatch of one pass on its radix.  The four source/destination combinations are
// instantiated by the callers below.  Radices above RM_, a compile-time bound of the kernel, are
// left out: a kernel built for lines of short radices holds fewer registers per lane, so more
// waves fit a SIMD.
#define CFFT_RADIX_CASE(RM_, RR, CALL) \
    case RR:                          \
        if constexpr ((RM_) >= RR) {  \
            CALL(RR);                 \
        }                             \
        break;
#define CFFT_RADIX_SWITCH(RM_, R_, CALL)   \
    switch (R_) {                          \
        CFFT_RADIX_CASE(RM_, 16, CALL)     \
        CFFT_RADIX_CASE(RM_, 13, CALL)     \
        CFFT_RADIX_CASE(RM_, 11, CALL)     \
        CFFT_RADIX_CASE(RM_, 9, CALL)      \
        CFFT_RADIX_CASE(RM_, 8, CALL)      \
        CFFT_RADIX_CASE(RM_, 7, CALL)      \
        CFFT_RADIX_CASE(RM_, 5, CALL)      \
        CFFT_RADIX_CASE(RM_, 4, CALL)      \
        CFFT_RADIX_CASE(RM_, 3, CALL)      \
        CFFT_RADIX_CASE(RM_, 2, CALL)      \
        default: break;                    \
    }

// Penultimate radices a wave tail takes (powers of two, within the kernel's bound RM_).
#define CFFT_WAVE_SWITCH(RM_, R_, CALL) \
    switch (R_) {                       \
        CFFT_RADIX_CASE(RM_, 16, CALL)  \
        CFFT_RADIX_CASE(RM_, 8, CALL)   \
        CFFT_RADIX_CASE(RM_, 4, CALL)   \
        CFFT_RADIX_CASE(RM_, 2, CALL)   \
        default: break;                 \
    }

// The largest radix a kernel must hold for a pass table (16: every radix).
inline int pass_table_rmax(const PassTable& pt) {
    int r = 2;
    for (int i = 0; i < pt.npass; ++i) r = pt.radix[i] > r ? pt.radix[i] : r;
    return r;
}

// Runs passes ip0.. of one line (s0 = product of the radices before ip0).  GSRC / GDST are the
// input and output functors of the line (HBM, or a gather / scatter over another LDS buffer);
// when from_global (to_global) is false the line's input is already in LDS (its output is left
// in LDS).  dst_aliases_lds asks for a barrier between the reads and writes of the last pass
// because GDST writes the line's own LDS.  The caller provides the barrier before the first pass
// when it staged the input into LDS.  src_aliases_lds asks for a barrier between the reads and
// writes of the first pass because GSRC reads the line's own LDS.  TWT is a twiddle table
// (TwGlobal or TwLds).
template <typename T, int E, int RM = 16, bool WT = false, class TWT, class GSRC, class GDST>
__device__ __forceinline__ void run_line_t(const PassTable& pt, const int N, const int t,
                                           const int Tt, const bool active, const TWT& twt,
                                           const GSRC& gsrc, const GDST& gdst,
                                           const LdsLine<T>& lds, const bool from_global,
                                           const bool to_global, const bool dst_aliases_lds,
                                           const int ip0, const int s0,
                                           const bool src_aliases_lds = false) {
    int s = s0;
    for (int ip = ip0; ip < pt.npass; ++ip) {
        const int R = pt.radix[ip];
        const bool first = ip == ip0, last = ip == pt.npass - 1;
        const auto twa = twt.pass(pt.twoff[ip], N / (s * R));
        const bool sg = first && from_global;
        if constexpr (WT) {
          if (pt.wave && ip == pt.npass - 2) {
            // the last two passes: the last one across the lanes of a wavefront
            const int RL = pt.radix[pt.npass - 1];
            const bool reads_lds = !sg || src_aliases_lds;
            const bool writes_lds = !to_global || dst_aliases_lds;
            const bool bar = reads_lds && writes_lds;
            if (sg && to_global) {
#define CFFT_CALL_WGG(RR) wave_tail_pass<T, RR>(N, s, pt.sdiv[ip], RL, t, active, twa, gsrc, gdst, bar)
                CFFT_WAVE_SWITCH(RM, R, CFFT_CALL_WGG)
#undef CFFT_CALL_WGG
            } else if (sg) {
#define CFFT_CALL_WGL(RR) wave_tail_pass<T, RR>(N, s, pt.sdiv[ip], RL, t, active, twa, gsrc, lds, bar)
                CFFT_WAVE_SWITCH(RM, R, CFFT_CALL_WGL)
#undef CFFT_CALL_WGL
            } else if (to_global) {
#define CFFT_CALL_WLG(RR) wave_tail_pass<T, RR>(N, s, pt.sdiv[ip], RL, t, active, twa, lds, gdst, bar)
                CFFT_WAVE_SWITCH(RM, R, CFFT_CALL_WLG)
#undef CFFT_CALL_WLG
            } else {
#define CFFT_CALL_WLL(RR) wave_tail_pass<T, RR>(N, s, pt.sdiv[ip], RL, t, active, twa, lds, lds, bar)
                CFFT_WAVE_SWITCH(RM, R, CFFT_CALL_WLL)
#undef CFFT_CALL_WLL
            }
            if (!to_global) __syncthreads();
            return;
          }
        }
        const bool dg = last && to_global;
        if (sg && dg) {
#define CFFT_CALL_GG(RR) stockham_pass<T, RR, bfly_per_thread(E, RR)>(N, s, pt.sdiv[ip], t, Tt, active, twa, gsrc, gdst, dst_aliases_lds || src_aliases_lds)
            CFFT_RADIX_SWITCH(RM, R, CFFT_CALL_GG)
#undef CFFT_CALL_GG
        } else if (sg) {
#define CFFT_CALL_GL(RR) stockham_pass<T, RR, bfly_per_thread(E, RR)>(N, s, pt.sdiv[ip], t, Tt, active, twa, gsrc, lds, src_aliases_lds)
            CFFT_RADIX_SWITCH(RM, R, CFFT_CALL_GL)
#undef CFFT_CALL_GL
        } else if (dg) {
#define CFFT_CALL_LG(RR) stockham_pass<T, RR, bfly_per_thread(E, RR)>(N, s, pt.sdiv[ip], t, Tt, active, twa, lds, gdst, dst_aliases_lds)
            CFFT_RADIX_SWITCH(RM, R, CFFT_CALL_LG)
#undef CFFT_CALL_LG
        } else {
#define CFFT_CALL_LL(RR) stockham_pass<T, RR, bfly_per_thread(E, RR)>(N, s, pt.sdiv[ip], t, Tt, active, twa, lds, lds, true)
            CFFT_RADIX_SWITCH(RM, R, CFFT_CALL_LL)
#undef CFFT_CALL_LL
        }
        // The pass output sits in LDS for the next pass (or for the caller's staged store).
        if (!dg) __syncthreads();
        s *= R;
    }
}

// Every pass of one line with the plan's twiddle table in global memory.  WT: the kernel
// instance holds the wave-tail path (pt.wave selects it at run time).
template <typename T, int E, int RM = 16, bool WT = false, class GSRC, class GDST>
__device__ __forceinline__ void run_line(const PassTable& pt, const int N, const int t,
                                         const int Tt, const bool active,
                                         const typename Vec2<T>::type* __restrict__ tw,
                                         const GSRC& gsrc, const GDST& gdst,
                                         const LdsLine<T>& lds, const bool from_global,
                                         const bool to_global,
                                         const bool dst_aliases_lds = false) {
    run_line_t<T, E, RM, WT>(pt, N, t, Tt, active, TwGlobal<typename Vec2<T>::type>{tw}, gsrc, gdst, lds,
                     from_global, to_global, dst_aliases_lds, 0, 1);
}

}  // namespace cfft
