// kernels_pair_f32.hip -- ComplexF32 four-step passes that carry two adjacent lines per thread.
//
// A four-step pass reads its lines as runs of C adjacent elements per matrix row (C columns in
// pass A, C values of k1 in pass B).  In F32 a run of 4 is 32 bytes, so each 64-byte L2 request
// carries 32 useful bytes, and a wider tile needs more LDS, which leaves fewer workgroups per CU.
// Here each thread computes two adjacent lines in lockstep (the engine on f2p / cp2): a tile of
// C lines has C / 2 line pairs, loads and stores two elements (16 bytes) per lane, and the same
// LDS bytes as the float kernel for the same tile.  The arithmetic per line is the float
// kernel's, so both give the same bits; the float kernels run the same tiles when an address or
// distance is not 16-byte aligned.
#include "kernels_impl.h"
#include "kernels_prec.h"

namespace cfft {

// Two adjacent ComplexF32 elements (16 bytes).
struct alignas(16) c32x2 {
    float r0, i0, r1, i1;
};

__device__ __forceinline__ cp2 load_pair(const float2* p) {
    const c32x2 g = *reinterpret_cast<const c32x2*>(p);
    return cp2{f2p(g.r0, g.r1), f2p(g.i0, g.i1)};
}
__device__ __forceinline__ void store_pair(float2* p, cp2 v) {
    *reinterpret_cast<c32x2*>(p) = c32x2{v.x.a, v.y.a, v.x.b, v.y.b};
}

// Pass A input: element e of columns j2 and j2 + 1 (p = in + j2), `stride` = N2 apart; sg = -1
// conjugates.
struct GPairStrideSrc {
    const float2* p;
    int64_t stride;
    float sg;
    __device__ __forceinline__ cp2 operator()(int e) const {
        cp2 v = load_pair(p + (int64_t)e * stride);
        v.y *= f2p(sg, sg);
        return v;
    }
};

// Pass A output: column j2's line at p, column j2 + 1's at p + N1 (S: streaming stores).
template <bool S>
struct GPairColsDst {
    float2* p;
    int N1;
    __device__ __forceinline__ void put(int e, cp2 v) const {
        store_line_elem<float, S>(p[e], float2{v.x.a, v.y.a});
        store_line_elem<float, S>(p[N1 + e], float2{v.x.b, v.y.b});
    }
};

// Pass B input: element e of lines k1 and k1 + 1 of the middle (p = middle + k1, `stride` = N1),
// then, once the pass's loads are issued, times the four-step twiddles w_n^{e k1} and
// w_n^{e (k1 + 1)} from the three-level table in LDS (t3 == nullptr: no twiddle); sg = -1
// conjugates.
struct GTwPairColSrc {
    const float2* p;
    int64_t stride;
    const float2* t3;
    int tbits, k1;
    float sg;
    __device__ __forceinline__ cp2 operator()(int e) const {
        return load_pair(p + (int64_t)e * stride);
    }
    __device__ __forceinline__ cp2 post(int e, cp2 v) const {
        if (t3) {
            const float2 a = cmul(float2{v.x.a, v.y.a}, tw3_get<float>(t3, tbits, e * k1));
            const float2 b = cmul(float2{v.x.b, v.y.b}, tw3_get<float>(t3, tbits, e * (k1 + 1)));
            v = cp2{f2p(a.x, b.x), f2p(a.y, b.y)};
        }
        v.y *= f2p(sg, sg);
        return v;
    }
};

// Pass B output: element e of lines k1 and k1 + 1 at p + e * stride (p = out + k1), scaled
// (re by sx, im by sy).
struct GPairStrideDst {
    float2* p;
    int64_t stride;
    float sx, sy;
    __device__ __forceinline__ void put(int e, cp2 v) const {
        store_pair(p + (int64_t)e * stride,
                   cp2{v.x * f2p(sx, sx), v.y * f2p(sy, sy)});
    }
};

// Pass A: columns j20 .. j20 + 2 Cp - 1 of one transform, as Cp pairs (threads pair first:
// lane = c + Cp t, so a wave's access to a row covers 2 Cp adjacent columns).
template <int E, bool S>
__device__ __forceinline__ void fs_a_pair_body(unsigned char* smem, const float2* src,
                                               float2* dst, int N1, int N2, int Cp, int Tt,
                                               int Ls, int j20, int sgn, const PassTable& pt,
                                               const float2* tw) {
    const int cshift = __ffs(Cp) - 1;
    const int tid = threadIdx.x;
    const int c = tid & (Cp - 1);
    const int t = tid >> cshift;
    const int j2 = j20 + 2 * c;
    const bool active = j2 < N2;  // N2 is even, so j2 + 1 < N2 as well
    const int j2c = active ? j2 : 0;
    const LdsLine<f2p> L = lds_line<f2p>(smem, Cp, Ls, c);
    GPairStrideSrc gs{src + j2c, N2, (float)sgn};
    GPairColsDst<S> d{dst + (int64_t)j2c * N1, N1};
    run_line_t<f2p, E>(pt, N1, t, Tt, active, TwGlobal<float2>{tw}, gs, d, L, true, true, false,
                       0, 1);
}

// Pass B: lines k10 .. k10 + 2 Cp - 1 (no convolution epilogue); in == out allowed, as in the
// float kernel (every element of the tile is read by the first pass before any write).
template <int E>
__device__ __forceinline__ void fs_b_pair_body(unsigned char* smem, const float2* src,
                                               float2* dst, int N1, int N2, int Cp, int Tt,
                                               int Ls, int k10, int sgn_in, int sgn,
                                               double scale, const float2* tw3, int tbits,
                                               int tnc, const PassTable& pt, const float2* tw) {
    const int cshift = __ffs(Cp) - 1;
    float2* t3 = reinterpret_cast<float2*>(smem + fsb_tw_offset(sizeof(f2p), Cp, Ls));
    if (tw3) stage_table<float>(tw3, t3, tw3_entries(tbits, tnc));
    __syncthreads();
    const int tid = threadIdx.x;
    const int c = tid & (Cp - 1);
    const int t = tid >> cshift;
    const bool active = k10 + 2 * c < N1;  // N1 is even
    const int k1 = active ? k10 + 2 * c : 0;
    const LdsLine<f2p> L = lds_line<f2p>(smem, Cp, Ls, c);
    GTwPairColSrc gs{src + k1, N1, tw3 ? t3 : nullptr, tbits, k1, (float)sgn_in};
    GPairStrideDst d{dst + k1, N1, (float)scale, (float)(scale * sgn)};
    run_line_t<f2p, E>(pt, N2, t, Tt, active, TwGlobal<float2>{tw}, gs, d, L, true, true,
                       pt.npass == 1, 0, 1);
}

// a.C = lines per tile (even), a.Ls = the LDS line stride of a line pair; S: the middle is
// written with streaming stores.
template <int E, bool S = true>
__global__ void __launch_bounds__(1024)
k_fs_a_pair(const float2* __restrict__ in, float2* __restrict__ out, FourStepAArgs a,
            PassTable pt, const float2* __restrict__ tw) {
    CFFT_DYN_LDS(smem);
    const int tiles = (a.N2 + a.C - 1) / a.C;
    const int bid = xcd_block(a.xcd);
    const int64_t bt = bid / tiles;
    const int tile = bid - (int)(bt * tiles);
    fs_a_pair_body<E, S>(smem, in + bt * a.idist, out + bt * a.odist, a.N1, a.N2, a.C / 2, a.Tt,
                      a.Ls, tile * a.C, a.sgn_in, pt, tw);
}

template <int E>
__global__ void __launch_bounds__(1024)
k_fs_b_pair(const float2* in, float2* out, FourStepBArgs a, PassTable pt,
            const float2* __restrict__ tw) {
    CFFT_DYN_LDS(smem);
    const int tiles = (a.N1 + a.C - 1) / a.C;
    const int bid = xcd_block(a.xcd);
    const int64_t bt = bid / tiles;
    const int tile = bid - (int)(bt * tiles);
    fs_b_pair_body<E>(smem, in + bt * a.idist, out + bt * a.odist, a.N1, a.N2, a.C / 2, a.Tt,
                      a.Ls, tile * a.C, a.sgn_in, a.sgn_out, a.scale, (const float2*)a.tw3,
                      a.tbits, a.tnc, pt, tw);
}

size_t fs_pair_lds(int pass, int C, int Ls, int tbits, int tnc, bool tw3) {
    const size_t lines = fsb_tw_offset(sizeof(f2p), C / 2, Ls);
    return pass && tw3 ? lines + (size_t)tw3_entries(tbits, tnc) * sizeof(float2) : lines;
}

template <int E>
static hipError_t fsa_pair_t(const void* in, void* out, const FourStepAArgs& a,
                             const PassTable& pt, const void* tw, hipStream_t st) {
    const size_t lds = fs_pair_lds(0, a.C, a.Ls, 0, 0, false);
    const int64_t blocks = (int64_t)((a.N2 + a.C - 1) / a.C) * a.batch;
    if (a.mid_plain)
        hipLaunchKernelGGL((k_fs_a_pair<E, false>), dim3((unsigned)blocks), dim3(a.C / 2 * a.Tt),
                           lds, st, (const float2*)in, (float2*)out, a, pt, (const float2*)tw);
    else
        hipLaunchKernelGGL((k_fs_a_pair<E>), dim3((unsigned)blocks), dim3(a.C / 2 * a.Tt), lds,
                           st, (const float2*)in, (float2*)out, a, pt, (const float2*)tw);
    return hipGetLastError();
}

template <int E>
static hipError_t fsb_pair_t(const void* in, void* out, const FourStepBArgs& a,
                             const PassTable& pt, const void* tw, hipStream_t st) {
    const size_t lds = fs_pair_lds(1, a.C, a.Ls, a.tbits, a.tnc, a.tw3 != nullptr);
    const int64_t blocks = (int64_t)((a.N1 + a.C - 1) / a.C) * a.batch;
    hipLaunchKernelGGL((k_fs_b_pair<E>), dim3((unsigned)blocks), dim3(a.C / 2 * a.Tt), lds, st,
                       (const float2*)in, (float2*)out, a, pt, (const float2*)tw);
    return hipGetLastError();
}

hipError_t fsa_pair_f32(const void* in, void* out, const FourStepAArgs& a, const PassTable& pt,
                        const void* tw, hipStream_t st) {
    return a.E <= 8 ? fsa_pair_t<8>(in, out, a, pt, tw, st) : fsa_pair_t<16>(in, out, a, pt, tw, st);
}
hipError_t fsb_pair_f32(const void* in, void* out, const FourStepBArgs& a, const PassTable& pt,
                        const void* tw, hipStream_t st) {
    return a.E <= 8 ? fsb_pair_t<8>(in, out, a, pt, tw, st) : fsb_pair_t<16>(in, out, a, pt, tw, st);
}

void init_kernels_pair_f32() {
    detail::allow_big_lds(k_fs_a_pair<8>);
    detail::allow_big_lds(k_fs_a_pair<16>);
    detail::allow_big_lds(k_fs_a_pair<8, false>);
    detail::allow_big_lds(k_fs_a_pair<16, false>);
    detail::allow_big_lds(k_fs_b_pair<8>);
    detail::allow_big_lds(k_fs_b_pair<16>);
}

}  // namespace cfft
