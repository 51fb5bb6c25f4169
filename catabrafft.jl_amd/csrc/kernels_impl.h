// kernels_impl.h -- the HIP kernels of the library and their templated launchers (gfx950).
// Instantiated per precision by kernels_f32.hip and kernels_f64.hip.
//
//   k_rows         batched contiguous line FFTs, one HBM read + one HBM write per element
//                  (radix_exec.jl's pass sequence for a whole transform, in one launch)
//   k_fs_a/k_fs_b  the two passes of the four-step n = N1*N2 (mixed_radix.jl:44-78), twiddle
//                  w_n^{j2 k1} fused into pass A's store, transpose fused into pass B's store
//   k_conv_small   Rader (prime.jl:5-82) or Bluestein fused in LDS
//   k_conv_gather / k_conv_scatter
//                  the permutation / chirp ends of a large convolution whose FFTs run on the
//                  kernels above
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace cfft {

template <typename T> using V_t = typename Vec2<T>::type;

// Global-memory line input; sg = -1 conjugates.
template <typename T>
struct GSrc {
    const V_t<T>* p;
    T sg;
    __device__ __forceinline__ V_t<T> operator()(int e) const {
        V_t<T> v = p[e];
        v.y *= sg;
        return v;
    }
};

// Global-memory line output with scale, conjugation and the optional convolution epilogue.
template <typename T>
struct GRowDst {
    V_t<T>* p;
    T sx, sy;
    const V_t<T>* K;
    V_t<T>* dc;
    __device__ __forceinline__ void put(int e, V_t<T> v) const {
        if (K) {
            if (dc && e == 0) *dc = v;
            v = cconj(cmul(v, K[e]));
        }
        p[e] = V_t<T>{v.x * sx, v.y * sy};
    }
};

template <typename T>
__device__ __forceinline__ LdsLine<T> lds_line(unsigned char* smem, int C, int Ls, int c) {
    T* re = reinterpret_cast<T*>(smem);
    return LdsLine<T>{re + c * Ls, re + C * Ls + c * Ls};
}

// ------------------------------------------------------------------------------------------
// k_rows
// ------------------------------------------------------------------------------------------
template <typename T, int E>
__global__ void __launch_bounds__(1024)
k_rows(const V_t<T>* __restrict__ in, V_t<T>* __restrict__ out, RowsArgs a, PassTable pt,
       const V_t<T>* __restrict__ tw) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x;
    const int c = tid / a.Tt;
    const int t = tid - c * a.Tt;
    const int64_t line = (int64_t)blockIdx.x * a.C + c;
    const bool active = (c < a.C) && (line < a.nlines);
    const int64_t li = active ? line : 0;
    GSrc<T> gs{in + li * a.idist, (T)a.sgn_in};
    GRowDst<T> gd{out + li * a.odist, (T)a.scale, (T)(a.scale * a.sgn_out),
                  (const V_t<T>*)a.epi.K, a.epi.dc ? (V_t<T>*)a.epi.dc + li : nullptr};
    const LdsLine<T> L = lds_line<T>(smem, a.C, a.Ls, c < a.C ? c : 0);
    run_line<T, E>(pt, a.N, t, a.Tt, active, tw, gs, gd, L, true, true);
}

// ------------------------------------------------------------------------------------------
// Four-step pass A.  C is a power of two (cshift = log2 C).
// ------------------------------------------------------------------------------------------
template <typename T>
struct TwiddleDst {
    V_t<T>* p;              // this column's output row: p[k1]
    const V_t<T>* T1;
    const V_t<T>* T2;
    int j2, shift, mask;
    __device__ __forceinline__ void put(int k1, V_t<T> v) const {
        const int e = j2 * k1;
        const V_t<T> w = cmul(T1[e & mask], T2[e >> shift]);
        p[k1] = cmul(v, w);
    }
};

template <typename T, int E>
__global__ void __launch_bounds__(1024)
k_fs_a(const V_t<T>* __restrict__ in, V_t<T>* __restrict__ out, FourStepAArgs a, PassTable pt,
       const V_t<T>* __restrict__ tw) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int tiles = (a.N2 + a.C - 1) / a.C;
    const int64_t bt = blockIdx.x / tiles;
    const int tile = blockIdx.x - (int)(bt * tiles);
    const int j20 = tile * a.C;
    const V_t<T>* src = in + bt * a.idist;
    const int cshift = __ffs(a.C) - 1;
    const T sg = (T)a.sgn_in;
    // Staged load: rows j1 of C contiguous columns (one run of C elements per row).
    const int total = a.C * a.N1;
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        const int c = idx & (a.C - 1);
        const int e = idx >> cshift;
        const int j2 = j20 + c;
        V_t<T> v = V_t<T>{0, 0};
        if (j2 < a.N2) v = src[(int64_t)e * a.N2 + j2];
        v.y *= sg;
        lds_line<T>(smem, a.C, a.Ls, c).put(e, v);
    }
    __syncthreads();
    const int tid = threadIdx.x;
    const int c = tid / a.Tt;
    const int t = tid - c * a.Tt;
    const int j2 = j20 + c;
    const bool active = (c < a.C) && (j2 < a.N2);
    const LdsLine<T> L = lds_line<T>(smem, a.C, a.Ls, c < a.C ? c : 0);
    TwiddleDst<T> dst{out + bt * a.odist + (int64_t)(active ? j2 : 0) * a.N1,
                      (const V_t<T>*)a.T1, (const V_t<T>*)a.T2, j2, a.shift, a.mask};
    run_line<T, E>(pt, a.N1, t, a.Tt, active, tw, L, dst, L, false, true);
}

// ------------------------------------------------------------------------------------------
// Four-step pass B.  Staged load and staged (transposing) store; in == out allowed.
// ------------------------------------------------------------------------------------------
template <typename T, int E>
__global__ void __launch_bounds__(1024)
k_fs_b(const V_t<T>* in, V_t<T>* out, FourStepBArgs a, PassTable pt,
       const V_t<T>* __restrict__ tw) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int tiles = (a.N1 + a.C - 1) / a.C;
    const int64_t bt = blockIdx.x / tiles;
    const int tile = blockIdx.x - (int)(bt * tiles);
    const int k10 = tile * a.C;
    const V_t<T>* src = in + bt * a.idist;
    V_t<T>* dst = out + bt * a.odist;
    const int cshift = __ffs(a.C) - 1;
    const int total = a.C * a.N2;
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        const int c = idx & (a.C - 1);
        const int e = idx >> cshift;
        const int k1 = k10 + c;
        V_t<T> v = V_t<T>{0, 0};
        if (k1 < a.N1) v = src[(int64_t)e * a.N1 + k1];
        lds_line<T>(smem, a.C, a.Ls, c).put(e, v);
    }
    __syncthreads();
    const int tid = threadIdx.x;
    const int c = tid / a.Tt;
    const int t = tid - c * a.Tt;
    const bool active = (c < a.C) && (k10 + c < a.N1);
    const LdsLine<T> L = lds_line<T>(smem, a.C, a.Ls, c < a.C ? c : 0);
    run_line<T, E>(pt, a.N2, t, a.Tt, active, tw, L, L, L, false, false);
    // run_line ended with a barrier: the lines' outputs are in LDS.
    const T sx = (T)a.scale, sy = (T)(a.scale * a.sgn_out);
    const V_t<T>* K = (const V_t<T>*)a.epi.K;
    V_t<T>* dc = (V_t<T>*)a.epi.dc;
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        const int cc = idx & (a.C - 1);
        const int e = idx >> cshift;
        const int k1 = k10 + cc;
        if (k1 < a.N1) {
            V_t<T> v = lds_line<T>(smem, a.C, a.Ls, cc)(e);
            const int64_t o = (int64_t)k1 + (int64_t)a.N1 * e;
            if (K) {
                if (dc && o == 0) dc[bt] = v;
                v = cconj(cmul(v, K[o]));
            }
            dst[o] = V_t<T>{v.x * sx, v.y * sy};
        }
    }
}

// ------------------------------------------------------------------------------------------
// k_conv_small: Rader / Bluestein, whole transform in LDS.
// LDS: X (input, then output; C lines of LsX) | A (convolution work, C lines of LsA) | dc[C]
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

template <typename T, int E>
__global__ void __launch_bounds__(1024)
k_conv_small(const V_t<T>* __restrict__ in, V_t<T>* __restrict__ out, ConvSmallArgs a,
             PassTable pt, const V_t<T>* __restrict__ tw) {
    extern __shared__ __align__(16) unsigned char smem[];
    T* base = reinterpret_cast<T*>(smem);
    T* xre = base;
    T* xim = xre + a.C * a.LsX;
    T* are = xim + a.C * a.LsX;
    T* aim = are + a.C * a.LsA;
    V_t<T>* sdc = reinterpret_cast<V_t<T>*>(aim + a.C * a.LsA);
    const T sg = (T)a.sgn_in;
    // 1. staged, coalesced load of the C input lines (flat over lines x elements)
    const int64_t line0 = (int64_t)blockIdx.x * a.C;
    const int nl = (int)((a.nlines - line0) < a.C ? (a.nlines - line0) : a.C);
    for (int idx = threadIdx.x; idx < nl * a.n; idx += blockDim.x) {
        const int c = idx / a.n;
        const int e = idx - c * a.n;
        V_t<T> v = in[(line0 + c) * a.idist + e];
        v.y *= sg;
        LdsLine<T>{xre + c * a.LsX, xim + c * a.LsX}.put(e, v);
    }
    __syncthreads();
    const int tid = threadIdx.x;
    const int c = tid / a.Tt;
    const int t = tid - c * a.Tt;
    const int64_t line = (int64_t)blockIdx.x * a.C + c;
    const bool active = (c < a.C) && (line < a.nlines);
    const int cc = c < a.C ? c : 0;
    const LdsLine<T> X{xre + cc * a.LsX, xim + cc * a.LsX};
    const LdsLine<T> A{are + cc * a.LsA, aim + cc * a.LsA};
    const V_t<T>* chirp = (const V_t<T>*)a.chirp;
    // 2. forward FFT_L of the gathered / chirped sequence; pointwise product with K on the way out
    GatherSrc<T> gsrc{X, a.perm, chirp, a.n, a.mode};
    ConvEpiDst<T> edst{A, (const V_t<T>*)a.K, sdc + cc};
    run_line<T, E>(pt, a.L, t, a.Tt, active, tw, gsrc, edst, A, true, true, true);
    __syncthreads();
    // 3. second forward FFT_L of the conjugated product = conj of the inverse; scatter into X
    const V_t<T> x0 = X(0);
    ScatterDst<T> sdst{X, a.iperm, chirp, x0, (T)(1.0 / a.L), a.n, a.mode};
    run_line<T, E>(pt, a.L, t, a.Tt, active, tw, A, sdst, A, false, true);
    __syncthreads();
    if (a.mode == (int)ConvMode::Rader && active && t == 0) {
        const V_t<T> d = sdc[cc];
        X.put(0, V_t<T>{x0.x + d.x, x0.y + d.y});
    }
    __syncthreads();
    // 4. staged, coalesced store
    const T sx = (T)a.scale, sy = (T)(a.scale * a.sgn_out);
    for (int idx = threadIdx.x; idx < nl * a.n; idx += blockDim.x) {
        const int c2 = idx / a.n;
        const int e = idx - c2 * a.n;
        const V_t<T> v = LdsLine<T>{xre + c2 * a.LsX, xim + c2 * a.LsX}(e);
        out[(line0 + c2) * a.odist + e] = V_t<T>{v.x * sx, v.y * sy};
    }
}

// ------------------------------------------------------------------------------------------
// Large convolution gather / scatter (grid: x over elements, y over lines).
// ------------------------------------------------------------------------------------------
template <typename Ti, typename Tc>
__global__ void __launch_bounds__(256)
k_conv_gather(const V_t<Ti>* __restrict__ x, V_t<Tc>* __restrict__ w, ConvGatherArgs a) {
  for (int64_t line = blockIdx.y; line < a.nlines; line += gridDim.y) {
    const V_t<Ti>* xs = x + line * a.idist;
    V_t<Tc>* ws = w + line * a.wdist;
    const V_t<Tc>* chirp = (const V_t<Tc>*)a.chirp;
    const Tc sg = (Tc)a.sgn_in;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < a.L;
         e += (int64_t)gridDim.x * blockDim.x) {
        V_t<Tc> v = V_t<Tc>{0, 0};
        if (a.mode == (int)ConvMode::Rader) {
            if (e < a.n - 1) {
                const V_t<Ti> u = xs[a.perm[e]];
                v = V_t<Tc>{(Tc)u.x, (Tc)u.y * sg};
            }
        } else if (e < a.n) {
            const V_t<Ti> u = xs[e];
            v = cmul(V_t<Tc>{(Tc)u.x, (Tc)u.y * sg}, chirp[e]);
        }
        ws[e] = v;
    }
  }
}

template <typename Ti, typename Tc>
__global__ void __launch_bounds__(256)
k_conv_scatter(const V_t<Tc>* __restrict__ w, V_t<Ti>* __restrict__ y, ConvScatterArgs a) {
  for (int64_t line = blockIdx.y; line < a.nlines; line += gridDim.y) {
    const V_t<Tc>* ws = w + line * a.wdist;
    V_t<Ti>* ys = y + line * a.odist;
    const V_t<Tc>* chirp = (const V_t<Tc>*)a.chirp;
    const Tc invL = (Tc)a.invL;
    const Tc sx = (Tc)a.scale, sy = (Tc)(a.scale * a.sgn_out);
    if (a.mode == (int)ConvMode::Rader) {
        const V_t<Ti> u0 = ((const V_t<Ti>*)a.x)[line * a.idist];
        const V_t<Tc> x0 = V_t<Tc>{(Tc)u0.x, (Tc)u0.y * (Tc)a.sgn_in};
        for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < a.n - 1;
             k += (int64_t)gridDim.x * blockDim.x) {
            const V_t<Tc> v = ws[k];
            const V_t<Tc> r = V_t<Tc>{x0.x + v.x * invL, x0.y - v.y * invL};
            ys[a.iperm[k]] = V_t<Ti>{(Ti)(r.x * sx), (Ti)(r.y * sy)};
            if (k == 0) {
                const V_t<Tc> d = ((const V_t<Tc>*)a.dc)[line];
                ys[0] = V_t<Ti>{(Ti)((x0.x + d.x) * sx), (Ti)((x0.y + d.y) * sy)};
            }
        }
    } else {
        for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < a.n;
             k += (int64_t)gridDim.x * blockDim.x) {
            const V_t<Tc> v = ws[k];
            const V_t<Tc> r = cmul(V_t<Tc>{v.x * invL, -v.y * invL}, chirp[k]);
            ys[k] = V_t<Ti>{(Ti)(r.x * sx), (Ti)(r.y * sy)};
        }
    }
  }
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

template <typename T, int E>
hipError_t rows_t(const void* in, void* out, const RowsArgs& a, const PassTable& pt,
                  const void* tw, hipStream_t st) {
    static bool once = (allow_big_lds(k_rows<T, E>), true);
    (void)once;
    const size_t lds = 2 * sizeof(T) * (size_t)a.C * a.Ls;
    const int64_t blocks = (a.nlines + a.C - 1) / a.C;
    hipLaunchKernelGGL((k_rows<T, E>), dim3((unsigned)blocks), dim3(a.C * a.Tt), lds, st,
                       (const V_t<T>*)in, (V_t<T>*)out, a, pt, (const V_t<T>*)tw);
    return hipGetLastError();
}

template <typename T, int E>
hipError_t fsa_t(const void* in, void* out, const FourStepAArgs& a, const PassTable& pt,
                 const void* tw, hipStream_t st) {
    static bool once = (allow_big_lds(k_fs_a<T, E>), true);
    (void)once;
    const size_t lds = 2 * sizeof(T) * (size_t)a.C * a.Ls;
    const int64_t blocks = (int64_t)((a.N2 + a.C - 1) / a.C) * a.batch;
    hipLaunchKernelGGL((k_fs_a<T, E>), dim3((unsigned)blocks), dim3(a.C * a.Tt), lds, st,
                       (const V_t<T>*)in, (V_t<T>*)out, a, pt, (const V_t<T>*)tw);
    return hipGetLastError();
}

template <typename T, int E>
hipError_t fsb_t(const void* in, void* out, const FourStepBArgs& a, const PassTable& pt,
                 const void* tw, hipStream_t st) {
    static bool once = (allow_big_lds(k_fs_b<T, E>), true);
    (void)once;
    const size_t lds = 2 * sizeof(T) * (size_t)a.C * a.Ls;
    const int64_t blocks = (int64_t)((a.N1 + a.C - 1) / a.C) * a.batch;
    hipLaunchKernelGGL((k_fs_b<T, E>), dim3((unsigned)blocks), dim3(a.C * a.Tt), lds, st,
                       (const V_t<T>*)in, (V_t<T>*)out, a, pt, (const V_t<T>*)tw);
    return hipGetLastError();
}

template <typename T, int E>
hipError_t conv_t(const void* in, void* out, const ConvSmallArgs& a, const PassTable& pt,
                  const void* tw, hipStream_t st) {
    static bool once = (allow_big_lds(k_conv_small<T, E>), true);
    (void)once;
    const size_t lds =
        2 * sizeof(T) * (size_t)a.C * (a.LsX + a.LsA) + sizeof(V_t<T>) * (size_t)a.C;
    const int64_t blocks = (a.nlines + a.C - 1) / a.C;
    hipLaunchKernelGGL((k_conv_small<T, E>), dim3((unsigned)blocks), dim3(a.C * a.Tt), lds, st,
                       (const V_t<T>*)in, (V_t<T>*)out, a, pt, (const V_t<T>*)tw);
    return hipGetLastError();
}

}  // namespace detail

namespace detail {
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
}  // namespace detail

}  // namespace cfft
