// kernels.h -- host-visible launch interface of the HIP kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "engine.h"

namespace cfft {

enum class Prec : int { F32 = 0, F64 = 1 };

// Epilogue applied by the last store of a line FFT that is the forward half of a convolution:
// v -> conj(v * K[e]) and slot[0] = v when e == 0 (Rader needs sum(x) = x0 + A^[0]).
struct ConvEpilogue {
    const void* K;   // kernel spectrum, length = line length, same precision as the line
    void* dc;        // two complex slots per line {A^[0], x0}, or nullptr
};

// Batched contiguous lines: line b reads in[b*idist + e], writes out[b*odist + e].
struct RowsArgs {
    int N, C, Tt, Ls, E;
    int64_t nlines, idist, odist;
    int sgn_in, sgn_out;   // -1 conjugates on load / store (backward = conj(fft(conj x)))
    double scale;          // applied on store
    ConvEpilogue epi;      // epi.K == nullptr: none
    int grid;              // pipelined kernel: workgroups launched (each loops over line groups)
    int nopad;             // plain rows kernel: LDS lines without pad words (Ls from N alone)
};

// Four-step pass A for n = N1*N2 (lines = the N2 columns of the N1 x N2 row-major input):
//   out[b*odist + j2*N1 + k1] = DFT_N1(in[b*idist + j1*N2 + j2])[k1]
// (the twiddle w_n^{j2 k1} between the two passes is applied by pass B as it loads).
struct FourStepAArgs {
    int N1, N2, C, Tt, Ls, E;
    int64_t batch, idist, odist;
    int sgn_in;
    int grid;              // pipelined kernel: workgroups launched (each loops over tiles)
    int xcd;               // nonzero: consecutive tiles on one XCD
    // nonzero: the middle is written with ordinary stores (kept in the caches for pass B, as in
    // chunks whose middle fits the Infinity Cache); zero: streaming stores
    int mid_plain;
};

// Four-step pass B (lines = the N1 values of k1, elements j2 at stride N1):
//   out[b*odist + k1 + N1*k2] = DFT_N2(w_n^{j2 k1} in[b*idist + j2*N1 + k1])[k2]
// with w_n^e = (A[e & m] * B[(e >> tbits) & m]) * C[e >> 2 tbits], the tables back to back in
// tw3 (tw3 == nullptr: no twiddle, a plain strided-line pass), copied into LDS after the lines.
// Reads and writes the same index set per tile, so in == out is allowed.
struct FourStepBArgs {
    int N1, N2, C, Tt, Ls, E;
    int64_t batch, idist, odist;
    int sgn_in;            // -1 conjugates on load (a strided-axis step of an N-D transform)
    int sgn_out;
    double scale;
    ConvEpilogue epi;
    int tbits, tnc;
    const void* tw3;
    int grid;              // pipelined kernel: workgroups launched (each loops over tiles)
    int xcd;               // nonzero: consecutive tiles on one XCD
};

// Entries of a three-level twiddle table, and the LDS bytes of a pass-B workgroup: C lines of
// stride Ls, then the table (16-byte aligned).
__host__ __device__ constexpr int tw3_entries(int tbits, int tnc) { return (2 << tbits) + tnc; }
__host__ __device__ constexpr size_t fsb_tw_offset(size_t elem, int C, int Ls) {
    return (2 * elem * (size_t)C * Ls + 15) & ~(size_t)15;
}

// Fused small convolution FFT (Rader or Bluestein) entirely in LDS: one launch, one HBM read
// and one HBM write per transform.
enum class ConvMode : int { Rader = 0, Bluestein = 1 };
struct ConvSmallArgs {
    int n, L, C, Tt, Ls, E;  // Ls: line stride of the LDS region the input and the work share
    int mode;              // ConvMode
    int64_t nlines, idist, odist;
    int sgn_in, sgn_out;
    double scale;
    const int* perm;       // Rader: g^i mod n, i < n-1
    const int* iperm;      // Rader: (g^-1)^i mod n, i < n-1
    const void* chirp;     // Bluestein: e^{-pi i m^2/n}, m < n (line precision)
    const void* K;         // FFT_L of the convolution kernel (line precision)
};

// Large convolution: gather/pre-multiply into the (zero-padded) length-L work buffer, and the
// final scatter/post-multiply.  Tio = precision of x / y, Tc = precision of the work buffer.
struct ConvGatherArgs {
    int64_t n, L, nlines, idist, wdist;
    int mode, sgn_in;
    const int* perm;
    const void* chirp;     // Tc precision
    void* dc;              // per line slots {A^[0], x0} (Tc precision); the gather writes x0
};
struct ConvScatterArgs {
    int64_t n, L, nlines, odist, idist, wdist;
    int mode, sgn_in, sgn_out;
    double invL;           // 1/L of the inverse transform
    double scale;          // output scale (1/n for ifft)
    const int* iperm;
    const void* chirp;     // Tc precision
    const void* dc;        // per line slots {A^[0], x0} (Tc precision)
};

// Large convolution (Rader / Bluestein whose length-L FFTs run as a four-step L = N1 x N2) in
// three launches instead of gather, two four-steps and scatter.  The second FFT uses the
// transposed split (N2 x N1), so the lines of the first FFT's pass B are exactly the columns of
// the second FFT's pass A and one workgroup does both:
//   A  (k_convl_a):  the gathered (permuted / chirped, zero-padded) sequence read straight from
//                    x as N2 columns of N1 points, FFT_N1 each -> m1[j2 N1 + k1];
//   BA (k_convl_ba): lines k1 < N1: w_L^{j2 k1} m1[j2 N1 + k1], FFT_N2 -> X1[k1 + N1 k2], then
//                    conj(X1 K) (and the slot A^[0] = X1[0]), FFT_N2 again -> m2[k1 N2 + k1'];
//   C  (k_convl_c):  lines k1' < N2: w_L^{j2 k1'} m2[j2 N2 + k1'], FFT_N1 -> X2[k1' + N2 k2],
//                    scattered into y as Rader (y[iperm[i]] = x0 + conj(X2[i]) / L) or Bluestein
//                    (y[i] = chirp[i] conj(X2[i]) / L), scaled and conjugated for backward.
// Ti = precision of x and y, Tc = precision of the convolution (and of m1, m2, K, chirp, dc).
struct ConvLArgs {
    int64_t n;             // transform length
    int N1, N2;            // L = N1 * N2
    int CA, CB, CC;        // lines (columns) per workgroup of the three kernels
    int TtA, LsA, TtB, LsB, TtC, LsC;  // A and C run line a (N1 points), BA line b (N2)
    int E;
    int64_t nlines, idist, odist;  // rows of x and y
    int mode, sgn;         // ConvMode; sgn = -1 conjugates the input and the output (backward)
    double scale;          // output scale (1/n for ifft)
    const int* perm;       // Rader: g^i mod n
    const int* iperm;      // Rader: (g^-1)^i mod n
    const void* chirp;     // Bluestein chirp (Tc)
    const void* K;         // FFT_L of the convolution kernel (Tc)
    void* m1;              // nlines x L work (Tc)
    void* m2;              // nlines x L work (Tc)
    void* dc;              // per line slots {A^[0], x0} (Tc)
    const void* tw3;       // w_L^e, three-level table (Tc)
    int tbits, tnc;
};

// Real-input transforms (n = 2m even), fused into one rows launch: forward reads `rdist`-spaced
// rows of n reals and writes `cdist`-spaced rows of m+1 complex; backward the reverse.
struct RealArgs {
    int m, C, Tt, Ls;
    int64_t nlines, rdist, cdist;  // rdist in reals (even), cdist in complex elements
    int forward;
    double scale;
    const void* w;                 // w[k] = e^{-2 pi i k / n}, k < m
};

// Elementwise steps of the general real path (see k_real_aux): mode 0 r2c post, 1 c2r pre,
// 2 reals -> complex, 3 half of a full spectrum, 4 half -> full Hermitian, 5 real part.
struct RealAuxArgs {
    int mode;
    int64_t n, m, nlines, count, idist, odist;
    double scale;
    const void* w;
};

// Sets per-kernel attributes (large LDS) once; called when a context is created.
void init_kernels();

hipError_t launch_rows(Prec p, const void* in, void* out, const RowsArgs& a, const PassTable& pt,
                       const void* tw, hipStream_t st);
// ComplexF16 storage, float arithmetic: the rows and four-step kernels with half-precision
// global lines (the four-step middle stays float), and conversions of rows x n elements between
// half and float storage (row r at r * sdist in the source, r * ddist in the destination).
hipError_t launch_rows_f16(const void* in, void* out, const RowsArgs& a, const PassTable& pt,
                           const void* tw, hipStream_t st);
hipError_t launch_fourstep_a_f16(const void* in, void* out, const FourStepAArgs& a,
                                 const PassTable& pt, const void* tw, hipStream_t st);
hipError_t launch_fourstep_b_f16(const void* in, void* out, const FourStepBArgs& a,
                                 const PassTable& pt, const void* tw, hipStream_t st);
hipError_t launch_convert_h2f(const void* src, void* dst, int64_t rows, int64_t n, int64_t sdist,
                              int64_t ddist, hipStream_t st);
hipError_t launch_convert_f2h(const void* src, void* dst, int64_t rows, int64_t n, int64_t sdist,
                              int64_t ddist, hipStream_t st);
// Persistent pipelined rows kernel (needs npass >= 2, no epilogue); a.grid workgroups.  twn =
// entries of the twiddle table, which the kernel copies into LDS after the work lines.
hipError_t launch_rows_pipe(Prec p, const void* in, void* out, const RowsArgs& a,
                            const PassTable& pt, const void* tw, size_t twn, hipStream_t st);
// LDS bytes of a pipelined rows launch and its resident workgroups per CU (0: cannot launch).
// Its workgroups have at most kPipeMaxThreads threads.
constexpr int kPipeMaxThreads = 512;
size_t rows_pipe_lds(Prec p, int C, int Ls, size_t twn);
int rows_pipe_blocks_per_cu(Prec p, int E, int R0, int block, size_t lds);
hipError_t launch_fourstep_a(Prec p, const void* in, void* out, const FourStepAArgs& a,
                             const PassTable& pt, const void* tw, hipStream_t st);
// Transpose of a rows x cols matrix: dst[c rows + r] = src[r cols + c], times w_n^{r c} from the
// three-level table tw3 (nullptr: no twiddle); then, at output index o, the convolution
// epilogue when K is set (dc = the value at o = 0, then conj(v K[o])), then (re sx, im sy).
// The steps between the passes of transforms too long for the four-step (the six-step path),
// and of N-D axes that one workgroup cannot take as strided lines.  batch > 1: that many
// matrices back to back (no twiddle or epilogue then).
hipError_t launch_transpose(Prec p, const void* src, void* dst, int64_t rows, int64_t cols,
                            int64_t batch, const void* tw3, int tbits, double sx, double sy,
                            const void* K, void* dc, hipStream_t st);
// Persistent pipelined four-step passes (a.grid workgroups of at most kPipeMaxThreads, lines of
// at least two passes, no epilogue); twn = entries of the line's twiddle table, copied into LDS.
// fs_pipe_lds gives the LDS bytes (pass 1 = B: lines, pass twiddles, then the three-level table
// of t3n entries); fs_pipe_blocks_per_cu the resident workgroups per CU for a first radix R0.
hipError_t launch_fourstep_a_pipe(Prec p, const void* in, void* out, const FourStepAArgs& a,
                                  const PassTable& pt, const void* tw, size_t twn,
                                  hipStream_t st);
hipError_t launch_fourstep_b_pipe(Prec p, const void* in, void* out, const FourStepBArgs& a,
                                  const PassTable& pt, const void* tw, size_t twn,
                                  hipStream_t st);
size_t fs_pipe_lds(Prec p, int C, int Ls, size_t twn, size_t t3n);
int fs_pipe_blocks_per_cu(Prec p, int E, int pass, int R0, int block, size_t lds);
hipError_t launch_fourstep_b(Prec p, const void* in, void* out, const FourStepBArgs& a,
                             const PassTable& pt, const void* tw, hipStream_t st);
// ComplexF32 four-step passes with two adjacent lines per thread (kernels_pair_f32.hip): a.C
// lines per tile (even), a.Ls the LDS stride of a line pair; in, out, the distances and the
// line length across the pairs (N2 for pass A, N1 for pass B) must keep every pair 16-byte
// aligned.  fs_pair_lds gives the LDS bytes (pass 1 = B, with the three-level table).
hipError_t launch_fourstep_a_pair(const void* in, void* out, const FourStepAArgs& a,
                                  const PassTable& pt, const void* tw, hipStream_t st);
hipError_t launch_fourstep_b_pair(const void* in, void* out, const FourStepBArgs& a,
                                  const PassTable& pt, const void* tw, hipStream_t st);
size_t fs_pair_lds(int pass, int C, int Ls, int tbits, int tnc, bool tw3);
hipError_t launch_conv_small(Prec p, const void* in, void* out, const ConvSmallArgs& a,
                             const PassTable& pt, const void* tw, hipStream_t st);
hipError_t launch_real_rows(Prec p, const void* in, void* out, const RealArgs& a,
                            const PassTable& pt, const void* tw, int E, hipStream_t st);
hipError_t launch_real_aux(Prec p, const void* in, void* out, const RealAuxArgs& a,
                           hipStream_t st);
hipError_t launch_conv_gather(Prec pio, Prec pc, const void* x, void* w, const ConvGatherArgs& a,
                              hipStream_t st);
hipError_t launch_conv_scatter(Prec pio, Prec pc, const void* w, void* y,
                               const ConvScatterArgs& a, hipStream_t st);
// The three launches of a large convolution over a four-step (ConvLArgs); ptA / ptB are the pass
// tables of the N1- and N2-point lines, twA / twB their twiddles.
hipError_t launch_convl(Prec pio, Prec pc, const void* x, void* y, const ConvLArgs& a,
                        const PassTable& ptA, const void* twA, const PassTable& ptB,
                        const void* twB, hipStream_t st);

// Largest per-thread complex capacity compiled (template E values).
constexpr int kEValues[] = {8, 16};

}  // namespace cfft
