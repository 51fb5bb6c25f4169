// library.cpp -- context, planner, plan cache, execution and the C ABI (include/catabrafft.h).
//
// Strategy per length n (the role of kernel.jl:72-92's size classification):
//   n == 1 ............................ copy
//   n smooth (factors 2,3,5,7,11,13), one line fits a workgroup's LDS
//                                       DIRECT: one k_rows launch (all passes in LDS)
//   n smooth, larger .................. FOURSTEP n = N1*N2 (mixed_radix.jl:44-78): k_fs_a, k_fs_b
//   n prime ........................... Rader (prime.jl:5-82): cyclic convolution of length
//                                       n-1, zero-padded to a smooth L >= 2n-3 when n-1 is not
//                                       smooth; fused in one k_conv_small launch when it fits LDS,
//                                       else gather -> FFT_L -> FFT_L -> scatter
//   other n ........................... Bluestein with a smooth L >= 2n-1, same two shapes
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

#include "catabrafft.h"
#include "hostmath.h"
#include "kernels.h"
#include "library.h"

namespace cfft {

// ------------------------------------------------------------------------------- errors
static thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define CFFT_HIP(expr)                                                                   \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(CFFT_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// ------------------------------------------------------------------------------- no-alloc work area
// The work area of cfft_exec_noalloc: a file-scope device array of kNoAllocScratchBytes, one copy
// per GPU (each at its own device address), held until the process ends.  Its host-side name is
// not a device address, so host code never passes `scratch` to a kernel or a copy: ctx_create
// looks up the current device's address once (hipGetSymbolAddress) and every use goes through
// that pointer.  Calls of cfft_exec_noalloc share their GPU's array.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
bool g_scratch_looked_up[kMaxDevices] = {};
void* g_scratch_ptr[kMaxDevices] = {};
std::string g_scratch_err[kMaxDevices];
}  // namespace

// Looks up the work area's address on `device` (the current device) once per device.
static void noalloc_scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (g_scratch_looked_up[device]) return;
    g_scratch_looked_up[device] = true;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        (void)hipGetLastError();
        g_scratch_err[device] = std::string("hipGetSymbolAddress(scratch) failed: ") +
                                hipGetErrorString(e);
        return;
    }
    g_scratch_ptr[device] = ptr;
}

// The work area's device address on `device` (nullptr, with the reason in *err, if unavailable).
static void* noalloc_scratch_ptr(int device, std::string* err) {
    if (device < 0 || device >= kMaxDevices) {
        *err = "device index beyond the work-area table";
        return nullptr;
    }
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[device])
        *err = g_scratch_looked_up[device] ? g_scratch_err[device]
                                            : std::string("no context was opened on this device");
    return g_scratch_ptr[device];
}

// ------------------------------------------------------------------------------- device buffers
DevBuf::~DevBuf() { reset(); }
void DevBuf::reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}
int DevBuf::alloc(size_t nbytes) {
    reset();
    if (nbytes == 0) return CFFT_OK;
    if (hipMalloc(&p, nbytes) != hipSuccess) {
        p = nullptr;
        (void)hipGetLastError();
        return fail(CFFT_ERR_NOMEM, "device allocation of " + std::to_string(nbytes) + " bytes failed");
    }
    bytes = nbytes;
    return CFFT_OK;
}

// Uploads long double complex values rounded once to the plan precision.
static int upload(DevBuf& buf, const std::vector<cld>& v, Prec pr) {
    if (v.empty()) return CFFT_OK;
    const size_t esz = pr == Prec::F32 ? 8 : 16;
    int rc = buf.alloc(v.size() * esz);
    if (rc) return rc;
    if (pr == Prec::F32) {
        std::vector<float> h(2 * v.size());
        for (size_t i = 0; i < v.size(); ++i) {
            h[2 * i] = (float)v[i].real();
            h[2 * i + 1] = (float)v[i].imag();
        }
        CFFT_HIP(hipMemcpy(buf.p, h.data(), buf.bytes, hipMemcpyHostToDevice));
    } else {
        std::vector<double> h(2 * v.size());
        for (size_t i = 0; i < v.size(); ++i) {
            h[2 * i] = (double)v[i].real();
            h[2 * i + 1] = (double)v[i].imag();
        }
        CFFT_HIP(hipMemcpy(buf.p, h.data(), buf.bytes, hipMemcpyHostToDevice));
    }
    return CFFT_OK;
}

int upload_cld(DevBuf& buf, const std::vector<cld>& v, Prec pr) { return upload(buf, v, pr); }
int upload_ints(DevBuf& buf, const std::vector<int>& v);
int upload_ints(DevBuf& buf, const std::vector<int>& v) {
    if (v.empty()) return CFFT_OK;
    int rc = buf.alloc(v.size() * sizeof(int));
    if (rc) return rc;
    CFFT_HIP(hipMemcpy(buf.p, v.data(), buf.bytes, hipMemcpyHostToDevice));
    return CFFT_OK;
}

static int upload_int(DevBuf& buf, const std::vector<int>& v) {
    if (v.empty()) return CFFT_OK;
    int rc = buf.alloc(v.size() * sizeof(int));
    if (rc) return rc;
    CFFT_HIP(hipMemcpy(buf.p, v.data(), buf.bytes, hipMemcpyHostToDevice));
    return CFFT_OK;
}

// ------------------------------------------------------------------------------- factorisation
static const int kPrimes[] = {2, 3, 5, 7, 11, 13};

bool is_smooth(int64_t n) {
    if (n < 1) return false;
    for (int p : kPrimes)
        while (n % p == 0) n /= p;
    return n == 1;
}

static bool is_smooth7(int64_t n) {
    for (int p : {2, 3, 5, 7})
        while (n % p == 0) n /= p;
    return n == 1;
}

// Smallest 7-smooth length >= m (convolution lengths; 11 and 13 are kept out so that the
// padded FFTs stay on the cheapest codelets).
int64_t next_smooth(int64_t m) {
    int64_t best = 1;
    while (best < m) best *= 2;
    for (int64_t a = 1; a <= best; a *= 2)
        for (int64_t b = a; b <= best; b *= 3)
            for (int64_t c = b; c <= best; c *= 5)
                for (int64_t d = c; d <= best; d *= 7)
                    if (d >= m && d < best) best = d;
    return best;
}

// Radix sequence of one line: 16s (with one 8, 4 or 2 for the remainder), 9s and a 3, then 13, 11,
// 7, 5; largest first.  Empty when n is not smooth over {2,3,5,7,11,13}.
std::vector<int> line_radices(int64_t n) {
    std::vector<int> r;
    if (!is_smooth(n) || n < 2) return r;
    int cnt[14] = {0};
    int64_t m = n;
    for (int p : kPrimes)
        while (m % p == 0) {
            cnt[p]++;
            m /= p;
        }
    const int a = cnt[2], q = a / 4, rem = a % 4;
    if (rem == 1 && q >= 1) {
        for (int i = 0; i < q - 1; ++i) r.push_back(16);
        r.push_back(8);
        r.push_back(4);
    } else {
        for (int i = 0; i < q; ++i) r.push_back(16);
        if (rem == 1) r.push_back(2);
        if (rem == 2) r.push_back(4);
        if (rem == 3) r.push_back(8);
    }
    for (int i = 0; i < cnt[3] / 2; ++i) r.push_back(9);
    if (cnt[3] % 2) r.push_back(3);
    for (int p : {13, 11, 7, 5})
        for (int i = 0; i < cnt[p]; ++i) r.push_back(p);
    std::sort(r.begin(), r.end(), std::greater<int>());
    return r;
}

// Radix sequence of a four-step line.  A power of two 2^k takes the fewest passes of radix <= 16
// with the bits spread as evenly as possible, smallest radix first (1024 = 8·8·16, 2048 =
// 8·16·16): timed over 2^16 .. 2^24 in both precisions on MI355X, these beat the largest-first
// sequences by 1-10 %.  Other lengths take line_radices.
std::vector<int> fourstep_line_radices(int64_t n) {
    if (n < 2 || (n & (n - 1))) return line_radices(n);
    int k = 0;
    while (((int64_t)1 << k) < n) ++k;
    const int passes = (k + 3) / 4, base = k / passes, extra = k % passes;
    std::vector<int> r;
    for (int i = 0; i < passes; ++i) r.push_back(1 << (base + (i >= passes - extra ? 1 : 0)));
    return r;
}

int threads_for(const std::vector<int>& rad, int64_t N, int E) {
    int tt = 1;
    for (int R : rad) {
        const int64_t nb = N / R;
        const int mb = bfly_per_thread(E, R);
        tt = std::max<int>(tt, (int)((nb + mb - 1) / mb));
    }
    return tt;
}

template <typename T>
static int line_stride(int64_t N, bool pad) {
    int ls = pad ? lds_index<T>((int)N - 1) + 1 : (int)N;
    if ((ls & 1) == 0) ++ls;  // odd line stride: consecutive lines fall in different banks
    return ls;
}

int line_stride_for(Prec p, int64_t N, bool pad) {
    return p == Prec::F32 ? line_stride<float>(N, pad) : line_stride<double>(N, pad);
}

bool radices_valid(const std::vector<int>& rad, int64_t N) {
    if (rad.empty() || (int)rad.size() > kMaxPasses) return false;
    int64_t prod = 1;
    for (int R : rad) {
        bool known = false;
        for (int k : {2, 3, 4, 5, 7, 8, 9, 11, 13, 16}) known = known || R == k;
        if (!known || prod > N) return false;
        prod *= R;
    }
    return prod == N;
}

// Builds the pass table, thread count and twiddles of one line FFT.  E = 0 picks the default;
// `radices` (any order of supported radices whose product is N) replaces the default sequence.
int make_line(LinePlan& lp, int64_t N, Prec pr, int E, int max_threads,
              const std::vector<int>* radices) {
    lp.N = (int)N;
    if (N == 1) {  // no passes: the rows kernel copies
        lp.E = E ? E : 16;
        lp.Tt = 1;
        lp.pt.npass = 0;
        lp.Ls = line_stride_for(pr, 1);
        return upload(lp.dtw, std::vector<cld>{cld(1, 0)}, pr);
    }
    std::vector<int> rad = line_radices(N);
    if (radices && !radices->empty()) {
        if (!radices_valid(*radices, N))
            return fail(CFFT_ERR_UNSUPPORTED,
                        "radix sequence does not factor the line length " + std::to_string(N));
        rad = *radices;
    }
    if (rad.empty() || (int)rad.size() > kMaxPasses)
        return fail(CFFT_ERR_UNSUPPORTED, "line length " + std::to_string(N) + " is not smooth");
    if (N >= (1 << 16))  // the stride multipliers (stride_divisor) are exact below 2^16
        return fail(CFFT_ERR_UNSUPPORTED, "line length " + std::to_string(N) + " too long");
    // The kernels are instantiated for E = 8 and E = 16 (a launch picks E <= 8 ? 8 : 16), so
    // the thread count is computed for the instance that will run.
    E = (E > 0 && E <= 8) ? 8 : 16;
    lp.E = E;
    lp.Tt = threads_for(rad, N, E);
    if (lp.Tt > max_threads)
        return fail(CFFT_ERR_UNSUPPORTED, "line length " + std::to_string(N) + " needs " +
                                              std::to_string(lp.Tt) + " threads");
    lp.pt.npass = (int)rad.size();
    lp.pt.wave = 0;
    std::vector<cld> tw;
    const OmegaTable om(N);  // w_g^{pk} = w_N^{s p k}
    int64_t s = 1;
    for (size_t i = 0; i < rad.size(); ++i) {
        const int R = rad[i];
        lp.pt.radix[i] = R;
        lp.pt.twoff[i] = (int)tw.size();
        lp.pt.sdiv[i] = stride_divisor((int)s);
        const int64_t g = N / s, n1 = g / R;
        for (int64_t p = 0; p < n1; ++p)
            for (int k = 1; k < R; ++k) tw.push_back(om(s * p * k));
        s *= R;
    }
    if (tw.empty()) tw.push_back(cld(1, 0));
    lp.Ls = line_stride_for(pr, N);
    lp.pad = 1;
    return upload(lp.dtw, tw, pr);
}

// Threads per line when its last pass runs across the lanes of a wavefront (wave_tail_pass in
// engine.h), or 0 when the line does not allow it.  Conditions: the last radix RL and the
// penultimate radix R are powers of two up to 16 (the kernels' wave-tail instances); the stride
// s before the penultimate pass is a power of two with s RL <= 64, so a digit group (lanes s
// apart) lies in one wavefront and xor on the lane index flips digit bits; every thread holds at
// most one penultimate butterfly (b = t); and the thread count is a multiple of s RL, so the
// groups of every line of a workgroup start on lane multiples of s RL.  The last pass then needs
// no threads of its own, so the count is the largest of the other passes' (s RL <= 64 keeps the
// lines to N <= 1024).
static int wave_tail_threads(const LinePlan& lp) {
    const PassTable& pt = lp.pt;
    if (pt.npass < 2) return 0;
    const int RL = pt.radix[pt.npass - 1], R = pt.radix[pt.npass - 2];
    auto pow2 = [](int64_t v) { return v > 0 && (v & (v - 1)) == 0; };
    if (!pow2(RL) || RL > 16 || !pow2(R) || R > 16) return 0;
    int64_t s = 1;
    for (int i = 0; i < pt.npass - 2; ++i) s *= pt.radix[i];
    if (!pow2(s) || s * RL > 64) return 0;
    int tt = lp.N / R;
    for (int i = 0; i < pt.npass - 2; ++i) {
        const int nb = lp.N / pt.radix[i];
        const int mb = bfly_per_thread(lp.E, pt.radix[i]);
        tt = std::max(tt, (nb + mb - 1) / mb);
    }
    if (tt % (s * RL)) return 0;
    return tt;
}

bool apply_wave_tail(LinePlan& lp) {
    const int tt = wave_tail_threads(lp);
    if (tt <= 0) return false;
    lp.Tt = tt;
    lp.pt.wave = 1;
    return true;
}

// Whether a direct or fused-convolution line runs its last pass across wavefront lanes by
// default (PlanOptions::wave = -1).
constexpr bool kWaveDefault = false;

static int64_t esize(Prec p) { return p == Prec::F32 ? 8 : 16; }
// Bytes of one element of the plan's input and output (ComplexF16: 4).
static int64_t io_esize(const Plan& p) { return p.half ? 4 : esize(p.prec); }

// ------------------------------------------------------------------------------- planning
static int pow2_floor(int64_t v) {
    int r = 1;
    while ((int64_t)r * 2 <= v) r *= 2;
    return r;
}

// Lines per workgroup for the rows kernel.
static int rows_lines_per_block(const LinePlan& lp, Prec pr, int64_t nlines, size_t lds_budget,
                                int target_threads) {
    const size_t per = 2 * (size_t)(pr == Prec::F32 ? 4 : 8) * lp.Ls;
    int c = std::max(1, target_threads / lp.Tt);
    c = std::min<int64_t>(c, std::max<size_t>(1, lds_budget / per));
    c = std::min<int64_t>(c, std::max<int64_t>(1, nlines));
    c = std::min(c, std::max(1, 1024 / lp.Tt));
    return c;
}

// Workgroups of the pipelined rows kernel that are resident at once on the device (0 = the plan
// cannot use it: one pass only, or the lines plus the LDS twiddle copy exceed the LDS).
static int pipe_resident(Ctx* ctx, const Plan& p) {
    if (p.line.pt.npass < 2 || p.C * p.line.Tt > kPipeMaxThreads) return 0;
    const size_t twn = p.line.dtw.bytes / (size_t)esize(p.prec);
    const size_t lds = rows_pipe_lds(p.prec, p.C, p.line.Ls, twn);
    if (lds > ctx->lds_max) return 0;
    const int bpc = rows_pipe_blocks_per_cu(p.prec, p.line.E, p.line.pt.radix[0],
                                            p.C * p.line.Tt, lds);
    return bpc > 0 ? bpc * ctx->num_cu : 0;
}

// pipe: 1 = the pipelined rows kernel, 0 (or -1) = the plain one.
static int plan_direct(Ctx* ctx, Plan& p, int E, int target_threads, int pipe = -1,
                       const std::vector<int>* rad = nullptr, int wave = -1, bool pad = true) {
    int rc = make_line(p.line, p.n, p.prec, E, 1024, rad);
    if (rc) return rc;
    // the wave-tail kernels are instantiated for float / double lines and the plain rows kernel
    if (!p.half && (wave > 0 || (wave < 0 && kWaveDefault)) && apply_wave_tail(p.line)) pipe = 0;
    // unpadded LDS lines: the plain rows kernel only (the pipelined one keeps the pad words)
    if (!pad && pipe <= 0) {
        p.line.pad = 0;
        p.line.Ls = line_stride_for(p.prec, p.n, false);
    }
    const size_t per = 2 * (size_t)esize(p.prec) / 2 * p.line.Ls;
    if (per > ctx->lds_max) return fail(CFFT_ERR_UNSUPPORTED, "line does not fit LDS");
    p.C = rows_lines_per_block(p.line, p.prec, p.batch, std::max<size_t>(per, ctx->lds_rows_budget),
                               target_threads);
    p.pipe_res = pipe > 0 ? pipe_resident(ctx, p) : 0;
    p.kind = Kind::Direct;
    return CFFT_OK;
}

// Can a line of length N run as one workgroup (threads and LDS)?
static bool line_fits(Ctx* ctx, int64_t N, Prec pr, int E);
bool line_fits_lds(Ctx* ctx, int64_t N, Prec pr, int E) { return line_fits(ctx, N, pr, E); }
static bool line_fits(Ctx* ctx, int64_t N, Prec pr, int E) {
    if (N < 2 || !is_smooth(N) || N > (1 << 24)) return false;
    std::vector<int> rad = line_radices(N);
    if ((int)rad.size() > kMaxPasses) return false;
    if (threads_for(rad, N, E) > 1024) return false;
    const size_t per = 2 * (size_t)(pr == Prec::F32 ? 4 : 8) * line_stride_for(pr, N);
    return per <= ctx->lds_max;
}

// Lines per four-step workgroup: a power of two, at most `want`, within 1024 threads and the LDS
// left after `reserve` bytes (pass B's twiddle table).
// LDS line stride of a four-step workgroup of C lines, for its column-first thread mapping
// (lane = c + C t): of the strides >= the line's minimum, the one whose bank conflicts are
// lowest, the smallest on ties.  The model replays every LDS instruction of the workgroup's
// first wave -- each pass's reads of its inputs and writes of its outputs, except the first
// pass's reads and the last pass's writes, which go to global memory -- and charges each
// instruction the bank conflicts of 32 four-byte banks per half wave (weight 4) plus those of 64
// banks over the whole wave (tie-break).  Results are cached per (precision, N, E, C, LDS room).
static int bank_cost(const int* words, int n, int banks) {
    // words: the 4-byte LDS words of one group of lanes (distinct: a pass reads or writes each
    // element once); cost = most words on one bank (banks: a power of two)
    int cnt[64] = {0}, worst = 0;
    for (int i = 0; i < n; ++i) worst = std::max(worst, ++cnt[words[i] & (banks - 1)]);
    return worst;
}

static int fs_line_stride(Ctx* ctx, Prec pr, const LinePlan& lp, int C, size_t reserve) {
    // the model replays the line's passes: the key holds everything it reads
    static std::mutex mu;
    static std::map<std::tuple<int, int, int, int, int, size_t, std::vector<int>>, int> cache;
    const auto key = std::make_tuple((int)pr, lp.N, lp.E, lp.Tt, C,
                                     ctx->lds_max - std::min(ctx->lds_max, reserve),
                                     std::vector<int>(lp.pt.radix, lp.pt.radix + lp.pt.npass));
    {
        std::lock_guard<std::mutex> g(mu);
        auto it = cache.find(key);
        if (it != cache.end()) return it->second;
    }
    const int ew = pr == Prec::F32 ? 1 : 2;
    auto lidx = [pr](int e) { return pr == Prec::F32 ? lds_index<float>(e) : lds_index<double>(e); };
    const int N = lp.N, Tt = lp.Tt, E = lp.E, threads = C * Tt;
    const int waves = 1;
    const int base = lidx(N - 1) + 1;
    int best = base;
    long best_cost = -1;
    int wl[64][2], acc[128], half[64];
    for (int Ls = base; Ls < base + 64; ++Ls) {
        if (fsb_tw_offset(ew * 4, C, Ls) + reserve > ctx->lds_max) break;
        long cost = 0;
        // one instruction: lane L touches element e(L) of column c(L) (e < 0: inactive)
        auto charge = [&](const int (*lane_ce)[2]) {
            int n = 0, nh = 0;
            for (int L = 0; L < 64; ++L) {
                if (lane_ce[L][1] < 0) continue;
                const int w0 = (lane_ce[L][0] * Ls + lidx(lane_ce[L][1])) * ew;
                for (int x = 0; x < ew; ++x) acc[n++] = w0 + x;
            }
            if (!n) return;
            for (int h = 0; h < 2; ++h) {
                nh = 0;
                for (int L = 32 * h; L < 32 * h + 32; ++L) {
                    if (lane_ce[L][1] < 0) continue;
                    const int w0 = (lane_ce[L][0] * Ls + lidx(lane_ce[L][1])) * ew;
                    for (int x = 0; x < ew; ++x) half[nh++] = w0 + x;
                }
                if (nh) cost += 4 * bank_cost(half, nh, 32);
            }
            cost += bank_cost(acc, n, 64);
        };
        // a candidate whose partial cost reaches the best so far cannot win: stop early
        auto over = [&] { return best_cost >= 0 && cost >= best_cost; };
        for (int w = 0; w < waves && !over(); ++w) {
            int s = 1;
            for (int ip = 0; ip < lp.pt.npass && !over(); ++ip) {
                const int R = lp.pt.radix[ip], nb = N / R, MB = std::max(1, E / R);
                const bool rd = ip > 0, wr = ip + 1 < lp.pt.npass;
                for (int ib = 0; ib < MB && !over(); ++ib)
                    for (int j = 0; j < R && !over(); ++j)
                        for (int pass_wr = 0; pass_wr < 2 && !over(); ++pass_wr) {
                            if (pass_wr ? !wr : !rd) continue;
                            for (int L = 0; L < 64; ++L) {
                                const int tid = 64 * w + L, c = tid % C;
                                const int b = tid / C + ib * Tt;
                                wl[L][0] = c;
                                if (tid >= threads || b >= nb) {
                                    wl[L][1] = -1;
                                    continue;
                                }
                                const int p = b / s, q = b - p * s;
                                wl[L][1] = pass_wr ? q + s * R * p + s * j : b + j * nb;
                            }
                            charge(wl);
                        }
                s *= R;
            }
        }
        if (best_cost < 0 || cost < best_cost) {
            best_cost = cost;
            best = Ls;
        }
    }
    std::lock_guard<std::mutex> g(mu);
    cache[key] = best;
    return best;
}

static int pick_tile(Ctx* ctx, const LinePlan& lp, Prec pr, int want, size_t reserve = 0) {
    const size_t per = 2 * (size_t)(pr == Prec::F32 ? 4 : 8) * lp.Ls;
    const size_t room = ctx->lds_max > reserve + 16 ? ctx->lds_max - reserve - 16 : 0;
    int c = pow2_floor(std::max<int64_t>(1, (int64_t)(room / per)));
    c = std::min(c, pow2_floor(std::max(1, 1024 / lp.Tt)));
    c = std::min(c, want);
    return std::max(c, 1);
}

// w_n^e for e < n as (A[e & m] * B[(e >> b) & m]) * C[e >> 2b], b = ceil(log2(n) / 3): three
// tables of about n^(1/3) entries, small enough for pass B's LDS.
static void set_tw3_shape(FourStepPlan& fs, int64_t n) {
    int tb = 0;
    while (((int64_t)1 << (3 * tb)) < n) ++tb;
    fs.tbits = tb;
    fs.tnc = (int)(((n - 1) >> (2 * tb)) + 1);
}
static int upload_tw3(FourStepPlan& fs, int64_t n, Prec pr) {
    const int tb = fs.tbits;
    std::vector<cld> t3;
    t3.reserve(tw3_entries(fs.tbits, fs.tnc));
    for (int64_t i = 0; i < ((int64_t)1 << tb); ++i) t3.push_back(omega(i, n));
    for (int64_t i = 0; i < ((int64_t)1 << tb); ++i) t3.push_back(omega(i << tb, n));
    for (int64_t i = 0; i < fs.tnc; ++i) t3.push_back(omega(i << (2 * tb), n));
    return upload(fs.tw3, t3, pr);
}

static int plan_fft_core(Ctx* ctx, Plan& p, const PlanOptions& o);

static int plan_fourstep_factors(Ctx* ctx, FourStepPlan& fs, int64_t n, Prec pr, int E,
                                 int64_t n1_force, int ca, int cb, int64_t chunk,
                                 const std::vector<int>* rad_a = nullptr,
                                 const std::vector<int>* rad_b = nullptr, int Eb = 0,
                                 int min_threads = 0, double ratio = 1.0, int cb_want = 0) {
    fs.chunk = chunk;
    const int E2 = Eb > 0 ? Eb : E;  // pass B's capacity
    int64_t best1 = 0;
    if (n1_force > 0) {
        best1 = n1_force;
    } else {
        // the split closest to N1 / N2 = ratio (1: balanced) with both factors fitting one
        // workgroup
        double bestscore = 1e300;
        const double lr = std::log(ratio);
        for (int64_t a = 2; a * a <= n * 64 && a < n; ++a) {
            if (n % a) continue;
            const int64_t b = n / a;
            if (!line_fits(ctx, a, pr, E) || !line_fits(ctx, b, pr, E2)) continue;
            const double score =
                std::fabs(std::log((double)a / (double)b) - lr) + (a < b ? 0.01 : 0.0);
            if (score < bestscore) {
                bestscore = score;
                best1 = a;
            }
        }
    }
    if (best1 == 0) return fail(CFFT_ERR_UNSUPPORTED, "no four-step split for n=" + std::to_string(n));
    fs.N1 = (int)best1;
    fs.N2 = (int)(n / best1);
    const std::vector<int> da = fourstep_line_radices(fs.N1), db = fourstep_line_radices(fs.N2);
    int rc = make_line(fs.a, fs.N1, pr, E, 1024, rad_a ? rad_a : &da);
    if (rc) return rc;
    rc = make_line(fs.b, fs.N2, pr, E2, 1024, rad_b ? rad_b : &db);
    if (rc) return rc;
    set_tw3_shape(fs, n);
    const size_t es = (size_t)esize(pr);
    const size_t table = (size_t)tw3_entries(fs.tbits, fs.tnc) * es;
    // Narrow tiles (4 columns; 8 in F32 pass B) let several workgroups share a CU, so one
    // loads while another computes; consecutive tiles run on one XCD (xcd_block), which shares
    // the partly covered 128-byte lines in its L2.  Measured best on MI355X for n = 2^20.
    // (min_threads: a default tile grows until its workgroup has that many threads)
    auto want = [min_threads](int c, const LinePlan& lp) {
        while (c * lp.Tt < min_threads && 2 * c * lp.Tt <= 1024) c *= 2;
        return c;
    };
    fs.CA = ca > 0 ? ca : pick_tile(ctx, fs.a, pr, want(4, fs.a));
    fs.CB = cb > 0 ? cb
                   : pick_tile(ctx, fs.b, pr,
                               want(cb_want > 0 ? cb_want : (pr == Prec::F32 ? 8 : 4), fs.b), table);
    fs.a.Ls = fs_line_stride(ctx, pr, fs.a, fs.CA, 16);
    fs.b.Ls = fs_line_stride(ctx, pr, fs.b, fs.CB, table + 16);
    // a requested tile must be a power of two that fits 1024 threads and the LDS
    for (int pass = 0; pass < 2; ++pass) {
        const int C = pass ? fs.CB : fs.CA;
        const LinePlan& lp = pass ? fs.b : fs.a;
        const size_t lds = pass ? fsb_tw_offset(es / 2, C, lp.Ls) + table
                                : 2 * (es / 2) * lp.Ls * (size_t)C;
        if (C < 1 || (C & (C - 1)) || C * lp.Tt > 1024 || lds > ctx->lds_max)
            return fail(CFFT_ERR_UNSUPPORTED, "four-step tile does not fit a workgroup");
    }
    return upload_tw3(fs, n, pr);
}

// Plans n = N1 * M too long for the four-step (its factors do not fit a workgroup), or forced
// by o.large: the six-step path.
//   1. pass A: the N1-point FFTs of the M columns of x (N1 x M row-major), column j2 written as
//      row j2 of a middle M x N1 (k_fs_a);
//   2. transpose to N1 x M, times w_n^{j2 k1} (k_transpose);
//   3. the M-point FFTs of the N1 contiguous rows (sub-plan: direct or four-step, in place);
//   4. transpose to M x N1: X[k1 + N1 k2], conjugated and scaled for backward transforms.
// Five reads and writes per element instead of the four-step's two; the sizes it serves cannot
// run otherwise.
static int plan_large(Ctx* ctx, Plan& p, const PlanOptions& o) {
    const int64_t n = p.n;
    const int E = o.E ? o.E : (p.prec == Prec::F64 ? 8 : 16);
    int64_t N1 = 0;
    if (o.n1 > 0) {
        N1 = o.n1;
        if (n % N1 || !line_fits(ctx, N1, p.prec, E))
            return fail(CFFT_ERR_UNSUPPORTED, "six-step split does not fit a workgroup");
    } else {  // the largest factor that fits a workgroup, so that M is as short as possible
        for (int64_t a = 2; a * a <= n; ++a) {
            if (n % a) continue;
            if (line_fits(ctx, a, p.prec, E)) N1 = std::max(N1, a);
            if (line_fits(ctx, n / a, p.prec, E)) N1 = std::max(N1, n / a);
        }
    }
    if (N1 < 2 || N1 == n) return fail(CFFT_ERR_UNSUPPORTED, "no six-step split for n=" + std::to_string(n));
    const int64_t M = n / N1;
    FourStepPlan& fs = p.fs;
    fs.N1 = (int)N1;
    fs.N2 = (int)M;
    int rc = make_line(fs.a, N1, p.prec, E, 1024);
    if (rc) return rc;
    fs.CA = pick_tile(ctx, fs.a, p.prec, 4);
    fs.a.Ls = fs_line_stride(ctx, p.prec, fs.a, fs.CA, 16);
    if (2 * (size_t)(esize(p.prec) / 2) * fs.a.Ls * fs.CA > ctx->lds_max || fs.CA * fs.a.Tt > 1024)
        return fail(CFFT_ERR_UNSUPPORTED, "six-step tile does not fit a workgroup");
    set_tw3_shape(fs, n);
    if ((rc = upload_tw3(fs, n, p.prec))) return rc;
    std::unique_ptr<Plan> s(new Plan());
    s->ctx = ctx;
    s->n = M;
    s->batch = N1;
    s->idist = s->odist = M;
    s->prec = p.prec;
    s->dtype = p.prec == Prec::F32 ? CFFT_C32 : CFFT_C64;
    s->flags = CFFT_NO_FLAG;
    if ((rc = plan_fft_core(ctx, *s, PlanOptions{}))) return rc;
    if (s->kind != Kind::Direct && s->kind != Kind::FourStep)
        return fail(CFFT_ERR_UNSUPPORTED, "six-step rows of length " + std::to_string(M) +
                                              " are not plannable");
    p.sub = std::move(s);
    p.kind = Kind::Large;
    return CFFT_OK;
}


// Convolution data of a Rader (prime n) or Bluestein transform: permutations or chirp, and the
// zero-padded kernel b whose spectrum the convolution multiplies by.
int conv_data(int64_t n, ConvData& d) {
    const bool rader = is_prime(n);
    d.mode = rader ? ConvMode::Rader : ConvMode::Bluestein;
    d.perm.clear();
    d.iperm.clear();
    d.chirp.clear();
    if (rader) {
        const int64_t N = n - 1;
        d.L = is_smooth(N) ? N : next_smooth(2 * N - 1);
        const int64_t g = primitive_root(n);
        const int64_t ginv = powmod(g, n - 2, n);
        d.perm.resize(N);
        d.iperm.resize(N);
        const OmegaTable om(n);
        std::vector<cld> w(N);
        int64_t gp = 1, gi = 1;
        for (int64_t i = 0; i < N; ++i) {
            d.perm[i] = (int)gp;
            d.iperm[i] = (int)gi;
            w[i] = om(gi);  // w[i] = w_n^{(g^-1)^i}  (prime.jl:36-37, rounded once)
            gp = gp * g % n;
            gi = gi * ginv % n;
        }
        d.b.assign(d.L, cld(0, 0));
        for (int64_t m = 0; m < N; ++m) d.b[m] = w[m];
        for (int64_t m = d.L - N + 1; m < d.L; ++m) d.b[m] = w[m - (d.L - N)];
        d.g = g;
    } else {
        d.L = next_smooth(2 * n - 1);
        d.chirp.resize(n);
        const OmegaTable om(2 * n);
        for (int64_t m = 0; m < n; ++m) {
            const int64_t e = (int64_t)((unsigned __int128)m * m % (unsigned __int128)(2 * n));
            d.chirp[m] = om(e);  // e^{-pi i m^2 / n}
        }
        d.b.assign(d.L, cld(0, 0));
        for (int64_t m = 0; m < n; ++m) d.b[m] = std::conj(d.chirp[m]);
        for (int64_t m = 1; m < n; ++m) d.b[d.L - m] = std::conj(d.chirp[m]);
        d.g = 0;
    }
    return CFFT_OK;
}

// K = FFT_L(b), rounded once to cprec.  Small L: long double on the host; larger L: this
// library's own FP64 FFT on the device (FP64 is kept for kernel generation only).
// A default (NO_FLAG) plan of `batch` contiguous transforms of length n from the context's plan
// cache, created on first use.
static int cached_plan_1d(Ctx* c, int64_t n, int64_t batch, int dtype, std::shared_ptr<Plan>* out) {
    const PlanKey key{n, batch, n, n, dtype, CFFT_NO_FLAG};
    {
        std::lock_guard<std::mutex> g(c->mu);
        auto it = c->cache.find(key);
        if (it != c->cache.end()) {
            *out = it->second;
            return CFFT_OK;
        }
    }
    std::shared_ptr<Plan> p;
    int rc = plan_create(c, n, batch, n, n, dtype, CFFT_NO_FLAG, PlanOptions{}, &p);
    if (rc) return rc;
    std::lock_guard<std::mutex> g(c->mu);
    auto ins = c->cache.emplace(key, p);  // another thread may have made it meanwhile
    *out = ins.first->second;
    return CFFT_OK;
}

int conv_spectrum(Ctx* ctx, const std::vector<cld>& b, Prec cprec, DevBuf& K) {
    const int64_t L = (int64_t)b.size();
    if (L <= 4096) {
        std::vector<cld> h = b;
        if (!host_fft_ld(h)) return fail(CFFT_ERR_UNSUPPORTED, "kernel spectrum length not smooth");
        return upload(K, h, cprec);
    }
    // the FP64 plan comes from the context's cache: convolutions of nearby lengths share L
    std::shared_ptr<Plan> t;
    int rc = cached_plan_1d(ctx, L, 1, CFFT_C64, &t);
    if (rc) return rc;
    DevBuf in, out;
    if ((rc = upload(in, b, Prec::F64))) return rc;
    if ((rc = out.alloc((size_t)L * 16))) return rc;
    if ((rc = plan_exec(*t, in.p, out.p, CFFT_FORWARD, 0, ctx->stream))) return rc;
    CFFT_HIP(hipStreamSynchronize(ctx->stream));
    if (cprec == Prec::F64) {
        std::swap(K.p, out.p);
        std::swap(K.bytes, out.bytes);
        return CFFT_OK;
    }
    std::vector<double> h((size_t)2 * L);
    CFFT_HIP(hipMemcpy(h.data(), out.p, (size_t)L * 16, hipMemcpyDeviceToHost));
    std::vector<float> f((size_t)2 * L);
    for (size_t i = 0; i < f.size(); ++i) f[i] = (float)h[i];
    if ((rc = K.alloc((size_t)L * 8))) return rc;
    CFFT_HIP(hipMemcpy(K.p, f.data(), (size_t)L * 8, hipMemcpyHostToDevice));
    return CFFT_OK;
}

// Precision of a convolution's work: a Rader convolution of F32 data with p > 2^16 runs in FP64
// (the precision policy for long Rader convolutions); Bluestein convolves in the data precision
// at every length (its kernel spectrum is FP64-generated and rounded once: relative L2 error
// 2.7e-7 to 3.0e-7 in F32 for n = 131074 .. 262202).
Prec conv_precision(int64_t n, Prec io, ConvMode mode) {
    return (io == Prec::F32 && mode == ConvMode::Rader && n > 65536) ? Prec::F64 : io;
}

// Rader / Bluestein planning.
static int plan_conv(Ctx* ctx, Plan& p, const PlanOptions& o) {
    const int64_t n = p.n;
    ConvData d;
    conv_data(n, d);
    p.mode = d.mode;
    p.L = d.L;
    p.g = d.g;
    p.cprec = conv_precision(n, p.prec, p.mode);
    const std::vector<int>& perm = d.perm;
    const std::vector<int>& iperm = d.iperm;
    const std::vector<cld>& chirp = d.chirp;
    const bool rader = p.mode == ConvMode::Rader;
    int rc;
    // fused small shape?
    const int E = o.E ? o.E : 16;
    bool small = false;
    if (p.cprec == p.prec && line_fits(ctx, p.L, p.prec, E)) {
        rc = make_line(p.line, p.L, p.prec, E, 1024);
        if (rc) return rc;
        if (!p.half && (o.wave > 0 || (o.wave < 0 && kWaveDefault))) apply_wave_tail(p.line);
        // the input lines and the convolution work share one LDS line (the kernel saves x0 and
        // separates reads from writes where they overlap)
        p.LsC = std::max(line_stride_for(p.prec, n), p.line.Ls);
        const size_t per = 2 * (size_t)(esize(p.prec) / 2) * p.LsC + esize(p.prec);
        if (per <= ctx->lds_max) {
            small = true;
            // small workgroups by default (PlanOptions::target_threads = 64): more of them share a
            // CU, which measured 13-14 % faster than 256 for p = 17 and 257 on MI355X
            int c = std::max(1, o.target_threads / p.line.Tt);
            c = std::min<int64_t>(c, std::max<size_t>(1, std::max(ctx->lds_rows_budget, per) / per));
            c = std::min<int64_t>(c, std::max<int64_t>(1, p.batch));
            c = std::min(c, std::max(1, 1024 / p.line.Tt));
            p.C = c;
        }
    }
    if (rader) {
        if ((rc = upload_int(p.perm, perm))) return rc;
        if ((rc = upload_int(p.iperm, iperm))) return rc;
    } else {
        if ((rc = upload(p.chirp, chirp, p.cprec))) return rc;
    }
    if ((rc = conv_spectrum(ctx, d.b, p.cprec, p.K))) return rc;
    if (small) {
        p.kind = Kind::ConvSmall;
        return CFFT_OK;
    }
    // large shape: sub-plan for the length-L FFT in cprec, batch lines, contiguous
    p.kind = Kind::ConvLarge;
    p.sub.reset(new Plan());
    Plan& s = *p.sub;
    s.ctx = ctx;
    s.n = p.L;
    s.batch = p.batch;
    s.idist = s.odist = p.L;
    s.prec = p.cprec;
    s.dtype = p.cprec == Prec::F32 ? CFFT_C32 : CFFT_C64;
    PlanOptions so = o;
    so.n1 = 0;
    so.rad_a.clear();  // radix sequences and the wave tail were chosen for the transform itself
    so.rad_b.clear();
    so.wave = -1;
    // the fused convolution kernels have no line-pair variant: the FFT_L keeps the float
    // kernels' default lines (E = 16 in F32)
    if (so.pair < 0) so.pair = 0;
    // and its own tiles: widening them (as an uneven batch does for its four-step plans) made
    // the uneven batch's large convolutions slower (9.10 -> 9.34 ms)
    so.min_threads = 0;
    rc = plan_fft_core(ctx, s, so);
    if (rc) return rc;
    if (s.kind != Kind::Direct && s.kind != Kind::FourStep && s.kind != Kind::Large)
        return fail(CFFT_ERR_UNSUPPORTED, "convolution length not plannable");
    if (s.kind == Kind::FourStep && o.fuse_conv != 0) {
        // Three fused launches: A runs the N1-point lines (columns), BA the N2-point lines, C the
        // N1-point lines over the second FFT's middle (rows of N2).  Each takes up to 4 lines per
        // workgroup that fit the LDS: 4 columns read 64-byte runs of F64 rows, which measured
        // faster than 2 (32-byte runs) with twice the workgroups per CU (p = 1048573, L = 2^21:
        // 1.37 vs 1.43 ms).
        const size_t table = (size_t)tw3_entries(s.fs.tbits, s.fs.tnc) * esize(p.cprec);
        const size_t es = (size_t)esize(p.cprec) / 2;
        const size_t budget = ctx->lds_max;
        auto tile = [&](const LinePlan& lp, size_t reserve) {
            int c = 4;
            while (c > 1 && (2 * es * (size_t)c * lp.Ls + reserve > budget || c * lp.Tt > 1024))
                c /= 2;
            return c;
        };
        const int ca = tile(s.fs.a, 0), cb = tile(s.fs.b, table), cc = tile(s.fs.a, table);
        const int lsa = fs_line_stride(ctx, p.cprec, s.fs.a, ca, 16);
        const int lsb = fs_line_stride(ctx, p.cprec, s.fs.b, cb, table + 16);
        const int lsc = fs_line_stride(ctx, p.cprec, s.fs.a, cc, table + 16);
        if (ca * s.fs.a.Tt <= 1024 && cb * s.fs.b.Tt <= 1024 && cc * s.fs.a.Tt <= 1024 &&
            2 * es * (size_t)ca * lsa <= ctx->lds_max &&
            fsb_tw_offset(es, cb, lsb) + table <= ctx->lds_max &&
            fsb_tw_offset(es, cc, lsc) + table <= ctx->lds_max) {
            p.convl = true;
            p.convl_ca = ca;
            p.convl_lsa = lsa;
            p.convl_cb = cb;
            p.convl_lsb = lsb;
            p.convl_cc = cc;
            p.convl_lsc = lsc;
        }
    }
    if (!o.alloc_work) return CFFT_OK;  // the caller supplies per-transform work buffers
    const size_t wbytes = (size_t)p.batch * p.L * esize(p.cprec);
    if ((rc = p.w1.alloc(wbytes))) return rc;
    if ((rc = p.w2.alloc(wbytes))) return rc;
    if ((rc = p.dc.alloc((size_t)p.batch * 2 * esize(p.cprec)))) return rc;
    return CFFT_OK;
}

// Transforms per four-step round by default: the whole batch (measured fastest on MI355X for
// 2^20 x 64 in F32 and F64 once pass B's twiddles come from LDS); smaller rounds, whose middle
// stays in the Infinity Cache, are tuning candidates.  CFFT_FOURSTEP_CHUNK overrides it.
static int64_t default_chunk(const Plan& p) {
    if (const char* e = std::getenv("CFFT_FOURSTEP_CHUNK")) return std::max<int64_t>(0, std::atoll(e));
    (void)p;
    return 0;
}

// Resident workgroups of the pipelined four-step passes (0 = a pass cannot use it).
static void fs_pipe_setup(Ctx* ctx, Plan& p, int mask) {
    FourStepPlan& fs = p.fs;
    fs.pipe_a = fs.pipe_b = 0;
    for (int pass = 0; pass < 2; ++pass) {
        if (!((mask >> pass) & 1)) continue;
        const LinePlan& lp = pass ? fs.b : fs.a;
        const int C = pass ? fs.CB : fs.CA;
        if (C * lp.Tt > kPipeMaxThreads || lp.pt.npass < 2) continue;
        const size_t twn = lp.dtw.bytes / (size_t)esize(p.prec);
        const size_t t3n = pass ? (size_t)tw3_entries(fs.tbits, fs.tnc) : 0;
        const size_t lds = fs_pipe_lds(p.prec, C, lp.Ls, twn, t3n);
        if (lds > ctx->lds_max) continue;
        const int bpc = fs_pipe_blocks_per_cu(p.prec, lp.E, pass, lp.pt.radix[0], C * lp.Tt, lds);
        (pass ? fs.pipe_b : fs.pipe_a) = bpc > 0 ? bpc * ctx->num_cu : 0;
    }
}

// Two lines per thread for the ComplexF32 four-step passes in `mask` (bit 0: A, bit 1: B) whose
// tile holds an even number of lines of an even-length stride (N2 for A, N1 for B) and whose
// pair layout fits the LDS.  The pair LDS line (C / 2 lines of 8-byte words) gets its own
// stride from the bank-conflict model of 8-byte words.
static void fs_pair_setup(Ctx* ctx, Plan& p, int mask) {
    FourStepPlan& fs = p.fs;
    fs.pair_a = fs.pair_b = 0;
    if (p.prec != Prec::F32) return;
    if ((mask & 1) && fs.pipe_a == 0 && fs.CA >= 2 && fs.CA % 2 == 0 && fs.N2 % 2 == 0) {
        fs.LsA2 = fs_line_stride(ctx, Prec::F64, fs.a, fs.CA / 2, 16);
        fs.pair_a = fs_pair_lds(0, fs.CA, fs.LsA2, 0, 0, false) <= ctx->lds_max;
    }
    if ((mask & 2) && fs.pipe_b == 0 && fs.CB >= 2 && fs.CB % 2 == 0 && fs.N1 % 2 == 0) {
        const size_t table = (size_t)tw3_entries(fs.tbits, fs.tnc) * sizeof(float2);
        fs.LsB2 = fs_line_stride(ctx, Prec::F64, fs.b, fs.CB / 2, table + 16);
        fs.pair_b = fs_pair_lds(1, fs.CB, fs.LsB2, fs.tbits, fs.tnc, true) <= ctx->lds_max;
    }
}

// Four-step plans of short power-of-two lengths (see plan_fft_core): N1 / N2 near this ratio,
// and by default pass-B tiles of this many lines (where they fit).
constexpr double kSkewSplitRatio = 4.0;
constexpr int kSkewTileB = 16;

static int plan_fft_core(Ctx* ctx, Plan& p, const PlanOptions& o) {
    if (p.n == 1) {
        p.kind = Kind::Trivial;
        return CFFT_OK;
    }
    const int E = o.E ? o.E : 16;
    if (o.force_conv == 0 && is_smooth(p.n)) {
        // Direct lines default to E = 8 (more threads per line) when the line fits a workgroup
        // that way, except F32 lines of 512 points or more with a radix above 9, which take
        // E = 16: measured cold on MI355X (512 MiB batches), E = 16 ran those lines 10-40 %
        // faster than E = 8 (512 .. 8192 points; 128 points favour E = 8) and the pipelined
        // rows kernel (their former default) 20-80 % slower than either; it stays a MEASURE
        // candidate.
        // a forced radix sequence for the whole length makes a direct line; for a factor of n
        // it is the four-step's first line and fixes the split
        int64_t rad_prod = 1;
        for (int R : o.rad_a) rad_prod *= R;
        const bool rad_direct = !o.rad_a.empty() && rad_prod == p.n;
        const int64_t n1 = o.n1 ? o.n1 : (!o.rad_a.empty() && !rad_direct ? rad_prod : 0);
        if (n1 == 0) {
            int Ed = o.E ? o.E : (line_fits(ctx, p.n, p.prec, 8) ? 8 : 16);
            const std::vector<int> dr = rad_direct ? o.rad_a : line_radices(p.n);
            const int rmax = dr.empty() ? 0 : *std::max_element(dr.begin(), dr.end());
            if (!o.E && p.prec == Prec::F32 && rmax > 9 && p.n >= 512 &&
                line_fits(ctx, p.n, p.prec, 16))
                Ed = 16;
            const int pipe = o.pipe >= 0 ? o.pipe : 0;
            if (line_fits(ctx, p.n, p.prec, Ed) || rad_direct)
                return plan_direct(ctx, p, Ed, o.target_threads, pipe,
                                   rad_direct ? &o.rad_a : nullptr, o.wave, o.lds_pad != 0);
        }
        if (o.large) return plan_large(ctx, p, o);
        // Four-step lines default to E = 8 in F64 (twice the threads per tile), 16 in F32.
        const int64_t chunk = o.chunk >= 0 ? o.chunk : default_chunk(p);
        // default tiles grow to this many threads per workgroup (0: pick_tile's tiles; as a
        // rule for every plan, 256 gained up to 3 % on some lengths and lost up to 5 % on others
        // on MI355X, so only uneven batches ask for it)
        const int min_thr = std::max(o.min_threads, 0);
        const std::vector<int>* ra = o.rad_a.empty() ? nullptr : &o.rad_a;
        const std::vector<int>* rb = o.rad_b.empty() ? nullptr : &o.rad_b;
        // Powers of two up to 2^22 in ComplexF32 (and ComplexF16), up to 2^17 in ComplexF64, split
        // N1 = kSkewSplitRatio N2, and their default F64 and line-pair plans take pass-B tiles of
        // kSkewTileB lines: MEASURE's choice there.  NO_FLAG on MI355X, cold, 512 MiB against the
        // balanced split: F32 2^15 / 2^16 / 2^17 / 2^20 / 2^22 9 / 13 / 6 / 11 / 3 % faster (2^14,
        // 2^18 equal; 2^24 36 % and 10^5, 10^6, 4 5^8, 3 2^19 3-11 % slower); F64 2^14 .. 2^17
        // 8-12 % faster (2^18 .. 2^21 and 2^24 2-9 % slower).  Every plan of such a length and
        // precision takes the same split, so a transform keeps its bits alone, in a work list or
        // as ComplexF16.
        const bool skew = (p.n & (p.n - 1)) == 0 &&
                          p.n <= ((int64_t)1 << (p.prec == Prec::F32 ? 22 : 17));
        const double ratio = skew ? kSkewSplitRatio : 1.0;
        int rc = CFFT_ERR_UNSUPPORTED;
        if (!o.E && p.prec == Prec::F64)
            rc = plan_fourstep_factors(ctx, p.fs, p.n, p.prec, 8, n1, o.ca, o.cb, chunk, ra, rb,
                                       o.Eb, min_thr, ratio, skew ? kSkewTileB : 0);
        // ComplexF32 lines of even length on both sides default to E = 8 with both passes on
        // line pairs (MEASURE's choice for every power of two 2^16 .. 2^24 and 4 5^8; E = 8
        // without pairs is slower than 16).  ComplexF16 plans run the float kernels: E = 16.
        const bool f32_pairs = !o.E && p.prec == Prec::F32 && p.dtype != CFFT_C16 &&
                               o.pair != 0 && p.n % 4 == 0;
        if (f32_pairs) {
            rc = plan_fourstep_factors(ctx, p.fs, p.n, p.prec, 8, n1, o.ca, o.cb, chunk, ra, rb,
                                       o.Eb, min_thr, ratio, skew ? kSkewTileB : 0);
            if (rc == CFFT_OK && (p.fs.N1 % 2 || p.fs.N2 % 2)) rc = CFFT_ERR_UNSUPPORTED;
        }
        if (rc)
            rc = plan_fourstep_factors(ctx, p.fs, p.n, p.prec, E, n1, o.ca, o.cb, chunk, ra, rb,
                                       o.Eb, min_thr, ratio);
        if (rc && n1 == 0) {  // too long for the four-step: six steps
            p.fs.b.dtw.reset();
            p.fs.tw3.reset();
            p.fs.CB = 1;
            p.fs.chunk = 0;
            p.fs.pipe_a = p.fs.pipe_b = 0;
            return plan_large(ctx, p, o);
        }
        if (rc) return rc;
        p.kind = Kind::FourStep;
        const int pipe_mask = o.pipe_mask >= 0 ? o.pipe_mask : (o.pipe > 0 ? 3 : 0);
        if (pipe_mask) fs_pipe_setup(ctx, p, pipe_mask);
        // ComplexF32 defaults to line pairs (same bits as the float kernels): both passes with
        // the E = 8 lines above, pass A otherwise (MEASURE chose pass A pairs for every even
        // split it timed with E = 16)
        const int pair = o.pair >= 0 ? o.pair
                                     : (p.prec != Prec::F32 ? 0 : (p.fs.a.E == 8 ? 3 : 1));
        if (pair > 0) fs_pair_setup(ctx, p, pair);
        return CFFT_OK;
    }
    return plan_conv(ctx, p, o);
}

int plan_build(Ctx* ctx, Plan& p, const PlanOptions& o) { return plan_fft_core(ctx, p, o); }

static std::string describe(const Plan& p) {
    std::ostringstream os;
    auto rad = [](const LinePlan& l) {
        std::ostringstream r;
        r << "[";
        for (int i = 0; i < l.pt.npass; ++i) r << (i ? "," : "") << l.pt.radix[i];
        r << "] Tt=" << l.Tt << " E=" << l.E;
        return r.str();
    };
    os << "n=" << p.n << " batch=" << p.batch << " prec=" << (p.prec == Prec::F32 ? "f32" : "f64")
       << " ";
    switch (p.kind) {
        case Kind::Trivial: os << "trivial"; break;
        case Kind::Direct:
            os << "direct " << rad(p.line) << " C=" << p.C;
            if (p.line.pt.wave) os << " wave";
            if (!p.line.pad) os << " nopad";
            if (p.pipe_res) os << " pipe";
            break;
        case Kind::FourStep:
            os << "fourstep N1=" << p.fs.N1 << " " << rad(p.fs.a) << " CA=" << p.fs.CA
               << " Ls=" << p.fs.a.Ls << " N2=" << p.fs.N2 << " " << rad(p.fs.b)
               << " CB=" << p.fs.CB << " Ls=" << p.fs.b.Ls;
            if (p.fs.chunk) os << " chunk=" << p.fs.chunk;
            if (p.fs.pipe_a || p.fs.pipe_b)
                os << " pipe=" << (p.fs.pipe_a ? "A" : "") << (p.fs.pipe_b ? "B" : "");
            if (p.fs.pair_a || p.fs.pair_b)
                os << " pair=" << (p.fs.pair_a ? "A" : "") << (p.fs.pair_b ? "B" : "");
            break;
        case Kind::Large:
            os << "sixstep N1=" << p.fs.N1 << " " << rad(p.fs.a) << " CA=" << p.fs.CA
               << " Ls=" << p.fs.a.Ls << " M=" << p.fs.N2 << " {" << describe(*p.sub) << "}";
            break;
        case Kind::ConvSmall:
            os << (p.mode == ConvMode::Rader ? "rader" : "bluestein") << "-fused L=" << p.L << " "
               << rad(p.line) << " C=" << p.C;
            if (p.line.pt.wave) os << " wave";
            if (p.mode == ConvMode::Rader) os << " g=" << p.g;
            break;
        case Kind::ConvLarge:
            os << (p.mode == ConvMode::Rader ? "rader" : "bluestein") << "-large L=" << p.L
               << " conv=" << (p.cprec == Prec::F32 ? "f32" : "f64");
            if (p.convl)
                os << " fused3 CA=" << p.convl_ca << " CB=" << p.convl_cb << " CC=" << p.convl_cc;
            os << " {" << describe(*p.sub) << "}";
            if (p.mode == ConvMode::Rader) os << " g=" << p.g;
            break;
        default: os << "uneven"; break;
    }
    if (p.half) os << " io=f16";
    return os.str();
}

// ------------------------------------------------------------------------------- execution
static int run_fft_s(const Plan& p, const void* in, void* out, int sgn, double scale,
                     ConvEpilogue epi, hipStream_t st, const ExecShape& sh);
static int run_fft(const Plan& p, const void* in, void* out, int sgn, double scale,
                   ConvEpilogue epi, hipStream_t st) {
    ExecShape sh;
    sh.batch = p.batch;
    return run_fft_s(p, in, out, sgn, scale, epi, st, sh);
}

// Runs the two passes over the batch in chunks of fs.chunk transforms (0 = the whole batch, the
// default: chunks small enough for the middle to stay in the Infinity Cache measured slower on
// MI355X, their launch tails outweighing the saved HBM traffic).  Chunking stays a tuning knob.
static int run_fourstep(const Plan& p, const void* in, void* mid, void* out, int64_t idist,
                        int64_t mdist, int64_t odist, int sgn, double scale, ConvEpilogue epi,
                        hipStream_t st, int64_t B) {
    const FourStepPlan& fs = p.fs;
    const int64_t chunk = fs.chunk > 0 ? std::min<int64_t>(fs.chunk, B) : B;
    const size_t es = esize(p.prec), eio = io_esize(p);
    for (int64_t b0 = 0; b0 < B; b0 += chunk) {
        const int64_t nb = std::min<int64_t>(chunk, B - b0);
        const char* inb = (const char*)in + (size_t)b0 * idist * eio;
        char* midb = (char*)mid + (size_t)b0 * mdist * es;
        char* outb = (char*)out + (size_t)b0 * odist * eio;
        FourStepAArgs a{};
        a.N1 = fs.N1;
        a.N2 = fs.N2;
        a.C = fs.CA;
        a.Tt = fs.a.Tt;
        a.Ls = fs.a.Ls;
        a.E = fs.a.E;
        a.batch = nb;
        a.idist = idist;
        a.odist = mdist;
        a.sgn_in = sgn;
        a.xcd = 1;
        a.mid_plain = fs.chunk > 0;  // a chunk's middle is read back while still cached
        const int64_t es_ = esize(p.prec);
        auto al16 = [](const void* q) { return ((uintptr_t)q & 15) == 0; };
        if (p.half) {
            CFFT_HIP(launch_fourstep_a_f16(inb, midb, a, fs.a.pt, fs.a.dtw.p, st));
        } else if (fs.pair_a && al16(inb) && al16(midb) && idist % 2 == 0 && mdist % 2 == 0) {
            a.Ls = fs.LsA2;
            CFFT_HIP(launch_fourstep_a_pair(inb, midb, a, fs.a.pt, fs.a.dtw.p, st));
        } else if (fs.pipe_a > 0) {
            a.grid = (int)std::min<int64_t>(fs.pipe_a, (int64_t)((fs.N2 + fs.CA - 1) / fs.CA) * nb);
            CFFT_HIP(launch_fourstep_a_pipe(p.prec, inb, midb, a, fs.a.pt, fs.a.dtw.p,
                                            fs.a.dtw.bytes / (size_t)es_, st));
        } else {
            CFFT_HIP(launch_fourstep_a(p.prec, inb, midb, a, fs.a.pt, fs.a.dtw.p, st));
        }
        FourStepBArgs b{};
        b.N1 = fs.N1;
        b.N2 = fs.N2;
        b.C = fs.CB;
        b.Tt = fs.b.Tt;
        b.Ls = fs.b.Ls;
        b.E = fs.b.E;
        b.batch = nb;
        b.idist = mdist;
        b.odist = odist;
        b.sgn_in = 1;
        b.sgn_out = sgn;
        b.scale = scale;
        b.epi = epi;
        if (epi.dc) b.epi.dc = (char*)epi.dc + (size_t)b0 * 2 * es;
        b.tbits = fs.tbits;
        b.tnc = fs.tnc;
        b.tw3 = fs.tw3.p;
        b.xcd = 1;
        if (p.half) {
            CFFT_HIP(launch_fourstep_b_f16(midb, outb, b, fs.b.pt, fs.b.dtw.p, st));
        } else if (fs.pair_b && !epi.K && al16(midb) && al16(outb) && mdist % 2 == 0 &&
                   odist % 2 == 0) {
            b.Ls = fs.LsB2;
            CFFT_HIP(launch_fourstep_b_pair(midb, outb, b, fs.b.pt, fs.b.dtw.p, st));
        } else if (fs.pipe_b > 0 && !epi.K) {
            b.grid = (int)std::min<int64_t>(fs.pipe_b, (int64_t)((fs.N1 + fs.CB - 1) / fs.CB) * nb);
            CFFT_HIP(launch_fourstep_b_pipe(p.prec, midb, outb, b, fs.b.pt, fs.b.dtw.p,
                                            fs.b.dtw.bytes / (size_t)es_, st));
        } else {
            CFFT_HIP(launch_fourstep_b(p.prec, midb, outb, b, fs.b.pt, fs.b.dtw.p, st));
        }
    }
    return CFFT_OK;
}

static int run_fft_s(const Plan& p, const void* in, void* out, int sgn, double scale,
                     ConvEpilogue epi, hipStream_t st, const ExecShape& sh) {
    const int64_t B = sh.batch;
    switch (p.kind) {
        case Kind::Trivial: {
            const int64_t es = io_esize(p);
            if (B == 0) return CFFT_OK;
            if (in == out && p.idist == p.odist) return CFFT_OK;
            CFFT_HIP(hipMemcpy2DAsync(out, p.odist * es, in, p.idist * es, es, B,
                                      hipMemcpyDeviceToDevice, st));
            return CFFT_OK;
        }
        case Kind::Direct: {
            RowsArgs a{};
            a.N = (int)p.n;
            a.C = p.C;
            a.Tt = p.line.Tt;
            a.Ls = p.line.Ls;
            a.E = p.line.E;
            a.nlines = B;
            a.idist = p.idist;
            a.odist = p.odist;
            a.sgn_in = sgn;
            a.sgn_out = sgn;
            a.scale = scale;
            a.epi = epi;
            a.nopad = p.line.pad ? 0 : 1;
            if (p.half) {
                CFFT_HIP(launch_rows_f16(in, out, a, p.line.pt, p.line.dtw.p, st));
                return CFFT_OK;
            }
            if (p.pipe_res > 0 && !epi.K) {
                const int64_t groups = (B + p.C - 1) / p.C;
                a.grid = (int)std::min<int64_t>(groups, p.pipe_res);
                if (a.grid == 0) return CFFT_OK;
                CFFT_HIP(launch_rows_pipe(p.prec, in, out, a, p.line.pt, p.line.dtw.p,
                                          p.line.dtw.bytes / (size_t)esize(p.prec), st));
                return CFFT_OK;
            }
            CFFT_HIP(launch_rows(p.prec, in, out, a, p.line.pt, p.line.dtw.p, st));
            return CFFT_OK;
        }
        case Kind::FourStep: {
            if (in == out || p.half) {  // the middle in scratch (float for ComplexF16 data)
                std::lock_guard<std::mutex> g(p.scratch_mu);
                const size_t need = (size_t)B * p.n * esize(p.prec);
                if (p.scratch.bytes < need) {
                    int rc = p.scratch.alloc(need);
                    if (rc) return rc;
                }
                return run_fourstep(p, in, p.scratch.p, out, p.idist, p.n, p.odist, sgn, scale,
                                    epi, st, B);
            }
            return run_fourstep(p, in, out, out, p.idist, p.odist, p.odist, sgn, scale, epi, st, B);
        }
        case Kind::Large: {  // six steps per transform (plan_large)
            const FourStepPlan& fs = p.fs;
            const int64_t n = p.n, M = fs.N2, N1 = fs.N1;
            const size_t es = esize(p.prec);
            std::lock_guard<std::mutex> g(p.scratch_mu);
            if (p.scratch.bytes < (size_t)n * es) {
                int rc = p.scratch.alloc((size_t)n * es);
                if (rc) return rc;
            }
            for (int64_t b = 0; b < B; ++b) {
                const char* inb = (const char*)in + (size_t)b * p.idist * es;
                char* outb = (char*)out + (size_t)b * p.odist * es;
                // in place: pass A into the scratch, then out holds the rows and the scratch the
                // result, copied back at the end
                const bool inplace = inb == outb;
                char* mid = inplace ? (char*)p.scratch.p : outb;
                char* rowsb = inplace ? outb : (char*)p.scratch.p;
                FourStepAArgs a{};
                a.N1 = (int)N1;
                a.N2 = (int)M;
                a.C = fs.CA;
                a.Tt = fs.a.Tt;
                a.Ls = fs.a.Ls;
                a.E = fs.a.E;
                a.batch = 1;
                a.idist = n;
                a.odist = n;
                a.sgn_in = sgn;
                a.xcd = 1;
                CFFT_HIP(launch_fourstep_a(p.prec, inb, mid, a, fs.a.pt, fs.a.dtw.p, st));
                CFFT_HIP(launch_transpose(p.prec, mid, rowsb, M, N1, 1, fs.tw3.p, fs.tbits, 1.0,
                                          1.0, nullptr, nullptr, st));
                int rc = run_fft(*p.sub, rowsb, rowsb, 1, 1.0, ConvEpilogue{nullptr, nullptr}, st);
                if (rc) return rc;
                char* res = inplace ? (char*)p.scratch.p : outb;
                // a convolution's epilogue (dc: two slots per line) goes into the last step
                void* dcb = epi.dc ? (char*)epi.dc + (size_t)b * 2 * es : nullptr;
                CFFT_HIP(launch_transpose(p.prec, rowsb, res, N1, M, 1, nullptr, 0, scale,
                                          scale * sgn, epi.K, dcb, st));
                if (inplace)
                    CFFT_HIP(hipMemcpyAsync(outb, res, (size_t)n * es, hipMemcpyDeviceToDevice, st));
            }
            return CFFT_OK;
        }
        case Kind::ColsT: {  // B blocks of n x inner: transpose, row FFTs, transpose back
            const int64_t n = p.n, inner = p.idist / p.n;
            const size_t es = esize(p.prec), bytes = (size_t)B * n * inner * es;
            std::lock_guard<std::mutex> g(p.scratch_mu);
            if (p.scratch.bytes < bytes) {
                int rc = p.scratch.alloc(bytes);
                if (rc) return rc;
            }
            CFFT_HIP(launch_transpose(p.prec, in, p.scratch.p, n, inner, B, nullptr, 0, 1.0, 1.0,
                                      nullptr, nullptr, st));
            ExecShape rs;
            rs.batch = B * inner;
            int rc = run_fft_s(*p.sub, p.scratch.p, p.scratch.p, sgn, scale, epi, st, rs);
            if (rc) return rc;
            CFFT_HIP(launch_transpose(p.prec, p.scratch.p, out, inner, n, B, nullptr, 0, 1.0, 1.0,
                                      nullptr, nullptr, st));
            return CFFT_OK;
        }
        case Kind::Cols: {  // fs.N1 adjacent lines of length fs.N2 (stride N1), batch blocks
            FourStepBArgs b{};
            b.N1 = p.fs.N1;
            b.N2 = p.fs.N2;
            b.C = p.fs.CB;
            b.Tt = p.fs.b.Tt;
            b.Ls = p.fs.b.Ls;
            b.E = p.fs.b.E;
            b.batch = B;
            b.idist = p.idist;
            b.odist = p.odist;
            b.sgn_in = sgn;
            b.sgn_out = sgn;
            b.scale = scale;
            b.epi = epi;
            b.xcd = 1;
            CFFT_HIP(launch_fourstep_b(p.prec, in, out, b, p.fs.b.pt, p.fs.b.dtw.p, st));
            return CFFT_OK;
        }
        case Kind::ND: {
            const size_t nparts = p.parts.size();
            for (size_t i = 0; i < nparts; ++i) {
                const Plan& s = *p.parts[i];
                const bool last = i + 1 == nparts;
                int rc = run_fft(s, i == 0 ? in : out, out, sgn, last ? scale : 1.0, epi, st);
                if (rc) return rc;
            }
            return CFFT_OK;
        }
        case Kind::ConvSmall: {
            ConvSmallArgs a{};
            a.n = (int)p.n;
            a.L = (int)p.L;
            a.C = p.C;
            a.Tt = p.line.Tt;
            a.Ls = p.LsC;
            a.E = p.line.E;
            a.mode = (int)p.mode;
            a.nlines = B;
            a.idist = p.idist;
            a.odist = p.odist;
            a.sgn_in = sgn;
            a.sgn_out = sgn;
            a.scale = scale;
            a.perm = (const int*)p.perm.p;
            a.iperm = (const int*)p.iperm.p;
            a.chirp = p.chirp.p;
            a.K = p.K.p;
            CFFT_HIP(launch_conv_small(p.prec, in, out, a, p.line.pt, p.line.dtw.p, st));
            return CFFT_OK;
        }
        case Kind::ConvLarge: {
            void* w1 = sh.w1 ? sh.w1 : p.w1.p;
            void* w2 = sh.w1 ? sh.w2 : p.w2.p;
            void* dc = sh.w1 ? sh.dc : p.dc.p;
            if (!w1 || !w2 || !dc) return fail(CFFT_ERR_INVALID, "no convolution work buffers");
            if (p.convl) {
                const Plan& s = *p.sub;
                const FourStepPlan& f = s.fs;
                ConvLArgs a{};
                a.n = p.n;
                a.N1 = f.N1;
                a.N2 = f.N2;
                a.CA = p.convl_ca;
                a.CB = p.convl_cb;
                a.CC = p.convl_cc;
                a.TtA = f.a.Tt;
                a.LsA = p.convl_lsa;
                a.TtB = f.b.Tt;
                a.LsB = p.convl_lsb;
                a.TtC = f.a.Tt;
                a.LsC = p.convl_lsc;
                a.E = f.a.E;
                a.nlines = B;
                a.idist = p.idist;
                a.odist = p.odist;
                a.mode = (int)p.mode;
                a.sgn = sgn;
                a.scale = scale;
                a.perm = (const int*)p.perm.p;
                a.iperm = (const int*)p.iperm.p;
                a.chirp = p.chirp.p;
                a.K = p.K.p;
                a.m1 = w1;
                a.m2 = w2;
                a.dc = dc;
                a.tw3 = f.tw3.p;
                a.tbits = f.tbits;
                a.tnc = f.tnc;
                CFFT_HIP(launch_convl(p.prec, p.cprec, in, out, a, f.a.pt, f.a.dtw.p, f.b.pt,
                                      f.b.dtw.p, st));
                return CFFT_OK;
            }
            ConvGatherArgs g{};
            g.n = p.n;
            g.L = p.L;
            g.nlines = B;
            g.idist = p.idist;
            g.wdist = p.L;
            g.mode = (int)p.mode;
            g.sgn_in = sgn;
            g.perm = (const int*)p.perm.p;
            g.chirp = p.chirp.p;
            g.dc = dc;
            CFFT_HIP(launch_conv_gather(p.prec, p.cprec, in, w1, g, st));
            ExecShape ss;
            ss.batch = B;
            int rc = run_fft_s(*p.sub, w1, w2, 1, 1.0, ConvEpilogue{p.K.p, dc}, st, ss);
            if (rc) return rc;
            rc = run_fft_s(*p.sub, w2, w1, 1, 1.0, ConvEpilogue{nullptr, nullptr}, st, ss);
            if (rc) return rc;
            ConvScatterArgs s{};
            s.n = p.n;
            s.L = p.L;
            s.nlines = B;
            s.odist = p.odist;
            s.idist = p.idist;
            s.wdist = p.L;
            s.mode = (int)p.mode;
            s.sgn_in = sgn;
            s.sgn_out = sgn;
            s.invL = 1.0 / (double)p.L;
            s.scale = scale;
            s.iperm = (const int*)p.iperm.p;
            s.chirp = p.chirp.p;
            s.dc = dc;
            CFFT_HIP(launch_conv_scatter(p.prec, p.cprec, w1, out, s, st));
            return CFFT_OK;
        }
        default: break;
    }
    return fail(CFFT_ERR_UNSUPPORTED, "plan kind not executable");
}

static int real_exec(const Plan& p, const void* in, void* out, int direction, int normalize,
                     hipStream_t st);

int plan_exec_shaped(const Plan& p, const void* in, void* out, int direction, int normalize,
                     hipStream_t st, const ExecShape& sh) {
    if (p.kind == Kind::Uneven || p.real || p.kind == Kind::ND || p.kind == Kind::Cols || p.half ||
        p.kind == Kind::Large)
        return fail(CFFT_ERR_UNSUPPORTED, "shaped execution needs a 1-D float/double plan");
    if (sh.batch <= 0) return CFFT_OK;
    if (!in || !out) return fail(CFFT_ERR_INVALID, "null buffer");
    if (in == out && p.kind == Kind::FourStep)
        return fail(CFFT_ERR_UNSUPPORTED, "shaped four-step execution is out of place");
    const int sgn = direction > 0 ? -1 : 1;
    const double scale = normalize ? 1.0 / (double)p.n : 1.0;
    return run_fft_s(p, in, out, sgn, scale, ConvEpilogue{nullptr, nullptr}, st, sh);
}

// ComplexF16 data on a plan kind without fused half-precision kernels: the input converted to
// float in a work buffer, the float transform, the result rounded into the output.
static int half_exec_staged(const Plan& p, const void* in, void* out, int sgn, double scale,
                            hipStream_t st) {
    const bool nd = p.kind == Kind::ND;
    const int64_t rows = nd ? 1 : p.batch, n = nd ? p.idist : p.n;
    const int64_t idist = p.idist, odist = p.odist;
    std::lock_guard<std::mutex> g(p.hbuf_mu);
    const size_t ib = (size_t)rows * idist * 8, ob = (size_t)rows * odist * 8;
    int rc;
    if (p.hbuf_in.bytes < ib && (rc = p.hbuf_in.alloc(ib))) return rc;
    if (p.hbuf_out.bytes < ob && (rc = p.hbuf_out.alloc(ob))) return rc;
    CFFT_HIP(launch_convert_h2f(in, p.hbuf_in.p, rows, n, idist, idist, st));
    if ((rc = run_fft(p, p.hbuf_in.p, p.hbuf_out.p, sgn, scale, ConvEpilogue{nullptr, nullptr}, st)))
        return rc;
    CFFT_HIP(launch_convert_f2h(p.hbuf_out.p, out, rows, n, odist, odist, st));
    return CFFT_OK;
}

// Whether an execution uses buffers the plan owns (scratch middles, convolution work lines,
// staging for ComplexF16, real-transform temporaries, uneven tables).
static bool uses_work_buffers(const Plan& p, const void* in, const void* out) {
    switch (p.kind) {
        case Kind::FourStep: return in == out || p.half;
        case Kind::Large:
        case Kind::ConvLarge:
        case Kind::Uneven:
        case Kind::ND: return true;
        default: return p.real >= 2 || (p.half && p.kind != Kind::Direct && p.kind != Kind::Trivial);
    }
}

static int plan_exec_body(const Plan& p, const void* in, void* out, int direction, int normalize,
                          hipStream_t st);

// Executions of one plan on different streams may overlap on the device, except those that use
// the plan's work buffers: they are ordered (each waits on the device for the previous one).
int plan_exec(const Plan& p, const void* in, void* out, int direction, int normalize,
              hipStream_t st) {
    if (!uses_work_buffers(p, in, out)) return plan_exec_body(p, in, out, direction, normalize, st);
    std::lock_guard<std::mutex> g(p.exec_mu);
    if (p.work_ev) CFFT_HIP(hipStreamWaitEvent(st, p.work_ev, 0));
    const int rc = plan_exec_body(p, in, out, direction, normalize, st);
    if (rc) return rc;
    if (!p.work_ev) CFFT_HIP(hipEventCreateWithFlags(&p.work_ev, hipEventDisableTiming));
    CFFT_HIP(hipEventRecord(p.work_ev, st));
    return CFFT_OK;
}

static int plan_exec_body(const Plan& p, const void* in, void* out, int direction, int normalize,
                          hipStream_t st) {
    if (p.kind == Kind::Uneven) return uneven_exec(p, in, out, direction, normalize, st);
    if (p.real) return real_exec(p, in, out, direction, normalize, st);
    if (p.batch == 0) return CFFT_OK;
    if (!in || !out) return fail(CFFT_ERR_INVALID, "null buffer");
    const int sgn = direction > 0 ? -1 : 1;
    const double scale = normalize ? 1.0 / (double)p.n : 1.0;
    if (p.half && p.kind != Kind::Direct && p.kind != Kind::FourStep && p.kind != Kind::Trivial)
        return half_exec_staged(p, in, out, sgn, scale, st);
    return run_fft(p, in, out, sgn, scale, ConvEpilogue{nullptr, nullptr}, st);
}

// Execution without any allocation (cfft_exec_noalloc): plan kinds whose kernels need no work
// buffer run as plan_exec runs them; a four-step whose middle needs a work buffer (in place, or
// ComplexF16 data) takes it from the static work area instead of the plan's lazily allocated
// scratch.  Other kinds hold work buffers of their own and are refused.  No event or buffer is
// created here, and executions on one GPU share its work area (overlapping calls race).
static int plan_exec_noalloc(const Plan& p, const void* in, void* out, int direction,
                             int normalize, hipStream_t st) {
    if (p.kind == Kind::Uneven || p.real || !(p.kind == Kind::Trivial || p.kind == Kind::Direct ||
                                              p.kind == Kind::FourStep ||
                                              (p.kind == Kind::ConvSmall && !p.half)))
        return fail(CFFT_ERR_UNSUPPORTED,
                    "cfft_exec_noalloc: this plan kind owns work buffers (only trivial, direct, "
                    "four-step and fused-convolution 1-D plans run without allocating)");
    if (p.batch == 0) return CFFT_OK;
    if (!in || !out) return fail(CFFT_ERR_INVALID, "null buffer");
    const int sgn = direction > 0 ? -1 : 1;
    const double scale = normalize ? 1.0 / (double)p.n : 1.0;
    const ConvEpilogue none{nullptr, nullptr};
    if (p.kind != Kind::FourStep || (in != out && !p.half))
        return run_fft(p, in, out, sgn, scale, none, st);
    const size_t need = (size_t)p.batch * (size_t)p.n * (size_t)esize(p.prec);
    if (need > kNoAllocScratchBytes)
        return fail(CFFT_ERR_UNSUPPORTED,
                    "cfft_exec_noalloc: the four-step middle needs " + std::to_string(need) +
                        " bytes; the work area holds " + std::to_string(kNoAllocScratchBytes));
    std::string err;
    void* work = noalloc_scratch_ptr(p.ctx->device, &err);
    if (!work) return fail(CFFT_ERR_HIP, "cfft_exec_noalloc: " + err);
    return run_fourstep(p, in, work, out, p.idist, p.n, p.odist, sgn, scale, none, st, p.batch);
}

// ------------------------------------------------------------------------------- N-D
// Multi-dimensional transforms over the axes in axes_mask (bit a = axis a, row-major dims, the
// last axis contiguous): the role of plan_fft's `region` (CatabraFFT.jl:169, accepted there and
// ignored).  The contiguous axis runs as a batch of 1-D transforms; every other axis runs as
// four-step pass B over its adjacent strided lines, in place on the output.
int plan_nd_create(Ctx* ctx, int rank, const int64_t* dims, int64_t axes_mask, int dtype,
                   int flags, std::shared_ptr<Plan>* out) {
    if (rank < 1 || rank > 8 || !dims) return fail(CFFT_ERR_INVALID, "rank must be 1..8");
    if (dtype != CFFT_C32 && dtype != CFFT_C64 && dtype != CFFT_C16)
        return fail(CFFT_ERR_INVALID, "bad dtype");
    if (flags < CFFT_NO_FLAG || flags > CFFT_ENCHANT) return fail(CFFT_ERR_INVALID, "bad flags");
    int64_t total = 1, ntrans = 1;
    for (int a = 0; a < rank; ++a) {
        if (dims[a] < 1) return fail(CFFT_ERR_INVALID, "dims must be >= 1");
        total *= dims[a];
        if ((axes_mask >> a) & 1) ntrans *= dims[a];
    }
    CFFT_HIP(hipSetDevice(ctx->device));
    auto p = std::make_shared<Plan>();
    p->ctx = ctx;
    p->kind = Kind::ND;
    p->n = ntrans;  // normalisation divides by the product of the transformed lengths
    p->batch = 1;
    p->idist = p->odist = total;
    p->dtype = dtype;
    p->flags = flags;
    p->prec = dtype == CFFT_C64 ? Prec::F64 : Prec::F32;
    p->half = dtype == CFFT_C16;  // float parts between two conversions (half_exec_staged)
    const int sdtype = dtype == CFFT_C16 ? CFFT_C32 : dtype;
    std::ostringstream os;
    os << "nd dims=[";
    for (int a = 0; a < rank; ++a) os << (a ? "," : "") << dims[a];
    os << "] axes=0x" << std::hex << axes_mask << std::dec;
    int rc;
    for (int a = rank - 1; a >= 0; --a) {
        if (!((axes_mask >> a) & 1) || dims[a] == 1) continue;
        int64_t inner = 1, outer = 1;
        for (int b = a + 1; b < rank; ++b) inner *= dims[b];
        for (int b = 0; b < a; ++b) outer *= dims[b];
        std::unique_ptr<Plan> s(new Plan());
        s->ctx = ctx;
        s->prec = p->prec;
        s->dtype = sdtype;
        s->flags = flags;
        if (inner == 1) {  // contiguous axis: any 1-D plan
            s->n = dims[a];
            s->batch = outer;
            s->idist = s->odist = dims[a];
            PlanOptions o;
            if ((rc = plan_build(ctx, *s, o))) return rc;
            os << " | axis " << a << ": " << describe(*s);
        } else if (!line_fits(ctx, dims[a], p->prec, 16) || inner > (int64_t)1 << 30) {
            // a strided axis that one workgroup cannot take (not smooth, or too long): each
            // outer block (dims[a] x inner) is transposed so that the axis is contiguous, one
            // batched 1-D plan of any kind runs over the outer * inner rows, and the rows are
            // transposed back
            const int64_t N = dims[a];
            s->kind = Kind::ColsT;
            s->n = N;
            s->batch = outer;
            s->idist = s->odist = inner * N;
            s->fs.N1 = (int)std::min<int64_t>(inner, INT32_MAX);
            s->sub.reset(new Plan());
            Plan& r = *s->sub;
            r.ctx = ctx;
            r.n = N;
            r.batch = outer * inner;
            r.idist = r.odist = N;
            r.prec = p->prec;
            r.dtype = sdtype;
            r.flags = flags;
            PlanOptions o;
            if ((rc = plan_build(ctx, r, o))) return rc;
            os << " | axis " << a << ": transposed lines N=" << N << " x " << inner << " {"
               << describe(r) << "}";
        } else {
            const int64_t N2 = dims[a];
            s->kind = Kind::Cols;
            s->n = N2;
            s->batch = outer;
            s->idist = s->odist = inner * N2;
            s->fs.N1 = (int)inner;
            s->fs.N2 = (int)N2;
            if ((rc = make_line(s->fs.b, N2, p->prec, 16, 1024))) return rc;
            s->fs.CB = pick_tile(ctx, s->fs.b, p->prec, p->prec == Prec::F32 ? 16 : 8);
            s->fs.b.Ls = fs_line_stride(ctx, p->prec, s->fs.b, s->fs.CB, 16);
            os << " | axis " << a << ": strided lines N=" << N2 << " x " << inner << " C="
               << s->fs.CB;
        }
        p->parts.push_back(std::move(s));
    }
    if (p->parts.empty()) {  // nothing to transform: a copy
        std::unique_ptr<Plan> s(new Plan());
        s->ctx = ctx;
        s->prec = p->prec;
        s->kind = Kind::Trivial;
        s->n = 1;
        s->batch = total;
        s->idist = s->odist = 1;
        p->parts.push_back(std::move(s));
    }
    if (p->half) os << " io=f16";
    p->desc = os.str();
    *out = p;
    return CFFT_OK;
}

// ------------------------------------------------------------------------------- real input
// rfft / brfft / irfft (AbstractFFTs' real transforms; the reference's plan_rfft / plan_brfft
// stubs at CatabraFFT.jl:199-208 build nothing usable).  Even n reads the reals as n/2 complex
// values and untangles the spectrum with w^k = e^{-2 pi i k/n}; odd n goes through a full
// complex transform.
int plan_real_create(Ctx* ctx, int64_t n, int64_t batch, int64_t rdist, int64_t cdist, int dtype,
                     int flags, std::shared_ptr<Plan>* out) {
    if (n < 2) return fail(CFFT_ERR_INVALID, "real transforms need n >= 2");
    if (batch < 0) return fail(CFFT_ERR_INVALID, "batch must be >= 0");
    if (dtype != CFFT_C32 && dtype != CFFT_C64) return fail(CFFT_ERR_INVALID, "bad dtype");
    if (flags < CFFT_NO_FLAG || flags > CFFT_ENCHANT) return fail(CFFT_ERR_INVALID, "bad flags");
    if (rdist < n || cdist < n / 2 + 1)
        return fail(CFFT_ERR_INVALID, "rdist must be >= n and cdist >= n/2 + 1");
    CFFT_HIP(hipSetDevice(ctx->device));
    auto p = std::make_shared<Plan>();
    p->ctx = ctx;
    p->n = n;
    p->batch = batch;
    p->idist = rdist;
    p->odist = cdist;
    p->dtype = dtype;
    p->flags = flags;
    p->prec = dtype == CFFT_C32 ? Prec::F32 : Prec::F64;
    const int64_t m = n / 2;
    const size_t es = esize(p->prec);
    int rc;
    auto sub = [&](int64_t len, int64_t idist, int64_t odist, std::unique_ptr<Plan>& s) {
        s.reset(new Plan());
        s->ctx = ctx;
        s->n = len;
        s->batch = batch;
        s->idist = idist;
        s->odist = odist;
        s->prec = p->prec;
        s->dtype = dtype;
        s->flags = flags;
        PlanOptions o;
        return plan_build(ctx, *s, o);
    };
    if (n % 2 == 0 && rdist % 2 == 0) {
        std::vector<cld> w(m);
        const OmegaTable om(n);
        for (int64_t k = 0; k < m; ++k) w[k] = om(k);
        if ((rc = upload(p->rw, w, p->prec))) return rc;
        if (m >= 2 && line_fits(ctx, m, p->prec, 16)) {
            if ((rc = make_line(p->line, m, p->prec, 16, 1024))) return rc;
            p->C = rows_lines_per_block(p->line, p->prec, batch, ctx->lds_rows_budget, 256);
            p->real = 1;
        } else {
            if ((rc = sub(m, rdist / 2, m, p->rsub_fwd))) return rc;
            if ((rc = sub(m, m, rdist / 2, p->rsub_bwd))) return rc;
            if ((rc = p->rs1.alloc((size_t)std::max<int64_t>(batch, 1) * m * es))) return rc;
            p->real = 2;
        }
    } else {
        if ((rc = sub(n, n, n, p->rsub_fwd))) return rc;
        if ((rc = p->rs1.alloc((size_t)std::max<int64_t>(batch, 1) * n * es))) return rc;
        if ((rc = p->rs2.alloc((size_t)std::max<int64_t>(batch, 1) * n * es))) return rc;
        p->real = 3;
    }
    p->kind = Kind::Direct;
    std::ostringstream os;
    os << "real n=" << n << " batch=" << batch << " prec=" << (p->prec == Prec::F32 ? "f32" : "f64")
       << (p->real == 1 ? " fused half-length " : p->real == 2 ? " half-length via " : " full-length via ");
    if (p->real == 1) {
        os << "[";
        for (int i = 0; i < p->line.pt.npass; ++i) os << (i ? "," : "") << p->line.pt.radix[i];
        os << "] C=" << p->C;
    } else {
        os << "{" << plan_description(*p->rsub_fwd) << "}";
    }
    p->desc = os.str();
    *out = p;
    return CFFT_OK;
}

static int real_aux(const Plan& p, int mode, const void* in, void* out, int64_t count,
                    int64_t idist, int64_t odist, double scale, hipStream_t st) {
    RealAuxArgs a{};
    a.mode = mode;
    a.n = p.n;
    a.m = p.n / 2;
    a.nlines = p.batch;
    a.count = count;
    a.idist = idist;
    a.odist = odist;
    a.scale = scale;
    a.w = p.rw.p;
    CFFT_HIP(launch_real_aux(p.prec, in, out, a, st));
    return CFFT_OK;
}

static int real_exec(const Plan& p, const void* in, void* out, int direction, int normalize,
                     hipStream_t st) {
    if (p.batch == 0) return CFFT_OK;
    if (!in || !out) return fail(CFFT_ERR_INVALID, "null buffer");
    if (in == out) return fail(CFFT_ERR_INVALID, "real transforms are out of place");
    const bool fwd = direction < 0;
    const double scale = (!fwd && normalize) ? 1.0 / (double)p.n : 1.0;
    const int64_t n = p.n, m = n / 2, rdist = p.idist, cdist = p.odist;
    const ConvEpilogue none{nullptr, nullptr};
    if (p.real == 1) {
        RealArgs a{};
        a.m = (int)m;
        a.C = p.C;
        a.Tt = p.line.Tt;
        a.Ls = p.line.Ls;
        a.nlines = p.batch;
        a.rdist = rdist;
        a.cdist = cdist;
        a.forward = fwd ? 1 : 0;
        a.scale = scale;
        a.w = p.rw.p;
        CFFT_HIP(launch_real_rows(p.prec, in, out, a, p.line.pt, p.line.dtw.p, p.line.E, st));
        return CFFT_OK;
    }
    int rc;
    if (p.real == 2) {
        if (fwd) {
            if ((rc = run_fft(*p.rsub_fwd, in, p.rs1.p, 1, 1.0, none, st))) return rc;
            return real_aux(p, 0, p.rs1.p, out, m / 2 + 1, m, cdist, 1.0, st);
        }
        if ((rc = real_aux(p, 1, in, p.rs1.p, m, cdist, m, scale, st))) return rc;
        return run_fft(*p.rsub_bwd, p.rs1.p, out, -1, 1.0, none, st);
    }
    if (fwd) {
        if ((rc = real_aux(p, 2, in, p.rs1.p, n, rdist, n, 1.0, st))) return rc;
        if ((rc = run_fft(*p.rsub_fwd, p.rs1.p, p.rs2.p, 1, 1.0, none, st))) return rc;
        return real_aux(p, 3, p.rs2.p, out, n / 2 + 1, n, cdist, 1.0, st);
    }
    if ((rc = real_aux(p, 4, in, p.rs1.p, n, cdist, n, scale, st))) return rc;
    if ((rc = run_fft(*p.rsub_fwd, p.rs1.p, p.rs2.p, -1, 1.0, none, st))) return rc;
    return real_aux(p, 5, p.rs2.p, out, n, n, rdist, 1.0, st);
}

size_t plan_device_bytes(const Plan& p) {
    size_t b = p.line.dtw.bytes + p.fs.a.dtw.bytes + p.fs.b.dtw.bytes + p.fs.tw3.bytes +
               p.scratch.bytes + p.perm.bytes + p.iperm.bytes + p.chirp.bytes +
               p.K.bytes + p.dc.bytes + p.w1.bytes + p.w2.bytes;
    if (p.sub) b += plan_device_bytes(*p.sub);
    b += p.rw.bytes + p.rs1.bytes + p.rs2.bytes;
    if (p.rsub_fwd) b += plan_device_bytes(*p.rsub_fwd);
    if (p.rsub_bwd) b += plan_device_bytes(*p.rsub_bwd);
    for (const auto& s : p.parts) b += plan_device_bytes(*s);
    b += p.uneven_bytes + p.hbuf_in.bytes + p.hbuf_out.bytes;
    return b;
}

int plan_create(Ctx* ctx, int64_t n, int64_t batch, int64_t idist, int64_t odist, int dtype,
                int flags, const PlanOptions& opts, std::shared_ptr<Plan>* out) {
    if (n < 1) return fail(CFFT_ERR_INVALID, "n must be >= 1");
    if (batch < 0) return fail(CFFT_ERR_INVALID, "batch must be >= 0");
    if (dtype != CFFT_C32 && dtype != CFFT_C64 && dtype != CFFT_C16)
        return fail(CFFT_ERR_INVALID, "bad dtype");
    if (flags < CFFT_NO_FLAG || flags > CFFT_ENCHANT) return fail(CFFT_ERR_INVALID, "bad flags");
    if (idist < n || odist < n) return fail(CFFT_ERR_INVALID, "idist/odist must be >= n");
    if (n > ((int64_t)1 << 30)) return fail(CFFT_ERR_UNSUPPORTED, "n above 2^30");
    auto p = std::make_shared<Plan>();
    p->ctx = ctx;
    p->n = n;
    p->batch = batch;
    p->idist = idist;
    p->odist = odist;
    p->dtype = dtype;
    p->flags = flags;
    p->prec = dtype == CFFT_C64 ? Prec::F64 : Prec::F32;
    CFFT_HIP(hipSetDevice(ctx->device));
    int rc;
    if (flags == CFFT_NO_FLAG || opts.explicit_opts) {
        rc = plan_fft_core(ctx, *p, opts);
    } else {
        rc = autotune(ctx, *p, flags, opts);
    }
    if (rc) return rc;
    if (dtype == CFFT_C16) {  // the float plan with half-precision lines (plain rows kernel)
        p->half = true;
        p->pipe_res = 0;
        p->fs.pipe_a = p->fs.pipe_b = 0;
        p->fs.pair_a = p->fs.pair_b = 0;
        p->desc.clear();
    }
    if (p->desc.empty()) p->desc = describe(*p);
    *out = p;
    return CFFT_OK;
}

std::string plan_description(const Plan& p) { return p.desc.empty() ? describe(p) : p.desc; }

// ------------------------------------------------------------------------------- context
int ctx_create(int device, Ctx** out) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        (void)hipGetLastError();
        return fail(CFFT_ERR_HIP, "no HIP device available");
    }
    if (device < 0 || device >= count) return fail(CFFT_ERR_INVALID, "bad device index");
    CFFT_HIP(hipSetDevice(device));
    std::unique_ptr<Ctx> c(new Ctx());
    c->device = device;
    int v = 0;
    CFFT_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
    c->lds_max = std::min<size_t>((size_t)v, 160 * 1024);
    CFFT_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, device));
    c->num_cu = v;
    CFFT_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    init_kernels();
    noalloc_scratch_setup(device);
    *out = c.release();
    return CFFT_OK;
}

void ctx_destroy(Ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    c->cache.clear();
    c->stage_in.reset();
    c->stage_out.reset();
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

}  // namespace cfft

// =============================================================================== C ABI
using namespace cfft;

struct cfft_plan_s {
    std::shared_ptr<Plan> p;
    cfft_ctx_s* ctx;
};

extern "C" {

const char* cfft_last_error(void) { return g_err.c_str(); }

const char* cfft_version(void) { return "catabrafft-hip 0.1 (gfx950)"; }

int cfft_init(int device, cfft_ctx* ctx) {
    if (!ctx) return fail(CFFT_ERR_INVALID, "null ctx pointer");
    Ctx* c = nullptr;
    int rc = ctx_create(device, &c);
    if (rc) return rc;
    *ctx = new cfft_ctx_s{c};
    g_err.clear();
    return CFFT_OK;
}

int cfft_finalize(cfft_ctx ctx) {
    if (!ctx) return fail(CFFT_ERR_INVALID, "null ctx");
    ctx_destroy(ctx->c);
    delete ctx;
    return CFFT_OK;
}

int cfft_plan_1d(cfft_ctx ctx, int64_t n, int64_t batch, int64_t idist, int64_t odist, int dtype,
                 int flags, cfft_plan* plan) {
    if (!ctx || !plan) return fail(CFFT_ERR_INVALID, "null argument");
    Ctx* c = ctx->c;
    PlanKey key{n, batch, idist, odist, dtype, flags};
    std::shared_ptr<Plan> p;
    {
        std::lock_guard<std::mutex> g(c->mu);
        auto it = c->cache.find(key);
        if (it != c->cache.end()) p = it->second;
    }
    if (!p) {
        PlanOptions o;
        int rc = plan_create(c, n, batch, idist, odist, dtype, flags, o, &p);
        if (rc) return rc;
        std::lock_guard<std::mutex> g(c->mu);
        c->cache[key] = p;
    }
    *plan = new cfft_plan_s{p, ctx};
    return CFFT_OK;
}

int cfft_plan_nd(cfft_ctx ctx, int rank, const int64_t* dims, int64_t axes_mask, int dtype,
                 int flags, cfft_plan* plan) {
    if (!ctx || !plan) return fail(CFFT_ERR_INVALID, "null argument");
    std::shared_ptr<Plan> p;
    int rc = plan_nd_create(ctx->c, rank, dims, axes_mask, dtype, flags, &p);
    if (rc) return rc;
    *plan = new cfft_plan_s{p, ctx};
    return CFFT_OK;
}

int cfft_plan_1d_real(cfft_ctx ctx, int64_t n, int64_t batch, int64_t rdist, int64_t cdist,
                      int dtype, int flags, cfft_plan* plan) {
    if (!ctx || !plan) return fail(CFFT_ERR_INVALID, "null argument");
    std::shared_ptr<Plan> p;
    int rc = plan_real_create(ctx->c, n, batch, rdist, cdist, dtype, flags, &p);
    if (rc) return rc;
    *plan = new cfft_plan_s{p, ctx};
    return CFFT_OK;
}

int cfft_plan_many_uneven(cfft_ctx ctx, int64_t count, const int64_t* lengths,
                          const int64_t* offsets, int dtype, int flags, cfft_plan* plan) {
    if (!ctx || !plan || (count > 0 && (!lengths || !offsets)))
        return fail(CFFT_ERR_INVALID, "null argument");
    std::shared_ptr<Plan> p;
    int rc = uneven_create(ctx->c, count, lengths, offsets, dtype, flags, &p);
    if (rc) return rc;
    *plan = new cfft_plan_s{p, ctx};
    return CFFT_OK;
}

int cfft_exec(cfft_plan plan, const void* in, void* out, int direction, int normalize,
              void* stream) {
    if (!plan) return fail(CFFT_ERR_INVALID, "null plan");
    if (direction != CFFT_FORWARD && direction != CFFT_BACKWARD)
        return fail(CFFT_ERR_INVALID, "direction must be -1 or +1");
    CFFT_HIP(hipSetDevice(plan->p->ctx->device));
    return plan_exec(*plan->p, in, out, direction, normalize, (hipStream_t)stream);
}

int cfft_exec_noalloc(cfft_plan plan, const void* in, void* out, int direction, int normalize,
                      void* stream) {
    if (!plan) return fail(CFFT_ERR_INVALID, "null plan");
    if (direction != CFFT_FORWARD && direction != CFFT_BACKWARD)
        return fail(CFFT_ERR_INVALID, "direction must be -1 or +1");
    CFFT_HIP(hipSetDevice(plan->p->ctx->device));
    return plan_exec_noalloc(*plan->p, in, out, direction, normalize, (hipStream_t)stream);
}

int cfft_exec_host(cfft_plan plan, const void* in, void* out, int direction, int normalize) {
    if (!plan || !in || !out) return fail(CFFT_ERR_INVALID, "null argument");
    if (direction != CFFT_FORWARD && direction != CFFT_BACKWARD)
        return fail(CFFT_ERR_INVALID, "direction must be -1 or +1");
    const Plan& p = *plan->p;
    Ctx* c = p.ctx;
    CFFT_HIP(hipSetDevice(c->device));
    const int64_t es = io_esize(p);
    // input and output spans, and the output rows the transform writes (rows of `row` bytes,
    // `pitch` bytes apart): the gaps between output rows keep the caller's contents
    size_t ib, ob, rows = 1, row = 0, pitch = 0;
    if (p.kind == Kind::Uneven) {
        ib = ob = (size_t)p.uneven_span * es;
    } else if (p.kind == Kind::ND) {
        ib = ob = (size_t)p.idist * es;
    } else if (p.real) {
        if (p.batch == 0) return CFFT_OK;
        const size_t rb = (size_t)((p.batch - 1) * p.idist + p.n) * (es / 2);
        const size_t cb = (size_t)((p.batch - 1) * p.odist + p.n / 2 + 1) * es;
        ib = direction < 0 ? rb : cb;
        ob = direction < 0 ? cb : rb;
        rows = (size_t)p.batch;
        row = direction < 0 ? (size_t)(p.n / 2 + 1) * es : (size_t)p.n * (es / 2);
        pitch = direction < 0 ? (size_t)p.odist * es : (size_t)p.idist * (es / 2);
    } else {
        if (p.batch == 0) return CFFT_OK;
        ib = (size_t)((p.batch - 1) * p.idist + p.n) * es;
        ob = (size_t)((p.batch - 1) * p.odist + p.n) * es;
        rows = (size_t)p.batch;
        row = (size_t)p.n * es;
        pitch = (size_t)p.odist * es;
    }
    const bool whole = rows == 1 || row == pitch || p.kind == Kind::Uneven || p.kind == Kind::ND;
    auto copy_out = [&](const void* from) {  // host copy of the written rows
        if (whole) {
            std::memcpy(out, from, ob);
            return;
        }
        for (size_t r = 0; r < rows; ++r)
            std::memcpy((char*)out + r * pitch, (const char*)from + r * pitch, row);
    };
    std::lock_guard<std::mutex> g(c->stage_mu);
    int rc;
    // Small plans of one kernel that reads its input once and writes its output once (direct
    // rows, fused convolutions) run on fine-grained pinned copies of the host buffers, read and
    // written by the kernel over PCIe: no DMA transfers, whose latency dominates small transforms.
    constexpr size_t kZeroCopyMax = (size_t)256 << 10;
    if ((p.kind == Kind::Direct || p.kind == Kind::ConvSmall) && !p.real &&
        std::max(ib, ob) <= kZeroCopyMax && c->zc_in.reserve(ib) == 0 &&
        c->zc_out.reserve(ob) == 0) {
        void *din = nullptr, *dout = nullptr;
        CFFT_HIP(hipHostGetDevicePointer(&din, c->zc_in.p, 0));
        CFFT_HIP(hipHostGetDevicePointer(&dout, c->zc_out.p, 0));
        std::memcpy(c->zc_in.p, in, ib);
        rc = plan_exec(p, din, dout, direction, normalize, c->stream);
        if (rc) return rc;
        CFFT_HIP(hipStreamSynchronize(c->stream));
        copy_out(c->zc_out.p);
        return CFFT_OK;
    }
    if (c->stage_in.bytes < ib && (rc = c->stage_in.alloc(ib))) return rc;
    if (c->stage_out.bytes < ob && (rc = c->stage_out.alloc(ob))) return rc;
    // Transfers up to kPinnedMax go through pinned copies (a host memcpy each way): the DMA then
    // runs asynchronously instead of through the runtime's pageable staging, which dominates the
    // latency of small transforms.
    constexpr size_t kPinnedMax = (size_t)4 << 20;
    const bool pinned = std::max(ib, ob) <= kPinnedMax && c->pin_in.reserve(ib) == 0 &&
                        c->pin_out.reserve(ob) == 0;
    const void* src = in;
    if (pinned) {
        std::memcpy(c->pin_in.p, in, ib);
        src = c->pin_in.p;
    }
    CFFT_HIP(hipMemcpyAsync(c->stage_in.p, src, ib, hipMemcpyHostToDevice, c->stream));
    if (p.kind == Kind::Uneven)  // untouched gaps come back as they went in
        CFFT_HIP(hipMemcpyAsync(c->stage_out.p, src, ob, hipMemcpyHostToDevice, c->stream));
    rc = plan_exec(p, c->stage_in.p, c->stage_out.p, direction, normalize, c->stream);
    if (rc) return rc;
    if (pinned) {
        CFFT_HIP(hipMemcpyAsync(c->pin_out.p, c->stage_out.p, ob, hipMemcpyDeviceToHost,
                                c->stream));
        CFFT_HIP(hipStreamSynchronize(c->stream));
        copy_out(c->pin_out.p);
    } else {
        if (whole)
            CFFT_HIP(hipMemcpyAsync(out, c->stage_out.p, ob, hipMemcpyDeviceToHost, c->stream));
        else
            CFFT_HIP(hipMemcpy2DAsync(out, pitch, c->stage_out.p, pitch, row, rows,
                                      hipMemcpyDeviceToHost, c->stream));
        CFFT_HIP(hipStreamSynchronize(c->stream));
    }
    return CFFT_OK;
}

int cfft_plan_device_bytes(cfft_plan plan, int64_t* bytes) {
    if (!plan || !bytes) return fail(CFFT_ERR_INVALID, "null argument");
    *bytes = (int64_t)plan_device_bytes(*plan->p);
    return CFFT_OK;
}

int cfft_plan_describe(cfft_plan plan, char* buf, int64_t buflen) {
    if (!plan || !buf || buflen <= 0) return fail(CFFT_ERR_INVALID, "null argument");
    const std::string d = plan_description(*plan->p);
    std::snprintf(buf, (size_t)buflen, "%s", d.c_str());
    return CFFT_OK;
}

int cfft_plan_destroy(cfft_plan plan) {
    if (!plan) return fail(CFFT_ERR_INVALID, "null plan");
    (void)hipSetDevice(plan->p->ctx->device);
    delete plan;
    return CFFT_OK;
}

int cfft_empty_cache(cfft_ctx ctx) {
    if (!ctx) return fail(CFFT_ERR_INVALID, "null ctx");
    Ctx* c = ctx->c;
    (void)hipSetDevice(c->device);
    {
        std::lock_guard<std::mutex> g(c->mu);
        c->cache.clear();
    }
    std::lock_guard<std::mutex> g(c->stage_mu);
    c->stage_in.reset();
    c->stage_out.reset();
    return CFFT_OK;
}

int cfft_wisdom_export(cfft_ctx ctx, const char* path) {
    if (!ctx || !path) return fail(CFFT_ERR_INVALID, "null argument");
    return wisdom_export(ctx->c, path);
}

int cfft_wisdom_import(cfft_ctx ctx, const char* path) {
    if (!ctx || !path) return fail(CFFT_ERR_INVALID, "null argument");
    const int n = wisdom_import(ctx->c, path);
    return n < 0 ? n : CFFT_OK;
}

int cfft_malloc(cfft_ctx ctx, int64_t bytes, void** ptr) {
    if (!ctx || !ptr || bytes < 0) return fail(CFFT_ERR_INVALID, "bad argument");
    CFFT_HIP(hipSetDevice(ctx->c->device));
    if (hipMalloc(ptr, (size_t)bytes) != hipSuccess) {
        (void)hipGetLastError();
        *ptr = nullptr;
        return fail(CFFT_ERR_NOMEM, "hipMalloc failed");
    }
    return CFFT_OK;
}

int cfft_free(cfft_ctx ctx, void* ptr) {
    if (!ctx) return fail(CFFT_ERR_INVALID, "null ctx");
    CFFT_HIP(hipSetDevice(ctx->c->device));
    CFFT_HIP(hipFree(ptr));
    return CFFT_OK;
}

int cfft_memcpy_h2d(cfft_ctx ctx, void* dst, const void* src, int64_t bytes) {
    if (!ctx || bytes < 0) return fail(CFFT_ERR_INVALID, "bad argument");
    CFFT_HIP(hipSetDevice(ctx->c->device));
    CFFT_HIP(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice));
    return CFFT_OK;
}

int cfft_memcpy_d2h(cfft_ctx ctx, void* dst, const void* src, int64_t bytes) {
    if (!ctx || bytes < 0) return fail(CFFT_ERR_INVALID, "bad argument");
    CFFT_HIP(hipSetDevice(ctx->c->device));
    CFFT_HIP(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost));
    return CFFT_OK;
}

int cfft_synchronize(cfft_ctx ctx) {
    if (!ctx) return fail(CFFT_ERR_INVALID, "null ctx");
    CFFT_HIP(hipSetDevice(ctx->c->device));
    CFFT_HIP(hipDeviceSynchronize());
    return CFFT_OK;
}

}  // extern "C"
