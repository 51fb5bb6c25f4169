// library.h -- host-side types shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <complex>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "kernels.h"

namespace cfft {

int fail(int code, const std::string& msg);

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf();
    int alloc(size_t nbytes);
    void reset();
};

// One line FFT of length N: radix sequence, threads per line, LDS line stride, twiddles.
struct LinePlan {
    int N = 0, E = 16, Tt = 1, Ls = 1;
    int pad = 1;  // LDS lines with pad words (0: unpadded, direct lines on the plain rows kernel)
    PassTable pt{};
    DevBuf dtw;
};

struct FourStepPlan {
    int N1 = 0, N2 = 0, CA = 1, CB = 1;
    LinePlan a, b;
    // w_n^e = (A[e & m] * B[(e >> tbits) & m]) * C[e >> 2 tbits], m = 2^tbits - 1: the three
    // tables back to back in tw3 (2^tbits, 2^tbits and tnc entries), staged into LDS by pass B
    DevBuf tw3;
    int tbits = 0, tnc = 0;
    int pipe_a = 0, pipe_b = 0;  // resident workgroups of the pipelined passes (0 = not used)
    // ComplexF32 passes with two lines per thread (kernels_pair_f32.hip) and their LDS strides
    // of a line pair; a call whose addresses or distances are not 16-byte aligned runs the float
    // kernels on the same tiles
    int pair_a = 0, pair_b = 0, LsA2 = 0, LsB2 = 0;
    int64_t chunk = 0;  // transforms per pass-A/pass-B round (0 = the whole batch)
};

enum class Kind : int { Trivial, Direct, FourStep, ConvSmall, ConvLarge, Uneven, Cols, ND, Large, ColsT };

struct Ctx;

struct Plan {
    Ctx* ctx = nullptr;
    Kind kind = Kind::Trivial;
    int64_t n = 0, batch = 0, idist = 0, odist = 0;
    int dtype = 0, flags = 0;
    Prec prec = Prec::F32;
    // direct and fused convolution: the line FFT and lines per workgroup
    LinePlan line;
    int C = 1;
    int pipe_res = 0;  // direct: resident workgroups of the pipelined rows kernel (0 = not used)
    // four-step
    FourStepPlan fs;
    mutable std::mutex scratch_mu;
    mutable DevBuf scratch;
    // Executions that use the plan's own work buffers are ordered on the device: each waits for
    // work_ev (recorded by the previous one, on whatever stream) before it starts.
    mutable std::mutex exec_mu;
    mutable hipEvent_t work_ev = nullptr;
    Plan() = default;
    Plan(const Plan&) = delete;
    Plan& operator=(const Plan&) = delete;
    ~Plan() {
        if (work_ev) (void)hipEventDestroy(work_ev);
    }
    // ComplexF16 input and output (dtype CFFT_C16): the plan computes in float; direct and
    // four-step plans convert inside their kernels, other kinds through hbuf_in / hbuf_out
    // (guarded by hbuf_mu, apart from scratch_mu: the float transform between the two
    // conversions may itself take scratch_mu, e.g. a six-step plan)
    bool half = false;
    mutable std::mutex hbuf_mu;
    mutable DevBuf hbuf_in, hbuf_out;
    // convolution (Rader / Bluestein)
    ConvMode mode = ConvMode::Rader;
    int64_t L = 0, g = 0;
    int LsC = 0;  // fused convolution: LDS line stride (input and work lines share it)
    Prec cprec = Prec::F32;
    DevBuf perm, iperm, chirp, K, dc, w1, w2;
    std::unique_ptr<Plan> sub;  // large convolution: FFT_L; six-step: the M-point row FFTs
    // large convolution whose FFT_L is a four-step, run as three fused launches: the tile and
    // LDS line stride of the last launch (lines of N1 points over the second FFT's middle)
    bool convl = false;
    int convl_ca = 0, convl_lsa = 0, convl_cb = 0, convl_lsb = 0, convl_cc = 0, convl_lsc = 0;
    // real-input transforms: 1 = fused (n even, n/2 fits a workgroup), 2 = even via a complex
    // sub-plan of n/2, 3 = odd via a complex sub-plan of n; idist / odist hold rdist (reals) and
    // cdist (complex elements)
    int real = 0;
    DevBuf rw, rs1, rs2;
    std::unique_ptr<Plan> rsub_fwd, rsub_bwd;
    // uneven batch
    std::vector<std::unique_ptr<Plan>> parts;
    std::shared_ptr<void> uneven;  // opaque state owned by uneven.cpp
    size_t uneven_bytes = 0;
    int64_t uneven_span = 0;       // elements covered by the uneven layout
    std::string desc;
};

// Execution of a 1-D complex plan on `batch` contiguous rows (idist / odist of the plan), other
// than the batch it was made for; a large convolution then uses the caller's work buffers
// (w1, w2: batch * L elements of its convolution precision each, dc: 2 per line).
struct ExecShape {
    int64_t batch = 0;
    void* w1 = nullptr;
    void* w2 = nullptr;
    void* dc = nullptr;
};

// Planner knobs; the autotuner enumerates them, NO_FLAG uses the defaults.
struct PlanOptions {
    int E = 0;                 // per-thread complex capacity (8 or 16), 0 = default
    int target_threads = 64;   // threads per workgroup the rows kernels aim for (short lines
                               // share a workgroup up to this; small workgroups measured fastest)
    int64_t n1 = 0;            // four-step split (0 = balanced / direct when it fits)
    int ca = 0, cb = 0;        // four-step column tiles (0 = default)
    int64_t chunk = -1;        // four-step transforms per round (-1 = default, 0 = whole batch)
    int pipe = -1;             // direct: pipelined rows kernel (0 = off, 1 = on, -1 = default)
    int large = 0;             // 1: the six-step path with N1 = n1 even where the four-step fits
    int force_conv = 0;
    // Radix sequences (the reference's MEASURE searches them, radix_plan.jl:23-68): rad_a for a
    // direct line or the four-step's first line (its product then sets the split), rad_b for the
    // four-step's second line; empty = the default sequence (line_radices).
    std::vector<int> rad_a, rad_b;
    // Direct and fused-convolution lines: 1 = the last pass runs across the lanes of a wavefront
    // where the line allows it (wave_tail_threads), 0 = through LDS, -1 = the default.
    int wave = -1;
    // Large convolutions over a four-step FFT_L: 1 = the three fused launches (k_convl_*), 0 =
    // gather, two four-steps and scatter, -1 = the default (fused).
    int fuse_conv = -1;
    // ComplexF32 four-step: passes run with two lines per thread (bit 0: pass A, bit 1: pass
    // B), -1 = the default (both with E = 8 lines, pass A with E = 16)
    int pair = -1;
    // Four-step: pass B's per-thread capacity (0 = E), and the passes that run the pipelined
    // kernels (bit 0: pass A, bit 1: pass B; -1 = both when pipe > 0, else neither)
    int Eb = 0;
    int pipe_mask = -1;
    // Direct lines: 0 = LDS lines without pad words (more workgroups per CU where LDS bounds
    // them), 1 / -1 = padded (the default)
    int lds_pad = -1;
    // Four-step tiles of the plan's own transform (not of a convolution's FFT_L): widened from
    // the default until a workgroup has this many threads (0 = the default tiles)
    int min_threads = 0;
    bool explicit_opts = false;
    bool alloc_work = true;    // large convolutions: allocate work buffers for `batch` lines
};

// Pinned host staging buffer (hipHostMalloc, grown on demand with headroom): asynchronous
// table uploads, and small host-pointer transfers.
struct HostStage {
    void* p = nullptr;
    size_t cap = 0;
    unsigned flags = hipHostMallocDefault;  // hipHostMallocCoherent: fine-grained (zero-copy)
    HostStage() = default;
    explicit HostStage(unsigned f) : flags(f) {}
    HostStage(const HostStage&) = delete;
    HostStage& operator=(const HostStage&) = delete;
    ~HostStage() {
        if (p) (void)hipHostFree(p);
    }
    int reserve(size_t n) {
        if (n <= cap) return 0;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
        const size_t want = n + n / 2;
        if (hipHostMalloc(&p, want, flags) != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            return -1;
        }
        cap = want;
        return 0;
    }
};

using PlanKey = std::tuple<int64_t, int64_t, int64_t, int64_t, int, int>;

struct Ctx {
    int device = 0;
    size_t lds_max = 64 * 1024;
    size_t lds_rows_budget = 64 * 1024;
    int num_cu = 256;
    hipStream_t stream = nullptr;
    std::mutex mu;
    std::map<PlanKey, std::shared_ptr<Plan>> cache;
    std::mutex stage_mu;
    DevBuf stage_in, stage_out;
    HostStage pin_in, pin_out;  // pinned copies of small host-pointer transfers
    // fine-grained pinned buffers that small single-kernel plans read and write directly
    HostStage zc_in{hipHostMallocCoherent}, zc_out{hipHostMallocCoherent};
    std::map<std::string, PlanOptions> wisdom;  // tuned choices (autotune.cpp), guarded by mu
};

int ctx_create(int device, Ctx** out);
void ctx_destroy(Ctx* c);

bool is_smooth(int64_t n);
int64_t next_smooth(int64_t m);
std::vector<int> line_radices(int64_t n);
std::vector<int> fourstep_line_radices(int64_t n);
int threads_for(const std::vector<int>& rad, int64_t N, int E);
int line_stride_for(Prec p, int64_t N, bool pad = true);
int make_line(LinePlan& lp, int64_t N, Prec pr, int E, int max_threads,
              const std::vector<int>* radices = nullptr);
// Whether rad is a radix sequence the kernels implement whose product is N.
bool radices_valid(const std::vector<int>& rad, int64_t N);
// Switches the line to a last pass across wavefront lanes when it allows it (returns whether it
// did; the line's thread count may shrink).
bool apply_wave_tail(LinePlan& lp);

// Plans p (fields n, batch, idist, odist, prec already set) with the given knobs.
int plan_build(Ctx* ctx, Plan& p, const PlanOptions& o);
int plan_create(Ctx* ctx, int64_t n, int64_t batch, int64_t idist, int64_t odist, int dtype,
                int flags, const PlanOptions& opts, std::shared_ptr<Plan>* out);
int plan_exec(const Plan& p, const void* in, void* out, int direction, int normalize,
              hipStream_t st);
int plan_exec_shaped(const Plan& p, const void* in, void* out, int direction, int normalize,
                     hipStream_t st, const ExecShape& sh);
int plan_nd_create(Ctx* ctx, int rank, const int64_t* dims, int64_t axes_mask, int dtype,
                   int flags, std::shared_ptr<Plan>* out);
int plan_real_create(Ctx* ctx, int64_t n, int64_t batch, int64_t rdist, int64_t cdist, int dtype,
                     int flags, std::shared_ptr<Plan>* out);
size_t plan_device_bytes(const Plan& p);
std::string plan_description(const Plan& p);

struct ConvData {
    ConvMode mode = ConvMode::Rader;
    int64_t L = 0, g = 0;
    std::vector<int> perm, iperm;
    std::vector<std::complex<long double>> chirp, b;
};
int conv_data(int64_t n, ConvData& d);
int conv_spectrum(Ctx* ctx, const std::vector<std::complex<long double>>& b, Prec cprec, DevBuf& K);
Prec conv_precision(int64_t n, Prec io, ConvMode mode);
int upload_cld(DevBuf& buf, const std::vector<std::complex<long double>>& v, Prec pr);
int upload_ints(DevBuf& buf, const std::vector<int>& v);
bool line_fits_lds(Ctx* ctx, int64_t N, Prec pr, int E);

// autotune.cpp: MEASURE / ENCHANT planning (times candidate PlanOptions on the device).
int autotune(Ctx* ctx, Plan& p, int flags, const PlanOptions& base);
int wisdom_export(Ctx* ctx, const char* path);
int wisdom_import(Ctx* ctx, const char* path);  // returns the number of entries read

// uneven.cpp: many transforms of different lengths in one plan.
int uneven_create(Ctx* ctx, int64_t count, const int64_t* lengths, const int64_t* offsets,
                  int dtype, int flags, std::shared_ptr<Plan>* out);
int uneven_exec(const Plan& p, const void* in, void* out, int direction, int normalize,
                hipStream_t st);

}  // namespace cfft

// The C ABI's context handle (include/catabrafft.h).
struct cfft_ctx_s {
    cfft::Ctx* c;
};
