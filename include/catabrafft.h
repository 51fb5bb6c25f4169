/*
 * catabrafft.h -- C ABI of the MI355X (gfx950) FFT library that backs CatabraFFT's API.
 *
 * Every entry point takes plain pointers and sizes, returns an int status (0 = OK, < 0 = error,
 * message in cfft_last_error()) and never aborts.  Complex data is interleaved (re, im) of
 * float (CFFT_C32, Julia ComplexF32), double (CFFT_C64, Julia ComplexF64) or, for 1-D and N-D
 * complex plans, half (CFFT_C16, Julia ComplexF16), rows contiguous,
 * 0-based.  Forward uses e^{-2 pi i jk/n}; backward is unnormalised (CatabraFFT.bfft) and
 * normalize = 1 divides by n (CatabraFFT.ifft).  Plans are immutable once created and may be
 * executed concurrently on distinct streams; executions that use the plan's own work buffers
 * (in-place four-step, six-step, large convolutions, ComplexF16 staging, N-D, real and uneven
 * plans) are ordered on the device in call order.  Out-of-place execution never modifies the input;
 * in == out is allowed.
 *
 * Reference interfaces these entry points replace (CatabraFFT.jl sources):
 *   cfft_init / cfft_finalize ...... the module's global state (CatabraFFT.jl:20, kernel.jl:20-22)
 *   cfft_plan_1d ................... plan_fft(x, flags) CatabraFFT.jl:149-171 and the cached
 *                                    closure of generate_and_cache_fft! kernel.jl:72-92
 *   cfft_exec ...................... mul!(y, p, x) CatabraFFT.jl:186-189 / p * x :192-196 on
 *                                    device pointers, batched
 *   cfft_exec_host ................. fft / ifft / bfft CatabraFFT.jl:65-115 on host arrays
 *   cfft_plan_destroy .............. (Julia GC of a Spell)
 *   cfft_empty_cache ............... empty_cache() CatabraFFT.jl:39-43
 *   flags .......................... NO_FLAG / MEASURE / ENCHANT spells.jl:4-8
 */
#ifndef CATABRAFFT_H
#define CATABRAFFT_H

#include <stddef.h>
#include <stdint.h>

#if defined(__GNUC__)
#define CFFT_API __attribute__((visibility("default")))
#else
#define CFFT_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define CFFT_OK 0
#define CFFT_ERR_INVALID (-1)     /* bad argument (n < 1, bad dtype, null pointer, ...)    */
#define CFFT_ERR_NOMEM (-2)       /* device or host allocation failed                       */
#define CFFT_ERR_HIP (-3)         /* HIP runtime error                                       */
#define CFFT_ERR_UNSUPPORTED (-4) /* size / layout the library cannot plan                   */
#define CFFT_ERR_ADMISSION (-5)   /* server: request larger than the HBM cap                 */
#define CFFT_ERR_NOTREADY (-6)    /* server: ticket not finished yet                          */

/* dtypes */
#define CFFT_C32 0
#define CFFT_C64 1
/* ComplexF16: (re, im) IEEE binary16 pairs, 4 bytes per element; computed in float and rounded
 * to nearest even on output.  cfft_plan_1d and cfft_plan_nd only (runtests.jl:39-46 round-trips
 * ComplexF16 through fft / ifft). */
#define CFFT_C16 2

/* planner flags (spells.jl:4-8) */
#define CFFT_NO_FLAG 0
#define CFFT_MEASURE 1
#define CFFT_ENCHANT 2

/* directions */
#define CFFT_FORWARD (-1)
#define CFFT_BACKWARD 1

typedef struct cfft_ctx_s* cfft_ctx;
typedef struct cfft_plan_s* cfft_plan;
typedef struct cfft_server_s* cfft_server;

/* Thread-local text of the last error on this thread ("" if none). */
CFFT_API const char* cfft_last_error(void);

/* Library version string. */
CFFT_API const char* cfft_version(void);

/* Opens a context on HIP device `device`. */
CFFT_API int cfft_init(int device, cfft_ctx* ctx);
/* Releases every cached plan, buffer and stream of the context. */
CFFT_API int cfft_finalize(cfft_ctx ctx);

/* A plan for `batch` transforms of length n: row b of the input starts at element b*idist,
 * row b of the output at element b*odist (idist, odist >= n).  dtype CFFT_C32 / CFFT_C64 /
 * CFFT_C16.  n up to 2^30: lengths the four-step cannot split into two workgroup-sized
 * factors (lines of at most 16384 points in C32, about 9800 in C64) run as six steps (about
 * five reads and writes per element), as do the convolution FFTs of such lengths.  flags CFFT_NO_FLAG (default
 * heuristics) / CFFT_MEASURE (time the candidate kernel configurations on the device and keep
 * the fastest) / CFFT_ENCHANT (wider search). */
CFFT_API int cfft_plan_1d(cfft_ctx ctx, int64_t n, int64_t batch, int64_t idist, int64_t odist,
                 int dtype, int flags, cfft_plan* plan);

/* Multi-dimensional transform of one row-major array dims[0] x ... x dims[rank-1] (the last
 * axis contiguous) over the axes whose bit is set in axes_mask (bit a = axis a); normalize
 * divides by the product of the transformed lengths.  Replaces plan_fft's `region` argument
 * (CatabraFFT.jl:169, accepted and ignored there).  Any axis length: strided axes that one
 * workgroup cannot take as strided lines (not smooth, or longer than a workgroup's line) are
 * transposed, transformed as contiguous rows and transposed back. */
CFFT_API int cfft_plan_nd(cfft_ctx ctx, int rank, const int64_t* dims, int64_t axes_mask,
                          int dtype, int flags, cfft_plan* plan);

/* Real-input transforms (rfft / brfft / irfft): `batch` rows of n reals (row b at element
 * b*rdist of a float / double array, rdist >= n) and their half spectra of n/2+1 complex
 * values (row b at element b*cdist, cdist >= n/2+1).  cfft_exec with CFFT_FORWARD maps reals
 * to half spectra (rfft); CFFT_BACKWARD maps half spectra to reals (brfft, and irfft with
 * normalize = 1).  Out of place only.  dtype CFFT_C32 / CFFT_C64 selects float / double.
 * No counterpart in the reference: its plan_rfft / plan_brfft stubs (CatabraFFT.jl:199-208)
 * construct nothing usable. */
CFFT_API int cfft_plan_1d_real(cfft_ctx ctx, int64_t n, int64_t batch, int64_t rdist,
                               int64_t cdist, int dtype, int flags, cfft_plan* plan);

/* A plan for `count` transforms of uneven lengths: transform i has lengths[i] elements at
 * element offset offsets[i] of both the input and the output buffer. */
CFFT_API int cfft_plan_many_uneven(cfft_ctx ctx, int64_t count, const int64_t* lengths,
                          const int64_t* offsets, int dtype, int flags, cfft_plan* plan);

/* Executes on device pointers, enqueued on `stream` (a hipStream_t; NULL = the null stream).
 * direction CFFT_FORWARD / CFFT_BACKWARD; normalize = 1 scales by 1/n. */
CFFT_API int cfft_exec(cfft_plan plan, const void* in, void* out, int direction, int normalize,
              void* stream);

/* cfft_exec without any allocation.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  A four-step execution whose middle needs a work buffer (in == out, or
 * ComplexF16 data) takes it from a static device work area of 64 MiB per GPU (batch * n * 8
 * bytes in C32 and C16, * 16 in C64, must fit; larger is CFFT_ERR_UNSUPPORTED); direct,
 * fused-convolution and trivial plans need none.  Plan kinds that own work buffers (six-step,
 * large convolutions, N-D, real, uneven) are refused with CFFT_ERR_UNSUPPORTED.  Every call on a
 * GPU shares that GPU's work area: calls that overlap on one GPU (other streams, other threads)
 * race and must be ordered by the caller.  Results are bitwise those of cfft_exec. */
CFFT_API int cfft_exec_noalloc(cfft_plan plan, const void* in, void* out, int direction,
                               int normalize, void* stream);

/* Executes on host pointers: stages through device buffers owned by the context and returns
 * when the result is in `out`.  This is what CatabraFFT.fft / ifft / bfft call. */
CFFT_API int cfft_exec_host(cfft_plan plan, const void* in, void* out, int direction, int normalize);

/* Device bytes of scratch the plan holds (tables + work buffers). */
CFFT_API int cfft_plan_device_bytes(cfft_plan plan, int64_t* bytes);

/* Human-readable description of the plan (strategy, factors, kernel shapes). */
CFFT_API int cfft_plan_describe(cfft_plan plan, char* buf, int64_t buflen);

CFFT_API int cfft_plan_destroy(cfft_plan plan);

/* Drops the context's plan cache and staging buffers (empty_cache, CatabraFFT.jl:39-43).
 * Plans still held by callers stay valid. */
CFFT_API int cfft_empty_cache(cfft_ctx ctx);

/* Wisdom: the choices MEASURE / ENCHANT planning made, as a text file, so a later process
 * reuses them without timing again (the plan serialisation drafted in spells.jl:180-374).
 * Import replaces the context's cached plans of the imported keys (n, batch, idist, odist,
 * dtype, flags): later MEASURE / ENCHANT requests for them are built from the imported
 * choices.  Plans already returned stay valid and unchanged. */
CFFT_API int cfft_wisdom_export(cfft_ctx ctx, const char* path);
CFFT_API int cfft_wisdom_import(cfft_ctx ctx, const char* path);

/* ---- device memory helpers for hosts without their own allocator (Julia shim, C drivers) */
CFFT_API int cfft_malloc(cfft_ctx ctx, int64_t bytes, void** ptr);
CFFT_API int cfft_free(cfft_ctx ctx, void* ptr);
CFFT_API int cfft_memcpy_h2d(cfft_ctx ctx, void* dst, const void* src, int64_t bytes);
CFFT_API int cfft_memcpy_d2h(cfft_ctx ctx, void* dst, const void* src, int64_t bytes);
CFFT_API int cfft_synchronize(cfft_ctx ctx);

/* ---- serving loop: continuous batching of FFT requests of mixed sizes ------------------- */
/* A server with its own worker thread and streams; admission keeps the resident device bytes
 * of all in-flight requests (their buffers + plan tables) under hbm_cap_bytes. */
CFFT_API int cfft_server_create(cfft_ctx ctx, int64_t hbm_cap_bytes, int nstreams, cfft_server* srv);
/* Submits `batch` contiguous transforms of length n.  on_device = 0: `in` / `out` are host
 * buffers (the server stages them through device memory it admits under the cap);
 * on_device = 1: they are device buffers already resident in HBM (their bytes still count
 * toward the cap while the request is in flight).  Both stay owned by the caller and must live
 * until the ticket completes.  in == out is allowed (device lengths that run as a four-step go
 * through a pooled output buffer and a device copy).  A request that fails reports its own
 * status; the other requests of its round are unaffected. */
CFFT_API int cfft_server_submit(cfft_server srv, int64_t n, int64_t batch, int dtype,
                                int direction, int normalize, const void* in, void* out,
                                int on_device, int64_t* ticket);
/* 0 when done, CFFT_ERR_NOTREADY when still queued / running, < 0 on failure. */
CFFT_API int cfft_server_poll(cfft_server srv, int64_t ticket);
/* Blocks until the ticket is done. */
CFFT_API int cfft_server_wait(cfft_server srv, int64_t ticket);
/* Latency percentiles (seconds, submit to completion) and throughput (transforms / s, and
 * GFLOP/s at 5 n log2 n) over every completed request so far. */
CFFT_API int cfft_server_stats(cfft_server srv, double* p50, double* p99, double* transforms_per_s,
                      double* gflops);
/* Forgets the latency / throughput history (resident plans stay). */
CFFT_API int cfft_server_reset_stats(cfft_server srv);
CFFT_API int cfft_server_destroy(cfft_server srv);

#ifdef __cplusplus
}
#endif

#endif /* CATABRAFFT_H */
