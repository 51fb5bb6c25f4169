/*
 * one_per_launch.c -- the plain one-transform-per-launch caller that the serving loop is
 * measured against: one cfft_exec of a batch-1 plan per transform, issued from C, the way a
 * Julia program that loops over CatabraFFT.fft-style calls (CatabraFFT.jl:65-72, one transform per
 * call, no batch API) would issue them through ccall.  Built by build.py into
 * lib/libcfft_one_per_launch.so; bench.py calls it through ctypes so that no interpreter
 * overhead sits between the launches.
 */
#include <stdint.h>

#include "catabrafft.h"

/* For i < count: cfft_exec(plans[which[i]], in_base + in_off[i] * esize, out_base + out_off[i] *
 * esize, direction, normalize, stream).  Returns 0, or the first failing call's status. */
CFFT_API int cfft_one_per_launch(const cfft_plan* plans, const int64_t* which,
                                 const int64_t* in_off, const int64_t* out_off, int64_t count,
                                 const void* in_base, void* out_base, int64_t esize,
                                 int direction, int normalize, void* stream) {
    const char* in = (const char*)in_base;
    char* out = (char*)out_base;
    for (int64_t i = 0; i < count; ++i) {
        const int rc = cfft_exec(plans[which[i]], in + in_off[i] * esize, out + out_off[i] * esize,
                                 direction, normalize, stream);
        if (rc != CFFT_OK) return rc;
    }
    return CFFT_OK;
}
