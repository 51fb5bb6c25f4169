/*
 * catabra_oracle.c -- CPU ORACLE for the batched 1-D complex FFT path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load this library, and only as the checker (or as the timed CPU baseline).  The
 * product (catabrafft.jl_amd/) never links, loads or calls it.
 *
 * It is a plain-C restatement of CatabraFFT.jl's algorithm, instantiated for float, double
 * and long double (oracle_body.h, suffixes _f/_d/_l):
 *   size classes and factor choice ........ kernel.jl:72-92, 102-110, 146-191
 *   greedy Stockham radix plan ............ radix_plan.jl:100-138 with the NO_FLAG radix sets
 *                                            [8,4,2] / [9,3] / [5] / [7] of kernel.jl:131-139
 *   Stockham passes and codelets ........... radix_2/3/5/7_codelets.jl, radix_exec.jl:49-70,101
 *   four-step mixed radix .................. mixed_radix.jl:19-78
 *   Rader .................................. prime.jl:5-82
 *   fft / ifft / bfft conventions .......... CatabraFFT.jl:65-115
 * plus a naive long-double DFT (orc_dft_l) used to pin the restatement.
 *
 * Deliberate deviations, each a reference defect rather than behaviour:
 *   - twiddles (layered passes, the four-step matrix D, Rader's W) are computed in long double
 *     from the exactly reduced index and rounded once, instead of cispi + power chains /
 *     repeated products (radix_2_codelets.jl:423-429, mixed_radix.jl:36-38, prime.jl:36-37);
 *   - Rader's FFT(W) is computed once at plan time and W is never overwritten (prime.jl:61);
 *   - the input is never overwritten (radix_exec.jl:63-68 lets passes write into x);
 *   - the defective layered fft4/fft2/fft3 variants are not used; every pass uses the correct
 *     pass formula out[q+s(Rp+k)] = w_g^{pk} sum_j in[q+s(p+j n1)] w_R^{jk}.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum { ORC_TRIVIAL = 0, ORC_RADIX = 1, ORC_MIXED = 2, ORC_RADER = 3 };

/* sin and cos of 2 pi m / n in long double, using the octant symmetry so that the argument
 * passed to sinl/cosl is at most pi/4 (keeps the result accurate to long-double precision). */
static void orc_sincos_2pi_frac(long long m, long long n, long double* s, long double* c) {
    m %= n;
    if (m < 0) m += n;
    /* reduce to the first octant: angle = 2 pi m / n; work with 8m vs n */
    long long m8 = 8 * m;  /* angle / (pi/4) = 8m/n */
    int oct = (int)(m8 / n);
    long long rem = m8 - (long long)oct * n;  /* angle = (oct + rem/n) * pi/4 */
    const long double pi4 = 0.785398163397448309615660845819875721L;
    long double a = pi4 * (long double)rem / (long double)n;
    long double b = pi4 - a;
    long double sa = sinl(a), ca = cosl(a), sb = sinl(b), cb = cosl(b);
    long double ss = 0, cc = 0;
    switch (oct) {
        case 0: ss = sa; cc = ca; break;            /* a             */
        case 1: ss = cb; cc = sb; break;            /* pi/2 - b      */
        case 2: ss = ca; cc = -sa; break;           /* pi/2 + a      */
        case 3: ss = sb; cc = -cb; break;           /* pi - b        */
        case 4: ss = -sa; cc = -ca; break;          /* pi + a        */
        case 5: ss = -cb; cc = -sb; break;          /* 3pi/2 - b     */
        case 6: ss = -ca; cc = sa; break;           /* 3pi/2 + a     */
        default: ss = -sb; cc = cb; break;          /* 2pi - b       */
    }
    *s = ss;
    *c = cc;
}

/* kernel.jl:102-110 */
int orc_is_power_of(long n, long p) {
    if (n < 1) return 0;
    while (n > 1) {
        if (n % p != 0) return 0;
        n /= p;
    }
    return 1;
}

/* Primes.isprime, used by kernel.jl:81 and prime.jl:6 (trial division is exact for our sizes). */
int orc_isprime(long n) {
    if (n < 2) return 0;
    if (n < 4) return 1;
    if (n % 2 == 0) return 0;
    for (long d = 3; d * d <= n; d += 2)
        if (n % d == 0) return 0;
    return 1;
}

static long orc_isqrt(long n) {
    long r = (long)sqrt((double)n);
    while (r * r > n) --r;
    while ((r + 1) * (r + 1) <= n) ++r;
    return r;
}

/* kernel.jl:146-191: largest power of 2,3,5,7 dividing n first, else the divisor closest
 * below sqrt(n).  Returns p (first factor) and m = n / p. */
void orc_closest_factors(long n, long* p, long* m) {
    if (orc_isprime(n)) {
        *p = 1;
        *m = n;
        return;
    }
    long powers[256];
    int np = 0;
    const long primes[4] = {2, 3, 5, 7};
    for (int i = 0; i < 4; ++i) {
        long pw = primes[i];
        while (pw <= n) {
            powers[np++] = pw;
            if (pw > (1L << 61) / primes[i]) break;
            pw *= primes[i];
        }
    }
    /* descending sort (kernel.jl:162) */
    for (int i = 1; i < np; ++i) {
        long v = powers[i];
        int j = i - 1;
        while (j >= 0 && powers[j] < v) {
            powers[j + 1] = powers[j];
            --j;
        }
        powers[j + 1] = v;
    }
    for (int i = 0; i < np; ++i)
        if (n % powers[i] == 0) {
            *p = powers[i];
            *m = n / powers[i];
            return;
        }
    long q = orc_isqrt(n);
    while (n % q != 0) --q;
    *p = q;
    *m = n / q;
}

/* radix_plan.jl:100-138 with the radix set chosen by kernel.jl:131-139.  Writes the radices
 * in pass order and returns the pass count (0 when n is not a power of 2, 3, 5 or 7). */
int orc_std_radix_plan(long n, int* radices, int maxlen) {
    int set[3];
    int nset;
    if (orc_is_power_of(n, 2)) {
        set[0] = 8; set[1] = 4; set[2] = 2; nset = 3;
    } else if (orc_is_power_of(n, 3)) {
        set[0] = 9; set[1] = 3; nset = 2;
    } else if (orc_is_power_of(n, 5)) {
        set[0] = 5; nset = 1;
    } else if (orc_is_power_of(n, 7)) {
        set[0] = 7; nset = 1;
    } else {
        return 0;
    }
    int k = 0;
    long rem = n;
    while (rem > 1) {
        int r = 0;
        for (int i = 0; i < nset; ++i)
            if (rem % set[i] == 0) {
                r = set[i];
                break;
            }
        if (r == 0 || k >= maxlen) return 0;
        radices[k++] = r;
        rem /= r;
    }
    return k;
}

long orc_powmod(long b, long e, long m) {
    unsigned __int128 r = 1, x = (unsigned __int128)(b % m);
    while (e > 0) {
        if (e & 1) r = (r * x) % (unsigned __int128)m;
        x = (x * x) % (unsigned __int128)m;
        e >>= 1;
    }
    return (long)r;
}

/* prime.jl:9-28: the smallest g in 2..p-1 whose powers g^1..g^(p-1) are all distinct. */
long orc_primitive_root(long p) {
    if (p == 2) return 1;
    /* distinct powers <=> order p-1 <=> g^((p-1)/q) != 1 for every prime q | p-1 */
    long f[64];
    int nf = 0;
    long t = p - 1;
    for (long d = 2; d * d <= t; ++d)
        if (t % d == 0) {
            f[nf++] = d;
            while (t % d == 0) t /= d;
        }
    if (t > 1) f[nf++] = t;
    for (long g = 2; g < p; ++g) {
        int ok = 1;
        for (int i = 0; i < nf; ++i)
            if (orc_powmod(g, (p - 1) / f[i], p) == 1) {
                ok = 0;
                break;
            }
        if (ok) return g;
    }
    return 0;
}

/* Naive DFT in long double, X_k = sum_j x_j e^{sign 2 pi i jk/n}, the pinning reference for
 * the restatement (sign = -1 forward, +1 backward, unnormalised). */
void orc_dft_l(long n, const long double* x, long double* y, int sign) {
    long double* c = (long double*)malloc(sizeof(long double) * (size_t)n);
    long double* s = (long double*)malloc(sizeof(long double) * (size_t)n);
    for (long m = 0; m < n; ++m) orc_sincos_2pi_frac(m, n, &s[m], &c[m]);
    for (long k = 0; k < n; ++k) {
        long double re = 0, im = 0;
        long idx = 0;
        for (long j = 0; j < n; ++j) {
            long double xr = x[2 * j], xi = x[2 * j + 1];
            long double cr = c[idx], ci = (long double)sign * s[idx];
            re += xr * cr - xi * ci;
            im += xr * ci + xi * cr;
            idx += k;
            if (idx >= n) idx -= n;
        }
        y[2 * k] = re;
        y[2 * k + 1] = im;
    }
    free(c);
    free(s);
}

/* Conversions used by the ctypes wrapper for long-double buffers. */
void orc_d_to_l(const double* a, long double* b, long count) {
    for (long i = 0; i < count; ++i) b[i] = (long double)a[i];
}
void orc_l_to_d(const long double* a, double* b, long count) {
    for (long i = 0; i < count; ++i) b[i] = (double)a[i];
}
/* Splits a long double array into hi + lo doubles so callers keep the extra precision. */
void orc_l_to_dd(const long double* a, double* hi, double* lo, long count) {
    for (long i = 0; i < count; ++i) {
        double h = (double)a[i];
        hi[i] = h;
        lo[i] = (double)(a[i] - (long double)h);
    }
}
int orc_sizeof_long_double(void) { return (int)sizeof(long double); }

#define OT float
#define OSFX f
#include "oracle_body.h"
#undef OT
#undef OSFX

#define OT double
#define OSFX d
#include "oracle_body.h"
#undef OT
#undef OSFX

#define OT long double
#define OSFX l
#include "oracle_body.h"
#undef OT
#undef OSFX
