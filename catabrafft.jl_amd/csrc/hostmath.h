// hostmath.h -- host-side table generation in long double, rounded once to the device precision.
//
// Twiddles, four-step tables, Bluestein chirps and convolution-kernel spectra are computed from
// exactly reduced integer indices in long double (64-bit mantissa on x86-64) and converted to
// float / double only when uploaded, so every device constant carries a single rounding.
// (CatabraFFT.jl instead forms them with cispi in the working precision plus power chains,
// radix_2_codelets.jl:423-429, mixed_radix.jl:36-38, prime.jl:36-37.)
#pragma once
#include <cmath>
#include <complex>
#include <cstdint>
#include <vector>

namespace cfft {

using cld = std::complex<long double>;

// sin and cos of 2 pi m / n with the octant reduction so sinl/cosl only see |arg| <= pi/4.
inline void sincos_2pi_frac(int64_t m, int64_t n, long double* s, long double* c) {
    m %= n;
    if (m < 0) m += n;
    const __int128 m8 = (__int128)8 * m;
    const int oct = (int)(m8 / n);
    const __int128 rem = m8 - (__int128)oct * n;
    const long double pi4 = 0.785398163397448309615660845819875721L;
    const long double a = pi4 * (long double)rem / (long double)n;
    const long double b = pi4 - a;
    const long double sa = sinl(a), ca = cosl(a), sb = sinl(b), cb = cosl(b);
    long double ss, cc;
    switch (oct) {
        case 0: ss = sa; cc = ca; break;
        case 1: ss = cb; cc = sb; break;
        case 2: ss = ca; cc = -sa; break;
        case 3: ss = sb; cc = -cb; break;
        case 4: ss = -sa; cc = -ca; break;
        case 5: ss = -cb; cc = -sb; break;
        case 6: ss = -ca; cc = sa; break;
        default: ss = -sb; cc = cb; break;
    }
    *s = ss;
    *c = cc;
}

// w_n^m = e^{-2 pi i m / n}
inline cld omega(int64_t m, int64_t n) {
    long double s, c;
    sincos_2pi_frac(m, n, &s, &c);
    return cld(c, -s);
}

// w_n^m for any integer m from two tables of ~sqrt(n) long-double entries:
// w_n^m = t1[m mod 2^shift] * t2[m >> shift] (one extra long-double rounding, ~1e-19).
struct OmegaTable {
    int64_t n = 1;
    int shift = 0;
    std::vector<cld> t1, t2;
    explicit OmegaTable(int64_t n_) : n(n_) {
        while (((int64_t)1 << (2 * shift)) < n) ++shift;
        t1.resize((size_t)1 << shift);
        t2.resize((size_t)((n >> shift) + 1));
        for (size_t i = 0; i < t1.size(); ++i) t1[i] = omega((int64_t)i, n);
        for (size_t i = 0; i < t2.size(); ++i) t2[i] = omega((int64_t)i << shift, n);
    }
    cld operator()(int64_t m) const {
        m %= n;
        if (m < 0) m += n;
        return t1[m & (((int64_t)1 << shift) - 1)] * t2[m >> shift];
    }
};

inline bool is_prime(int64_t n) {
    if (n < 2) return false;
    if (n < 4) return true;
    if (n % 2 == 0) return false;
    for (int64_t d = 3; d * d <= n; d += 2)
        if (n % d == 0) return false;
    return true;
}

inline int64_t powmod(int64_t b, int64_t e, int64_t m) {
    unsigned __int128 r = 1, x = (unsigned __int128)(b % m);
    while (e > 0) {
        if (e & 1) r = (r * x) % (unsigned __int128)m;
        x = (x * x) % (unsigned __int128)m;
        e >>= 1;
    }
    return (int64_t)r;
}

// Smallest primitive root (prime.jl:9-28 searches g = 2, 3, ... for distinct powers).
inline int64_t primitive_root(int64_t p) {
    if (p == 2) return 1;
    std::vector<int64_t> f;
    int64_t t = p - 1;
    for (int64_t d = 2; d * d <= t; ++d)
        if (t % d == 0) {
            f.push_back(d);
            while (t % d == 0) t /= d;
        }
    if (t > 1) f.push_back(t);
    for (int64_t g = 2; g < p; ++g) {
        bool ok = true;
        for (int64_t q : f)
            if (powmod(g, (p - 1) / q, p) == 1) {
                ok = false;
                break;
            }
        if (ok) return g;
    }
    return 0;
}

// Forward DFT of a smooth length in long double (iterative Stockham, generic radix), used only
// at plan time for convolution-kernel spectra.
inline bool host_fft_ld(std::vector<cld>& x) {
    const int64_t n = (int64_t)x.size();
    std::vector<int> rad;
    int64_t r = n;
    for (int p : {4, 2, 3, 5, 7, 11, 13}) {
        while (r % p == 0) {
            rad.push_back(p);
            r /= p;
        }
    }
    if (r != 1) return false;
    std::vector<cld> y(n);
    cld* src = x.data();
    cld* dst = y.data();
    int64_t s = 1;
    std::vector<cld> wR;
    const OmegaTable om(n);  // w_g^{pk} = w_n^{s p k}
    for (int R : rad) {
        const int64_t g = n / s, n1 = g / R;
        wR.resize((size_t)R * R);
        for (int j = 0; j < R; ++j)
            for (int k = 0; k < R; ++k) wR[j * R + k] = omega((int64_t)j * k, R);
        cld v[16], o[16];
        for (int64_t p = 0; p < n1; ++p) {
            cld tw[16];
            for (int k = 0; k < R; ++k) tw[k] = om(s * p * k);
            for (int64_t q = 0; q < s; ++q) {
                for (int j = 0; j < R; ++j) v[j] = src[q + s * (p + j * n1)];
                for (int k = 0; k < R; ++k) {
                    cld acc = 0;
                    for (int j = 0; j < R; ++j) acc += v[j] * wR[j * R + k];
                    o[k] = acc;
                }
                for (int k = 0; k < R; ++k) dst[q + s * (R * p + k)] = o[k] * tw[k];
            }
        }
        std::swap(src, dst);
        s *= R;
    }
    if (src != x.data()) std::copy(src, src + n, x.data());
    return true;
}

}  // namespace cfft
