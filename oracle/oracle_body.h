/*
 * oracle_body.h -- type-generic body of the CPU oracle (TEST INFRASTRUCTURE ONLY).
 *
 * Included three times by catabra_oracle.c with
 *   OT   = float | double | long double   (arithmetic type)
 *   OSFX = f | d | l                      (name suffix)
 * Every function restates one piece of CatabraFFT.jl's algorithm; the reference file:line
 * it follows is cited on each function.  Deliberate deviations from the reference (all are
 * reference defects, not behaviour) are listed in the header of catabra_oracle.c.
 */

#define OCAT2(a, b) a##_##b
#define OCAT(a, b) OCAT2(a, b)
#define ON(name) OCAT(name, OSFX)

typedef struct { OT re, im; } ON(ocpx);
#define OC ON(ocpx)

static inline OC ON(cmk)(OT r, OT i) { OC z; z.re = r; z.im = i; return z; }
static inline OC ON(cadd)(OC a, OC b) { return ON(cmk)(a.re + b.re, a.im + b.im); }
static inline OC ON(csub)(OC a, OC b) { return ON(cmk)(a.re - b.re, a.im - b.im); }
static inline OC ON(cmul)(OC a, OC b) {
    return ON(cmk)(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline OC ON(cscale)(OC a, OT s) { return ON(cmk)(a.re * s, a.im * s); }
/* multiply by -i */
static inline OC ON(cmnj)(OC a) { return ON(cmk)(a.im, -a.re); }
static inline OC ON(cconj)(OC a) { return ON(cmk)(a.re, -a.im); }

/* e^{-2 pi i k / n}, computed in long double with the index reduced mod n, rounded once to OT. */
static OC ON(omega)(long long k, long long n) {
    long long m = k % n;
    if (m < 0) m += n;
    long double c, s;
    orc_sincos_2pi_frac(m, n, &s, &c);
    return ON(cmk)((OT)c, (OT)-s);
}

/* Codelet constants cos/sin(2 pi m / R) for R <= 16, filled once by ON(consts_init) before any
 * transform runs (plan creation calls it), so the butterflies below only read them. */
static OT ON(kcos)[17][17], ON(ksin)[17][17];
static int ON(kready);
static void ON(consts_init)(void) {
    if (ON(kready)) return;
    for (int R = 1; R <= 16; ++R)
        for (int m = 0; m < R; ++m) {
            long double c, s;
            orc_sincos_2pi_frac(m, R, &s, &c);
            ON(kcos)[R][m] = (OT)c;
            ON(ksin)[R][m] = (OT)s;
        }
    ON(kready) = 1;
}
/* e^{-2 pi i m / R} from the constant table, R <= 16. */
static inline OC ON(kw)(int m, int R) { m %= R; return ON(cmk)(ON(kcos)[R][m], -ON(ksin)[R][m]); }

/* ---------------------------------------------------------------------------------------
 * Codelets: one radix-R butterfly, in place on v[0..R-1], forward direction.
 * Radix-2 family: radix_2_codelets.jl:450-488 (fft8 shell), :572-588 (fft4), :642-656 (fft2),
 *   :123-199 (fft16, computed as 4x4 with internal rotations).
 * Radix-3 family: radix_3_codelets.jl:12-74 (fft9 via pairs), :338-366 (fft3).
 * Radix-5: radix_5_codelets.jl:27-44.  Radix-7: radix_7_codelets.jl:30-51.
 * The formulas are the DFT definition X_k = sum_j v_j w_R^{jk}, factored the way those
 * codelets factor it (pairs of conjugate-symmetric terms for the odd primes).
 * ------------------------------------------------------------------------------------- */
static void ON(bfly2)(OC* v) {
    OC a = v[0], b = v[1];
    v[0] = ON(cadd)(a, b);
    v[1] = ON(csub)(a, b);
}

static void ON(bfly4)(OC* v) {
    OC t0 = ON(cadd)(v[0], v[2]), t1 = ON(csub)(v[0], v[2]);
    OC t2 = ON(cadd)(v[1], v[3]), t3 = ON(cmnj)(ON(csub)(v[1], v[3]));
    v[0] = ON(cadd)(t0, t2);
    v[2] = ON(csub)(t0, t2);
    v[1] = ON(cadd)(t1, t3);
    v[3] = ON(csub)(t1, t3);
}

/* radix_2_codelets.jl:470-488: 8 = 2 x 4 split with the 1/sqrt(2)(1-i) rotations. */
static void ON(bfly8)(OC* v) {
    const OT r = (OT)0.707106781186547524400844362104849039L;
    OC t13 = ON(cadd)(v[0], v[4]), t14 = ON(csub)(v[0], v[4]);
    OC t15 = ON(cadd)(v[2], v[6]), t16 = ON(cmnj)(ON(csub)(v[2], v[6]));
    OC t5 = ON(cadd)(t13, t15), t6 = ON(cadd)(t14, t16);
    OC t7 = ON(csub)(t13, t15), t8 = ON(csub)(t14, t16);
    OC t17 = ON(cadd)(v[1], v[5]), t18 = ON(csub)(v[1], v[5]);
    OC t19 = ON(cadd)(v[3], v[7]), t20 = ON(cmnj)(ON(csub)(v[3], v[7]));
    OC t9 = ON(cadd)(t17, t19);
    OC u = ON(cadd)(t18, t20);
    OC t10 = ON(cmk)(r * (u.re + u.im), r * (u.im - u.re));          /* r(1-i)u  */
    OC t11 = ON(cmnj)(ON(csub)(t17, t19));
    OC w = ON(csub)(t18, t20);
    OC t12 = ON(cmk)(r * (w.im - w.re), -r * (w.re + w.im));         /* r(-1-i)w */
    v[0] = ON(cadd)(t5, t9);   v[4] = ON(csub)(t5, t9);
    v[1] = ON(cadd)(t6, t10);  v[5] = ON(csub)(t6, t10);
    v[2] = ON(cadd)(t7, t11);  v[6] = ON(csub)(t7, t11);
    v[3] = ON(cadd)(t8, t12);  v[7] = ON(csub)(t8, t12);
}

/* radix_2_codelets.jl:123-199: 16-point as 4x4 (column DFT4, rotation w16^{jk}, row DFT4). */
static void ON(bfly16)(OC* v) {
    OC a[16];
    for (int j1 = 0; j1 < 4; ++j1) {
        OC c[4] = {v[j1], v[j1 + 4], v[j1 + 8], v[j1 + 12]};
        ON(bfly4)(c);
        for (int k2 = 0; k2 < 4; ++k2) a[j1 * 4 + k2] = ON(cmul)(c[k2], ON(kw)(j1 * k2, 16));
    }
    for (int k2 = 0; k2 < 4; ++k2) {
        OC c[4] = {a[k2], a[4 + k2], a[8 + k2], a[12 + k2]};
        ON(bfly4)(c);
        for (int k1 = 0; k1 < 4; ++k1) v[k2 + 4 * k1] = c[k1];
    }
}

/* radix_3_codelets.jl:338-366. */
static void ON(bfly3)(OC* v) {
    const OT h = (OT)0.5L, s3 = (OT)0.866025403784438646763723170752936183L;
    OC a = v[0], b = v[1], c = v[2];
    OC sp = ON(cadd)(b, c), sm = ON(csub)(b, c);
    OC m = ON(csub)(a, ON(cscale)(sp, h));
    OC d = ON(cmnj)(ON(cscale)(sm, s3));
    v[0] = ON(cadd)(a, sp);
    v[1] = ON(cadd)(m, d);
    v[2] = ON(csub)(m, d);
}

/* Odd prime radix R by conjugate pairs (radix_5_codelets.jl:27-44, radix_7_codelets.jl:30-51):
 * X_k = v0 + sum_{j=1}^{(R-1)/2} [ (v_j+v_{R-j}) cos(2pi jk/R) - i (v_j - v_{R-j}) sin(2pi jk/R) ] */
static void ON(bfly_oddprime)(OC* v, int R) {
    OC sp[8], sm[8], out[17];
    int h = (R - 1) / 2;
    for (int j = 1; j <= h; ++j) {
        sp[j] = ON(cadd)(v[j], v[R - j]);
        sm[j] = ON(csub)(v[j], v[R - j]);
    }
    OC dc = v[0];
    for (int j = 1; j <= h; ++j) dc = ON(cadd)(dc, sp[j]);
    out[0] = dc;
    for (int k = 1; k <= h; ++k) {
        OC re = v[0], im = ON(cmk)(0, 0);
        for (int j = 1; j <= h; ++j) {
            int jk = (j * k) % R;
            re = ON(cadd)(re, ON(cscale)(sp[j], ON(kcos)[R][jk]));
            im = ON(cadd)(im, ON(cscale)(sm[j], ON(ksin)[R][jk]));
        }
        OC t = ON(cmnj)(im);
        out[k] = ON(cadd)(re, t);
        out[R - k] = ON(csub)(re, t);
    }
    for (int k = 0; k < R; ++k) v[k] = out[k];
}

/* radix_3_codelets.jl:12-74: 9 = 3 x 3 with rotations w9^{jk}. */
static void ON(bfly9)(OC* v) {
    OC a[9];
    for (int j1 = 0; j1 < 3; ++j1) {
        OC c[3] = {v[j1], v[j1 + 3], v[j1 + 6]};
        ON(bfly3)(c);
        for (int k2 = 0; k2 < 3; ++k2) a[j1 * 3 + k2] = ON(cmul)(c[k2], ON(kw)(j1 * k2, 9));
    }
    for (int k2 = 0; k2 < 3; ++k2) {
        OC c[3] = {a[k2], a[3 + k2], a[6 + k2]};
        ON(bfly3)(c);
        for (int k1 = 0; k1 < 3; ++k1) v[k2 + 3 * k1] = c[k1];
    }
}

static void ON(bfly)(OC* v, int R) {
    switch (R) {
        case 2: ON(bfly2)(v); break;
        case 3: ON(bfly3)(v); break;
        case 4: ON(bfly4)(v); break;
        case 5: ON(bfly_oddprime)(v, 5); break;
        case 7: ON(bfly_oddprime)(v, 7); break;
        case 8: ON(bfly8)(v); break;
        case 9: ON(bfly9)(v); break;
        case 16: ON(bfly16)(v); break;
        default: abort();
    }
}

/* ---------------------------------------------------------------------------------------
 * Plan tree, restating the closures that kernel.jl:72-92 caches.
 * ------------------------------------------------------------------------------------- */
typedef struct ON(onode) ON(onode);
struct ON(onode) {
    int kind;                 /* ORC_TRIVIAL / ORC_RADIX / ORC_MIXED / ORC_RADER */
    long n;
    /* radix */
    int nops;
    int radix[64];
    OC* tw[64];               /* per-op twiddle table, [p][k-1], w_g^{p k} */
    /* mixed: n = p * m */
    long p, m;
    OC* dtw;                  /* four-step twiddles w_n^{i k2}, [i][k2] */
    ON(onode) *fm, *fp;
    /* rader */
    long g;
    long* perm;               /* g^i mod n, i < n-1 */
    long* iperm;              /* (g^{-1})^i mod n, i < n-1 */
    OC* what;                 /* FFT_{n-1}(W), computed once */
    ON(onode)* fsub;
    size_t work;              /* workspace elements needed by exec */
};

static ON(onode)* ON(node_build)(long n);

/* radix_plan.jl:100-138 + kernel.jl:131-139; twiddles rounded once from long double. */
static ON(onode)* ON(build_radix)(long n) {
    ON(onode)* nd = (ON(onode)*)calloc(1, sizeof(ON(onode)));
    nd->kind = ORC_RADIX;
    nd->n = n;
    nd->nops = orc_std_radix_plan(n, nd->radix, 64);
    long s = 1;
    for (int i = 0; i < nd->nops; ++i) {
        int R = nd->radix[i];
        long g = n / s, n1 = g / R;
        nd->tw[i] = (OC*)malloc(sizeof(OC) * (size_t)(n1 * (R - 1) + 1));
        for (long p = 0; p < n1; ++p)
            for (int k = 1; k < R; ++k) nd->tw[i][p * (R - 1) + (k - 1)] = ON(omega)(p * k, g);
        s *= R;
    }
    nd->work = (size_t)n;
    return nd;
}

/* mixed_radix.jl:19-41 (plan + twiddle matrix D, here rounded once instead of chained). */
static ON(onode)* ON(build_mixed)(long n) {
    ON(onode)* nd = (ON(onode)*)calloc(1, sizeof(ON(onode)));
    nd->kind = ORC_MIXED;
    nd->n = n;
    orc_closest_factors(n, &nd->p, &nd->m);
    long p = nd->p, m = nd->m;
    nd->dtw = (OC*)malloc(sizeof(OC) * (size_t)(p * m));
    for (long i = 0; i < p; ++i)
        for (long k2 = 0; k2 < m; ++k2) nd->dtw[i * m + k2] = ON(omega)(i * k2, n);
    nd->fm = ON(node_build)(m);
    nd->fp = ON(node_build)(p);
    size_t sub = nd->fm->work > nd->fp->work ? nd->fm->work : nd->fp->work;
    nd->work = 2 * (size_t)n + 2 * (size_t)(m > p ? m : p) + sub;
    return nd;
}

static void ON(node_exec)(const ON(onode)* nd, const OC* x, OC* y, OC* work);

/* prime.jl:5-40 (generator, permutations, W) with FFT(W) precomputed once. */
static ON(onode)* ON(build_rader)(long n) {
    ON(onode)* nd = (ON(onode)*)calloc(1, sizeof(ON(onode)));
    nd->kind = ORC_RADER;
    nd->n = n;
    long N = n - 1;
    nd->g = orc_primitive_root(n);
    long ginv = orc_powmod(nd->g, n - 2, n);
    nd->perm = (long*)malloc(sizeof(long) * (size_t)N);
    nd->iperm = (long*)malloc(sizeof(long) * (size_t)N);
    for (long i = 0; i < N; ++i) {
        nd->perm[i] = orc_powmod(nd->g, i, n);
        nd->iperm[i] = orc_powmod(ginv, i, n);
    }
    nd->fsub = ON(node_build)(N);
    /* W[i] = w_n^{-(-iperm[i])}... prime.jl:36-37: W[i] = cispi(2/n)^(-inv_seq[i]) = e^{-2 pi i inv_seq[i]/n} */
    OC* W = (OC*)malloc(sizeof(OC) * (size_t)N);
    for (long i = 0; i < N; ++i) W[i] = ON(omega)(nd->iperm[i], n);
    nd->what = (OC*)malloc(sizeof(OC) * (size_t)N);
    OC* wk = (OC*)malloc(sizeof(OC) * (nd->fsub->work + 1));
    ON(node_exec)(nd->fsub, W, nd->what, wk);
    free(wk);
    free(W);
    nd->work = 2 * (size_t)N + nd->fsub->work;
    return nd;
}

/* kernel.jl:72-92 size classification. */
static ON(onode)* ON(node_build)(long n) {
    if (n == 1) {
        ON(onode)* nd = (ON(onode)*)calloc(1, sizeof(ON(onode)));
        nd->kind = ORC_TRIVIAL;
        nd->n = 1;
        return nd;
    }
    if (orc_is_power_of(n, 2) || orc_is_power_of(n, 3) || orc_is_power_of(n, 5) ||
        orc_is_power_of(n, 7))
        return ON(build_radix)(n);
    if (orc_isprime(n)) return ON(build_rader)(n);
    return ON(build_mixed)(n);
}

static void ON(node_free)(ON(onode)* nd) {
    if (!nd) return;
    for (int i = 0; i < nd->nops; ++i) free(nd->tw[i]);
    free(nd->dtw);
    free(nd->perm);
    free(nd->iperm);
    free(nd->what);
    ON(node_free)(nd->fm);
    ON(node_free)(nd->fp);
    ON(node_free)(nd->fsub);
    free(nd);
}

/* One layered Stockham pass (radix_2_codelets.jl:402-447 and its radix-3/5/7 twins):
 *   out[q + s(Rp + k)] = w_g^{pk} * sum_j in[q + s(p + j n1)] w_R^{jk},  q < s, p < n1. */
static void ON(stockham_pass)(const OC* in, OC* out, long s, long n1, int R, const OC* tw) {
    OC v[16];
    for (long p = 0; p < n1; ++p) {
        const OC* twp = tw + p * (R - 1);
        for (long q = 0; q < s; ++q) {
            for (int j = 0; j < R; ++j) v[j] = in[q + s * (p + j * n1)];
            ON(bfly)(v, R);
            out[q + s * (R * p)] = v[0];
            if (p == 0) {
                for (int k = 1; k < R; ++k) out[q + s * (R * p + k)] = v[k];
            } else {
                for (int k = 1; k < R; ++k) out[q + s * (R * p + k)] = ON(cmul)(v[k], twp[k - 1]);
            }
        }
    }
}

/* radix_exec.jl:49-70,101: passes ping-pong between two buffers; the last pass lands in y. */
static void ON(exec_radix)(const ON(onode)* nd, const OC* x, OC* y, OC* work) {
    long n = nd->n, s = 1;
    const OC* src = x;
    /* choose the first destination so that the final pass writes y */
    OC* bufs[2] = {y, work};
    int dst = (nd->nops % 2 == 1) ? 0 : 1;
    for (int i = 0; i < nd->nops; ++i) {
        int R = nd->radix[i];
        long g = n / s, n1 = g / R;
        OC* out = bufs[dst];
        ON(stockham_pass)(src, out, s, n1, R, nd->tw[i]);
        src = out;
        dst ^= 1;
        s *= R;
    }
}

/* mixed_radix.jl:44-78: four-step n = p*m. */
static void ON(exec_mixed)(const ON(onode)* nd, const OC* x, OC* y, OC* work) {
    long p = nd->p, m = nd->m;
    long mx = m > p ? m : p;
    OC* Y = work;                 /* p x m, row i at Y[i*m]  */
    OC* rowin = work + p * m;     /* gathered row / column   */
    OC* rowout = rowin + mx;
    OC* sub = rowout + mx;
    /* Step 1: rows i of X[i, j] = x[i + p j], DFT_m over j. */
    for (long i = 0; i < p; ++i) {
        for (long j = 0; j < m; ++j) rowin[j] = x[i + p * j];
        ON(node_exec)(nd->fm, rowin, Y + i * m, sub);
    }
    /* Step 2: twiddle w_n^{i k2} (mixed_radix.jl:61-65). */
    for (long i = 1; i < p; ++i)
        for (long k2 = 1; k2 < m; ++k2) Y[i * m + k2] = ON(cmul)(Y[i * m + k2], nd->dtw[i * m + k2]);
    /* Step 3+4: column k2, DFT_p over i, written to y[k2 + m k1]. */
    for (long k2 = 0; k2 < m; ++k2) {
        for (long i = 0; i < p; ++i) rowin[i] = Y[i * m + k2];
        ON(node_exec)(nd->fp, rowin, rowout, sub);
        for (long k1 = 0; k1 < p; ++k1) y[k2 + m * k1] = rowout[k1];
    }
}

/* prime.jl:42-81: Rader's algorithm with the conj trick for the inverse sub-FFT. */
static void ON(exec_rader)(const ON(onode)* nd, const OC* x, OC* y, OC* work) {
    long n = nd->n, N = n - 1;
    OC* b1 = work;
    OC* b2 = work + N;
    OC* sub = work + 2 * N;
    OC x0 = x[0];
    OC sum = x0;
    for (long i = 0; i < N; ++i) {
        b1[i] = x[nd->perm[i]];
        sum = ON(cadd)(sum, b1[i]);
    }
    ON(node_exec)(nd->fsub, b1, b2, sub);
    for (long i = 0; i < N; ++i) b2[i] = ON(cconj)(ON(cmul)(b2[i], nd->what[i]));
    ON(node_exec)(nd->fsub, b2, b1, sub);
    OT inv = (OT)1 / (OT)N;
    y[0] = sum;
    for (long j = 0; j < N; ++j) y[nd->iperm[j]] = ON(cadd)(x0, ON(cscale)(ON(cconj)(b1[j]), inv));
}

static void ON(node_exec)(const ON(onode)* nd, const OC* x, OC* y, OC* work) {
    switch (nd->kind) {
        case ORC_TRIVIAL: y[0] = x[0]; break;
        case ORC_RADIX: ON(exec_radix)(nd, x, y, work); break;
        case ORC_MIXED: ON(exec_mixed)(nd, x, y, work); break;
        case ORC_RADER: ON(exec_rader)(nd, x, y, work); break;
    }
}

/* ---------------------------------------------------------------------------------------
 * Exported API (suffix f/d/l).
 * ------------------------------------------------------------------------------------- */
void* ON(orc_plan_create)(long n) {
    if (n < 1) return NULL;
    ON(consts_init)();
    return ON(node_build)(n);
}
void ON(orc_plan_destroy)(void* h) { ON(node_free)((ON(onode)*)h); }
long ON(orc_plan_work)(void* h) { return (long)((ON(onode)*)h)->work + 1; }

/* CatabraFFT.jl:65-115 conventions: dir=-1 forward; dir=+1 backward = conj(fft(conj x));
 * normalize=1 divides by n (ifft).  Batched rows at distance n.  nthreads>1 splits rows. */
int ON(orc_exec_batch)(void* h, const OT* xin, OT* yout, long batch, int dir, int normalize,
                       int nthreads) {
    const ON(onode)* nd = (const ON(onode)*)h;
    long n = nd->n;
    const OC* x = (const OC*)xin;
    OC* y = (OC*)yout;
    size_t wk = nd->work + 1;
    OT scale = normalize ? (OT)1 / (OT)n : (OT)1;
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
    {
        OC* work = (OC*)malloc(sizeof(OC) * wk);
        OC* tmp = (dir > 0) ? (OC*)malloc(sizeof(OC) * (size_t)n) : NULL;
#pragma omp for schedule(static)
        for (long b = 0; b < batch; ++b) {
            const OC* xb = x + b * n;
            OC* yb = y + b * n;
            if (dir > 0) {
                for (long i = 0; i < n; ++i) tmp[i] = ON(cconj)(xb[i]);
                ON(node_exec)(nd, tmp, yb, work);
                for (long i = 0; i < n; ++i) yb[i] = ON(cscale)(ON(cconj)(yb[i]), scale);
            } else {
                ON(node_exec)(nd, xb, yb, work);
                if (normalize)
                    for (long i = 0; i < n; ++i) yb[i] = ON(cscale)(yb[i], scale);
            }
        }
        free(work);
        free(tmp);
    }
    return 0;
}

/* One Stockham pass exposed for pass-level tests (radix_2_codelets.jl:402-447 semantics). */
void ON(orc_stockham_pass)(const OT* in, OT* out, long n, long s, int R) {
    ON(consts_init)();
    long g = n / s, n1 = g / R;
    OC* tw = (OC*)malloc(sizeof(OC) * (size_t)(n1 * (R - 1) + 1));
    for (long p = 0; p < n1; ++p)
        for (int k = 1; k < R; ++k) tw[p * (R - 1) + (k - 1)] = ON(omega)(p * k, g);
    ON(stockham_pass)((const OC*)in, (OC*)out, s, n1, R, tw);
    free(tw);
}

/* One codelet exposed for codelet tests. */
void ON(orc_codelet)(OT* v, int R) {
    ON(consts_init)();
    ON(bfly)((OC*)v, R);
}

#undef OC
#undef ON
#undef OCAT
#undef OCAT2
