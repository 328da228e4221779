/*
 * kfo.c -- CPU restatement of the KISS FFT algorithm, used as the oracle.
 *
 * TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load it as the checker; the product never links or calls it.
 *
 * What is restated, and where the reference states it:
 *   radix schedule ................ kiss_fft.c:306-328        (kfo_factor)
 *   twiddle tables ................ kiss_fft.c:361-367, _kiss_fft_guts.h:127-145
 *   recursive decimation in time .. kiss_fft.c:235-300        (dit)
 *   radix 2/4/3/5/generic stages .. kiss_fft.c:15-233         (stage2..stage_generic)
 *   fixed-point arithmetic ........ _kiss_fft_guts.h:43-124   (fx_* helpers)
 *   real split / merge ............ kiss_fftr.c:21-155        (real_fwd / real_inv)
 *   N-D column passes ............. kiss_fftnd.c:156-188      (nd_one)
 *   real N-D ...................... kiss_fftndr.c:86-132      (ndr_fwd_one / ndr_inv_one)
 *
 * Arithmetic model for the fixed-point builds: every value that the reference holds in
 * a kiss_fft_scalar is truncated to the scalar width when it is stored (two's-complement
 * wrap), every expression is evaluated in the C integer type the reference's macros
 * produce (int for Q15 and for Q31 sums, int64 for Q31 products), and rounding is
 * sround: add half an LSB, arithmetic shift right by the fraction bits.  The float build
 * evaluates each float expression in the same association order as the reference so that
 * its output can be compared with a reference float build bit for bit.
 */
#include "kfo.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#if defined(KFO_Q15) || defined(KFO_Q31)
#define KFO_FIXED 1
#endif

#if defined(KFO_Q15)
#define FRAC 15
#define SAMP_MAXV 32767
#elif defined(KFO_Q31)
#define FRAC 31
#define SAMP_MAXV 2147483647
#endif

typedef kfo_cpx cpx;
typedef int64_t wide; /* holds every fixed-point intermediate of either width exactly */

/* ------------------------------------------------------------------------------------ */
/* Scalar arithmetic layer                                                               */
/* ------------------------------------------------------------------------------------ */
#ifdef KFO_FIXED
/* A store into a kiss_fft_scalar: keep the low bits, two's complement. */
static inline kfo_scalar st(wide x) { return (kfo_scalar)x; }

/* The C `int` value of a sum of two scalars.  Q15 sums never leave int range; Q31 sums
 * of two int32 wrap in 32-bit int arithmetic. */
static inline wide as_int(wide x)
{
#if defined(KFO_Q31)
    return (int32_t)(uint32_t)(uint64_t)x;
#else
    return x;
#endif
}

/* sround(x) then store, _kiss_fft_guts.h:65 */
static inline kfo_scalar rnd(wide x) { return (kfo_scalar)((x + ((wide)1 << (FRAC - 1))) >> FRAC); }

/* S_MUL, _kiss_fft_guts.h:67 */
static inline kfo_scalar smul_r(kfo_scalar a, kfo_scalar b) { return rnd((wide)a * (wide)b); }

static inline kfo_scalar sadd(kfo_scalar a, kfo_scalar b) { return st((wide)a + b); }
static inline kfo_scalar ssub(kfo_scalar a, kfo_scalar b) { return st((wide)a - b); }
static inline kfo_scalar sneg(kfo_scalar a) { return st(-(wide)a); }
/* HALF_OF applied to the int value of a scalar expression (_kiss_fft_guts.h:130) */
static inline wide half_int(wide x) { return as_int(x) >> 1; }

/* C_MUL, _kiss_fft_guts.h:69-71: each component rounded once. */
static inline cpx cmul(cpx a, cpx b)
{
    cpx m;
    m.r = rnd((wide)a.r * b.r - (wide)a.i * b.i);
    m.i = rnd((wide)a.r * b.i + (wide)a.i * b.r);
    return m;
}

/* C_FIXDIV(c, p), _kiss_fft_guts.h:73-78: multiply by floor(SAMP_MAX/p), then sround. */
static inline cpx fixdiv(cpx c, int p)
{
    const wide k = SAMP_MAXV / p;
    c.r = rnd((wide)c.r * k);
    c.i = rnd((wide)c.i * k);
    return c;
}

/* C_MULBYSCALAR, _kiss_fft_guts.h:80-82 */
static inline cpx cscale(cpx c, kfo_scalar s)
{
    c.r = rnd((wide)c.r * s);
    c.i = rnd((wide)c.i * s);
    return c;
}
#else /* float */
static inline kfo_scalar smul_r(kfo_scalar a, kfo_scalar b) { return a * b; }
static inline kfo_scalar sadd(kfo_scalar a, kfo_scalar b) { return a + b; }
static inline kfo_scalar ssub(kfo_scalar a, kfo_scalar b) { return a - b; }
static inline kfo_scalar sneg(kfo_scalar a) { return -a; }
static inline cpx cmul(cpx a, cpx b)
{
    cpx m;
    m.r = a.r * b.r - a.i * b.i;
    m.i = a.r * b.i + a.i * b.r;
    return m;
}
static inline cpx fixdiv(cpx c, int p) { (void)p; return c; }
static inline cpx cscale(cpx c, kfo_scalar s)
{
    c.r *= s;
    c.i *= s;
    return c;
}
#endif

static inline cpx cadd(cpx a, cpx b)
{
    cpx r;
    r.r = sadd(a.r, b.r);
    r.i = sadd(a.i, b.i);
    return r;
}
static inline cpx csub(cpx a, cpx b)
{
    cpx r;
    r.r = ssub(a.r, b.r);
    r.i = ssub(a.i, b.i);
    return r;
}

/* ------------------------------------------------------------------------------------ */
/* Plan                                                                                 */
/* ------------------------------------------------------------------------------------ */
#define KFO_MAXPAIRS 32

typedef struct {
    int n;
    int inverse;
    int npairs;
    int pairs[2 * KFO_MAXPAIRS];
    cpx *tw;
} plan_t;

int kfo_factor(int n, int *pairs)
{
    /* Strip 4s, then 2s, then odd trial divisors up to floor(sqrt(n0)); a cofactor left
     * once the divisor passes that bound is taken whole.  kiss_fft.c:306-328. */
    const double lim = floor(sqrt((double)n));
    int p = 4, cnt = 0;
    if (n <= 1) {
        pairs[0] = n < 1 ? 1 : n;
        pairs[1] = 1;
        return 1;
    }
    while (n > 1) {
        while (n % p) {
            if (p == 4)
                p = 2;
            else if (p == 2)
                p = 3;
            else
                p += 2;
            if (p > lim)
                p = n;
        }
        n /= p;
        pairs[2 * cnt] = p;
        pairs[2 * cnt + 1] = n;
        ++cnt;
    }
    return cnt;
}

static void twiddle_at(cpx *dst, double phase)
{
#ifdef KFO_FIXED
    dst->r = (kfo_scalar)floor(.5 + SAMP_MAXV * cos(phase));
    dst->i = (kfo_scalar)floor(.5 + SAMP_MAXV * sin(phase));
#else
    dst->r = (kfo_scalar)cos(phase);
    dst->i = (kfo_scalar)sin(phase);
#endif
}

void kfo_twiddles(int n, int inverse, cpx *tw)
{
    /* The phase expression must round exactly as kiss_fft.c:362-365 does, because the
     * fixed-point tables hit x.5 ties that flip with the last bit of the phase. */
    const double pi = 3.141592653589793238462643383279502884197169399375105820974944;
    for (int k = 0; k < n; ++k) {
        double phase = -2 * pi * k / n;
        if (inverse)
            phase *= -1;
        twiddle_at(tw + k, phase);
    }
}

void kfo_super_twiddles(int n, int inverse, cpx *stw)
{
    /* kiss_fftr.c:53-59, with ncfft = n/2 */
    const int ncfft = n / 2;
    for (int k = 0; k < ncfft / 2; ++k) {
        double phase = -3.14159265358979323846264338327 * ((double)(k + 1) / ncfft + .5);
        if (inverse)
            phase *= -1;
        twiddle_at(stw + k, phase);
    }
}

int kfo_next_fast_size(int n)
{
    for (;; ++n) {
        int m = n;
        while (m % 2 == 0)
            m /= 2;
        while (m % 3 == 0)
            m /= 3;
        while (m % 5 == 0)
            m /= 5;
        if (m <= 1)
            return n;
    }
}

static int plan_init(plan_t *pl, int n, int inverse)
{
    pl->n = n;
    pl->inverse = inverse;
    pl->npairs = kfo_factor(n, pl->pairs);
    pl->tw = (cpx *)malloc(sizeof(cpx) * (size_t)(n > 0 ? n : 1));
    if (!pl->tw)
        return -1;
    kfo_twiddles(n, inverse, pl->tw);
    return 0;
}

static void plan_free(plan_t *pl)
{
    free(pl->tw);
    pl->tw = NULL;
}

/* ------------------------------------------------------------------------------------ */
/* Stage butterflies.  y holds p sub-transforms of length m back to back; stage u combines */
/* y[u + q*m], q < p, with twiddles tw[q*u*fs] (fs * p * m == n).                         */
/* ------------------------------------------------------------------------------------ */

/* kiss_fft.c:15-36 */
static void stage2(const plan_t *pl, cpx *y, int fs, int m)
{
    for (int u = 0; u < m; ++u) {
        cpx a = fixdiv(y[u], 2), b = fixdiv(y[u + m], 2);
        cpx t = cmul(b, pl->tw[(size_t)u * fs]);
        y[u + m] = csub(a, t);
        y[u] = cadd(a, t);
    }
}

/* kiss_fft.c:38-84 */
static void stage4(const plan_t *pl, cpx *y, int fs, int m)
{
    for (int u = 0; u < m; ++u) {
        cpx x0 = fixdiv(y[u], 4), x1 = fixdiv(y[u + m], 4);
        cpx x2 = fixdiv(y[u + 2 * m], 4), x3 = fixdiv(y[u + 3 * m], 4);
        cpx s0 = cmul(x1, pl->tw[(size_t)u * fs]);
        cpx s1 = cmul(x2, pl->tw[(size_t)2 * u * fs]);
        cpx s2 = cmul(x3, pl->tw[(size_t)3 * u * fs]);
        cpx s5 = csub(x0, s1);
        x0 = cadd(x0, s1);
        cpx s3 = cadd(s0, s2);
        cpx s4 = csub(s0, s2);
        y[u + 2 * m] = csub(x0, s3);
        y[u] = cadd(x0, s3);
        cpx o1, o3;
        if (pl->inverse) {
            o1.r = ssub(s5.r, s4.i);
            o1.i = sadd(s5.i, s4.r);
            o3.r = sadd(s5.r, s4.i);
            o3.i = ssub(s5.i, s4.r);
        } else {
            o1.r = sadd(s5.r, s4.i);
            o1.i = ssub(s5.i, s4.r);
            o3.r = ssub(s5.r, s4.i);
            o3.i = sadd(s5.i, s4.r);
        }
        y[u + m] = o1;
        y[u + 3 * m] = o3;
    }
}

/* kiss_fft.c:86-128 */
static void stage3(const plan_t *pl, cpx *y, int fs, int m)
{
    const cpx epi3 = pl->tw[(size_t)fs * m];
    for (int u = 0; u < m; ++u) {
        cpx x0 = fixdiv(y[u], 3), x1 = fixdiv(y[u + m], 3), x2 = fixdiv(y[u + 2 * m], 3);
        cpx s1 = cmul(x1, pl->tw[(size_t)u * fs]);
        cpx s2 = cmul(x2, pl->tw[(size_t)2 * u * fs]);
        cpx s3 = cadd(s1, s2);
        cpx s0 = csub(s1, s2);
        cpx o1;
#ifdef KFO_FIXED
        o1.r = st((wide)x0.r - ((wide)s3.r >> 1));
        o1.i = st((wide)x0.i - ((wide)s3.i >> 1));
#else
        o1.r = x0.r - s3.r * (kfo_scalar).5;
        o1.i = x0.i - s3.i * (kfo_scalar).5;
#endif
        s0 = cscale(s0, epi3.i);
        x0 = cadd(x0, s3);
        cpx o2;
        o2.r = sadd(o1.r, s0.i);
        o2.i = ssub(o1.i, s0.r);
        o1.r = ssub(o1.r, s0.i);
        o1.i = sadd(o1.i, s0.r);
        y[u] = x0;
        y[u + m] = o1;
        y[u + 2 * m] = o2;
    }
}

/* kiss_fft.c:130-189 */
static void stage5(const plan_t *pl, cpx *y, int fs, int m)
{
    const cpx ya = pl->tw[(size_t)fs * m];
    const cpx yb = pl->tw[(size_t)2 * fs * m];
    for (int u = 0; u < m; ++u) {
        cpx x0 = fixdiv(y[u], 5), x1 = fixdiv(y[u + m], 5), x2 = fixdiv(y[u + 2 * m], 5);
        cpx x3 = fixdiv(y[u + 3 * m], 5), x4 = fixdiv(y[u + 4 * m], 5);
        const cpx c0 = x0;
        cpx c1 = cmul(x1, pl->tw[(size_t)u * fs]);
        cpx c2 = cmul(x2, pl->tw[(size_t)2 * u * fs]);
        cpx c3 = cmul(x3, pl->tw[(size_t)3 * u * fs]);
        cpx c4 = cmul(x4, pl->tw[(size_t)4 * u * fs]);
        cpx c7 = cadd(c1, c4), c10 = csub(c1, c4);
        cpx c8 = cadd(c2, c3), c9 = csub(c2, c3);

        x0.r = sadd(x0.r, sadd(c7.r, c8.r));
        x0.i = sadd(x0.i, sadd(c7.i, c8.i));

        cpx c5, c6, c11, c12;
        c5.r = sadd(sadd(c0.r, smul_r(c7.r, ya.r)), smul_r(c8.r, yb.r));
        c5.i = sadd(sadd(c0.i, smul_r(c7.i, ya.r)), smul_r(c8.i, yb.r));
        c6.r = sadd(smul_r(c10.i, ya.i), smul_r(c9.i, yb.i));
        c6.i = ssub(sneg(smul_r(c10.r, ya.i)), smul_r(c9.r, yb.i));

        y[u + m] = csub(c5, c6);
        y[u + 4 * m] = cadd(c5, c6);

        c11.r = sadd(sadd(c0.r, smul_r(c7.r, yb.r)), smul_r(c8.r, ya.r));
        c11.i = sadd(sadd(c0.i, smul_r(c7.i, yb.r)), smul_r(c8.i, ya.r));
        c12.r = sadd(sneg(smul_r(c10.i, yb.i)), smul_r(c9.i, ya.i));
        c12.i = ssub(smul_r(c10.r, yb.i), smul_r(c9.r, ya.i));

        y[u + 2 * m] = cadd(c11, c12);
        y[u + 3 * m] = csub(c11, c12);
        y[u] = x0;
    }
}

/* kiss_fft.c:191-233: direct DFT of each column, twiddle and kernel folded into one
 * table index q*fs*k mod n (k = output index inside the length p*m sub-transform). */
static void stage_generic(const plan_t *pl, cpx *y, int fs, int m, int p)
{
    const int n = pl->n;
    cpx *col = (cpx *)malloc(sizeof(cpx) * (size_t)p);
    if (!col)
        abort();
    for (int u = 0; u < m; ++u) {
        for (int q = 0; q < p; ++q)
            col[q] = fixdiv(y[u + q * m], p);
        for (int q1 = 0; q1 < p; ++q1) {
            const int k = u + q1 * m;
            const long step = (long)fs * k; /* < n */
            long idx = 0;
            cpx acc = col[0];
            for (int q = 1; q < p; ++q) {
                idx += step;
                if (idx >= n)
                    idx -= n;
                acc = cadd(acc, cmul(col[q], pl->tw[idx]));
            }
            y[k] = acc;
        }
    }
    free(col);
}

/* kiss_fft.c:235-300.  Computes the length p*m transform of x[0], x[xstep], ... into y,
 * where (p, m) is the level-th pair; xstep = fs * in_stride. */
static void dit(const plan_t *pl, int level, cpx *y, const cpx *x, long xstep, int fs)
{
    const int p = pl->pairs[2 * level];
    const int m = pl->pairs[2 * level + 1];
    if (m == 1) {
        for (int q = 0; q < p; ++q)
            y[q] = x[q * xstep];
    } else {
        for (int q = 0; q < p; ++q)
            dit(pl, level + 1, y + (long)q * m, x + q * xstep, xstep * p, fs * p);
    }
    switch (p) {
    case 2: stage2(pl, y, fs, m); break;
    case 3: stage3(pl, y, fs, m); break;
    case 4: stage4(pl, y, fs, m); break;
    case 5: stage5(pl, y, fs, m); break;
    default: stage_generic(pl, y, fs, m, p); break;
    }
}

/* kiss_fft_stride semantics: out-of-place, or through a temporary when in == out. */
static void fft1(const plan_t *pl, const cpx *in, int in_stride, cpx *out)
{
    if (pl->n <= 0)
        return;
    if (in == out) {
        cpx *tmp = (cpx *)malloc(sizeof(cpx) * (size_t)pl->n);
        if (!tmp)
            abort();
        dit(pl, 0, tmp, in, in_stride, 1);
        memcpy(out, tmp, sizeof(cpx) * (size_t)pl->n);
        free(tmp);
    } else {
        dit(pl, 0, out, in, in_stride, 1);
    }
}

void kfo_fft(int n, int inverse, const cpx *in, int in_stride, cpx *out, long howmany,
             long idist, long odist)
{
    plan_t pl;
    if (plan_init(&pl, n, inverse))
        abort();
#pragma omp parallel for schedule(static)
    for (long b = 0; b < howmany; ++b)
        fft1(&pl, in + b * idist, in_stride, out + b * odist);
    plan_free(&pl);
}

/* ------------------------------------------------------------------------------------ */
/* Real transforms                                                                      */
/* ------------------------------------------------------------------------------------ */

/* kiss_fftr.c:63-117 for one item; z is scratch of n/2 complex. */
static void real_fwd(const plan_t *half, const cpx *stw, const kfo_scalar *x, cpx *X, cpx *z)
{
    const int nc = half->n;
    fft1(half, (const cpx *)x, 1, z);

    cpx dc = fixdiv(z[0], 2);
    X[0].r = sadd(dc.r, dc.i);
    X[nc].r = ssub(dc.r, dc.i);
    X[0].i = 0;
    X[nc].i = 0;

    for (int k = 1; k <= nc / 2; ++k) {
        cpx a = z[k], b;
        b.r = z[nc - k].r;
        b.i = sneg(z[nc - k].i);
        a = fixdiv(a, 2);
        b = fixdiv(b, 2);
        cpx f1 = cadd(a, b);
        cpx f2 = csub(a, b);
        cpx t = cmul(f2, stw[k - 1]);
#ifdef KFO_FIXED
        X[k].r = st(half_int((wide)f1.r + t.r));
        X[k].i = st(half_int((wide)f1.i + t.i));
        X[nc - k].r = st(half_int((wide)f1.r - t.r));
        X[nc - k].i = st(half_int((wide)t.i - f1.i));
#else
        X[k].r = (f1.r + t.r) * (kfo_scalar).5;
        X[k].i = (f1.i + t.i) * (kfo_scalar).5;
        X[nc - k].r = (f1.r - t.r) * (kfo_scalar).5;
        X[nc - k].i = (t.i - f1.i) * (kfo_scalar).5;
#endif
    }
}

/* kiss_fftr.c:119-155 for one item; z is scratch of n/2 complex. */
static void real_inv(const plan_t *half, const cpx *stw, const cpx *X, kfo_scalar *x, cpx *z)
{
    const int nc = half->n;
    z[0].r = sadd(X[0].r, X[nc].r);
    z[0].i = ssub(X[0].r, X[nc].r);
    z[0] = fixdiv(z[0], 2);
    for (int k = 1; k <= nc / 2; ++k) {
        cpx a = X[k], b;
        b.r = X[nc - k].r;
        b.i = sneg(X[nc - k].i);
        a = fixdiv(a, 2);
        b = fixdiv(b, 2);
        cpx e = cadd(a, b);
        cpx d = csub(a, b);
        cpx o = cmul(d, stw[k - 1]);
        z[k] = cadd(e, o);
        z[nc - k] = csub(e, o);
#ifdef KFO_FIXED
        z[nc - k].i = sneg(z[nc - k].i);
#else
        z[nc - k].i *= -1;
#endif
    }
    fft1(half, z, 1, (cpx *)x);
}

typedef struct {
    plan_t half;
    cpx *stw;
} rplan_t;

static int rplan_init(rplan_t *rp, int n, int inverse)
{
    if (n & 1)
        return -1;
    if (plan_init(&rp->half, n / 2, inverse))
        return -1;
    rp->stw = (cpx *)malloc(sizeof(cpx) * (size_t)(n / 4 + 1));
    if (!rp->stw)
        return -1;
    kfo_super_twiddles(n, inverse, rp->stw);
    return 0;
}

static void rplan_free(rplan_t *rp)
{
    plan_free(&rp->half);
    free(rp->stw);
}

void kfo_fftr(int n, const kfo_scalar *in, cpx *out, long howmany, long idist, long odist)
{
    rplan_t rp;
    if (rplan_init(&rp, n, 0))
        abort();
#pragma omp parallel
    {
        cpx *z = (cpx *)malloc(sizeof(cpx) * (size_t)(n / 2 + 1));
#pragma omp for schedule(static)
        for (long b = 0; b < howmany; ++b)
            real_fwd(&rp.half, rp.stw, in + b * idist, out + b * odist, z);
        free(z);
    }
    rplan_free(&rp);
}

void kfo_fftri(int n, const cpx *in, kfo_scalar *out, long howmany, long idist, long odist)
{
    rplan_t rp;
    if (rplan_init(&rp, n, 1))
        abort();
#pragma omp parallel
    {
        cpx *z = (cpx *)malloc(sizeof(cpx) * (size_t)(n / 2 + 1));
#pragma omp for schedule(static)
        for (long b = 0; b < howmany; ++b)
            real_inv(&rp.half, rp.stw, in + b * idist, out + b * odist, z);
        free(z);
    }
    rplan_free(&rp);
}

/* ------------------------------------------------------------------------------------ */
/* N-D transforms                                                                       */
/* ------------------------------------------------------------------------------------ */

/* kiss_fftnd.c:156-188 for one item.  Axis k (dims[0] first) runs `stride` column
 * transforms of length dims[k] reading every stride-th element and writing each result
 * contiguously, which transposes; the ping-pong is arranged so the last pass writes out. */
static void nd_one(const plan_t *axis, const int *dims, int ndims, const cpx *in, cpx *out,
                   cpx *tmp)
{
    long total = 1;
    for (int k = 0; k < ndims; ++k)
        total *= dims[k];
    const cpx *src = in;
    cpx *dst;
    if (ndims & 1) {
        dst = out;
        if (in == out) {
            memcpy(tmp, in, sizeof(cpx) * (size_t)total);
            src = tmp;
        }
    } else {
        dst = tmp;
    }
    for (int k = 0; k < ndims; ++k) {
        const int len = dims[k];
        const long stride = total / len;
        for (long c = 0; c < stride; ++c)
            fft1(&axis[k], src + c, (int)stride, dst + c * len);
        if (dst == tmp) {
            dst = out;
            src = tmp;
        } else {
            dst = tmp;
            src = out;
        }
    }
}

static long prod_dims(const int *dims, int ndims)
{
    long t = 1;
    for (int k = 0; k < ndims; ++k)
        t *= dims[k];
    return t;
}

void kfo_fftnd(const int *dims, int ndims, int inverse, const cpx *in, cpx *out, long howmany)
{
    plan_t axis[KFO_MAXPAIRS];
    const long total = prod_dims(dims, ndims);
    for (int k = 0; k < ndims; ++k)
        if (plan_init(&axis[k], dims[k], inverse))
            abort();
#pragma omp parallel
    {
        cpx *tmp = (cpx *)malloc(sizeof(cpx) * (size_t)total);
#pragma omp for schedule(static)
        for (long b = 0; b < howmany; ++b)
            nd_one(axis, dims, ndims, in + b * total, out + b * total, tmp);
        free(tmp);
    }
    for (int k = 0; k < ndims; ++k)
        plan_free(&axis[k]);
}

/* kiss_fftndr.c:86-110 for one item: rows of the last axis through the real transform,
 * bins scattered into columns, then one complex N-D transform of the leading axes per
 * bin, gathered back as [rows][bins].  A 1-D shape is the real transform alone. */
static void ndr_fwd_one(const rplan_t *rp, const plan_t *axis, const int *dims, int ndims,
                        const kfo_scalar *x, cpx *X, cpx *work)
{
    const int nreal = dims[ndims - 1];
    const int nbins = nreal / 2 + 1;
    const long rows = prod_dims(dims, ndims - 1);
    cpx *cols = work;                       /* nbins * rows */
    cpx *row = cols + (long)nbins * rows;   /* max(nbins, rows) */
    cpx *z = row + (nbins > rows ? nbins : rows); /* nreal/2 + rows (nd scratch) */
    for (long r = 0; r < rows; ++r) {
        real_fwd(&rp->half, rp->stw, x + r * nreal, row, z);
        for (int k = 0; k < nbins; ++k)
            cols[(long)k * rows + r] = row[k];
    }
    for (int k = 0; k < nbins; ++k) {
        if (ndims > 1)
            nd_one(axis, dims, ndims - 1, cols + (long)k * rows, row, z);
        else
            row[0] = cols[k];
        for (long r = 0; r < rows; ++r)
            X[r * nbins + k] = row[r];
    }
}

/* kiss_fftndr.c:112-132 for one item. */
static void ndr_inv_one(const rplan_t *rp, const plan_t *axis, const int *dims, int ndims,
                        const cpx *X, kfo_scalar *x, cpx *work)
{
    const int nreal = dims[ndims - 1];
    const int nbins = nreal / 2 + 1;
    const long rows = prod_dims(dims, ndims - 1);
    cpx *cols = work;
    cpx *row = cols + (long)nbins * rows;
    cpx *z = row + (nbins > rows ? nbins : rows);
    for (int k = 0; k < nbins; ++k) {
        for (long r = 0; r < rows; ++r)
            row[r] = X[r * nbins + k];
        if (ndims > 1)
            nd_one(axis, dims, ndims - 1, row, cols + (long)k * rows, z);
        else
            cols[k] = row[0];
    }
    for (long r = 0; r < rows; ++r) {
        for (int k = 0; k < nbins; ++k)
            row[k] = cols[(long)k * rows + r];
        real_inv(&rp->half, rp->stw, row, x + r * nreal, z);
    }
}

static void ndr_run(const int *dims, int ndims, int inverse, const void *in, void *out,
                    long howmany)
{
    rplan_t rp;
    plan_t axis[KFO_MAXPAIRS];
    const int nreal = dims[ndims - 1];
    const int nbins = nreal / 2 + 1;
    const long rows = prod_dims(dims, ndims - 1);
    const long work_len = (long)nbins * rows + (nbins > rows ? nbins : rows) + nreal / 2 + rows + 1;
    if (rplan_init(&rp, nreal, inverse))
        abort();
    for (int k = 0; k + 1 < ndims; ++k)
        if (plan_init(&axis[k], dims[k], inverse))
            abort();
#pragma omp parallel
    {
        cpx *work = (cpx *)malloc(sizeof(cpx) * (size_t)work_len);
#pragma omp for schedule(static)
        for (long b = 0; b < howmany; ++b) {
            if (inverse)
                ndr_inv_one(&rp, axis, dims, ndims, (const cpx *)in + b * rows * nbins,
                            (kfo_scalar *)out + b * rows * nreal, work);
            else
                ndr_fwd_one(&rp, axis, dims, ndims, (const kfo_scalar *)in + b * rows * nreal,
                            (cpx *)out + b * rows * nbins, work);
        }
        free(work);
    }
    for (int k = 0; k + 1 < ndims; ++k)
        plan_free(&axis[k]);
    rplan_free(&rp);
}

void kfo_fftndr(const int *dims, int ndims, const kfo_scalar *in, cpx *out, long howmany)
{
    ndr_run(dims, ndims, 0, in, out, howmany);
}

void kfo_fftndri(const int *dims, int ndims, const cpx *in, kfo_scalar *out, long howmany)
{
    ndr_run(dims, ndims, 1, in, out, howmany);
}
