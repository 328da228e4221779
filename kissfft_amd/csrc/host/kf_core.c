/*
 * kf_core.c -- complex 1-D plan and execution (reference kiss_fft.c:302-424).
 *
 * Plan building stays on the host and follows the reference exactly, because the
 * fixed-point results depend on it bit for bit: kf_factor's radix order
 * (kiss_fft.c:306-328) and the FP64 twiddle expression rounded once to the scalar
 * (kiss_fft.c:361-367).  This file is compiled without -ffast-math: the Q15 tables hit
 * x.5 ties that flip with the last bit of the phase.  Execution is handed to the HIP
 * engine through kfa_engine.h.
 */
#include <math.h>
#include <stdint.h>

#include "kf_internal.h"

void kf_cexp(kiss_fft_cpx *dst, double phase)
{
#ifdef FIXED_POINT
    dst->r = (kiss_fft_scalar)floor(.5 + KF_SAMP_MAX * cos(phase));
    dst->i = (kiss_fft_scalar)floor(.5 + KF_SAMP_MAX * sin(phase));
#else
    dst->r = (kiss_fft_scalar)cos(phase);
    dst->i = (kiss_fft_scalar)sin(phase);
#endif
}

/* Radix schedule: 4s first, then a 2, then odd trial divisors up to floor(sqrt(n));
 * a cofactor left when the divisor passes that bound is one (generic) radix.
 * Writes (p, m) pairs, m = remaining length.  (kiss_fft.c:306-328) */
static void kf_schedule(int n, int *pairs)
{
    const double bound = floor(sqrt((double)n));
    int p = 4;
    int k = 0;
    do {
        while (n % p) {
            p = (p == 4) ? 2 : (p == 2) ? 3 : p + 2;
            if (p > bound)
                p = n;
        }
        n /= p;
        pairs[k++] = p;
        pairs[k++] = n;
    } while (n > 1 && k < 2 * KF_MAXFACTORS);
}

kiss_fft_cfg kiss_fft_alloc(int nfft, int inverse_fft, void *mem, size_t *lenmem)
{
    kiss_fft_cfg st = NULL;
    if (nfft <= 0)
        return NULL;
    if ((size_t)nfft >= (SIZE_MAX - 2 * sizeof(struct kiss_fft_state)) / sizeof(kiss_fft_cpx))
        return NULL;
    const size_t need = sizeof(struct kiss_fft_state) + sizeof(kiss_fft_cpx) * (size_t)(nfft - 1);
    if (lenmem == NULL) {
        st = (kiss_fft_cfg)KISS_FFT_MALLOC(need);
    } else {
        if (mem != NULL && *lenmem >= need)
            st = (kiss_fft_cfg)mem;
        *lenmem = need;
    }
    if (!st)
        return NULL;
    st->nfft = nfft;
    st->inverse = inverse_fft;
    {
        /* same expression and association as kiss_fft.c:362-365 */
        const double pi = 3.141592653589793238462643383279502884197169399375105820974944;
        for (int i = 0; i < nfft; ++i) {
            double phase = -2 * pi * i / nfft;
            if (st->inverse)
                phase *= -1;
            kf_cexp(st->twiddles + i, phase);
        }
    }
    memset(st->factors, 0, sizeof st->factors);
    kf_schedule(nfft, st->factors);
    return st;
}

void kf_cplan_of(kiss_fft_cfg st, kfa_cplan *pl)
{
    pl->type = KF_TYPE;
    pl->nfft = st->nfft;
    pl->inverse = st->inverse;
    pl->factors = st->factors;
    pl->twiddles = st->twiddles;
}

void kiss_fft_stride(kiss_fft_cfg st, const kiss_fft_cpx *fin, kiss_fft_cpx *fout, int in_stride)
{
    kfa_cplan pl;
    if (st == NULL || fin == NULL || fout == NULL) {
        KF_ERROR("kiss_fft: NULL cfg or buffer.");
        return;
    }
    kf_cplan_of(st, &pl);
    if (kfa_c2c(&pl, fin, in_stride, 0, fout, 0, 1, NULL, 1))
        KF_ERROR("kiss_fft: %s", kfa_last_error());
}

void kiss_fft(kiss_fft_cfg cfg, const kiss_fft_cpx *fin, kiss_fft_cpx *fout)
{
    kiss_fft_stride(cfg, fin, fout, 1);
}

void kiss_fft_cleanup(void)
{
    kfa_cleanup();
}

int kiss_fft_next_fast_size(int n)
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

/* ---- batched / asynchronous (kiss_fft_batch.h) ---------------------------------- */

int kiss_fft_batch(kiss_fft_cfg cfg, const kiss_fft_cpx *fin, int in_stride, long long idist,
                   kiss_fft_cpx *fout, long long odist, long long howmany)
{
    kfa_cplan pl;
    if (!cfg || !fin || !fout)
        return -2;
    kf_cplan_of(cfg, &pl);
    return kfa_c2c(&pl, fin, in_stride, idist, fout, odist, howmany, NULL, 1);
}

int kiss_fft_batch_async(kiss_fft_cfg cfg, const kiss_fft_cpx *fin, int in_stride, long long idist,
                         kiss_fft_cpx *fout, long long odist, long long howmany, void *stream)
{
    kfa_cplan pl;
    if (!cfg || !fin || !fout)
        return -2;
    kf_cplan_of(cfg, &pl);
    return kfa_c2c(&pl, fin, in_stride, idist, fout, odist, howmany, stream, 0);
}

int kiss_fft_batch_noalloc(kiss_fft_cfg cfg, const kiss_fft_cpx *fin, int in_stride, long long idist,
                           kiss_fft_cpx *fout, long long odist, long long howmany, void *stream)
{
    kfa_cplan pl;
    if (!cfg || !fin || !fout)
        return -2;
    kf_cplan_of(cfg, &pl);
    return kfa_c2c_noalloc(&pl, fin, in_stride, idist, fout, odist, howmany, stream);
}

int kiss_fft_stream_sync(void *stream)
{
    return kfa_stream_sync(stream);
}

const char *kiss_fft_last_error(void)
{
    return kfa_last_error();
}

int kiss_fft_device_available(void)
{
    return kfa_device_available();
}

void kiss_fft_get_engine_stats(kiss_fft_engine_stats *stats)
{
    kfa_stats s;
    kfa_get_stats(&s);
    stats->plans_built = s.plans_built;
    stats->plan_hits = s.plan_hits;
    stats->launches = s.launches;
    stats->h2d_bytes = s.h2d_bytes;
    stats->d2h_bytes = s.d2h_bytes;
}
