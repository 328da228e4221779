/*
 * kf_real.c -- real-input 1-D transforms (reference kiss_fftr.c:21-155).
 *
 * The plan is a half-length complex cfg plus nfft/4 "super twiddles"
 * exp(-+i*pi*((k+1)/(nfft/2) + 1/2)) built in FP64 and rounded once, exactly as
 * kiss_fftr.c:53-59.  The GPU runs the half-length transform and the split (forward) or
 * merge (inverse) step in one fused pass; cfg->tmpbuf is kept for the reference layout
 * but never written, so a real cfg may be shared between threads.
 */
#include <math.h>

#include "kf_internal.h"

kiss_fftr_cfg kiss_fftr_alloc(int nfft, int inverse_fft, void *mem, size_t *lenmem)
{
    kiss_fftr_cfg st = NULL;
    size_t subsize = 0;
    if (nfft & 1) {
        KF_ERROR("Real FFT optimization must be even.");
        return NULL;
    }
    if (nfft <= 0)
        return NULL;
    const int half = nfft >> 1;
    kiss_fft_alloc(half, inverse_fft, NULL, &subsize);
    const size_t need = sizeof(struct kiss_fftr_state) + subsize + sizeof(kiss_fft_cpx) * (size_t)(half * 3 / 2);
    if (lenmem == NULL) {
        st = (kiss_fftr_cfg)KISS_FFT_MALLOC(need);
    } else {
        if (mem != NULL && *lenmem >= need)
            st = (kiss_fftr_cfg)mem;
        *lenmem = need;
    }
    if (!st)
        return NULL;
    st->substate = (kiss_fft_cfg)(st + 1);
    st->tmpbuf = (kiss_fft_cpx *)((char *)st->substate + subsize);
    st->super_twiddles = st->tmpbuf + half;
    kiss_fft_alloc(half, inverse_fft, st->substate, &subsize);
    for (int k = 0; k < half / 2; ++k) {
        double phase = -3.14159265358979323846264338327 * ((double)(k + 1) / half + .5);
        if (inverse_fft)
            phase *= -1;
        kf_cexp(st->super_twiddles + k, phase);
    }
    return st;
}

void kf_rplan_of(kiss_fftr_cfg st, kfa_rplan *pl)
{
    kf_cplan_of(st->substate, &pl->half);
    pl->super_twiddles = st->super_twiddles;
}

void kiss_fftr(kiss_fftr_cfg st, const kiss_fft_scalar *timedata, kiss_fft_cpx *freqdata)
{
    kfa_rplan pl;
    if (!st || !timedata || !freqdata) {
        KF_ERROR("kiss_fftr: NULL cfg or buffer.");
        return;
    }
    if (st->substate->inverse) {
        KF_ERROR("kiss fft usage error: improper alloc");
        return;
    }
    kf_rplan_of(st, &pl);
    const long long n = 2LL * st->substate->nfft;
    if (kfa_r2c(&pl, timedata, n, freqdata, n / 2 + 1, 1, NULL, 1))
        KF_ERROR("kiss_fftr: %s", kfa_last_error());
}

void kiss_fftri(kiss_fftr_cfg st, const kiss_fft_cpx *freqdata, kiss_fft_scalar *timedata)
{
    kfa_rplan pl;
    if (!st || !timedata || !freqdata) {
        KF_ERROR("kiss_fftri: NULL cfg or buffer.");
        return;
    }
    if (st->substate->inverse == 0) {
        KF_ERROR("kiss fft usage error: improper alloc");
        return;
    }
    kf_rplan_of(st, &pl);
    const long long n = 2LL * st->substate->nfft;
    if (kfa_c2r(&pl, freqdata, n / 2 + 1, timedata, n, 1, NULL, 1))
        KF_ERROR("kiss_fftri: %s", kfa_last_error());
}

static int real_batch(kiss_fftr_cfg st, const void *in, long long idist, void *out, long long odist,
                      long long howmany, void *stream, int sync, int inverse)
{
    kfa_rplan pl;
    if (!st || !in || !out)
        return -2;
    if ((st->substate->inverse != 0) != (inverse != 0))
        return -2;
    kf_rplan_of(st, &pl);
    return inverse ? kfa_c2r(&pl, in, idist, out, odist, howmany, stream, sync)
                   : kfa_r2c(&pl, in, idist, out, odist, howmany, stream, sync);
}

int kiss_fftr_batch(kiss_fftr_cfg cfg, const kiss_fft_scalar *timedata, long long idist, kiss_fft_cpx *freqdata,
                    long long odist, long long howmany)
{
    return real_batch(cfg, timedata, idist, freqdata, odist, howmany, NULL, 1, 0);
}

int kiss_fftri_batch(kiss_fftr_cfg cfg, const kiss_fft_cpx *freqdata, long long idist, kiss_fft_scalar *timedata,
                     long long odist, long long howmany)
{
    return real_batch(cfg, freqdata, idist, timedata, odist, howmany, NULL, 1, 1);
}

int kiss_fftr_batch_async(kiss_fftr_cfg cfg, const kiss_fft_scalar *timedata, long long idist,
                          kiss_fft_cpx *freqdata, long long odist, long long howmany, void *stream)
{
    return real_batch(cfg, timedata, idist, freqdata, odist, howmany, stream, 0, 0);
}

int kiss_fftri_batch_async(kiss_fftr_cfg cfg, const kiss_fft_cpx *freqdata, long long idist,
                           kiss_fft_scalar *timedata, long long odist, long long howmany, void *stream)
{
    return real_batch(cfg, freqdata, idist, timedata, odist, howmany, stream, 0, 1);
}
