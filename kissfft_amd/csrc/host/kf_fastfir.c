/*
 * kf_fastfir.c -- overlap-scrap fast FIR (include/kiss_fastfir.h), the plan and stream
 * bookkeeping of the reference's tools/kiss_fastfir.c:55-246 around one GPU launch per
 * call.
 *
 * Plan (kiss_fastfir.c:65-153): block length nfft, ngood = nfft - nh + 1 good samples per
 * block; the impulse response is rotated left by nh-1 (h[nh-1] first, then zeros, then
 * h[0..nh-2] at the end) so that the wrap-around lands in the block's tail; its spectrum,
 * scaled by 1/nfft, is computed here in FP64 and rounded once.  For the real filter the
 * full complex spectrum of the real response is kept: the GPU filters two real blocks per
 * complex transform.
 */
#include <math.h>
#include <string.h>

#include "kf_internal.h"
#include "kiss_fastfir.h"

struct kiss_fastfir_state {
    size_t nfft;
    size_t ngood;
    int real;
    kiss_fft_cpx *H;  /* nfft complex, spectrum / nfft */
    void *tmpbuf;     /* nfft samples, the flush block */
};

#ifndef FIXED_POINT
/* in-place radix-2 FFT in FP64, forward, n a power of two */
static void fft_f64(double *re, double *im, size_t n)
{
    size_t i, j, len, k;
    for (i = 1, j = 0; i < n; ++i) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1)
            j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i];
            re[i] = re[j];
            re[j] = t;
            t = im[i];
            im[i] = im[j];
            im[j] = t;
        }
    }
    for (len = 2; len <= n; len <<= 1)
        for (k = 0; k < len / 2; ++k) {
            const double ang = -2.0 * 3.14159265358979323846264338327 * (double)k / (double)len;
            const double wr = cos(ang), wi = sin(ang);
            for (i = 0; i < n; i += len) {
                const size_t a = i + k, b = i + k + len / 2;
                const double xr = re[b] * wr - im[b] * wi, xi = re[b] * wi + im[b] * wr;
                re[b] = re[a] - xr;
                im[b] = im[a] - xi;
                re[a] += xr;
                im[a] += xi;
            }
        }
}
#endif

static kiss_fastfir_cfg ff_alloc(const void *h, size_t nh, int real, size_t *pnfft, void *mem, size_t *lenmem)
{
#ifdef FIXED_POINT
    (void)h;
    (void)nh;
    (void)real;
    (void)pnfft;
    (void)mem;
    (void)lenmem;
    KF_ERROR("kiss_fastfir_alloc: float builds only (the spectrum is scaled by 1/nfft in float)");
    return NULL;
#else
    size_t nfft = pnfft ? *pnfft : 0, i;
    if (nh < 1)
        return NULL;
    if (nfft == 0) {
        /* next power of two at least 2x the impulse response length, kiss_fastfir.c:77-89 */
        size_t k = nh - 1;
        nfft = 2;
        do {
            nfft <<= 1;
        } while (k >>= 1);
        if (nfft < (real ? 2048u : 1024u))
            nfft = real ? 2048u : 1024u;
    }
    if (pnfft)
        *pnfft = nfft;
    if (nfft < 256 || nfft > 8192 || (nfft & (nfft - 1)) || nh > nfft) {
        KF_ERROR("kiss_fastfir_alloc: block length %zu unsupported (powers of two 256..8192, >= taps)", nfft);
        return NULL;
    }
    const size_t samp = real ? sizeof(kiss_fft_scalar) : sizeof(kiss_fft_cpx);
    const size_t need = sizeof(struct kiss_fastfir_state) + sizeof(kiss_fft_cpx) * nfft + samp * nfft;
    kiss_fastfir_cfg st = NULL;
    if (lenmem == NULL) {
        st = (kiss_fastfir_cfg)malloc(need);
    } else {
        if (mem && *lenmem >= need)
            st = (kiss_fastfir_cfg)mem;
        *lenmem = need;
    }
    if (!st)
        return NULL;
    st->nfft = nfft;
    st->ngood = nfft - nh + 1;
    st->real = real;
    st->H = (kiss_fft_cpx *)(st + 1);
    st->tmpbuf = (void *)(st->H + nfft);

    double *re = (double *)calloc(nfft, sizeof(double)), *im = (double *)calloc(nfft, sizeof(double));
    if (!re || !im) {
        free(re);
        free(im);
        if (!lenmem)
            free(st);
        return NULL;
    }
    /* the rotated impulse response, kiss_fastfir.c:137-141 */
    for (i = 0; i < nh; ++i) {
        const size_t at = (i == nh - 1) ? 0 : nfft - nh + 1 + i;
        if (real) {
            re[at] = ((const kiss_fft_scalar *)h)[i];
        } else {
            re[at] = ((const kiss_fft_cpx *)h)[i].r;
            im[at] = ((const kiss_fft_cpx *)h)[i].i;
        }
    }
    fft_f64(re, im, nfft);
    for (i = 0; i < nfft; ++i) {
        st->H[i].r = (kiss_fft_scalar)(re[i] / (double)nfft);
        st->H[i].i = (kiss_fft_scalar)(im[i] / (double)nfft);
    }
    free(re);
    free(im);
    return st;
#endif
}

/* every complete block of the n samples at in; returns the good samples written */
static size_t ff_blocks(kiss_fastfir_cfg st, const void *in, void *out, size_t n)
{
    if (n < st->nfft)
        return 0;
    const long long nb = (long long)((n - st->nfft) / st->ngood + 1);
    if (kfa_fastfir(KF_TYPE, (int)st->nfft, st->real, st->H, in, out, nb, (long long)st->ngood, NULL, 1) != 0) {
        KF_ERROR("kiss_fastfir: %s", kfa_last_error());
        return 0;
    }
    return (size_t)nb * st->ngood;
}

static size_t ff_run(kiss_fastfir_cfg st, void *inbuf, void *outbuf, size_t n_new, size_t *offset)
{
    const size_t samp = st->real ? sizeof(kiss_fft_scalar) : sizeof(kiss_fft_cpx);
    const size_t ntot = n_new + *offset;
    if (n_new == 0) {
        /* flush, kiss_fastfir.c:200-217: complete blocks, then the zero-padded rest */
        size_t done = ff_blocks(st, inbuf, outbuf, ntot), n = ntot - done, keep;
        char *tmp = (char *)st->tmpbuf;
        memset(tmp, 0, samp * st->nfft);
        memcpy(tmp, (const char *)inbuf + samp * done, samp * n);
        if (ff_blocks(st, tmp, tmp, st->nfft) == 0)
            return done;
        /* good samples of the padded block that come from real input */
        keep = n + st->ngood > st->nfft ? n + st->ngood - st->nfft : 0;
        memcpy((char *)outbuf + samp * done, tmp, samp * keep);
        return done + keep;
    }
    const size_t nwritten = ff_blocks(st, inbuf, outbuf, ntot);
    *offset = ntot - nwritten;
    memmove(inbuf, (const char *)inbuf + samp * nwritten, samp * *offset);
    return nwritten;
}

kiss_fastfir_cfg kiss_fastfir_alloc_cpx(const kiss_fft_cpx *imp_resp, size_t n_imp_resp, size_t *nfft, void *mem,
                                        size_t *lenmem)
{
    return ff_alloc(imp_resp, n_imp_resp, 0, nfft, mem, lenmem);
}

kiss_fastfir_cfg kiss_fastfir_alloc_real(const kiss_fft_scalar *imp_resp, size_t n_imp_resp, size_t *nfft, void *mem,
                                         size_t *lenmem)
{
    return ff_alloc(imp_resp, n_imp_resp, 1, nfft, mem, lenmem);
}

size_t kiss_fastfir_cpx(kiss_fastfir_cfg cfg, kiss_fft_cpx *inbuf, kiss_fft_cpx *outbuf, size_t n_new, size_t *offset)
{
    if (!cfg || cfg->real || !offset)
        return 0;
    return ff_run(cfg, inbuf, outbuf, n_new, offset);
}

size_t kiss_fastfir_real(kiss_fastfir_cfg cfg, kiss_fft_scalar *inbuf, kiss_fft_scalar *outbuf, size_t n_new,
                         size_t *offset)
{
    if (!cfg || !cfg->real || !offset)
        return 0;
    return ff_run(cfg, inbuf, outbuf, n_new, offset);
}
