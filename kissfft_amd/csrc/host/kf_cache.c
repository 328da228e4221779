/*
 * kf_cache.c -- kfc_fft / kfc_ifft / kfc_cleanup (reference kfc.c:21-83).
 *
 * A list of cfgs keyed by (nfft, direction), like the reference, but guarded by a mutex
 * so concurrent callers are safe (the reference's static list is not, kfc.c:21-22).
 * Each entry's cfg is built in the same block as its list node.
 */
#include <pthread.h>

#include "kf_internal.h"
#include "kfc.h"

struct kf_cached {
    int nfft;
    int inverse;
    kiss_fft_cfg cfg;
    struct kf_cached *next;
};

static struct kf_cached *kf_cache_head = NULL;
static pthread_mutex_t kf_cache_mu = PTHREAD_MUTEX_INITIALIZER;

static kiss_fft_cfg kf_cache_get(int nfft, int inverse)
{
    struct kf_cached *e;
    kiss_fft_cfg cfg = NULL;
    pthread_mutex_lock(&kf_cache_mu);
    for (e = kf_cache_head; e; e = e->next)
        if (e->nfft == nfft && e->inverse == inverse)
            break;
    if (!e) {
        size_t len = 0;
        kiss_fft_alloc(nfft, inverse, NULL, &len);
        if (len) {
            e = (struct kf_cached *)KISS_FFT_MALLOC(sizeof(struct kf_cached) + len);
            if (e) {
                e->cfg = kiss_fft_alloc(nfft, inverse, e + 1, &len);
                e->nfft = nfft;
                e->inverse = inverse;
                e->next = kf_cache_head;
                kf_cache_head = e;
            }
        }
    }
    if (e)
        cfg = e->cfg;
    pthread_mutex_unlock(&kf_cache_mu);
    return cfg;
}

void kfc_fft(int nfft, const kiss_fft_cpx *fin, kiss_fft_cpx *fout)
{
    kiss_fft(kf_cache_get(nfft, 0), fin, fout);
}

void kfc_ifft(int nfft, const kiss_fft_cpx *fin, kiss_fft_cpx *fout)
{
    kiss_fft(kf_cache_get(nfft, 1), fin, fout);
}

void kfc_cleanup(void)
{
    pthread_mutex_lock(&kf_cache_mu);
    while (kf_cache_head) {
        struct kf_cached *next = kf_cache_head->next;
        KISS_FFT_FREE(kf_cache_head);
        kf_cache_head = next;
    }
    pthread_mutex_unlock(&kf_cache_mu);
}

int kfc_cached_count(void)
{
    int n = 0;
    pthread_mutex_lock(&kf_cache_mu);
    for (struct kf_cached *e = kf_cache_head; e; e = e->next)
        ++n;
    pthread_mutex_unlock(&kf_cache_mu);
    return n;
}
