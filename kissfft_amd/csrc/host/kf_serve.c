/*
 * kf_serve.c -- C entry points of the request server (include/kiss_fft_serve.h).
 *
 * Supplies the engine with reference plans: the cfgs of each (kind, length) are built
 * once with kiss_fft_alloc / kiss_fftr_alloc (the reference schedule and twiddles) and
 * kept for the server's lifetime.
 */
#include <pthread.h>
#include <string.h>

#include "kf_internal.h"
#include "kiss_fft_serve.h"

struct kf_plan_entry {
    int kind;
    int nfft;
    int ndims;      /* 0: a 1-D plan of nfft points; else an N-D plan of dims */
    int dims[32];
    void *cfg;
    struct kf_plan_entry *next;
};

struct kiss_fft_server_s {
    void *engine;
    pthread_mutex_t mu;
    struct kf_plan_entry *plans;
};

kiss_fft_server kiss_fft_server_create(long long hbm_budget_bytes, int streams, int max_group)
{
    kiss_fft_server s = (kiss_fft_server)calloc(1, sizeof(*s));
    if (!s)
        return NULL;
    s->engine = kfa_server_create(KF_TYPE, hbm_budget_bytes, streams, max_group);
    if (!s->engine) {
        KF_ERROR("kiss_fft_server_create: %s", kfa_last_error());
        free(s);
        return NULL;
    }
    pthread_mutex_init(&s->mu, NULL);
    return s;
}

static void *plan_for(kiss_fft_server s, int kind, int nfft)
{
    struct kf_plan_entry *e;
    void *cfg = NULL;
    pthread_mutex_lock(&s->mu);
    for (e = s->plans; e; e = e->next)
        if (e->kind == kind && e->nfft == nfft && e->ndims == 0)
            break;
    if (!e) {
        if (kind == KISS_FFT_SERVE_R2C || kind == KISS_FFT_SERVE_C2R)
            cfg = kiss_fftr_alloc(nfft, kind == KISS_FFT_SERVE_C2R, NULL, NULL);
        else
            cfg = kiss_fft_alloc(nfft, kind == KISS_FFT_SERVE_C2C_INV, NULL, NULL);
        if (cfg) {
            e = (struct kf_plan_entry *)malloc(sizeof(*e));
            if (e) {
                e->kind = kind;
                e->nfft = nfft;
                e->ndims = 0;
                e->cfg = cfg;
                e->next = s->plans;
                s->plans = e;
            } else {
                free(cfg);
                cfg = NULL;
            }
        }
    }
    if (e)
        cfg = e->cfg;
    pthread_mutex_unlock(&s->mu);
    return cfg;
}

long long kiss_fft_server_submit(kiss_fft_server s, int kind, int nfft, const void *in, void *out, long long howmany)
{
    if (!s || kind < 0 || kind > 3)
        return -2;
    void *cfg = plan_for(s, kind, nfft);
    if (!cfg)
        return -2;
    if (kind == KISS_FFT_SERVE_R2C || kind == KISS_FFT_SERVE_C2R) {
        kfa_rplan rp;
        kf_rplan_of((kiss_fftr_cfg)cfg, &rp);
        return kfa_server_submit(s->engine, kind, NULL, &rp, nfft, in, out, howmany);
    }
    kfa_cplan cp;
    kf_cplan_of((kiss_fft_cfg)cfg, &cp);
    return kfa_server_submit(s->engine, kind, &cp, NULL, nfft, in, out, howmany);
}

static void *plan_for_nd(kiss_fft_server s, int kind, int ndims, const int *dims)
{
    struct kf_plan_entry *e;
    void *cfg = NULL;
    pthread_mutex_lock(&s->mu);
    for (e = s->plans; e; e = e->next)
        if (e->kind == kind && e->ndims == ndims && memcmp(e->dims, dims, sizeof(int) * (size_t)ndims) == 0)
            break;
    if (!e) {
        if (kind == KISS_FFT_SERVE_R2C || kind == KISS_FFT_SERVE_C2R)
            cfg = kiss_fftndr_alloc(dims, ndims, kind == KISS_FFT_SERVE_C2R, NULL, NULL);
        else
            cfg = kiss_fftnd_alloc(dims, ndims, kind == KISS_FFT_SERVE_C2C_INV, NULL, NULL);
        if (cfg) {
            e = (struct kf_plan_entry *)calloc(1, sizeof(*e));
            if (e) {
                e->kind = kind;
                e->ndims = ndims;
                memcpy(e->dims, dims, sizeof(int) * (size_t)ndims);
                e->cfg = cfg;
                e->next = s->plans;
                s->plans = e;
            } else {
                free(cfg);
                cfg = NULL;
            }
        }
    }
    if (e)
        cfg = e->cfg;
    pthread_mutex_unlock(&s->mu);
    return cfg;
}

long long kiss_fft_server_submit_nd(kiss_fft_server s, int kind, int ndims, const int *dims, const void *in, void *out,
                                    long long howmany)
{
    kfa_cplan axes[32];
    kfa_rplan rp;
    int k;
    if (!s || kind < 0 || kind > 3 || ndims < 1 || ndims > 32 || !dims)
        return -2;
    void *cfg = plan_for_nd(s, kind, ndims, dims);
    if (!cfg)
        return -2;
    if (kind == KISS_FFT_SERVE_R2C || kind == KISS_FFT_SERVE_C2R) {
        kiss_fftndr_cfg st = (kiss_fftndr_cfg)cfg;
        const int nlead = st->cfg_nd ? st->cfg_nd->ndims : 0;
        for (k = 0; k < nlead; ++k)
            kf_cplan_of(st->cfg_nd->states[k], &axes[k]);
        kf_rplan_of(st->cfg_r, &rp);
        return kfa_server_submit_nd(s->engine, kind, ndims, dims, axes, &rp, in, out, howmany);
    }
    kiss_fftnd_cfg st = (kiss_fftnd_cfg)cfg;
    for (k = 0; k < ndims; ++k)
        kf_cplan_of(st->states[k], &axes[k]);
    return kfa_server_submit_nd(s->engine, kind, ndims, dims, axes, NULL, in, out, howmany);
}

int kiss_fft_server_wait(kiss_fft_server s, long long ticket, double *latency_s)
{
    return s ? kfa_server_wait(s->engine, ticket, latency_s) : -2;
}

int kiss_fft_server_drain(kiss_fft_server s)
{
    return s ? kfa_server_drain(s->engine) : -2;
}

int kiss_fft_server_get_stats(kiss_fft_server s, kiss_fft_server_stats *st)
{
    kfa_server_stats e;
    if (!s || !st)
        return -2;
    kfa_server_get_stats(s->engine, &e);
    st->submitted = e.submitted;
    st->completed = e.completed;
    st->failed = e.failed;
    st->launches = e.launches;
    st->grouped_requests = e.grouped_requests;
    st->bytes = e.bytes;
    st->budget_bytes = e.budget_bytes;
    st->in_use_bytes = e.in_use_bytes;
    st->p50_s = e.p50_s;
    st->p99_s = e.p99_s;
    st->max_s = e.max_s;
    st->mean_s = e.mean_s;
    st->span_s = e.span_s;
    return 0;
}

void kiss_fft_server_reset_stats(kiss_fft_server s)
{
    if (s)
        kfa_server_reset_stats(s->engine);
}

void kiss_fft_server_destroy(kiss_fft_server s)
{
    if (!s)
        return;
    kfa_server_destroy(s->engine);
    while (s->plans) {
        struct kf_plan_entry *n = s->plans->next;
        free(s->plans->cfg);
        free(s->plans);
        s->plans = n;
    }
    pthread_mutex_destroy(&s->mu);
    free(s);
}
