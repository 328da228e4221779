/*
 * kf_multi.c -- C entry points of the multi-batch (include/kiss_fft_multi.h).
 *
 * Each item's plan is built with kiss_fft_alloc / kiss_fftr_alloc (the reference schedule
 * and twiddles) and handed to the engine, which keeps device copies; the host cfgs are
 * freed once the multi-batch exists.
 */
#include "kf_internal.h"
#include "kiss_fft_multi.h"

struct kiss_fft_multi_s {
    void *engine;
};

kiss_fft_multi kiss_fft_multi_alloc(const kiss_fft_multi_item *items, int nitems)
{
    kiss_fft_multi m = NULL;
    kfa_multi_item *ei = NULL;
    kfa_cplan *cp = NULL;
    kfa_rplan *rp = NULL;
    void **cfgs = NULL;
    int i, ok = 1;
    if (!items || nitems <= 0)
        return NULL;
    ei = (kfa_multi_item *)calloc((size_t)nitems, sizeof(*ei));
    cp = (kfa_cplan *)calloc((size_t)nitems, sizeof(*cp));
    rp = (kfa_rplan *)calloc((size_t)nitems, sizeof(*rp));
    cfgs = (void **)calloc((size_t)nitems, sizeof(*cfgs));
    if (!ei || !cp || !rp || !cfgs)
        ok = 0;
    for (i = 0; ok && i < nitems; ++i) {
        const kiss_fft_multi_item *it = &items[i];
        const int real = it->kind == KISS_FFT_MULTI_R2C || it->kind == KISS_FFT_MULTI_C2R;
        if (it->kind < 0 || it->kind > 3 || it->nfft <= 0) {
            KF_ERROR("kiss_fft_multi_alloc: item %d: bad kind or length", i);
            ok = 0;
            break;
        }
        if (real) {
            cfgs[i] = kiss_fftr_alloc(it->nfft, it->kind == KISS_FFT_MULTI_C2R, NULL, NULL);
            if (cfgs[i])
                kf_rplan_of((kiss_fftr_cfg)cfgs[i], &rp[i]);
        } else {
            cfgs[i] = kiss_fft_alloc(it->nfft, it->kind == KISS_FFT_MULTI_C2C_INV, NULL, NULL);
            if (cfgs[i])
                kf_cplan_of((kiss_fft_cfg)cfgs[i], &cp[i]);
        }
        if (!cfgs[i]) {
            ok = 0;
            break;
        }
        ei[i].kind = it->kind;
        ei[i].nfft = it->nfft;
        ei[i].cp = real ? NULL : &cp[i];
        ei[i].rp = real ? &rp[i] : NULL;
        ei[i].in = it->in;
        ei[i].out = it->out;
        ei[i].howmany = it->howmany;
    }
    if (ok) {
        void *e = kfa_multi_create(KF_TYPE, ei, nitems);
        if (e) {
            m = (kiss_fft_multi)calloc(1, sizeof(*m));
            if (m)
                m->engine = e;
            else
                kfa_multi_destroy(e);
        } else {
            KF_ERROR("kiss_fft_multi_alloc: %s", kfa_last_error());
        }
    }
    if (cfgs)
        for (i = 0; i < nitems; ++i)
            free(cfgs[i]);
    free(cfgs);
    free(ei);
    free(cp);
    free(rp);
    return m;
}

int kiss_fft_multi_exec(kiss_fft_multi m, void *stream)
{
    return m ? kfa_multi_exec(m->engine, stream, stream == NULL) : -2;
}

int kiss_fft_multi_capture(kiss_fft_multi m, void *stream)
{
    return m ? kfa_multi_capture(m->engine, stream) : -2;
}

int kiss_fft_multi_replay(kiss_fft_multi m, void *stream)
{
    return m ? kfa_multi_replay(m->engine, stream, stream == NULL) : -2;
}

int kiss_fft_multi_info(kiss_fft_multi m, long long *tiles, int *grid, int *persistent_items)
{
    return m ? kfa_multi_info(m->engine, tiles, grid, persistent_items) : -2;
}

int kiss_fft_multi_layout(kiss_fft_multi m, int *groups, int *group_items, int *bluestein_items,
                          int *multipass_items)
{
    return m ? kfa_multi_layout(m->engine, groups, group_items, bluestein_items, multipass_items) : -2;
}

int kiss_fft_multi_fused(kiss_fft_multi m, int *members, long long *workgroups, int *threads)
{
    return m ? kfa_multi_fused(m->engine, members, workgroups, threads) : -2;
}

void kiss_fft_multi_free(kiss_fft_multi m)
{
    if (!m)
        return;
    kfa_multi_destroy(m->engine);
    free(m);
}
