/*
 * kf_nd.c -- complex N-D transforms (reference kiss_fftnd.c:20-188).
 *
 * The cfg keeps the reference layout: per-axis 1-D cfgs, a copy of dims and the
 * reference-sized tmpbuf (unused here).  The GPU runs one LDS-tiled pass per axis,
 * dims[0] first like kiss_fftnd.c:181-186; fixed-point results round per axis, so the
 * order is part of the result.
 */
#include "kf_internal.h"

kiss_fftnd_cfg kiss_fftnd_alloc(const int *dims, int ndims, int inverse_fft, void *mem, size_t *lenmem)
{
    kiss_fftnd_cfg st = NULL;
    size_t need = sizeof(struct kiss_fftnd_state);
    long long dimprod = 1;
    char *ptr;
    if (ndims < 0 || (ndims > 0 && dims == NULL))
        return NULL;
    for (int k = 0; k < ndims; ++k) {
        size_t sub = 0;
        if (dims[k] <= 0)
            return NULL;
        kiss_fft_alloc(dims[k], inverse_fft, NULL, &sub);
        need += sub;
        dimprod *= dims[k];
        if (dimprod > 0x7fffffffLL)
            return NULL;
    }
    need += sizeof(int) * (size_t)ndims;
    need += sizeof(void *) * (size_t)ndims;
    need += sizeof(kiss_fft_cpx) * (size_t)dimprod;
    if (lenmem == NULL) {
        ptr = (char *)malloc(need);
    } else {
        ptr = (mem != NULL && *lenmem >= need) ? (char *)mem : NULL;
        *lenmem = need;
    }
    if (!ptr)
        return NULL;
    st = (kiss_fftnd_cfg)ptr;
    st->dimprod = (int)dimprod;
    st->ndims = ndims;
    ptr += sizeof(struct kiss_fftnd_state);
    st->states = (kiss_fft_cfg *)ptr;
    ptr += sizeof(void *) * (size_t)ndims;
    st->dims = (int *)ptr;
    ptr += sizeof(int) * (size_t)ndims;
    st->tmpbuf = (kiss_fft_cpx *)ptr;
    ptr += sizeof(kiss_fft_cpx) * (size_t)dimprod;
    for (int k = 0; k < ndims; ++k) {
        size_t len = 0;
        st->dims[k] = dims[k];
        kiss_fft_alloc(dims[k], inverse_fft, NULL, &len);
        st->states[k] = kiss_fft_alloc(dims[k], inverse_fft, ptr, &len);
        ptr += len;
    }
    return st;
}

static int nd_run(kiss_fftnd_cfg st, const kiss_fft_cpx *fin, kiss_fft_cpx *fout, long long howmany, void *stream,
                  int sync)
{
    kfa_cplan stack_axes[8];
    kfa_cplan *axes = stack_axes;
    int rc;
    memset(stack_axes, 0, sizeof stack_axes);
    if (!st || !fin || !fout)
        return -2;
    if (st->ndims > 8) {
        axes = (kfa_cplan *)malloc(sizeof(kfa_cplan) * (size_t)st->ndims);
        if (!axes)
            return -2;
    }
    for (int k = 0; k < st->ndims; ++k)
        kf_cplan_of(st->states[k], &axes[k]);
    rc = kfa_nd(st->ndims, st->dims, axes, fin, fout, howmany, stream, sync);
    if (axes != stack_axes)
        free(axes);
    return rc;
}

void kiss_fftnd(kiss_fftnd_cfg st, const kiss_fft_cpx *fin, kiss_fft_cpx *fout)
{
    if (nd_run(st, fin, fout, 1, NULL, 1))
        KF_ERROR("kiss_fftnd: %s", kfa_last_error());
}

int kiss_fftnd_batch(kiss_fftnd_cfg cfg, const kiss_fft_cpx *fin, kiss_fft_cpx *fout, long long howmany)
{
    return nd_run(cfg, fin, fout, howmany, NULL, 1);
}

int kiss_fftnd_batch_async(kiss_fftnd_cfg cfg, const kiss_fft_cpx *fin, kiss_fft_cpx *fout, long long howmany,
                           void *stream)
{
    return nd_run(cfg, fin, fout, howmany, stream, 0);
}
