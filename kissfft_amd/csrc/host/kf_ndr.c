/*
 * kf_ndr.c -- real N-D transforms (reference kiss_fftndr.c:30-132).
 *
 * Same composition as the reference: the real transform along the last axis
 * (kiss_fftr), then the complex N-D transform over the leading axes for every bin
 * (forward); the inverse reverses the two.  Deliberate deviation: the reference's size
 * check (kiss_fftndr.c:50-59) rejects nearly every shape above a few dozen points
 * (64x64 and 1024x1000 included); every shape with an even last axis is accepted here,
 * and a 1-D shape is the real transform alone.
 */
#include "kf_internal.h"

kiss_fftndr_cfg kiss_fftndr_alloc(const int *dims, int ndims, int inverse_fft, void *mem, size_t *lenmem)
{
    kiss_fftndr_cfg st = NULL;
    size_t nr = 0, nd = 0, need, ntmp;
    char *ptr;
    if (ndims < 1 || dims == NULL)
        return NULL;
    const int dimReal = dims[ndims - 1];
    size_t dimOther = 1;
    for (int k = 0; k + 1 < ndims; ++k) {
        if (dims[k] <= 0)
            return NULL;
        dimOther *= (size_t)dims[k];
    }
    if (dimReal <= 0 || (dimReal & 1)) {
        KF_ERROR("Real FFT optimization must be even.");
        return NULL;
    }
    (void)kiss_fftr_alloc(dimReal, inverse_fft, NULL, &nr);
    (void)kiss_fftnd_alloc(dims, ndims - 1, inverse_fft, NULL, &nd);
    ntmp = (2 * dimOther > (size_t)dimReal + 2 ? 2 * dimOther : (size_t)dimReal + 2) * sizeof(kiss_fft_scalar) +
           dimOther * ((size_t)dimReal + 2) * sizeof(kiss_fft_scalar);
    need = sizeof(struct kiss_fftndr_state) + nr + nd + ntmp;
    if (lenmem == NULL) {
        ptr = (char *)malloc(need);
    } else {
        ptr = (mem != NULL && *lenmem >= need) ? (char *)mem : NULL;
        *lenmem = need;
    }
    if (!ptr)
        return NULL;
    st = (kiss_fftndr_cfg)ptr;
    memset(st, 0, need);
    ptr += sizeof(struct kiss_fftndr_state);
    st->dimReal = dimReal;
    st->dimOther = dimOther;
    st->cfg_r = kiss_fftr_alloc(dimReal, inverse_fft, ptr, &nr);
    ptr += nr;
    st->cfg_nd = kiss_fftnd_alloc(dims, ndims - 1, inverse_fft, ptr, &nd);
    ptr += nd;
    st->tmpbuf = ptr;
    return st;
}

static int ndr_run(kiss_fftndr_cfg st, const void *in, void *out, long long howmany, void *stream, int sync,
                   int inverse)
{
    kfa_rplan rp;
    kfa_cplan stack_axes[8];
    int stack_dims[9];
    kfa_cplan *axes = stack_axes;
    int *dims = stack_dims;
    int rc;
    if (!st || !in || !out)
        return -2;
    if ((st->cfg_r->substate->inverse != 0) != (inverse != 0))
        return -2;
    const int nlead = st->cfg_nd->ndims;
    if (nlead > 8) {
        axes = (kfa_cplan *)malloc(sizeof(kfa_cplan) * (size_t)nlead);
        dims = (int *)malloc(sizeof(int) * (size_t)(nlead + 1));
        if (!axes || !dims) {
            free(axes);
            free(dims);
            return -2;
        }
    }
    for (int k = 0; k < nlead; ++k) {
        kf_cplan_of(st->cfg_nd->states[k], &axes[k]);
        dims[k] = st->cfg_nd->dims[k];
    }
    dims[nlead] = st->dimReal;
    kf_rplan_of(st->cfg_r, &rp);
    rc = kfa_ndr(nlead + 1, dims, &rp, axes, inverse, in, out, howmany, stream, sync);
    if (axes != stack_axes) {
        free(axes);
        free(dims);
    }
    return rc;
}

void kiss_fftndr(kiss_fftndr_cfg st, const kiss_fft_scalar *timedata, kiss_fft_cpx *freqdata)
{
    if (st && st->cfg_r->substate->inverse) {
        KF_ERROR("kiss fft usage error: improper alloc");
        return;
    }
    if (ndr_run(st, timedata, freqdata, 1, NULL, 1, 0))
        KF_ERROR("kiss_fftndr: %s", kfa_last_error());
}

void kiss_fftndri(kiss_fftndr_cfg st, const kiss_fft_cpx *freqdata, kiss_fft_scalar *timedata)
{
    if (st && st->cfg_r->substate->inverse == 0) {
        KF_ERROR("kiss fft usage error: improper alloc");
        return;
    }
    if (ndr_run(st, freqdata, timedata, 1, NULL, 1, 1))
        KF_ERROR("kiss_fftndri: %s", kfa_last_error());
}

int kiss_fftndr_batch(kiss_fftndr_cfg cfg, const kiss_fft_scalar *timedata, kiss_fft_cpx *freqdata,
                      long long howmany)
{
    return ndr_run(cfg, timedata, freqdata, howmany, NULL, 1, 0);
}

int kiss_fftndri_batch(kiss_fftndr_cfg cfg, const kiss_fft_cpx *freqdata, kiss_fft_scalar *timedata,
                       long long howmany)
{
    return ndr_run(cfg, freqdata, timedata, howmany, NULL, 1, 1);
}

int kiss_fftndr_batch_async(kiss_fftndr_cfg cfg, const kiss_fft_scalar *timedata, kiss_fft_cpx *freqdata,
                            long long howmany, void *stream)
{
    return ndr_run(cfg, timedata, freqdata, howmany, stream, 0, 0);
}

int kiss_fftndri_batch_async(kiss_fftndr_cfg cfg, const kiss_fft_cpx *freqdata, kiss_fft_scalar *timedata,
                             long long howmany, void *stream)
{
    return ndr_run(cfg, freqdata, timedata, howmany, stream, 0, 1);
}
