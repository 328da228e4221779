/*
 * kiss_fftndr.h -- real multi-dimensional FFT (reference kiss_fftndr.h:23-49).
 * The real axis is the LAST one, dims[ndims-1], which must be even: input
 * dims[0] x ... x dims[ndims-1] scalars, output dims[0] x ... x (dims[ndims-1]/2+1)
 * complex.  Unlike the reference, whose size check rejects most shapes beyond a few
 * dozen points (kiss_fftndr.c:50-59), every even real axis is accepted.
 */
#ifndef KISS_NDR_H
#define KISS_NDR_H

#include "kiss_fft.h"
#include "kiss_fftnd.h"
#include "kiss_fftr.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kiss_fftndr_state *kiss_fftndr_cfg;

kiss_fftndr_cfg KISS_FFT_API kiss_fftndr_alloc(const int *dims, int ndims, int inverse_fft, void *mem,
                                               size_t *lenmem);

void KISS_FFT_API kiss_fftndr(kiss_fftndr_cfg cfg, const kiss_fft_scalar *timedata, kiss_fft_cpx *freqdata);

void KISS_FFT_API kiss_fftndri(kiss_fftndr_cfg cfg, const kiss_fft_cpx *freqdata, kiss_fft_scalar *timedata);

#define kiss_fftndr_free free

#ifdef __cplusplus
}
#endif
#endif
