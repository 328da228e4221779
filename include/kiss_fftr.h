/*
 * kiss_fftr.h -- real-input FFT (reference kiss_fftr.h:29-49), served by the MI355X engine.
 * N real points <-> N/2+1 complex bins, computed through an N/2-point complex transform
 * whose split step runs fused in the same GPU pass.
 */
#ifndef KISS_FTR_H
#define KISS_FTR_H

#include "kiss_fft.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kiss_fftr_state *kiss_fftr_cfg;

/* nfft must be even (NULL otherwise).  Same mem/lenmem protocol as kiss_fft_alloc. */
kiss_fftr_cfg KISS_FFT_API kiss_fftr_alloc(int nfft, int inverse_fft, void *mem, size_t *lenmem);

/* timedata: nfft scalars -> freqdata: nfft/2+1 complex.  Needs a forward cfg. */
void KISS_FFT_API kiss_fftr(kiss_fftr_cfg cfg, const kiss_fft_scalar *timedata, kiss_fft_cpx *freqdata);

/* freqdata: nfft/2+1 complex -> timedata: nfft scalars (unnormalised).  Needs an
 * inverse cfg. */
void KISS_FFT_API kiss_fftri(kiss_fftr_cfg cfg, const kiss_fft_cpx *freqdata, kiss_fft_scalar *timedata);

#define kiss_fftr_free KISS_FFT_FREE

#ifdef __cplusplus
}
#endif
#endif
