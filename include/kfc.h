/*
 * kfc.h -- cached-plan convenience calls (reference kfc.h:43-48).
 * Plans are cached per (nfft, direction); unlike the reference's global list, the cache
 * is guarded by a lock, so concurrent callers are safe.
 */
#ifndef KFC_H
#define KFC_H

#include "kiss_fft.h"

#ifdef __cplusplus
extern "C" {
#endif

/* forward complex FFT */
void KISS_FFT_API kfc_fft(int nfft, const kiss_fft_cpx *fin, kiss_fft_cpx *fout);
/* inverse complex FFT */
void KISS_FFT_API kfc_ifft(int nfft, const kiss_fft_cpx *fin, kiss_fft_cpx *fout);
/* free all cached plans */
void KISS_FFT_API kfc_cleanup(void);

/* Number of plans currently cached.  No counterpart in the reference (whose count is a
 * static used only by its KFC_TEST self-test, kfc.c:85-109); exported for that test. */
int KISS_FFT_API kfc_cached_count(void);

#ifdef __cplusplus
}
#endif
#endif
