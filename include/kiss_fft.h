/*
 * kiss_fft.h -- KISS FFT complex 1-D API, served by the MI355X engine (libkissfft-<type>.so).
 *
 * Drop-in for the reference header (kiss_fft.h:117-150 of mborgerding/kissfft 131.1.0):
 * same type names, same prototypes, same scalar selection by FIXED_POINT, same
 * allocation protocol.  Transforms run on the GPU; host pointers are staged through HBM
 * and the call returns when the result is in place, device pointers are used in place.
 */
#ifndef KISS_FFT_H
#define KISS_FFT_H

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* Symbol visibility, as kiss_fft.h:17-30 of the reference. */
#ifdef KISS_FFT_SHARED
#define KISS_FFT_API __attribute__((visibility("default")))
#else
#define KISS_FFT_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

#ifndef KISS_FFT_MALLOC
#define KISS_FFT_MALLOC malloc
#endif
#ifndef KISS_FFT_FREE
#define KISS_FFT_FREE free
#endif

/* Scalar selection (reference kiss_fft.h:73-85): FIXED_POINT=16 -> Q15 int16,
 * FIXED_POINT=32 -> Q31 int32, otherwise float.  The engine serves these three. */
#ifdef FIXED_POINT
#include <stdint.h>
#if (FIXED_POINT == 32)
#define kiss_fft_scalar int32_t
#else
#define kiss_fft_scalar int16_t
#endif
#else
#ifndef kiss_fft_scalar
#define kiss_fft_scalar float
#endif
#endif

/* Interleaved complex, reference kiss_fft.h:87-90. */
typedef struct {
    kiss_fft_scalar r;
    kiss_fft_scalar i;
} kiss_fft_cpx;

typedef struct kiss_fft_state *kiss_fft_cfg;

/* Plan a length-nfft transform (inverse_fft != 0: positive exponent).
 *   lenmem == NULL                -> malloc'd cfg, release with free() / kiss_fft_free
 *   mem != NULL, *lenmem large    -> cfg placed in mem, *lenmem = bytes used
 *   otherwise                     -> NULL, *lenmem = bytes needed
 * Returns NULL for nfft <= 0.  (reference kiss_fft.h:94-117, kiss_fft.c:337-372) */
kiss_fft_cfg KISS_FFT_API kiss_fft_alloc(int nfft, int inverse_fft, void *mem, size_t *lenmem);

/* Unnormalised DFT of nfft complex points, fin -> fout (fin == fout allowed).
 * Fixed-point builds scale by about 1/nfft, bit-identical to the reference. */
void KISS_FFT_API kiss_fft(kiss_fft_cfg cfg, const kiss_fft_cpx *fin, kiss_fft_cpx *fout);

/* As kiss_fft, reading every fin_stride-th input sample; output is contiguous. */
void KISS_FFT_API kiss_fft_stride(kiss_fft_cfg cfg, const kiss_fft_cpx *fin, kiss_fft_cpx *fout,
                                  int fin_stride);

#define kiss_fft_free KISS_FFT_FREE

/* Releases device plans and twiddle tables cached for every cfg (optional at exit). */
void KISS_FFT_API kiss_fft_cleanup(void);

/* Smallest k >= n whose only prime factors are 2, 3 and 5. */
int KISS_FFT_API kiss_fft_next_fast_size(int n);

/* Real transforms need an even size. */
#define kiss_fftr_next_fast_size_real(n) (kiss_fft_next_fast_size(((n) + 1) >> 1) << 1)

#ifdef __cplusplus
}
#endif

#endif
