/*
 * kiss_fft_multi.h -- several batches of different lengths and kinds in one launch.
 *
 * No counterpart in the reference, whose calls each transform one length; the nearest
 * reference analogues are kiss_fftnd's loop over many columns (kiss_fftnd.c:176-177) and
 * the SIMD build's transforms in lockstep (README.simd:12-20).
 *
 * A multi-batch fixes a list of items -- kind, length, device buffers, count -- and
 * executes all of them with a single persistent kernel launch: mixed lengths (any
 * factorisation, generic prime radices included) and mixed complex / real kinds together.
 * The launch can be captured into a hipGraph once and replayed.  Results equal the
 * per-item batched calls of the fixed-point builds bit for bit, and the float builds'
 * within the float tolerance (the float batched calls may use other radix schedules).
 *
 * Items whose length exceeds one LDS line (float 8192 complex / 16384 real, Q15 16384 /
 * 32768) run as their own multi-pass launches after the persistent one; a multi-batch that
 * holds such items cannot be captured.  kiss_fft_cleanup() invalidates every multi-batch.
 */
#ifndef KISS_FFT_MULTI_H
#define KISS_FFT_MULTI_H

#include "kiss_fft.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kiss_fft_multi_s *kiss_fft_multi;

enum {
    KISS_FFT_MULTI_C2C_FWD = 0, /* complex forward */
    KISS_FFT_MULTI_C2C_INV = 1, /* complex inverse (unnormalised) */
    KISS_FFT_MULTI_R2C = 2,     /* real forward, nfft even: nfft scalars -> nfft/2+1 complex */
    KISS_FFT_MULTI_C2R = 3      /* real inverse, nfft even: nfft/2+1 complex -> nfft scalars */
};

typedef struct {
    int kind;
    int nfft;
    const void *in;    /* device memory: howmany dense transforms */
    void *out;         /* device memory */
    long long howmany;
} kiss_fft_multi_item;

/* Builds the plans and the launch table.  NULL on failure (see kiss_fft_last_error). */
kiss_fft_multi KISS_FFT_API kiss_fft_multi_alloc(const kiss_fft_multi_item *items, int nitems);

/* Runs every item: one persistent launch on `stream` (a hipStream_t; NULL: the calling
 * thread's stream, synchronised before returning). 0 or a negative code. */
int KISS_FFT_API kiss_fft_multi_exec(kiss_fft_multi m, void *stream);

/* Captures the launch on `stream` (not NULL) into a graph and instantiates it. */
int KISS_FFT_API kiss_fft_multi_capture(kiss_fft_multi m, void *stream);

/* Launches the captured graph on `stream` (NULL: the calling thread's stream, synchronised). */
int KISS_FFT_API kiss_fft_multi_replay(kiss_fft_multi m, void *stream);

/* Tiles of the persistent launch, its grid, and the items it covers. */
int KISS_FFT_API kiss_fft_multi_info(kiss_fft_multi m, long long *tiles, int *grid, int *persistent_items);

void KISS_FFT_API kiss_fft_multi_free(kiss_fft_multi m);

#ifdef __cplusplus
}
#endif
#endif
