/*
 * kiss_fft_multi.h -- several batches of different lengths and kinds in one launch.
 *
 * No counterpart in the reference, whose calls each transform one length; the nearest
 * reference analogues are kiss_fftnd's loop over many columns (kiss_fftnd.c:176-177) and
 * the SIMD build's transforms in lockstep (README.simd:12-20).
 *
 * A multi-batch fixes a list of items -- kind, length, device buffers, count -- and runs
 * all of them with one call (or one replay of a captured hipGraph).  At creation each item
 * is assigned, once, to a launch:
 *   - items whose (kind, length) has a compiled row kernel are grouped by (kind, length)
 *     into one launch of that kernel (up to 64 items per launch, one segment each);
 *   - float lengths with a prime factor above 16 (e.g. 1009) run Bluestein's kernel,
 *     grouped by length the same way;
 *   - every other item of at most one LDS line (generic prime radices, fixed-point lengths
 *     without a compiled kernel) joins a single persistent launch that walks a table of
 *     all of them, any mix of lengths and kinds together;
 *   - items longer than one LDS line (float 8192 complex / 16384 real, Q15 16384 / 32768)
 *     run as their own multi-pass launches, their temporaries in a workspace the
 *     multi-batch allocates at creation;
 *   - the compiled-kernel and Bluestein launches of the most common workgroup size are
 *     fused into one launch of a kernel compiled at run time from their own bodies (one
 *     grid, the groups in workgroup ranges, largest first); KISSFFT_AMD_MULTI_FUSED=0 at
 *     creation keeps separate launches.
 * The launches run back to back on the caller's stream and capture records them into one
 * graph (KISSFFT_AMD_MULTI_HELPERS=n at creation forks them over n streams of the
 * multi-batch, joined on the caller's stream; the hand-overs between queues cost more than
 * the launches overlap).
 * KISSFFT_AMD_MULTI_PERSISTENT=1 at creation puts every single-pass item into the
 * persistent launch.  Results equal the per-item batched calls of the fixed-point builds
 * bit for bit, and the float builds' within the float tolerance.  Items must not share
 * output memory.  kiss_fft_cleanup() invalidates every multi-batch.
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

/* Runs every item: the multi-batch's launches (the fused launch of its compiled-kernel
 * groups, the persistent launch of the other single-pass items, multi-pass launches) back
 * to back on `stream` (a hipStream_t; NULL: the calling thread's stream, synchronised
 * before returning). 0 or a negative code. */
int KISS_FFT_API kiss_fft_multi_exec(kiss_fft_multi m, void *stream);

/* Captures the launches on `stream` (not NULL) into a graph and instantiates it. */
int KISS_FFT_API kiss_fft_multi_capture(kiss_fft_multi m, void *stream);

/* Launches the captured graph on `stream` (NULL: the calling thread's stream, synchronised). */
int KISS_FFT_API kiss_fft_multi_replay(kiss_fft_multi m, void *stream);

/* Tiles of the persistent launch, its grid, and the items it covers. */
int KISS_FFT_API kiss_fft_multi_info(kiss_fft_multi m, long long *tiles, int *grid, int *persistent_items);

/* Compiled-kernel launches and the items they cover, Bluestein items, multi-pass items. */
int KISS_FFT_API kiss_fft_multi_layout(kiss_fft_multi m, int *groups, int *group_items, int *bluestein_items,
                                       int *multipass_items);

/* The fused launch: the compiled-kernel / Bluestein groups it runs, its workgroups and
 * threads per workgroup (all 0: no fused launch). */
int KISS_FFT_API kiss_fft_multi_fused(kiss_fft_multi m, int *members, long long *workgroups, int *threads);

void KISS_FFT_API kiss_fft_multi_free(kiss_fft_multi m);

#ifdef __cplusplus
}
#endif
#endif
