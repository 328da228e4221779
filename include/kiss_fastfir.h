/*
 * kiss_fastfir.h -- fast FIR filtering by overlap-scrap on the GPU.
 *
 * The interface of the reference's tools/kiss_fastfir.c:31-42,55-153,229-246 (which is a
 * tool source, not a library header): a filter is planned once from its impulse response,
 * then a stream is filtered in pieces.  The reference compiles either the complex or the
 * real filter under the same names (REAL_FASTFIR); this library provides both, and
 * defining REAL_FASTFIR before including this header maps the names to the real variant
 * so code written for the reference's real build compiles unchanged.
 *
 *   kiss_fastfir_alloc(h, nh, &nfft, mem, lenmem)
 *       nfft <= 0 (or NULL): the next power of two at least twice nh-1, at least 1024
 *       (complex) or 2048 (real) -- kiss_fastfir.c:77-89.  Supported block lengths are
 *       powers of two from 256 to 8192 (one block is transformed in LDS).  mem/lenmem:
 *       the reference's size-query protocol; the cfg is one free()-able block.
 *   kiss_fastfir(cfg, inbuf, outbuf, n_new, &offset)
 *       filters the offset + n_new samples at the front of inbuf: every complete block
 *       (nfft samples in, nfft-nh+1 good samples out) in one GPU launch; returns the
 *       number of output samples written and moves the unprocessed input samples to the
 *       front of inbuf, setting offset to their count.  n_new == 0 flushes: the remaining
 *       samples are zero-padded and filtered.  Buffers may be host or device memory.
 *
 * Differences: float builds only (the reference scales the spectrum by 1/nfft in float,
 * kiss_fastfir.c:145-152, which does not work for fixed point); outbuf receives only the
 * good samples (the reference also writes the discarded tail of the last block).
 */
#ifndef KISS_FASTFIR_H
#define KISS_FASTFIR_H

#include <stddef.h>

#include "kiss_fft.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kiss_fastfir_state *kiss_fastfir_cfg;

kiss_fastfir_cfg KISS_FFT_API kiss_fastfir_alloc_cpx(const kiss_fft_cpx *imp_resp, size_t n_imp_resp, size_t *nfft,
                                                     void *mem, size_t *lenmem);
size_t KISS_FFT_API kiss_fastfir_cpx(kiss_fastfir_cfg cfg, kiss_fft_cpx *inbuf, kiss_fft_cpx *outbuf, size_t n_new,
                                     size_t *offset);

kiss_fastfir_cfg KISS_FFT_API kiss_fastfir_alloc_real(const kiss_fft_scalar *imp_resp, size_t n_imp_resp,
                                                      size_t *nfft, void *mem, size_t *lenmem);
size_t KISS_FFT_API kiss_fastfir_real(kiss_fastfir_cfg cfg, kiss_fft_scalar *inbuf, kiss_fft_scalar *outbuf,
                                      size_t n_new, size_t *offset);

#ifdef REAL_FASTFIR
#define kiss_fastfir_alloc kiss_fastfir_alloc_real
#define kiss_fastfir kiss_fastfir_real
#else
#define kiss_fastfir_alloc kiss_fastfir_alloc_cpx
#define kiss_fastfir kiss_fastfir_cpx
#endif

#ifdef __cplusplus
}
#endif
#endif
