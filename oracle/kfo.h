/*
 * kfo.h -- interface of the KISS FFT oracle (CPU restatement).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py load this library, and only as the checker.  The product library
 * (libkissfft-*.so) never links or calls it.
 *
 * One source, three builds (selected by a macro, as the reference selects its scalar by
 * FIXED_POINT, kiss_fft.h:73-85):
 *   -DKFO_FLOAT  float32 scalar           -> liboracle-float.so
 *   -DKFO_Q15    int16 Q15 fixed point    -> liboracle-int16_t.so  (reference FIXED_POINT=16)
 *   -DKFO_Q31    int32 Q31 fixed point    -> liboracle-int32_t.so  (reference FIXED_POINT=32)
 *
 * All buffers are interleaved {re, im} pairs of the scalar type, like kiss_fft_cpx
 * (kiss_fft.h:87-90).  Batched entry points loop the single transform over `howmany`
 * independent items, `idist`/`odist` elements apart (OpenMP over the items).
 */
#ifndef KFO_H
#define KFO_H

#include <stddef.h>
#include <stdint.h>

#if defined(KFO_Q15)
typedef int16_t kfo_scalar;
#elif defined(KFO_Q31)
typedef int32_t kfo_scalar;
#else
typedef float kfo_scalar;
#endif

typedef struct {
    kfo_scalar r;
    kfo_scalar i;
} kfo_cpx;

#ifdef __cplusplus
extern "C" {
#endif

/* Radix schedule, kiss_fft.c:306-328.  Writes (p, m) pairs; returns the pair count. */
int kfo_factor(int n, int *pairs);

/* Full-circle twiddle table exp(-+2*pi*i*k/n), kiss_fft.c:361-367 with _kiss_fft_guts.h:127-145. */
void kfo_twiddles(int n, int inverse, kfo_cpx *tw);

/* Real-FFT split twiddles for a real length n (n even): n/4 entries, kiss_fftr.c:53-59. */
void kfo_super_twiddles(int n, int inverse, kfo_cpx *stw);

/* kiss_fft_next_fast_size, kiss_fft.c:412-424. */
int kfo_next_fast_size(int n);

/* Complex 1-D transform, kiss_fft_stride semantics (kiss_fft.c:375-399):
 * item b reads in[b*idist + k*in_stride], writes out[b*odist + k]. */
void kfo_fft(int n, int inverse, const kfo_cpx *in, int in_stride, kfo_cpx *out,
             long howmany, long idist, long odist);

/* Real forward, kiss_fftr (kiss_fftr.c:63-117): n reals -> n/2+1 complex per item. */
void kfo_fftr(int n, const kfo_scalar *in, kfo_cpx *out, long howmany, long idist, long odist);

/* Real inverse, kiss_fftri (kiss_fftr.c:119-155): n/2+1 complex -> n reals per item. */
void kfo_fftri(int n, const kfo_cpx *in, kfo_scalar *out, long howmany, long idist, long odist);

/* Complex N-D, kiss_fftnd (kiss_fftnd.c:156-188); items are dense and back to back. */
void kfo_fftnd(const int *dims, int ndims, int inverse, const kfo_cpx *in, kfo_cpx *out,
               long howmany);

/* Real N-D, kiss_fftndr / kiss_fftndri (kiss_fftndr.c:86-132).  The real axis is the last. */
void kfo_fftndr(const int *dims, int ndims, const kfo_scalar *in, kfo_cpx *out, long howmany);
void kfo_fftndri(const int *dims, int ndims, const kfo_cpx *in, kfo_scalar *out, long howmany);

#ifdef __cplusplus
}
#endif
#endif
