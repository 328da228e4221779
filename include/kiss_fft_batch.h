/*
 * kiss_fft_batch.h -- batched and asynchronous calls of the MI355X engine.
 *
 * No counterpart in the reference: these are the additive "batched plan call" of the
 * engine.  Each takes an ordinary cfg from the reference API and runs `howmany`
 * independent transforms in one GPU launch.  Its nearest reference analogues are the
 * column loop of kiss_fftnd (kiss_fftnd.c:176-177) and the 4-lane lockstep of USE_SIMD
 * builds; a batched call gives the same result as a loop of single calls (bit for bit in
 * the fixed-point builds).
 *
 * Distances are in elements of the side they describe: complex values for complex data,
 * scalars for real data (real distances must be even).
 *
 * Synchronous forms accept host or device pointers and return once results are in place.
 * _async forms take device (or managed) pointers and enqueue on `stream`, a hipStream_t
 * (NULL: the calling thread's own stream of this library); kiss_fft_stream_sync waits.
 *
 * Every call returns 0 on success or a negative code; kiss_fft_last_error() describes the
 * calling thread's last failure.
 */
#ifndef KISS_FFT_BATCH_H
#define KISS_FFT_BATCH_H

#include "kiss_fft.h"
#include "kiss_fftnd.h"
#include "kiss_fftndr.h"
#include "kiss_fftr.h"

#ifdef __cplusplus
extern "C" {
#endif

int KISS_FFT_API kiss_fft_batch(kiss_fft_cfg cfg, const kiss_fft_cpx *fin, int in_stride, long long idist,
                                kiss_fft_cpx *fout, long long odist, long long howmany);
int KISS_FFT_API kiss_fftr_batch(kiss_fftr_cfg cfg, const kiss_fft_scalar *timedata, long long idist,
                                 kiss_fft_cpx *freqdata, long long odist, long long howmany);
int KISS_FFT_API kiss_fftri_batch(kiss_fftr_cfg cfg, const kiss_fft_cpx *freqdata, long long idist,
                                  kiss_fft_scalar *timedata, long long odist, long long howmany);
int KISS_FFT_API kiss_fftnd_batch(kiss_fftnd_cfg cfg, const kiss_fft_cpx *fin, kiss_fft_cpx *fout,
                                  long long howmany);
int KISS_FFT_API kiss_fftndr_batch(kiss_fftndr_cfg cfg, const kiss_fft_scalar *timedata,
                                   kiss_fft_cpx *freqdata, long long howmany);
int KISS_FFT_API kiss_fftndri_batch(kiss_fftndr_cfg cfg, const kiss_fft_cpx *freqdata,
                                    kiss_fft_scalar *timedata, long long howmany);

int KISS_FFT_API kiss_fft_batch_async(kiss_fft_cfg cfg, const kiss_fft_cpx *fin, int in_stride,
                                      long long idist, kiss_fft_cpx *fout, long long odist, long long howmany,
                                      void *stream);
int KISS_FFT_API kiss_fftr_batch_async(kiss_fftr_cfg cfg, const kiss_fft_scalar *timedata, long long idist,
                                       kiss_fft_cpx *freqdata, long long odist, long long howmany, void *stream);
int KISS_FFT_API kiss_fftri_batch_async(kiss_fftr_cfg cfg, const kiss_fft_cpx *freqdata, long long idist,
                                        kiss_fft_scalar *timedata, long long odist, long long howmany,
                                        void *stream);
int KISS_FFT_API kiss_fftnd_batch_async(kiss_fftnd_cfg cfg, const kiss_fft_cpx *fin, kiss_fft_cpx *fout,
                                        long long howmany, void *stream);
int KISS_FFT_API kiss_fftndr_batch_async(kiss_fftndr_cfg cfg, const kiss_fft_scalar *timedata,
                                         kiss_fft_cpx *freqdata, long long howmany, void *stream);
int KISS_FFT_API kiss_fftndri_batch_async(kiss_fftndr_cfg cfg, const kiss_fft_cpx *freqdata,
                                          kiss_fft_scalar *timedata, long long howmany, void *stream);

/* kiss_fft_batch_noalloc: kiss_fft_batch_async that allocates nothing.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid allocation).
 *   - fin / fout must be device memory; NULL `stream` is the null stream.
 *   - The cfg's device plan and kernel choice must exist on this device: run the same cfg
 *     once through kiss_fft_batch / kiss_fft_batch_async first.  Otherwise the call fails.
 *   - Lengths longer than one LDS line (float / Q31 above 8192 points, Q15 above 16384)
 *     take their temporaries from a static 64 MiB device array: one transform's
 *     temporary must fit it (float 2^22 points at most); larger batches run in chunks.
 *     Longer lengths are rejected with an error.
 *   - Every call on one GPU shares that array: calls that overlap on one GPU race.
 * Returns 0 or a negative code (kiss_fft_last_error()). */
int KISS_FFT_API kiss_fft_batch_noalloc(kiss_fft_cfg cfg, const kiss_fft_cpx *fin, int in_stride,
                                        long long idist, kiss_fft_cpx *fout, long long odist,
                                        long long howmany, void *stream);

/* Waits for everything enqueued on `stream` (NULL: the calling thread's stream). */
int KISS_FFT_API kiss_fft_stream_sync(void *stream);

/* Message of the calling thread's last failure ("" if none). */
const char KISS_FFT_API *kiss_fft_last_error(void);

/* 1 if a HIP device is visible to this process, else 0. */
int KISS_FFT_API kiss_fft_device_available(void);

/* Engine counters: device plans built / reused, kernel launches, staged bytes. */
typedef struct {
    long long plans_built;
    long long plan_hits;
    long long launches;
    long long h2d_bytes;
    long long d2h_bytes;
} kiss_fft_engine_stats;
void KISS_FFT_API kiss_fft_get_engine_stats(kiss_fft_engine_stats *stats);

#ifdef __cplusplus
}
#endif
#endif
