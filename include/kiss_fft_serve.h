/*
 * kiss_fft_serve.h -- a request server over the MI355X engine (continuous batching).
 *
 * No counterpart in the reference.  The nearest reference piece is the plan cache of
 * kfc.c (kfc.c:21-83), which this generalises to a serving loop: requests of mixed kinds
 * and lengths are submitted from any thread; a scheduler admits them in arrival order as
 * the server's HBM budget allows, groups admitted requests of the same kind and length
 * into one kernel launch, overlaps launches on several streams, and keeps admitting while
 * earlier work runs.  Plans and twiddles are cached on the device.
 *
 * Buffers: device (or managed) pointers are used in place and cost the server no memory;
 * host pointers are staged through server-owned HBM that counts against the budget, as
 * does the scratch of lengths longer than one LDS line.  `in` holds `howmany` dense
 * transforms (complex: nfft values; real forward: nfft scalars; real inverse: nfft/2+1
 * complex), `out` likewise.  Results equal the batched calls' (bit for bit in the
 * fixed-point builds).
 */
#ifndef KISS_FFT_SERVE_H
#define KISS_FFT_SERVE_H

#include "kiss_fft.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kiss_fft_server_s *kiss_fft_server;

enum {
    KISS_FFT_SERVE_C2C_FWD = 0, /* complex forward */
    KISS_FFT_SERVE_C2C_INV = 1, /* complex inverse (unnormalised) */
    KISS_FFT_SERVE_R2C = 2,     /* real forward, nfft even */
    KISS_FFT_SERVE_C2R = 3      /* real inverse, nfft even (unnormalised) */
};

typedef struct {
    long long submitted, completed, failed;
    long long launches, grouped_requests; /* kernel launch groups and the requests in them */
    long long bytes;                      /* algorithmic bytes (input + output) completed */
    long long budget_bytes, in_use_bytes;
    double p50_s, p99_s, max_s, mean_s;   /* request latency: submit to completion */
    double span_s;                        /* first submit to last completion */
} kiss_fft_server_stats;

/* hbm_budget_bytes <= 0: 80% of the free HBM; streams <= 0: 4; max_group <= 0: 64. NULL on failure. */
kiss_fft_server KISS_FFT_API kiss_fft_server_create(long long hbm_budget_bytes, int streams, int max_group);

/* Queues one request; returns a ticket > 0, or a negative code (see kiss_fft_last_error). */
long long KISS_FFT_API kiss_fft_server_submit(kiss_fft_server srv, int kind, int nfft, const void *in, void *out,
                                              long long howmany);

/* Queues one N-D request: `howmany` arrays of dims[0] x ... x dims[ndims-1] (dims[0]
 * slowest, as kiss_fftnd / kiss_fftndr).  Kinds: complex forward / inverse, real forward
 * (last dimension even; its nfft/2+1 bins on the output side) / real inverse.  N-D
 * requests are admitted, staged and overlapped like 1-D ones and launched one by one. */
long long KISS_FFT_API kiss_fft_server_submit_nd(kiss_fft_server srv, int kind, int ndims, const int *dims,
                                                 const void *in, void *out, long long howmany);

/* Blocks until the request is done (0) or failed (< 0); releases the ticket. */
int KISS_FFT_API kiss_fft_server_wait(kiss_fft_server srv, long long ticket, double *latency_s);

/* Blocks until nothing is queued or running. */
int KISS_FFT_API kiss_fft_server_drain(kiss_fft_server srv);

int KISS_FFT_API kiss_fft_server_get_stats(kiss_fft_server srv, kiss_fft_server_stats *stats);
void KISS_FFT_API kiss_fft_server_reset_stats(kiss_fft_server srv);

/* Drains and frees the server. */
void KISS_FFT_API kiss_fft_server_destroy(kiss_fft_server srv);

#ifdef __cplusplus
}
#endif
#endif
