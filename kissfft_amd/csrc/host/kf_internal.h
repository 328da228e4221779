/*
 * kf_internal.h -- host-side plan layouts and helpers of the C layer.
 *
 * The cfg blocks keep the reference layouts (_kiss_fft_guts.h:27-32, kiss_fftr.c:12-19,
 * kiss_fftnd.c:12-18, kiss_fftndr.c:13-20), so the sizes reported through the
 * mem/lenmem protocol are the reference's.  Device state never lives in a cfg: the
 * engine keeps it in its own cache, keyed by the plan parameters.
 */
#ifndef KF_INTERNAL_H
#define KF_INTERNAL_H

#include <stdint.h>

#include "../kfa_engine.h"
#include "kiss_fft.h"
#include "kiss_fft_batch.h"
#include "kiss_fftnd.h"
#include "kiss_fftndr.h"
#include "kiss_fftr.h"

#if defined(FIXED_POINT) && (FIXED_POINT == 32)
#define KF_TYPE KFA_Q31
#define KF_SAMP_MAX 2147483647.0
#elif defined(FIXED_POINT)
#define KF_TYPE KFA_Q15
#define KF_SAMP_MAX 32767.0
#else
#define KF_TYPE KFA_F32
#endif

#define KF_MAXFACTORS 32

struct kiss_fft_state {
    int nfft;
    int inverse;
    int factors[2 * KF_MAXFACTORS];
    kiss_fft_cpx twiddles[1];
};

struct kiss_fftr_state {
    kiss_fft_cfg substate;
    kiss_fft_cpx *tmpbuf;
    kiss_fft_cpx *super_twiddles;
};

struct kiss_fftnd_state {
    int dimprod;
    int ndims;
    int *dims;
    kiss_fft_cfg *states;
    kiss_fft_cpx *tmpbuf;
};

struct kiss_fftndr_state {
    int dimReal;
    size_t dimOther;
    kiss_fftr_cfg cfg_r;
    kiss_fftnd_cfg cfg_nd;
    void *tmpbuf;
};

/* [ERROR] file:line message on stderr, as KISS_FFT_ERROR (kiss_fft_log.h:20-32);
 * silent under NDEBUG. */
#ifdef NDEBUG
#define KF_ERROR(...) ((void)0)
#else
#define KF_ERROR(...)                                                   \
    do {                                                                \
        fprintf(stderr, "[ERROR] %s:%d ", __FILE__, __LINE__);          \
        fprintf(stderr, __VA_ARGS__);                                   \
        fprintf(stderr, "\n");                                          \
    } while (0)
#endif

/* Fill one twiddle entry exp(i*phase) the way the reference does (_kiss_fft_guts.h:127-145):
 * phase in FP64, rounded once to the scalar type. */
void kf_cexp(kiss_fft_cpx *dst, double phase);

/* Engine view of a cfg. */
void kf_cplan_of(kiss_fft_cfg st, kfa_cplan *pl);
void kf_rplan_of(kiss_fftr_cfg st, kfa_rplan *pl);

#endif
