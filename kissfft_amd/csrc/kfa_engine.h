/*
 * kfa_engine.h -- the thin C ABI between the C host layer (the files under kissfft_amd/csrc/host)
 * and the HIP engine (kissfft_amd/csrc/engine/kfa_engine.hip).
 *
 * Internal to libkissfft-<type>.so: the symbols are hidden.  The public entry points are
 * the reference API (include/kiss_fft*.h, include/kfc.h) and the batched API
 * (include/kiss_fft_batch.h).
 *
 * Plain C types only.  Data pointers may be host memory (staged through HBM, the call is
 * synchronous) or device/managed memory (used in place).  `stream` is a hipStream_t or
 * NULL for the calling thread's own stream; `sync` != 0 waits for completion.
 * Every entry returns 0 on success and a negative code on failure; kfa_last_error()
 * gives the message for the calling thread.
 */
#ifndef KFA_ENGINE_H
#define KFA_ENGINE_H

#ifdef __cplusplus
extern "C" {
#endif

#define KFA_HIDDEN __attribute__((visibility("hidden")))

enum { KFA_F32 = 0, KFA_Q15 = 1, KFA_Q31 = 2 };

/* One 1-D complex plan as the host cfg holds it (struct kiss_fft_state,
 * _kiss_fft_guts.h:27-32): length, direction, (p, m) radix pairs ending with m == 1,
 * and the host twiddle table of nfft complex scalars. */
typedef struct {
    int type;
    int nfft;
    int inverse;
    const int *factors;
    const void *twiddles;
} kfa_cplan;

/* One real plan (struct kiss_fftr_state, kiss_fftr.c:12-19): the half-length complex
 * plan plus nfft/4 split twiddles. */
typedef struct {
    kfa_cplan half;
    const void *super_twiddles;
} kfa_rplan;

KFA_HIDDEN int kfa_c2c(const kfa_cplan *pl, const void *in, long long istride, long long idist,
                       void *out, long long odist, long long howmany, void *stream, int sync);

/* kfa_c2c without any allocation: device buffers only, the plan prepared by an earlier
 * kfa_c2c on this device, multi-pass temporaries in a static device array. */
KFA_HIDDEN int kfa_c2c_noalloc(const kfa_cplan *pl, const void *in, long long istride, long long idist,
                               void *out, long long odist, long long howmany, void *stream);

KFA_HIDDEN int kfa_r2c(const kfa_rplan *pl, const void *in, long long idist, void *out,
                       long long odist, long long howmany, void *stream, int sync);

KFA_HIDDEN int kfa_c2r(const kfa_rplan *pl, const void *in, long long idist, void *out,
                       long long odist, long long howmany, void *stream, int sync);

/* Complex N-D over `ndims` axes (dims[0] slowest), items dense and back to back. */
KFA_HIDDEN int kfa_nd(int ndims, const int *dims, const kfa_cplan *axes, const void *in,
                      void *out, long long howmany, void *stream, int sync);

/* Real N-D: real transform on the last axis, complex N-D over the others. */
KFA_HIDDEN int kfa_ndr(int ndims, const int *dims, const kfa_rplan *rpl, const kfa_cplan *axes,
                       int inverse, const void *in, void *out, long long howmany, void *stream,
                       int sync);

/* Waits for `stream` (NULL: the calling thread's stream). */
KFA_HIDDEN int kfa_stream_sync(void *stream);

KFA_HIDDEN void kfa_cleanup(void);
KFA_HIDDEN const char *kfa_last_error(void);
KFA_HIDDEN int kfa_device_available(void);

/* ---- request server (kfa_serve.hip) ---------------------------------------------------- */
enum { KFA_SERVE_C2C_FWD = 0, KFA_SERVE_C2C_INV = 1, KFA_SERVE_R2C = 2, KFA_SERVE_C2R = 3 };

typedef struct {
    long long submitted, completed, failed;
    long long launches, grouped_requests;
    long long bytes;          /* algorithmic bytes of completed requests */
    long long budget_bytes, in_use_bytes;
    double p50_s, p99_s, max_s, mean_s;
    double span_s;            /* first submit to last completion */
} kfa_server_stats;

/* budget_bytes <= 0: 80% of the HBM free at creation.  nstreams <= 0: 4.  max_group <= 0: 64. */
KFA_HIDDEN void *kfa_server_create(int type, long long budget_bytes, int nstreams, int max_group);
/* cp for complex kinds, rp for real kinds; nfft is the transform length (real length for
 * real kinds).  Returns a ticket > 0 or a negative code. */
KFA_HIDDEN long long kfa_server_submit(void *srv, int kind, const kfa_cplan *cp, const kfa_rplan *rp, int nfft,
                                       const void *in, void *out, long long howmany);
/* N-D request: complex kinds take one plan per axis in `axes`; real kinds take the real
 * plan `rp` of the last axis and the leading axes' plans in `axes`. */
KFA_HIDDEN long long kfa_server_submit_nd(void *srv, int kind, int ndims, const int *dims, const kfa_cplan *axes,
                                          const kfa_rplan *rp, const void *in, void *out, long long howmany);
KFA_HIDDEN int kfa_server_wait(void *srv, long long ticket, double *latency_s);
KFA_HIDDEN int kfa_server_drain(void *srv);
KFA_HIDDEN int kfa_server_get_stats(void *srv, kfa_server_stats *st);
KFA_HIDDEN void kfa_server_reset_stats(void *srv);
KFA_HIDDEN void kfa_server_destroy(void *srv);

/* Counters for tests and the serving layer. */
typedef struct {
    long long plans_built;
    long long plan_hits;
    long long launches;
    long long h2d_bytes;
    long long d2h_bytes;
} kfa_stats;
KFA_HIDDEN void kfa_get_stats(kfa_stats *s);

/* Multi-batch: several batches of different lengths / kinds in one launch sequence. */
typedef struct {
    int kind;              /* KFA_SERVE_* */
    int nfft;
    const kfa_cplan *cp;   /* complex kinds */
    const kfa_rplan *rp;   /* real kinds */
    const void *in;        /* device pointers, howmany dense transforms */
    void *out;
    long long howmany;
} kfa_multi_item;

KFA_HIDDEN void *kfa_multi_create(int type, const kfa_multi_item *items, int nitems);
KFA_HIDDEN int kfa_multi_exec(void *m, void *stream, int sync);
KFA_HIDDEN int kfa_multi_capture(void *m, void *stream);
KFA_HIDDEN int kfa_multi_replay(void *m, void *stream, int sync);
KFA_HIDDEN int kfa_multi_info(void *m, long long *tiles, int *grid, int *persistent_items);
KFA_HIDDEN int kfa_multi_layout(void *m, int *groups, int *group_items, int *bluestein_items,
                                int *multipass_items);
KFA_HIDDEN int kfa_multi_fused(void *m, int *members, long long *workgroups, int *threads);
KFA_HIDDEN void kfa_multi_destroy(void *m);

/* Fast FIR by overlap-scrap: nblocks windows in[k*ngood .. +nfft) -> out[k*ngood .. +ngood).
 * H: nfft complex (host), the filter spectrum scaled by 1/nfft.  real: samples are scalars. */
KFA_HIDDEN int kfa_fastfir(int type, int nfft, int real, const void *H, const void *in, void *out, long long nblocks,
                           long long ngood, void *stream, int sync);

#ifdef __cplusplus
}
#endif
#endif
