/*
 * cpu_baseline.c -- times the reference KISS FFT on the host cores (reported baseline).
 *
 * TEST/MEASUREMENT INFRASTRUCTURE ONLY.  Built by oracle/Makefile against the reference
 * sources into oracle/_ref/ (never linked into the product).  bench.py runs it as the
 * cpu_baseline leg.  Two builds of the reference at its test flags (-O3 -ffast-math):
 *   cpu_baseline-<type>         with -fopenmp (the reference's KISSFFT_OPENMP code path,
 *                               kiss_fft.c:250-271, is compiled in)
 *   cpu_baseline_serial-<type>  without OpenMP: one thread, the plain serial library
 * and two modes:
 *   batch  an OpenMP `parallel for` over the batch with one cfg per thread (kiss_fftr /
 *          kiss_fftnd / kiss_fftndr write cfg->tmpbuf, kiss_fftr.c:77, kiss_fftnd.c:166-170);
 *          nested regions are off, so each transform runs serially on its thread
 *   intra  one calling thread, transforms one after another; the library's own OpenMP
 *          region splits each transform's top-level butterflies over THREADS threads (the
 *          as-shipped OpenMP build)
 * The cfgs are allocated before the timed passes; one warm-up pass, then the median of
 * `reps` timed passes.
 *
 * usage: cpu_baseline KIND N HOWMANY THREADS REPS [batch|intra]
 *   KIND: c2c | c2r | r2c | nd (N is "d0xd1x...") | ndr | ndri
 * prints one JSON object: transforms/s, GB/s (algorithmic bytes: one read of the input,
 * one write of the output), the median seconds per pass and the thread count.
 */
#ifdef _OPENMP
#include <omp.h>
#else
static int omp_get_thread_num(void) { return 0; }
static int omp_get_max_threads(void) { return 1; }
static void omp_set_num_threads(int n) { (void)n; }
#endif
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "kiss_fft.h"
#include "kiss_fftnd.h"
#include "kiss_fftndr.h"
#include "kiss_fftr.h"

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static int cmp_d(const void *a, const void *b)
{
    double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

static int parse_dims(const char *s, int *dims)
{
    int nd = 0;
    const char *p = s;
    while (*p && nd < 8) {
        dims[nd++] = atoi(p);
        while (*p && *p != 'x')
            ++p;
        if (*p == 'x')
            ++p;
    }
    return nd;
}

static void fill(kiss_fft_scalar *v, size_t n)
{
    uint64_t s = 0x243F6A8885A308D3ull;
    for (size_t k = 0; k < n; ++k) {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
#ifdef FIXED_POINT
        v[k] = (kiss_fft_scalar)((int64_t)(s >> 49) - 16384);
#else
        v[k] = (kiss_fft_scalar)((double)(s >> 40) * (1.0 / 8388608.0) - 1.0);
#endif
    }
}

static void run_one(const char *kind, void *cfg, kiss_fft_scalar *src, kiss_fft_scalar *dst)
{
    if (!strcmp(kind, "c2c"))
        kiss_fft((kiss_fft_cfg)cfg, (kiss_fft_cpx *)src, (kiss_fft_cpx *)dst);
    else if (!strcmp(kind, "r2c"))
        kiss_fftr((kiss_fftr_cfg)cfg, src, (kiss_fft_cpx *)dst);
    else if (!strcmp(kind, "c2r"))
        kiss_fftri((kiss_fftr_cfg)cfg, (kiss_fft_cpx *)src, dst);
    else if (!strcmp(kind, "nd"))
        kiss_fftnd((kiss_fftnd_cfg)cfg, (kiss_fft_cpx *)src, (kiss_fft_cpx *)dst);
    else if (!strcmp(kind, "ndr"))
        kiss_fftndr((kiss_fftndr_cfg)cfg, src, (kiss_fft_cpx *)dst);
    else
        kiss_fftndri((kiss_fftndr_cfg)cfg, (kiss_fft_cpx *)src, dst);
}

int main(int argc, char **argv)
{
    if (argc < 6) {
        fprintf(stderr, "usage: %s KIND N HOWMANY THREADS REPS [batch|intra]\n", argv[0]);
        return 2;
    }
    const char *kind = argv[1];
    int dims[8];
    const int nd = parse_dims(argv[2], dims);
    const long howmany = atol(argv[3]);
    const int threads = atoi(argv[4]);
    const int reps = atoi(argv[5]) > 0 ? atoi(argv[5]) : 1;
    long total = 1;
    for (int k = 0; k < nd; ++k)
        total *= dims[k];
    const long last = dims[nd - 1];
    const long nbins_total = total / last * (last / 2 + 1);

    omp_set_num_threads(threads);

    long in_sc, out_sc; /* scalars per item */
    if (!strcmp(kind, "c2c") || !strcmp(kind, "nd")) {
        in_sc = 2 * total;
        out_sc = 2 * total;
    } else if (!strcmp(kind, "r2c") || !strcmp(kind, "ndr")) {
        in_sc = total;
        out_sc = 2 * nbins_total;
    } else if (!strcmp(kind, "c2r") || !strcmp(kind, "ndri")) {
        in_sc = 2 * nbins_total;
        out_sc = total;
    } else {
        fprintf(stderr, "unknown kind %s\n", kind);
        return 2;
    }
    kiss_fft_scalar *in = malloc(sizeof(kiss_fft_scalar) * in_sc * howmany);
    kiss_fft_scalar *out = malloc(sizeof(kiss_fft_scalar) * out_sc * howmany);
    if (!in || !out) {
        fprintf(stderr, "allocation failed\n");
        return 1;
    }
    fill(in, (size_t)in_sc * howmany);
    memset(out, 0, sizeof(kiss_fft_scalar) * out_sc * howmany);

    const int intra = argc > 6 && !strcmp(argv[6], "intra");
    /* cfgs before the timed passes: one per thread in batch mode, one in intra mode */
    const int ncfg = intra ? 1 : omp_get_max_threads();
    void **cfgs = calloc((size_t)ncfg, sizeof(void *));
    int failed = 0;
    for (int c = 0; c < ncfg; ++c) {
        if (!strcmp(kind, "c2c"))
            cfgs[c] = kiss_fft_alloc(dims[0], 0, NULL, NULL);
        else if (!strcmp(kind, "r2c"))
            cfgs[c] = kiss_fftr_alloc(dims[0], 0, NULL, NULL);
        else if (!strcmp(kind, "c2r"))
            cfgs[c] = kiss_fftr_alloc(dims[0], 1, NULL, NULL);
        else if (!strcmp(kind, "nd"))
            cfgs[c] = kiss_fftnd_alloc(dims, nd, 0, NULL, NULL);
        else if (!strcmp(kind, "ndr"))
            cfgs[c] = kiss_fftndr_alloc(dims, nd, 0, NULL, NULL);
        else
            cfgs[c] = kiss_fftndr_alloc(dims, nd, 1, NULL, NULL);
        if (!cfgs[c])
            failed = 1;
    }
    double *t = malloc(sizeof(double) * (reps + 1));
    for (int r = 0; r <= reps && !failed; ++r) {
        const double t0 = now_s();
        if (intra) {
            for (long b = 0; b < howmany; ++b)
                run_one(kind, cfgs[0], in + b * in_sc, out + b * out_sc);
        } else {
#pragma omp parallel
            {
                void *cfg = cfgs[omp_get_thread_num()];
#pragma omp for schedule(static)
                for (long b = 0; b < howmany; ++b)
                    run_one(kind, cfg, in + b * in_sc, out + b * out_sc);
            }
        }
        t[r] = now_s() - t0;
    }
    for (int c = 0; c < ncfg; ++c)
        free(cfgs[c]);
    free(cfgs);
    if (failed) {
        printf("{\"error\": \"reference alloc returned NULL for %s %s\"}\n", kind, argv[2]);
        return 1;
    }
    qsort(t + 1, reps, sizeof(double), cmp_d);
    const double med = t[1 + reps / 2];
    const double bytes = (double)sizeof(kiss_fft_scalar) * (in_sc + out_sc) * howmany;
    printf("{\"kind\": \"%s\", \"n\": \"%s\", \"howmany\": %ld, \"threads\": %d, \"mode\": \"%s\", "
           "\"openmp\": %d, \"median_s\": %.6g, \"transforms_per_s\": %.6g, \"GBps\": %.6g, "
           "\"scalar_bytes\": %d}\n",
           kind, argv[2], howmany, omp_get_max_threads(), intra ? "intra" : "batch",
#ifdef _OPENMP
           1,
#else
           0,
#endif
           med, howmany / med, bytes / med * 1e-9, (int)sizeof(kiss_fft_scalar));
    free(in);
    free(out);
    free(t);
    return 0;
}
