/*
 * fftutil.c -- stream transformer over the drop-in library, the protocol of the
 * reference's tools/fftutil.c:18-126 (built there as "fft" and, for its tests, as
 * "fastfilt-<type>"):
 *
 *     fftutil-<type> [-n d1[,d2,...]] [-i] [-R] [-b records] [infile|- [outfile|-]]
 *
 * Input and output are raw interleaved scalars of the build's type.  A record is one
 * transform: complex 1-D (nfft complex in and out), real 1-D (-R: nfft scalars to
 * nfft/2+1 complex, or back with -i), complex N-D (prod(dims) complex), real N-D (-R with
 * several dims: the last dimension real).  Default: one 1024-point complex transform per
 * record, forward.
 *
 * Unlike the reference, which transforms one record per call, records are read in batches
 * (-b, default: as many as fit 256 MiB of input and output) and each batch is one batched
 * library call on the GPU; the output is the same as record-by-record processing (bit for
 * bit in the fixed-point builds).  A trailing partial record is dropped in every mode (the
 * reference drops it except in the real N-D mode, where it transforms the partial record
 * padded with the previous record's data).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "kiss_fft.h"
#include "kiss_fft_batch.h"
#include "kiss_fftnd.h"
#include "kiss_fftndr.h"
#include "kiss_fftr.h"

static int parse_dims(char *arg, int *dims, int cap)
{
    int nd = 0;
    char *p = arg;
    while (p && *p && nd < cap) {
        char *c = strchr(p, ',');
        if (c)
            *c++ = '\0';
        dims[nd++] = atoi(p);
        p = c;
    }
    return nd;
}

static void die(const char *what)
{
    const char *e = kiss_fft_last_error();
    fprintf(stderr, "fftutil: %s%s%s\n", what, e && *e ? ": " : "", e ? e : "");
    exit(1);
}

int main(int argc, char **argv)
{
    int inverse = 0, real = 0, ndims = 1, dims[32], c, i;
    long long batch = 0;
    FILE *fin = stdin, *fout = stdout;
    dims[0] = 1024;
    while ((c = getopt(argc, argv, "n:iRb:")) != -1) {
        switch (c) {
        case 'n': ndims = parse_dims(optarg, dims, 32); break;
        case 'i': inverse = 1; break;
        case 'R': real = 1; break;
        case 'b': batch = atoll(optarg); break;
        default:
            fprintf(stderr, "usage options:\n"
                            "\t-n d1[,d2,d3...]: fft dimension(s)\n"
                            "\t-i : inverse\n"
                            "\t-R : real input samples, not complex\n"
                            "\t-b records : records per GPU batch\n");
            return 1;
        }
    }
    if (ndims < 1) {
        fprintf(stderr, "fftutil: bad -n\n");
        return 1;
    }
    if (optind < argc) {
        if (strcmp(argv[optind], "-") != 0 && !(fin = fopen(argv[optind], "rb"))) {
            perror(argv[optind]);
            return 1;
        }
        ++optind;
    }
    if (optind < argc) {
        if (strcmp(argv[optind], "-") != 0 && !(fout = fopen(argv[optind], "wb"))) {
            perror(argv[optind]);
            return 1;
        }
        ++optind;
    }

    long long prod = 1;
    for (i = 0; i < ndims; ++i) {
        if (dims[i] <= 0) {
            fprintf(stderr, "fftutil: bad dimension %d\n", dims[i]);
            return 1;
        }
        prod *= dims[i];
    }
    const long long last = dims[ndims - 1];
    const long long bins = real ? prod / last * (last / 2 + 1) : prod;  /* complex values, spectrum side */
    const size_t sb = sizeof(kiss_fft_scalar), cb = sizeof(kiss_fft_cpx);
    const size_t in_rec = real ? (inverse ? bins * cb : prod * sb) : prod * cb;
    const size_t out_rec = real ? (inverse ? prod * sb : bins * cb) : prod * cb;
    if (batch <= 0) {
        batch = (long long)((256u << 20) / (in_rec + out_rec));
        if (batch < 1)
            batch = 1;
    }

    void *cfg;
    if (ndims == 1)
        cfg = real ? (void *)kiss_fftr_alloc(dims[0], inverse, NULL, NULL)
                   : (void *)kiss_fft_alloc(dims[0], inverse, NULL, NULL);
    else
        cfg = real ? (void *)kiss_fftndr_alloc(dims, ndims, inverse, NULL, NULL)
                   : (void *)kiss_fftnd_alloc(dims, ndims, inverse, NULL, NULL);
    if (!cfg)
        die("cannot plan this size");

    char *ib = (char *)malloc(in_rec * (size_t)batch), *ob = (char *)malloc(out_rec * (size_t)batch);
    if (!ib || !ob)
        die("out of host memory");
    size_t got;
    while ((got = fread(ib, in_rec, (size_t)batch, fin)) > 0) {
        const long long n = (long long)got;
        int rc;
        if (ndims == 1 && !real)
            rc = kiss_fft_batch((kiss_fft_cfg)cfg, (const kiss_fft_cpx *)ib, 1, prod, (kiss_fft_cpx *)ob, prod, n);
        else if (ndims == 1 && !inverse)
            rc = kiss_fftr_batch((kiss_fftr_cfg)cfg, (const kiss_fft_scalar *)ib, prod, (kiss_fft_cpx *)ob, bins, n);
        else if (ndims == 1)
            rc = kiss_fftri_batch((kiss_fftr_cfg)cfg, (const kiss_fft_cpx *)ib, bins, (kiss_fft_scalar *)ob, prod, n);
        else if (!real)
            rc = kiss_fftnd_batch((kiss_fftnd_cfg)cfg, (const kiss_fft_cpx *)ib, (kiss_fft_cpx *)ob, n);
        else if (!inverse)
            rc = kiss_fftndr_batch((kiss_fftndr_cfg)cfg, (const kiss_fft_scalar *)ib, (kiss_fft_cpx *)ob, n);
        else
            rc = kiss_fftndri_batch((kiss_fftndr_cfg)cfg, (const kiss_fft_cpx *)ib, (kiss_fft_scalar *)ob, n);
        if (rc != 0)
            die("transform failed");
        if (fwrite(ob, out_rec, got, fout) != got)
            die("write failed");
    }
    free(cfg);
    free(ib);
    free(ob);
    if (fout != stdout)
        fclose(fout);
    if (fin != stdin)
        fclose(fin);
    return 0;
}
