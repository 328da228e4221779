// kfa_blue.hpp -- one-kernel convolutions in LDS: Bluestein's chirp-z transform for float
// lengths with large prime factors, and overlap-scrap fast FIR filtering.
//
// The reference transforms a prime factor p by the direct O(p^2) butterfly
// (kf_bfly_generic, kiss_fft.c:191-233).  For float lines whose schedule contains a large
// prime this kernel computes the same DFT through a cyclic convolution of power-of-two
// length M >= 2N-1 (Bluestein, "A linear filtering approach to the computation of the
// discrete Fourier transform", 1970):
//   w[n] = exp(-+ i*pi*n^2/N),   X[k] = w[k] * sum_n (x[n] w[n]) conj(w[k-n])
// i.e. X = w * IFFT_M( FFT_M(x*w, zero-padded) * B ),  B = FFT_M(conj(w) wrapped) / M.
// One line per T threads, entirely in LDS: the chirp multiply on load, the forward M-point
// schedule, the pointwise product with B, the inverse schedule, the chirp on store.
// w (N entries) and B (M entries) are computed in FP64 on the host and rounded once.
#pragma once

#include "kfa_tile.hpp"

namespace kfa {

struct BlueArgs {
    const void *in;
    void *out;
    const void *stw_f, *stw_i;  // stage tables of the M-point schedule, forward and inverse
    const void *chirp;          // w[n], n < N
    const void *bfft;           // B[m], m < M
    int n;
    long long nlines;
    long long in_hi, in_el, out_hi, out_el;
    int nseg;                   // 0: one buffer (in / out); else segments as in FastArgs
    FastSeg seg[KFA_MAX_SEG];
};

// The Bluestein body for workgroup `blk` of its launch (see k_fast_body).
template <int T, int LPB, class RS>
__device__ __forceinline__ void k_blue_body(const BlueArgs &a, long long blk)
{
    typedef F32::store2 St;
    typedef Ops<F32> O;
    typedef Cx<F32> C;
    constexpr int M = RS::N();
    constexpr int IT = M / T;
    static_assert(M % T == 0, "threads per line must divide M");
    extern __shared__ __align__(16) unsigned char kfa_blue_smem[];
    const int line = threadIdx.x / T, t = threadIdx.x - line * T;
    St *buf = reinterpret_cast<St *>(kfa_blue_smem) + line * (M + 1);
    const long long gl = blk * LPB + line;
    const bool ok = gl < a.nlines;
    const int n = a.n;
    const St *chirp = reinterpret_cast<const St *>(a.chirp);
    const St *bf = reinterpret_cast<const St *>(a.bfft);

    const St *src;
    St *dst;
    line_base<St>(a, ok ? gl : 0, src, dst);
    {
        C v[IT];
#pragma unroll
        for (int k = 0; k < IT; ++k) {
            const int i = t + k * T;
            v[k] = (ok && i < n) ? O::cmul(O::load(src[(long long)i * a.in_el]), O::load(chirp[i])) : O::zero();
        }
#pragma unroll
        for (int k = 0; k < IT; ++k)
            buf[t + k * T] = O::pack(v[k]);
    }
    __syncthreads();
    tile_stages<F32, false, T, RS>(std::make_integer_sequence<int, RS::K>{}, buf,
                                   reinterpret_cast<const St *>(a.stw_f), t);
#pragma unroll
    for (int k = 0; k < IT; ++k) {
        const int i = t + k * T;
        buf[i] = O::pack(O::cmul(O::load(buf[i]), O::load(bf[i])));
    }
    __syncthreads();
    tile_stages<F32, true, T, RS>(std::make_integer_sequence<int, RS::K>{}, buf,
                                  reinterpret_cast<const St *>(a.stw_i), t);
    if (!ok)
        return;
#pragma unroll
    for (int k = 0; k < IT; ++k) {
        const int i = t + k * T;
        if (i < n)
            dst[(long long)i * a.out_el] = O::pack(O::cmul(O::load(buf[i]), O::load(chirp[i])));
    }
}

template <int T, int LPB, class RS>
__global__ void __launch_bounds__(T *LPB) k_blue(BlueArgs a)
{
    k_blue_body<T, LPB, RS>(a, blockIdx.x);
}

}  // namespace kfa

namespace kfa {

// Fast convolution by overlap-scrap (the reference's tools/kiss_fastfir.c:167-239): block
// k reads the window in[k*ngood .. k*ngood + N), is transformed, multiplied by the filter
// spectrum H (already scaled by 1/N), transformed back, and its first ngood samples --
// the ones the rotated impulse response leaves free of wrap-around -- are written to
// out[k*ngood ..].  One block per T threads, entirely in LDS.  REAL: the samples are
// real and two consecutive blocks travel as the real and imaginary parts of one complex
// block; with a real impulse response the two convolutions stay separate.
struct ConvArgs {
    const void *in;
    void *out;
    const void *stw_f, *stw_i;  // stage tables of the N-point schedule
    const void *H;              // N complex
    long long nlines;           // complex blocks (REAL: ceil(nblocks / 2))
    long long nblocks;          // blocks of samples
    long long ngood;
};

template <int T, int LPB, bool REAL, class RS>
__global__ void __launch_bounds__(T *LPB) k_conv(ConvArgs a)
{
    typedef F32::store2 St;
    typedef Ops<F32> O;
    typedef Cx<F32> C;
    constexpr int N = RS::N();
    constexpr int IT = N / T;
    static_assert(N % T == 0, "threads per line must divide N");
    extern __shared__ __align__(16) unsigned char kfa_conv_smem[];
    const int line = threadIdx.x / T, t = threadIdx.x - line * T;
    St *buf = reinterpret_cast<St *>(kfa_conv_smem) + line * (N + 1);
    const long long gl = (long long)blockIdx.x * LPB + line;
    const bool ok = gl < a.nlines;
    const long long b0 = REAL ? 2 * gl : gl;          // first block of this line
    const bool two = REAL && b0 + 1 < a.nblocks;      // second block (imaginary part)
    const St *H = reinterpret_cast<const St *>(a.H);

    {
        C v[IT];
        if constexpr (REAL) {
            const float *x = reinterpret_cast<const float *>(a.in);
            const float *xa = x + (ok ? b0 : 0) * a.ngood, *xb = x + (two ? b0 + 1 : 0) * a.ngood;
#pragma unroll
            for (int k = 0; k < IT; ++k) {
                const int i = t + k * T;
                v[k] = C{ok ? xa[i] : 0.0f, two ? xb[i] : 0.0f};
            }
        } else {
            const St *x = reinterpret_cast<const St *>(a.in) + (ok ? gl : 0) * a.ngood;
#pragma unroll
            for (int k = 0; k < IT; ++k)
                v[k] = ok ? O::load(x[t + k * T]) : O::zero();
        }
#pragma unroll
        for (int k = 0; k < IT; ++k)
            buf[t + k * T] = O::pack(v[k]);
    }
    __syncthreads();
    tile_stages<F32, false, T, RS>(std::make_integer_sequence<int, RS::K>{}, buf,
                                   reinterpret_cast<const St *>(a.stw_f), t);
#pragma unroll
    for (int k = 0; k < IT; ++k) {
        const int i = t + k * T;
        buf[i] = O::pack(O::cmul(O::load(buf[i]), O::load(H[i])));
    }
    __syncthreads();
    tile_stages<F32, true, T, RS>(std::make_integer_sequence<int, RS::K>{}, buf,
                                  reinterpret_cast<const St *>(a.stw_i), t);
    if (!ok)
        return;
    if constexpr (REAL) {
        float *y = reinterpret_cast<float *>(a.out);
        float *ya = y + b0 * a.ngood, *yb = y + (b0 + 1) * a.ngood;
#pragma unroll
        for (int k = 0; k < IT; ++k) {
            const int i = t + k * T;
            if (i < a.ngood) {
                const C c = O::load(buf[i]);
                ya[i] = c.r;
                if (two)
                    yb[i] = c.i;
            }
        }
    } else {
        St *y = reinterpret_cast<St *>(a.out) + gl * a.ngood;
#pragma unroll
        for (int k = 0; k < IT; ++k) {
            const int i = t + k * T;
            if (i < a.ngood)
                y[i] = buf[i];
        }
    }
}

}  // namespace kfa
