// kfa_dft.hpp -- one radix stage too long for LDS, as a global-memory pass.
//
// kf_factor (kiss_fft.c:306-328) takes a cofactor above sqrt(n) whole, as the last factor,
// so a length with a large prime factor p (above one LDS line: 8192 points for float and
// Q31, 16384 for Q15) starts with p-point DFTs at m = 1: kf_bfly_generic at the leaf of the
// recursion (kiss_fft.c:191-233).  Column o (< Q = n/p) of a transform reads the inputs
// x[r] = in[(o + r*Q) * in_el], r < p, scales each by C_FIXDIV(., p) (fixed point), and
// output q is
//     X[q] = x'[0] + sum_{r=1}^{p-1} x'[r] * tw[(r*Q*q) mod n]
// written to out[(o + q*Q) * out_el]: the layout the line kernel's second four-step pass
// reads.  Fixed point reproduces the reference exactly: every product is the reference's
// rounded C_MUL, and the sum of kiss_fft_scalar values wraps modulo 2^16 / 2^32 as the
// reference's repeated C_ADDTO into a kiss_fft_scalar does, so the order of the additions
// does not change the result.  Float accumulates in FP64 (O(p) terms; the result is within
// the float tolerance, not bit-equal to the reference's order).
//
// One thread per output q; a workgroup covers 256 outputs of one column and stages the
// column's scaled inputs through LDS 256 at a time.  The twiddle index advances by
// Q*q (mod n) per input.
#pragma once

#include <hip/hip_runtime.h>

#include "kfa_arith.hpp"

namespace kfa {

struct DftStageArgs {
    const void *in;
    void *out;
    const void *tw;        // the plan's full twiddle table (n entries)
    long long n, Q;        // n = p * Q
    int p;
    long long ncols;       // transforms * Q
    long long in_hi, in_el, out_hi, out_el;
    long long qtiles;      // ceil(p / 256)
};

template <class A>
struct DftAcc;

template <>
struct DftAcc<F32> {
    double r = 0, i = 0;
    __device__ __forceinline__ void first(Cx<F32> x)
    {
        r = x.r;
        i = x.i;
    }
    __device__ __forceinline__ void mac(Cx<F32> x, float2 w)
    {
        r += (double)x.r * w.x - (double)x.i * w.y;
        i += (double)x.r * w.y + (double)x.i * w.x;
    }
    __device__ __forceinline__ float2 out() const { return make_float2((float)r, (float)i); }
};

template <int BITS>
struct DftAcc<FX<BITS>> {
    typedef FX<BITS> A;
    unsigned r = 0, i = 0;  // sums of stored scalars, modulo 2^32 (truncated on output)
    __device__ __forceinline__ void first(Cx<A> x)
    {
        r = (unsigned)x.r;
        i = (unsigned)x.i;
    }
    __device__ __forceinline__ void mac(Cx<A> x, typename A::store2 w)
    {
        const Cx<A> t = Ops<A>::cmul(x, Ops<A>::load(w));
        r += (unsigned)t.r;
        i += (unsigned)t.i;
    }
    __device__ __forceinline__ typename A::store2 out() const
    {
        typename A::store2 s;
        s.r = (typename A::scalar)(int)r;
        s.i = (typename A::scalar)(int)i;
        return s;
    }
};

template <class A>
__global__ void __launch_bounds__(256) k_dft_stage(DftStageArgs a)
{
    typedef typename A::store2 St;
    __shared__ Cx<A> xs[256];
    const long long blk = blockIdx.x;
    const long long col = blk / a.qtiles;
    const long long qt = blk - col * a.qtiles;
    const long long b = col / a.Q, o = col - b * a.Q;
    const int t = threadIdx.x;
    const long long q = qt * 256 + t;
    const St *in = static_cast<const St *>(a.in) + b * a.in_hi + o * a.in_el;
    const St *tw = static_cast<const St *>(a.tw);
    const long long step = q < a.p ? a.Q * q : 0;  // < n
    long long idx = 0;
    DftAcc<A> acc;
    for (int r0 = 0; r0 < a.p; r0 += 256) {
        __syncthreads();  // the previous chunk is consumed
        if (r0 + t < a.p)
            xs[t] = Ops<A>::fixdiv(Ops<A>::load(in[(long long)(r0 + t) * a.Q * a.in_el]), a.p);
        __syncthreads();
        const int cnt = a.p - r0 < 256 ? a.p - r0 : 256;
        int j = 0;
        if (r0 == 0) {
            acc.first(xs[0]);
            j = 1;
        }
        for (; j < cnt; ++j) {
            idx += step;
            if (idx >= a.n)
                idx -= a.n;
            acc.mac(xs[j], tw[idx]);
        }
    }
    if (q < a.p)
        static_cast<St *>(a.out)[b * a.out_hi + (o + q * a.Q) * a.out_el] = acc.out();
}

}  // namespace kfa
