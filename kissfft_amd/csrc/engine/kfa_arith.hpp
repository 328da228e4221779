// kfa_arith.hpp -- scalar kinds and radix butterflies for the MI355X engine.
//
// Three scalar kinds, matching the reference's datatype builds (kiss_fft.h:73-85):
//   F32    float32, interleaved float2 storage (8 B per complex)
//   FX<16> Q15 fixed point, int16 pairs (4 B per complex), int32 products
//   FX<32> Q31 fixed point, int32 pairs (8 B per complex), int64 products
//
// The fixed-point butterflies reproduce the reference arithmetic operation for operation
// (kiss_fft.c:15-233 with the macros of _kiss_fft_guts.h:43-124): inputs scaled by
// C_FIXDIV, products rounded once by sround, and every value the reference keeps in a
// kiss_fft_scalar truncated to the scalar width.  Add/sub chains whose results are only
// stored are evaluated in 32-bit integers and truncated once, which gives the same bits
// as truncating after every step because two's-complement truncation is a ring
// homomorphism; truncation is kept explicit wherever a value feeds a non-linear step
// (a rounded product or a shift).  The float butterflies follow the same formulas.
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include <stdint.h>
#else  // compiled at run time (kfa_rtc.hip): the runtime compiler's own integer types
using __hip_internal::int16_t;
using __hip_internal::int32_t;
using __hip_internal::int64_t;
using __hip_internal::uint16_t;
using __hip_internal::uint32_t;
using __hip_internal::uint64_t;
#endif

namespace kfa {

struct alignas(4) q15x2 {
    int16_t r, i;
};
struct alignas(8) q31x2 {
    int32_t r, i;
};

struct F32 {
    typedef float2 store2;  // one complex in HBM / LDS
    typedef float scalar;
    typedef float reg;
    static constexpr bool fixed = false;
};

template <int BITS>
struct FX;

template <>
struct FX<16> {
    typedef q15x2 store2;
    typedef int16_t scalar;
    typedef int reg;      // holds a sign-extended int16
    typedef int32_t wide; // SAMPPROD
    static constexpr bool fixed = true;
    static constexpr int frac = 15;
    static constexpr int64_t samp_max = 32767;
};

template <>
struct FX<32> {
    typedef q31x2 store2;
    typedef int32_t scalar;
    typedef int reg;
    typedef int64_t wide;
    static constexpr bool fixed = true;
    static constexpr int frac = 31;
    static constexpr int64_t samp_max = 2147483647;
};

typedef FX<16> Q15;
typedef FX<32> Q31;

template <class A>
struct Cx {
    typename A::reg r, i;
};

// ------------------------------------------------------------------------------------
// Scalar operations per kind
// ------------------------------------------------------------------------------------
template <class A>
struct Ops;

template <>
struct Ops<F32> {
    typedef float R;
    typedef Cx<F32> C;
    __device__ __forceinline__ static C load(const float2 &s) { return C{s.x, s.y}; }
    __device__ __forceinline__ static float2 pack(C c) { return make_float2(c.r, c.i); }
    __device__ __forceinline__ static R add(R a, R b) { return a + b; }
    __device__ __forceinline__ static R sub(R a, R b) { return a - b; }
    __device__ __forceinline__ static R neg(R a) { return -a; }
    __device__ __forceinline__ static R smul(R a, R b) { return a * b; }
    __device__ __forceinline__ static R half(R a) { return a * 0.5f; }
    // HALF_OF(a + b) / HALF_OF(a - b) on unstored sums (kiss_fftr.c:112-115)
    __device__ __forceinline__ static R half_add(R a, R b) { return (a + b) * 0.5f; }
    __device__ __forceinline__ static R half_sub(R a, R b) { return (a - b) * 0.5f; }
    __device__ __forceinline__ static C cmul(C a, C b)
    {
        return C{a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r};
    }
    __device__ __forceinline__ static C fixdiv(C c, int) { return c; }
    __device__ __forceinline__ static C scale(C c, R s) { return C{c.r * s, c.i * s}; }
    __device__ __forceinline__ static C cadd(C a, C b) { return C{a.r + b.r, a.i + b.i}; }
    __device__ __forceinline__ static C csub(C a, C b) { return C{a.r - b.r, a.i - b.i}; }
    __device__ __forceinline__ static C zero() { return C{0.f, 0.f}; }
};

template <int BITS>
struct Ops<FX<BITS>> {
    typedef FX<BITS> A;
    typedef int R;
    typedef typename A::wide W;
    typedef Cx<A> C;
    // store into a kiss_fft_scalar: keep the low BITS bits, two's complement
    __device__ __forceinline__ static R st(W x) { return (R)(typename A::scalar)x; }
    // sround, _kiss_fft_guts.h:65
    __device__ __forceinline__ static R rnd(W x)
    {
        return (R)(typename A::scalar)((x + ((W)1 << (A::frac - 1))) >> A::frac);
    }
    __device__ __forceinline__ static C load(const typename A::store2 &s) { return C{s.r, s.i}; }
    __device__ __forceinline__ static typename A::store2 pack(C c)
    {
        typename A::store2 s;
        s.r = (typename A::scalar)c.r;
        s.i = (typename A::scalar)c.i;
        return s;
    }
    // sums wrap in 32-bit int arithmetic and are then stored (truncated)
    __device__ __forceinline__ static R add(R a, R b) { return st((W)(int)((unsigned)a + (unsigned)b)); }
    __device__ __forceinline__ static R sub(R a, R b) { return st((W)(int)((unsigned)a - (unsigned)b)); }
    __device__ __forceinline__ static R neg(R a) { return st((W)(int)(0u - (unsigned)a)); }
    __device__ __forceinline__ static R smul(R a, R b) { return rnd((W)a * (W)b); }
    // HALF_OF of a stored scalar: arithmetic shift (_kiss_fft_guts.h:130)
    __device__ __forceinline__ static R half(R a) { return a >> 1; }
    // HALF_OF of the int value of an unstored sum, then stored (kiss_fftr.c:112-115)
    __device__ __forceinline__ static R half_add(R a, R b) { return st((W)((int)((unsigned)a + (unsigned)b) >> 1)); }
    __device__ __forceinline__ static R half_sub(R a, R b) { return st((W)((int)((unsigned)a - (unsigned)b) >> 1)); }
    // C_MUL, _kiss_fft_guts.h:69-71
    __device__ __forceinline__ static C cmul(C a, C b)
    {
        return C{rnd((W)a.r * b.r - (W)a.i * b.i), rnd((W)a.r * b.i + (W)a.i * b.r)};
    }
    // C_FIXDIV, _kiss_fft_guts.h:73-78: times floor(SAMP_MAX/p), then sround
    __device__ __forceinline__ static C fixdiv(C c, int p)
    {
        const W k = (W)(A::samp_max / p);
        return C{rnd((W)c.r * k), rnd((W)c.i * k)};
    }
    // C_MULBYSCALAR, _kiss_fft_guts.h:80-82
    __device__ __forceinline__ static C scale(C c, R s) { return C{rnd((W)c.r * s), rnd((W)c.i * s)}; }
    __device__ __forceinline__ static C cadd(C a, C b) { return C{add(a.r, b.r), add(a.i, b.i)}; }
    __device__ __forceinline__ static C csub(C a, C b) { return C{sub(a.r, b.r), sub(a.i, b.i)}; }
    __device__ __forceinline__ static C zero() { return C{0, 0}; }
};

// ------------------------------------------------------------------------------------
// Radix butterflies.  v[q] holds input q of the butterfly (sub-transform q at column u);
// w[r-1] is the stage twiddle for input r (table entry r*u*fstride); ya/yb are the
// fixed rotations table[fstride*m] and table[2*fstride*m] (radix 3 uses ya only).
// Results overwrite v in output order.
//
// ONE: every twiddle of the butterfly is table entry 0, (SAMP_MAX, 0) in fixed point.
// C_MUL by it returns its operand for every component in (-2^(B-1), 2^(B-1)] (B the
// fraction bits), and C_FIXDIV by p >= 2 always lands there, so the product is left out
// with the same bits.  Only fixed-point stages pass it.
// ------------------------------------------------------------------------------------

// kiss_fft.c:15-36
template <class A, bool ONE = false>
__device__ __forceinline__ void bfly2(Cx<A> *v, const Cx<A> *w)
{
    typedef Ops<A> O;
    Cx<A> a = O::fixdiv(v[0], 2), b = O::fixdiv(v[1], 2);
    Cx<A> t = ONE ? b : O::cmul(b, w[0]);
    v[1] = O::csub(a, t);
    v[0] = O::cadd(a, t);
}

// kiss_fft.c:38-84
template <class A, bool ONE = false>
__device__ __forceinline__ void bfly4(Cx<A> *v, const Cx<A> *w, bool inverse)
{
    typedef Ops<A> O;
    Cx<A> x0 = O::fixdiv(v[0], 4), x1 = O::fixdiv(v[1], 4);
    Cx<A> x2 = O::fixdiv(v[2], 4), x3 = O::fixdiv(v[3], 4);
    Cx<A> s0 = ONE ? x1 : O::cmul(x1, w[0]);
    Cx<A> s1 = ONE ? x2 : O::cmul(x2, w[1]);
    Cx<A> s2 = ONE ? x3 : O::cmul(x3, w[2]);
    Cx<A> s5 = O::csub(x0, s1);
    x0 = O::cadd(x0, s1);
    Cx<A> s3 = O::cadd(s0, s2);
    Cx<A> s4 = O::csub(s0, s2);
    v[2] = O::csub(x0, s3);
    v[0] = O::cadd(x0, s3);
    if (inverse) {
        v[1] = Cx<A>{O::sub(s5.r, s4.i), O::add(s5.i, s4.r)};
        v[3] = Cx<A>{O::add(s5.r, s4.i), O::sub(s5.i, s4.r)};
    } else {
        v[1] = Cx<A>{O::add(s5.r, s4.i), O::sub(s5.i, s4.r)};
        v[3] = Cx<A>{O::sub(s5.r, s4.i), O::add(s5.i, s4.r)};
    }
}

// kiss_fft.c:86-128
template <class A, bool ONE = false>
__device__ __forceinline__ void bfly3(Cx<A> *v, const Cx<A> *w, Cx<A> epi3)
{
    typedef Ops<A> O;
    Cx<A> x0 = O::fixdiv(v[0], 3), x1 = O::fixdiv(v[1], 3), x2 = O::fixdiv(v[2], 3);
    Cx<A> s1 = ONE ? x1 : O::cmul(x1, w[0]);
    Cx<A> s2 = ONE ? x2 : O::cmul(x2, w[1]);
    Cx<A> s3 = O::cadd(s1, s2);
    Cx<A> s0 = O::csub(s1, s2);
    Cx<A> o1{O::sub(x0.r, O::half(s3.r)), O::sub(x0.i, O::half(s3.i))};
    s0 = O::scale(s0, epi3.i);
    x0 = O::cadd(x0, s3);
    Cx<A> o2{O::add(o1.r, s0.i), O::sub(o1.i, s0.r)};
    o1 = Cx<A>{O::sub(o1.r, s0.i), O::add(o1.i, s0.r)};
    v[0] = x0;
    v[1] = o1;
    v[2] = o2;
}

// kiss_fft.c:130-189
template <class A, bool ONE = false>
__device__ __forceinline__ void bfly5(Cx<A> *v, const Cx<A> *w, Cx<A> ya, Cx<A> yb)
{
    typedef Ops<A> O;
    Cx<A> x0 = O::fixdiv(v[0], 5);
    const Cx<A> c0 = x0;
    Cx<A> c1 = O::fixdiv(v[1], 5), c2 = O::fixdiv(v[2], 5);
    Cx<A> c3 = O::fixdiv(v[3], 5), c4 = O::fixdiv(v[4], 5);
    if constexpr (!ONE) {
        c1 = O::cmul(c1, w[0]);
        c2 = O::cmul(c2, w[1]);
        c3 = O::cmul(c3, w[2]);
        c4 = O::cmul(c4, w[3]);
    }
    Cx<A> c7 = O::cadd(c1, c4), c10 = O::csub(c1, c4);
    Cx<A> c8 = O::cadd(c2, c3), c9 = O::csub(c2, c3);

    x0.r = O::add(x0.r, O::add(c7.r, c8.r));
    x0.i = O::add(x0.i, O::add(c7.i, c8.i));

    Cx<A> c5{O::add(O::add(c0.r, O::smul(c7.r, ya.r)), O::smul(c8.r, yb.r)),
             O::add(O::add(c0.i, O::smul(c7.i, ya.r)), O::smul(c8.i, yb.r))};
    Cx<A> c6{O::add(O::smul(c10.i, ya.i), O::smul(c9.i, yb.i)),
             O::sub(O::neg(O::smul(c10.r, ya.i)), O::smul(c9.r, yb.i))};
    Cx<A> c11{O::add(O::add(c0.r, O::smul(c7.r, yb.r)), O::smul(c8.r, ya.r)),
              O::add(O::add(c0.i, O::smul(c7.i, yb.r)), O::smul(c8.i, ya.r))};
    Cx<A> c12{O::add(O::neg(O::smul(c10.i, yb.i)), O::smul(c9.i, ya.i)),
              O::sub(O::smul(c10.r, yb.i), O::smul(c9.r, ya.i))};

    v[0] = x0;
    v[1] = O::csub(c5, c6);
    v[4] = O::cadd(c5, c6);
    v[2] = O::cadd(c11, c12);
    v[3] = O::csub(c11, c12);
}

}  // namespace kfa
