// kfa_q15p.hpp -- Q15 butterflies on packed complex words.
//
// A Q15 complex is one 32-bit word: real part in the low 16 bits, imaginary part in the
// high 16 bits -- the in-memory int16 pair (kiss_fft_cpx, kiss_fft.h:87-90) loaded as is.
// Each reference operation maps onto CDNA packed-integer instructions with the same bits:
//   C_ADD / C_SUB with truncation to int16   -> v_pk_add_u16 / v_pk_sub_u16 (wrap per half)
//   C_MUL, one sround per component          -> v_dot2_i32_i16 (a.r*b.r - a.i*b.i + 2^14 in
//                                               one instruction, no intermediate rounding),
//                                               then bits 15..30 of each sum repacked
//   C_FIXDIV / C_MULBYSCALAR / S_MUL         -> the same dot2 form with one zero lane
//   HALF_OF of a stored scalar               -> v_pk_ashrrev_i16
//   multiplication by -i / +i                -> a 16-bit rotate plus v_pk_mul_lo_u16 by (1,-1)
// Operation order per butterfly is the reference's (kiss_fft.c:15-189, macros of
// _kiss_fft_guts.h:43-124): sums of stored values may be regrouped because int16
// truncation is a ring homomorphism; every rounded product sees exactly the int16 operands
// the reference rounds.
//
// Twiddles are stored pre-arranged: entry w becomes two words (w.r, -w.i) and (w.i, w.r),
// so the real and imaginary dot products read them directly (|w.i| <= 32767, so -w.i fits).
#pragma once

#include "kfa_arith.hpp"

namespace kfa {

struct Q15P {
    typedef q15x2 store2;
    typedef unsigned reg;
    static constexpr bool fixed = true;
    static constexpr bool packed = true;
};

template <>
struct Cx<Q15P> {
    unsigned w;
};

namespace q15p {

typedef short v2s __attribute__((ext_vector_type(2)));
typedef unsigned u32;

__device__ __forceinline__ v2s as_v(u32 w) { return __builtin_bit_cast(v2s, w); }
__device__ __forceinline__ u32 as_u(v2s v) { return __builtin_bit_cast(u32, v); }

__device__ __forceinline__ u32 add(u32 a, u32 b) { return as_u(as_v(a) + as_v(b)); }
__device__ __forceinline__ u32 sub(u32 a, u32 b) { return as_u(as_v(a) - as_v(b)); }
__device__ __forceinline__ u32 half(u32 a) { return as_u(as_v(a) >> (v2s){1, 1}); }
__device__ __forceinline__ u32 mul_lo(u32 a, u32 m) { return as_u(as_v(a) * as_v(m)); }
// (a.i, a.r)
__device__ __forceinline__ u32 swap(u32 a) { return (a >> 16) | (a << 16); }
// (x.lo, y.hi)
__device__ __forceinline__ u32 lohi(u32 x, u32 y) { return (x & 0xffffu) | (y & 0xffff0000u); }

// Instruction forms (compile-time; both forms compute the same bits; on random data the
// pair is 9 % faster than neither, at 2^20 transforms of 512 points):
//   KFA_Q15_DOT_ASM   dot products as the three-address VOP3P v_dot2_i32_i16 with the
//                     rounding constant in a scalar register (else the builtin, which
//                     selects the accumulate form and copies the constant per product)
//   KFA_Q15_SDWA_PACK the high half of a packed result written by a word-select shift
//                     (two instructions instead of shift, shift, bit-field insert)
#ifndef KFA_Q15_DOT_ASM
#define KFA_Q15_DOT_ASM 1
#endif
#ifndef KFA_Q15_SDWA_PACK
#define KFA_Q15_SDWA_PACK 1
#endif

// bits 15..30 of two sums, packed (sround of each component, truncated to int16)
__device__ __forceinline__ u32 round_pack(int re, int im)
{
#if KFA_Q15_SDWA_PACK
    u32 r;
    asm("v_lshrrev_b32 %0, 15, %1" : "=v"(r) : "v"(re));
    asm("v_lshrrev_b32_sdwa %0, 15, %1 dst_sel:WORD_1 dst_unused:UNUSED_PRESERVE src0_sel:DWORD src1_sel:DWORD"
        : "+v"(r)
        : "v"(im));
    return r;
#else
    return lohi((u32)(re >> 15), (u32)im << 1);
#endif
}

// a.lo*b.lo + a.hi*b.hi + 2^14 (signed 16-bit halves, 32-bit wrapping sum).  Written as
// the three-address VOP3P form with the rounding constant in a scalar register: the
// builtin selects the two-address accumulate form, which needs a register copy of the
// constant before every product.
__device__ __forceinline__ int dot(u32 a, u32 b)
{
#if KFA_Q15_DOT_ASM
    int r;
    asm("v_dot2_i32_i16 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(1 << 14));
    return r;
#else
    return __builtin_amdgcn_sdot2(as_v(a), as_v(b), 1 << 14, false);
#endif
}

// The same product with the rounding constant in a vector register and b in a scalar or a
// vector register: a twiddle read with a scalar load (a row that is the same on every
// lane) then feeds the product directly instead of through a register copy (one scalar
// operand per instruction).
__device__ __forceinline__ int dot_tw(u32 a, u32 b)
{
#if KFA_Q15_DOT_ASM
    int rc = 1 << 14;
    asm("" : "+v"(rc));
    int r;
    asm("v_dot2_i32_i16 %0, %1, %2, %3" : "=v"(r) : "v"(a), "r"(b), "v"(rc));
    return r;
#else
    return __builtin_amdgcn_sdot2(as_v(a), as_v(b), 1 << 14, false);
#endif
}

// complex product with a pre-arranged twiddle pair (w.r, -w.i), (w.i, w.r)
__device__ __forceinline__ u32 cmul(u32 a, u32 wr, u32 wi) { return round_pack(dot_tw(a, wr), dot_tw(a, wi)); }

// both components times the scalar s: pass (s, 0) and (0, s)
__device__ __forceinline__ u32 scale(u32 a, u32 s_lo, u32 s_hi) { return round_pack(dot(a, s_lo), dot(a, s_hi)); }

struct Scal {
    u32 lo, hi;  // (s, 0) and (0, s)
};
__device__ __forceinline__ Scal scal(int s) { return Scal{(u32)s & 0xffffu, (u32)s << 16}; }
__device__ __forceinline__ u32 scale(u32 a, Scal s) { return scale(a, s.lo, s.hi); }

// -i * z = (z.i, -z.r) and +i * z = (-z.i, z.r): one packed multiply by (1, -1) / (-1, 1)
// whose source halves are exchanged by its operand selection (written as a vector shuffle,
// which the compiler folds into op_sel of the multiply or of the add that uses it)
__device__ __forceinline__ u32 mul_mi(u32 z)
{
    const v2s zz = as_v(z);
    return as_u(__builtin_shufflevector(zz, zz, 1, 0) * (v2s){1, -1});
}
__device__ __forceinline__ u32 mul_pi(u32 z)
{
    const v2s zz = as_v(z);
    return as_u(__builtin_shufflevector(zz, zz, 1, 0) * (v2s){-1, 1});
}

__device__ __forceinline__ int lo16(u32 w) { return (int)(short)(w & 0xffffu); }

// a.lo*b.lo + a.hi*b.hi + 2^15 (the dot product above with twice its rounding constant)
__device__ __forceinline__ int dot_h(u32 a, u32 b)
{
#if KFA_Q15_DOT_ASM
    int r;
    asm("v_dot2_i32_i16 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(1 << 15));
    return r;
#else
    return __builtin_amdgcn_sdot2(as_v(a), as_v(b), 1 << 15, false);
#endif
}

// C_FIXDIV(c, p): K = floor(32767 / p), each component sround(c * K) truncated to int16.
// 2K < 32768 for every p >= 2, so the products are taken with 2K and 2^15: the sum is
// 2 * (c*K + 2^14) (no overflow: |c * 2K| < 2^30), whose high 16 bits are exactly the
// rounded, truncated component; one byte permute packs the two high halves.
template <int P>
__device__ __forceinline__ u32 fixdiv(u32 c)
{
    constexpr u32 K2 = 2 * (32767 / P);
    static_assert(K2 < 32768, "2K must fit a signed 16-bit lane");
    const int re = dot_h(c, K2), im = dot_h(c, K2 << 16);
    return __builtin_amdgcn_perm((u32)im, (u32)re, 0x07060302u);  // (re.hi, im.hi)
}

// C_FIXDIV of K values and complex products of K values, each written as all of its dot
// products first, then all of the packing: a packing instruction right behind the dot
// product it reads waits a cycle (an s_nop the compiler inserts), one K - 1 products later
// does not.  Same operations, same bits as fixdiv / cmul one at a time.
// The two multipliers are held in registers for the K products: the dot product reads one
// scalar register (the rounding constant) at most, so a multiplier left to the compiler
// is copied into a vector register before every product.
template <int P, int K>
__device__ __forceinline__ void fixdiv_n(u32 (&x)[K], const u32 *c)
{
    constexpr u32 K2 = 2 * (32767 / P);
    u32 klo = K2, khi = K2 << 16;
    asm("" : "+v"(klo), "+v"(khi));
    int re[K], im[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        re[k] = dot_h(c[k], klo);
        im[k] = dot_h(c[k], khi);
    }
#pragma unroll
    for (int k = 0; k < K; ++k)
        x[k] = __builtin_amdgcn_perm((u32)im[k], (u32)re[k], 0x07060302u);
}

template <int K>
__device__ __forceinline__ void cmul_n(u32 (&s)[K], const u32 *a, const uint2 *tw)
{
    int re[K], im[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        re[k] = dot_tw(a[k], tw[k].x);
        im[k] = dot_tw(a[k], tw[k].y);
    }
#if KFA_Q15_SDWA_PACK
    // round_pack's two shifts, each kind for all K values in a row
#pragma unroll
    for (int k = 0; k < K; ++k)
        asm("v_lshrrev_b32 %0, 15, %1" : "=v"(s[k]) : "v"(re[k]));
#pragma unroll
    for (int k = 0; k < K; ++k)
        asm("v_lshrrev_b32_sdwa %0, 15, %1 dst_sel:WORD_1 dst_unused:UNUSED_PRESERVE src0_sel:DWORD src1_sel:DWORD"
            : "+v"(s[k])
            : "v"(im[k]));
#else
#pragma unroll
    for (int k = 0; k < K; ++k)
        s[k] = round_pack(re[k], im[k]);
#endif
}

// ONE: every twiddle of the butterfly is table entry 0, (32767, 0).  C_MUL by it returns
// its operand for every component in (-2^14, 2^14], and C_FIXDIV by p >= 2 always lands
// there (|sround(c * floor(32767 / p))| <= 16383), so the product is left out: same bits.

// kiss_fft.c:15-36
template <bool ONE = false>
__device__ __forceinline__ void bfly2(u32 *v, const uint2 *tw)
{
    u32 x[2];
    fixdiv_n<2, 2>(x, v);
    const u32 a = x[0];
    u32 t = ONE ? x[1] : cmul(x[1], tw[0].x, tw[0].y);
    v[1] = sub(a, t);
    v[0] = add(a, t);
}

// kiss_fft.c:38-84
template <bool INV, bool ONE = false>
__device__ __forceinline__ void bfly4(u32 *v, const uint2 *tw)
{
    u32 x[4], s[3];
    fixdiv_n<4, 4>(x, v);
    if constexpr (ONE) {
        s[0] = x[1];
        s[1] = x[2];
        s[2] = x[3];
    } else {
        cmul_n<3>(s, x + 1, tw);
    }
    u32 x0 = x[0];
    u32 s0 = s[0], s1 = s[1], s2 = s[2];
    u32 s5 = sub(x0, s1);
    x0 = add(x0, s1);
    u32 s3 = add(s0, s2);
    u32 s4 = sub(s0, s2);
    v[2] = sub(x0, s3);
    v[0] = add(x0, s3);
    const u32 u = mul_mi(s4);  // (s4.i, -s4.r)
    if (INV) {
        v[1] = sub(s5, u);  // (s5.r - s4.i, s5.i + s4.r)
        v[3] = add(s5, u);  // (s5.r + s4.i, s5.i - s4.r)
    } else {
        v[1] = add(s5, u);
        v[3] = sub(s5, u);
    }
}

// kiss_fft.c:86-128; e = (epi3.i, 0), (0, epi3.i)
template <bool ONE = false>
__device__ __forceinline__ void bfly3(u32 *v, const uint2 *tw, Scal e)
{
    u32 x0 = fixdiv<3>(v[0]), x1 = fixdiv<3>(v[1]), x2 = fixdiv<3>(v[2]);
    u32 s1 = ONE ? x1 : cmul(x1, tw[0].x, tw[0].y);
    u32 s2 = ONE ? x2 : cmul(x2, tw[1].x, tw[1].y);
    u32 s3 = add(s1, s2);
    u32 s0 = sub(s1, s2);
    u32 o1 = sub(x0, half(s3));
    s0 = scale(s0, e);
    x0 = add(x0, s3);
    const u32 u = mul_mi(s0);  // (s0.i, -s0.r)
    v[0] = x0;
    v[1] = sub(o1, u);  // (o1.r - s0.i, o1.i + s0.r)
    v[2] = add(o1, u);  // (o1.r + s0.i, o1.i - s0.r)
}

// kiss_fft.c:130-189
struct Y5 {
    Scal yar, yai, ybr, ybi;
};

template <bool ONE = false>
__device__ __forceinline__ void bfly5(u32 *v, const uint2 *tw, const Y5 &y)
{
    const u32 x0 = fixdiv<5>(v[0]);
    u32 c1 = fixdiv<5>(v[1]), c2 = fixdiv<5>(v[2]), c3 = fixdiv<5>(v[3]), c4 = fixdiv<5>(v[4]);
    if constexpr (!ONE) {
        c1 = cmul(c1, tw[0].x, tw[0].y);
        c2 = cmul(c2, tw[1].x, tw[1].y);
        c3 = cmul(c3, tw[2].x, tw[2].y);
        c4 = cmul(c4, tw[3].x, tw[3].y);
    }
    u32 c7 = add(c1, c4), c10 = sub(c1, c4);
    u32 c8 = add(c2, c3), c9 = sub(c2, c3);
    u32 c5 = add(add(x0, scale(c7, y.yar)), scale(c8, y.ybr));
    u32 c6 = mul_mi(add(scale(c10, y.yai), scale(c9, y.ybi)));
    u32 c11 = add(add(x0, scale(c7, y.ybr)), scale(c8, y.yar));
    u32 c12 = mul_pi(sub(scale(c10, y.ybi), scale(c9, y.yai)));
    v[0] = add(x0, add(c7, c8));
    v[1] = sub(c5, c6);
    v[4] = add(c5, c6);
    v[2] = add(c11, c12);
    v[3] = sub(c11, c12);
}

}  // namespace q15p

template <class A>
struct IsPacked {
    static constexpr bool value = false;
};
template <>
struct IsPacked<Q15P> {
    static constexpr bool value = true;
};

template <>
struct Ops<Q15P> {
    typedef Cx<Q15P> C;
    __device__ __forceinline__ static C load(const q15x2 &s) { return C{__builtin_bit_cast(unsigned, s)}; }
    __device__ __forceinline__ static q15x2 pack(C c) { return __builtin_bit_cast(q15x2, c.w); }
    __device__ __forceinline__ static C zero() { return C{0u}; }
};

}  // namespace kfa
