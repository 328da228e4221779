// kfa_tile.hpp -- the tile kernel: compile-time radix schedules on strided lines.
//
// A workgroup owns a tile of TL lines of N points.  Line l = (hi, lo), lo = l % nlo, reads
// element e at in[hi*in_hi + lo*in_lo + e*in_el] and writes it at
// out[hi*out_hi + lo*out_lo + e*out_el].  The tile moves between HBM and LDS in two
// phases of its own, each with the thread mapping that makes wave accesses contiguous:
//   * "column" phases (LCOL / SCOL): consecutive lanes take consecutive lines at the same
//     element -- the layout of the non-last axes of an N-D array (kiss_fftnd.c:172-177
//     walks the same columns) and of the first pass of the four-step split;
//   * "row" phases: consecutive lanes take consecutive elements of one line.
// In between, the stages of the radix schedule RS run LDS -> registers -> LDS exactly as
// in the fast line kernel (FastStage, T threads per line).  Lines sit tile_pitch() values
// apart in LDS so that lanes working on neighbouring lines fall in different banks.
//
// TWID multiplies element e of line l by W^(lo*e), W = exp(-+2*pi*i/tw_n), on the way out:
// the twiddle step between the two passes of the four-step split (Bailey, "FFTs in
// external or hierarchical memory", 1990):  with n = n1 + N1*n2 and k = k2 + N2*k1,
//   X[k2 + N2*k1] = sum_n1 W_N1^(n1*k1) * W_N^(n1*k2) * sum_n2 x[n1 + N1*n2] W_N2^(n2*k2).
// W^a is read as hi[a >> 10] * lo[a & 1023] from two small tables (computed in FP64,
// rounded once); a thread's IT elements share one such product and a per-line table of
// W^(lo*T*k) in LDS.  The float path only: fixed-point long lines keep the reference's own
// stage order (the general line kernel's two-pass split).
#pragma once

#include "kfa_fast.hpp"

namespace kfa {

struct TileArgs {
    const void *in;
    void *out;
    const void *stw;     // stage table of the schedule (the fast kernels' layout)
    const void *tw_hi;   // TWID: W^(q*1024), q <= tw_n/1024
    const void *tw_lo;   // TWID: W^r, r < 1024
    long long nlines, nlo;
    long long in_hi, in_lo, in_el;
    long long out_hi, out_lo, out_el;
    int xcd_group;       // > 1: tiles g*G .. g*G+G-1 (which share cache lines) run on one XCD
};

template <class A, bool INV, int T, class RS, int S>
__device__ __forceinline__ void tile_stage(typename A::store2 *lb, const typename A::store2 *stw, int t)
{
    typedef FastStage<A, INV, T, RS, S, false> F;
    typename F::C v[F::J][F::p];
    F::template gather<false>(v, lb, 1, t, true);
    line_sync<T>();  // every input read before the buffer is overwritten
    F::compute(v, stw, t);
    F::template store<false>(v, lb, 1, t, true);
    line_sync<T>();
}

template <class A, bool INV, int T, class RS, int... S>
__device__ __forceinline__ void tile_stages(std::integer_sequence<int, S...>, typename A::store2 *lb,
                                            const typename A::store2 *stw, int t)
{
    (tile_stage<A, INV, T, RS, S>(lb, stw, t), ...);
}

// Workgroup -> tile.  Workgroups are dealt to the 8 XCDs round-robin; with G tiles sharing
// each cache line, the G tiles of a group go to one XCD back to back, so the line is
// fetched into that XCD's L2 once.
__device__ __forceinline__ long long tile_of_block(long long b, long long ntiles, int G)
{
    const long long full = ntiles / (8LL * G) * (8LL * G);  // complete groups; the tail in order
    if (G <= 1 || b >= full)
        return b;
    const long long x = b % 8, s = b / 8;
    return ((s / G) * 8 + x) * G + s % G;
}

// Line pitch in LDS.  The column phases put lanes ln = 0..TL-1 on neighbouring lines and
// the rest of a wave's lanes on the next elements e0: lane (ln, e0) hits bank pair
// (ln * pitch + e0) mod 32, all different when pitch = N + 32 / TL (mod 32) for TL < 32
// (N + 1, the usual odd pitch, puts lanes (ln, e0) and (ln + 1, e0 - 1) on one bank).
template <int N, int TL>
constexpr int tile_pitch()
{
    return N + (TL < 32 ? 32 / TL : 1);
}

template <class A, bool INV, int T, int TL, bool LCOL, bool SCOL, bool TWID, class RS>
__global__ void __launch_bounds__(T *TL) k_tile(TileArgs a)
{
    typedef typename A::store2 St;
    typedef Ops<A> O;
    typedef Cx<A> C;
    constexpr int N = RS::N();
    constexpr int NS = tile_pitch<N, TL>();
    constexpr int IT = N / T;  // elements per thread in the HBM phases
    static_assert(N % T == 0, "threads per line must divide the length");
    static_assert(!TWID || SCOL, "the four-step twiddle is applied on column stores");
    extern __shared__ __align__(16) unsigned char kfa_tile_smem[];
    St *buf = reinterpret_cast<St *>(kfa_tile_smem);
    // TWID: per line, W^(lo*T*k) for k < IT; a thread's element e = e0 + k*T then takes
    // W^(lo*e) = W^(lo*e0) * W^(lo*T*k) (one table product per thread, one per line and k)
    __shared__ St s_step[TWID ? TL * IT : 1];

    const int tid = threadIdx.x;
    const long long ntiles = (a.nlines + TL - 1) / TL;
    const long long l0 = tile_of_block(blockIdx.x, ntiles, a.xcd_group) * TL;
    const int nl = (int)((a.nlines - l0) < TL ? (a.nlines - l0) : TL);

    // HBM phases: each thread owns one line and IT of its elements.  Column mapping:
    // lanes take consecutive lines (line = tid % TL), element e0 + k*T; row mapping:
    // T lanes take consecutive elements of one line (line = tid / T).
    auto line_of = [&](bool col) { return col ? tid % TL : tid / T; };
    auto elem0 = [&](bool col) { return col ? tid / TL : tid % T; };

    {
        const int ln = line_of(LCOL), e0 = elem0(LCOL);
        const bool ok = ln < nl;
        const long long gl = l0 + (ok ? ln : 0);
        const long long hi = gl / a.nlo, lo = gl - hi * a.nlo;
        const St *src = reinterpret_cast<const St *>(a.in) + hi * a.in_hi + lo * a.in_lo + e0 * a.in_el;
        const long long step = (long long)T * a.in_el;
        St v[IT];
#pragma unroll
        for (int k = 0; k < IT; ++k)
            if (ok)
                v[k] = src[k * step];
#pragma unroll
        for (int k = 0; k < IT; ++k)
            if (ok)
                buf[ln * NS + e0 + k * T] = v[k];
    }
    if constexpr (TWID) {
        const St *whi = reinterpret_cast<const St *>(a.tw_hi);
        const St *wlo = reinterpret_cast<const St *>(a.tw_lo);
        for (int w = tid; w < TL * IT; w += T * TL) {
            const int ln = w / IT, k = w - (w / IT) * IT;
            const long long gl = l0 + (ln < nl ? ln : 0);
            const unsigned long long lo = (unsigned long long)(gl % a.nlo);
            const unsigned long long ex = lo * (unsigned)(k * T);  // < tw_n
            s_step[w] = O::pack(O::cmul(O::load(whi[ex >> 10]), O::load(wlo[ex & 1023])));
        }
    }
    __syncthreads();

    {
        const int line = tid / T, t = tid - line * T;
        tile_stages<A, INV, T, RS>(std::make_integer_sequence<int, RS::K>{}, buf + line * NS,
                                   reinterpret_cast<const St *>(a.stw), t);
    }
    if constexpr (64 % T == 0)
        __syncthreads();  // every line done before the store phase reads across lines

    {
        const int ln = line_of(SCOL), e0 = elem0(SCOL);
        if (ln >= nl)
            return;
        const long long gl = l0 + ln;
        const long long hi = gl / a.nlo, lo = gl - hi * a.nlo;
        St *dst = reinterpret_cast<St *>(a.out) + hi * a.out_hi + lo * a.out_lo + e0 * a.out_el;
        const long long step = (long long)T * a.out_el;
        C v[IT];
#pragma unroll
        for (int k = 0; k < IT; ++k)
            v[k] = O::load(buf[ln * NS + e0 + k * T]);
        if constexpr (TWID) {
            const St *whi = reinterpret_cast<const St *>(a.tw_hi);
            const St *wlo = reinterpret_cast<const St *>(a.tw_lo);
            const unsigned long long ex = (unsigned long long)lo * (unsigned)e0;  // < tw_n
            const C base = O::cmul(O::load(whi[ex >> 10]), O::load(wlo[ex & 1023]));
#pragma unroll
            for (int k = 0; k < IT; ++k)
                v[k] = O::cmul(v[k], O::cmul(base, O::load(s_step[ln * IT + k])));
        }
#pragma unroll
        for (int k = 0; k < IT; ++k)
            dst[k * step] = O::pack(v[k]);
    }
}

}  // namespace kfa
