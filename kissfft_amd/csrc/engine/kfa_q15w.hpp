// kfa_q15w.hpp -- Q15 lines held by one wavefront: the stage exchanges are lane swaps.
//
// The radix-2 / radix-4 butterflies of kf_bfly2 / kf_bfly4 (kiss_fft.c:15-84, packed forms
// of kfa_q15p.hpp, the reference's rounding) on a power-of-two line of N = T * R points held
// by T lanes of one wavefront, R packed Q15 words per lane.  A point's index is spread over
// the lane id (log2 T bits) and the register index (log2 R bits).  A radix-2^c stage needs
// the c index bits of its butterfly inputs in register bits (the butterfly's inputs in one
// lane); Stockham's data movement between stages is only a relabelling of index bits, so
// the one exchange a stage may need is a swap of a lane bit with a register bit:
//   lane bit 0..3   DPP row shifts by 2^k (v_mov_dpp row_shl / row_shr) and a select,
//                   one pair of instructions per register pair
//   lane bit 4      v_permlane16_swap (rows of 16 lanes exchanged between two registers)
//   lane bit 5      v_permlane32_swap
// There is no LDS traffic and no barrier: lines never leave their wavefront.  A bit that
// leaves the registers is always one the remaining stages never butterfly (a bit above
// the stage's butterfly bits), and a compile-time plan (WavePlan) picks the swaps and
// tracks which index bit every lane and register bit holds.  Loads read 16-byte-aligned
// runs of consecutive points per register; the final swaps put the output's low index bits
// on lanes, so every store of a register covers T consecutive outputs.
//
// Fixed point keeps every rounding of the reference: the butterflies are the reference's
// operation order with the reference's twiddle table entries (the stage table), only the
// positions of the values move.
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>

#include <utility>
#endif

#include "kfa_fast.hpp"
#include "kfa_q15p.hpp"

namespace kfa {

constexpr int ilog2(int v)
{
    int b = 0;
    while ((1 << b) < v)
        ++b;
    return b;
}

// The swap plan of schedule RS on lines of T lanes.
template <class RS, int T>
struct WavePlan {
    static constexpr int N = RS::N();
    static constexpr int K = RS::K;
    static constexpr int NBITS = ilog2(N);
    static constexpr int TB = ilog2(T);
    static constexpr int RB = NBITS - TB;
    static constexpr int R = 1 << RB;
    static_assert((1 << NBITS) == N && (1 << TB) == T, "power-of-two lines and lanes");
    static_assert(RB >= 2, "a radix-4 stage needs two register bits");

    struct Data {
        int nswap[K + 1];       // swaps before stage s (s == K: after the last stage)
        int swap_lane[K + 1][4];
        int swap_reg[K + 1][4];
        int lane_lab[K + 1][6];  // index bit held by lane bit i when stage s computes
        int reg_lab[K + 1][8];   // ... and by register bit j (s == K: output index bits)
    };

    static constexpr Data make()
    {
        Data d{};
        int lane[6] = {0, 0, 0, 0, 0, 0}, reg[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int i = 0; i < TB; ++i)
            lane[i] = i;  // loads: lane bits are the low index bits
        for (int j = 0; j < RB; ++j)
            reg[j] = TB + j;
        for (int s = 0; s < K; ++s) {
            const int c = ilog2(RS::p(s));
            const int a = ilog2(N / (RS::p(s) * RS::m(s)));  // log2 of the stage's stride P
            int ns = 0;
            for (int b = a; b < a + c; ++b) {
                for (int i = 0; i < TB; ++i) {
                    if (lane[i] != b)
                        continue;
                    // out goes the register bit with the lowest label above the butterfly
                    // bits (labels above a + c are never butterflied again)
                    int best = -1;
                    for (int j = 0; j < RB; ++j)
                        if (reg[j] >= a + c && (best < 0 || reg[j] < reg[best]))
                            best = j;
                    d.swap_lane[s][ns] = i;
                    d.swap_reg[s][ns] = best;
                    ++ns;
                    const int t = lane[i];
                    lane[i] = reg[best];
                    reg[best] = t;
                }
            }
            d.nswap[s] = ns;
            for (int i = 0; i < TB; ++i)
                d.lane_lab[s][i] = lane[i];
            for (int j = 0; j < RB; ++j)
                d.reg_lab[s][j] = reg[j];
            // Stockham relabelling: butterfly bits go to the top, the bits above shift down
            auto rel = [&](int x) { return x < a ? x : (x < a + c ? NBITS - c + (x - a) : x - c); };
            for (int i = 0; i < TB; ++i)
                lane[i] = rel(lane[i]);
            for (int j = 0; j < RB; ++j)
                reg[j] = rel(reg[j]);
        }
        // output: the low TB index bits on lanes
        int ns = 0;
        for (int i = 0; i < TB; ++i) {
            if (lane[i] < TB)
                continue;
            int best = -1;
            for (int j = 0; j < RB; ++j)
                if (reg[j] < TB && best < 0)
                    best = j;
            d.swap_lane[K][ns] = i;
            d.swap_reg[K][ns] = best;
            ++ns;
            const int t = lane[i];
            lane[i] = reg[best];
            reg[best] = t;
        }
        d.nswap[K] = ns;
        for (int i = 0; i < TB; ++i)
            d.lane_lab[K][i] = lane[i];
        for (int j = 0; j < RB; ++j)
            d.reg_lab[K][j] = reg[j];
        return d;
    }
    static constexpr Data d = make();

    static constexpr bool valid()
    {
        for (int s = 0; s <= K; ++s)
            if (d.nswap[s] > 4)
                return false;
        for (int s = 0; s <= K; ++s)
            for (int k = 0; k < d.nswap[s]; ++k)
                if (d.swap_reg[s][k] < 0)
                    return false;
        return true;
    }
    static_assert(valid(), "no swap plan for this schedule");
};

namespace q15w {

typedef unsigned u32;

// Swaps lane bit L with register bit J across the R registers of a lane.
template <int L, int J, int R>
__device__ __forceinline__ void swap_bit(u32 (&v)[R], int lane)
{
#pragma unroll
    for (int x = 0; x < R; ++x) {
        if (x & (1 << J))
            continue;
        u32 &A = v[x];
        u32 &B = v[x | (1 << J)];
        if constexpr (L == 4) {
            // rows 1 / 3 of A trade places with rows 0 / 2 of B
            const auto r = __builtin_amdgcn_permlane16_swap(A, B, false, false);
            A = r[0];
            B = r[1];
        } else if constexpr (L == 5) {
            const auto r = __builtin_amdgcn_permlane32_swap(A, B, false, false);
            A = r[0];
            B = r[1];
        } else {
            // lanes with bit L set take B of lane - 2^L (row_shr), the others A of lane + 2^L
            // (row_shl); both partners sit in the lane's own row of 16: a DPP move and a
            // select per new value (the compiler keeps them apart: the select's lane mask
            // is a scalar register pair, which the DPP form of v_cndmask cannot take).
            const bool hi = (lane >> L) & 1;
            const u32 fromB = (u32)__builtin_amdgcn_mov_dpp((int)B, 0x110 + (1 << L), 0xf, 0xf, true);
            const u32 fromA = (u32)__builtin_amdgcn_mov_dpp((int)A, 0x100 + (1 << L), 0xf, 0xf, true);
            const u32 na = !hi ? A : fromB;
            const u32 nb = hi ? B : fromA;
            A = na;
            B = nb;
        }
    }
}

template <class PL, int S, int K, int R>
__device__ __forceinline__ void swaps(u32 (&v)[R], int lane)
{
    if constexpr (K < PL::d.nswap[S]) {
        swap_bit<PL::d.swap_lane[S][K], PL::d.swap_reg[S][K], R>(v, lane);
        swaps<PL, S, K + 1, R>(v, lane);
    }
}

// value of the index bits that register index x contributes to a field starting at bit
// `lo` of width `w` (stage labels of stage S)
template <class PL, int S>
__device__ __forceinline__ constexpr int reg_field(int x, int lo, int w)
{
    int f = 0;
    for (int j = 0; j < PL::RB; ++j) {
        const int lab = PL::d.reg_lab[S][j];
        if (lab >= lo && lab < lo + w && ((x >> j) & 1))
            f |= 1 << (lab - lo);
    }
    return f;
}

template <class PL, int S>
__device__ __forceinline__ int lane_field(int lane, int lo, int w)
{
    int f = 0;
#pragma unroll
    for (int i = 0; i < PL::TB; ++i) {
        const int lab = PL::d.lane_lab[S][i];
        if (lab >= lo && lab < lo + w)
            f |= ((lane >> i) & 1) << (lab - lo);
    }
    return f;
}

// Stage S: swaps, then every butterfly of the lane (its inputs are the registers whose
// butterfly bits vary; outputs replace them in place), twiddle rows u from the stage table.
template <bool INV, class RS, class PL, int S>
__device__ __forceinline__ void stage(u32 (&v)[PL::R], const uint2 *__restrict__ tab, int lane)
{
    constexpr int R = PL::R;
    swaps<PL, S, 0, R>(v, lane);
    constexpr int p = RS::p(S);
    constexpr int c = ilog2(p);
    constexpr int m = RS::m(S);
    constexpr int a = ilog2(RS::N() / (p * m));
    constexpr int PR = RS::PR(S);
    constexpr int TWO = RS::tw_off(S);
    // register bit positions of the butterfly bits a, a+1
    constexpr int rb0 = [] {
        for (int j = 0; j < PL::RB; ++j)
            if (PL::d.reg_lab[S][j] == a)
                return j;
        return -1;
    }();
    constexpr int rb1 = [] {
        for (int j = 0; j < PL::RB; ++j)
            if (PL::d.reg_lab[S][j] == a + 1)
                return j;
        return -1;
    }();
    static_assert(rb0 >= 0 && (c == 1 || rb1 >= 0), "butterfly bits not in registers");
    constexpr int mask = (1 << rb0) | (c == 2 ? (1 << rb1) : 0);
    // the twiddle row u is index bits a + c and up; when no lane bit holds one of them, the
    // butterflies whose register bits of u are all clear use row 0 only: their twiddles are
    // table entry 0 and the products are left out (q15p::bfly2 / bfly4, ONE)
    constexpr bool u_on_lanes = [] {
        for (int i = 0; i < PL::TB; ++i)
            if (PL::d.lane_lab[S][i] >= a + c)
                return true;
        return false;
    }();
    const int u_lane = lane_field<PL, S>(lane, a + c, PL::NBITS - a - c);
#pragma unroll
    for (int x = 0; x < R; ++x) {
        if (x & mask)
            continue;
        const int u_reg = reg_field<PL, S>(x, a + c, PL::NBITS - a - c);
        const bool one = !u_on_lanes && u_reg == 0;  // a constant once the loop is unrolled
        const unsigned u = (unsigned)u_lane + (unsigned)u_reg;
        const uint2 *trow = tab + (TWO + u * PR);
        if constexpr (c == 1) {
            u32 y[2] = {v[x], v[x | (1 << rb0)]};
            if (one) {
                q15p::bfly2<true>(y, nullptr);
            } else {
                uint2 w[1] = {trow[0]};
                q15p::bfly2(y, w);
            }
            v[x] = y[0];
            v[x | (1 << rb0)] = y[1];
        } else {
            // input r = r0 + 2 r1 sits at register x | r0 << rb0 | r1 << rb1
            constexpr int o1 = 1 << rb0, o2 = 1 << rb1;
            u32 y[4] = {v[x], v[x | o1], v[x | o2], v[x | o1 | o2]};
            if (one) {
                q15p::bfly4<INV, true>(y, nullptr);
            } else {
                uint2 w[3] = {trow[0], trow[1], trow[2]};
                q15p::bfly4<INV>(y, w);
            }
            v[x] = y[0];
            v[x | o1] = y[1];
            v[x | o2] = y[2];
            v[x | o1 | o2] = y[3];
        }
    }
}

template <bool INV, class RS, class PL, int... S>
__device__ __forceinline__ void stages(std::integer_sequence<int, S...>, u32 (&v)[PL::R], const uint2 *tab, int lane)
{
    (stage<INV, RS, PL, S>(v, tab, lane), ...);
}

}  // namespace q15w

// The plan of 16-lane lines that exchange through LDS once instead of swapping lane bits.
// Loads put index bits 0-3 on the lanes (coalesced rows), so the stages that butterfly
// those bits would need four lane swaps (four DPP exchanges of every register).  Instead,
// at the first stage whose butterfly bits are not all in registers, the line is transposed
// through LDS: the four lane labels go to register bits 0-3, the four largest register
// labels (bits already butterflied: labels at or above a + c never are again) go to the
// lanes, the other register labels stay in register bits 4 and up.  The remaining stages
// must then find their butterfly bits in registers (checked at compile time).
template <class RS, int T>
struct XposePlan {
    static constexpr int N = RS::N();
    static constexpr int K = RS::K;
    static constexpr int NBITS = ilog2(N);
    static constexpr int TB = ilog2(T);
    static constexpr int RB = NBITS - TB;
    static constexpr int R = 1 << RB;
    static_assert(T == 16, "the LDS transpose exchanges 16 x 16 blocks");
    static_assert((1 << NBITS) == N && RB >= 4 && RB <= 5, "256- or 512-point lines");

    struct Data {
        int nswap[K + 1];  // always 0: the stage code of k_q15w runs unchanged
        int swap_lane[K + 1][4];
        int swap_reg[K + 1][4];
        int lane_lab[K + 1][6];
        int reg_lab[K + 1][8];
        int xs;            // the transpose runs before stage xs (K: never)
        int pre_lane[4];   // labels on lane bits / register bits just before it
        int pre_reg[8];
        bool ok;
    };

    static constexpr Data make()
    {
        Data d{};
        d.ok = true;
        d.xs = K;
        int lane[6] = {0, 1, 2, 3, 0, 0}, reg[8] = {};
        for (int j = 0; j < RB; ++j)
            reg[j] = TB + j;
        for (int s = 0; s < K; ++s) {
            const int c = ilog2(RS::p(s));
            const int a = ilog2(N / (RS::p(s) * RS::m(s)));
            bool on_lane = false;
            for (int i = 0; i < TB; ++i)
                if (lane[i] >= a && lane[i] < a + c)
                    on_lane = true;
            if (on_lane) {
                if (d.xs != K)
                    d.ok = false;  // a second transpose would be needed
                d.xs = s;
                for (int i = 0; i < TB; ++i)
                    d.pre_lane[i] = lane[i];
                for (int j = 0; j < RB; ++j)
                    d.pre_reg[j] = reg[j];
                // lanes take the four largest register labels, which must be done
                int big[8] = {}, nb = 0;
                for (int j = 0; j < RB; ++j)
                    big[nb++] = reg[j];
                for (int x = 0; x < nb; ++x)  // descending
                    for (int y = x + 1; y < nb; ++y)
                        if (big[y] > big[x]) {
                            const int t = big[x];
                            big[x] = big[y];
                            big[y] = t;
                        }
                int nl[4] = {big[3], big[2], big[1], big[0]};  // ascending on lane bits 0-3
                for (int i = 0; i < 4; ++i)
                    if (nl[i] < a + c)
                        d.ok = false;
                int nr[8] = {};
                for (int i = 0; i < TB; ++i)
                    nr[i] = lane[i];
                int k = TB;
                for (int x = nb - 1; x >= 4; --x)  // the rest ascending on bits 4 and up
                    nr[k++] = big[x];
                for (int i = 0; i < TB; ++i)
                    lane[i] = nl[i];
                for (int j = 0; j < RB; ++j)
                    reg[j] = nr[j];
            }
            for (int i = 0; i < TB; ++i)
                if (lane[i] >= a && lane[i] < a + c)
                    d.ok = false;
            d.nswap[s] = 0;
            for (int i = 0; i < TB; ++i)
                d.lane_lab[s][i] = lane[i];
            for (int j = 0; j < RB; ++j)
                d.reg_lab[s][j] = reg[j];
            auto rel = [&](int x) { return x < a ? x : (x < a + c ? NBITS - c + (x - a) : x - c); };
            for (int i = 0; i < TB; ++i)
                lane[i] = rel(lane[i]);
            for (int j = 0; j < RB; ++j)
                reg[j] = rel(reg[j]);
        }
        d.nswap[K] = 0;
        for (int i = 0; i < TB; ++i)
            d.lane_lab[K][i] = lane[i];
        for (int j = 0; j < RB; ++j)
            d.reg_lab[K][j] = reg[j];
        if (d.xs == K)
            d.ok = false;  // nothing to transpose: the plain lane-swap kernel covers it
        return d;
    }
    static constexpr Data d = make();
    static_assert(d.ok, "no single-transpose plan for this schedule");
};

namespace q15w {

// LDS word of a transposed element within its line's block: row j (the four labels that go
// to the lanes), column k (the four labels that go to registers 0-3).  Columns are taken
// in 16-byte groups g = k >> 2 swizzled by t(j) = (j >> 1 ^ j >> 2) & 3: writes (16
// lanes over k, one j) cover one row, reads (lanes over j, four consecutive k per
// ds_read_b128) hit distinct 16-byte bank groups for 32 or 64 banks of 4 bytes.
__device__ __forceinline__ constexpr int xpose_t(int j) { return ((j >> 1) ^ (j >> 2)) & 3; }
constexpr int kXposeStride = 272;  // words per line: 256 + 16, so the four lines of a
                                   // wave start in different 16-bank groups

// The transpose of one lane's registers v (labels pre_lane / pre_reg) into w (labels
// lane_lab / reg_lab of stage xs), through blk (this line's kXposeStride words).  Chunks
// of the register labels that stay in registers (bit 4 of the register index, if any)
// go one after the other through the same words; the line's lanes are one wavefront's, so
// wavefront-scope ordering is all it needs (LDS operations of a wave run in order).
template <class PL>
__device__ __forceinline__ void xpose(const u32 (&v)[PL::R], u32 (&w)[PL::R], u32 *blk, int lane)
{
    constexpr int R = PL::R;
    constexpr int NC = R / 16;
    // register bit of each pre-transpose register label that goes to lane bit q / stays
    constexpr auto pos = [] {
        struct P {
            int jpos[4];
            int cpos[4];
        } p{};
        for (int q = 0; q < 4; ++q)
            for (int j = 0; j < PL::RB; ++j)
                if (PL::d.pre_reg[j] == PL::d.lane_lab[PL::d.xs][q])
                    p.jpos[q] = j;
        for (int r = 4; r < PL::RB; ++r)
            for (int j = 0; j < PL::RB; ++j)
                if (PL::d.pre_reg[j] == PL::d.reg_lab[PL::d.xs][r])
                    p.cpos[r - 4] = j;
        return p;
    }();
    // write: k = lane (pre lane bits, in order, are the new register bits 0-3)
    const int kg = lane >> 2, kw = lane & 3;
    u32 *wa[4];
#pragma unroll
    for (int t = 0; t < 4; ++t)
        wa[t] = blk + 4 * (kg ^ t) + kw;
    // read: j = lane (new lane bit q holds pre register bit jpos[q])
    const int tr = xpose_t(lane);
    const u32 *ra[4];
#pragma unroll
    for (int g = 0; g < 4; ++g)
        ra[g] = blk + 16 * lane + 4 * (g ^ tr);
#pragma unroll
    for (int c = 0; c < NC; ++c) {
#pragma unroll
        for (int x = 0; x < R; ++x) {
            int cx = 0, j = 0;
            for (int r = 4; r < PL::RB; ++r)
                cx |= ((x >> pos.cpos[r - 4]) & 1) << (r - 4);
            if (cx != c)
                continue;
            for (int q = 0; q < 4; ++q)
                j |= ((x >> pos.jpos[q]) & 1) << q;
            wa[xpose_t(j)][16 * j] = v[x];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const uint4 q = *reinterpret_cast<const uint4 *>(__builtin_assume_aligned(ra[g], 16));
            w[16 * c + 4 * g + 0] = q.x;
            w[16 * c + 4 * g + 1] = q.y;
            w[16 * c + 4 * g + 2] = q.z;
            w[16 * c + 4 * g + 3] = q.w;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

template <bool INV, class RS, class PL, int... S>
__device__ __forceinline__ void xstages(std::integer_sequence<int, S...>, u32 (&v)[PL::R], const uint2 *tab,
                                        int lane, u32 *blk)
{
    auto one = [&](auto sc) {
        constexpr int s = decltype(sc)::value;
        if constexpr (s == PL::d.xs) {
            u32 w[PL::R];
            xpose<PL>(v, w, blk, lane);
#pragma unroll
            for (int x = 0; x < PL::R; ++x)
                v[x] = w[x];
        }
        stage<INV, RS, PL, s>(v, tab, lane);
    };
    (one(std::integral_constant<int, S>{}), ...);
}

}  // namespace q15w

// Q15 c2c lines of the power-of-two schedule RS held by 16 lanes each, LPB lines per
// workgroup, with the one LDS transpose of XposePlan in place of lane swaps: the body for
// workgroup `blk` of its launch (k_q15x below, or a range of a fused launch).
template <bool INV, int LPB, class RS>
__device__ __forceinline__ void k_q15x_body(const FastArgs &a, long long blk)
{
    constexpr int T = 16;
    typedef XposePlan<RS, T> PL;
    typedef q15w::u32 u32;
    constexpr int R = PL::R;
    __shared__ __align__(16) u32 lds[LPB * q15w::kXposeStride];
    const int line = threadIdx.x / T;
    const int lane = threadIdx.x - line * T;
    const long long gl = blk * LPB + line;
    const bool valid = gl < a.nlines;
    const q15x2 *gin;
    q15x2 *gout;
    line_base<q15x2>(a, valid ? gl : a.nlines - 1, gin, gout);
    const uint2 *tab = reinterpret_cast<const uint2 *>(__builtin_assume_aligned(a.stw, 16));
    const long long ie = a.in_el, oe = a.out_el;
    const bool unit = ie == 1 && oe == 1;
    u32 v[R];
    if (unit) {
#pragma unroll
        for (int x = 0; x < R; ++x)
            v[x] = __builtin_bit_cast(u32, hbm_ld(gin + lane + T * x));
    } else {
#pragma unroll
        for (int x = 0; x < R; ++x)
            v[x] = __builtin_bit_cast(u32, hbm_ld(gin + (long long)(lane + T * x) * ie));
    }
    q15w::xstages<INV, RS, PL>(std::make_integer_sequence<int, RS::K>{}, v, tab, lane,
                               lds + line * q15w::kXposeStride);
    if (!valid)
        return;
    const int y_lane = q15w::lane_field<PL, RS::K>(lane, 0, PL::NBITS);
    if (unit) {
#pragma unroll
        for (int x = 0; x < R; ++x)
            hbm_st(gout + y_lane + q15w::reg_field<PL, RS::K>(x, 0, PL::NBITS), __builtin_bit_cast(q15x2, v[x]));
    } else {
#pragma unroll
        for (int x = 0; x < R; ++x)
            hbm_st(gout + (long long)(y_lane + q15w::reg_field<PL, RS::K>(x, 0, PL::NBITS)) * oe,
                   __builtin_bit_cast(q15x2, v[x]));
    }
}

// WPE: the wavefronts per SIMD the register allocation must leave room for.
template <bool INV, int LPB, class RS, int WPE>
__global__ void __launch_bounds__(16 * LPB, WPE) k_q15x(FastArgs a)
{
    k_q15x_body<INV, LPB, RS>(a, blockIdx.x);
}

// Q15 c2c lines of the power-of-two schedule RS, T lanes per line, LPB lines per workgroup.
template <bool INV, int T, int LPB, class RS>
__global__ void __launch_bounds__(T *LPB) k_q15w(FastArgs a)
{
    typedef WavePlan<RS, T> PL;
    typedef q15w::u32 u32;
    constexpr int R = PL::R;
    static_assert(64 % T == 0 || T % 64 == 0, "lines must not straddle wavefronts");
    static_assert(T <= 64, "one wavefront per line at most");
    const int line = threadIdx.x / T;
    const int lane = threadIdx.x - line * T;
    const long long gl = (long long)blockIdx.x * LPB + line;
    const bool valid = gl < a.nlines;
    const q15x2 *gin;
    q15x2 *gout;
    // lines past the end compute on the last line's data and store nothing
    line_base<q15x2>(a, valid ? gl : a.nlines - 1, gin, gout);
    const uint2 *tab = reinterpret_cast<const uint2 *>(__builtin_assume_aligned(a.stw, 16));
    const long long ie = a.in_el, oe = a.out_el;
    const bool unit = ie == 1 && oe == 1;
    u32 v[R];
    // index x = lane + T * register
    if (unit) {
#pragma unroll
        for (int x = 0; x < R; ++x)
            v[x] = __builtin_bit_cast(u32, hbm_ld(gin + lane + T * x));
    } else {
#pragma unroll
        for (int x = 0; x < R; ++x)
            v[x] = __builtin_bit_cast(u32, hbm_ld(gin + (long long)(lane + T * x) * ie));
    }
    q15w::stages<INV, RS, PL>(std::make_integer_sequence<int, RS::K>{}, v, tab, lane);
    q15w::swaps<PL, RS::K, 0, R>(v, lane);
    if (!valid)
        return;
    const int y_lane = q15w::lane_field<PL, RS::K>(lane, 0, PL::NBITS);
    if (unit) {
#pragma unroll
        for (int x = 0; x < R; ++x)
            hbm_st(gout + y_lane + q15w::reg_field<PL, RS::K>(x, 0, PL::NBITS),
                   __builtin_bit_cast(q15x2, v[x]));
    } else {
#pragma unroll
        for (int x = 0; x < R; ++x)
            hbm_st(gout + (long long)(y_lane + q15w::reg_field<PL, RS::K>(x, 0, PL::NBITS)) * oe,
                   __builtin_bit_cast(q15x2, v[x]));
    }
}

}  // namespace kfa
