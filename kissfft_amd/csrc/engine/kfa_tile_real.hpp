// kfa_tile_real.hpp -- the real split / merge fused into the two passes of the four-step
// split (float), for real transforms whose half length nc = N1*N2 is longer than one LDS
// line.
//
// Forward (kiss_fftr.c:88-116).  The half-length transform Z = FFT(z), z[n] = x[2n] +
// i x[2n+1], leaves the second four-step pass row by row: the transform of row k2 is
// Z[k2 + N2*k1], k1 < N1.  The split combines Z[k] with Z[nc-k], and
//     nc - (q + N2*k1) = (N2-q) + N2*(N1-1-k1)      (0 < q < N2),
// so the partner of row q is row N2-q read backwards.  k_tile_rsplit gives a workgroup
// pairs of rows (q, N2-q); pair 0 holds rows 0 and N2/2, which pair with themselves
// (row 0: k1 <-> N1-k1; row N2/2: k1 <-> N1-1-k1).  After the row transforms every pair
// (k, nc-k) is in the workgroup's LDS and the bins X[k] go straight to HBM: two HBM
// passes instead of three.
//
// Inverse (kiss_fftr.c:131-153).  The merge z[k] = f(X[k], X[nc-k]) feeds the first
// four-step pass, which transforms column n1 of z (elements k = n1 + N1*n2).  With
//     nc - (p + N1*n2) = (N1-p) + N1*(N2-1-n2)      (0 < p < N1),
// the partner of column p is column N1-p read backwards; pair 0 holds columns 0 and N1/2.
// k_tile_rmerge loads X by column pairs, merges in LDS, transforms the columns and
// applies the four-step twiddle W_nc^(n1*k2) on the way out, as the column pass of the
// complex four-step does.
#pragma once

#include "kfa_lines.hpp"
#include "kfa_tile.hpp"

namespace kfa {

// The split / merge twiddle of bin k (kiss_fftr.c:48-53), exp(-+i*pi*(k/nc + 1/2)), is
// read as the product of two small tables: with k = a + L*e (a < L = `other`, e < the
// line length n),
//     exp(-+i*pi*(k/nc + 1/2)) = exp(-+i*pi*(a/nc + 1/2)) * exp(-+i*pi*e/n),
// rtw_a[a] and rtw_b[e] (FP64, rounded once).  The split's bins are k = q + N2*k1 (a = q,
// e = k1), the merge's k = p + N1*n2 (a = p, e = n2).
struct RealTileArgs {
    const void *in;
    void *out;
    const void *stw;    // stage table of the line schedule
    const void *rtw_a;  // `other` entries
    const void *rtw_b;  // n entries
    const void *tw_hi;  // merge: four-step twiddles W^(q*1024), W^r (as TileArgs)
    const void *tw_lo;
    long long nc;       // half length N1*N2
    long long in_hi;    // distance between transforms, input side (complex values)
    long long out_hi;   // ... output side
    long long tiles;    // workgroups of the launch
    int other;          // split: N2 (the row count); merge: N1 (the column count)
    int xcd_group;      // > 1: that many neighbouring tiles (sharing cache lines) per XCD
};

// Line m (0 or 1) of pair j among L lines: pair 0 is (0, L/2), pair j > 0 is (j, L-j).
__device__ __forceinline__ int pair_line(int j, int m, int L)
{
    return j == 0 ? (m ? L / 2 : 0) : (m ? L - j : j);
}

// The bin array of transform b starts `e` complex values into a 128-byte line (rows of
// nc+1 bins drift by one value per transform).  Tiles take their pairs from index r on,
// r = -e mod 16, so that the run of lower bins of an even tile starts on a line: tile t
// holds pairs (t*NP + r + j) mod npairs (npairs >= 16).
__device__ __forceinline__ int bin_rotation(const void *bins, long long b, long long hi)
{
    const long long e = (long long)(reinterpret_cast<uintptr_t>(bins) / 8) + b * hi;
    return (int)(-e & 15);
}

__device__ __forceinline__ int rot_pair(int j0, int j, int npairs)
{
    const int J = j0 + j;
    return J >= npairs ? J - npairs : J;
}

// Forward: first-pass output rows -> bins.  Row length N1 = RS::N(); TL/2 row pairs per
// workgroup; N2/2 pairs per transform (TL/2 divides N2/2, checked by the launcher).
template <int T, int TL, class RS>
__global__ void __launch_bounds__(T *TL) k_tile_rsplit(RealTileArgs a)
{
    typedef F32::store2 St;
    typedef Ops<F32> O;
    typedef Cx<F32> C;
    constexpr int N1 = RS::N();
    constexpr int NS = N1 + 1;
    constexpr int IT = N1 / T;
    constexpr int NP = TL / 2;                // row pairs per tile
    constexpr int IS = NP * N1 / (T * TL);    // split items per thread
    static_assert(N1 % (2 * T) == 0 && TL % 2 == 0, "tile shape");
    extern __shared__ __align__(16) unsigned char kfa_rtile_smem[];
    St *buf = reinterpret_cast<St *>(kfa_rtile_smem);

    const int N2 = a.other;
    const long long tpt = N2 / 2 / NP;  // tiles per transform
    const long long tile = tile_of_block(blockIdx.x, a.tiles, a.xcd_group);
    const long long b = tile / tpt;
    const int j0 = (int)(tile - b * tpt) * NP + bin_rotation(a.out, b, a.out_hi);
    const int tid = threadIdx.x;

    {   // row mapping: T lanes per row, every load issued before the LDS writes
        const int s = tid / T, t = tid - (tid / T) * T;
        const int row = pair_line(rot_pair(j0, s >> 1, N2 / 2), s & 1, N2);
        const St *src = reinterpret_cast<const St *>(a.in) + b * a.in_hi + (long long)row * N1 + t;
        St v[IT];
#pragma unroll
        for (int k = 0; k < IT; ++k)
            v[k] = src[k * T];
#pragma unroll
        for (int k = 0; k < IT; ++k)
            buf[s * NS + t + k * T] = v[k];
    }
    __syncthreads();
    {
        const int line = tid / T, t = tid - line * T;
        tile_stages<F32, false, T, RS>(std::make_integer_sequence<int, RS::K>{}, buf + line * NS,
                                       reinterpret_cast<const St *>(a.stw), t);
    }
    // Split item (j, k1): lanes over j, so a wave's bins are NP-long runs.  Pair 0: items
    // k1 < N1/2 take row 0's pairs (k1, N1-k1) -- k1 = 0 is DC / Nyquist, which also takes
    // the middle bin nc/2 (row 0, element N1/2) -- and items k1 >= N1/2 take row N2/2's
    // pairs (m, N1-1-m), m = k1 - N1/2.  Every other item takes Z[q + N2*k1] (row q) and
    // its partner (row N2-q, element N1-1-k1).
    const St *rta = reinterpret_cast<const St *>(a.rtw_a);
    const St *rtb = reinterpret_cast<const St *>(a.rtw_b);
    St *out = reinterpret_cast<St *>(a.out) + b * a.out_hi;
    const long long nc = a.nc;
    C za[IS], zb[IS], ta[IS], tb[IS], zm{};
    long long kl[IS];
#pragma unroll
    for (int i = 0; i < IS; ++i) {
        const int w = tid + i * T * TL;
        const int j = w % NP, k1 = w / NP;
        int ea, eb, sa = 2 * j, sb = 2 * j + 1;
        int q = rot_pair(j0, j, N2 / 2), e = k1;  // the lower bin of the pair is q + N2*e
        bool swap = false;
        if (q == 0) {
            if (2 * k1 < N1) {  // row 0
                sb = sa;
                ea = k1;
                eb = k1 ? N1 - k1 : 0;
            } else {            // row N2/2
                sa = sb;
                ea = k1 - N1 / 2;
                eb = N1 - 1 - ea;
                q = N2 / 2;
                e = ea;
            }
        } else {
            ea = k1;
            eb = N1 - 1 - k1;
            swap = 2 * (q + (long long)N2 * k1) > nc;
            if (swap) {
                q = N2 - q;
                e = N1 - 1 - k1;
            }
        }
        const C z1 = O::load(buf[sa * NS + ea]), z2 = O::load(buf[sb * NS + eb]);
        za[i] = swap ? z2 : z1;  // Z[min(k, nc-k)]
        zb[i] = swap ? z1 : z2;
        kl[i] = q + (long long)N2 * e;
        if (kl[i] == 0)  // Z[nc/2] = row 0, element N1/2: its own partner
            zm = O::load(buf[sa * NS + N1 / 2]);
        ta[i] = O::load(rta[q]);
        tb[i] = O::load(rtb[e]);
    }
#pragma unroll
    for (int i = 0; i < IS; ++i) {
        C xa, xb;
        if (kl[i] == 0) {
            split_dc<F32>(za[i], xa, xb);
            out[0] = O::pack(xa);
            out[nc] = O::pack(xb);
            split_k<F32>(zm, zm, O::cmul(O::load(rta[0]), O::load(rtb[N1 / 2])), xa, xb);
            out[nc / 2] = O::pack(xb);
            continue;
        }
        split_k<F32>(za[i], zb[i], O::cmul(ta[i], tb[i]), xa, xb);
        out[kl[i]] = O::pack(xa);
        out[nc - kl[i]] = O::pack(xb);
    }
}

// Inverse: bins -> merged columns -> column transforms -> four-step twiddle -> first-pass
// output (element n1 + N1*k2).  Column length N2 = RS::N(); TL/2 column pairs per
// workgroup; N1/2 pairs per transform (TL/2 divides N1/2, checked by the launcher).
template <int T, int TL, class RS>
__global__ void __launch_bounds__(T *TL) k_tile_rmerge(RealTileArgs a)
{
    typedef F32::store2 St;
    typedef Ops<F32> O;
    typedef Cx<F32> C;
    constexpr int N2 = RS::N();
    constexpr int NS = N2 + 1;
    constexpr int IT = N2 / T;
    constexpr int NP = TL / 2;              // column pairs per tile
    constexpr int IM = NP * N2 / (T * TL);  // merge items per thread
    static_assert(N2 % (2 * T) == 0 && TL % 2 == 0, "tile shape");
    extern __shared__ __align__(16) unsigned char kfa_rtile_smem[];
    St *buf = reinterpret_cast<St *>(kfa_rtile_smem);
    __shared__ St s_step[TL * IT];  // per line: W^(col*T*k), as in k_tile
    __shared__ St s_nyq;

    const int N1 = a.other;
    const long long nc = a.nc;
    const long long tpt = N1 / 2 / NP;
    const long long tile = tile_of_block(blockIdx.x, a.tiles, a.xcd_group);
    const long long b = tile / tpt;
    const int j0 = (int)(tile - b * tpt) * NP + bin_rotation(a.in, b, a.in_hi);
    const int tid = threadIdx.x;
    const St *in = reinterpret_cast<const St *>(a.in) + b * a.in_hi;
    const St *whi = reinterpret_cast<const St *>(a.tw_hi);
    const St *wlo = reinterpret_cast<const St *>(a.tw_lo);

    {   // column mapping: lanes over lines, every load issued before the LDS writes
        const int s = tid % TL, e0 = tid / TL;
        const int col = pair_line(rot_pair(j0, s >> 1, N1 / 2), s & 1, N1);
        const St *src = in + col + (long long)e0 * N1;
        St v[IT];
#pragma unroll
        for (int k = 0; k < IT; ++k)
            v[k] = src[(long long)k * T * N1];
        if (tid == 0)
            s_nyq = in[nc];
#pragma unroll
        for (int k = 0; k < IT; ++k)
            buf[s * NS + e0 + k * T] = v[k];
    }
    for (int w = tid; w < TL * IT; w += T * TL) {
        const int s = w / IT, k = w - (w / IT) * IT;
        const unsigned long long col = (unsigned)pair_line(rot_pair(j0, s >> 1, N1 / 2), s & 1, N1);
        const unsigned long long ex = col * (unsigned)(k * T);  // < nc
        s_step[w] = O::pack(O::cmul(O::load(whi[ex >> 10]), O::load(wlo[ex & 1023])));
    }
    __syncthreads();
    {   // merge in place; item (j, n2) owns the two LDS slots it reads (mirror of the split)
        const St *rta = reinterpret_cast<const St *>(a.rtw_a);
        const St *rtb = reinterpret_cast<const St *>(a.rtw_b);
        C xa[IM], xb[IM], ta[IM], tb[IM], xm{};
        int ia[IM], ib[IM];  // LDS slots of z[kl] and z[nc-kl]
        bool dc[IM];
#pragma unroll
        for (int i = 0; i < IM; ++i) {
            const int w = tid + i * T * TL;
            const int j = w % NP, n2 = w / NP;
            int ea, eb, sa = 2 * j, sb = 2 * j + 1;
            int p = rot_pair(j0, j, N1 / 2), e = n2;  // the lower bin of the pair is p + N1*e
            bool swap = false;
            if (p == 0) {
                if (2 * n2 < N2) {  // column 0
                    sb = sa;
                    ea = n2;
                    eb = n2 ? N2 - n2 : 0;
                } else {            // column N1/2
                    sa = sb;
                    ea = n2 - N2 / 2;
                    eb = N2 - 1 - ea;
                    p = N1 / 2;
                    e = ea;
                }
            } else {
                ea = n2;
                eb = N2 - 1 - n2;
                swap = 2 * (p + (long long)N1 * n2) > nc;
                if (swap) {
                    p = N1 - p;
                    e = N2 - 1 - n2;
                }
            }
            const int s1 = sa * NS + ea, s2 = sb * NS + eb;
            const C x1 = O::load(buf[s1]), x2 = O::load(buf[s2]);
            xa[i] = swap ? x2 : x1;  // X[min(k, nc-k)]
            xb[i] = swap ? x1 : x2;
            ia[i] = swap ? s2 : s1;
            ib[i] = swap ? s1 : s2;
            dc[i] = p == 0 && e == 0;
            if (dc[i])  // z[nc/2] (column 0, element N2/2) pairs with itself; the DC item takes it
                xm = O::load(buf[s1 + N2 / 2]);
            ta[i] = O::load(rta[p]);
            tb[i] = O::load(rtb[e]);
        }
        const C twm = O::cmul(O::load(rta[0]), O::load(rtb[N2 / 2]));
        __syncthreads();  // every operand read before any slot is rewritten
#pragma unroll
        for (int i = 0; i < IM; ++i) {
            C zl, zh;
            if (dc[i]) {
                buf[ia[i]] = O::pack(merge_dc<F32>(xa[i], O::load(s_nyq)));
                merge_k<F32>(xm, xm, twm, zl, zh);
                buf[ia[i] + N2 / 2] = O::pack(zh);
                continue;
            }
            merge_k<F32>(xa[i], xb[i], O::cmul(ta[i], tb[i]), zl, zh);
            buf[ia[i]] = O::pack(zl);
            buf[ib[i]] = O::pack(zh);
        }
    }
    __syncthreads();
    {
        const int line = tid / T, t = tid - line * T;
        tile_stages<F32, true, T, RS>(std::make_integer_sequence<int, RS::K>{}, buf + line * NS,
                                      reinterpret_cast<const St *>(a.stw), t);
    }
    {   // column stores with W^(col*k2), k2 = e0 + k*T
        const int s = tid % TL, e0 = tid / TL;
        const int col = pair_line(rot_pair(j0, s >> 1, N1 / 2), s & 1, N1);
        const unsigned long long ex = (unsigned long long)col * (unsigned)e0;
        const C base = O::cmul(O::load(whi[ex >> 10]), O::load(wlo[ex & 1023]));
        St *dst = reinterpret_cast<St *>(a.out) + b * a.out_hi + col + (long long)e0 * N1;
        C v[IT];
#pragma unroll
        for (int k = 0; k < IT; ++k)
            v[k] = O::cmul(O::load(buf[s * NS + e0 + k * T]), O::cmul(base, O::load(s_step[s * IT + k])));
#pragma unroll
        for (int k = 0; k < IT; ++k)
            dst[(long long)k * T * N1] = O::pack(v[k]);
    }
}

}  // namespace kfa
