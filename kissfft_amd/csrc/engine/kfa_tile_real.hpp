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
    constexpr int NS = N1 + 1;  // row loads: the odd pitch
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

    // Split item (j, k1): lanes over j, so a wave's bins are NP-long runs.  Pair 0: items
    // k1 < N1/2 take row 0's pairs (k1, N1-k1) -- k1 = 0 is DC / Nyquist, which also takes
    // the middle bin nc/2 (row 0, element N1/2) -- and items k1 >= N1/2 take row N2/2's
    // pairs (m, N1-1-m), m = k1 - N1/2.  Every other item takes Z[q + N2*k1] (row q) and
    // its partner (row N2-q, element N1-1-k1).  The items' LDS slots and twiddle reads are
    // set up first, so that the table reads are in flight with the row loads.
    const St *rta = reinterpret_cast<const St *>(a.rtw_a);
    const St *rtb = reinterpret_cast<const St *>(a.rtw_b);
    const long long nc = a.nc;
    int s1[IS], s2[IS];  // LDS slots of Z[kl] and Z[nc-kl]
    long long kl[IS];
    St ta[IS], tb[IS], tm0{}, tm1{};
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
        s1[i] = swap ? sb * NS + eb : sa * NS + ea;
        s2[i] = swap ? sa * NS + ea : sb * NS + eb;
        kl[i] = q + (long long)N2 * e;
        ta[i] = rta[q];
        tb[i] = rtb[e];
        if (kl[i] == 0) {  // the middle bin's twiddle
            tm0 = rta[0];
            tm1 = rtb[N1 / 2];
        }
    }
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
    if constexpr (64 % T == 0)
        __syncthreads();  // the stages order each line only: the split reads across lines
    St *out = reinterpret_cast<St *>(a.out) + b * a.out_hi;
#pragma unroll
    for (int i = 0; i < IS; ++i) {
        const C za = O::load(buf[s1[i]]), zb = O::load(buf[s2[i]]);  // Z[kl], Z[nc-kl]
        C xa, xb;
        if (kl[i] == 0) {
            split_dc<F32>(za, xa, xb);
            out[0] = O::pack(xa);
            out[nc] = O::pack(xb);
            const C zm = O::load(buf[s1[i] + N1 / 2]);  // Z[nc/2] = row 0, element N1/2: its own partner
            split_k<F32>(zm, zm, O::cmul(O::load(tm0), O::load(tm1)), xa, xb);
            out[nc / 2] = O::pack(xb);
            continue;
        }
        split_k<F32>(za, zb, O::cmul(O::load(ta[i]), O::load(tb[i])), xa, xb);
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
    constexpr int NS = tile_pitch<N2, TL>();  // column loads: see k_tile
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

    // Twiddle reads first (they are in flight with the column loads): the per-line steps
    // W^(col*T*k), this thread's store base W^(col*e0), and the merge items' factors.
    constexpr int NSTEP = (TL * IT + T * TL - 1) / (T * TL);
    St sh[NSTEP], sl[NSTEP];
#pragma unroll
    for (int i = 0; i < NSTEP; ++i) {
        const int w = tid + i * T * TL;
        if (w < TL * IT) {
            const int s = w / IT, k = w - (w / IT) * IT;
            const unsigned long long col = (unsigned)pair_line(rot_pair(j0, s >> 1, N1 / 2), s & 1, N1);
            const unsigned long long ex = col * (unsigned)(k * T);  // < nc
            sh[i] = whi[ex >> 10];
            sl[i] = wlo[ex & 1023];
        }
    }
    const int scol = pair_line(rot_pair(j0, (tid % TL) >> 1, N1 / 2), (tid % TL) & 1, N1);
    St bh, bl;
    {
        const unsigned long long ex = (unsigned long long)scol * (unsigned)(tid / TL);
        bh = whi[ex >> 10];
        bl = wlo[ex & 1023];
    }
    // merge item (j, n2) owns the two LDS slots it reads (mirror of the split)
    const St *rta = reinterpret_cast<const St *>(a.rtw_a);
    const St *rtb = reinterpret_cast<const St *>(a.rtw_b);
    St ta[IM], tb[IM], tm0{}, tm1{};
    int ia[IM], ib[IM];  // LDS slots of X[kl] and X[nc-kl] (then z[kl], z[nc-kl])
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
        ia[i] = swap ? sb * NS + eb : sa * NS + ea;
        ib[i] = swap ? sa * NS + ea : sb * NS + eb;
        dc[i] = p == 0 && e == 0;
        ta[i] = rta[p];
        tb[i] = rtb[e];
        if (dc[i]) {  // the middle bin's twiddle
            tm0 = rta[0];
            tm1 = rtb[N2 / 2];
        }
    }
    {   // column mapping: lanes over lines, every load issued before the LDS writes
        const int s = tid % TL, e0 = tid / TL;
        const St *src = in + scol + (long long)e0 * N1;
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
#pragma unroll
    for (int i = 0; i < NSTEP; ++i) {
        const int w = tid + i * T * TL;
        if (w < TL * IT)
            s_step[w] = O::pack(O::cmul(O::load(sh[i]), O::load(sl[i])));
    }
    __syncthreads();
    {   // merge in place
        C xa[IM], xb[IM], xm{};
#pragma unroll
        for (int i = 0; i < IM; ++i) {
            xa[i] = O::load(buf[ia[i]]);  // X[kl]
            xb[i] = O::load(buf[ib[i]]);  // X[nc-kl]
            if (dc[i])  // z[nc/2] (column 0, element N2/2) pairs with itself; the DC item takes it
                xm = O::load(buf[ia[i] + N2 / 2]);
        }
        __syncthreads();  // every operand read before any slot is rewritten
#pragma unroll
        for (int i = 0; i < IM; ++i) {
            C zl, zh;
            if (dc[i]) {
                buf[ia[i]] = O::pack(merge_dc<F32>(xa[i], O::load(s_nyq)));
                merge_k<F32>(xm, xm, O::cmul(O::load(tm0), O::load(tm1)), zl, zh);
                buf[ia[i] + N2 / 2] = O::pack(zh);
                continue;
            }
            merge_k<F32>(xa[i], xb[i], O::cmul(O::load(ta[i]), O::load(tb[i])), zl, zh);
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
    if constexpr (64 % T == 0)
        __syncthreads();  // the stages order each line only: the stores read across lines
    {   // column stores with W^(col*k2), k2 = e0 + k*T
        const int s = tid % TL, e0 = tid / TL;
        const C base = O::cmul(O::load(bh), O::load(bl));
        St *dst = reinterpret_cast<St *>(a.out) + b * a.out_hi + scol + (long long)e0 * N1;
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
