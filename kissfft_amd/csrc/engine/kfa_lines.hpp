// kfa_lines.hpp -- the general LDS-resident Stockham "line" kernel.
//
// A launch transforms `nlines` independent lines.  Line l = (hi, lo) with lo = l % nlo
// reads element t at  in[hi*in_hi + lo*in_lo + t*in_el]  and writes element t at
// out[hi*out_hi + lo*out_lo + t*out_el].  With suitable strides one kernel serves a batch
// of 1-D transforms (kiss_fft_stride's in_stride included), the per-axis passes of the
// N-D transform (kiss_fftnd.c:172-177 walks the same columns), and both passes of the
// four-step split used for lines longer than LDS holds.
//
// A workgroup owns a tile of TL consecutive lines.  The tile is read once from HBM into
// LDS, every radix stage runs LDS -> LDS (ping-pong), and the result is written once.
// LDS layout: element (line, t) sits at t*TL + line in "column" tiles (lines adjacent in
// HBM, so a wave reads TL neighbours per row) and at line*L + t in "row" tiles.
//
// Stage arithmetic is the Stockham form of the reference's decimation in time: with
// radix p, sub-length m and fstride P = L/(p*m), butterfly b = u*P + o reads inputs
// u*P*p + r*P + o (r < p) and writes outputs b + q*(L/p) (q < p), using twiddles
// table[r*u*P] -- exactly the values kf_bfly* (kiss_fft.c:15-233) combines at column u of
// the recursion, so fixed-point results are bit-identical.  For the second pass of a
// four-step split the stage's column index is u_global = lo + u_mul*u and the twiddle
// index uses the global fstride Ptw.
#pragma once

#include "kfa_arith.hpp"

#define KFA_MAX_STAGES 32

namespace kfa {

struct StageDesc {
    int p;      // radix
    int ml;     // local sub-length m (in line elements)
    int Pl;     // local fstride = L / (p * ml)
    int Ptw;    // fstride in the full-length twiddle table
    int mg;     // global sub-length m (table index of ya is Ptw*mg)
};

struct LinePlan {
    int L;        // line length
    int N;        // full transform length (twiddle table length)
    int nst;      // number of stages
    int inverse;
    int u_mul;    // u_global = u_add + u_mul * u_local
    int u_add_lo; // 1: u_add is the line's lo index (four-step second pass), else 0
    StageDesc st[KFA_MAX_STAGES];
};

struct LineIO {
    long long nlines;
    long long nlo;
    long long in_hi, in_lo, in_el;
    long long out_hi, out_lo, out_el;
};

enum { EPI_NONE = 0, EPI_R2C = 1 };
enum { PRO_NONE = 0, PRO_C2R = 1 };

template <class A>
struct LineCtx {
    typedef typename A::store2 S;
    int TL, nl, L, colmode;
    __device__ __forceinline__ int idx(int line, int t) const { return colmode ? t * TL + line : line * L + t; }
    // flat work item -> (line, j) for j in [0, n)
    __device__ __forceinline__ void split(int w, int n, int &line, int &j) const
    {
        if (colmode) {
            line = w % TL;
            j = w / TL;
        } else {
            line = w / n;
            j = w - line * n;
        }
    }
};

template <class A>
__device__ __forceinline__ void run_stage(const LinePlan &pl, const StageDesc &sd, const LineCtx<A> &cx,
                                          const typename A::store2 *src, typename A::store2 *dst,
                                          const typename A::store2 *__restrict__ tw, long long l0,
                                          long long nlo)
{
    typedef Ops<A> O;
    typedef Cx<A> C;
    const int p = sd.p, L = cx.L, nb = L / p, Pl = sd.Pl;
    const int nthreads = blockDim.x;
    if (p >= 2 && p <= 5) {
        const C ya = O::load(tw[(long long)sd.Ptw * sd.mg % pl.N]);
        const C yb = O::load(tw[2LL * sd.Ptw * sd.mg % pl.N]);
        for (int w = threadIdx.x; w < cx.TL * nb; w += nthreads) {
            int line, b;
            cx.split(w, nb, line, b);
            if (line >= cx.nl)
                continue;
            const int u = b / Pl, o = b - (b / Pl) * Pl;
            const long long uadd = pl.u_add_lo ? (l0 + line) % nlo : 0;
            const long long ug = uadd + (long long)pl.u_mul * u;
            C v[5], wv[4];
            for (int r = 0; r < p; ++r)
                v[r] = O::load(src[cx.idx(line, u * Pl * p + r * Pl + o)]);
            for (int r = 1; r < p; ++r)
                wv[r - 1] = O::load(tw[(long long)r * ug * sd.Ptw]);
            switch (p) {
            case 2: bfly2<A>(v, wv); break;
            case 3: bfly3<A>(v, wv, ya); break;
            case 4: bfly4<A>(v, wv, pl.inverse != 0); break;
            default: bfly5<A>(v, wv, ya, yb); break;
            }
            for (int q = 0; q < p; ++q)
                dst[cx.idx(line, b + q * nb)] = O::pack(v[q]);
        }
    } else {
        // Generic radix (kiss_fft.c:191-233): one work item per output k = u + q1*m.
        for (int w = threadIdx.x; w < cx.TL * L; w += nthreads) {
            int line, j;
            cx.split(w, L, line, j);
            if (line >= cx.nl)
                continue;
            const int q1 = j / nb, b = j - (j / nb) * nb;
            const int u = b / Pl, o = b - (b / Pl) * Pl;
            const long long uadd = pl.u_add_lo ? (l0 + line) % nlo : 0;
            const long long ug = uadd + (long long)pl.u_mul * u;
            const long long kg = ug + (long long)q1 * sd.mg;
            const long long step = (long long)sd.Ptw * kg % pl.N;
            C acc = O::fixdiv(O::load(src[cx.idx(line, u * Pl * p + o)]), p);
            long long ti = 0;
            for (int q = 1; q < p; ++q) {
                ti += step;
                if (ti >= pl.N)
                    ti -= pl.N;
                C x = O::fixdiv(O::load(src[cx.idx(line, u * Pl * p + q * Pl + o)]), p);
                acc = O::cadd(acc, O::cmul(x, O::load(tw[ti])));
            }
            dst[cx.idx(line, j)] = O::pack(acc);
        }
    }
}

// Real-input split after the half-length transform, kiss_fftr.c:88-116.
// DC/Nyquist pair from Z[0]:
template <class A>
__device__ __forceinline__ void split_dc(Cx<A> z0, Cx<A> &x0, Cx<A> &xn)
{
    typedef Ops<A> O;
    const Cx<A> dc = O::fixdiv(z0, 2);
    x0 = Cx<A>{O::add(dc.r, dc.i), O::zero().r};
    xn = Cx<A>{O::sub(dc.r, dc.i), O::zero().r};
}

// X[k] and X[nc-k] (0 < k <= nc/2) from Z[k], Z[nc-k] and super twiddle tw = stw[k-1].
template <class A>
__device__ __forceinline__ void split_k(Cx<A> zk, Cx<A> znk, Cx<A> tw, Cx<A> &xk, Cx<A> &xnk)
{
    typedef Ops<A> O;
    Cx<A> a = O::fixdiv(zk, 2);
    Cx<A> b = O::fixdiv(Cx<A>{znk.r, O::neg(znk.i)}, 2);
    const Cx<A> f1 = O::cadd(a, b), f2 = O::csub(a, b);
    const Cx<A> t = O::cmul(f2, tw);
    xk = Cx<A>{O::half_add(f1.r, t.r), O::half_add(f1.i, t.i)};
    xnk = Cx<A>{O::half_sub(f1.r, t.r), O::half_sub(t.i, f1.i)};
}

// Hermitian merge before the half-length inverse transform, kiss_fftr.c:131-153.
template <class A>
__device__ __forceinline__ Cx<A> merge_dc(Cx<A> x0, Cx<A> xn)
{
    typedef Ops<A> O;
    return O::fixdiv(Cx<A>{O::add(x0.r, xn.r), O::sub(x0.r, xn.r)}, 2);
}

// Z[k] and Z[nc-k] (0 < k <= nc/2) from X[k], X[nc-k] and tw = stw[k-1].
template <class A>
__device__ __forceinline__ void merge_k(Cx<A> xk, Cx<A> xnk, Cx<A> tw, Cx<A> &zk, Cx<A> &znk)
{
    typedef Ops<A> O;
    Cx<A> a = O::fixdiv(xk, 2);
    Cx<A> b = O::fixdiv(Cx<A>{xnk.r, O::neg(xnk.i)}, 2);
    const Cx<A> e = O::cadd(a, b), d = O::csub(a, b);
    const Cx<A> o = O::cmul(d, tw);
    Cx<A> lo = O::csub(e, o);
    lo.i = O::neg(lo.i);
    zk = O::cadd(e, o);
    znk = lo;
}

template <class A>
__device__ __forceinline__ void r2c_epilogue(const LineCtx<A> &cx, const typename A::store2 *z,
                                             typename A::store2 *__restrict__ out,
                                             const typename A::store2 *__restrict__ stw,
                                             const LineIO &io, long long l0)
{
    typedef Ops<A> O;
    typedef Cx<A> C;
    const int nc = cx.L, nk = nc / 2 + 1;
    for (int w = threadIdx.x; w < cx.TL * nk; w += blockDim.x) {
        int line, k;
        cx.split(w, nk, line, k);
        if (line >= cx.nl)
            continue;
        const long long l = l0 + line;
        const long long base = (l / io.nlo) * io.out_hi + (l % io.nlo) * io.out_lo;
        C xa, xb;
        if (k == 0) {
            split_dc<A>(O::load(z[cx.idx(line, 0)]), xa, xb);
            out[base] = O::pack(xa);
            out[base + (long long)nc * io.out_el] = O::pack(xb);
            continue;
        }
        split_k<A>(O::load(z[cx.idx(line, k)]), O::load(z[cx.idx(line, nc - k)]), O::load(stw[k - 1]), xa, xb);
        if (2 * k != nc)
            out[base + (long long)k * io.out_el] = O::pack(xa);
        out[base + (long long)(nc - k) * io.out_el] = O::pack(xb);
    }
}

template <class A>
__device__ __forceinline__ void c2r_prologue(const LineCtx<A> &cx, typename A::store2 *z,
                                             const typename A::store2 *__restrict__ in,
                                             const typename A::store2 *__restrict__ stw,
                                             const LineIO &io, long long l0)
{
    typedef Ops<A> O;
    typedef Cx<A> C;
    const int nc = cx.L, nk = nc / 2 + 1;
    for (int w = threadIdx.x; w < cx.TL * nk; w += blockDim.x) {
        int line, k;
        cx.split(w, nk, line, k);
        if (line >= cx.nl)
            continue;
        const long long l = l0 + line;
        const long long base = (l / io.nlo) * io.in_hi + (l % io.nlo) * io.in_lo;
        if (k == 0) {
            z[cx.idx(line, 0)] = O::pack(merge_dc<A>(O::load(in[base]), O::load(in[base + (long long)nc * io.in_el])));
            continue;
        }
        C za, zb;
        merge_k<A>(O::load(in[base + (long long)k * io.in_el]), O::load(in[base + (long long)(nc - k) * io.in_el]),
                   O::load(stw[k - 1]), za, zb);
        if (2 * k != nc)
            z[cx.idx(line, k)] = O::pack(za);
        z[cx.idx(line, nc - k)] = O::pack(zb);
    }
}

// Standalone split / merge passes (HBM -> HBM) for half lengths longer than one LDS line:
// the half-length complex transform runs as a multi-pass transform into a temporary and
// these kernels apply kiss_fftr.c:88-116 (split) or kiss_fftr.c:131-153 (merge) to it.
// Thread k of line l handles the pair (k, nc-k), k in [0, nc/2]; grid.y strides lines.
// Distances are in complex values; z lines are nc apart.
template <class A>
__global__ void __launch_bounds__(256) k_real_split(const typename A::store2 *__restrict__ z,
                                                    typename A::store2 *__restrict__ out,
                                                    const typename A::store2 *__restrict__ stw, int nc,
                                                    long long out_hi, long long nlines)
{
    typedef Ops<A> O;
    typedef Cx<A> C;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > nc / 2)
        return;
    for (long long l = blockIdx.y; l < nlines; l += gridDim.y) {
        const typename A::store2 *zl = z + l * nc;
        typename A::store2 *ol = out + l * out_hi;
        C xa, xb;
        if (k == 0) {
            split_dc<A>(O::load(zl[0]), xa, xb);
            ol[0] = O::pack(xa);
            ol[nc] = O::pack(xb);
            continue;
        }
        split_k<A>(O::load(zl[k]), O::load(zl[nc - k]), O::load(stw[k - 1]), xa, xb);
        if (2 * k != nc)
            ol[k] = O::pack(xa);
        ol[nc - k] = O::pack(xb);
    }
}

template <class A>
__global__ void __launch_bounds__(256) k_real_merge(const typename A::store2 *__restrict__ in,
                                                    typename A::store2 *__restrict__ z,
                                                    const typename A::store2 *__restrict__ stw, int nc,
                                                    long long in_hi, long long nlines)
{
    typedef Ops<A> O;
    typedef Cx<A> C;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > nc / 2)
        return;
    for (long long l = blockIdx.y; l < nlines; l += gridDim.y) {
        const typename A::store2 *il = in + l * in_hi;
        typename A::store2 *zl = z + l * nc;
        if (k == 0) {
            zl[0] = O::pack(merge_dc<A>(O::load(il[0]), O::load(il[nc])));
            continue;
        }
        C za, zb;
        merge_k<A>(O::load(il[k]), O::load(il[nc - k]), O::load(stw[k - 1]), za, zb);
        if (2 * k != nc)
            zl[k] = O::pack(za);
        zl[nc - k] = O::pack(zb);
    }
}

// One tile of TL lines starting at line l0: HBM -> LDS, every stage, LDS -> HBM.
// bufA holds 2 * TL * L complex values.
template <class A>
__device__ __forceinline__ void lines_tile(const typename A::store2 *__restrict__ in,
                                           typename A::store2 *__restrict__ out,
                                           const typename A::store2 *__restrict__ tw,
                                           const typename A::store2 *__restrict__ stw, const LinePlan &pl,
                                           const LineIO &io, int TL, int colmode, int prologue, int epilogue,
                                           long long l0, typename A::store2 *bufA)
{
    typedef typename A::store2 S;
    S *bufB = bufA + (size_t)TL * pl.L;
    LineCtx<A> cx;
    cx.TL = TL;
    cx.L = pl.L;
    cx.colmode = colmode;
    cx.nl = (int)((io.nlines - l0) < TL ? (io.nlines - l0) : TL);

    if (prologue == PRO_C2R) {
        c2r_prologue<A>(cx, bufA, in, stw, io, l0);
    } else {
        const int L = pl.L;
        for (int w = threadIdx.x; w < TL * L; w += blockDim.x) {
            int line, t;
            cx.split(w, L, line, t);
            if (line >= cx.nl)
                continue;
            const long long l = l0 + line;
            const long long base = (l / io.nlo) * io.in_hi + (l % io.nlo) * io.in_lo;
            bufA[cx.idx(line, t)] = in[base + (long long)t * io.in_el];
        }
    }
    __syncthreads();

    S *src = bufA, *dst = bufB;
    for (int s = 0; s < pl.nst; ++s) {
        run_stage<A>(pl, pl.st[s], cx, src, dst, tw, l0, io.nlo);
        __syncthreads();
        S *t = src;
        src = dst;
        dst = t;
    }

    if (epilogue == EPI_R2C) {
        r2c_epilogue<A>(cx, src, out, stw, io, l0);
    } else {
        const int L = pl.L;
        for (int w = threadIdx.x; w < TL * L; w += blockDim.x) {
            int line, t;
            cx.split(w, L, line, t);
            if (line >= cx.nl)
                continue;
            const long long l = l0 + line;
            const long long base = (l / io.nlo) * io.out_hi + (l % io.nlo) * io.out_lo;
            out[base + (long long)t * io.out_el] = src[cx.idx(line, t)];
        }
    }
}

// The line kernel: one tile per workgroup.  Dynamic LDS: 2 * TL * L complex values.
template <class A>
__global__ void __launch_bounds__(256) k_lines(const typename A::store2 *__restrict__ in,
                                               typename A::store2 *__restrict__ out,
                                               const typename A::store2 *__restrict__ tw,
                                               const typename A::store2 *__restrict__ stw,
                                               LinePlan pl, LineIO io, int TL, int colmode, int prologue,
                                               int epilogue)
{
    extern __shared__ __align__(16) unsigned char kfa_smem[];
    lines_tile<A>(in, out, tw, stw, pl, io, TL, colmode, prologue, epilogue, (long long)blockIdx.x * TL,
                  reinterpret_cast<typename A::store2 *>(kfa_smem));
}

// Several independent batches of different lengths and kinds in one persistent launch:
// item i owns tiles [tile0, tile0 + ntiles) of TL lines each; every workgroup walks the
// global tile list with stride gridDim.x (no counters, so the launch can be captured and
// replayed as is).
struct MultiItem {
    LinePlan pl;
    const void *tw, *stw;
    const void *in;
    void *out;
    LineIO io;
    long long tile0;
    int TL, prologue, epilogue;
};

template <class A>
__global__ void __launch_bounds__(256) k_multi(const MultiItem *__restrict__ items, int nitems, long long ntiles)
{
    typedef typename A::store2 S;
    extern __shared__ __align__(16) unsigned char kfa_msmem[];
    S *buf = reinterpret_cast<S *>(kfa_msmem);
    for (long long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        int lo = 0, hi = nitems - 1;  // last item with tile0 <= tile
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (items[mid].tile0 <= tile)
                lo = mid;
            else
                hi = mid - 1;
        }
        const MultiItem &it = items[lo];
        lines_tile<A>(reinterpret_cast<const S *>(it.in), reinterpret_cast<S *>(it.out),
                      reinterpret_cast<const S *>(it.tw), reinterpret_cast<const S *>(it.stw), it.pl, it.io, it.TL,
                      0, it.prologue, it.epilogue, (tile - it.tile0) * it.TL, buf);
        __syncthreads();  // the next tile reuses the buffers
    }
}

}  // namespace kfa
