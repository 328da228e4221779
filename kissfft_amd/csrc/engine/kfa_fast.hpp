// kfa_fast.hpp -- compile-time specialised single-pass Stockham kernels for hot sizes.
//
// Same stage arithmetic as kfa_lines.hpp (butterfly b = u*P + o reads u*P*p + r*P + o,
// writes b + q*(N/p), twiddles table[r*u*P]), but with the radix schedule, the line length
// and the thread mapping fixed at compile time, so all index arithmetic folds into
// constants and every loop unrolls.  Data moves HBM -> registers (the first stage reads its
// inputs straight from HBM: for m == 1 the reads r*(N/p) + b are coalesced), then
// registers -> LDS -> registers between stages (ping-pong buffers, one barrier per stage),
// and the last stage writes registers -> HBM (outputs b + q*(N/p), coalesced).  The real
// transforms fuse the split (forward) or merge (inverse) of kiss_fftr.c into the same pass.
//
// Twiddles come from a per-stage table built on the host by gathering the reference's
// full-length table (entry (u, r) of stage s is table[r*u*P_s]; the last two entries are
// table[P_s*m_s] and table[2*P_s*m_s], the fixed rotations of radix 3 and 5), so fixed
// point uses exactly the values kf_bfly* uses.  Float stages use their own radix-p DFTs
// (the decimation is the same, the butterfly algebra is free); fixed-point stages use the
// reference-order butterflies of kfa_arith.hpp.
#pragma once

#ifndef __HIPCC_RTC__
#include <utility>
#else  // compiled at run time (kfa_rtc.hip): no standard library; the sequence type only
namespace std {
template <class T, T... I>
struct integer_sequence {
};
template <class T, T N>
using make_integer_sequence = __make_integer_seq<integer_sequence, T, N>;
using __hip_internal::integral_constant;
}  // namespace std
#endif

#include "kfa_arith.hpp"
#include "kfa_lines.hpp"
#include "kfa_q15p.hpp"

namespace kfa {

enum { FK_C2C = 0, FK_R2C = 1, FK_C2R = 2 };

// Stage-table layout: stage s holds m_s rows of PR(p_s) entries (entry r-1 of row u is the
// twiddle of input r at column u; rows padded to G entries so a row is a 16-byte aligned
// run), then the two fixed rotations padded to G entries.
constexpr int round_up(int v, int g) { return (v + g - 1) / g * g; }

template <int G, int... R>
struct Radices {
    static constexpr int K = sizeof...(R);
    static constexpr int r[K] = {R...};
    static constexpr int p(int s) { return r[s]; }
    static constexpr int m(int s)
    {
        int v = 1;
        for (int i = 0; i < s; ++i)
            v *= r[i];
        return v;
    }
    static constexpr int N() { return m(K); }
    static constexpr int PR(int s) { return round_up(r[s] - 1 > 0 ? r[s] - 1 : 1, G); }
    static constexpr int tw_off(int s)
    {
        int o = 0;
        for (int i = 0; i < s; ++i)
            o += m(i) * PR(i) + round_up(2, G);
        return o;
    }
    static constexpr int tw_len() { return tw_off(K); }
    // Generic radices (any p but 2, 3, 4, 5, 8, 16).  Fixed point runs them as
    // kf_bfly_generic does (kiss_fft.c:191-233): output q1 of column u sums
    // C_MUL(x'_q, table[q * P * (u + q1 * m) mod N]) over q, so after the table above a
    // fixed-point stage table holds, per generic stage, m rows of p * (p - 1) such
    // entries (row u: q1 major, q minor), padded to G.
    static constexpr bool gen(int s)
    {
        return !(r[s] == 2 || r[s] == 3 || r[s] == 4 || r[s] == 5 || r[s] == 8 || r[s] == 16);
    }
    static constexpr int GR(int s) { return round_up(r[s] * (r[s] - 1), G); }
    static constexpr int gen_off(int s)
    {
        int o = tw_len();
        for (int i = 0; i < s; ++i)
            if (gen(i))
                o += m(i) * GR(i);
        return o;
    }
};

// One launch may cover several separate buffers ("segments", e.g. the requests of a
// serving batch): segment s owns lines [first[s], first[s+1]) and reads / writes them at
// in[s] / out[s] with the common distances.  nseg == 0: one buffer, in / out below.
struct FastSeg {
    long long first;  // first line of the segment
    const void *in;
    void *out;
};

#define KFA_MAX_SEG 64

struct FastArgs {
    const void *in;
    void *out;
    const void *stw;  // per-stage twiddle tables
    const void *rtw;  // real split twiddles (nfft/4 entries)
    long long nlines;
    long long in_hi, in_el;
    long long out_hi, out_el;
    int nseg;                    // 0: one buffer (in / out); else segments below
    FastSeg seg[KFA_MAX_SEG];    // ascending `first`, seg[0].first == 0 (passed by value)
};

// The kernel argument of a fused launch (several row-kernel launches of one workgroup size
// as one grid): member i runs workgroups [end[i-1], end[i]) with the argument block at
// arg[i] (a FastArgs or a BlueArgs in device memory).
#define KFA_MAX_FUSED 16
struct FusedTab {
    const void *arg[KFA_MAX_FUSED];
    long long end[KFA_MAX_FUSED];
};

// HBM accesses of the fast kernels.  KFA_NT selects non-temporal (streaming) loads and
// stores: every input is read once and every output written once.
#ifndef KFA_NT
#define KFA_NT 0
#endif
template <class St>
__device__ __forceinline__ St hbm_ld(const St *p)
{
#if KFA_NT
    if constexpr (sizeof(St) == 8) {
        typedef unsigned u2 __attribute__((ext_vector_type(2)));
        return __builtin_bit_cast(St, __builtin_nontemporal_load(reinterpret_cast<const u2 *>(p)));
    } else if constexpr (sizeof(St) == 4) {
        return __builtin_bit_cast(St, __builtin_nontemporal_load(reinterpret_cast<const unsigned *>(p)));
    }
#endif
    return *p;
}
template <class St>
__device__ __forceinline__ void hbm_st(St *p, St v)
{
#if KFA_NT
    if constexpr (sizeof(St) == 8) {
        typedef unsigned u2 __attribute__((ext_vector_type(2)));
        __builtin_nontemporal_store(__builtin_bit_cast(u2, v), reinterpret_cast<u2 *>(p));
        return;
    } else if constexpr (sizeof(St) == 4) {
        __builtin_nontemporal_store(__builtin_bit_cast(unsigned, v), reinterpret_cast<unsigned *>(p));
        return;
    }
#endif
    *p = v;
}

// Base pointers of global line gl: one buffer (nseg == 0) or the segment holding gl, at
// gl's index inside it.  Any argument block with in / out / in_hi / out_hi / nseg / seg
// (FastArgs, BlueArgs).
template <class S, class Args>
__device__ __forceinline__ void line_base(const Args &a, long long gl, const S *&in, S *&out)
{
    if (a.nseg <= 0) {
        in = reinterpret_cast<const S *>(a.in) + gl * a.in_hi;
        out = reinterpret_cast<S *>(a.out) + gl * a.out_hi;
        return;
    }
    int lo = 0, hi = a.nseg - 1;  // last segment with first <= gl
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (a.seg[mid].first <= gl)
            lo = mid;
        else
            hi = mid - 1;
    }
    const long long k = gl - a.seg[lo].first;
    in = reinterpret_cast<const S *>(a.seg[lo].in) + k * a.in_hi;
    out = reinterpret_cast<S *>(a.seg[lo].out) + k * a.out_hi;
}

// ------------------------------------------------------------------------------------
// Float radix-p DFTs on registers.  Forward uses exp(-2*pi*i/p); INV the conjugate.
// ------------------------------------------------------------------------------------
typedef Cx<F32> Cf;

__device__ __forceinline__ Cf cf_add(Cf a, Cf b) { return Cf{a.r + b.r, a.i + b.i}; }
__device__ __forceinline__ Cf cf_sub(Cf a, Cf b) { return Cf{a.r - b.r, a.i - b.i}; }
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ Cf cf_rot(Cf a)
{
    return INV ? Cf{-a.i, a.r} : Cf{a.i, -a.r};
}

// Compile-time cos / sin of 2*pi*j/p in double, rounded once to float: Taylor series
// after reducing the angle to [-pi, pi] (30 terms: the series' remainder there is far
// below double precision).
constexpr double kfa_pi = 3.14159265358979323846264338327950288;
constexpr double taylor_cos(double x)
{
    double t = 1, s = 1;
    for (int k = 1; k < 30; ++k) {
        t *= -x * x / ((2.0 * k - 1) * (2.0 * k));
        s += t;
    }
    return s;
}
constexpr double taylor_sin(double x)
{
    double t = x, s = x;
    for (int k = 1; k < 30; ++k) {
        t *= -x * x / ((2.0 * k) * (2.0 * k + 1));
        s += t;
    }
    return s;
}
constexpr double angle_2pi(int j, int p)
{
    const double x = 2 * kfa_pi * (j % p) / p;
    return x > kfa_pi ? x - 2 * kfa_pi : x;
}
template <int P>
struct Roots {
    float c[P], s[P];  // cos / sin of 2*pi*j/P
};
template <int P>
constexpr Roots<P> make_roots()
{
    Roots<P> r{};
    for (int j = 0; j < P; ++j) {
        r.c[j] = (float)taylor_cos(angle_2pi(j, P));
        r.s[j] = (float)taylor_sin(angle_2pi(j, P));
    }
    return r;
}

// Any odd radix (7, 11, 13, ...; the schedules of lengths with such factors, kf_factor's
// generic radices, kiss_fft.c:191-233): the pairs j, p - j share a cosine and an opposite
// sine, so X[k] = x0 + sum_j c(jk) (x_j + x_{p-j}) -/+ i sum_j s(jk) (x_j - x_{p-j}) and
// X[p-k] takes the other sign: (p-1)^2 / 2 real products per component.
template <int P, bool INV>
struct Dft {
    static_assert(P % 2 == 1 && P >= 7, "generic radix: odd, 7 or more");
    __device__ __forceinline__ static void run(Cf *v)
    {
        constexpr int H = (P - 1) / 2;
        constexpr Roots<P> R = make_roots<P>();
        Cf t[H], d[H];
        const Cf x0 = v[0];
        Cf s0 = x0;
#pragma unroll
        for (int j = 1; j <= H; ++j) {
            t[j - 1] = cf_add(v[j], v[P - j]);
            d[j - 1] = cf_sub(v[j], v[P - j]);
            s0 = cf_add(s0, t[j - 1]);
        }
#pragma unroll
        for (int k = 1; k <= H; ++k) {
            Cf a = x0, b{0.f, 0.f};
#pragma unroll
            for (int j = 1; j <= H; ++j) {
                const float c = R.c[(j * k) % P], sn = R.s[(j * k) % P];
                a = Cf{a.r + c * t[j - 1].r, a.i + c * t[j - 1].i};
                b = Cf{b.r + sn * d[j - 1].r, b.i + sn * d[j - 1].i};
            }
            const Cf rb = cf_rot<INV>(b);
            v[k] = cf_add(a, rb);
            v[P - k] = cf_sub(a, rb);
        }
        v[0] = s0;
    }
};

template <bool INV>
struct Dft<2, INV> {
    __device__ __forceinline__ static void run(Cf *v)
    {
        Cf a = v[0], b = v[1];
        v[0] = cf_add(a, b);
        v[1] = cf_sub(a, b);
    }
};

template <bool INV>
struct Dft<4, INV> {
    __device__ __forceinline__ static void run(Cf *v)
    {
        Cf t0 = cf_add(v[0], v[2]), t1 = cf_sub(v[0], v[2]);
        Cf t2 = cf_add(v[1], v[3]), t3 = cf_rot<INV>(cf_sub(v[1], v[3]));
        v[0] = cf_add(t0, t2);
        v[2] = cf_sub(t0, t2);
        v[1] = cf_add(t1, t3);
        v[3] = cf_sub(t1, t3);
    }
};

template <bool INV>
struct Dft<3, INV> {
    __device__ __forceinline__ static void run(Cf *v)
    {
        const float s3 = 0.866025403784438646763723170752936183f;
        Cf t = cf_add(v[1], v[2]);
        Cf d = cf_sub(v[1], v[2]);
        Cf mid{v[0].r - 0.5f * t.r, v[0].i - 0.5f * t.i};
        Cf sd = cf_rot<INV>(Cf{s3 * d.r, s3 * d.i});
        v[0] = cf_add(v[0], t);
        v[1] = cf_add(mid, sd);
        v[2] = cf_sub(mid, sd);
    }
};

template <bool INV>
struct Dft<5, INV> {
    __device__ __forceinline__ static void run(Cf *v)
    {
        const float c1 = 0.309016994374947424102293417182819059f;   // cos(2pi/5)
        const float c2 = -0.809016994374947424102293417182819059f;  // cos(4pi/5)
        const float s1 = 0.951056516295153572116439333379382143f;   // sin(2pi/5)
        const float s2 = 0.587785252292473129168705954639072769f;   // sin(4pi/5)
        Cf t1 = cf_add(v[1], v[4]), d1 = cf_sub(v[1], v[4]);
        Cf t2 = cf_add(v[2], v[3]), d2 = cf_sub(v[2], v[3]);
        Cf x0 = v[0];
        Cf a1{x0.r + c1 * t1.r + c2 * t2.r, x0.i + c1 * t1.i + c2 * t2.i};
        Cf a2{x0.r + c2 * t1.r + c1 * t2.r, x0.i + c2 * t1.i + c1 * t2.i};
        Cf b1 = cf_rot<INV>(Cf{s1 * d1.r + s2 * d2.r, s1 * d1.i + s2 * d2.i});
        Cf b2 = cf_rot<INV>(Cf{s2 * d1.r - s1 * d2.r, s2 * d1.i - s1 * d2.i});
        v[0] = Cf{x0.r + t1.r + t2.r, x0.i + t1.i + t2.i};
        v[1] = cf_add(a1, b1);
        v[4] = cf_sub(a1, b1);
        v[2] = cf_add(a2, b2);
        v[3] = cf_sub(a2, b2);
    }
};

template <bool INV>
struct Dft<8, INV> {
    __device__ __forceinline__ static void run(Cf *v)
    {
        const float h = 0.707106781186547524400844362104849039f;
        Cf e[4] = {v[0], v[2], v[4], v[6]};
        Cf o[4] = {v[1], v[3], v[5], v[7]};
        Dft<4, INV>::run(e);
        Dft<4, INV>::run(o);
        // o[k] *= W8^k
        Cf o1 = INV ? Cf{h * (o[1].r - o[1].i), h * (o[1].r + o[1].i)}
                    : Cf{h * (o[1].r + o[1].i), h * (o[1].i - o[1].r)};
        Cf o2 = cf_rot<INV>(o[2]);
        Cf o3 = INV ? Cf{-h * (o[3].r + o[3].i), h * (o[3].r - o[3].i)}
                    : Cf{h * (o[3].i - o[3].r), -h * (o[3].r + o[3].i)};
        v[0] = cf_add(e[0], o[0]);
        v[4] = cf_sub(e[0], o[0]);
        v[1] = cf_add(e[1], o1);
        v[5] = cf_sub(e[1], o1);
        v[2] = cf_add(e[2], o2);
        v[6] = cf_sub(e[2], o2);
        v[3] = cf_add(e[3], o3);
        v[7] = cf_sub(e[3], o3);
    }
};

// multiply by the constant (c + i*s)
__device__ __forceinline__ Cf cf_mulc(Cf a, float c, float s) { return Cf{a.r * c - a.i * s, a.r * s + a.i * c}; }

// 16 = 4 x 4: column DFT4s, twiddles W16^(n2*k1), row DFT4s.
template <bool INV>
struct Dft<16, INV> {
    __device__ __forceinline__ static void run(Cf *v)
    {
        const float c1 = 0.923879532511286756128183189396788933f;  // cos(pi/8)
        const float s1 = 0.382683432365089771728459984030398867f;  // sin(pi/8)
        const float h = 0.707106781186547524400844362104849039f;
        const float sg = INV ? 1.f : -1.f;  // sign of the exponent
        Cf a[4][4];
#pragma unroll
        for (int n2 = 0; n2 < 4; ++n2) {
            Cf e[4] = {v[n2], v[4 + n2], v[8 + n2], v[12 + n2]};
            Dft<4, INV>::run(e);
#pragma unroll
            for (int k1 = 0; k1 < 4; ++k1)
                a[k1][n2] = e[k1];
        }
        // W16^j = cos(2 pi j/16) + sg*i*sin(2 pi j/16) for j = n2*k1 in {1,2,3,4,6,9}
        a[1][1] = cf_mulc(a[1][1], c1, sg * s1);
        a[1][2] = cf_mulc(a[1][2], h, sg * h);
        a[1][3] = cf_mulc(a[1][3], s1, sg * c1);
        a[2][1] = cf_mulc(a[2][1], h, sg * h);
        a[2][2] = cf_rot<INV>(a[2][2]);
        a[2][3] = cf_mulc(a[2][3], -h, sg * h);
        a[3][1] = cf_mulc(a[3][1], s1, sg * c1);
        a[3][2] = cf_mulc(a[3][2], -h, sg * h);
        a[3][3] = cf_mulc(a[3][3], -c1, -sg * s1);
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            Dft<4, INV>::run(a[k1]);
#pragma unroll
            for (int k2 = 0; k2 < 4; ++k2)
                v[k1 + 4 * k2] = a[k1][k2];
        }
    }
};

// ------------------------------------------------------------------------------------
// one stage
// ------------------------------------------------------------------------------------
// ------------------------------------------------------------------------------------
// Real split / merge of one bin pair (k, N-k), 1 <= k <= N/2 (kiss_fftr.c:101-116 and
// kiss_fftr.c:135-153).  For k == N/2 the pair is one bin and the second store wins, as in
// the reference loops.
// ------------------------------------------------------------------------------------
template <class A>
__device__ __forceinline__ void split_pair(Cx<A> zk, Cx<A> znk, Cx<A> w, Cx<A> &outk, Cx<A> &outnk)
{
    typedef Ops<A> O;
    typedef Cx<A> C;
    C a = O::fixdiv(zk, 2);
    C b = O::fixdiv(C{znk.r, O::neg(znk.i)}, 2);
    C f1 = O::cadd(a, b), f2 = O::csub(a, b);
    C tw = O::cmul(f2, w);
    outk = C{O::half_add(f1.r, tw.r), O::half_add(f1.i, tw.i)};
    outnk = C{O::half_sub(f1.r, tw.r), O::half_sub(tw.i, f1.i)};
}

template <class A>
__device__ __forceinline__ void merge_pair(Cx<A> xk, Cx<A> xnk, Cx<A> w, Cx<A> &zk, Cx<A> &znk)
{
    typedef Ops<A> O;
    typedef Cx<A> C;
    C a = O::fixdiv(xk, 2);
    C b = O::fixdiv(C{xnk.r, O::neg(xnk.i)}, 2);
    C e = O::cadd(a, b), d = O::csub(a, b);
    C o = O::cmul(d, w);
    zk = O::cadd(e, o);
    C lo = O::csub(e, o);
    znk = C{lo.r, O::neg(lo.i)};
}

// Stage S of schedule RS for one thread: gather its J butterflies' inputs (from HBM for
// the first stage, else from LDS), then twiddle, transform and scatter the outputs (to
// HBM for the last stage, else to LDS).  With a single LDS buffer a barrier separates
// the gather from the scatter.
//
// PAIRED stages (the last stage of a forward real transform, the first of an inverse one)
// give thread t the butterflies t and NB-t (thread 0: 0 and NB/2), so the thread holds
// both z[k] and z[N-k] of every bin pair its outputs (inputs) touch and the real split
// (merge) runs in registers, fused with the stage.
// Orders the LDS accesses of one line's threads around a stage: when a line's T threads
// are lanes of one wavefront (T divides 64), only that wavefront touches the line, and the
// in-order LDS operations of a wavefront need no workgroup barrier.
template <int T>
__device__ __forceinline__ void line_sync()
{
    if constexpr (64 % T == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

template <class A, bool INV, int T, class RS, int S, bool PAIRED>
struct FastStage {
    typedef typename A::store2 St;
    typedef Ops<A> O;
    typedef Cx<A> C;
    static constexpr int N = RS::N();
    static constexpr int p = RS::p(S);
    static constexpr int m = RS::m(S);
    static constexpr int NB = N / p;
    static constexpr int P = N / (p * m);
    static constexpr int J = PAIRED ? 2 : (NB + T - 1) / T;
    static constexpr int PR = RS::PR(S);
    static constexpr int TWO = RS::tw_off(S);
    static_assert(!PAIRED || (NB % 2 == 0 && T * 2 == NB), "paired stage needs N/p == 2*T");

    __device__ __forceinline__ static bool live(int t, int j)
    {
        return PAIRED || (NB % T) == 0 || t + j * T < NB;
    }
    __device__ __forceinline__ static int bidx(int t, int j)
    {
        if (PAIRED)
            return j == 0 ? t : (t == 0 ? NB / 2 : NB - t);
        return t + j * T;
    }

    template <bool GSRC>
    __device__ __forceinline__ static void gather(C (&v)[J][p], const St *__restrict__ src, long long src_el, int t,
                                                  bool valid)
    {
#pragma unroll
        for (int j = 0; j < J; ++j) {
            if (!live(t, j))
                break;
            const int b = bidx(t, j);
            const int u = b / P, o = b - (b / P) * P;
#pragma unroll
            for (int r = 0; r < p; ++r) {
                const int idx = u * P * p + r * P + o;
                if (GSRC)
                    v[j][r] = valid ? O::load(hbm_ld(src + idx * src_el)) : O::zero();
                else
                    v[j][r] = O::load(src[idx * src_el]);
            }
        }
    }

    // First stage of an inverse real transform: bins from HBM, Hermitian merge into the
    // stage inputs z[r*NB + b] (m == 1, so P == NB).
    __device__ __forceinline__ static void gather_merge(C (&v)[J][p], const St *__restrict__ bins, long long el,
                                                        const St *__restrict__ rtw, int t, bool valid)
    {
        static_assert(m == 1, "the merge feeds the first stage");
        auto X = [&](int k) { return valid ? O::load(hbm_ld(bins + k * el)) : O::zero(); };
        if (t == 0) {
            // butterfly 0: z[r*NB], pairs (r, p-r); z[0] from DC and Nyquist
            {
                C x0 = X(0), xn = X(N);
                v[0][0] = O::fixdiv(C{O::add(x0.r, xn.r), O::sub(x0.r, xn.r)}, 2);
            }
#pragma unroll
            for (int r = 1; 2 * r <= p; ++r) {
                const int k = r * NB;
                C zk, znk;
                merge_pair<A>(X(k), X(N - k), O::load(rtw[k - 1]), zk, znk);
                if (2 * r != p)
                    v[0][r] = zk;
                v[0][p - r] = znk;
            }
            // butterfly NB/2: z[NB/2 + r*NB], pairs (r, p-1-r)
#pragma unroll
            for (int r = 0; 2 * r <= p - 1; ++r) {
                const int k = NB / 2 + r * NB;
                C zk, znk;
                merge_pair<A>(X(k), X(N - k), O::load(rtw[k - 1]), zk, znk);
                if (2 * r != p - 1)
                    v[1][r] = zk;
                v[1][p - 1 - r] = znk;
            }
            return;
        }
#pragma unroll
        for (int r = 0; r < p; ++r) {
            // z[k] is input r of butterfly t; z[N-k] is input p-1-r of butterfly NB-t
            const int k = r * NB + t;
            const int lo = k < N - k ? k : N - k;
            C zlo, zhi;
            merge_pair<A>(X(lo), X(N - lo), O::load(rtw[lo - 1]), zlo, zhi);
            if (k == lo) {
                v[0][r] = zlo;
                v[1][p - 1 - r] = zhi;
            } else {
                v[0][r] = zhi;
                v[1][p - 1 - r] = zlo;
            }
        }
    }

    __device__ __forceinline__ static void compute(C (&v)[J][p], const St *__restrict__ stw, int t)
    {
        // m == 1: every butterfly's twiddles are table entry 0 (see ONE in kfa_arith.hpp)
        constexpr bool ONE = A::fixed && m == 1;
#pragma unroll
        for (int j = 0; j < J; ++j) {
            if (!live(t, j))
                break;
            const int b = bidx(t, j);
            const int u = b / P;
            const St *row = stw + TWO + u * PR;
            if constexpr (IsPacked<A>::value && RS::gen(S)) {
                // kf_bfly_generic (kiss_fft.c:191-233) on packed words: C_FIXDIV, then per
                // output the packed C_MULs with the combined entries (word pairs) summed by
                // packed 16-bit adds (C_ADDTO into kiss_fft_scalar)
                const uint2 *grow = reinterpret_cast<const uint2 *>(stw) + RS::gen_off(S) + u * RS::GR(S);
                unsigned x[p];
#pragma unroll
                for (int q = 0; q < p; ++q)
                    x[q] = q15p::fixdiv<p>(v[j][q].w);
#pragma unroll
                for (int q1 = 0; q1 < p; ++q1) {
                    unsigned acc = x[0];
#pragma unroll
                    for (int q = 1; q < p; ++q) {
                        const uint2 w = grow[q1 * (p - 1) + q - 1];
                        acc = q15p::add(acc, q15p::cmul(x[q], w.x, w.y));
                    }
                    v[j][q1].w = acc;
                }
            } else if constexpr (IsPacked<A>::value) {
                // stage table of pre-arranged twiddle word pairs (see kfa_q15p.hpp)
                const uint2 *tab = reinterpret_cast<const uint2 *>(stw);
                const uint2 *trow = tab + TWO + u * PR;
                uint2 w[p > 1 ? p - 1 : 1] = {};
                if constexpr (!ONE) {
#pragma unroll
                    for (int r = 1; r < p; ++r)
                        w[r - 1] = trow[r - 1];
                }
                unsigned x[p];
#pragma unroll
                for (int r = 0; r < p; ++r)
                    x[r] = v[j][r].w;
                if constexpr (p == 2) {
                    q15p::bfly2<ONE>(x, w);
                } else if constexpr (p == 4) {
                    q15p::bfly4<INV, ONE>(x, w);
                } else if constexpr (p == 3) {
                    const uint2 ya = tab[TWO + m * PR];
                    q15p::bfly3<ONE>(x, w, q15p::scal(q15p::lo16(ya.y)));
                } else {
                    const uint2 ya = tab[TWO + m * PR], yb = tab[TWO + m * PR + 1];
                    q15p::Y5 y{q15p::scal(q15p::lo16(ya.x)), q15p::scal(q15p::lo16(ya.y)),
                               q15p::scal(q15p::lo16(yb.x)), q15p::scal(q15p::lo16(yb.y))};
                    q15p::bfly5<ONE>(x, w, y);
                }
#pragma unroll
                for (int r = 0; r < p; ++r)
                    v[j][r].w = x[r];
            } else if constexpr (A::fixed && RS::gen(S)) {
                // kf_bfly_generic (kiss_fft.c:191-233): every input divided by p, then each
                // output the first input plus the rounded products of the others with
                // their table entries, summed in kiss_fft_scalar (C_ADDTO)
                const St *grow = stw + RS::gen_off(S) + u * RS::GR(S);
                C x[p];
#pragma unroll
                for (int q = 0; q < p; ++q)
                    x[q] = O::fixdiv(v[j][q], p);
#pragma unroll
                for (int q1 = 0; q1 < p; ++q1) {
                    C acc = x[0];
#pragma unroll
                    for (int q = 1; q < p; ++q)
                        acc = O::cadd(acc, O::cmul(x[q], O::load(grow[q1 * (p - 1) + q - 1])));
                    v[j][q1] = acc;
                }
            } else if constexpr (A::fixed) {
                C w[p > 1 ? p - 1 : 1] = {};
                if constexpr (!ONE) {
#pragma unroll
                    for (int r = 1; r < p; ++r)
                        w[r - 1] = O::load(row[r - 1]);
                }
                const C ya = O::load(stw[TWO + m * PR]);
                const C yb = O::load(stw[TWO + m * PR + 1]);
                if constexpr (p == 2)
                    bfly2<A, ONE>(v[j], w);
                else if constexpr (p == 3)
                    bfly3<A, ONE>(v[j], w, ya);
                else if constexpr (p == 4)
                    bfly4<A, ONE>(v[j], w, INV);
                else
                    bfly5<A, ONE>(v[j], w, ya, yb);
                (void)yb;
            } else {
                if constexpr (m > 1) {
#pragma unroll
                    for (int r = 1; r < p; ++r)
                        v[j][r] = O::cmul(v[j][r], O::load(row[r - 1]));
                }
                Dft<p, INV>::run(v[j]);
            }
        }
    }

    template <bool GDST>
    __device__ __forceinline__ static void store(C (&v)[J][p], St *__restrict__ dst, long long dst_el, int t,
                                                 bool valid)
    {
#pragma unroll
        for (int j = 0; j < J; ++j) {
            if (!live(t, j))
                break;
            const int b = bidx(t, j);
#pragma unroll
            for (int q = 0; q < p; ++q) {
                const int idx = b + q * NB;
                if (GDST) {
                    if (valid)
                        hbm_st(dst + idx * dst_el, O::pack(v[j][q]));
                } else {
                    dst[idx * dst_el] = O::pack(v[j][q]);
                }
            }
        }
    }

    // Last stage of a forward real transform: outputs z[b + q*NB] of butterflies t and
    // NB-t, split into bins k and N-k and written to HBM.
    __device__ __forceinline__ static void store_split(C (&v)[J][p], St *__restrict__ out, long long el,
                                                       const St *__restrict__ rtw, int t, bool valid)
    {
        if (!valid)
            return;
        auto put = [&](int k, C c) { hbm_st(out + k * el, O::pack(c)); };
        if (t == 0) {
            {
                C dc = O::fixdiv(v[0][0], 2);
                put(0, C{O::add(dc.r, dc.i), O::zero().r});
                put(N, C{O::sub(dc.r, dc.i), O::zero().r});
            }
#pragma unroll
            for (int q = 1; 2 * q <= p; ++q) {
                const int k = q * NB;
                C ok, onk;
                split_pair<A>(v[0][q], v[0][p - q], O::load(rtw[k - 1]), ok, onk);
                if (2 * q != p)
                    put(k, ok);
                put(N - k, onk);
            }
#pragma unroll
            for (int q = 0; 2 * q <= p - 1; ++q) {
                const int k = NB / 2 + q * NB;
                C ok, onk;
                split_pair<A>(v[1][q], v[1][p - 1 - q], O::load(rtw[k - 1]), ok, onk);
                if (2 * q != p - 1)
                    put(k, ok);
                put(N - k, onk);
            }
            return;
        }
#pragma unroll
        for (int q = 0; q < p; ++q) {
            const int k = t + q * NB;  // output q of butterfly t; N-k is output p-1-q of NB-t
            const int lo = k < N - k ? k : N - k;
            C zlo = (k == lo) ? v[0][q] : v[1][p - 1 - q];
            C zhi = (k == lo) ? v[1][p - 1 - q] : v[0][q];
            C ok, onk;
            split_pair<A>(zlo, zhi, O::load(rtw[lo - 1]), ok, onk);
            put(lo, ok);
            put(N - lo, onk);
        }
    }
};

// Runs stage S.  Double-buffered LDS (DB): stage S reads buffer S%2 and writes the other,
// one barrier per stage.  Single buffer: gather, barrier, scatter, barrier.  With FUSE the
// real split / merge runs inside the last / first stage (paired butterflies).
template <class A, bool INV, int T, int KIND, bool DB, bool FUSE, class RS, int S>
__device__ __forceinline__ void fast_run_one(typename A::store2 *bufA, typename A::store2 *bufB,
                                             const typename A::store2 *gin, long long in_el,
                                             typename A::store2 *gout, long long out_el,
                                             const typename A::store2 *stw, const typename A::store2 *rtw, int t,
                                             bool valid)
{
    typedef typename A::store2 Sx;
    constexpr int K = RS::K;
    constexpr bool SPLIT = FUSE && KIND == FK_R2C && S == K - 1;
    constexpr bool MERGE = FUSE && KIND == FK_C2R && S == 0;
    typedef FastStage<A, INV, T, RS, S, SPLIT || MERGE> F;
    constexpr bool GSRC = (S == 0) && (KIND != FK_C2R || MERGE);
    constexpr bool GDST = (S == K - 1) && (KIND != FK_R2C || SPLIT);
    const Sx *src = GSRC ? gin : ((!DB || S % 2 == 0) ? bufA : bufB);
    Sx *dst = GDST ? gout : ((!DB || S % 2 == 1) ? bufA : bufB);
    typename F::C v[F::J][F::p];
    if constexpr (MERGE)
        F::gather_merge(v, gin, in_el, rtw, t, valid);
    else
        F::template gather<GSRC>(v, src, GSRC ? in_el : 1, t, valid);
    if constexpr (!DB && !GSRC && !GDST)
        line_sync<T>();  // every input read before the buffer is overwritten
    F::compute(v, stw, t);
    if constexpr (SPLIT)
        F::store_split(v, gout, out_el, rtw, t, valid);
    else
        F::template store<GDST>(v, dst, GDST ? out_el : 1, t, valid);
    if constexpr (!GDST)
        line_sync<T>();
}

// Stages 0 and 1 without the LDS exchange between them.  Thread t owns the stage-0
// butterflies b = t + j*T, whose outputs sit at b + q*NB0; its stage-1 butterflies
// b' = t + j'*T read u*P1*p1 + r*P1 + o with u = b' / P1, o = b' % P1.  When T divides NB0
// and P1, u = j'*T / P1 and o = (j'*T) % P1 + t, so every stage-1 input of the thread is
// the output (j, q) = ((c % NB0) / T, c / NB0), c = u*P1*p1 + r*P1 + (j'*T) % P1, of one
// of its own stage-0 butterflies: a permutation of registers fixed at compile time.  The
// arithmetic of both stages is unchanged; one LDS store, one LDS gather and a barrier go.
// Used by the fixed-point complex kernels (bound by instruction issue), whose 512-point
// reference schedule (2, 4, 4, 4, 4) meets the condition at 16, 32 and 64 threads per line.
template <class A, bool INV, int T, int KIND, bool FUSE, class RS>
struct FirstPairInRegisters {
    typedef FastStage<A, INV, T, RS, 0, false> F0;
    typedef FastStage<A, INV, T, RS, (RS::K > 1 ? 1 : 0), false> F1;
    static constexpr bool value = A::fixed && KIND == FK_C2C && !FUSE && RS::K > 2 &&
                                  F0::NB % T == 0 && F1::P % T == 0 && F1::NB % T == 0;
};

template <class A, bool INV, int T, int KIND, bool DB, class RS>
__device__ __forceinline__ void fast_run_first_pair(typename A::store2 *bufA, typename A::store2 *bufB,
                                                    const typename A::store2 *gin, long long in_el,
                                                    const typename A::store2 *stw, int t, bool valid)
{
    typedef FastStage<A, INV, T, RS, 0, false> F0;
    typedef FastStage<A, INV, T, RS, 1, false> F1;
    typename F0::C w[F0::J][F0::p];
    F0::template gather<true>(w, gin, in_el, t, valid);
    F0::compute(w, stw, t);
    typename F1::C v[F1::J][F1::p];
#pragma unroll
    for (int j1 = 0; j1 < F1::J; ++j1) {
#pragma unroll
        for (int r = 0; r < F1::p; ++r) {
            const int c = (j1 * T / F1::P) * F1::P * F1::p + r * F1::P + (j1 * T) % F1::P;
            v[j1][r] = w[(c % F0::NB) / T][c / F0::NB];
        }
    }
    F1::compute(v, stw, t);
    // stage 1 writes where it always does: bufA, single-buffered or double-buffered (odd
    // stages write bufA), which stage 2 reads; stage 1 is not the last stage (RS::K > 2)
    F1::template store<false>(v, bufA, 1, t, valid);
    (void)bufB;
    line_sync<T>();
}

template <class A, bool INV, int T, int KIND, bool DB, bool FUSE, class RS, int S, int FROM>
__device__ __forceinline__ void fast_run_from(typename A::store2 *bufA, typename A::store2 *bufB,
                                              const typename A::store2 *gin, long long in_el,
                                              typename A::store2 *gout, long long out_el,
                                              const typename A::store2 *stw, const typename A::store2 *rtw, int t,
                                              bool valid)
{
    if constexpr (S >= FROM)
        fast_run_one<A, INV, T, KIND, DB, FUSE, RS, S>(bufA, bufB, gin, in_el, gout, out_el, stw, rtw, t, valid);
}

template <class A, bool INV, int T, int KIND, bool DB, bool FUSE, class RS, int... S>
__device__ __forceinline__ void fast_run_all(std::integer_sequence<int, S...>, typename A::store2 *bufA,
                                             typename A::store2 *bufB, const typename A::store2 *gin,
                                             long long in_el, typename A::store2 *gout, long long out_el,
                                             const typename A::store2 *stw, const typename A::store2 *rtw, int t,
                                             bool valid)
{
    if constexpr (FirstPairInRegisters<A, INV, T, KIND, FUSE, RS>::value) {
        fast_run_first_pair<A, INV, T, KIND, DB, RS>(bufA, bufB, gin, in_el, stw, t, valid);
        (fast_run_from<A, INV, T, KIND, DB, FUSE, RS, S, 2>(bufA, bufB, gin, in_el, gout, out_el, stw, rtw, t, valid),
         ...);
    } else {
        (fast_run_one<A, INV, T, KIND, DB, FUSE, RS, S>(bufA, bufB, gin, in_el, gout, out_el, stw, rtw, t, valid),
         ...);
    }
}

// buffer holding the result of the last stage (input of the real split)
template <bool DB, int K>
struct LastBuf {
    static constexpr bool is_a = !DB || (K % 2 == 0);
};

// The kernel body for workgroup `blk` of the launch (blockIdx.x of k_fast; a fused launch
// of several groups passes the workgroup's index within its group); UNIT: both element
// strides are 1.
template <class A, int KIND, bool INV, int T, int LPB, bool DB, bool FUSE, class RS, bool UNIT>
__device__ __forceinline__ void k_fast_body(const FastArgs &a, long long blk)
{
    typedef typename A::store2 S;
    typedef Ops<A> O;
    typedef Cx<A> C;
    constexpr int N = RS::N();
    constexpr int K = RS::K;
    extern __shared__ __align__(16) unsigned char kfa_fast_smem[];
    const int line = threadIdx.x / T;
    const int t = threadIdx.x - line * T;
    const long long gl = blk * LPB + line;
    const bool valid = gl < a.nlines;
    S *bufA = reinterpret_cast<S *>(kfa_fast_smem) + line * N;
    S *bufB = DB ? reinterpret_cast<S *>(kfa_fast_smem) + (LPB + line) * N : bufA;
    const S *gin;
    S *gout;
    line_base<S>(a, valid ? gl : 0, gin, gout);
    const S *stw = reinterpret_cast<const S *>(__builtin_assume_aligned(a.stw, 16));
    const S *rtw = reinterpret_cast<const S *>(a.rtw);
    const long long ie = UNIT ? 1 : a.in_el, oe = UNIT ? 1 : a.out_el;  // element strides

    if constexpr (KIND == FK_C2R && !FUSE) {
        // Hermitian merge into bufA (kiss_fftr.c:131-153); bins read from HBM.  Every load
        // is issued first (indices clamped into the row), then the pairs are combined:
        // loads behind the loop's exit test would wait one iteration at a time.
        constexpr int NK = N / 2 + 1;
        constexpr int JK = (NK + T - 1) / T;
        C xa[JK], bz[JK], tw[JK];
#pragma unroll
        for (int j = 0; j < JK; ++j) {
            const int k = t + j * T;
            const int kc = k < NK ? k : NK - 1;
            xa[j] = valid ? O::load(hbm_ld(gin + (long long)kc * ie)) : O::zero();
            bz[j] = valid ? O::load(hbm_ld(gin + (long long)(N - kc) * ie)) : O::zero();
            tw[j] = O::load(rtw[kc > 0 ? kc - 1 : 0]);
        }
#pragma unroll
        for (int j = 0; j < JK; ++j) {
            const int k = t + j * T;
            if (k >= NK)
                break;
            if (k == 0) {  // DC and Nyquist: xa = X[0], bz = X[N]
                C d{O::add(xa[j].r, bz[j].r), O::sub(xa[j].r, bz[j].r)};
                bufA[0] = O::pack(O::fixdiv(d, 2));
                continue;
            }
            C xb{bz[j].r, O::neg(bz[j].i)};
            const C x1 = O::fixdiv(xa[j], 2);
            xb = O::fixdiv(xb, 2);
            C e = O::cadd(x1, xb), d = O::csub(x1, xb);
            C o = O::cmul(d, tw[j]);
            C lo = O::csub(e, o);
            lo.i = O::neg(lo.i);
            if (2 * k != N)
                bufA[k] = O::pack(O::cadd(e, o));
            bufA[N - k] = O::pack(lo);
        }
        line_sync<T>();
    }

    fast_run_all<A, INV, T, KIND, DB, FUSE, RS>(std::make_integer_sequence<int, K>{}, bufA, bufB, gin, ie,
                                                gout, oe, stw, rtw, t, valid);

    if constexpr (KIND == FK_R2C && !FUSE) {
        // split (kiss_fftr.c:88-116) from the LDS buffer the last stage wrote
        const S *z = LastBuf<DB, K>::is_a ? bufA : bufB;
        constexpr int NK = N / 2 + 1;
        constexpr int JK = (NK + T - 1) / T;
        if (!valid)
            return;
        // operands first (LDS pairs and split twiddles, indices clamped), then the stores
        C za[JK], zb[JK], tw[JK];
#pragma unroll
        for (int j = 0; j < JK; ++j) {
            const int k = t + j * T;
            const int kc = k < NK ? k : NK - 1;
            za[j] = O::load(z[kc]);
            zb[j] = O::load(z[kc == 0 ? 0 : N - kc]);
            tw[j] = O::load(rtw[kc > 0 ? kc - 1 : 0]);
        }
#pragma unroll
        for (int j = 0; j < JK; ++j) {
            const int k = t + j * T;
            if (k >= NK)
                break;
            if (k == 0) {
                C dc = O::fixdiv(za[j], 2);
                hbm_st(gout, O::pack(C{O::add(dc.r, dc.i), O::zero().r}));
                hbm_st(gout + (long long)N * oe, O::pack(C{O::sub(dc.r, dc.i), O::zero().r}));
                continue;
            }
            C xa = O::fixdiv(za[j], 2);
            C xb = O::fixdiv(C{zb[j].r, O::neg(zb[j].i)}, 2);
            C f1 = O::cadd(xa, xb), f2 = O::csub(xa, xb);
            C w = O::cmul(f2, tw[j]);
            if (2 * k != N)
                hbm_st(gout + (long long)k * oe, O::pack(C{O::half_add(f1.r, w.r), O::half_add(f1.i, w.i)}));
            hbm_st(gout + (long long)(N - k) * oe, O::pack(C{O::half_sub(f1.r, w.r), O::half_sub(w.i, f1.i)}));
        }
    }
}

// Dynamic LDS: (DB ? 2 : 1) * LPB * N complex values.
// Fixed point: contiguous lines (both element strides 1, every call but kiss_fft_stride
// and strided batches) take an instance whose element offsets are compile-time multiples
// of the element size, folded into the loads' and stores' immediate offsets; any other
// stride multiplies 64-bit indices per element.  The fixed-point kernels are bound by
// vector-ALU issue (Q15 c2c 512: +9-11 % from the unit instance, Q31 +4 %); the float kernels
// stream from HBM and ran 1-3 % slower with the second instance (more registers), so they
// keep the one general body.
template <class A, int KIND, bool INV, int T, int LPB, bool DB, bool FUSE, class RS>
__global__ void __launch_bounds__(T *LPB) k_fast(FastArgs a)
{
    if constexpr (A::fixed) {
        if (a.in_el == 1 && a.out_el == 1) {
            k_fast_body<A, KIND, INV, T, LPB, DB, FUSE, RS, true>(a, blockIdx.x);
            return;
        }
    }
    k_fast_body<A, KIND, INV, T, LPB, DB, FUSE, RS, false>(a, blockIdx.x);
}

}  // namespace kfa
