// kfa_plane.hpp -- the plane kernel: the last two axes of an N-D transform in one HBM pass.
//
// kiss_fftnd transforms one axis at a time (kiss_fftnd.c:150-188), one pass over the whole
// array per axis.  When a plane of the last two axes (A x B complex values, contiguous in
// HBM) fits in one CU's LDS, a workgroup can do both axes of a plane between one load and
// one store: rows of B points, then columns of A points, all in LDS.  An array of d axes
// then costs d-1 HBM passes instead of d (64 x 64 x 64: 2 instead of 3).
//
// A "plane" may also be a slab of two non-adjacent axes: columns along an outer axis k,
// rows along the last (contiguous) axis, the axes between them fixed per plane.  Plane p
// then starts at (p / nsub) * po + (p % nsub) * pi and its rows are rs values apart (the
// last two axes: nsub = 1, po = A * B, rs = B).  Every row is still a contiguous run of B
// values, so loads and stores stay coalesced; a 3-axis array then costs two passes with a
// slab of axes 0 and 2 when the plane of axes 1 and 2 does not fit in LDS.
//
// Layout in LDS: row r at r*RSX (RSX = B+1 when B is even, so that a column's elements
// fall in different banks).  Row stages use TB threads per row and unit element stride;
// column stages TA threads per column and element stride RSX.  Both use the fast kernels'
// stage code (FastStage), so the butterflies and twiddle tables are those of the line
// kernels of lengths B and A.
#pragma once

#include "kfa_fast.hpp"

namespace kfa {

struct PlaneArgs {
    const void *in;
    void *out;
    const void *stw_a;  // stage table of the column schedule (length A)
    const void *stw_b;  // stage table of the row schedule (length B)
    long long planes;
    int rs;             // values between the starts of two rows of a plane ((A - 1) * rs + B
                        // must stay below 2^31)
    long long nsub;     // plane p starts at (p / nsub) * po + (p % nsub) * pi
    long long po;
    long long pi;
};

// the arguments of contiguous A x B planes
inline PlaneArgs plane_args(const void *in, void *out, const void *stw_a, const void *stw_b, long long planes,
                            long long A, long long B)
{
    return PlaneArgs{in, out, stw_a, stw_b, planes, (int)B, 1, A * B, 0};
}

template <bool INV, int T, class RS, int ES, int S>
__device__ __forceinline__ void plane_stage(F32::store2 *lb, const F32::store2 *stw, int t, bool active)
{
    typedef FastStage<F32, INV, T, RS, S, false> F;
    typename F::C v[F::J][F::p];
    if (active)
        F::template gather<false>(v, lb, ES, t, true);
    line_sync<T>();  // every input read before the buffer is overwritten
    if (active) {
        F::compute(v, stw, t);
        F::template store<false>(v, lb, ES, t, true);
    }
    line_sync<T>();
}

template <bool INV, int T, class RS, int ES, int... S>
__device__ __forceinline__ void plane_stages(std::integer_sequence<int, S...>, F32::store2 *lb,
                                             const F32::store2 *stw, int t, bool active)
{
    (plane_stage<INV, T, RS, ES, S>(lb, stw, t, active), ...);
}

// Persistent: workgroup g takes planes g, g + G, g + 2G, ... (G = gridDim.x).  PF: the
// next plane's loads are issued into registers as soon as the current plane is in LDS, so
// they are in flight while the current plane is transformed (for planes that leave room
// for one workgroup per CU); without PF a plane is loaded when its turn comes, which saves
// the registers of the prefetch so that two workgroups fit a CU and overlap each other.
// The stage tables are copied to LDS once, so that the stages issue no global loads (a
// global load's wait would also wait for the prefetch).
// SLAB: planes addressed through rs / nsub / po / pi (else contiguous A x B planes, p * A * B)
template <bool INV, int TA, int TB, int NTH, class RA, class RB, bool SLAB, bool PF>
__global__ void __launch_bounds__(NTH) k_plane(PlaneArgs a)
{
    typedef F32::store2 St;
    constexpr int A = RA::N(), B = RB::N();
    constexpr int RSX = (B % 2 == 0) ? B + 1 : B;
    constexpr int E = A * B;
    constexpr int IT = (E + NTH - 1) / NTH;
    constexpr int TWA = RA::tw_len(), TWB = RB::tw_len();
    static_assert(A * TB <= NTH && B * TA <= NTH, "threads per plane");
    extern __shared__ __align__(16) unsigned char kfa_plane_smem[];
    St *buf = reinterpret_cast<St *>(kfa_plane_smem);
    St *twa = buf + A * RSX;  // stage tables (after the plane)
    St *twb = twa + TWA;
    const int tid = threadIdx.x;
    const St *in = reinterpret_cast<const St *>(a.in);
    St *out = reinterpret_cast<St *>(a.out);

    for (int i = tid; i < TWA; i += NTH)
        twa[i] = reinterpret_cast<const St *>(a.stw_a)[i];
    for (int i = tid; i < TWB; i += NTH)
        twb[i] = reinterpret_cast<const St *>(a.stw_b)[i];

    long long pl = blockIdx.x;
    St v[IT];
    const int rs = SLAB ? a.rs : B;
    auto base = [&](long long p) { return SLAB ? (p / a.nsub) * a.po + (p % a.nsub) * a.pi : p * E; };
    auto at = [&](int e) { return SLAB ? (e / B) * rs + e % B : e; };
    auto load = [&](long long p) {
        const St *src = in + base(p);
#pragma unroll
        for (int k = 0; k < IT; ++k) {
            const int e = tid + k * NTH;
            if (E % NTH == 0 || e < E)
                v[k] = src[at(e)];
        }
    };
    if (PF && pl < a.planes)
        load(pl);
    for (; pl < a.planes; pl += gridDim.x) {
        if (!PF)
            load(pl);
#pragma unroll
        for (int k = 0; k < IT; ++k) {
            const int e = tid + k * NTH;
            if (E % NTH == 0 || e < E)
                buf[(e / B) * RSX + e % B] = v[k];
        }
        __syncthreads();
        if (PF && pl + gridDim.x < a.planes)
            load(pl + gridDim.x);  // in flight during the stages below
        {   // rows: B points, TB threads each
            const int line = tid / TB;
            plane_stages<INV, TB, RB, 1>(std::make_integer_sequence<int, RB::K>{},
                                         buf + (line < A ? line : 0) * RSX, twb, tid - line * TB, line < A);
        }
        if constexpr (64 % TB == 0)
            __syncthreads();  // every row done before the columns read them
        {   // columns: A points at stride RSX, TA threads each
            const int line = tid / TA;
            plane_stages<INV, TA, RA, RSX>(std::make_integer_sequence<int, RA::K>{}, buf + (line < B ? line : 0),
                                           twa, tid - line * TA, line < B);
        }
        if constexpr (64 % TA == 0)
            __syncthreads();  // every column done before the plane is read out
        {
            St *dst = out + base(pl);
            St w[IT];
#pragma unroll
            for (int k = 0; k < IT; ++k) {
                const int e = tid + k * NTH;
                if (E % NTH == 0 || e < E)
                    w[k] = buf[(e / B) * RSX + e % B];
            }
#pragma unroll
            for (int k = 0; k < IT; ++k) {
                const int e = tid + k * NTH;
                if (E % NTH == 0 || e < E)
                    dst[at(e)] = w[k];
            }
        }
        __syncthreads();  // the plane is read out before the next one is written
    }
}

}  // namespace kfa
