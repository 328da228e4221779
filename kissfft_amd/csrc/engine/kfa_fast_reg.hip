// kfa_fast_reg.hip -- the instantiated fast-kernel variants and their launchers.
//
// Each variant is (scalar kind, transform kind, line length, direction, radix schedule,
// threads per line, lines per workgroup, LDS buffering).  Fixed-point variants carry the
// reference radix schedule of their length (kf_factor order, executed innermost first) and
// are used only when the plan's schedule equals it; float variants may use any schedule.
// When several variants serve the same key the engine times them on first use and keeps
// the fastest; a later variant must win by a margin, so the first registered of a key --
// the fastest in full-size sweeps -- is kept on near-ties.
#include <hip/hip_runtime.h>

#include <mutex>

#include "kfa_fast.hpp"
#include "kfa_fast_reg.hpp"
#include "kfa_q15w.hpp"

namespace kfa {

template <class A, int KIND, bool INV, int T, int LPB, bool DB, bool FUSE, class RS>
static int launch_fast(const FastArgs &a, hipStream_t s)
{
    typedef typename A::store2 S;
    const size_t lds = (DB ? 2 : 1) * (size_t)LPB * RS::N() * sizeof(S);
    static std::once_flag once;
    std::call_once(once, [=] {
        (void)hipFuncSetAttribute((const void *)k_fast<A, KIND, INV, T, LPB, DB, FUSE, RS>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    });
    const long long blocks = (a.nlines + LPB - 1) / LPB;
    if (blocks <= 0)
        return 0;
    hipLaunchKernelGGL((k_fast<A, KIND, INV, T, LPB, DB, FUSE, RS>), dim3((unsigned)blocks), dim3(T * LPB), lds, s,
                       a);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

// twiddle-row granularity: 16-byte aligned rows
template <class A>
struct TwPad {
    static constexpr int G = 16 / (int)sizeof(typename A::store2);
    static constexpr int fmt = 0;
};
template <>
struct TwPad<Q15P> {
    static constexpr int G = 2;  // 8-byte word-pair entries
    static constexpr int fmt = 1;
};

// the arithmetic's name in kernel source (for fused launches compiled at run time)
template <class A>
struct ArithName;
template <>
struct ArithName<F32> {
    static constexpr const char *value = "kfa::F32";
};
template <>
struct ArithName<Q15> {
    static constexpr const char *value = "kfa::Q15";
};
template <>
struct ArithName<Q15P> {
    static constexpr const char *value = "kfa::Q15P";
};
template <>
struct ArithName<Q31> {
    static constexpr const char *value = "kfa::Q31";
};

template <class A, int KIND, int T, int LPB, bool DB, bool FUSE, int... R>
static void add_variant(std::vector<FastVariant> &v, int type, const char *name)
{
    constexpr int G = TwPad<A>::G;
    typedef Radices<G, R...> RS;
    static_assert(RS::K >= 2, "single-stage schedules read and write HBM in one stage (not in-place safe)");
    static_assert(T * LPB <= 1024, "workgroup too large");
    for (int inv = 0; inv < 2; ++inv) {
        if (KIND == FK_R2C && inv)
            continue;
        if (KIND == FK_C2R && !inv)
            continue;
        FastVariant f;
        f.type = type;
        f.kind = KIND;
        f.n = RS::N();
        f.inverse = inv;
        f.T = T;
        f.LPB = LPB;
        f.radices = {R...};
        f.tw_pad = G;
        f.tw_fmt = TwPad<A>::fmt;
        f.tw_len = RS::tw_len();
        f.name = name;
        f.arith = ArithName<A>::value;
        f.db = DB;
        f.fuse = FUSE;
        f.launch = inv ? &launch_fast<A, KIND, true, T, LPB, DB, FUSE, RS>
                       : &launch_fast<A, KIND, false, T, LPB, DB, FUSE, RS>;
        v.push_back(f);
    }
}

template <bool INV, int T, int LPB, class RS>
static int launch_wave(const FastArgs &a, hipStream_t s)
{
    const long long blocks = (a.nlines + LPB - 1) / LPB;
    if (blocks <= 0)
        return 0;
    hipLaunchKernelGGL((k_q15w<INV, T, LPB, RS>), dim3((unsigned)blocks), dim3(T * LPB), 0, s, a);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

// Q15 complex lines held by one wavefront (kfa_q15w.hpp): stage exchanges by lane swaps,
// no LDS; radix 2 / 4 power-of-two schedules
template <int T, int LPB, int... R>
static void add_wave(std::vector<FastVariant> &v, const char *name)
{
    typedef Radices<TwPad<Q15P>::G, R...> RS;
    static_assert(T * LPB <= 1024, "workgroup too large");
    for (int inv = 0; inv < 2; ++inv) {
        FastVariant f;
        f.type = KFA_Q15;
        f.kind = FK_C2C;
        f.n = RS::N();
        f.inverse = inv;
        f.T = T;
        f.LPB = LPB;
        f.radices = {R...};
        f.tw_pad = TwPad<Q15P>::G;
        f.tw_fmt = TwPad<Q15P>::fmt;
        f.tw_len = RS::tw_len();
        f.name = name;
        f.launch = inv ? &launch_wave<true, T, LPB, RS> : &launch_wave<false, T, LPB, RS>;
        v.push_back(f);
    }
}

template <bool INV, int LPB, class RS, int WPE>
static int launch_xpose(const FastArgs &a, hipStream_t s)
{
    const long long blocks = (a.nlines + LPB - 1) / LPB;
    if (blocks <= 0)
        return 0;
    hipLaunchKernelGGL((k_q15x<INV, LPB, RS, WPE>), dim3((unsigned)blocks), dim3(16 * LPB), 0, s, a);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

// Q15 complex lines of 16 lanes with one LDS transpose in place of the lane swaps
// (k_q15x, kfa_q15w.hpp); radix 2 / 4 schedules of 256 or 512 points; WPE: wavefronts per
// SIMD the registers must allow (1: the compiler's choice)
template <int LPB, int WPE, int... R>
static void add_xpose(std::vector<FastVariant> &v, const char *name)
{
    typedef Radices<TwPad<Q15P>::G, R...> RS;
    static_assert(16 * LPB <= 1024, "workgroup too large");
    for (int inv = 0; inv < 2; ++inv) {
        FastVariant f;
        f.type = KFA_Q15;
        f.kind = FK_C2C;
        f.n = RS::N();
        f.inverse = inv;
        f.T = 16;
        f.LPB = LPB;
        f.radices = {R...};
        f.tw_pad = TwPad<Q15P>::G;
        f.tw_fmt = TwPad<Q15P>::fmt;
        f.tw_len = RS::tw_len();
        f.name = name;
        f.arith = ArithName<Q15P>::value;
        f.xpose = true;
        f.launch = inv ? &launch_xpose<true, LPB, RS, WPE> : &launch_xpose<false, LPB, RS, WPE>;
        v.push_back(f);
    }
}

// forward and inverse real variants of one schedule, split / merge in a separate LDS pass
template <class A, int T, int LPB, bool DB, int... R>
static void add_real(std::vector<FastVariant> &v, int type, const char *name)
{
    add_variant<A, FK_R2C, T, LPB, DB, false, R...>(v, type, name);
    add_variant<A, FK_C2R, T, LPB, DB, false, R...>(v, type, name);
}

// a real variant with the split (forward) / merge (inverse) fused into the last / first
// stage; N / (that stage's radix) must equal 2*T
template <class A, int KIND, int T, int LPB, int... R>
static void add_fused(std::vector<FastVariant> &v, int type, const char *name)
{
    add_variant<A, KIND, T, LPB, false, true, R...>(v, type, name);
}

const std::vector<FastVariant> &fast_registry()
{
    static std::vector<FastVariant> *reg = [] {
        auto *v = new std::vector<FastVariant>();
        // ---- float real transforms; the line is the half length ----
        // N = 4096 (half 2048)
        // (full-size sweep, 65536 lines: t128x2_sb 5.0 TB/s both directions, the others
        // 4.4 - 4.8)
        add_real<F32, 128, 2, false, 16, 16, 8>(*v, KFA_F32, "real2048_16168_t128x2_sb");
        add_real<F32, 128, 1, false, 16, 16, 8>(*v, KFA_F32, "real2048_16168_t128_sb");
        add_real<F32, 256, 1, false, 8, 8, 8, 4>(*v, KFA_F32, "real2048_8884_t256_sb");
        add_real<F32, 256, 1, true, 8, 8, 8, 4>(*v, KFA_F32, "real2048_8884_t256_db");
        add_real<F32, 128, 1, true, 16, 16, 8>(*v, KFA_F32, "real2048_16168_t128_db");
        add_real<F32, 128, 1, false, 8, 16, 16>(*v, KFA_F32, "real2048_81616_t128_sb");
        // N = 4000 (half 2000); sweep, 67092 lines: t128x2_sb 4.88 / 4.82 TB/s (whole
        // waves: lines of 125 or 250 threads straddle waves), t125x2_sb 4.13 / 4.34,
        // t250_sb 4.35 / 4.01
        add_real<F32, 128, 2, false, 16, 5, 5, 5>(*v, KFA_F32, "real2000_16555_t128x2_sb");
        add_real<F32, 256, 1, false, 16, 5, 5, 5>(*v, KFA_F32, "real2000_16555_t256_sb");
        add_real<F32, 125, 2, false, 16, 5, 5, 5>(*v, KFA_F32, "real2000_16555_t125x2_sb");
        add_real<F32, 250, 1, false, 8, 2, 5, 5, 5>(*v, KFA_F32, "real2000_82555_t250_sb");
        add_real<F32, 100, 2, true, 4, 4, 5, 5, 5>(*v, KFA_F32, "real2000_44555_t100x2_db");
        add_real<F32, 100, 2, false, 4, 4, 5, 5, 5>(*v, KFA_F32, "real2000_44555_t100x2_sb");
        add_real<F32, 250, 1, false, 5, 5, 5, 4, 4>(*v, KFA_F32, "real2000_55544_t250_sb");
        add_fused<F32, FK_R2C, 128, 1, 16, 16, 8>(*v, KFA_F32, "r2c2048_16168_t128_fused");
        add_fused<F32, FK_C2R, 128, 1, 8, 16, 16>(*v, KFA_F32, "c2r2048_81616_t128_fused");
        add_fused<F32, FK_R2C, 128, 2, 16, 16, 8>(*v, KFA_F32, "r2c2048_16168_t128x2_fused");
        add_fused<F32, FK_C2R, 128, 2, 8, 16, 16>(*v, KFA_F32, "c2r2048_81616_t128x2_fused");
        add_fused<F32, FK_R2C, 256, 1, 8, 8, 8, 4>(*v, KFA_F32, "r2c2048_8884_t256_fused");
        add_fused<F32, FK_C2R, 256, 1, 4, 8, 8, 8>(*v, KFA_F32, "c2r2048_4888_t256_fused");
        add_fused<F32, FK_R2C, 250, 1, 5, 5, 5, 4, 4>(*v, KFA_F32, "r2c2000_55544_t250_fused");
        add_fused<F32, FK_C2R, 250, 1, 4, 4, 5, 5, 5>(*v, KFA_F32, "c2r2000_44555_t250_fused");
        add_fused<F32, FK_R2C, 200, 1, 16, 5, 5, 5>(*v, KFA_F32, "r2c2000_16555_t200_fused");
        add_fused<F32, FK_C2R, 200, 1, 5, 5, 5, 16>(*v, KFA_F32, "c2r2000_55516_t200_fused");
        add_fused<F32, FK_R2C, 200, 1, 8, 2, 5, 5, 5>(*v, KFA_F32, "r2c2000_82555_t200_fused");
        add_fused<F32, FK_C2R, 200, 1, 5, 5, 5, 2, 8>(*v, KFA_F32, "c2r2000_55528_t200_fused");
        // N = 1000 (half 500), the real axis of the 2-D workload
        // (sweep, 268167 lines: t64x4 4.70 / 4.58 TB/s, t50x4 4.02 / 3.68 -- whole waves)
        add_real<F32, 64, 4, false, 4, 5, 5, 5>(*v, KFA_F32, "real500_4555_t64x4_sb");
        add_real<F32, 64, 2, false, 4, 5, 5, 5>(*v, KFA_F32, "real500_4555_t64x2_sb");
        add_real<F32, 50, 4, false, 4, 5, 5, 5>(*v, KFA_F32, "real500_4555_t50x4_sb");
        add_real<F32, 125, 2, false, 4, 5, 5, 5>(*v, KFA_F32, "real500_4555_t125x2_sb");
        add_fused<F32, FK_R2C, 50, 4, 4, 5, 5, 5>(*v, KFA_F32, "r2c500_4555_t50x4_fused");
        add_fused<F32, FK_C2R, 50, 4, 5, 5, 5, 4>(*v, KFA_F32, "c2r500_5554_t50x4_fused");
        // N = 256 ... 2048 (halves 128 ... 1024), the shapes of the complex variants of the
        // half lengths: common real sizes start without a runtime compile
        add_real<F32, 16, 16, false, 8, 16>(*v, KFA_F32, "real128_816_t16x16_sb");
        add_real<F32, 16, 16, false, 16, 16>(*v, KFA_F32, "real256_1616_t16x16_sb");
        add_real<F32, 64, 4, false, 8, 8, 8>(*v, KFA_F32, "real512_888_t64x4_sb");
        add_real<F32, 64, 2, false, 16, 16, 4>(*v, KFA_F32, "real1024_16164_t64x2_sb");
        // N = 3072 (half 1536), a size of the serving mix
        add_real<F32, 192, 1, false, 8, 8, 8, 3>(*v, KFA_F32, "real1536_8883_t192_sb");
        add_real<F32, 96, 2, false, 16, 16, 3, 2>(*v, KFA_F32, "real1536_161632_t96x2_sb");
        // ---- float complex ----
        add_variant<F32, FK_C2C, 240, 1, true, false, 4, 4, 3, 5, 5>(*v, KFA_F32, "c2c1200_44355_t240_db");
        add_variant<F32, FK_C2C, 240, 1, false, false, 4, 4, 3, 5, 5>(*v, KFA_F32, "c2c1200_44355_t240_sb");
        add_variant<F32, FK_C2C, 120, 2, false, false, 5, 5, 3, 16>(*v, KFA_F32, "c2c1200_55316_t120x2_sb");
        add_variant<F32, FK_C2C, 128, 1, false, false, 16, 16, 8>(*v, KFA_F32, "c2c2048_16168_t128_sb");
        add_variant<F32, FK_C2C, 256, 1, false, false, 8, 8, 8, 4>(*v, KFA_F32, "c2c2048_8884_t256_sb");
        add_variant<F32, FK_C2C, 64, 2, false, false, 16, 16, 4>(*v, KFA_F32, "c2c1024_16164_t64x2_sb");
        add_variant<F32, FK_C2C, 128, 2, false, false, 8, 8, 4, 4>(*v, KFA_F32, "c2c1024_8844_t128x2_sb");
        add_variant<F32, FK_C2C, 64, 4, false, false, 8, 8, 8>(*v, KFA_F32, "c2c512_888_t64x4_sb");
        add_variant<F32, FK_C2C, 16, 16, false, false, 16, 16>(*v, KFA_F32, "c2c256_1616_t16x16_sb");
        add_variant<F32, FK_C2C, 64, 4, false, false, 4, 4, 4, 4>(*v, KFA_F32, "c2c256_4444_t64x4_sb");
        // whole waves per line first (sweeps: c2c1000 t64x4 5.06 TB/s, t125x2 4.67;
        // c2c120 t16x16 5.21, t40x8 3.79; c2c500 t64x4 4.97, t125x2 4.87)
        add_variant<F32, FK_C2C, 64, 4, false, false, 5, 5, 5, 8>(*v, KFA_F32, "c2c1000_5558_t64x4_sb");
        add_variant<F32, FK_C2C, 128, 2, false, false, 5, 5, 5, 8>(*v, KFA_F32, "c2c1000_5558_t128x2_sb");
        add_variant<F32, FK_C2C, 125, 2, false, false, 5, 5, 5, 8>(*v, KFA_F32, "c2c1000_5558_t125x2_sb");
        add_variant<F32, FK_C2C, 100, 2, false, false, 5, 5, 5, 4, 2>(*v, KFA_F32, "c2c1000_55542_t100x2_sb");
        // N-D axis lengths (contiguous last axes; the non-last axes run as tiles)
        add_variant<F32, FK_C2C, 16, 16, false, false, 8, 5, 3>(*v, KFA_F32, "c2c120_853_t16x16_sb");
        add_variant<F32, FK_C2C, 24, 8, false, false, 8, 5, 3>(*v, KFA_F32, "c2c120_853_t24x8_sb");
        add_variant<F32, FK_C2C, 32, 8, false, false, 8, 5, 3>(*v, KFA_F32, "c2c120_853_t32x8_sb");
        add_variant<F32, FK_C2C, 40, 8, false, false, 8, 3, 5>(*v, KFA_F32, "c2c120_835_t40x8_sb");
        add_variant<F32, FK_C2C, 16, 16, false, false, 8, 4, 3>(*v, KFA_F32, "c2c96_843_t16x16_sb");
        add_variant<F32, FK_C2C, 8, 32, false, false, 8, 8>(*v, KFA_F32, "c2c64_88_t8x32_sb");
        add_variant<F32, FK_C2C, 16, 16, false, false, 8, 16>(*v, KFA_F32, "c2c128_816_t16x16_sb");
        add_variant<F32, FK_C2C, 64, 4, false, false, 5, 5, 5, 4>(*v, KFA_F32, "c2c500_5554_t64x4_sb");
        add_variant<F32, FK_C2C, 128, 2, false, false, 5, 5, 5, 4>(*v, KFA_F32, "c2c500_5554_t128x2_sb");
        add_variant<F32, FK_C2C, 125, 2, false, false, 5, 5, 5, 4>(*v, KFA_F32, "c2c500_5554_t125x2_sb");
        add_variant<F32, FK_C2C, 192, 1, false, false, 3, 16, 8, 8>(*v, KFA_F32, "c2c3072_31688_t192_sb");
        // (65536 lines: 8888_t256 5.11-5.14 TB/s, 8888_t512 4.92, 161616_t256 4.82-4.88,
        // 161616_t256 double-buffered 4.59: not registered)
        add_variant<F32, FK_C2C, 256, 1, false, false, 8, 8, 8, 8>(*v, KFA_F32, "c2c4096_8888_t256_sb");
        add_variant<F32, FK_C2C, 256, 1, false, false, 16, 16, 16>(*v, KFA_F32, "c2c4096_161616_t256_sb");
        add_variant<F32, FK_C2C, 512, 1, false, false, 8, 8, 8, 8>(*v, KFA_F32, "c2c4096_8888_t512_sb");
        // ---- Q15 (reference schedules), wavefront-resident lines: one LDS transpose ----
        add_xpose<8, 1, 2, 4, 4, 4, 4>(*v, "q15x_c2c512_t16x8");
        add_xpose<16, 1, 2, 4, 4, 4, 4>(*v, "q15x_c2c512_t16x16");
        add_xpose<8, 6, 2, 4, 4, 4, 4>(*v, "q15x_c2c512_t16x8_o6");
        add_xpose<4, 6, 2, 4, 4, 4, 4>(*v, "q15x_c2c512_t16x4_o6");
        add_xpose<16, 1, 4, 4, 4, 4>(*v, "q15x_c2c256_t16x16");
        // ---- Q15 (reference schedules), wavefront-resident lines: lane-swap exchanges ----
        add_wave<16, 16, 2, 4, 4, 4, 4>(*v, "q15w_c2c512_t16x16");
        add_wave<32, 8, 2, 4, 4, 4, 4>(*v, "q15w_c2c512_t32x8");
        add_wave<16, 16, 4, 4, 4, 4>(*v, "q15w_c2c256_t16x16");
        add_wave<32, 8, 4, 4, 4, 4, 4>(*v, "q15w_c2c1024_t32x8");
        add_wave<16, 16, 4, 4, 4, 4, 4>(*v, "q15w_c2c1024_t16x16");
        add_wave<64, 4, 2, 4, 4, 4, 4, 4>(*v, "q15w_c2c2048_t64x4");
        // ---- Q15 (reference schedules), packed-word butterflies through LDS ----
        add_variant<Q15P, FK_C2C, 16, 16, false, false, 2, 4, 4, 4, 4>(*v, KFA_Q15, "q15p_c2c512_t16x16_sb");
        add_variant<Q15P, FK_C2C, 32, 8, false, false, 2, 4, 4, 4, 4>(*v, KFA_Q15, "q15p_c2c512_t32x8_sb");
        add_variant<Q15P, FK_C2C, 240, 1, false, false, 5, 5, 3, 4, 4>(*v, KFA_Q15, "q15p_c2c1200_t240_sb");
        add_variant<Q15P, FK_C2C, 64, 4, false, false, 4, 4, 4, 4>(*v, KFA_Q15, "q15p_c2c256_t64x4_sb");
        add_variant<Q15P, FK_C2C, 128, 2, false, false, 4, 4, 4, 4, 4>(*v, KFA_Q15, "q15p_c2c1024_t128x2_sb");
        // ---- Q15 (reference schedules), one value per component (the real kinds' arithmetic) ----
        add_variant<Q15, FK_C2C, 128, 2, false, false, 2, 4, 4, 4, 4>(*v, KFA_Q15, "q15c2c512_t128x2_sb");
        add_real<Q15, 128, 2, false, 2, 4, 4, 4, 4>(*v, KFA_Q15, "q15real512_t128x2_sb");
        add_real<Q15, 256, 1, false, 2, 4, 4, 4, 4, 4>(*v, KFA_Q15, "q15real2048_t256_sb");
        add_fused<Q15, FK_R2C, 64, 4, 2, 4, 4, 4, 4>(*v, KFA_Q15, "q15r2c512_t64x4_fused");
        add_fused<Q15, FK_C2R, 128, 2, 2, 4, 4, 4, 4>(*v, KFA_Q15, "q15c2r512_t128x2_fused");
        add_fused<Q15, FK_R2C, 256, 1, 2, 4, 4, 4, 4, 4>(*v, KFA_Q15, "q15r2c2048_t256_fused");
        add_fused<Q15, FK_C2R, 512, 1, 2, 4, 4, 4, 4, 4>(*v, KFA_Q15, "q15c2r2048_t512_fused");
        // ---- Q31 (reference schedules; int64 products, 8 bytes per complex) ----
        // (2^19 lines: t64x4 875-880 us, t128x2 949-1003 us)
        add_variant<Q31, FK_C2C, 64, 4, false, false, 2, 4, 4, 4, 4>(*v, KFA_Q31, "q31c2c512_t64x4_sb");
        add_variant<Q31, FK_C2C, 128, 2, false, false, 2, 4, 4, 4, 4>(*v, KFA_Q31, "q31c2c512_t128x2_sb");
        add_variant<Q31, FK_C2C, 256, 1, false, false, 5, 5, 3, 4, 4>(*v, KFA_Q31, "q31c2c1200_t256_sb");
        add_real<Q31, 256, 1, false, 2, 4, 4, 4, 4, 4>(*v, KFA_Q31, "q31real2048_t256_sb");
        add_real<Q31, 256, 1, false, 5, 5, 5, 4, 4>(*v, KFA_Q31, "q31real2000_t256_sb");
        return v;
    }();
    return *reg;
}

}  // namespace kfa
