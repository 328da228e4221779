// kfa_fast_reg.hip -- the instantiated fast-kernel variants and their launchers.
//
// Each variant is (scalar kind, transform kind, line length, direction, radix schedule,
// threads per line, lines per workgroup).  Fixed-point variants carry the reference radix
// schedule of their length (kf_factor order, executed innermost first) and are used only
// when the plan's schedule equals it; float variants may use any schedule.  When several
// variants serve the same key the engine times them on first use and keeps the fastest.
#include <hip/hip_runtime.h>

#include <mutex>

#include "kfa_fast.hpp"
#include "kfa_fast_reg.hpp"

namespace kfa {

template <class A, int KIND, bool INV, int T, int LPB, class RS>
static int launch_fast(const FastArgs &a, hipStream_t s)
{
    typedef typename A::store2 S;
    const size_t lds = 2 * (size_t)LPB * RS::N() * sizeof(S);
    static std::once_flag once;
    std::call_once(once, [=] {
        (void)hipFuncSetAttribute((const void *)k_fast<A, KIND, INV, T, LPB, RS>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    });
    const long long blocks = (a.nlines + LPB - 1) / LPB;
    if (blocks <= 0)
        return 0;
    hipLaunchKernelGGL((k_fast<A, KIND, INV, T, LPB, RS>), dim3((unsigned)blocks), dim3(T * LPB), lds, s, a);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

template <class A, int KIND, int T, int LPB, int... R>
static void add_variant(std::vector<FastVariant> &v, int type, const char *name)
{
    typedef Radices<R...> RS;
    static_assert(RS::K >= 2, "single-stage schedules read and write HBM in one stage (not in-place safe)");
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
        f.tw_len = RS::tw_len();
        f.name = name;
        f.launch = inv ? &launch_fast<A, KIND, true, T, LPB, RS> : &launch_fast<A, KIND, false, T, LPB, RS>;
        v.push_back(f);
    }
}

const std::vector<FastVariant> &fast_registry()
{
    static std::vector<FastVariant> *reg = [] {
        auto *v = new std::vector<FastVariant>();
        // ---- float, real transforms: half lengths of the real workloads ----
        add_variant<F32, FK_R2C, 256, 1, 8, 8, 8, 4>(*v, KFA_F32, "r2c2048_8884_t256");
        add_variant<F32, FK_R2C, 256, 1, 4, 8, 8, 8>(*v, KFA_F32, "r2c2048_4888_t256");
        add_variant<F32, FK_R2C, 128, 2, 8, 8, 8, 4>(*v, KFA_F32, "r2c2048_8884_t128");
        add_variant<F32, FK_C2R, 256, 1, 8, 8, 8, 4>(*v, KFA_F32, "c2r2048_8884_t256");
        add_variant<F32, FK_C2R, 256, 1, 4, 8, 8, 8>(*v, KFA_F32, "c2r2048_4888_t256");
        add_variant<F32, FK_C2R, 128, 2, 8, 8, 8, 4>(*v, KFA_F32, "c2r2048_8884_t128");
        add_variant<F32, FK_R2C, 100, 2, 4, 4, 5, 5, 5>(*v, KFA_F32, "r2c2000_44555_t100");
        add_variant<F32, FK_R2C, 200, 1, 5, 5, 5, 4, 4>(*v, KFA_F32, "r2c2000_55544_t200");
        add_variant<F32, FK_C2R, 100, 2, 4, 4, 5, 5, 5>(*v, KFA_F32, "c2r2000_44555_t100");
        add_variant<F32, FK_C2R, 200, 1, 5, 5, 5, 4, 4>(*v, KFA_F32, "c2r2000_55544_t200");
        add_variant<F32, FK_R2C, 50, 4, 4, 5, 5, 5>(*v, KFA_F32, "r2c500_4555_t50");
        add_variant<F32, FK_C2R, 50, 4, 4, 5, 5, 5>(*v, KFA_F32, "c2r500_4555_t50");
        // ---- float, complex ----
        add_variant<F32, FK_C2C, 240, 1, 4, 4, 3, 5, 5>(*v, KFA_F32, "c2c1200_44355_t240");
        add_variant<F32, FK_C2C, 120, 2, 5, 5, 3, 4, 4>(*v, KFA_F32, "c2c1200_55344_t120");
        add_variant<F32, FK_C2C, 256, 1, 8, 8, 8, 4>(*v, KFA_F32, "c2c2048_8884_t256");
        add_variant<F32, FK_C2C, 128, 2, 8, 8, 4, 4>(*v, KFA_F32, "c2c1024_8844_t128");
        add_variant<F32, FK_C2C, 64, 4, 8, 8, 8>(*v, KFA_F32, "c2c512_888_t64");
        add_variant<F32, FK_C2C, 64, 4, 4, 4, 4, 4>(*v, KFA_F32, "c2c256_4444_t64");
        add_variant<F32, FK_C2C, 100, 2, 5, 5, 5, 4, 2>(*v, KFA_F32, "c2c1000_55542_t100");
        // ---- Q15 (reference schedules) ----
        add_variant<Q15, FK_C2C, 128, 2, 2, 4, 4, 4, 4>(*v, KFA_Q15, "q15c2c512_t128");
        add_variant<Q15, FK_C2C, 64, 4, 2, 4, 4, 4, 4>(*v, KFA_Q15, "q15c2c512_t64");
        add_variant<Q15, FK_R2C, 128, 2, 2, 4, 4, 4, 4>(*v, KFA_Q15, "q15r2c1024_t128");
        add_variant<Q15, FK_C2R, 128, 2, 2, 4, 4, 4, 4>(*v, KFA_Q15, "q15c2r1024_t128");
        add_variant<Q15, FK_R2C, 256, 1, 2, 4, 4, 4, 4, 4>(*v, KFA_Q15, "q15r2c4096_t256");
        add_variant<Q15, FK_C2R, 256, 1, 2, 4, 4, 4, 4, 4>(*v, KFA_Q15, "q15c2r4096_t256");
        return v;
    }();
    return *reg;
}

}  // namespace kfa
