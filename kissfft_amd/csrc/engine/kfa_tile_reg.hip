// kfa_tile_reg.hip -- the instantiated tile-kernel and Bluestein variants (float) and
// their launchers.
//
// Lengths cover the non-last axes of the N-D workloads (64, 96, 1024, ...) and the halves
// of the four-step split of long lines (128 .. 1024, N = N1 * N2 up to 2^20 and beyond).
// Each length is registered in three HBM mappings (column / column + twiddle / row-in
// column-out) and both directions; the halves of long real transforms also have the
// fused real split (forward) and merge (inverse) passes.
#include <hip/hip_runtime.h>

#include <mutex>

#include "kfa_fast_reg.hpp"
#include "kfa_blue.hpp"
#include "kfa_tile.hpp"
#include "kfa_tile_real.hpp"
#include "kfa_plane.hpp"

namespace kfa {

template <bool INV, int T, int TL, bool LCOL, bool SCOL, bool TWID, class RS>
static int launch_tile(const TileArgs &a, long long blocks, hipStream_t s)
{
    const size_t lds = (size_t)TL * tile_pitch<RS::N(), TL>() * sizeof(float2);
    static std::once_flag once;
    std::call_once(once, [=] {
        (void)hipFuncSetAttribute((const void *)k_tile<F32, INV, T, TL, LCOL, SCOL, TWID, RS>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    });
    if (blocks <= 0)
        return 0;
    hipLaunchKernelGGL((k_tile<F32, INV, T, TL, LCOL, SCOL, TWID, RS>), dim3((unsigned)blocks), dim3(T * TL), lds,
                       s, a);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

template <int T, int TL, int... R>
static void add_tile(std::vector<TileVariant> &v, const char *name)
{
    constexpr int G = 2;  // 16-byte stage-table rows of float2
    typedef Radices<G, R...> RS;
    static_assert(T * TL <= 1024, "workgroup too large");
    static_assert((size_t)TL * tile_pitch<RS::N(), TL>() * 8 <= 160 * 1024, "tile exceeds LDS");
    for (int mode = 0; mode < 3; ++mode)
        for (int inv = 0; inv < 2; ++inv) {
            TileVariant f;
            f.n = RS::N();
            f.inverse = inv;
            f.mode = mode;
            f.T = T;
            f.TL = TL;
            f.radices = {R...};
            f.tw_pad = G;
            f.tw_len = RS::tw_len();
            f.lds = (size_t)TL * tile_pitch<RS::N(), TL>() * 8;
            f.name = name;
            f.launch_real = nullptr;
            if (mode == TILE_COLS)
                f.launch = inv ? &launch_tile<true, T, TL, true, true, false, RS>
                               : &launch_tile<false, T, TL, true, true, false, RS>;
            else if (mode == TILE_COLS_TWID)
                f.launch = inv ? &launch_tile<true, T, TL, true, true, true, RS>
                               : &launch_tile<false, T, TL, true, true, true, RS>;
            else
                f.launch = inv ? &launch_tile<true, T, TL, false, true, false, RS>
                               : &launch_tile<false, T, TL, false, true, false, RS>;
            v.push_back(f);
        }
}

template <bool MERGE, int T, int TL, class RS>
static int launch_rtile(const RealTileArgs &a, long long blocks, hipStream_t s)
{
    const size_t lds = (size_t)TL * (MERGE ? tile_pitch<RS::N(), TL>() : RS::N() + 1) * sizeof(float2);
    const void *fn = MERGE ? (const void *)k_tile_rmerge<T, TL, RS> : (const void *)k_tile_rsplit<T, TL, RS>;
    static std::once_flag once;
    std::call_once(once, [=] { (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); });
    if (blocks <= 0)
        return 0;
    if (MERGE)
        hipLaunchKernelGGL((k_tile_rmerge<T, TL, RS>), dim3((unsigned)blocks), dim3(T * TL), lds, s, a);
    else
        hipLaunchKernelGGL((k_tile_rsplit<T, TL, RS>), dim3((unsigned)blocks), dim3(T * TL), lds, s, a);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

// The fused real passes of one length: split (forward, row pairs) and merge (inverse,
// column pairs).
template <int T, int TL, int... R>
static void add_rtile(std::vector<TileVariant> &v, const char *name)
{
    constexpr int G = 2;
    typedef Radices<G, R...> RS;
    static_assert(T * TL <= 1024, "workgroup too large");
    static_assert((size_t)TL * tile_pitch<RS::N(), TL>() * 8 + (size_t)TL * (RS::N() / T) * 8 + 8 <= 160 * 1024,
                  "tile exceeds LDS");
    for (int inv = 0; inv < 2; ++inv) {
        TileVariant f;
        f.n = RS::N();
        f.inverse = inv;
        f.mode = inv ? TILE_RMERGE : TILE_RSPLIT;
        f.T = T;
        f.TL = TL;
        f.radices = {R...};
        f.tw_pad = G;
        f.tw_len = RS::tw_len();
        f.lds = (size_t)TL * (inv ? tile_pitch<RS::N(), TL>() : RS::N() + 1) * 8;
        f.name = name;
        f.launch = nullptr;
        f.launch_real = inv ? &launch_rtile<true, T, TL, RS> : &launch_rtile<false, T, TL, RS>;
        v.push_back(f);
    }
}

const std::vector<TileVariant> &tile_registry()
{
    static std::vector<TileVariant> *reg = [] {
        auto *v = new std::vector<TileVariant>();
        add_tile<8, 64, 8, 8>(*v, "tile64_88_t8x64");
        add_tile<16, 32, 8, 4, 3>(*v, "tile96_843_t16x32");
        add_tile<16, 32, 8, 16>(*v, "tile128_816_t16x32");
        add_tile<16, 16, 16, 16>(*v, "tile256_1616_t16x16");
        // first registered = preferred unless another wins the HBM-sized tuning by 4 %
        // (t64x8 is the fastest of the 512 and 1024 tiles on multi-GiB batches)
        add_tile<64, 8, 8, 8, 8>(*v, "tile512_888_t64x8");
        add_tile<64, 16, 8, 8, 8>(*v, "tile512_888_t64x16");
        add_tile<64, 8, 16, 16, 4>(*v, "tile1024_16164_t64x8");
        add_tile<64, 4, 16, 16, 4>(*v, "tile1024_16164_t64x4");
        add_tile<64, 16, 16, 16, 4>(*v, "tile1024_16164_t64x16");
        add_tile<16, 32, 16, 16>(*v, "tile256_1616_t16x32");
        add_tile<16, 16, 8, 16>(*v, "tile128_816_t16x16");
        add_tile<8, 32, 8, 8>(*v, "tile64_88_t8x32");
        add_tile<4, 64, 8, 4>(*v, "tile32_84_t4x64");
        add_tile<1, 256, 16>(*v, "tile16_16_t1x256");
        add_rtile<16, 32, 8, 16>(*v, "rtile128_816_t16x32");
        add_rtile<16, 16, 16, 16>(*v, "rtile256_1616_t16x16");
        add_rtile<16, 32, 16, 16>(*v, "rtile256_1616_t16x32");
        // real 2^20 x512 (r2c / c2r TB/s): t32x16 1.78 / 1.93, t64x8 1.69 / 1.84, t64x16
        // 1.60 / 1.65; t32x32, t16x16, t32x8 no better than t64x8
        add_rtile<32, 16, 8, 8, 8>(*v, "rtile512_888_t32x16");
        add_rtile<64, 8, 8, 8, 8>(*v, "rtile512_888_t64x8");
        add_rtile<64, 16, 8, 8, 8>(*v, "rtile512_888_t64x16");
        // (1024 threads x 16 lines of 1024 points spills 40 bytes per lane: not registered)
        add_rtile<64, 8, 16, 16, 4>(*v, "rtile1024_16164_t64x8");
        return v;
    }();
    return *reg;
}

template <class RA, class RB>
constexpr size_t plane_lds()
{
    constexpr int B = RB::N();
    return ((size_t)RA::N() * (B % 2 == 0 ? B + 1 : B) + RA::tw_len() + RB::tw_len()) * sizeof(float2);
}

// Planes of which one workgroup fills a CU's LDS run persistently (as many workgroups as
// fit on the device, each looping over its share of the planes, so that loads of the next
// plane overlap the transform of the current one); smaller planes run one per workgroup,
// several workgroups per CU overlapping each other.
template <bool INV, int TA, int TB, int NTH, class RA, class RB, bool SLAB, bool PF>
static int launch_plane_as(const PlaneArgs &a, hipStream_t s)
{
    const void *fn = (const void *)k_plane<INV, TA, TB, NTH, RA, RB, SLAB, PF>;
    constexpr size_t lds = plane_lds<RA, RB>();
    static long long resident = 0;  // 0: one workgroup per plane
    static std::once_flag once;
    std::call_once(once, [=] {
        (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        int dev = 0, cus = 0, per = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, NTH, lds) == hipSuccess && per == 1 && cus > 0)
            resident = cus;
    });
    if (a.planes <= 0)
        return 0;
    const long long grid = (resident > 0 && a.planes > resident) ? resident : a.planes;
    if (grid > 0x7fffffffLL)
        return -2;
    hipLaunchKernelGGL((k_plane<INV, TA, TB, NTH, RA, RB, SLAB, PF>), dim3((unsigned)grid), dim3(NTH), lds, s, a);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

// contiguous planes through the plain instance; slabs through the slab instance when the
// variant has one (else -2)
template <bool INV, int TA, int TB, int NTH, class RA, class RB, bool SLAB, bool PF>
static int launch_plane(const PlaneArgs &a, hipStream_t s)
{
    constexpr int A = RA::N(), B = RB::N();
    if (a.nsub == 1 && a.rs == B && a.po == (long long)A * B)
        return launch_plane_as<INV, TA, TB, NTH, RA, RB, false, PF>(a, s);
    if constexpr (SLAB)
        return launch_plane_as<INV, TA, TB, NTH, RA, RB, true, PF>(a, s);
    return -2;
}

template <class RS>
static std::vector<int> radix_list()
{
    return std::vector<int>(RS::r, RS::r + RS::K);
}

template <int TA, int TB, int NTH, class RA, class RB, bool SLAB = true, bool PF = true>
static void add_plane(std::vector<PlaneVariant> &v, const char *name)
{
    constexpr int B = RB::N();
    constexpr size_t lds = plane_lds<RA, RB>();
    static_assert(lds <= 160 * 1024, "plane exceeds LDS");
    static_assert(NTH <= 1024, "workgroup too large");
    for (int inv = 0; inv < 2; ++inv) {
        PlaneVariant f;
        f.A = RA::N();
        f.B = B;
        f.inverse = inv;
        f.TA = TA;
        f.TB = TB;
        f.NTH = NTH;
        f.rad_a = radix_list<RA>();
        f.rad_b = radix_list<RB>();
        f.tw_pad = 2;
        f.lds = lds;
        f.name = name;
        f.slab = SLAB;
        f.launch = inv ? &launch_plane<true, TA, TB, NTH, RA, RB, SLAB, PF>
                       : &launch_plane<false, TA, TB, NTH, RA, RB, SLAB, PF>;
        v.push_back(f);
    }
}

const std::vector<PlaneVariant> &plane_registry()
{
    static std::vector<PlaneVariant> *reg = [] {
        auto *v = new std::vector<PlaneVariant>();
        // Planes that leave room for several workgroups per CU.  A 96 x 120 plane (92 KB,
        // one workgroup per CU, 12 barriers per plane) measured slower than its two axis
        // passes on the 64x96x120 x16 array, which fits in the Infinity Cache (108.7 against
        // 101.1 us), and 3 % faster on 96x120 x1024: not registered.
        add_plane<8, 8, 512, Radices<2, 8, 8>, Radices<2, 8, 8>>(*v, "plane64x64_t8x8");
        // (no slab instance: its addressing spills registers at 1024 threads)
        add_plane<8, 8, 1024, Radices<2, 16, 8>, Radices<2, 16, 8>, false>(*v, "plane128x128_t8x8");
        add_plane<4, 4, 128, Radices<2, 8, 4>, Radices<2, 8, 4>>(*v, "plane32x32_t4x4");
        // 64 x 120 (63 KB): the slab of the first and last axes of the 64 x 96 x 120 array,
        // whose 96 x 120 plane does not fit twice in a CU; two passes instead of three,
        // 103-111 us against 109-119 us for the three axis passes (x16, alternating runs)
        add_plane<8, 16, 1024, Radices<2, 8, 8>, Radices<2, 8, 5, 3>>(*v, "plane64x120_t8x16");
        add_plane<4, 8, 512, Radices<2, 8, 8>, Radices<2, 8, 5, 3>>(*v, "plane64x120_t4x8");
        // without the next plane's prefetch: 121 registers instead of 151, two workgroups
        // per CU
        add_plane<4, 8, 512, Radices<2, 8, 8>, Radices<2, 8, 5, 3>, true, false>(*v, "plane64x120_t4x8_np");
        return v;
    }();
    return *reg;
}

template <int T, int LPB, class RS>
static int launch_blue(const BlueArgs &a, hipStream_t s)
{
    const size_t lds = (size_t)LPB * (RS::N() + 1) * sizeof(float2);
    static std::once_flag once;
    std::call_once(once, [=] {
        (void)hipFuncSetAttribute((const void *)k_blue<T, LPB, RS>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds);
    });
    const long long blocks = (a.nlines + LPB - 1) / LPB;
    if (blocks <= 0)
        return 0;
    hipLaunchKernelGGL((k_blue<T, LPB, RS>), dim3((unsigned)blocks), dim3(T * LPB), lds, s, a);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

template <int T, int LPB, int... R>
static void add_blue(std::vector<BlueVariant> &v, const char *name)
{
    typedef Radices<2, R...> RS;
    static_assert(T * LPB <= 1024, "workgroup too large");
    BlueVariant f;
    f.M = RS::N();
    f.T = T;
    f.LPB = LPB;
    f.radices = {R...};
    f.tw_pad = 2;
    f.lds = (size_t)LPB * (RS::N() + 1) * 8;
    f.name = name;
    f.launch = &launch_blue<T, LPB, RS>;
    v.push_back(f);
}

const std::vector<BlueVariant> &blue_registry()
{
    static std::vector<BlueVariant> *reg = [] {
        auto *v = new std::vector<BlueVariant>();
        add_blue<8, 8, 8, 8>(*v, "blue64_88_t8x8");
        add_blue<16, 4, 8, 16>(*v, "blue128_816_t16x4");
        add_blue<16, 4, 16, 16>(*v, "blue256_1616_t16x4");
        add_blue<64, 2, 8, 8, 8>(*v, "blue512_888_t64x2");
        add_blue<64, 2, 16, 16, 4>(*v, "blue1024_16164_t64x2");
        add_blue<128, 2, 16, 16, 8>(*v, "blue2048_16168_t128x2");
        add_blue<256, 1, 16, 16, 16>(*v, "blue4096_161616_t256");
        return v;
    }();
    return *reg;
}

template <int T, int LPB, bool REAL, class RS>
static int launch_conv(const ConvArgs &a, hipStream_t s)
{
    const size_t lds = (size_t)LPB * (RS::N() + 1) * sizeof(float2);
    static std::once_flag once;
    std::call_once(once, [=] {
        (void)hipFuncSetAttribute((const void *)k_conv<T, LPB, REAL, RS>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds);
    });
    const long long blocks = (a.nlines + LPB - 1) / LPB;
    if (blocks <= 0)
        return 0;
    hipLaunchKernelGGL((k_conv<T, LPB, REAL, RS>), dim3((unsigned)blocks), dim3(T * LPB), lds, s, a);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

template <int T, int LPB, int... R>
static void add_conv(std::vector<ConvVariant> &v, const char *name)
{
    typedef Radices<2, R...> RS;
    for (int real = 0; real < 2; ++real) {
        ConvVariant f;
        f.N = RS::N();
        f.real = real;
        f.T = T;
        f.LPB = LPB;
        f.radices = {R...};
        f.tw_pad = 2;
        f.lds = (size_t)LPB * (RS::N() + 1) * 8;
        f.name = name;
        f.launch = real ? &launch_conv<T, LPB, true, RS> : &launch_conv<T, LPB, false, RS>;
        v.push_back(f);
    }
}

const std::vector<ConvVariant> &conv_registry()
{
    static std::vector<ConvVariant> *reg = [] {
        auto *v = new std::vector<ConvVariant>();
        add_conv<16, 4, 16, 16>(*v, "conv256_1616_t16x4");
        add_conv<64, 2, 8, 8, 8>(*v, "conv512_888_t64x2");
        add_conv<64, 2, 16, 16, 4>(*v, "conv1024_16164_t64x2");
        add_conv<128, 2, 16, 16, 8>(*v, "conv2048_16168_t128x2");
        add_conv<256, 1, 16, 16, 16>(*v, "conv4096_161616_t256");
        add_conv<512, 1, 16, 16, 8, 4>(*v, "conv8192_161684_t512");
        add_conv<256, 1, 16, 16, 8, 4>(*v, "conv8192_161684_t256");
        return v;
    }();
    return *reg;
}

}  // namespace kfa
