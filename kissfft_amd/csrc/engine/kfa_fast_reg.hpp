// kfa_fast_reg.hpp -- registry of compile-time specialised kernels (see kfa_fast_reg.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "../kfa_engine.h"

namespace kfa {

struct FastArgs;

struct FastVariant {
    int type;     // KFA_F32 / KFA_Q15 / KFA_Q31
    int kind;     // FK_C2C / FK_R2C / FK_C2R
    int n;        // complex line length (the half length for real transforms)
    int inverse;
    int T;        // threads per line
    int LPB;      // lines per workgroup
    std::vector<int> radices;  // execution order, innermost first
    int tw_pad;   // stage-table rows padded to a multiple of this many entries
    int tw_fmt;   // 0: entries as stored complex values; 1: Q15 word pairs (w.r,-w.i),(w.i,w.r)
    int tw_len;   // entries of the per-stage twiddle table
    const char *name;
    int (*launch)(const FastArgs &, hipStream_t);
    const void *rtc = nullptr;  // runtime-compiled variant (kfa_rtc.hip): launched through it
    // k_fast variants (both built-in and runtime-compiled): the template arguments a fused
    // launch needs to name the same kernel body; arith == nullptr: another kernel
    const char *arith = nullptr;
    bool db = false;
    bool fuse = false;
    bool xpose = false;  // k_q15x (kfa_q15w.hpp: k_q15x_body<INV, LPB, Radices<tw_pad, radices...>>)
};

const std::vector<FastVariant> &fast_registry();
// launches variant v (built in or runtime-compiled)
int fast_launch(const FastVariant &v, const FastArgs &a, hipStream_t s);

}  // namespace kfa

namespace kfa {

struct TileArgs;
struct RealTileArgs;

// Tile-kernel variants (kfa_tile.hpp, kfa_tile_real.hpp): float, one length, direction
// and HBM mapping.
enum { TILE_COLS = 0, TILE_COLS_TWID = 1, TILE_ROWS_TO_COLS = 2, TILE_RSPLIT = 3, TILE_RMERGE = 4 };

struct TileVariant {
    int n;
    int inverse;
    int mode;     // TILE_COLS: column load / column store (N-D axes)
                  // TILE_COLS_TWID: the same with the four-step twiddle (first pass)
                  // TILE_ROWS_TO_COLS: row load / column store (four-step second pass)
                  // TILE_RSPLIT: row pairs, real split fused (forward real second pass)
                  // TILE_RMERGE: column pairs, real merge fused (inverse real first pass)
    int T, TL;    // threads per line, lines per workgroup
    std::vector<int> radices;
    int tw_pad;
    int tw_len;
    size_t lds;   // dynamic LDS bytes
    const char *name;
    int (*launch)(const TileArgs &, long long blocks, hipStream_t);            // modes 0-2
    int (*launch_real)(const RealTileArgs &, long long blocks, hipStream_t);  // modes 3-4
};

const std::vector<TileVariant> &tile_registry();

}  // namespace kfa

namespace kfa {

struct BlueArgs;

// Bluestein variants (kfa_blue.hpp): one convolution length M, float.
struct BlueVariant {
    int M;
    int T, LPB;  // (k_blue<T, LPB, Radices<tw_pad, radices...>>)
    std::vector<int> radices;
    int tw_pad;
    size_t lds;
    const char *name;
    int (*launch)(const BlueArgs &, hipStream_t);
};

const std::vector<BlueVariant> &blue_registry();

}  // namespace kfa

namespace kfa {

struct ConvArgs;

// Fast-convolution variants (k_conv in kfa_blue.hpp): one block length N, float.
struct ConvVariant {
    int N;
    int real;
    int T, LPB;
    std::vector<int> radices;
    int tw_pad;
    size_t lds;
    const char *name;
    int (*launch)(const ConvArgs &, hipStream_t);
};

const std::vector<ConvVariant> &conv_registry();

}  // namespace kfa

namespace kfa {

struct PlaneArgs;

// Plane variants (kfa_plane.hpp): float, the last two axes A x B of an N-D transform in
// one pass, one direction.
struct PlaneVariant {
    int A, B;
    int inverse;
    int TA, TB, NTH;  // threads per column, per row, per workgroup
    std::vector<int> rad_a, rad_b;
    int tw_pad;
    size_t lds;
    bool slab;    // also runs slabs of two non-adjacent axes (PlaneArgs rs / nsub / po / pi)
    const char *name;
    int (*launch)(const PlaneArgs &, hipStream_t);
};

const std::vector<PlaneVariant> &plane_registry();

}  // namespace kfa
