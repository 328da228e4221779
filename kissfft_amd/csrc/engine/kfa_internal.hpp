// kfa_internal.hpp -- engine internals shared by the engine's translation units
// (kfa_engine.hip: plans, staging, launches; kfa_serve.hip: the request server).
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <memory>
#include <mutex>
#include <vector>

#include "../kfa_engine.h"
#include "kfa_fast_reg.hpp"
#include "kfa_lines.hpp"

namespace kfa {

// The fast variant chosen for one transform kind of a plan, with its stage table.
struct FastChoice {
    bool decided = false;
    const FastVariant *v = nullptr;
    void *d_stw = nullptr;
};

struct DevPlan {
    int type = 0, n = 0, inverse = 0;
    std::vector<int> radix;  // execution order (innermost factor first)
    std::vector<unsigned char> h_tw;  // host copy of the reference twiddle table
    void *d_tw = nullptr;    // n complex
    bool single = true;
    LinePlan lp{};           // single pass
    int L1 = 0, Q = 0;       // four-step: L1 * Q == n
    LinePlan lp1{}, lp2{};
    std::mutex mu;           // guards `fast`
    FastChoice fast[3];      // per FK_C2C / FK_R2C / FK_C2R
};

struct RealTables {
    void *d_stw = nullptr;  // nfft/4 complex
};

// records a message for kfa_last_error() and returns `code`
int fail(int code, const char *fmt, ...);
size_t cpx_bytes(int type);
int line_cap(int type);
int current_device(int *dev);
// stream-ordered scratch from the engine's per-device pool (free with hipFreeAsync)
hipError_t scratch_alloc(void **p, size_t bytes, hipStream_t s);
// multi-pass transforms process batches in chunks whose temporary fits this many bytes
constexpr size_t kScratchChunk = (size_t)1 << 30;
bool device_accessible(const void *p);

// device plans (cached; thread-safe)
int build_cplan(const kfa_cplan *hp, int dev, std::shared_ptr<DevPlan> &out);
int build_rtab(const kfa_rplan *hp, int dev, std::shared_ptr<RealTables> &out);
// the fast variant for one transform kind of a plan (decided once per plan and kind, timed
// on a batch sized from lines_hint, the number of lines of the call that asks)
int choose_fast(DevPlan &P, int kind, hipStream_t s, const FastChoice **out, long long lines_hint = 0);

// launches between device buffers, enqueued on s
int run_c2c(DevPlan &P, const void *in, long long istride, long long idist, void *out, long long ostride,
            long long odist, long long howmany, hipStream_t s);
int run_real_rows(DevPlan &H, const RealTables &R, int inverse, const void *in, long long in_hi, void *out,
                  long long out_hi, long long howmany, hipStream_t s);
// N-D transforms of `howmany` arrays between device buffers (P: one plan per axis; for the
// real transform P holds the leading axes and H / R the real last axis)
int run_nd(const std::vector<std::shared_ptr<DevPlan>> &P, const int *dims, const void *in, void *out,
           long long howmany, hipStream_t s);
int run_ndr(DevPlan &H, const RealTables &R, const std::vector<std::shared_ptr<DevPlan>> &P, const int *dims,
            int inverse, const void *in, void *out, long long howmany, hipStream_t s);

// A caller buffer: used in place when the device can access it, else staged through a
// stream-ordered device copy (stage_in copies host -> device; finish_out device -> host).
struct Operand {
    const void *host = nullptr;  // non-null when staged
    void *dev = nullptr;
    size_t bytes = 0;
    bool staged = false;
};
int stage_in(const void *p, size_t bytes, hipStream_t s, Operand &op);
// gaps: the output span holds values the call must not touch (rows further apart than
// their length); a staged copy then starts from the caller's bytes.
int stage_out(void *p, size_t bytes, hipStream_t s, Operand &op, bool gaps = false);
int finish_out(Operand &op, hipStream_t s);
void release(Operand &op, hipStream_t s);

bool env_flag(const char *name, bool dflt);
// float twiddles exp(-+2*pi*i*(j*step mod n)/n), j < count, FP64 rounded once (8 bytes each)
std::vector<unsigned char> float_twiddles(long long n, int inverse, long long count, long long step);
// per-stage twiddle table of schedule `rad` (the fast kernels' layout) from a full table
int make_stage_table_raw(const unsigned char *h_tw, int type, long long n, const std::vector<int> &rad, int pad,
                         int fmt, void **d_out);

// whether a float plan runs Bluestein's kernel (its tables built on return)
int blue_prepare(const DevPlan &P, bool *applies);
// one Bluestein launch for plan P; b carries buffers, segments, counts and distances
struct BlueArgs;
int blue_launch(const DevPlan &P, BlueArgs &b, hipStream_t s);

// releases the fast-convolution device plans (kfa_conv.hip)
void conv_release();

// entry / exit of a C-ABI call: the stream to use (the caller's, or the calling thread's
// own blocking stream for NULL) and the optional synchronisation at the end
int call_begin(void *stream, hipStream_t *s, int *dev);
int call_end(hipStream_t s, int sync, int rc);

extern std::atomic<long long> g_plans_built, g_plan_hits, g_launches, g_h2d, g_d2h;

}  // namespace kfa
