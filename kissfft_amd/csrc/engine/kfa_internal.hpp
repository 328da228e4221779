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
// the fast variant for one transform kind of a plan (decided once per plan and kind)
int choose_fast(DevPlan &P, int kind, hipStream_t s, const FastChoice **out);

// launches between device buffers, enqueued on s
int run_c2c(DevPlan &P, const void *in, long long istride, long long idist, void *out, long long ostride,
            long long odist, long long howmany, hipStream_t s);
int run_real_rows(DevPlan &H, const RealTables &R, int inverse, const void *in, long long in_hi, void *out,
                  long long out_hi, long long howmany, hipStream_t s);

// entry / exit of a C-ABI call: the stream to use (the caller's, or the calling thread's
// own blocking stream for NULL) and the optional synchronisation at the end
int call_begin(void *stream, hipStream_t *s, int *dev);
int call_end(hipStream_t s, int sync, int rc);

extern std::atomic<long long> g_plans_built, g_plan_hits, g_launches, g_h2d, g_d2h;

}  // namespace kfa
