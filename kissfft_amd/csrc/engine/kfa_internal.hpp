// kfa_internal.hpp -- engine internals shared by the engine's translation units
// (kfa_engine.hip: plans, staging, launches; kfa_serve.hip: the request server).
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../kfa_engine.h"
#include "kfa_fast_reg.hpp"
#include "kfa_lines.hpp"

namespace kfa {

// Device memory with shared ownership: the plan cache, calls in flight, queued server
// requests and multi-batches each hold a reference.  When the last holder lets go the
// pointer is queued, and free_deferred() -- called by kfa_cleanup -- frees the queue
// (hipFree waits for the device first).  So a cleanup never frees memory that a queued or
// running kernel still reads, and no HIP call is made from a destructor, which may run
// inside a stream callback.
using DevBuf = std::shared_ptr<void>;
int dev_upload(const void *host, size_t bytes, DevBuf &out);
int dev_alloc(size_t bytes, DevBuf &out);
void free_deferred();

// The fast variant chosen for one transform kind of a plan, with its stage table.
struct FastChoice {
    bool decided = false;
    const FastVariant *v = nullptr;
    DevBuf stw_mem;
    void *d_stw = nullptr;
};

// A float tile variant (kfa_tile.hpp) with its stage table.
struct TileChoice {
    const TileVariant *v = nullptr;
    DevBuf stw_mem;
    void *d_stw = nullptr;
};

// Float Bailey four-step split of one multi-pass line: pass 1 the N1 columns of length N2
// (with the four-step twiddle), pass 2 the N2 rows of length N1.
struct BaileyPlan {
    long long N1 = 0, N2 = 0;
    std::shared_ptr<const TileChoice> c1, c2;
    DevBuf tw_hi, tw_lo;
};

// Float real lines whose half is multi-pass, the split / merge fused into the four-step
// passes: the fused tile (length f) and the plain tile (length other = nc / f).
struct RealFusedPlan {
    long long f = 0, other = 0;
    std::shared_ptr<const TileChoice> cf, cp;
    DevBuf tw_hi, tw_lo, rtw_a, rtw_b;
};

struct BlueChoice;

struct DevPlan {
    int type = 0, n = 0, inverse = 0;
    std::vector<int> radix;  // execution order (innermost factor first)
    std::vector<unsigned char> h_tw;  // host copy of the reference twiddle table
    DevBuf tw_mem;
    void *d_tw = nullptr;    // n complex
    bool single = true;
    LinePlan lp{};           // single pass
    int L1 = 0, Q = 0;       // four-step: L1 * Q == n
    bool bigp = false;       // first stage radix L1 above one LDS line: a global-memory DFT pass
    LinePlan lp1{}, lp2{};
    std::mutex mu;           // guards the choices below (each decided once)
    FastChoice fast[3];      // per FK_C2C / FK_R2C / FK_C2R
    bool bailey_decided = false;
    std::shared_ptr<const BaileyPlan> bailey;
    bool rfused_decided[2] = {false, false};
    std::shared_ptr<const RealFusedPlan> rfused[2];  // per direction
    bool blue_decided = false;
    std::shared_ptr<const BlueChoice> blue;
};

struct RealTables {
    DevBuf stw_mem;
    void *d_stw = nullptr;  // nfft/4 complex
};

// records a message for kfa_last_error() and returns `code`
int fail(int code, const char *fmt, ...);
size_t cpx_bytes(int type);
int line_cap(int type);
int current_device(int *dev);
// stream-ordered scratch from the engine's per-device pool (free with hipFreeAsync)
hipError_t scratch_alloc(void **p, size_t bytes, hipStream_t s);

// Scratch of one call: carved from the calling thread's current workspace when one is set
// (a multi-batch's, allocated at its creation; the no-allocation entry point's static
// device array), else stream-ordered from the device pool.  Released when the object
// goes out of scope (workspace carving is a stack: releases come in reverse order).
struct Workspace {
    char *base = nullptr;
    size_t size = 0, used = 0;
};
Workspace *&thread_workspace();
// set while a no-allocation entry point runs: every allocation site fails instead
bool &thread_noalloc();
int noalloc_refuse(const char *what);
struct Scratch {
    void *p = nullptr;
    Scratch() = default;
    Scratch(const Scratch &) = delete;
    Scratch &operator=(const Scratch &) = delete;
    int get(size_t bytes, hipStream_t s);
    ~Scratch();

  private:
    hipStream_t s_ = nullptr;
    Workspace *ws_ = nullptr;
    size_t mark_ = 0;
    bool pooled_ = false;
};
// Multi-pass batches run in chunks whose temporaries fit this many bytes: the free part
// of the thread's workspace, else the chunk limit (KISSFFT_AMD_CHUNK_MB, default below:
// 2 GiB; against 1 GiB, 2^20-point batches ran 1-3 % faster with fewer, larger launches).
constexpr size_t kScratchChunk = (size_t)2 << 30;
size_t scratch_room();
// scratch bytes one transform of plan P needs on the multi-pass paths (complex: kind
// FK_C2C; real: FK_R2C / FK_C2R with P the half-length plan), 0 for single-pass plans
size_t multipass_bytes_per_transform(DevPlan &P, int fk);
// decides every lazily built choice a multi-pass transform of P will use (tile variants,
// four-step tables), so that later calls allocate nothing but their scratch
int prepare_multipass(DevPlan &P, int fk, hipStream_t s);
bool device_accessible(const void *p);

// device plans (cached; thread-safe)
int build_cplan(const kfa_cplan *hp, int dev, std::shared_ptr<DevPlan> &out);
int build_rtab(const kfa_rplan *hp, int dev, std::shared_ptr<RealTables> &out);
// the fast variant for one transform kind of a plan (decided once per plan and kind, timed
// on a batch sized from lines_hint, the number of lines of the call that asks)
int choose_fast(DevPlan &P, int kind, hipStream_t s, const FastChoice **out, long long lines_hint = 0);

// runtime-compiled variants of a plan's length and kind (kfa_rtc.hip), when none is built in
int rtc_candidates(const DevPlan &P, int kind, std::vector<const FastVariant *> &out);
// the persisted table of tuned choices (kfa_rtc.hip): key of a decision, lookup, store
std::string tuned_key(const DevPlan &P, int kind, long long tune_lines);
bool tuned_lookup(const std::string &key, std::string &name);
void tuned_store(const std::string &key, const std::string &name);

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
// The staging copy is returned to the pool by release() or, on an early return, by the
// destructor.
struct Operand {
    const void *host = nullptr;  // non-null when staged
    void *dev = nullptr;
    size_t bytes = 0;
    bool staged = false;
    hipStream_t s = nullptr;     // the stream the staging copy is ordered on
    Operand() = default;
    Operand(const Operand &) = delete;
    Operand &operator=(const Operand &) = delete;
    ~Operand()
    {
        if (staged && dev)
            (void)hipFreeAsync(dev, s);
    }
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
                         int fmt, DevBuf &out);

// whether a float plan runs Bluestein's kernel (its tables built on return; the caller may
// hold them in `keep`)
int blue_prepare(DevPlan &P, bool *applies, std::shared_ptr<const BlueChoice> *keep = nullptr);
// one Bluestein launch for plan P; b carries buffers, segments, counts and distances
struct BlueArgs;
int blue_launch(DevPlan &P, BlueArgs &b, hipStream_t s);
// fills b's tables from choice bc (plan length n) and returns the variant that runs it
struct BlueVariant;
const BlueVariant *blue_fill(const BlueChoice &bc, int n, BlueArgs &b);

// One member of a fused launch: a k_fast body (fv, a variant with `arith` set) or a
// Bluestein body (bv).
struct FastVariant;
struct FusedMember {
    const FastVariant *fv;
    const BlueVariant *bv;
};
// The runtime-compiled kernel that runs members m[0..] in one grid of NT-thread workgroups
// (member i on the workgroups of FusedTab entry i); 0 and *fn on success, nonzero when
// the runtime compiler is unavailable or the kernel would spill.
int rtc_fused(const std::vector<FusedMember> &m, int NT, hipFunction_t *fn);

// Tuning (first-use timing of variants) runs on a private stream of the device, after the
// work queued on the caller's stream, never on that stream itself, and not at all while the
// caller's stream is being captured.
int tune_stream(hipStream_t caller, hipStream_t *s);
bool stream_capturing(hipStream_t s);
// Plan tables are uploaded with synchronous copies; while this object lives the calling
// thread's capture mode is relaxed, so such uploads do not invalidate a capture in progress.
struct RelaxCapture {
    hipStreamCaptureMode mode = hipStreamCaptureModeRelaxed;
    RelaxCapture() { (void)hipThreadExchangeStreamCaptureMode(&mode); }
    ~RelaxCapture() { (void)hipThreadExchangeStreamCaptureMode(&mode); }
};

// releases the fast-convolution device plans (kfa_conv.hip)
void conv_release();

// entry / exit of a C-ABI call: the stream to use (the caller's, or the calling thread's
// own blocking stream for NULL) and the optional synchronisation at the end
int call_begin(void *stream, hipStream_t *s, int *dev);
int call_end(hipStream_t s, int sync, int rc);

extern std::atomic<long long> g_plans_built, g_plan_hits, g_launches, g_h2d, g_d2h;

}  // namespace kfa
