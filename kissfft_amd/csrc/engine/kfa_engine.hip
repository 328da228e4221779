// kfa_engine.hip -- device plans, pointer staging and launches behind kfa_engine.h.
//
// Device plan = the reference plan moved to HBM: the radix schedule of kf_factor
// (kiss_fft.c:306-328) executed innermost-first as Stockham stages, and the host-built
// twiddle table (kiss_fft.c:361-367) uploaded once.  Plans live in a process-wide,
// mutex-guarded cache keyed by (device, scalar kind, length, direction); the host cfg
// stays a free()-able block with the reference layout, so it cannot own device memory
// (kiss_fft.h:136-138) -- kiss_fft_cleanup() releases the cache instead
// (kiss_fft.c:407-410 is that hook in the reference).
//
// Lines up to the LDS capacity run as one pass; longer 1-D lines run as two passes of
// the same kernel (four-step split of the Stockham schedule at a stage boundary, which
// keeps every butterfly and twiddle of the reference schedule, so fixed point stays
// bit-exact).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../kfa_engine.h"
#include "kfa_fast.hpp"
#include "kfa_fast_reg.hpp"
#include "kfa_internal.hpp"
#include "kfa_lines.hpp"
#include "kfa_blue.hpp"
#include "kfa_tile.hpp"
#include "kfa_tile_real.hpp"
#include "kfa_plane.hpp"
#include "kfa_dft.hpp"

namespace kfa {

// ------------------------------------------------------------------------------------
// errors
// ------------------------------------------------------------------------------------
static thread_local char g_err[512];

int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

#define KFA_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(-3, "%s failed: %s", #call, hipGetErrorString(e_));             \
    } while (0)

std::atomic<long long> g_plans_built{0}, g_plan_hits{0}, g_launches{0}, g_h2d{0}, g_d2h{0};

size_t cpx_bytes(int type) { return type == KFA_Q15 ? 4 : 8; }

// ------------------------------------------------------------------------------------
// scratch memory: one stream-ordered pool per device that keeps what it has mapped
// (release threshold: unlimited), so repeated large temporaries -- four-step passes,
// staging copies, the request server's working set -- are not mapped again on every
// call.  kfa_cleanup trims it.
// ------------------------------------------------------------------------------------
static std::mutex g_pool_mu;
static std::map<int, hipMemPool_t> g_pools;

hipError_t scratch_alloc(void **p, size_t bytes, hipStream_t s)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
        return e;
    hipMemPool_t pool;
    {
        std::lock_guard<std::mutex> g(g_pool_mu);
        auto it = g_pools.find(dev);
        if (it == g_pools.end()) {
            hipMemPoolProps props = {};
            props.allocType = hipMemAllocationTypePinned;
            props.handleTypes = hipMemHandleTypeNone;
            props.location.type = hipMemLocationTypeDevice;
            props.location.id = dev;
            if ((e = hipMemPoolCreate(&pool, &props)) != hipSuccess)
                return e;
            uint64_t keep = ~0ULL;
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
            g_pools[dev] = pool;
        } else {
            pool = it->second;
        }
    }
    return hipMallocFromPoolAsync(p, bytes ? bytes : 1, pool, s);
}

static void scratch_trim()
{
    std::lock_guard<std::mutex> g(g_pool_mu);
    for (auto &kv : g_pools)
        (void)hipMemPoolTrimTo(kv.second, 0);
}

// ---- shared device memory: deferred release ----------------------------------------------
static std::mutex g_defer_mu;
static std::vector<void *> g_deferred;

static void defer_free(void *p)
{
    if (!p)
        return;
    std::lock_guard<std::mutex> g(g_defer_mu);
    g_deferred.push_back(p);
}

void free_deferred()
{
    std::vector<void *> v;
    {
        std::lock_guard<std::mutex> g(g_defer_mu);
        v.swap(g_deferred);
    }
    for (void *p : v)
        (void)hipFree(p);  // waits for the device before freeing
}

int dev_upload(const void *host, size_t bytes, DevBuf &out)
{
    if (thread_noalloc())
        return noalloc_refuse("a device table");
    void *p = nullptr;
    KFA_HIP(hipMalloc(&p, bytes ? bytes : 1));
    out = DevBuf(p, defer_free);
    if (bytes && host)
        KFA_HIP(hipMemcpy(p, host, bytes, hipMemcpyHostToDevice));
    return 0;
}

int dev_alloc(size_t bytes, DevBuf &out)
{
    return dev_upload(nullptr, bytes, out);
}

// ---- per-call scratch: workspace carving or the pool ------------------------------------
Workspace *&thread_workspace()
{
    static thread_local Workspace *ws = nullptr;
    return ws;
}

bool &thread_noalloc()
{
    static thread_local bool on = false;
    return on;
}

int noalloc_refuse(const char *what)
{
    return fail(-2, "no-allocation call: %s would have to be allocated (run the same transform once through "
                    "the allocating entry point first, with device buffers)", what);
}

int Scratch::get(size_t bytes, hipStream_t s)
{
    s_ = s;
    if (Workspace *ws = thread_workspace()) {
        const size_t off = (ws->used + 255) & ~(size_t)255;
        if (off + bytes > ws->size)
            return fail(-2, "scratch of %zu bytes does not fit the %zu-byte workspace (%zu in use)", bytes, ws->size,
                        ws->used);
        ws_ = ws;
        mark_ = ws->used;
        ws->used = off + bytes;
        p = ws->base + off;
        return 0;
    }
    if (thread_noalloc())
        return noalloc_refuse("scratch memory");
    KFA_HIP(scratch_alloc(&p, bytes, s));
    pooled_ = true;
    return 0;
}

Scratch::~Scratch()
{
    if (ws_)
        ws_->used = mark_;
    else if (pooled_ && p)
        (void)hipFreeAsync(p, s_);
}

static size_t chunk_limit()
{
    static const size_t lim = [] {
        const char *v = getenv("KISSFFT_AMD_CHUNK_MB");
        const long long mb = v && *v ? atoll(v) : 0;
        return mb > 0 ? (size_t)mb << 20 : kScratchChunk;
    }();
    return lim;
}

size_t scratch_room()
{
    if (Workspace *ws = thread_workspace()) {
        const size_t off = (ws->used + 255) & ~(size_t)255;
        // room for a few aligned carvings of one chunk
        return off + 1024 < ws->size ? ws->size - off - 1024 : 0;
    }
    return chunk_limit();
}

// ---- tuning stream ---------------------------------------------------------------------------
static std::mutex g_tune_mu;
static std::map<int, hipStream_t> g_tune_streams;

// The device's tuning stream, made to wait for the work already queued on the caller's
// stream: candidates are timed after it, not beside it (kernels running alongside would
// decide the ranking).
int tune_stream(hipStream_t caller, hipStream_t *s)
{
    int dev = 0;
    KFA_HIP(hipGetDevice(&dev));
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        auto it = g_tune_streams.find(dev);
        if (it == g_tune_streams.end()) {
            hipStream_t t = nullptr;
            KFA_HIP(hipStreamCreateWithFlags(&t, hipStreamNonBlocking));  // default priority
            it = g_tune_streams.emplace(dev, t).first;
        }
        *s = it->second;
    }
    hipEvent_t ev = nullptr;
    KFA_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    const hipError_t e1 = hipEventRecord(ev, caller);
    const hipError_t e2 = e1 == hipSuccess ? hipStreamWaitEvent(*s, ev, 0) : e1;
    (void)hipEventDestroy(ev);
    return e2 == hipSuccess ? 0 : fail(-3, "tuning stream: %s", hipGetErrorString(e2));
}

bool stream_capturing(hipStream_t s)
{
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &st) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return st != hipStreamCaptureStatusNone;
}

// largest line kept whole in LDS (two ping-pong buffers must fit 160 KiB)
int line_cap(int type) { return type == KFA_Q15 ? 16384 : 8192; }

static const size_t kLdsMax = 160 * 1024;

// ------------------------------------------------------------------------------------
// device plans
// ------------------------------------------------------------------------------------
struct Cache {
    std::mutex mu;
    std::map<std::tuple<int, int, int, int>, std::shared_ptr<DevPlan>> cplans;   // dev,type,n,inv
    std::map<std::tuple<int, int, int, int>, std::shared_ptr<RealTables>> rtabs; // dev,type,nreal,inv
};

static Cache &cache()
{
    static Cache *c = new Cache();  // never destroyed: device memory is released by kfa_cleanup
    return *c;
}

static void fill_stages(LinePlan &lp, const std::vector<int> &radix, int s0, int s1, int n, int L,
                        int div_P, int div_m, int u_mul, int u_add_lo, int inverse)
{
    // global sub-length m before stage s is the product of the radices executed before it
    long long mg = 1;
    for (int s = 0; s < s0; ++s)
        mg *= radix[s];
    lp.L = L;
    lp.N = n;
    lp.nst = s1 - s0;
    lp.inverse = inverse;
    lp.u_mul = u_mul;
    lp.u_add_lo = u_add_lo;
    for (int s = s0; s < s1; ++s) {
        const int p = radix[s];
        const long long Pg = n / (p * mg);
        StageDesc &d = lp.st[s - s0];
        d.p = p;
        d.mg = (int)mg;
        d.Ptw = (int)Pg;
        d.ml = (int)(mg / div_m);
        d.Pl = (int)(Pg / div_P);
        mg *= p;
    }
}

int build_cplan(const kfa_cplan *hp, int dev, std::shared_ptr<DevPlan> &out)
{
    auto &C = cache();
    const auto key = std::make_tuple(dev, hp->type, hp->nfft, hp->inverse ? 1 : 0);
    {
        std::lock_guard<std::mutex> g(C.mu);
        auto it = C.cplans.find(key);
        if (it != C.cplans.end()) {
            out = it->second;
            ++g_plan_hits;
            return 0;
        }
    }
    if (thread_noalloc())
        return noalloc_refuse("the device plan");
    auto P = std::make_shared<DevPlan>();
    P->type = hp->type;
    P->n = hp->nfft;
    P->inverse = hp->inverse ? 1 : 0;
    // factors are (p, m) pairs from the outermost level; execution is innermost first
    std::vector<int> outer;
    for (int k = 0; k < 64; ++k) {
        outer.push_back(hp->factors[2 * k]);
        if (hp->factors[2 * k + 1] <= 1)
            break;
    }
    P->radix.assign(outer.rbegin(), outer.rend());
    const int K = (int)P->radix.size();
    const int n = P->n;
    const int cap = line_cap(hp->type);
    if (n <= cap) {
        P->single = true;
        fill_stages(P->lp, P->radix, 0, K, n, n, 1, 1, 1, 0, P->inverse);
    } else {
        int best = -1;
        long long bestmax = 0;
        long long L1 = 1;
        for (int j = 1; j < K; ++j) {
            L1 *= P->radix[j - 1];
            const long long Q = n / L1;
            if (L1 <= cap && Q <= cap) {
                const long long mx = L1 > Q ? L1 : Q;
                if (best < 0 || mx < bestmax) {
                    best = j;
                    bestmax = mx;
                }
            }
        }
        if (best < 0 && P->radix[0] > cap && (long long)n / P->radix[0] <= cap) {
            // kf_factor takes a cofactor above sqrt(n) whole, as the last factor: it runs
            // first (m = 1), as one DFT of radix[0] points per column.  That stage runs as
            // a global-memory pass (k_dft_stage), the remaining stages as the second pass.
            best = 1;
            P->bigp = true;
        }
        if (best < 0)
            return fail(-2, "nfft=%d: radix factors exceed the %d-point LDS line in more than the first stage; "
                            "not supported", n, cap);
        long long l1 = 1;
        for (int j = 0; j < best; ++j)
            l1 *= P->radix[j];
        P->single = false;
        P->L1 = (int)l1;
        P->Q = (int)(n / l1);
        // pass 1: lines are the Q decimated columns, local fstride P/Q
        if (!P->bigp)
            fill_stages(P->lp1, P->radix, 0, best, n, P->L1, P->Q, 1, 1, 0, P->inverse);
        // pass 2: lines are the L1 rows of length Q, column index u = lo + L1*u_local
        fill_stages(P->lp2, P->radix, best, K, n, P->Q, 1, P->L1, P->L1, 1, P->inverse);
    }
    const size_t tw_bytes = cpx_bytes(hp->type) * (size_t)n;
    P->h_tw.assign((const unsigned char *)hp->twiddles, (const unsigned char *)hp->twiddles + tw_bytes);
    {
        RelaxCapture relax;
        if (int rc = dev_upload(hp->twiddles, tw_bytes, P->tw_mem))
            return rc;
    }
    P->d_tw = P->tw_mem.get();
    ++g_plans_built;
    std::lock_guard<std::mutex> g(C.mu);
    auto it = C.cplans.find(key);
    if (it != C.cplans.end()) {  // another thread won the race (this copy is released)
        out = it->second;
        return 0;
    }
    C.cplans[key] = P;
    out = P;
    return 0;
}

int build_rtab(const kfa_rplan *hp, int dev, std::shared_ptr<RealTables> &out)
{
    auto &C = cache();
    const int nreal = 2 * hp->half.nfft;
    const auto key = std::make_tuple(dev, hp->half.type, nreal, hp->half.inverse ? 1 : 0);
    {
        std::lock_guard<std::mutex> g(C.mu);
        auto it = C.rtabs.find(key);
        if (it != C.rtabs.end()) {
            out = it->second;
            return 0;
        }
    }
    auto R = std::make_shared<RealTables>();
    const size_t cnt = (size_t)(hp->half.nfft / 2);
    {
        RelaxCapture relax;
        if (int rc = dev_upload(hp->super_twiddles, cpx_bytes(hp->half.type) * cnt, R->stw_mem))
            return rc;
    }
    R->d_stw = R->stw_mem.get();
    std::lock_guard<std::mutex> g(C.mu);
    auto it = C.rtabs.find(key);
    if (it != C.rtabs.end()) {
        out = it->second;
        return 0;
    }
    C.rtabs[key] = R;
    out = R;
    return 0;
}

// ------------------------------------------------------------------------------------
// fast variants: per-stage twiddle tables and first-use selection
// ------------------------------------------------------------------------------------
bool env_flag(const char *name, bool dflt)
{
    const char *v = getenv(name);
    if (!v || !*v)
        return dflt;
    return !(v[0] == '0' || v[0] == 'n' || v[0] == 'N' || v[0] == 'f' || v[0] == 'F');
}

// Gathers the stage tables of schedule `rad` from the full-length table: row u of stage s
// holds table[r*u*P_s] for r = 1..p-1, padded to a multiple of `pad` entries; then
// table[P_s*m_s] and table[2*P_s*m_s], padded likewise (the layout of Radices<G, ...>).
int make_stage_table_raw(const unsigned char *h_tw, int type, long long n, const std::vector<int> &rad,
                                int pad, int fmt, DevBuf &out)
{
    // fmt 1: a Q15 entry w becomes the word pair (w.r, -w.i), (w.i, w.r) (8 bytes)
    const size_t cb = fmt == 1 ? 8 : cpx_bytes(type);
    std::vector<unsigned char> h;
    long long m = 1;
    auto put = [&](long long idx) {
        const unsigned char *src = h_tw + cpx_bytes(type) * (size_t)(idx % n);
        if (fmt == 1) {
            int16_t wr, wi;
            memcpy(&wr, src, 2);
            memcpy(&wi, src + 2, 2);
            const uint32_t w0 = (uint32_t)(uint16_t)wr | ((uint32_t)(uint16_t)(int16_t)(-wi) << 16);
            const uint32_t w1 = (uint32_t)(uint16_t)wi | ((uint32_t)(uint16_t)wr << 16);
            unsigned char buf[8];
            memcpy(buf, &w0, 4);
            memcpy(buf + 4, &w1, 4);
            h.insert(h.end(), buf, buf + 8);
        } else {
            h.insert(h.end(), src, src + cb);
        }
    };
    auto fill_to = [&](size_t entries) {
        while (h.size() < entries * cb)
            put(0);
    };
    auto rup = [](int v, int g) { return (v + g - 1) / g * g; };
    for (int p : rad) {
        const long long Pst = n / (p * m);
        const int PR = rup(p - 1 > 0 ? p - 1 : 1, pad);
        for (long long u = 0; u < m; ++u) {
            const size_t row0 = h.size() / cb;
            for (int r = 1; r < p; ++r)
                put((long long)r * u * Pst);
            fill_to(row0 + PR);
        }
        const size_t c0 = h.size() / cb;
        put(Pst * m);
        put(2 * Pst * m);
        fill_to(c0 + rup(2, pad));
        m *= p;
    }
    // fixed point: the generic stages' combined entries (Radices::gen_off / GR in
    // kfa_fast.hpp: row u of stage s holds table[q * P * (u + q1 * m) mod n], q1 major)
    if (type != KFA_F32) {
        m = 1;
        for (int p : rad) {
            const bool gen = !(p == 2 || p == 3 || p == 4 || p == 5 || p == 8 || p == 16);
            const long long Pst = n / (p * m);
            if (gen)
                for (long long u = 0; u < m; ++u) {
                    const size_t row0 = h.size() / cb;
                    for (int q1 = 0; q1 < p; ++q1)
                        for (int q = 1; q < p; ++q)
                            put((long long)q * Pst * (u + (long long)q1 * m));
                    fill_to(row0 + rup(p * (p - 1), pad));
                }
            m *= p;
        }
    }
    RelaxCapture relax;
    return dev_upload(h.data(), h.size(), out);
}

static int make_stage_table(const DevPlan &P, const std::vector<int> &rad, int pad, int fmt, DevBuf &out)
{
    return make_stage_table_raw(P.h_tw.data(), P.type, P.n, rad, pad, fmt, out);
}

// Float twiddles of length n, the host library's expression (kiss_fft.c:361-367): phase
// -2*pi*i/n in FP64 (negated for the inverse), cos / sin rounded once to float.
std::vector<unsigned char> float_twiddles(long long n, int inverse, long long count, long long step)
{
    const double pi = 3.14159265358979323846264338327;
    std::vector<unsigned char> h((size_t)count * 8);
    for (long long i = 0; i < count; ++i) {
        double phase = -2 * pi * (double)(i * step % n) / (double)n;
        if (inverse)
            phase *= -1;
        const float v[2] = {(float)cos(phase), (float)sin(phase)};
        memcpy(h.data() + 8 * i, v, 8);
    }
    return h;
}

// ------------------------------------------------------------------------------------
// tile variants (float): choice per (length, direction, mapping), four-step twiddles
// ------------------------------------------------------------------------------------
struct PlaneChoice {
    const PlaneVariant *v = nullptr;
    DevBuf mem_a, mem_b;
    void *stw_a = nullptr, *stw_b = nullptr;
};

typedef std::pair<DevBuf, DevBuf> BufPair;

struct TileCache {
    std::mutex mu;
    // dev, n, inv, mode, inner, small (see tile_choice)
    std::map<std::tuple<int, int, int, int, long long, int>, std::shared_ptr<const TileChoice>> choice;
    std::map<std::tuple<int, long long, int>, BufPair> fs;                                // dev, N, inv -> hi, lo
    std::map<std::tuple<int, long long, int, int>, BufPair> rs;                           // dev, nc, n, inv
    // dev, A, B, inv, slab, small (see plane_choice)
    std::map<std::tuple<int, int, int, int, int, int>, std::shared_ptr<const PlaneChoice>> plane;
};

static TileCache &tile_cache()
{
    static TileCache *c = new TileCache();
    return *c;
}

// The tile variant for (n, inverse, mode) on the current device, or v == nullptr.
// Tuning input: random-looking data of the scalar kind (float uniform in [-1, 1); Q15 /
// Q31 halves uniform over half their range).  The fixed-point kernels are arithmetic
// bound, and their clock -- hence the ranking of variants -- depends on the data.
__global__ void k_fill_tune(unsigned *p, size_t nwords, int type)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += (size_t)gridDim.x * blockDim.x) {
        unsigned z = (unsigned)i * 0x9E3779B9u + 0x7F4A7C15u;
        z = (z ^ (z >> 16)) * 0x85EBCA6Bu;
        z = (z ^ (z >> 13)) * 0xC2B2AE35u;
        z ^= z >> 16;
        unsigned w;
        if (type == KFA_F32)
            w = __float_as_uint((float)(z >> 8) * (1.0f / 8388608.0f) - 1.0f);
        else if (type == KFA_Q15)
            w = (((z & 0x7fffu) - 0x4000u) & 0xffffu) | ((((z >> 16) & 0x7fffu) - 0x4000u) << 16);
        else
            w = (unsigned)((int)z >> 1);
        p[i] = w;
    }
}

static int fill_tune(void *p, size_t bytes, int type, hipStream_t s)
{
    const size_t nw = bytes / 4;
    hipLaunchKernelGGL(k_fill_tune, dim3(2048), dim3(256), 0, s, (unsigned *)p, nw, type);
    KFA_HIP(hipGetLastError());
    return 0;
}

static int launch_tile(const TileChoice &tc, TileArgs a, hipStream_t s);
static int launch_rtile(const TileChoice &tc, RealTileArgs a, long long howmany, hipStream_t s);

// Times candidates: `run(i)` enqueues candidate i on the tuning stream; returns the index
// of the fastest (a later candidate must win by 4 %: registration order, fastest in
// full-size sweeps first, breaks near-ties so the choice does not flip with timing noise).
// Interleaved rounds, each candidate's best round counts: a clock or power transient during
// one candidate's turn does not decide the choice.  Any failure returns candidate 0.
template <class Run>
static size_t time_candidates(size_t ncand, hipStream_t ts, Run run, const char *what)
{
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        (void)hipGetLastError();
        if (e0)
            (void)hipEventDestroy(e0);
        return 0;
    }
    std::vector<float> t(ncand, 1e30f);
    bool ok = true;
    for (size_t i = 0; i < ncand && ok; ++i)
        ok = run(i) == 0;  // warm-up (clocks, code objects)
    for (int round = 0; round < 3 && ok; ++round)
        for (size_t i = 0; i < ncand && ok; ++i) {
            ok = hipEventRecord(e0, ts) == hipSuccess;
            for (int r = 0; r < 3 && ok; ++r)
                ok = run(i) == 0;
            ok = ok && hipEventRecord(e1, ts) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
            float ms = 0;
            ok = ok && hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
            if (ok)
                t[i] = std::min(t[i], ms);
        }
    (void)hipStreamSynchronize(ts);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    size_t best = 0;
    if (!ok) {
        (void)hipGetLastError();
        if (env_flag("KISSFFT_AMD_VERBOSE", false))
            fprintf(stderr, "[kissfft_amd] %s: timing failed, taking the first registered variant\n", what);
        return 0;
    }
    for (size_t i = 1; i < ncand; ++i)
        if (t[i] < 0.96f * t[best])
            best = i;
    if (env_flag("KISSFFT_AMD_VERBOSE", false))
        for (size_t i = 0; i < ncand; ++i)
            fprintf(stderr, "[kissfft_amd]   %s candidate %zu %8.1f us\n", what, i, 1e3f * t[i] / 3);
    return best;
}

// Scratch for timing on the tuning stream, returned to the pool when the object goes.
struct TuneBuf {
    void *p = nullptr;
    hipStream_t s = nullptr;
    TuneBuf(size_t bytes, hipStream_t ts) : s(ts)
    {
        if (scratch_alloc(&p, bytes, ts) != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
        }
    }
    ~TuneBuf()
    {
        if (p)
            (void)hipFreeAsync(p, s);
    }
};

// Times the tile candidates of one (length, direction, mapping) on a scratch layout of
// that mapping (1024 lines per group, kTuneValues complex values per side) and returns
// the index of the fastest.  The fused real mappings time transforms of n * 1024 complex
// values.  Tile and plane candidates are timed on 512 MiB per side: larger than the 256 MB
// Infinity Cache, so the ranking is the one of HBM-resident batches (on a cache-resident
// 128 MiB the ranking differed: a variant 18 % slower on a 4 GiB batch won).  When the
// buffers cannot be had the smaller sizes are tried, then the first candidate is taken.
constexpr long long kTuneValues = 64LL << 20;

// Times the tile candidates of (n, mode) on up to 512 MiB per side: 1024 lines per slice,
// or for column passes with a caller's geometry (inner > 0) that geometry, inner lines
// per slice and at most the caller's `outer` slices.
static size_t tune_tiles(int n, int mode, const std::vector<std::shared_ptr<TileChoice>> &cand, hipStream_t s,
                         long long inner = 0, long long outer_max = 0)
{
    hipStream_t ts;
    if (tune_stream(s, &ts))
        return 0;
    const bool geo = mode == TILE_COLS && inner > 0 && outer_max > 0;
    const long long C = geo ? inner : 1024;
    const bool real = mode == TILE_RSPLIT || mode == TILE_RMERGE;
    for (long long values = kTuneValues; values >= (kTuneValues >> 4); values >>= 2) {
        long long outer = std::max(1LL, values / ((long long)n * C));
        if (geo)
            outer = std::min(outer, outer_max);
        const size_t bytes = 8 * (size_t)(n * C + 1) * (size_t)outer;
        const size_t hib = 8 * (size_t)((long long)n * C / 1024 + 1);
        const size_t rtb = 8 * (size_t)(n + C);
        TuneBuf bin(bytes, ts), bout(bytes, ts), whi(hib, ts), wlo(8 * 1024, ts), rtw(rtb, ts);
        if (!bin.p || !bout.p || !whi.p || !wlo.p || !rtw.p)
            continue;
        if (fill_tune(bin.p, bytes, KFA_F32, ts) || hipMemsetAsync(whi.p, 0x3c, hib, ts) != hipSuccess ||
            hipMemsetAsync(wlo.p, 0x3c, 8 * 1024, ts) != hipSuccess || hipMemsetAsync(rtw.p, 0x3c, rtb, ts) != hipSuccess)
            return 0;
        TileArgs a{};
        a.in = bin.p;
        a.out = bout.p;
        a.tw_hi = whi.p;
        a.tw_lo = wlo.p;
        a.nlines = outer * C;
        a.nlo = C;
        a.in_hi = a.out_hi = (long long)n * C;
        a.out_lo = 1;
        a.out_el = C;
        a.in_lo = mode == TILE_ROWS_TO_COLS ? n : 1;
        a.in_el = mode == TILE_ROWS_TO_COLS ? 1 : C;
        RealTileArgs ra{};
        ra.in = bin.p;
        ra.out = bout.p;
        ra.rtw_a = rtw.p;
        ra.rtw_b = (const char *)rtw.p + 8 * C;
        ra.tw_hi = whi.p;
        ra.tw_lo = wlo.p;
        ra.nc = (long long)n * C;
        ra.in_hi = ra.out_hi = ra.nc + 1;
        ra.other = (int)C;
        char what[64];
        snprintf(what, sizeof what, "tile n=%d mode=%d", n, mode);
        return time_candidates(cand.size(), ts, [&](size_t i) {
            return real ? launch_rtile(*cand[i], ra, outer, ts) : launch_tile(*cand[i], a, ts);
        }, what);
    }
    return 0;
}

// The tile variant for (n, inverse, mode) on the current device (v == nullptr: none);
// several candidates are timed once (KISSFFT_AMD_TILE forces one by name,
// KISSFFT_AMD_TUNE=0 takes the first registered; while the caller's stream `s` is being
// captured the first registered is taken untimed).  A column pass of an N-D axis passes
// its geometry (inner lines per slice, outer slices): its candidates are timed in that
// geometry and the choice is kept per inner count and size class (up to 128 MiB per
// side, or more), since the row pitch and the Infinity Cache change their ranking.
static int tile_choice(int n, int inverse, int mode, hipStream_t s, std::shared_ptr<const TileChoice> &out,
                       long long inner = 0, long long outer = 0)
{
    int dev;
    int rc = current_device(&dev);
    if (rc)
        return rc;
    auto &C = tile_cache();
    std::lock_guard<std::mutex> g(C.mu);
    const bool geo = mode == TILE_COLS && inner > 0 && outer > 0;
    const int small = geo && 8.0 * n * (double)inner * (double)outer <= (double)(128 << 20);
    const auto key = std::make_tuple(dev, n, inverse, mode, geo ? inner : 0LL, small);
    auto it = C.choice.find(key);
    if (it != C.choice.end()) {
        out = it->second;
        return 0;
    }
    if (thread_noalloc())
        return noalloc_refuse("a tile kernel's tables");
    std::vector<std::shared_ptr<TileChoice>> cand;
    if (env_flag("KISSFFT_AMD_FAST", true)) {
        const char *force = getenv("KISSFFT_AMD_TILE");
        for (const TileVariant &v : tile_registry())
            if (v.n == n && v.inverse == inverse && v.mode == mode && (!force || !*force || !strcmp(force, v.name)))
                cand.push_back(std::make_shared<TileChoice>(TileChoice{&v, nullptr, nullptr}));
        if (cand.empty() && force && *force)  // forced name of another length: the first candidate
            for (const TileVariant &v : tile_registry())
                if (v.n == n && v.inverse == inverse && v.mode == mode) {
                    cand.push_back(std::make_shared<TileChoice>(TileChoice{&v, nullptr, nullptr}));
                    break;
                }
    }
    const std::vector<unsigned char> tw = float_twiddles(n, inverse, n, 1);
    for (auto &c : cand) {
        if ((rc = make_stage_table_raw(tw.data(), KFA_F32, n, c->v->radices, c->v->tw_pad, 0, c->stw_mem)))
            return rc;
        c->d_stw = c->stw_mem.get();
    }
    size_t best = 0;
    if (cand.size() > 1 && env_flag("KISSFFT_AMD_TUNE", true) && !stream_capturing(s))
        best = geo ? tune_tiles(n, mode, cand, s, inner, outer) : tune_tiles(n, mode, cand, s);
    std::shared_ptr<const TileChoice> tc = cand.empty() ? std::make_shared<const TileChoice>() : cand[best];
    if (tc->v && env_flag("KISSFFT_AMD_VERBOSE", false))
        fprintf(stderr, "[kissfft_amd] tile n=%d inverse=%d mode=%d: %s\n", n, inverse, mode, tc->v->name);
    C.choice[key] = tc;
    out = tc;
    return 0;
}

// W^(q*1024) for q <= N/1024 and W^r for r < 1024, W = exp(-+2*pi*i/N)
static int fourstep_twiddles(long long N, int inverse, BufPair &out)
{
    int dev;
    int rc = current_device(&dev);
    if (rc)
        return rc;
    auto &C = tile_cache();
    std::lock_guard<std::mutex> g(C.mu);
    const auto key = std::make_tuple(dev, N, inverse);
    auto it = C.fs.find(key);
    if (it == C.fs.end()) {
        const std::vector<unsigned char> h = float_twiddles(N, inverse, N / 1024 + 1, 1024);
        const std::vector<unsigned char> l = float_twiddles(N, inverse, 1024, 1);
        BufPair bp;
        RelaxCapture relax;
        if ((rc = dev_upload(h.data(), h.size(), bp.first)) || (rc = dev_upload(l.data(), l.size(), bp.second)))
            return rc;
        it = C.fs.emplace(key, bp).first;
    }
    out = it->second;
    return 0;
}

// The fused real passes' split / merge twiddles for half length nc, line length n and
// other = nc / n lines: a[i] = exp(-+i*pi*(i/nc + 1/2)), i < other, and
// b[e] = exp(-+i*pi*e/n), e < n (kiss_fftr.c:48-53 restated as a product).
static int real_split_twiddles(long long nc, int n, int inverse, BufPair &out)
{
    int dev;
    int rc = current_device(&dev);
    if (rc)
        return rc;
    auto &C = tile_cache();
    std::lock_guard<std::mutex> g(C.mu);
    const auto key = std::make_tuple(dev, nc, n, inverse);
    auto it = C.rs.find(key);
    if (it == C.rs.end()) {
        const double pi = 3.14159265358979323846264338327950288;
        const double sg = inverse ? 1.0 : -1.0;
        const long long other = nc / n;
        std::vector<float> ha(2 * other), hb(2 * (size_t)n);
        for (long long i = 0; i < other; ++i) {
            const double ph = sg * pi * ((double)i / (double)nc + 0.5);
            ha[2 * i] = (float)cos(ph);
            ha[2 * i + 1] = (float)sin(ph);
        }
        for (int e = 0; e < n; ++e) {
            const double ph = sg * pi * (double)e / (double)n;
            hb[2 * e] = (float)cos(ph);
            hb[2 * e + 1] = (float)sin(ph);
        }
        BufPair bp;
        RelaxCapture relax;
        if ((rc = dev_upload(ha.data(), ha.size() * 4, bp.first)) ||
            (rc = dev_upload(hb.data(), hb.size() * 4, bp.second)))
            return rc;
        it = C.rs.emplace(key, bp).first;
    }
    out = it->second;
    return 0;
}

static int launch_tile(const TileChoice &tc, TileArgs a, hipStream_t s)
{
    a.stw = tc.d_stw;
    const long long tiles = (a.nlines + tc.v->TL - 1) / tc.v->TL;
    // Column phases (consecutive lines = neighbouring column segments of the same rows):
    // groups of 16 neighbouring tiles run on one XCD.  Segments narrower than a 128-byte
    // line, or rows that do not start on a line (spectra of nfft/2+1 bins), share lines
    // with the neighbouring tile, which then hit in that XCD's L2; aligned segments gain
    // from neighbours sharing DRAM pages (measured against no grouping: 2-D real 1024x1000
    // +13 % forward and +17 % inverse, c2c 2^20 +10 % with groups of 8; groups of 16
    // against 8: c2c / r2c / c2r 2^20 +3 / +5 / +2 %, 2-D real +1-2 %; 32: no further gain).
    const bool cols = a.in_lo == 1 || a.out_lo == 1;
    a.xcd_group = cols ? 16 : 1;
    if (tiles > 0x7fffffffLL)
        return fail(-2, "too many lines (%lld) for one launch", a.nlines);
    ++g_launches;
    return tc.v->launch(a, tiles, s) ? fail(-3, "tile kernel launch failed") : 0;
}

// A fused real pass over `howmany` transforms: a.other / 2 line pairs each, TL / 2 per
// workgroup.
static int launch_rtile(const TileChoice &tc, RealTileArgs a, long long howmany, hipStream_t s)
{
    a.stw = tc.d_stw;
    const int np = tc.v->TL / 2;
    if (a.other < 32 || a.other % 2 || (a.other / 2) % np)
        return fail(-2, "fused real pass: %d lines do not tile by %d pairs", a.other, np);
    const long long blocks = howmany * (a.other / 2 / np);
    if (blocks > 0x7fffffffLL)
        return fail(-2, "too many transforms (%lld) for one launch", howmany);
    a.tiles = blocks;
    // neighbouring tiles share the cache lines of their bin runs and DRAM pages (groups of
    // 16, as for the column phases of launch_tile)
    a.xcd_group = 16;
    ++g_launches;
    return tc.v->launch_real(a, blocks, s) ? fail(-3, "fused real tile kernel launch failed") : 0;
}

// ------------------------------------------------------------------------------------
// Bluestein (float lengths with a large prime factor)
// ------------------------------------------------------------------------------------
struct BlueChoice {
    const BlueVariant *v = nullptr;
    DevBuf mem_f, mem_i, mem_chirp, mem_bfft;
};

// In-place radix-2 FFT in FP64 (forward: exp(-2*pi*i*j*k/M)), M a power of two.
static void fft_pow2_f64(std::vector<double> &re, std::vector<double> &im)
{
    const size_t M = re.size();
    for (size_t i = 1, j = 0; i < M; ++i) {
        size_t bit = M >> 1;
        for (; j & bit; bit >>= 1)
            j ^= bit;
        j ^= bit;
        if (i < j) {
            std::swap(re[i], re[j]);
            std::swap(im[i], im[j]);
        }
    }
    const double pi = 3.14159265358979323846264338327;
    for (size_t len = 2; len <= M; len <<= 1) {
        for (size_t k = 0; k < len / 2; ++k) {
            const double ang = -2 * pi * (double)k / (double)len;
            const double wr = cos(ang), wi = sin(ang);
            for (size_t i = 0; i < M; i += len) {
                const size_t a = i + k, b = i + k + len / 2;
                const double xr = re[b] * wr - im[b] * wi, xi = re[b] * wi + im[b] * wr;
                re[b] = re[a] - xr;
                im[b] = im[a] - xi;
                re[a] += xr;
                im[a] += xi;
            }
        }
    }
}

// The Bluestein variant and tables for a float plan (v == nullptr when the plan has no
// prime factor above 16 or no registered convolution length fits), decided once per plan.
static int blue_choice(DevPlan &P, std::shared_ptr<const BlueChoice> &out)
{
    static const std::shared_ptr<const BlueChoice> none = std::make_shared<const BlueChoice>();
    out = none;
    if (P.type != KFA_F32 || !P.single || !env_flag("KISSFFT_AMD_BLUESTEIN", true))
        return 0;
    int pmax = 0;
    for (int p : P.radix)
        pmax = std::max(pmax, p);
    if (pmax <= 16)
        return 0;
    std::lock_guard<std::mutex> g(P.mu);
    if (P.blue_decided) {
        out = P.blue;
        return 0;
    }
    if (thread_noalloc())
        return noalloc_refuse("Bluestein tables");
    auto bc = std::make_shared<BlueChoice>();
    const long long n = P.n;
    for (const BlueVariant &v : blue_registry())
        if (v.M >= 2 * n - 1 && (!bc->v || v.M < bc->v->M))
            bc->v = &v;
    if (bc->v) {
        const int M = bc->v->M;
        const double pi = 3.14159265358979323846264338327;
        const double sg = P.inverse ? 1.0 : -1.0;
        std::vector<double> wr(n), wi(n);
        std::vector<float> hw(2 * n);
        for (long long i = 0; i < n; ++i) {
            const long long q = (i * i) % (2 * n);  // n^2 mod 2N: exact phase reduction
            const double ph = sg * pi * (double)q / (double)n;
            wr[i] = cos(ph);
            wi[i] = sin(ph);
            hw[2 * i] = (float)wr[i];
            hw[2 * i + 1] = (float)wi[i];
        }
        std::vector<double> br(M, 0.0), bi(M, 0.0);
        for (long long m = 0; m < n; ++m) {
            br[m] = wr[m];
            bi[m] = -wi[m];
            if (m) {
                br[M - m] = wr[m];
                bi[M - m] = -wi[m];
            }
        }
        fft_pow2_f64(br, bi);
        std::vector<float> hb(2 * M);
        for (int m = 0; m < M; ++m) {
            hb[2 * m] = (float)(br[m] / M);
            hb[2 * m + 1] = (float)(bi[m] / M);
        }
        const std::vector<unsigned char> tf = float_twiddles(M, 0, M, 1), ti = float_twiddles(M, 1, M, 1);
        int rc;
        RelaxCapture relax;
        if ((rc = make_stage_table_raw(tf.data(), KFA_F32, M, bc->v->radices, bc->v->tw_pad, 0, bc->mem_f)) ||
            (rc = make_stage_table_raw(ti.data(), KFA_F32, M, bc->v->radices, bc->v->tw_pad, 0, bc->mem_i)) ||
            (rc = dev_upload(hw.data(), hw.size() * 4, bc->mem_chirp)) ||
            (rc = dev_upload(hb.data(), hb.size() * 4, bc->mem_bfft)))
            return rc;
        if (env_flag("KISSFFT_AMD_VERBOSE", false))
            fprintf(stderr, "[kissfft_amd] n=%lld inverse=%d: Bluestein %s\n", n, P.inverse, bc->v->name);
    }
    P.blue = bc;
    P.blue_decided = true;
    out = bc;
    return 0;
}

static void blue_args(const BlueChoice &bc, int n, BlueArgs &b)
{
    b.stw_f = bc.mem_f.get();
    b.stw_i = bc.mem_i.get();
    b.chirp = bc.mem_chirp.get();
    b.bfft = bc.mem_bfft.get();
    b.n = n;
}

const BlueVariant *blue_fill(const BlueChoice &bc, int n, BlueArgs &b)
{
    blue_args(bc, n, b);
    return bc.v;
}

// Whether float plan P runs Bluestein's kernel; its tables are built now rather than at
// the first launch (the multi-batch decides at creation, before any capture).
int blue_prepare(DevPlan &P, bool *applies, std::shared_ptr<const BlueChoice> *keep)
{
    std::shared_ptr<const BlueChoice> bc;
    const int rc = blue_choice(P, bc);
    *applies = rc == 0 && bc->v != nullptr;
    if (keep)
        *keep = bc;
    return rc;
}

// Bluestein launch for float plan P with b's buffers / segments / distances filled in by
// the caller (the multi-batch's grouped items); the tables come from P's choice.
int blue_launch(DevPlan &P, BlueArgs &b, hipStream_t s)
{
    std::shared_ptr<const BlueChoice> bc;
    int rc = blue_choice(P, bc);
    if (rc)
        return rc;
    if (!bc->v)
        return fail(-2, "no Bluestein kernel for n=%d", P.n);
    blue_args(*bc, P.n, b);
    ++g_launches;
    return bc->v->launch(b, s) ? fail(-3, "Bluestein kernel launch failed") : 0;
}

// The plane variant for A x B planes (float) on the current device, or v == nullptr:
// contiguous planes of the last two axes, or (slab) an outer axis of length A with the
// last axis, rows rs values apart and nsub planes per value of the axes before it (the
// plane geometry of PlaneArgs).  Several candidates are timed once, in the caller's
// geometry, on the caller's planes up to 512 MiB; the choice is kept per geometry kind
// and per size class (up to 128 MiB per side, which the 256 MB Infinity Cache holds
// between passes, or more: the ranking differs between the two).
// (KISSFFT_AMD_PLANE=0 turns the plane pass off, KISSFFT_AMD_TILE forces a variant by name.)
static int plane_choice(int A, int B, int inverse, bool slab, int rs, long long nsub, long long planes,
                        hipStream_t s, std::shared_ptr<const PlaneChoice> &out)
{
    int dev;
    int rc = current_device(&dev);
    if (rc)
        return rc;
    auto &C = tile_cache();
    std::lock_guard<std::mutex> g(C.mu);
    const bool small = 8.0 * A * B * (double)planes <= (double)(128 << 20);
    const auto key = std::make_tuple(dev, A, B, inverse, (int)slab, (int)small);
    auto it = C.plane.find(key);
    if (it != C.plane.end()) {
        out = it->second;
        return 0;
    }
    if (thread_noalloc())
        return noalloc_refuse("a plane kernel's tables");
    std::vector<std::shared_ptr<PlaneChoice>> cand;
    if (env_flag("KISSFFT_AMD_FAST", true) && env_flag("KISSFFT_AMD_PLANE", true)) {
        const char *force = getenv("KISSFFT_AMD_TILE");
        bool named = false;
        for (const PlaneVariant &v : plane_registry())
            named |= force && *force && !strcmp(force, v.name) && v.A == A && v.B == B;
        for (const PlaneVariant &v : plane_registry())
            if (v.A == A && v.B == B && v.inverse == inverse && (!slab || v.slab) &&
                (!named || !strcmp(force, v.name))) {
                auto c = std::make_shared<PlaneChoice>();
                c->v = &v;
                cand.push_back(c);
            }
    }
    for (auto &c : cand) {
        const std::vector<unsigned char> ta = float_twiddles(A, inverse, A, 1), tb = float_twiddles(B, inverse, B, 1);
        if ((rc = make_stage_table_raw(ta.data(), KFA_F32, A, c->v->rad_a, c->v->tw_pad, 0, c->mem_a)) ||
            (rc = make_stage_table_raw(tb.data(), KFA_F32, B, c->v->rad_b, c->v->tw_pad, 0, c->mem_b)))
            return rc;
        c->stw_a = c->mem_a.get();
        c->stw_b = c->mem_b.get();
    }
    size_t best = 0;
    hipStream_t ts;
    if (cand.size() > 1 && env_flag("KISSFFT_AMD_TUNE", true) && !stream_capturing(s) && tune_stream(s, &ts) == 0) {
        const long long sub = slab ? std::max(1LL, nsub) : 1;
        for (long long values = kTuneValues; values >= (kTuneValues >> 4); values >>= 2) {
            // whole groups of `sub` planes (a slab's planes tile its buffer only so)
            const long long want = std::min(planes, std::max(1LL, values / ((long long)A * B)));
            const long long tp = std::max(1LL, want / sub) * sub;
            const size_t bytes = 8 * (size_t)A * B * (size_t)tp;
            TuneBuf bin(bytes, ts), bout(bytes, ts);
            if (!bin.p || !bout.p)
                continue;
            if (fill_tune(bin.p, bytes, KFA_F32, ts))
                break;
            char what[64];
            snprintf(what, sizeof what, "plane %dx%d%s", A, B, slab ? " slab" : "");
            best = time_candidates(cand.size(), ts, [&](size_t i) {
                const PlaneArgs pa =
                    slab ? PlaneArgs{bin.p, bout.p, cand[i]->stw_a, cand[i]->stw_b, tp, rs, sub, (long long)A * rs, B}
                         : plane_args(bin.p, bout.p, cand[i]->stw_a, cand[i]->stw_b, tp, A, B);
                return cand[i]->v->launch(pa, ts);
            }, what);
            break;
        }
    }
    std::shared_ptr<const PlaneChoice> pc = cand.empty() ? std::make_shared<const PlaneChoice>() : cand[best];
    if (pc->v && env_flag("KISSFFT_AMD_VERBOSE", false))
        fprintf(stderr, "[kissfft_amd] plane %dx%d inverse=%d: %s\n", A, B, inverse, pc->v->name);
    C.plane[key] = pc;
    out = pc;
    return 0;
}

// Float four-step (Bailey) for n = N1*N2 with both halves registered as tile variants,
// decided once per plan: pass 1 transforms the N1 columns (length N2, stride N1) and
// applies W_N^(n1*k2); pass 2 transforms the N2 rows (length N1) and writes them as
// columns (k = k2 + N2*k1).  P.bailey stays null when no such split exists.
static int bailey_plan(DevPlan &P, hipStream_t s, std::shared_ptr<const BaileyPlan> &out)
{
    std::lock_guard<std::mutex> g(P.mu);
    if (P.bailey_decided) {
        out = P.bailey;
        return 0;
    }
    out = nullptr;
    if (P.type != KFA_F32 || P.single || P.bigp || !env_flag("KISSFFT_AMD_FAST", true)) {
        P.bailey_decided = true;
        return 0;
    }
    if (thread_noalloc())
        return noalloc_refuse("the four-step plan");
    const long long n = P.n;
    long long best1 = 0;
    for (const TileVariant &v : tile_registry()) {
        if (v.mode != TILE_ROWS_TO_COLS || v.inverse != P.inverse || n % v.n != 0)
            continue;
        const long long N1 = v.n, N2 = n / N1;
        bool have2 = false;
        for (const TileVariant &w : tile_registry())
            have2 |= w.mode == TILE_COLS_TWID && w.inverse == P.inverse && w.n == N2;
        if (!have2)
            continue;
        // most balanced split
        auto score = [&](long long a) { return a > n / a ? a : n / a; };
        if (best1 == 0 || score(N1) < score(best1))
            best1 = N1;
    }
    if (best1) {
        auto bp = std::make_shared<BaileyPlan>();
        bp->N1 = best1;
        bp->N2 = n / best1;
        BufPair tw;
        int rc;
        if ((rc = tile_choice((int)bp->N1, P.inverse, TILE_ROWS_TO_COLS, s, bp->c2)) ||
            (rc = tile_choice((int)bp->N2, P.inverse, TILE_COLS_TWID, s, bp->c1)) ||
            (rc = fourstep_twiddles(n, P.inverse, tw)))
            return rc;
        bp->tw_hi = tw.first;
        bp->tw_lo = tw.second;
        if (bp->c1->v && bp->c2->v)
            P.bailey = bp;
    }
    P.bailey_decided = true;
    out = P.bailey;
    return 0;
}

static int run_c2c_bailey(const BaileyPlan &bp, long long n, const void *in, long long istride, long long idist,
                          void *out, long long ostride, long long odist, long long howmany, hipStream_t s)
{
    const long long N1 = bp.N1, N2 = bp.N2;
    TileArgs a{};
    a.tw_hi = bp.tw_hi.get();
    a.tw_lo = bp.tw_lo.get();
    Scratch tmp;
    if (int rc = tmp.get(8 * (size_t)n * (size_t)howmany, s))
        return rc;
    // pass 1: line (b, n1): element n2 at in[b*idist + (n1 + N1*n2)*istride] -> tmp
    a.in = in;
    a.out = tmp.p;
    a.nlines = howmany * N1;
    a.nlo = N1;
    a.in_hi = idist;
    a.in_lo = istride;
    a.in_el = N1 * istride;
    a.out_hi = n;
    a.out_lo = 1;
    a.out_el = N1;
    int rc = launch_tile(*bp.c1, a, s);
    if (rc == 0) {
        // pass 2: line (b, k2): element n1 at tmp[b*n + n1 + N1*k2] -> out[k2 + N2*k1]
        TileArgs b{};
        b.in = tmp.p;
        b.out = out;
        b.nlines = howmany * N2;
        b.nlo = N2;
        b.in_hi = n;
        b.in_lo = N1;
        b.in_el = 1;
        b.out_hi = odist;
        b.out_lo = ostride;
        b.out_el = N2 * ostride;
        rc = launch_tile(*bp.c2, b, s);
    }
    return rc;
}

// Chooses (once per plan and kind) the fastest registered variant, or none.  Timing runs
// on the device's tuning stream with its own scratch; when that scratch cannot be had,
// smaller batches are tried and then registration order decides.  While the caller's
// stream is being captured the first registered variant is taken untimed (the capture is
// not disturbed).  The choice is recorded only once its table exists.
int choose_fast(DevPlan &P, int kind, hipStream_t s, const FastChoice **out, long long lines_hint)
{
    std::lock_guard<std::mutex> g(P.mu);
    FastChoice &fc = P.fast[kind];
    *out = &fc;
    if (fc.decided)
        return 0;
    if (!P.single || !env_flag("KISSFFT_AMD_FAST", true)) {
        fc.decided = true;
        return 0;
    }
    std::vector<const FastVariant *> cand;
    for (const FastVariant &v : fast_registry()) {
        if (v.type != P.type || v.kind != kind || v.n != P.n || v.inverse != P.inverse)
            continue;
        if (P.type != KFA_F32 && v.radices != P.radix)
            continue;  // fixed point must keep the reference schedule
        cand.push_back(&v);
    }
    if (thread_noalloc())
        return noalloc_refuse("the kernel choice");
    // lengths without a built-in variant: compiled at run time (not while the caller's
    // stream is being captured: such a plan keeps the general kernel)
    if (cand.empty() && !stream_capturing(s))
        (void)rtc_candidates(P, kind, cand);
    if (cand.empty()) {
        fc.decided = true;
        return 0;
    }
    std::vector<DevBuf> tabs(cand.size());
    for (size_t i = 0; i < cand.size(); ++i) {
        if ((long long)(cand[i]->tw_len) <= 0)
            return fail(-2, "bad fast variant table");
        if (int rc = make_stage_table(P, cand[i]->radices, cand[i]->tw_pad, cand[i]->tw_fmt, tabs[i]))
            return rc;
    }
    size_t best = 0;
    const char *force = getenv("KISSFFT_AMD_VARIANT");
    bool forced = false;
    if (force && *force) {
        for (size_t i = 0; i < cand.size(); ++i)
            if (strcmp(force, cand[i]->name) == 0) {
                best = i;
                forced = true;
            }
    }
    // time every candidate on a scratch batch large enough to stream from HBM, as large as
    // the call that asks (between 256 MiB and 1 GiB per side, beyond the last-level cache):
    // the ranking of variants depends on the batch size.  Earlier decisions for the same
    // device, length, kind and timing batch come from the persisted table.
    const size_t cb = cpx_bytes(P.type);
    const long long per_line = (long long)cb * (P.n + 1);
    long long tune_lines = std::max(lines_hint, (256LL << 20) / per_line);
    tune_lines = std::max(64LL, std::min(tune_lines, (1LL << 30) / per_line));
    const bool tune = !forced && cand.size() > 1 && env_flag("KISSFFT_AMD_TUNE", true);
    const std::string tkey = tune ? tuned_key(P, kind, tune_lines) : std::string();
    std::string tuned;
    bool from_table = false;
    if (tune && tuned_lookup(tkey, tuned))
        for (size_t i = 0; i < cand.size(); ++i)
            if (tuned == cand[i]->name) {
                best = i;
                from_table = true;
            }
    hipStream_t ts;
    if (tune && !from_table && !stream_capturing(s) && tune_stream(s, &ts) == 0) {
        long long lines = tune_lines;
        for (; lines >= 64; lines /= 4) {
            const size_t bytes = (size_t)per_line * (size_t)lines;
            const size_t rtb = cb * (size_t)(P.n / 2 + 1);
            TuneBuf bin(bytes, ts), bout(bytes, ts), brtw(rtb, ts);
            if (!bin.p || !bout.p || !brtw.p)
                continue;
            // random-looking data of the plan's kind: representative power draw and clock
            if (fill_tune(bin.p, bytes, P.type, ts) || hipMemsetAsync(brtw.p, 0x3c, rtb, ts) != hipSuccess)
                break;
            auto args_for = [&](size_t i) {
                FastArgs a{};
                a.in = bin.p;
                a.out = bout.p;
                a.stw = tabs[i].get();
                a.rtw = brtw.p;
                a.nlines = lines;
                a.in_hi = kind == FK_C2R ? P.n + 1 : P.n;
                a.in_el = 1;
                a.out_hi = kind == FK_R2C ? P.n + 1 : P.n;
                a.out_el = 1;
                return a;
            };
            char what[64];
            snprintf(what, sizeof what, "n=%d kind=%d", P.n, kind);
            best = time_candidates(cand.size(), ts, [&](size_t i) { return fast_launch(*cand[i], args_for(i), ts); },
                                   what);
            tuned_store(tkey, cand[best]->name);
            break;
        }
    }
    fc.v = cand[best];
    fc.stw_mem = tabs[best];
    fc.d_stw = fc.stw_mem.get();
    fc.decided = true;
    if (env_flag("KISSFFT_AMD_VERBOSE", false))
        fprintf(stderr, "[kissfft_amd] n=%d kind=%d inverse=%d: fast variant %s\n", P.n, kind, P.inverse,
                fc.v->name);
    return 0;
}

// ------------------------------------------------------------------------------------
// streams, pointers, staging
// ------------------------------------------------------------------------------------
static thread_local hipStream_t t_stream = nullptr;

// The calling thread's stream: a blocking stream, so synchronous calls are ordered after
// work the caller queued on the legacy default stream (e.g. whatever produced a device
// input) and before work queued there afterwards.  Default priority.
static int thread_stream(hipStream_t *s)
{
    if (!t_stream)
        KFA_HIP(hipStreamCreateWithFlags(&t_stream, hipStreamDefault));
    *s = t_stream;
    return 0;
}

bool device_accessible(const void *p)
{
    if (!p)
        return false;
    hipPointerAttribute_t a;
    hipError_t e = hipPointerGetAttributes(&a, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged ||
           a.type == hipMemoryTypeUnified;
}

int current_device(int *dev)
{
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) {
        (void)hipGetLastError();
        return fail(-1, "no HIP device available: the MI355X engine cannot run (no CPU fallback)");
    }
    KFA_HIP(hipGetDevice(dev));
    return 0;
}

// A call's view of one operand: the caller's pointer, or a stream-ordered HBM copy.
int stage_in(const void *p, size_t bytes, hipStream_t s, Operand &op)
{
    op.bytes = bytes;
    if (device_accessible(p)) {
        op.dev = const_cast<void *>(p);
        return 0;
    }
    if (thread_noalloc())
        return noalloc_refuse("a staging copy of a host buffer");
    KFA_HIP(scratch_alloc(&op.dev, bytes ? bytes : 1, s));
    op.staged = true;
    op.s = s;
    op.host = p;
    if (bytes) {
        KFA_HIP(hipMemcpyAsync(op.dev, p, bytes, hipMemcpyHostToDevice, s));
        g_h2d += (long long)bytes;
    }
    return 0;
}

int stage_out(void *p, size_t bytes, hipStream_t s, Operand &op, bool gaps)
{
    op.bytes = bytes;
    if (device_accessible(p)) {
        op.dev = p;
        return 0;
    }
    if (thread_noalloc())
        return noalloc_refuse("a staging copy of a host buffer");
    KFA_HIP(scratch_alloc(&op.dev, bytes ? bytes : 1, s));
    op.staged = true;
    op.s = s;
    op.host = p;
    if (gaps && bytes) {  // the copy back covers the whole span
        KFA_HIP(hipMemcpyAsync(op.dev, p, bytes, hipMemcpyHostToDevice, s));
        g_h2d += (long long)bytes;
    }
    return 0;
}

int finish_out(Operand &op, hipStream_t s)
{
    if (op.staged && op.bytes) {
        KFA_HIP(hipMemcpyAsync(const_cast<void *>(op.host), op.dev, op.bytes, hipMemcpyDeviceToHost, s));
        g_d2h += (long long)op.bytes;
    }
    return 0;
}

void release(Operand &op, hipStream_t s)
{
    if (op.staged && op.dev)
        (void)hipFreeAsync(op.dev, s);
    op.dev = nullptr;
}


// ------------------------------------------------------------------------------------
// launches of the line kernel
// ------------------------------------------------------------------------------------
struct LineLaunch {
    const void *in;
    void *out;
    const void *tw;
    const void *stw;
    const LinePlan *pl;
    LineIO io;
    int colmode;
    int prologue;
    int epilogue;
};

template <class A>
static int launch_lines_t(const LineLaunch &ll, hipStream_t s)
{
    typedef typename A::store2 S;
    const int L = ll.pl->L;
    const size_t cb = sizeof(S);
    const size_t per_line = 2 * (size_t)L * cb;
    if (per_line > kLdsMax)
        return fail(-2, "line of %d points exceeds LDS", L);
    int tl;
    if (ll.colmode) {
        tl = L >= 256 ? 16 : 64;
    } else {
        tl = 2048 / L;
        if (tl < 1)
            tl = 1;
        if (tl > 64)
            tl = 64;
    }
    while (tl > 1 && per_line * tl > 64 * 1024)
        tl >>= 1;
    if ((long long)tl > ll.io.nlines)
        tl = (int)ll.io.nlines;
    if (tl < 1)
        tl = 1;
    const size_t lds = per_line * tl;
    static std::once_flag once;
    std::call_once(once, [] {
        (void)hipFuncSetAttribute((const void *)k_lines<A>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)kLdsMax);
    });
    const long long blocks = (ll.io.nlines + tl - 1) / tl;
    if (blocks <= 0)
        return 0;
    if (blocks > 0x7fffffffLL)
        return fail(-2, "too many lines (%lld) for one launch", ll.io.nlines);
    hipLaunchKernelGGL(k_lines<A>, dim3((unsigned)blocks), dim3(256), lds, s, (const S *)ll.in, (S *)ll.out,
                       (const S *)ll.tw, (const S *)ll.stw, *ll.pl, ll.io, tl, ll.colmode, ll.prologue,
                       ll.epilogue);
    KFA_HIP(hipGetLastError());
    ++g_launches;
    return 0;
}

static int launch_lines(int type, const LineLaunch &ll, hipStream_t s)
{
    switch (type) {
    case KFA_F32: return launch_lines_t<F32>(ll, s);
    case KFA_Q15: return launch_lines_t<Q15>(ll, s);
    case KFA_Q31: return launch_lines_t<Q31>(ll, s);
    default: return fail(-2, "unknown scalar kind %d", type);
    }
}

// Whether device ranges [a, a+na) and [b, b+nb) overlap.
static bool overlap(const void *a, size_t na, const void *b, size_t nb)
{
    const uintptr_t x = reinterpret_cast<uintptr_t>(a), y = reinterpret_cast<uintptr_t>(b);
    return x < y + nb && y < x + na;
}

template <class A>
static int launch_dft_stage_t(DftStageArgs a, hipStream_t s)
{
    a.qtiles = (a.p + 255) / 256;
    const long long blocks = a.ncols * a.qtiles;
    if (blocks <= 0)
        return 0;
    if (blocks > 0x7fffffffLL)
        return fail(-2, "too many columns (%lld) for one DFT-stage launch", a.ncols);
    hipLaunchKernelGGL(k_dft_stage<A>, dim3((unsigned)blocks), dim3(256), 0, s, a);
    KFA_HIP(hipGetLastError());
    ++g_launches;
    return 0;
}

static int launch_dft_stage(int type, const DftStageArgs &a, hipStream_t s)
{
    switch (type) {
    case KFA_F32: return launch_dft_stage_t<F32>(a, s);
    case KFA_Q15: return launch_dft_stage_t<Q15>(a, s);
    case KFA_Q31: return launch_dft_stage_t<Q31>(a, s);
    default: return fail(-2, "unknown scalar kind %d", type);
    }
}

// Chunks of a batch that does not fit one temporary: as many as the room requires, of
// nearly equal size (256 transforms at 127 per chunk run as 86 + 86 + 84, not 127 + 127 + 2:
// a launch of a few transforms costs almost as much as a full one).
static long long balanced_chunk(long long howmany, long long max_chunk)
{
    const long long n = (howmany + max_chunk - 1) / max_chunk;
    return (howmany + n - 1) / n;
}

size_t multipass_bytes_per_transform(DevPlan &P, int fk)
{
    if (P.single)
        return 0;
    const size_t cb = cpx_bytes(P.type);
    const size_t c2c = (P.type == KFA_F32 ? 8 : cb) * (size_t)P.n + 256;  // the four-step temporary
    if (fk == FK_C2C)
        return c2c;
    // real: the temporary of the fused passes or of the split / merge pass, plus the complex
    // pass's own when the passes are not fused
    return cb * (size_t)(P.n + 16) + 256 + c2c;
}

// Runs `howmany` complex transforms of plan P between two device buffers.
// in: element t of item b at in + (b*idist + t*istride); out: out + (b*odist + t*ostride).
int run_c2c(DevPlan &P, const void *in, long long istride, long long idist, void *out,
                   long long ostride, long long odist, long long howmany, hipStream_t s)
{
    const int n = P.n;
    if (howmany <= 0 || n <= 0)
        return 0;
    if (!P.single) {
        // multi-pass lines use a temporary as large as their data: run large batches in
        // chunks whose temporary fits the scratch room (workspace or chunk limit)
        const size_t cb = cpx_bytes(P.type);
        const size_t per = multipass_bytes_per_transform(P, FK_C2C);
        const size_t room = scratch_room();
        if (room < per)
            return fail(-2, "nfft=%d needs %zu bytes of scratch per transform; %zu available", n, per, room);
        long long chunk = std::max(1LL, (long long)(room / per));
        if (howmany > chunk) {
            chunk = balanced_chunk(howmany, chunk);
            for (long long b = 0; b < howmany; b += chunk) {
                const int rc = run_c2c(P, (const char *)in + (size_t)(b * idist) * cb, istride, idist,
                                       (char *)out + (size_t)(b * odist) * cb, ostride, odist,
                                       std::min(chunk, howmany - b), s);
                if (rc)
                    return rc;
            }
            return 0;
        }
    }
    if (P.single) {
        const FastChoice *fc = nullptr;
        int rc = choose_fast(P, FK_C2C, s, &fc, howmany);
        if (rc)
            return rc;
        if (fc->v) {
            FastArgs a{};
            a.in = in;
            a.out = out;
            a.stw = fc->d_stw;
            a.nlines = howmany;
            a.in_hi = idist;
            a.in_el = istride;
            a.out_hi = odist;
            a.out_el = ostride;
            ++g_launches;
            return fast_launch(*fc->v, a, s) ? fail(-3, "fast kernel launch failed") : 0;
        }
        std::shared_ptr<const BlueChoice> bc;
        if ((rc = blue_choice(P, bc)))
            return rc;
        if (bc->v) {
            BlueArgs b{};
            b.in = in;
            b.out = out;
            blue_args(*bc, n, b);
            b.nlines = howmany;
            b.in_hi = idist;
            b.in_el = istride;
            b.out_hi = odist;
            b.out_el = ostride;
            ++g_launches;
            return bc->v->launch(b, s) ? fail(-3, "Bluestein kernel launch failed") : 0;
        }
        LineLaunch ll{};
        ll.in = in;
        ll.out = out;
        ll.tw = P.d_tw;
        ll.pl = &P.lp;
        ll.io.nlines = howmany;
        ll.io.nlo = 1;
        ll.io.in_hi = idist;
        ll.io.in_lo = 0;
        ll.io.in_el = istride;
        ll.io.out_hi = odist;
        ll.io.out_lo = 0;
        ll.io.out_el = ostride;
        ll.colmode = 0;
        return launch_lines(P.type, ll, s);
    }
    {
        std::shared_ptr<const BaileyPlan> bp;
        if (int rc = bailey_plan(P, s, bp))
            return rc;
        if (bp)
            return run_c2c_bailey(*bp, n, in, istride, idist, out, ostride, odist, howmany, s);
    }
    // four-step in the reference's stage order: pass 1 into a temporary, pass 2 into out
    DftStageArgs d{};
    if (P.bigp) {
        d.tw = P.d_tw;
        d.n = n;
        d.Q = P.Q;
        d.p = P.L1;
        d.ncols = howmany * P.Q;
        d.in = in;
        d.in_hi = idist;
        d.in_el = istride;
        if (P.Q == 1) {  // n is the prime itself: one pass
            d.out = out;
            d.out_hi = odist;
            d.out_el = ostride;
            const size_t cb = cpx_bytes(P.type);
            const size_t ispan = cb * (size_t)((howmany - 1) * idist + (n - 1) * istride + 1);
            const size_t ospan = cb * (size_t)((howmany - 1) * odist + (n - 1) * ostride + 1);
            if (!overlap(in, ispan, out, ospan))
                return launch_dft_stage(P.type, d, s);
            // in place: the workgroups of a transform read all of it but write parts of it,
            // so the stage writes a copy of the output span (holding its other values too)
            Scratch mirror;
            int rc = mirror.get(ospan, s);
            if (rc)
                return rc;
            if (hipMemcpyAsync(mirror.p, out, ospan, hipMemcpyDeviceToDevice, s) != hipSuccess)
                return fail(-3, "device copy failed");
            d.out = mirror.p;
            if ((rc = launch_dft_stage(P.type, d, s)))
                return rc;
            return hipMemcpyAsync(out, mirror.p, ospan, hipMemcpyDeviceToDevice, s) == hipSuccess
                       ? 0 : fail(-3, "device copy failed");
        }
    }
    Scratch tmp;
    if (int rc = tmp.get(cpx_bytes(P.type) * (size_t)n * (size_t)howmany, s))
        return rc;
    int rc;
    if (P.bigp) {
        d.out = tmp.p;
        d.out_hi = n;
        d.out_el = 1;
        rc = launch_dft_stage(P.type, d, s);
    } else {
        LineLaunch a{};
        a.in = in;
        a.out = tmp.p;
        a.tw = P.d_tw;
        a.pl = &P.lp1;
        a.io.nlines = howmany * P.Q;
        a.io.nlo = P.Q;
        a.io.in_hi = idist;
        a.io.in_lo = istride;
        a.io.in_el = (long long)P.Q * istride;
        a.io.out_hi = n;
        a.io.out_lo = 1;
        a.io.out_el = P.Q;
        a.colmode = 1;
        rc = launch_lines(P.type, a, s);
    }
    if (rc == 0) {
        LineLaunch b{};
        b.in = tmp.p;
        b.out = out;
        b.tw = P.d_tw;
        b.pl = &P.lp2;
        b.io.nlines = howmany * P.L1;
        b.io.nlo = P.L1;
        b.io.in_hi = n;
        b.io.in_lo = P.Q;
        b.io.in_el = 1;
        b.io.out_hi = odist;
        b.io.out_lo = ostride;
        b.io.out_el = (long long)P.L1 * ostride;
        b.colmode = 1;
        rc = launch_lines(P.type, b, s);
    }
    return rc;
}

template <class A>
static int launch_real_pass(int merge, const void *src, void *dst, const void *stw, int nc, long long hi,
                            long long howmany, hipStream_t s)
{
    typedef typename A::store2 S;
    const unsigned bx = (unsigned)((nc / 2 + 1 + 255) / 256);
    const unsigned by = (unsigned)(howmany < 65535 ? howmany : 65535);
    if (merge)
        hipLaunchKernelGGL(k_real_merge<A>, dim3(bx, by), dim3(256), 0, s, (const S *)src, (S *)dst, (const S *)stw,
                           nc, hi, howmany);
    else
        hipLaunchKernelGGL(k_real_split<A>, dim3(bx, by), dim3(256), 0, s, (const S *)src, (S *)dst, (const S *)stw,
                           nc, hi, howmany);
    KFA_HIP(hipGetLastError());
    ++g_launches;
    return 0;
}

// Float real rows of half length nc = N1*N2 in two HBM passes, the split / merge fused
// into the four-step passes (kfa_tile_real.hpp), decided once per plan and direction:
// the fused tile (length f) and the plain tile of the other half, with their tables.
// H.rfused[inverse] stays null when no registered split of nc fits.
static int real_fused_plan(DevPlan &H, int inverse, hipStream_t s, std::shared_ptr<const RealFusedPlan> &out)
{
    std::lock_guard<std::mutex> g(H.mu);
    if (H.rfused_decided[inverse]) {
        out = H.rfused[inverse];
        return 0;
    }
    out = nullptr;
    if (H.type != KFA_F32 || H.single || !env_flag("KISSFFT_AMD_FAST", true) ||
        !env_flag("KISSFFT_AMD_REAL_FUSED", true)) {
        H.rfused_decided[inverse] = true;
        return 0;
    }
    if (thread_noalloc())
        return noalloc_refuse("the fused real four-step plan");
    const long long nc = H.n;
    // the fused mapping sets the length of its own half: N1 (split rows) or N2 (merge columns)
    const int fused_mode = inverse ? TILE_RMERGE : TILE_RSPLIT;
    const int plain_mode = inverse ? TILE_ROWS_TO_COLS : TILE_COLS_TWID;
    long long bestf = 0;
    for (const TileVariant &v : tile_registry()) {
        if (v.mode != fused_mode || v.inverse != inverse || nc % v.n != 0)
            continue;
        const long long other = nc / v.n;
        if (other < 32 || other % 2 || (other / 2) % (v.TL / 2))
            continue;
        bool have = false;
        for (const TileVariant &w : tile_registry())
            have |= w.mode == plain_mode && w.inverse == inverse && w.n == other;
        if (!have)
            continue;
        auto score = [&](long long f) { return f > nc / f ? f : nc / f; };
        if (bestf == 0 || score(v.n) < score(bestf))
            bestf = v.n;
    }
    if (bestf) {
        auto rp = std::make_shared<RealFusedPlan>();
        rp->f = bestf;
        rp->other = nc / bestf;
        BufPair fs, rs;
        int rc;
        if ((rc = tile_choice((int)bestf, inverse, fused_mode, s, rp->cf)) ||
            (rc = tile_choice((int)rp->other, inverse, plain_mode, s, rp->cp)) ||
            (rc = fourstep_twiddles(nc, inverse, fs)) || (rc = real_split_twiddles(nc, (int)bestf, inverse, rs)))
            return rc;
        rp->tw_hi = fs.first;
        rp->tw_lo = fs.second;
        rp->rtw_a = rs.first;
        rp->rtw_b = rs.second;
        if (rp->cf->v && rp->cp->v && (rp->other / 2) % (rp->cf->v->TL / 2) == 0)
            H.rfused[inverse] = rp;
    }
    H.rfused_decided[inverse] = true;
    out = H.rfused[inverse];
    return 0;
}

// Forward: the column pass of z (length N2, four-step twiddle) into tmp, then row pairs of
// tmp (length N1) split into the bins.  Inverse: column pairs of the bins merged and
// transformed (length N2, twiddle) into tmp, then the rows of tmp (length N1) written as
// columns of z.  tmp holds (nc + 16) * howmany + 16 values.
static int run_real_fourstep(const RealFusedPlan &fp, long long nc, int inverse, const void *in, long long in_hi,
                             void *out, long long out_hi, long long howmany, void *tmp, hipStream_t s)
{
    const long long N1 = inverse ? fp.other : fp.f, N2 = inverse ? fp.f : fp.other;
    const TileChoice &cf = *fp.cf, &cp = *fp.cp;
    RealTileArgs r{};
    r.nc = nc;
    r.rtw_a = fp.rtw_a.get();
    r.rtw_b = fp.rtw_b.get();
    TileArgs a{};
    a.tw_hi = fp.tw_hi.get();
    a.tw_lo = fp.tw_lo.get();
    int rc;
    if (!inverse) {
        // pass 1: line (b, n1): element n2 at in[b*in_hi + n1 + N1*n2] -> tmp[b*nc + n1 + N1*k2]
        a.in = in;
        a.out = tmp;
        a.nlines = howmany * N1;
        a.nlo = N1;
        a.in_hi = in_hi;
        a.in_lo = 1;
        a.in_el = N1;
        a.out_hi = nc;
        a.out_lo = 1;
        a.out_el = N1;
        rc = launch_tile(cp, a, s);
        // pass 2: rows k2 and N2-k2 of tmp -> bins
        r.in = tmp;
        r.out = out;
        r.in_hi = nc;
        r.out_hi = out_hi;
        r.other = (int)N2;
        if (rc == 0)
            rc = launch_rtile(cf, r, howmany, s);
    } else {
        // pass 1: column pairs (n1, N1-n1) of the bins -> tmp[b*tp + n1 + N1*k2].  Each
        // transform of tmp sits at the same offset in its 128-byte line as its bins, so
        // the kernel's line-aligned runs are aligned on both sides (tp = in_hi mod 16).
        const long long tp = nc + ((in_hi - nc) & 15);
        const uintptr_t want = (reinterpret_cast<uintptr_t>(in) / 8) & 15;
        void *tmp_al = (char *)tmp + 8 * ((want - reinterpret_cast<uintptr_t>(tmp) / 8) & 15);
        r.in = in;
        r.out = tmp_al;
        r.tw_hi = a.tw_hi;
        r.tw_lo = a.tw_lo;
        r.in_hi = in_hi;
        r.out_hi = tp;
        r.other = (int)N1;
        rc = launch_rtile(cf, r, howmany, s);
        // pass 2: line (b, k2): element n1 at tmp[b*tp + n1 + N1*k2] -> z[b*out_hi + k2 + N2*k1]
        a.in = tmp_al;
        a.out = out;
        a.nlines = howmany * N2;
        a.nlo = N2;
        a.in_hi = tp;
        a.in_lo = N1;
        a.in_el = 1;
        a.out_hi = out_hi;
        a.out_lo = 1;
        a.out_el = N2;
        if (rc == 0)
            rc = launch_tile(cp, a, s);
    }
    return rc;
}

int prepare_multipass(DevPlan &P, int fk, hipStream_t s)
{
    if (P.single)
        return 0;
    std::shared_ptr<const BaileyPlan> bp;
    std::shared_ptr<const RealFusedPlan> fp;
    int rc = bailey_plan(P, s, bp);
    if (rc == 0 && fk != FK_C2C)
        rc = real_fused_plan(P, fk == FK_C2R, s, fp);
    return rc;
}

// Real rows whose half length needs the multi-pass complex transform.  Float lengths with
// a registered four-step split take the two fused passes of run_real_fourstep; otherwise
// the complex pass runs into a temporary (nc values per line) and the split (forward) or
// merge (inverse) runs as its own HBM pass on that temporary.
static int run_real_multipass(DevPlan &H, const RealTables &R, int inverse, const void *in, long long in_hi,
                              void *out, long long out_hi, long long howmany, hipStream_t s)
{
    const int nc = H.n;
    const size_t cb = cpx_bytes(H.type);
    const size_t per = multipass_bytes_per_transform(H, inverse ? FK_C2R : FK_R2C) + 1024;
    const size_t room = scratch_room();
    if (room < per)
        return fail(-2, "real nfft=%d needs %zu bytes of scratch per transform; %zu available", 2 * nc, per, room);
    long long chunk = std::max(1LL, (long long)(room / per));
    if (howmany > chunk) {  // bounded temporaries, as for multi-pass complex lines
        chunk = balanced_chunk(howmany, chunk);
        for (long long b = 0; b < howmany; b += chunk) {
            const int rc = run_real_multipass(H, R, inverse, (const char *)in + (size_t)(b * in_hi) * cb, in_hi,
                                              (char *)out + (size_t)(b * out_hi) * cb, out_hi,
                                              std::min(chunk, howmany - b), s);
            if (rc)
                return rc;
        }
        return 0;
    }
    std::shared_ptr<const RealFusedPlan> fp;
    int rc = real_fused_plan(H, inverse, s, fp);
    if (rc)
        return rc;
    Scratch tmp;  // room for the fused inverse's aligned layout (16 more values per transform)
    if ((rc = tmp.get(cb * ((size_t)(nc + 16) * (size_t)howmany + 16), s)))
        return rc;
    if (fp)
        return run_real_fourstep(*fp, nc, inverse, in, in_hi, out, out_hi, howmany, tmp.p, s);
    auto pass = [&](int merge, const void *src, void *dst, long long hi) {
        switch (H.type) {
        case KFA_F32: return launch_real_pass<F32>(merge, src, dst, R.d_stw, nc, hi, howmany, s);
        case KFA_Q15: return launch_real_pass<Q15>(merge, src, dst, R.d_stw, nc, hi, howmany, s);
        default: return launch_real_pass<Q31>(merge, src, dst, R.d_stw, nc, hi, howmany, s);
        }
    };
    if (!inverse) {
        rc = run_c2c(H, in, 1, in_hi, tmp.p, 1, nc, howmany, s);
        if (rc == 0)
            rc = pass(0, tmp.p, out, out_hi);
    } else {
        rc = pass(1, in, tmp.p, in_hi);
        if (rc == 0)
            rc = run_c2c(H, tmp.p, 1, nc, out, 1, out_hi, howmany, s);
    }
    return rc;
}

// Real rows: `howmany` lines, distances in complex units (real side counted in pairs).
int run_real_rows(DevPlan &H, const RealTables &R, int inverse, const void *in, long long in_hi,
                         void *out, long long out_hi, long long howmany, hipStream_t s)
{
    if (howmany <= 0)
        return 0;
    if (!H.single)
        return run_real_multipass(H, R, inverse, in, in_hi, out, out_hi, howmany, s);
    const FastChoice *fc = nullptr;
    int rc = choose_fast(H, inverse ? FK_C2R : FK_R2C, s, &fc, howmany);
    if (rc)
        return rc;
    if (fc->v) {
        FastArgs a{};
        a.in = in;
        a.out = out;
        a.stw = fc->d_stw;
        a.rtw = R.d_stw;
        a.nlines = howmany;
        a.in_hi = in_hi;
        a.in_el = 1;
        a.out_hi = out_hi;
        a.out_el = 1;
        ++g_launches;
        return fast_launch(*fc->v, a, s) ? fail(-3, "fast kernel launch failed") : 0;
    }
    LineLaunch ll{};
    ll.in = in;
    ll.out = out;
    ll.tw = H.d_tw;
    ll.stw = R.d_stw;
    ll.pl = &H.lp;
    ll.io.nlines = howmany;
    ll.io.nlo = 1;
    ll.io.in_hi = in_hi;
    ll.io.in_el = 1;
    ll.io.out_hi = out_hi;
    ll.io.out_el = 1;
    ll.colmode = 0;
    if (inverse)
        ll.prologue = PRO_C2R;
    else
        ll.epilogue = EPI_R2C;
    return launch_lines(H.type, ll, s);
}

static int begin(hipStream_t *s, void *stream, int *dev)
{
    int rc = current_device(dev);
    if (rc)
        return rc;
    if (stream) {
        *s = (hipStream_t)stream;
        return 0;
    }
    return thread_stream(s);
}

static int end_call(hipStream_t s, int sync, int rc)
{
    if (sync || rc) {
        hipError_t e = hipStreamSynchronize(s);
        if (rc == 0 && e != hipSuccess)
            return fail(-3, "stream synchronize: %s", hipGetErrorString(e));
    }
    return rc;
}

int call_begin(void *stream, hipStream_t *s, int *dev) { return begin(s, stream, dev); }
int call_end(hipStream_t s, int sync, int rc) { return end_call(s, sync, rc); }

}  // namespace kfa

using namespace kfa;

// ------------------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------------------
extern "C" {

const char *kfa_last_error(void) { return g_err; }

int kfa_device_available(void)
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return count > 0;
}

void kfa_get_stats(kfa_stats *st)
{
    st->plans_built = g_plans_built.load();
    st->plan_hits = g_plan_hits.load();
    st->launches = g_launches.load();
    st->h2d_bytes = g_h2d.load();
    st->d2h_bytes = g_d2h.load();
}

static int c2c_call(const kfa_cplan *pl, const void *in, long long istride, long long idist, void *out,
                    long long odist, long long howmany, hipStream_t s, int dev, int sync)
{
    int rc;
    std::shared_ptr<DevPlan> P;
    if ((rc = build_cplan(pl, dev, P)))
        return rc;
    const size_t cb = cpx_bytes(pl->type);
    const long long n = pl->nfft;
    if (howmany <= 0)
        return 0;
    const size_t in_bytes = cb * (size_t)((howmany - 1) * idist + (n - 1) * istride + 1);
    const size_t out_bytes = cb * (size_t)((howmany - 1) * odist + n);
    Operand I, O;
    if ((rc = stage_in(in, in_bytes, s, I)) == 0 &&
        (rc = stage_out(out, out_bytes, s, O, howmany > 1 && odist > n)) == 0) {
        // in place is safe when every transform reads and writes its own span, except for
        // the one-pass global DFT stage (its workgroups read a whole transform but write
        // parts of it)
        const bool aliased = !I.staged && !O.staged && overlap(I.dev, in_bytes, O.dev, out_bytes);
        const bool inplace_ok = !aliased || (I.dev == O.dev && istride == 1 && idist == odist &&
                                             !(P->bigp && P->Q == 1));
        if (inplace_ok) {
            rc = run_c2c(*P, I.dev, istride, idist, O.dev, 1, odist, howmany, s);
        } else {
            // overlapping device buffers with layouts that cannot be overwritten in place
            Scratch tmp;
            rc = tmp.get(out_bytes, s);
            // values between the rows (odist > n) keep the caller's bytes
            if (rc == 0 && howmany > 1 && odist > n &&
                hipMemcpyAsync(tmp.p, O.dev, out_bytes, hipMemcpyDeviceToDevice, s) != hipSuccess)
                rc = fail(-3, "device copy failed");
            if (rc == 0)
                rc = run_c2c(*P, I.dev, istride, idist, tmp.p, 1, odist, howmany, s);
            if (rc == 0 && hipMemcpyAsync(O.dev, tmp.p, out_bytes, hipMemcpyDeviceToDevice, s) != hipSuccess)
                rc = fail(-3, "device copy failed");
        }
        if (rc == 0)
            rc = finish_out(O, s);
    }
    release(I, s);
    release(O, s);
    return end_call(s, sync || I.staged || O.staged, rc);
}

int kfa_c2c(const kfa_cplan *pl, const void *in, long long istride, long long idist, void *out,
            long long odist, long long howmany, void *stream, int sync)
{
    hipStream_t s;
    int dev;
    int rc = begin(&s, stream, &dev);
    if (rc)
        return rc;
    return c2c_call(pl, in, istride, idist, out, odist, howmany, s, dev, sync);
}

// ------------------------------------------------------------------------------------
// The no-allocation form of kfa_c2c: the temporaries of multi-pass lengths (and of
// overlapping buffers) come from `scratch`, a file-scope device array each GPU holds for
// the life of the process; every other allocation site refuses (device plans, kernel
// choices and tables must exist from an earlier allocating call; host buffers would need
// staging).  The array's device address is looked up once per device with that device
// current (the host symbol is not a device address).  Every call on a GPU shares that
// GPU's array: calls that overlap on one GPU race.  The temporaries are written in full
// before they are read, so nothing needs clearing between calls.
// ------------------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static std::mutex g_scratch_mu;
static std::map<int, void *> g_scratch_ptr;  // per device: the address of `scratch`

static int noalloc_scratch(int dev, void **ptr)
{
    std::lock_guard<std::mutex> g(g_scratch_mu);
    auto it = g_scratch_ptr.find(dev);
    if (it == g_scratch_ptr.end()) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(-3, "no-allocation scratch: hipGetSymbolAddress failed: %s", hipGetErrorString(e));
        it = g_scratch_ptr.emplace(dev, p).first;
    }
    *ptr = it->second;
    return 0;
}

int kfa_c2c_noalloc(const kfa_cplan *pl, const void *in, long long istride, long long idist, void *out,
                    long long odist, long long howmany, void *stream)
{
    int dev;
    int rc = current_device(&dev);
    if (rc)
        return rc;
    // NULL is the null stream here: the calling thread's own stream may not exist yet
    const hipStream_t s = (hipStream_t)stream;
    void *base = nullptr;
    if ((rc = noalloc_scratch(dev, &base)))
        return rc;
    if (!device_accessible(in) || !device_accessible(out))
        return fail(-2, "no-allocation call: buffers must be device memory");
    Workspace ws;
    ws.base = static_cast<char *>(base);
    ws.size = kNoAllocScratchBytes;
    struct Mode {  // the no-allocation state of this thread, restored on every return
        Workspace *prev_ws;
        bool prev_noalloc;
        Mode(Workspace *w) : prev_ws(thread_workspace()), prev_noalloc(thread_noalloc())
        {
            thread_workspace() = w;
            thread_noalloc() = true;
        }
        ~Mode()
        {
            thread_workspace() = prev_ws;
            thread_noalloc() = prev_noalloc;
        }
    } mode(&ws);
    // a multi-pass transform must fit its temporary in the array (the batch runs in chunks)
    {
        std::shared_ptr<DevPlan> P;
        if ((rc = build_cplan(pl, dev, P)))
            return rc;
        const size_t per = multipass_bytes_per_transform(*P, FK_C2C);
        if (per && per + 4096 > kNoAllocScratchBytes)
            return fail(-2, "no-allocation call: nfft=%d needs %zu bytes of scratch per transform; the static "
                            "scratch holds %zu", pl->nfft, per, kNoAllocScratchBytes);
    }
    return c2c_call(pl, in, istride, idist, out, odist, howmany, s, dev, 0);
}

static int real_common(const kfa_rplan *pl, const void *in, long long idist, void *out, long long odist,
                       long long howmany, void *stream, int sync, int inverse)
{
    hipStream_t s;
    int dev;
    int rc = begin(&s, stream, &dev);
    if (rc)
        return rc;
    std::shared_ptr<DevPlan> P;
    std::shared_ptr<RealTables> R;
    if ((rc = build_cplan(&pl->half, dev, P)))
        return rc;
    if ((rc = build_rtab(pl, dev, R)))
        return rc;
    if (howmany <= 0)
        return 0;
    const long long nc = pl->half.nfft;
    const size_t sb = cpx_bytes(pl->half.type) / 2;  // real scalar bytes
    const size_t cb = 2 * sb;
    // real side counts scalars, complex side counts complex values
    const size_t real_bytes_in = sb * (size_t)((howmany - 1) * idist + 2 * nc);
    const size_t cpx_bytes_in = cb * (size_t)((howmany - 1) * idist + nc + 1);
    const size_t real_bytes_out = sb * (size_t)((howmany - 1) * odist + 2 * nc);
    const size_t cpx_bytes_out = cb * (size_t)((howmany - 1) * odist + nc + 1);
    Operand I, O;
    rc = stage_in(in, inverse ? cpx_bytes_in : real_bytes_in, s, I);
    if (rc == 0)
        rc = stage_out(out, inverse ? real_bytes_out : cpx_bytes_out, s, O,
                       howmany > 1 && odist > (inverse ? 2 * nc : nc + 1));
    if (rc == 0) {
        const size_t in_b = inverse ? cpx_bytes_in : real_bytes_in, out_b = inverse ? real_bytes_out : cpx_bytes_out;
        // rows of different length on the two sides: overlapping buffers are only safe for
        // one single-pass transform in place (its line is read whole before it is written)
        const bool aliased = !I.staged && !O.staged && overlap(I.dev, in_b, O.dev, out_b) &&
                             !(howmany == 1 && I.dev == O.dev && P->single);
        Scratch tmp;
        void *dst = O.dev;
        if (aliased) {
            rc = tmp.get(out_b, s);
            if (rc == 0 && howmany > 1 && odist > (inverse ? 2 * nc : nc + 1) &&
                hipMemcpyAsync(tmp.p, O.dev, out_b, hipMemcpyDeviceToDevice, s) != hipSuccess)
                rc = fail(-3, "device copy failed");
            dst = tmp.p;
        }
        if (rc)
            ;
        else if (!inverse && (idist & 1))
            rc = fail(-2, "real input distance must be even (complex-pair loads)");
        else if (inverse && (odist & 1))
            rc = fail(-2, "real output distance must be even (complex-pair stores)");
        else if (!inverse)
            rc = run_real_rows(*P, *R, 0, I.dev, idist / 2, dst, odist, howmany, s);
        else
            rc = run_real_rows(*P, *R, 1, I.dev, idist, dst, odist / 2, howmany, s);
        if (rc == 0 && aliased && hipMemcpyAsync(O.dev, tmp.p, out_b, hipMemcpyDeviceToDevice, s) != hipSuccess)
            rc = fail(-3, "device copy failed");
        if (rc == 0)
            rc = finish_out(O, s);
    }
    release(I, s);
    release(O, s);
    return end_call(s, sync || I.staged || O.staged, rc);
}

int kfa_r2c(const kfa_rplan *pl, const void *in, long long idist, void *out, long long odist,
            long long howmany, void *stream, int sync)
{
    return real_common(pl, in, idist, out, odist, howmany, stream, sync, 0);
}

int kfa_c2r(const kfa_rplan *pl, const void *in, long long idist, void *out, long long odist,
            long long howmany, void *stream, int sync)
{
    return real_common(pl, in, idist, out, odist, howmany, stream, sync, 1);
}

// One axis of an N-D array batch, dims[k] long, between device buffers (in may equal out).
static int run_axis(DevPlan &P, const void *in, void *out, long long outer, long long inner,
                    hipStream_t s)
{
    const long long len = P.n;
    if (inner == 1)  // the last axis: contiguous lines
        return run_c2c(P, in, 1, len, out, 1, len, outer, s);
    if (!P.single) {
        // an axis longer than one LDS line: the strided lines of each outer slice through the
        // multi-pass path (element e of line j at j + e*inner)
        const size_t cb = cpx_bytes(P.type);
        for (long long o = 0; o < outer; ++o) {
            const size_t off = cb * (size_t)(o * len * inner);
            if (int rc = run_c2c(P, (const char *)in + off, inner, 1, (char *)out + off, inner, 1, inner, s))
                return rc;
        }
        return 0;
    }
    if (P.type == KFA_F32) {
        std::shared_ptr<const TileChoice> tc;
        int rc = tile_choice((int)len, P.inverse, TILE_COLS, s, tc, inner, outer);
        if (rc)
            return rc;
        if (tc->v) {
            TileArgs a{};
            a.in = in;
            a.out = out;
            a.nlines = outer * inner;
            a.nlo = inner;
            a.in_hi = len * inner;
            a.in_lo = 1;
            a.in_el = inner;
            a.out_hi = len * inner;
            a.out_lo = 1;
            a.out_el = inner;
            return launch_tile(*tc, a, s);
        }
    }
    LineLaunch ll{};
    ll.in = in;
    ll.out = out;
    ll.tw = P.d_tw;
    ll.pl = &P.lp;
    ll.io.nlines = outer * inner;
    ll.io.nlo = inner;
    ll.io.in_hi = len * inner;
    ll.io.in_lo = 1;
    ll.io.in_el = inner;
    ll.io.out_hi = len * inner;
    ll.io.out_lo = 1;
    ll.io.out_el = inner;
    ll.colmode = inner > 1 ? 1 : 0;
    return launch_lines(P.type, ll, s);
}

}  // extern "C" (reopened below)

namespace kfa {

// Complex N-D transform of `howmany` arrays between device buffers: axis 0 first, as
// kiss_fftnd.c:181-186 does (fixed point rounds per axis, so the order is part of the
// result); the first axis reads the input, later axes run in place on the output.
int run_nd(const std::vector<std::shared_ptr<DevPlan>> &P, const int *dims, const void *in, void *out,
           long long howmany, hipStream_t s)
{
    const int ndims = (int)P.size();
    long long total = 1;
    for (int k = 0; k < ndims; ++k)
        total *= dims[k];
    long long outer = howmany, inner = total;
    const void *src = in;
    int rc = 0;
    int paired = -1;  // float: the axis done together with the last one in a plane pass
    std::shared_ptr<const PlaneChoice> pc;
    if (ndims >= 2 && P[ndims - 1]->type == KFA_F32) {
        // the last axis and one other in one pass when a plane (or slab) of them fits in
        // LDS: the axis before the last first (contiguous planes), then the outer axes from
        // the first (float results do not depend on the order of the axes beyond rounding)
        const int nb = dims[ndims - 1];
        std::vector<int> order{ndims - 2};
        for (int k = 0; k < ndims - 2; ++k)
            order.push_back(k);
        for (int k : order) {
            long long rs = 1;  // in-plane offsets are 32-bit
            for (int j = k + 1; j < ndims; ++j)
                rs *= dims[j];
            if ((long long)(dims[k] - 1) * rs + nb >= (1LL << 31))
                continue;
            long long mid = 1;
            for (int j = k + 1; j < ndims - 1; ++j)
                mid *= dims[j];
            const long long planes = howmany * (total / ((long long)dims[k] * nb));
            if ((rc = plane_choice(dims[k], nb, P[ndims - 1]->inverse, k != ndims - 2, (int)rs, mid, planes, s, pc)))
                return rc;
            if (pc->v && (k == ndims - 2 || pc->v->slab)) {
                paired = k;
                break;
            }
        }
        if (paired >= 0) {
            long long rs = 1, mid = 1;  // row stride; planes per value of the axes before `paired`
            for (int k = paired + 1; k < ndims; ++k)
                rs *= dims[k];
            for (int k = paired + 1; k < ndims - 1; ++k)
                mid *= dims[k];
            const long long planes = howmany * (total / ((long long)dims[paired] * nb));
            const PlaneArgs pa{in, out, pc->stw_a, pc->stw_b, planes, (int)rs, mid, (long long)dims[paired] * rs, nb};
            ++g_launches;
            if (pc->v->launch(pa, s))
                return fail(-3, "plane kernel launch failed");
            src = out;
        }
    }
    for (int k = 0; k < ndims && rc == 0; ++k) {
        inner /= dims[k];
        if (paired < 0 || (k != paired && k != ndims - 1)) {
            rc = run_axis(*P[k], src, out, outer, inner, s);
            src = out;
        }
        outer *= dims[k];
    }
    return rc;
}

// Real N-D transform between device buffers (kiss_fftndr.c:99-109 forward: real rows,
// then the complex axes per bin column; kiss_fftndr.c:121-131 inverse: the complex axes
// into a temporary, then the real rows).  P holds the leading axes, H/R the real axis.
int run_ndr(DevPlan &H, const RealTables &R, const std::vector<std::shared_ptr<DevPlan>> &P, const int *dims,
            int inverse, const void *in, void *out, long long howmany, hipStream_t s)
{
    const int nlead = (int)P.size();
    long long rows = 1;
    for (int k = 0; k < nlead; ++k)
        rows *= dims[k];
    const long long nreal = dims[nlead], nbins = nreal / 2 + 1;
    int rc = 0;
    if (!inverse) {
        rc = run_real_rows(H, R, 0, in, nreal / 2, out, nbins, rows * howmany, s);
        long long outer = howmany, inner = rows * nbins;
        for (int k = 0; k < nlead && rc == 0; ++k) {
            inner /= dims[k];
            rc = run_axis(*P[k], out, out, outer, inner, s);
            outer *= dims[k];
        }
        return rc;
    }
    // the input is left untouched: the axes run out of place into a temporary
    Scratch tmp;
    const size_t tb = cpx_bytes(H.type) * (size_t)(rows * nbins * howmany);
    if ((rc = tmp.get(tb, s)))
        return rc;
    long long outer = howmany, inner = rows * nbins;
    const void *src = in;
    for (int k = 0; k < nlead && rc == 0; ++k) {
        inner /= dims[k];
        rc = run_axis(*P[k], src, tmp.p, outer, inner, s);
        outer *= dims[k];
        src = tmp.p;
    }
    if (rc == 0 && nlead == 0)
        rc = hipMemcpyAsync(tmp.p, in, tb, hipMemcpyDeviceToDevice, s) == hipSuccess ? 0 : fail(-3, "copy failed");
    if (rc == 0)
        rc = run_real_rows(H, R, 1, tmp.p, nbins, out, nreal / 2, rows * howmany, s);
    return rc;
}

}  // namespace kfa

extern "C" {

int kfa_nd(int ndims, const int *dims, const kfa_cplan *axes, const void *in, void *out,
           long long howmany, void *stream, int sync)
{
    hipStream_t s;
    int dev;
    int rc = begin(&s, stream, &dev);
    if (rc)
        return rc;
    if (ndims <= 0 || howmany <= 0)
        return 0;
    std::vector<std::shared_ptr<DevPlan>> P(ndims);
    long long total = 1;
    for (int k = 0; k < ndims; ++k) {
        if ((rc = build_cplan(&axes[k], dev, P[k])))
            return rc;
        total *= dims[k];
    }
    const size_t bytes = cpx_bytes(axes[0].type) * (size_t)(total * howmany);
    Operand I, O;
    rc = stage_in(in, bytes, s, I);
    if (rc == 0)
        rc = stage_out(out, bytes, s, O);
    if (rc == 0) {
        // overlapping but not identical buffers: the passes run into a temporary
        const bool aliased = !I.staged && !O.staged && I.dev != O.dev && overlap(I.dev, bytes, O.dev, bytes);
        Scratch tmp;
        if (aliased)
            rc = tmp.get(bytes, s);
        if (rc == 0)
            rc = run_nd(P, dims, I.dev, aliased ? tmp.p : O.dev, howmany, s);
        if (rc == 0 && aliased && hipMemcpyAsync(O.dev, tmp.p, bytes, hipMemcpyDeviceToDevice, s) != hipSuccess)
            rc = fail(-3, "device copy failed");
        if (rc == 0)
            rc = finish_out(O, s);
    }
    release(I, s);
    release(O, s);
    return end_call(s, sync || I.staged || O.staged, rc);
}

int kfa_ndr(int ndims, const int *dims, const kfa_rplan *rpl, const kfa_cplan *axes, int inverse,
            const void *in, void *out, long long howmany, void *stream, int sync)
{
    hipStream_t s;
    int dev;
    int rc = begin(&s, stream, &dev);
    if (rc)
        return rc;
    if (ndims <= 0 || howmany <= 0)
        return 0;
    std::shared_ptr<DevPlan> H;
    std::shared_ptr<RealTables> R;
    if ((rc = build_cplan(&rpl->half, dev, H)) || (rc = build_rtab(rpl, dev, R)))
        return rc;
    std::vector<std::shared_ptr<DevPlan>> P(ndims > 1 ? ndims - 1 : 0);
    long long rows = 1;
    for (int k = 0; k + 1 < ndims; ++k) {
        if ((rc = build_cplan(&axes[k], dev, P[k])))
            return rc;
        rows *= dims[k];
    }
    const long long nreal = dims[ndims - 1];
    const long long nbins = nreal / 2 + 1;
    const size_t sb = cpx_bytes(rpl->half.type) / 2;
    const size_t real_bytes = sb * (size_t)(rows * nreal * howmany);
    const size_t cpx_bytes_ = 2 * sb * (size_t)(rows * nbins * howmany);
    Operand I, O;
    rc = stage_in(in, inverse ? cpx_bytes_ : real_bytes, s, I);
    if (rc == 0)
        rc = stage_out(out, inverse ? real_bytes : cpx_bytes_, s, O);
    if (rc == 0) {
        // the forward real rows write rows of nbins complex over the following rows' real
        // input: overlapping buffers run into a temporary (the inverse reads its input
        // whole into its own temporary first)
        const size_t in_b = inverse ? cpx_bytes_ : real_bytes, out_b = inverse ? real_bytes : cpx_bytes_;
        const bool aliased = !inverse && !I.staged && !O.staged && overlap(I.dev, in_b, O.dev, out_b);
        Scratch tmp;
        if (aliased)
            rc = tmp.get(out_b, s);
        if (rc == 0)
            rc = run_ndr(*H, *R, P, dims, inverse, I.dev, aliased ? tmp.p : O.dev, howmany, s);
        if (rc == 0 && aliased && hipMemcpyAsync(O.dev, tmp.p, out_b, hipMemcpyDeviceToDevice, s) != hipSuccess)
            rc = fail(-3, "device copy failed");
        if (rc == 0)
            rc = finish_out(O, s);
    }
    release(I, s);
    release(O, s);
    return end_call(s, sync || I.staged || O.staged, rc);
}

int kfa_stream_sync(void *stream)
{
    hipStream_t s;
    int dev;
    int rc = begin(&s, stream, &dev);
    if (rc)
        return rc;
    KFA_HIP(hipStreamSynchronize(s));
    return 0;
}

// Drops the engine's caches.  Plans and tables still held elsewhere -- calls in flight on
// other threads, requests queued in a server, multi-batches -- stay valid until their last
// holder lets go; everything released is freed here (hipFree waits for the device), or at
// a later cleanup for what is released later.  (Graphs that a caller captured from *_async
// calls hold no reference: they must not be replayed after a cleanup.)
void kfa_cleanup(void)
{
    {
        auto &C = cache();
        std::lock_guard<std::mutex> g(C.mu);
        C.cplans.clear();
        C.rtabs.clear();
    }
    {
        auto &T = tile_cache();
        std::lock_guard<std::mutex> gt(T.mu);
        T.plane.clear();
        T.choice.clear();
        T.fs.clear();
        T.rs.clear();
    }
    conv_release();
    free_deferred();
    scratch_trim();
}

}  // extern "C"
