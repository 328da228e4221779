// kfa_multi.hip -- several batches of different lengths and kinds in one launch sequence.
//
// The batches (items) are fixed when the multi-batch is created: kind, length, device
// buffers and count.  Creation builds (or finds) each length's device plan and decides, once,
// how each item runs:
//   * items whose (kind, length) has a compiled row kernel are grouped by (kind, length):
//     one launch of that kernel per group, its segment table mapping lines to the items'
//     buffers (as the request server does), the variant chosen by the engine's tuner;
//   * float lengths with a prime factor above 16 run the Bluestein kernel, grouped by
//     (kind, length) the same way;
//   * every other single-pass item (generic radices, fixed-point lengths without a compiled
//     kernel) joins one persistent launch of k_multi (kfa_lines.hpp): a grid sized to the
//     device walks a global list of tiles with the general LDS-resident line code, any mix of
//     radices and kinds together, bit-exact in the fixed-point builds;
//   * items longer than one LDS line run as their own multi-pass launches; their temporaries
//     are carved from one workspace allocated at creation (sized for the largest item's
//     chunk), and their lazily built choices (tile variants, four-step tables) are decided
//     at creation too, so these launches are capturable as well.
// The compiled-kernel and Bluestein groups are then fused into one launch where the
// runtime compiler allows (fuse_groups below): one grid of one workgroup size running
// every group's own kernel body on its range of workgroups.
// KISSFFT_AMD_MULTI_PERSISTENT=1 puts every single-pass item into the persistent launch.
// Execution issues the launches back to back on the caller's stream (largest group first);
// capture records them into one hipGraph.  KISSFFT_AMD_MULTI_HELPERS=n (at most 8) forks
// them round robin over n helper streams joined on the caller's stream instead: measured
// on 64 uneven batches as 8 separate launches (KISSFFT_AMD_MULTI_FUSED=0), they did not overlap and every hand-over
// between queues cost 10-30 us (graph replay 163-170 us with three helpers, 109-114 us
// with one, 78 us with none).  Nothing is allocated and no memset is issued on the
// capturable path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "../kfa_engine.h"
#include "kfa_blue.hpp"
#include "kfa_fast.hpp"
#include "kfa_internal.hpp"

namespace kfa {

struct Multi {
    int type = 0;
    int dev = 0;
    std::vector<std::shared_ptr<DevPlan>> plans;
    std::vector<std::shared_ptr<RealTables>> rtabs;
    std::vector<MultiItem> h;           // persistent items
    MultiItem *d = nullptr;
    long long ntiles = 0;
    size_t lds = 0;
    int grid = 0;
    struct Long {                        // items launched through the engine's batched path
        int kind;
        std::shared_ptr<DevPlan> P;
        std::shared_ptr<RealTables> R;
        const void *in;
        void *out;
        long long howmany;
        int n;
    };
    std::vector<Long> longs;             // multi-pass (their temporaries in `ws`)
    DevBuf ws_mem;
    Workspace ws;
    struct Group {                       // one segmented launch of a compiled row kernel
        const FastVariant *v;
        FastArgs a;
        size_t bytes;
        bool fused;                      // runs inside the fused launch
    };
    std::vector<Group> groups;
    int group_items = 0;
    struct BlueGroup {                   // one segmented Bluestein launch (float primes > 16)
        std::shared_ptr<DevPlan> P;
        std::shared_ptr<const BlueChoice> bc;  // the tables the launches (and a graph) use
        BlueArgs a;
        bool fused;
    };
    std::vector<BlueGroup> blue;
    int blue_items = 0;
    struct Fused {                       // groups of one workgroup size as one grid
        hipFunction_t fn = nullptr;
        FusedTab tab{};
        long long blocks = 0;
        int NT = 0, members = 0;
        size_t lds = 0;
        DevBuf args;                     // the members' argument blocks
    } fused;
    std::vector<std::unique_ptr<FastVariant>> shaped;  // reshaped members of the fused launch
    std::vector<hipStream_t> helpers;    // fork / join streams (default priority)
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
};

static constexpr int kMultiHelpers = 0;  // default helper streams (see the file comment)

template <class A>
static int launch_multi(Multi &M, hipStream_t s)
{
    static std::once_flag once;
    std::call_once(once, [] {
        (void)hipFuncSetAttribute((const void *)k_multi<A>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024);
    });
    hipLaunchKernelGGL(k_multi<A>, dim3((unsigned)M.grid), dim3(256), M.lds, s, (const MultiItem *)M.d,
                       (int)M.h.size(), M.ntiles);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : fail(-3, "multi-batch launch failed: %s", hipGetErrorString(e));
}

static int run_long(const Multi::Long &L, hipStream_t s)
{
    const long long nc = L.P->n;
    if (L.kind == KFA_SERVE_R2C)
        return run_real_rows(*L.P, *L.R, 0, L.in, nc, L.out, nc + 1, L.howmany, s);
    if (L.kind == KFA_SERVE_C2R)
        return run_real_rows(*L.P, *L.R, 1, L.in, nc + 1, L.out, nc, L.howmany, s);
    return run_c2c(*L.P, L.in, 1, nc, L.out, 1, nc, L.howmany, s);
}

// The row kernel of variant v with another thread shape: NT threads per workgroup as
// T x LPB lines, T chosen to keep v's points per thread as nearly as NT allows.  The
// radix schedule, the stage tables and every butterfly stay v's (the thread count only
// deals the butterflies of each stage): fixed-point results are v's bit for bit, float
// results may round differently in the last bit where the compiler forms multiply-adds
// differently for the new shape.  Variants with the real split or merge fused into a stage
// (which ties T to that stage) do not reshape.
static bool reshape(const FastVariant &v, int NT, FastVariant &out, size_t &lds)
{
    if (!v.arith || v.fuse || NT <= 0)
        return false;
    if (v.xpose) {  // k_q15x: 16 lanes per line; lines per workgroup re-dealt (static LDS)
        if (NT % 16 || NT / 16 > 32)
            return false;
        out = v;
        out.LPB = NT / 16;
        out.launch = nullptr;
        out.rtc = nullptr;
        lds = 0;
        return true;
    }
    const int pmin = *std::min_element(v.radices.begin(), v.radices.end());
    const double want = std::log((double)v.n / v.T);
    bool found = false;
    double err = 0;
    for (int lpb = 1; lpb <= NT; ++lpb) {
        if (NT % lpb)
            continue;
        const int T = NT / lpb;
        if (T > v.n / pmin)  // every thread owns a butterfly of the widest stage
            continue;
        const size_t l = (v.db ? 2 : 1) * (size_t)lpb * v.n * (16 / v.tw_pad);
        const double e = std::fabs(std::log((double)v.n / T) - want);
        if (!found || e < err - 1e-9 || (e < err + 1e-9 && l < lds)) {
            found = true;
            err = e;
            lds = l;
            out = v;
            out.T = T;
            out.LPB = lpb;
            out.launch = nullptr;
            out.rtc = nullptr;
        }
    }
    return found;
}

// Puts the compiled-kernel groups and Bluestein groups of one workgroup size into one
// fused launch: a kernel compiled at run time from those groups' own bodies (rtc_fused),
// one grid whose workgroup ranges run the groups in order of size, largest first, so the
// small groups fill the large ones' tail instead of each draining the device on its own.
// Row-kernel groups (k_fast, and the Q15 wavefront-transpose kernel k_q15x) of another
// size join reshaped to that size (reshape above); the size
// is the one whose members carry the most work.  Groups that cannot join (Bluestein of
// another size, fused real split / merge variants of another size, more than 64 KiB of
// LDS), and every group when the runtime compiler is unavailable, keep their own launches.
// KISSFFT_AMD_MULTI_FUSED=0 turns fusing off, KISSFFT_AMD_MULTI_RESHAPE=0 reshaping.
static void fuse_groups(Multi &M)
{
    if (!env_flag("KISSFFT_AMD_MULTI_FUSED", true))
        return;
    struct Cand {
        int g;
        bool blue;
        int NT;
        size_t lds, work;
        long long blocks;
        FusedMember m;
    };
    std::vector<Cand> c;
    constexpr size_t kMaxLds = 64 * 1024;  // dynamic LDS of a module kernel without attributes
    for (size_t i = 0; i < M.groups.size(); ++i) {
        const FastVariant *v = M.groups[i].v;
        if (!v->arith)
            continue;  // another kernel than k_fast / k_q15x
        // (k_q15x: its transpose buffer is static LDS)
        const size_t lds = v->xpose ? 0 : (v->db ? 2 : 1) * (size_t)v->LPB * v->n * (16 / v->tw_pad);
        if (lds > kMaxLds)
            continue;
        const long long blocks = (M.groups[i].a.nlines + v->LPB - 1) / v->LPB;
        c.push_back({(int)i, false, v->T * v->LPB, lds, M.groups[i].bytes, blocks, {v, nullptr}});
    }
    for (size_t i = 0; i < M.blue.size(); ++i) {
        Multi::BlueGroup &G = M.blue[i];
        const BlueVariant *bv = blue_fill(*G.bc, G.P->n, G.a);
        if (!bv || bv->lds > kMaxLds)
            continue;
        const long long blocks = (G.a.nlines + bv->LPB - 1) / bv->LPB;
        // work: the two convolution transforms of M points per line
        const size_t work = (size_t)G.a.nlines * bv->M * sizeof(float2) * 2;
        c.push_back({(int)i, true, bv->T * bv->LPB, bv->lds, work, blocks, {nullptr, bv}});
    }
    // the workgroup size whose members (groups of that size, and row-kernel groups that
    // reshape to it) carry the most work, at least two members
    const bool reshape_on = env_flag("KISSFFT_AMD_MULTI_RESHAPE", true);
    auto member = [&](const Cand &x, int NT, Cand &out) {
        out = x;
        if (x.NT == NT)
            return true;
        if (x.blue || !reshape_on)
            return false;
        FastVariant r;
        if (!reshape(*x.m.fv, NT, r, out.lds) || out.lds > kMaxLds)
            return false;
        out.NT = NT;
        out.blocks = (M.groups[x.g].a.nlines + r.LPB - 1) / r.LPB;
        M.shaped.push_back(std::make_unique<FastVariant>(r));
        out.m.fv = M.shaped.back().get();
        return true;
    };
    int NT = 0;
    size_t best = 0;
    std::vector<int> sizes;
    for (const Cand &x : c)
        if (std::find(sizes.begin(), sizes.end(), x.NT) == sizes.end())
            sizes.push_back(x.NT);
    for (int nt : sizes) {
        int count = 0;
        size_t work = 0;
        Cand y;
        for (const Cand &x : c)
            if (member(x, nt, y)) {
                ++count;
                work += x.work;
            }
        if (count >= 2 && work > best) {
            best = work;
            NT = nt;
        }
    }
    M.shaped.clear();
    if (!NT)
        return;
    std::vector<Cand> mem;
    for (const Cand &x : c) {
        Cand y;
        if (mem.size() < KFA_MAX_FUSED && member(x, NT, y))
            mem.push_back(y);
    }
    std::stable_sort(mem.begin(), mem.end(), [](const Cand &x, const Cand &y) { return x.work > y.work; });
    std::vector<FusedMember> fm;
    for (const Cand &x : mem)
        fm.push_back(x.m);
    hipFunction_t fn = nullptr;
    if (rtc_fused(fm, NT, &fn)) {
        (void)hipGetLastError();
        return;
    }
    // the members' argument blocks, each at a 256-byte boundary
    std::vector<size_t> off;
    size_t total = 0;
    for (const Cand &x : mem) {
        off.push_back(total);
        total += ((x.blue ? sizeof(BlueArgs) : sizeof(FastArgs)) + 255) / 256 * 256;
    }
    std::vector<unsigned char> h(total, 0);
    for (size_t i = 0; i < mem.size(); ++i) {
        if (mem[i].blue)
            memcpy(h.data() + off[i], &M.blue[mem[i].g].a, sizeof(BlueArgs));
        else
            memcpy(h.data() + off[i], &M.groups[mem[i].g].a, sizeof(FastArgs));
    }
    Multi::Fused F;
    if (dev_upload(h.data(), total, F.args)) {
        (void)hipGetLastError();
        return;
    }
    long long end = 0;
    for (size_t i = 0; i < mem.size(); ++i) {
        F.tab.arg[i] = static_cast<const unsigned char *>(F.args.get()) + off[i];
        end += mem[i].blocks;
        F.tab.end[i] = end;
        F.lds = std::max(F.lds, mem[i].lds);
    }
    if (end <= 0 || end > 0x7fffffffLL)
        return;
    F.fn = fn;
    F.blocks = end;
    F.NT = NT;
    F.members = (int)mem.size();
    for (const Cand &x : mem) {
        if (x.blue)
            M.blue[x.g].fused = true;
        else
            M.groups[x.g].fused = true;
    }
    M.fused = std::move(F);
    if (env_flag("KISSFFT_AMD_VERBOSE", false))
        fprintf(stderr, "[kissfft_amd] multi-batch: %d groups fused, %d threads, %lld workgroups\n", M.fused.members,
                NT, end);
}

// Every launch of the multi-batch on s (forked over s and the helper streams and joined
// on s when there are helpers).
static int run_multi(Multi &M, hipStream_t s)
{
    std::vector<hipStream_t> st{s};
    st.insert(st.end(), M.helpers.begin(), M.helpers.end());
    const bool fork = st.size() > 1;
    std::vector<hipEvent_t> ev(fork ? st.size() : 0, nullptr);
    for (auto &e : ev)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            for (auto &x : ev)
                if (x)
                    (void)hipEventDestroy(x);
            return fail(-3, "multi-batch: event creation failed");
        }
    int rc = 0;
    auto check = [&](hipError_t e) {
        if (e != hipSuccess && rc == 0)
            rc = fail(-3, "multi-batch fork / join: %s", hipGetErrorString(e));
    };
    if (fork) {
        check(hipEventRecord(ev[0], s));
        for (size_t i = 1; i < st.size(); ++i)
            check(hipStreamWaitEvent(st[i], ev[0], 0));
    }
    size_t next = 0;
    auto pick = [&]() { return st[next++ % st.size()]; };
    if (M.fused.fn) {
        ++g_launches;
        FusedTab tab = M.fused.tab;
        void *params[] = {&tab};
        const hipError_t e = hipModuleLaunchKernel(M.fused.fn, (unsigned)M.fused.blocks, 1, 1, (unsigned)M.fused.NT, 1,
                                                   1, (unsigned)M.fused.lds, pick(), params, nullptr);
        if (e != hipSuccess)
            rc = fail(-3, "fused multi-batch launch failed: %s", hipGetErrorString(e));
    }
    // groups are sorted largest first at creation
    for (size_t i = 0; i < M.groups.size() && rc == 0; ++i) {
        if (M.groups[i].fused)
            continue;
        ++g_launches;
        if (fast_launch(*M.groups[i].v, M.groups[i].a, pick()))
            rc = fail(-3, "fast kernel launch failed");
    }
    for (size_t i = 0; i < M.blue.size() && rc == 0; ++i)
        if (!M.blue[i].fused)
            rc = blue_launch(*M.blue[i].P, M.blue[i].a, pick());
    if (rc == 0 && M.ntiles > 0) {
        hipStream_t sp = pick();
        switch (M.type) {
        case KFA_F32: rc = launch_multi<F32>(M, sp); break;
        case KFA_Q15: rc = launch_multi<Q15>(M, sp); break;
        default: rc = launch_multi<Q31>(M, sp); break;
        }
    }
    if (!M.longs.empty()) {
        // the long items' temporaries come from the multi-batch's workspace
        Workspace *prev = thread_workspace();
        M.ws.used = 0;
        thread_workspace() = &M.ws;
        for (size_t i = 0; i < M.longs.size() && rc == 0; ++i)
            rc = run_long(M.longs[i], s);
        thread_workspace() = prev;
    }
    // join (also after a failure: the helper streams must leave a capture)
    for (size_t i = 1; i < st.size(); ++i) {
        check(hipEventRecord(ev[i], st[i]));
        check(hipStreamWaitEvent(s, ev[i], 0));
    }
    for (auto &e : ev)
        (void)hipEventDestroy(e);
    return rc;
}

}  // namespace kfa

using namespace kfa;

extern "C" {

void *kfa_multi_create(int type, const kfa_multi_item *items, int nitems)
{
    if (!items || nitems <= 0) {
        fail(-2, "multi-batch: no items");
        return nullptr;
    }
    int dev;
    if (current_device(&dev))
        return nullptr;
    auto M = std::make_unique<Multi>();
    M->type = type;
    M->dev = dev;
    const size_t cb = cpx_bytes(type);
    const bool all_persistent = env_flag("KISSFFT_AMD_MULTI_PERSISTENT", false);
    // decisions (variant tuning) happen here, on a stream of their own, never during capture
    hipStream_t s0 = nullptr;
    if (hipStreamCreateWithFlags(&s0, hipStreamNonBlocking) != hipSuccess) {
        fail(-3, "multi-batch: stream creation failed");
        return nullptr;
    }
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard()
        {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
    } guard{s0};
    std::map<std::pair<int, int>, std::vector<int>> by_key;  // (kind, nfft) -> compiled items
    std::map<std::pair<int, int>, std::vector<int>> by_blue; // (kind, nfft) -> Bluestein items
    std::vector<const FastChoice *> choice(nitems, nullptr);
    std::vector<std::shared_ptr<DevPlan>> iplan(nitems);
    std::vector<std::shared_ptr<RealTables>> irtab(nitems);
    long long tiles = 0;
    size_t ws_bytes = 0;
    for (int i = 0; i < nitems; ++i) {
        const kfa_multi_item &it = items[i];
        const bool real = it.kind == KFA_SERVE_R2C || it.kind == KFA_SERVE_C2R;
        if (it.kind < 0 || it.kind > 3 || it.howmany < 0 || it.nfft <= 0 || (real && (it.nfft & 1))) {
            fail(-2, "multi-batch item %d: bad kind / length / count", i);
            return nullptr;
        }
        if (it.howmany == 0)
            continue;
        if (!device_accessible(it.in) || !device_accessible(it.out)) {
            fail(-2, "multi-batch item %d: buffers must be device memory", i);
            return nullptr;
        }
        std::shared_ptr<DevPlan> P;
        std::shared_ptr<RealTables> R;
        int rc = build_cplan(real ? &it.rp->half : it.cp, dev, P);
        if (rc == 0 && real)
            rc = build_rtab(it.rp, dev, R);
        if (rc)
            return nullptr;
        M->plans.push_back(P);
        if (R)
            M->rtabs.push_back(R);
        iplan[i] = P;
        irtab[i] = R;
        const long long nc = P->n;
        if (!P->single) {
            const int fk = it.kind == KFA_SERVE_R2C ? FK_R2C : it.kind == KFA_SERVE_C2R ? FK_C2R : FK_C2C;
            if ((rc = prepare_multipass(*P, fk, s0)))
                return nullptr;
            // workspace: the item's chunk of transforms (as the call without a workspace
            // would run it), at least one transform
            const size_t per = multipass_bytes_per_transform(*P, fk) + 1024;
            const size_t want = std::min((size_t)it.howmany * per, std::max(per, scratch_room()));
            ws_bytes = std::max(ws_bytes, want + 4096);
            M->longs.push_back(Multi::Long{it.kind, P, R, it.in, it.out, it.howmany, it.nfft});
            continue;
        }
        if (!all_persistent) {
            const int fk = it.kind == KFA_SERVE_R2C ? FK_R2C : it.kind == KFA_SERVE_C2R ? FK_C2R : FK_C2C;
            const FastChoice *fc = nullptr;
            if ((rc = choose_fast(*P, fk, s0, &fc, it.howmany)))
                return nullptr;
            if (fc->v) {
                choice[i] = fc;
                by_key[{it.kind, it.nfft}].push_back(i);
                continue;
            }
            bool blue = false;
            if (type == KFA_F32 && !real && (rc = blue_prepare(*P, &blue)))
                return nullptr;
            if (blue) {
                by_blue[{it.kind, it.nfft}].push_back(i);
                continue;
            }
        }
        // lines per tile as in the line kernel's row mode
        const size_t per_line = 2 * (size_t)nc * cb;
        int tl = (int)std::max(1LL, std::min(64LL, 2048 / nc));
        while (tl > 1 && per_line * tl > 64 * 1024)
            tl >>= 1;
        if ((long long)tl > it.howmany)
            tl = (int)it.howmany;
        MultiItem mi{};
        mi.pl = P->lp;
        mi.tw = P->d_tw;
        mi.stw = R ? R->d_stw : nullptr;
        mi.in = it.in;
        mi.out = it.out;
        mi.io.nlines = it.howmany;
        mi.io.nlo = 1;
        mi.io.in_hi = it.kind == KFA_SERVE_C2R ? nc + 1 : nc;
        mi.io.in_el = 1;
        mi.io.out_hi = it.kind == KFA_SERVE_R2C ? nc + 1 : nc;
        mi.io.out_el = 1;
        mi.tile0 = tiles;
        mi.TL = tl;
        mi.prologue = it.kind == KFA_SERVE_C2R ? PRO_C2R : PRO_NONE;
        mi.epilogue = it.kind == KFA_SERVE_R2C ? EPI_R2C : EPI_NONE;
        tiles += (it.howmany + tl - 1) / tl;
        M->lds = std::max(M->lds, per_line * tl);
        M->h.push_back(mi);
    }
    M->ntiles = tiles;
    if (ws_bytes) {
        if (dev_alloc(ws_bytes, M->ws_mem))
            return nullptr;
        M->ws.base = static_cast<char *>(M->ws_mem.get());
        M->ws.size = ws_bytes;
    }
    // one segmented launch per (kind, length) and at most KFA_MAX_SEG items
    for (auto &kv : by_key) {
        const std::vector<int> &idx = kv.second;
        for (size_t g0 = 0; g0 < idx.size(); g0 += KFA_MAX_SEG) {
            const size_t g1 = std::min(idx.size(), g0 + KFA_MAX_SEG);
            const kfa_multi_item &f = items[idx[g0]];
            const bool real = f.kind == KFA_SERVE_R2C || f.kind == KFA_SERVE_C2R;
            const long long nc = iplan[idx[g0]]->n;
            Multi::Group G{};
            G.v = choice[idx[g0]]->v;
            G.a.stw = choice[idx[g0]]->d_stw;
            G.a.rtw = real ? irtab[idx[g0]]->d_stw : nullptr;
            G.a.in_hi = f.kind == KFA_SERVE_C2R ? nc + 1 : nc;
            G.a.out_hi = f.kind == KFA_SERVE_R2C ? nc + 1 : nc;
            G.a.in_el = G.a.out_el = 1;
            G.a.nseg = (int)(g1 - g0);
            long long line = 0;
            for (size_t j = g0; j < g1; ++j) {
                const kfa_multi_item &it = items[idx[j]];
                G.a.seg[j - g0].first = line;
                G.a.seg[j - g0].in = it.in;
                G.a.seg[j - g0].out = it.out;
                line += it.howmany;
            }
            G.a.in = items[idx[g0]].in;
            G.a.out = items[idx[g0]].out;
            G.a.nlines = line;
            G.bytes = (size_t)line * cb * (size_t)(G.a.in_hi + G.a.out_hi);
            M->groups.push_back(G);
            M->group_items += (int)(g1 - g0);
        }
    }
    for (auto &kv : by_blue) {
        const std::vector<int> &idx = kv.second;
        for (size_t g0 = 0; g0 < idx.size(); g0 += KFA_MAX_SEG) {
            const size_t g1 = std::min(idx.size(), g0 + KFA_MAX_SEG);
            Multi::BlueGroup G{};
            G.P = iplan[idx[g0]];
            bool applies = false;
            if (blue_prepare(*G.P, &applies, &G.bc) || !applies)
                return nullptr;
            const long long n = G.P->n;
            G.a.in_hi = G.a.out_hi = n;
            G.a.in_el = G.a.out_el = 1;
            G.a.nseg = (int)(g1 - g0);
            long long line = 0;
            for (size_t j = g0; j < g1; ++j) {
                const kfa_multi_item &it = items[idx[j]];
                G.a.seg[j - g0].first = line;
                G.a.seg[j - g0].in = it.in;
                G.a.seg[j - g0].out = it.out;
                line += it.howmany;
            }
            G.a.in = items[idx[g0]].in;
            G.a.out = items[idx[g0]].out;
            G.a.nlines = line;
            M->blue.push_back(G);
            M->blue_items += (int)(g1 - g0);
        }
    }
    std::stable_sort(M->groups.begin(), M->groups.end(),
                     [](const Multi::Group &x, const Multi::Group &y) { return x.bytes > y.bytes; });
    fuse_groups(*M);
    int nhelpers = kMultiHelpers;
    if (const char *e = getenv("KISSFFT_AMD_MULTI_HELPERS"))
        nhelpers = std::max(0, std::min(atoi(e), 8));
    for (int i = 0; i < nhelpers; ++i) {
        hipStream_t hs = nullptr;
        if (hipStreamCreateWithFlags(&hs, hipStreamNonBlocking) != hipSuccess) {
            fail(-3, "multi-batch: stream creation failed");
            for (hipStream_t x : M->helpers)
                (void)hipStreamDestroy(x);
            return nullptr;
        }
        M->helpers.push_back(hs);
    }
    if (!M->h.empty()) {
        const size_t bytes = sizeof(MultiItem) * M->h.size();
        hipError_t e = hipMalloc(&M->d, bytes);
        if (e == hipSuccess)
            e = hipMemcpy(M->d, M->h.data(), bytes, hipMemcpyHostToDevice);
        int cus = 0;
        if (e == hipSuccess)
            e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess) {
            fail(-3, "multi-batch setup: %s", hipGetErrorString(e));
            if (M->d)
                (void)hipFree(M->d);
            for (hipStream_t x : M->helpers)
                (void)hipStreamDestroy(x);
            return nullptr;
        }
        // enough resident workgroups per CU to overlap memory and arithmetic
        const long long per_cu = std::max<long long>(1, std::min<long long>(8, (160 * 1024) / (long long)M->lds));
        M->grid = (int)std::min<long long>(tiles, (long long)cus * per_cu);
    }
    return M.release();
}

int kfa_multi_exec(void *m, void *stream, int sync)
{
    Multi *M = static_cast<Multi *>(m);
    if (!M)
        return fail(-2, "multi-batch: NULL");
    hipStream_t s;
    int dev;
    int rc = call_begin(stream, &s, &dev);
    if (rc)
        return rc;
    if (dev != M->dev)
        return fail(-2, "multi-batch created on device %d, called on device %d", M->dev, dev);
    return call_end(s, sync, run_multi(*M, s));
}

int kfa_multi_capture(void *m, void *stream)
{
    Multi *M = static_cast<Multi *>(m);
    if (!M || !stream)
        return fail(-2, "multi-batch capture needs a stream");
    hipStream_t s = (hipStream_t)stream;
    if (M->gexec) {
        (void)hipGraphExecDestroy(M->gexec);
        M->gexec = nullptr;
    }
    if (M->graph) {
        (void)hipGraphDestroy(M->graph);
        M->graph = nullptr;
    }
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess)
        return fail(-3, "hipStreamBeginCapture: %s", hipGetErrorString(e));
    const int rc = run_multi(*M, s);
    e = hipStreamEndCapture(s, &M->graph);
    if (rc)
        return rc;
    if (e != hipSuccess)
        return fail(-3, "hipStreamEndCapture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&M->gexec, M->graph, nullptr, nullptr, 0);
    return e == hipSuccess ? 0 : fail(-3, "hipGraphInstantiate: %s", hipGetErrorString(e));
}

int kfa_multi_replay(void *m, void *stream, int sync)
{
    Multi *M = static_cast<Multi *>(m);
    if (!M || !M->gexec)
        return fail(-2, "multi-batch: nothing captured");
    hipStream_t s;
    int dev;
    int rc = call_begin(stream, &s, &dev);
    if (rc)
        return rc;
    const hipError_t e = hipGraphLaunch(M->gexec, s);
    return call_end(s, sync, e == hipSuccess ? 0 : fail(-3, "hipGraphLaunch: %s", hipGetErrorString(e)));
}

int kfa_multi_info(void *m, long long *tiles, int *grid, int *persistent_items)
{
    Multi *M = static_cast<Multi *>(m);
    if (!M)
        return -2;
    if (tiles)
        *tiles = M->ntiles;
    if (grid)
        *grid = M->grid;
    if (persistent_items)
        *persistent_items = (int)M->h.size();
    return 0;
}

int kfa_multi_layout(void *m, int *groups, int *group_items, int *bluestein_items, int *multipass_items)
{
    Multi *M = static_cast<Multi *>(m);
    if (!M)
        return -2;
    if (groups)
        *groups = (int)M->groups.size();
    if (group_items)
        *group_items = M->group_items;
    if (bluestein_items)
        *bluestein_items = M->blue_items;
    if (multipass_items)
        *multipass_items = (int)M->longs.size();
    return 0;
}

int kfa_multi_fused(void *m, int *members, long long *workgroups, int *threads)
{
    Multi *M = static_cast<Multi *>(m);
    if (!M)
        return -2;
    const bool f = M->fused.fn != nullptr;
    if (members)
        *members = f ? M->fused.members : 0;
    if (workgroups)
        *workgroups = f ? M->fused.blocks : 0;
    if (threads)
        *threads = f ? M->fused.NT : 0;
    return 0;
}

void kfa_multi_destroy(void *m)
{
    Multi *M = static_cast<Multi *>(m);
    if (!M)
        return;
    if (M->gexec)
        (void)hipGraphExecDestroy(M->gexec);
    if (M->graph)
        (void)hipGraphDestroy(M->graph);
    if (M->d)
        (void)hipFree(M->d);
    for (hipStream_t x : M->helpers)
        (void)hipStreamDestroy(x);
    delete M;
}

}  // extern "C"
