// kfa_multi.hip -- several batches of different lengths and kinds in one persistent launch.
//
// The batches (items) are fixed when the multi-batch is created: kind, length, device
// buffers and count.  Creation builds (or finds) each length's device plan and uploads one
// table describing every item: its stage list, twiddles, buffers and its range in a global
// list of tiles.  Execution is a single launch of k_multi (kfa_lines.hpp): a grid sized to
// the device walks the tile list, each workgroup transforming whole tiles with the
// general LDS-resident line code, so any mix of radices -- including generic primes -- and
// of complex / real kinds runs together, bit-exact in the fixed-point builds.  Items longer
// than one LDS line run as their own multi-pass launches after it.  The launch uses no
// counters or memsets and allocates nothing, so it can be captured into a hipGraph once and
// replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <mutex>
#include <vector>

#include "../kfa_engine.h"
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
    struct Long {                        // multi-pass items, launched on their own
        int kind;
        std::shared_ptr<DevPlan> P;
        std::shared_ptr<RealTables> R;
        const void *in;
        void *out;
        long long howmany;
        int n;
    };
    std::vector<Long> longs;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
};

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

static int run_multi(Multi &M, hipStream_t s)
{
    int rc = 0;
    if (M.ntiles > 0) {
        switch (M.type) {
        case KFA_F32: rc = launch_multi<F32>(M, s); break;
        case KFA_Q15: rc = launch_multi<Q15>(M, s); break;
        default: rc = launch_multi<Q31>(M, s); break;
        }
    }
    for (size_t i = 0; i < M.longs.size() && rc == 0; ++i) {
        const Multi::Long &L = M.longs[i];
        const long long nc = L.P->n;
        if (L.kind == KFA_SERVE_R2C)
            rc = run_real_rows(*L.P, *L.R, 0, L.in, nc, L.out, nc + 1, L.howmany, s);
        else if (L.kind == KFA_SERVE_C2R)
            rc = run_real_rows(*L.P, *L.R, 1, L.in, nc + 1, L.out, nc, L.howmany, s);
        else
            rc = run_c2c(*L.P, L.in, 1, nc, L.out, 1, nc, L.howmany, s);
    }
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
    long long tiles = 0;
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
        const long long nc = P->n;
        if (!P->single) {
            M->longs.push_back(Multi::Long{it.kind, P, R, it.in, it.out, it.howmany, it.nfft});
            continue;
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
    if (!M->longs.empty())
        return fail(-2, "multi-batch capture: items longer than one LDS line allocate scratch; not capturable");
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
    delete M;
}

}  // extern "C"
