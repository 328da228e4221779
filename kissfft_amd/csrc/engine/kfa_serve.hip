// kfa_serve.hip -- the request server: continuous batching of mixed FFT requests.
//
// Requests (complex forward/inverse, real forward/inverse; any length; any batch count)
// arrive through kfa_server_submit from any thread.  A scheduler thread
//   * admits them in arrival order while the HBM they need fits the server's budget (the
//     server's own memory: staging copies of host buffers and the scratch of multi-pass
//     lengths; caller device buffers are used in place and cost nothing),
//   * groups admitted requests of the same (kind, length) into one launch of the size's
//     compiled kernel (up to KFA_MAX_SEG requests, each its own buffers: the kernel's
//     segment table maps lines to requests),
//   * spreads launches over several streams so work from different groups overlaps on
//     the device, and keeps admitting while earlier launches run.
// Completion is signalled by a host function enqueued behind each launch; it records the
// request's latency (submit to completion), returns its memory to the budget and wakes
// waiters and the scheduler.  Plans and twiddles come from the engine's device cache.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../kfa_engine.h"
#include "kfa_fast.hpp"
#include "kfa_internal.hpp"

namespace kfa {

static double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct ServeReq {
    long long id = 0;
    int kind = 0;  // KFA_SERVE_*
    int n = 0;
    std::shared_ptr<DevPlan> P;     // complex plan (half length for real kinds)
    std::shared_ptr<RealTables> R;  // real kinds only
    bool nd = false;                // N-D request: dims, one plan per (leading) axis in axes
    std::vector<int> dims;
    std::vector<std::shared_ptr<DevPlan>> axes;
    const void *in = nullptr;
    void *out = nullptr;
    long long howmany = 0;
    size_t in_bytes = 0, out_bytes = 0;
    bool stage_in = false, stage_out = false;
    void *d_in = nullptr, *d_out = nullptr;
    size_t need = 0;  // server memory held while in flight
    double t_submit = 0, t_done = 0;
    int state = 0;  // 0 pending, 1 launched, 2 done, <0 failed
    bool launch_failed = false;
    std::string err;
};

struct Server;

struct LaunchRec {
    Server *srv;
    std::vector<std::shared_ptr<ServeReq>> reqs;
    double t_enq0 = 0, t_enq1 = 0;  // host time around enqueueing the group
};

struct Server {
    int dev = 0;
    int type = KFA_F32;
    size_t budget = 0, in_use = 0;
    int max_group = KFA_MAX_SEG;
    std::vector<hipStream_t> streams;
    hipStream_t plan_stream = nullptr;
    size_t next_stream = 0;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<ServeReq>> pending;
    std::unordered_map<long long, std::shared_ptr<ServeReq>> live;
    long long next_id = 1;
    bool stopping = false;
    bool trace = false;  // KISSFFT_AMD_VERBOSE=2: one line per completed launch group
    int inflight_launches = 0;
    std::thread th;

    // stats
    long long submitted = 0, completed = 0, failed = 0, launches = 0, grouped = 0;
    long long bytes_done = 0;
    std::vector<double> lat;
    double t_first = 0, t_last = 0;

    void run();
    int launch_group(std::vector<std::shared_ptr<ServeReq>> &grp, hipStream_t s);
};

static void on_launch_done(void *p)
{
    LaunchRec *rec = static_cast<LaunchRec *>(p);
    Server *S = rec->srv;
    const double t = now_s();
    {
        std::lock_guard<std::mutex> g(S->mu);
        for (auto &r : rec->reqs) {
            r->t_done = t;
            if (!r->launch_failed) {
                r->state = 2;
                S->completed++;
                S->lat.push_back(t - r->t_submit);
                S->bytes_done += (long long)(r->in_bytes + r->out_bytes);
            } else {
                r->state = -1;
                S->failed++;
            }
            S->in_use -= r->need;
        }
        S->inflight_launches--;
        S->t_last = t;
    }
    S->cv_done.notify_all();
    S->cv_work.notify_all();
    if (S->trace) {
        const ServeReq &r0 = *rec->reqs[0];
        long long lines = 0;
        for (auto &r : rec->reqs)
            lines += r->howmany;
        fprintf(stderr, "[kissfft_amd serve] kind=%d n=%d reqs=%zu lines=%lld enqueue=%.3f ms done=+%.3f ms\n",
                r0.kind, r0.n, rec->reqs.size(), lines, 1e3 * (rec->t_enq1 - rec->t_enq0), 1e3 * (t - rec->t_enq1));
    }
    delete rec;
}

// Bytes of one side of a request: complex sides count n (or n/2+1) complex values.
static void side_bytes(int type, int kind, int n, long long howmany, size_t &in_b, size_t &out_b)
{
    const size_t cb = cpx_bytes(type), sb = cb / 2;
    const size_t cpx_full = cb * (size_t)n, real = sb * (size_t)n, half = cb * (size_t)(n / 2 + 1);
    switch (kind) {
    case KFA_SERVE_C2C_FWD:
    case KFA_SERVE_C2C_INV: in_b = out_b = cpx_full; break;
    case KFA_SERVE_R2C: in_b = real; out_b = half; break;
    default: in_b = half; out_b = real; break;
    }
    in_b *= (size_t)howmany;
    out_b *= (size_t)howmany;
}

// Enqueues one group of same-key requests on stream s (device buffers in d_in/d_out).
int Server::launch_group(std::vector<std::shared_ptr<ServeReq>> &grp, hipStream_t s)
{
    const ServeReq &r0 = *grp[0];
    DevPlan &P = *r0.P;
    const bool real = r0.kind == KFA_SERVE_R2C || r0.kind == KFA_SERVE_C2R;
    const int fk = r0.kind == KFA_SERVE_R2C ? FK_R2C : r0.kind == KFA_SERVE_C2R ? FK_C2R : FK_C2C;
    const FastChoice *fc = nullptr;
    int rc = (P.single && !r0.nd) ? choose_fast(P, fk, s, &fc) : 0;
    if (rc)
        return rc;
    if (fc && fc->v) {
        // one launch for the whole group, one segment per request
        FastArgs a{};
        a.stw = fc->d_stw;
        a.rtw = real ? r0.R->d_stw : nullptr;
        const long long nc = P.n;  // complex line length
        a.in_hi = r0.kind == KFA_SERVE_R2C ? nc : (r0.kind == KFA_SERVE_C2R ? nc + 1 : nc);
        a.out_hi = r0.kind == KFA_SERVE_R2C ? nc + 1 : nc;
        a.in_el = a.out_el = 1;
        long long line = 0;
        a.nseg = (int)grp.size();
        for (size_t i = 0; i < grp.size(); ++i) {
            a.seg[i].first = line;
            a.seg[i].in = grp[i]->d_in;
            a.seg[i].out = grp[i]->d_out;
            line += grp[i]->howmany;
        }
        a.in = grp[0]->d_in;
        a.out = grp[0]->d_out;
        a.nlines = line;
        ++g_launches;
        return fast_launch(*fc->v, a, s) ? fail(-3, "fast kernel launch failed") : 0;
    }
    // no compiled kernel for this key (or a multi-pass length, or N-D): one launch per request
    for (auto &r : grp) {
        if (r->nd) {
            if (r->kind == KFA_SERVE_R2C || r->kind == KFA_SERVE_C2R)
                rc = run_ndr(*r->P, *r->R, r->axes, r->dims.data(), r->kind == KFA_SERVE_C2R, r->d_in, r->d_out,
                             r->howmany, s);
            else
                rc = run_nd(r->axes, r->dims.data(), r->d_in, r->d_out, r->howmany, s);
            if (rc)
                return rc;
            continue;
        }
        switch (r->kind) {
        case KFA_SERVE_C2C_FWD:
        case KFA_SERVE_C2C_INV:
            rc = run_c2c(*r->P, r->d_in, 1, r->n, r->d_out, 1, r->n, r->howmany, s);
            break;
        case KFA_SERVE_R2C:
            rc = run_real_rows(*r->P, *r->R, 0, r->d_in, r->n / 2, r->d_out, r->n / 2 + 1, r->howmany, s);
            break;
        default:
            rc = run_real_rows(*r->P, *r->R, 1, r->d_in, r->n / 2 + 1, r->d_out, r->n / 2, r->howmany, s);
            break;
        }
        if (rc)
            return rc;
    }
    return 0;
}

void Server::run()
{
    (void)hipSetDevice(dev);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        // admit in arrival order while memory allows (head-of-line: large requests are not
        // starved by a stream of small ones)
        std::vector<std::shared_ptr<ServeReq>> ready;
        bool rejected = false;
        while (!pending.empty()) {
            auto &r = pending.front();
            if (r->need > budget) {
                r->state = -1;  // can never fit
                r->err = "needs " + std::to_string(r->need) + " bytes of server memory, budget is " +
                         std::to_string(budget);
                failed++;
                rejected = true;
                pending.pop_front();
                continue;
            }
            if (in_use + r->need > budget)
                break;
            in_use += r->need;
            ready.push_back(r);
            pending.pop_front();
        }
        if (rejected)
            cv_done.notify_all();
        if (ready.empty()) {
            if (stopping && pending.empty() && inflight_launches == 0)
                break;
            cv_work.wait_for(lk, std::chrono::milliseconds(2));
            continue;
        }
        lk.unlock();
        // group by (kind, length), keeping arrival order inside a group
        std::map<std::pair<int, int>, std::vector<std::shared_ptr<ServeReq>>> groups;
        for (auto &r : ready)  // N-D requests run alone (a key of their own)
            groups[r->nd ? std::make_pair(16 + r->kind, (int)(r->id & 0x7fffffff)) : std::make_pair(r->kind, r->n)]
                .push_back(r);
        for (auto &kv : groups) {
            auto &all = kv.second;
            for (size_t off = 0; off < all.size(); off += (size_t)max_group) {
                std::vector<std::shared_ptr<ServeReq>> grp(all.begin() + off,
                                                           all.begin() + std::min(all.size(), off + (size_t)max_group));
                hipStream_t s = streams[next_stream++ % streams.size()];
                const double t_enq0 = now_s();
                int rc = 0;
                for (auto &r : grp) {
                    r->d_in = const_cast<void *>(r->in);
                    r->d_out = r->out;
                    if (r->stage_in && rc == 0) {
                        rc = scratch_alloc(&r->d_in, r->in_bytes ? r->in_bytes : 1, s) == hipSuccess ? 0 : -3;
                        if (rc == 0 && r->in_bytes)
                            rc = hipMemcpyAsync(r->d_in, r->in, r->in_bytes, hipMemcpyHostToDevice, s) == hipSuccess
                                     ? 0
                                     : -3;
                    }
                    if (r->stage_out && rc == 0)
                        rc = scratch_alloc(&r->d_out, r->out_bytes ? r->out_bytes : 1, s) == hipSuccess ? 0 : -3;
                }
                if (rc == 0)
                    rc = launch_group(grp, s);
                for (auto &r : grp) {
                    if (r->stage_out && r->d_out && rc == 0 && r->out_bytes)
                        (void)hipMemcpyAsync(r->out, r->d_out, r->out_bytes, hipMemcpyDeviceToHost, s);
                    if (r->stage_in && r->d_in)
                        (void)hipFreeAsync(r->d_in, s);
                    if (r->stage_out && r->d_out)
                        (void)hipFreeAsync(r->d_out, s);
                }
                LaunchRec *rec = new LaunchRec{this, grp, t_enq0, now_s()};
                {
                    std::lock_guard<std::mutex> g(mu);
                    // a failed group still completes behind the stream (its staging copies may
                    // be in flight), and is reported failed then
                    for (auto &r : grp) {
                        r->state = 1;
                        r->launch_failed = rc != 0;
                        if (rc)
                            r->err = kfa_last_error();
                    }
                    inflight_launches++;
                    launches++;
                    grouped += (long long)grp.size();
                }
                if (hipLaunchHostFunc(s, on_launch_done, rec) != hipSuccess) {
                    (void)hipStreamSynchronize(s);
                    on_launch_done(rec);
                }
            }
        }
        lk.lock();
    }
}

}  // namespace kfa

using namespace kfa;

extern "C" {

void *kfa_server_create(int type, long long budget_bytes, int nstreams, int max_group)
{
    int dev;
    if (current_device(&dev))
        return nullptr;
    Server *S = new Server();
    S->dev = dev;
    S->type = type;
    if (budget_bytes <= 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
            delete S;
            fail(-3, "hipMemGetInfo failed");
            return nullptr;
        }
        budget_bytes = (long long)(fr / 10 * 8);  // 80% of the HBM free now
    }
    S->budget = (size_t)budget_bytes;
    {
        const char *v = getenv("KISSFFT_AMD_VERBOSE");
        S->trace = v && atoi(v) >= 2;
    }
    S->max_group = (max_group <= 0 || max_group > KFA_MAX_SEG) ? KFA_MAX_SEG : max_group;
    if (nstreams <= 0)
        nstreams = 4;
    for (int i = 0; i < nstreams; ++i) {
        hipStream_t s;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
            fail(-3, "stream creation failed");
            break;
        }
        S->streams.push_back(s);
    }
    if (S->streams.empty() || hipStreamCreateWithFlags(&S->plan_stream, hipStreamNonBlocking) != hipSuccess) {
        delete S;
        return nullptr;
    }
    S->th = std::thread([S] { S->run(); });
    return S;
}

long long kfa_server_submit(void *srv, int kind, const kfa_cplan *cp, const kfa_rplan *rp, int nfft, const void *in,
                            void *out, long long howmany)
{
    Server *S = static_cast<Server *>(srv);
    if (!S || !in || !out || howmany <= 0 || nfft <= 0)
        return fail(-2, "server submit: bad arguments");
    auto r = std::make_shared<ServeReq>();
    r->kind = kind;
    r->n = nfft;
    r->howmany = howmany;
    r->in = in;
    r->out = out;
    int rc;
    if (kind == KFA_SERVE_R2C || kind == KFA_SERVE_C2R) {
        if (!rp || (nfft & 1))
            return fail(-2, "server submit: real requests need an even length");
        if ((rc = build_cplan(&rp->half, S->dev, r->P)) || (rc = build_rtab(rp, S->dev, r->R)))
            return rc;
    } else {
        if (!cp)
            return fail(-2, "server submit: no plan");
        if ((rc = build_cplan(cp, S->dev, r->P)))
            return rc;
    }
    // decide the compiled variant now, on the server's plan stream, not in the scheduler
    if (r->P->single) {
        const FastChoice *fc;
        const int fk = kind == KFA_SERVE_R2C ? FK_R2C : kind == KFA_SERVE_C2R ? FK_C2R : FK_C2C;
        if ((rc = choose_fast(*r->P, fk, S->plan_stream, &fc, howmany)))
            return rc;
    }
    side_bytes(S->type, kind, nfft, howmany, r->in_bytes, r->out_bytes);
    r->stage_in = !device_accessible(in);
    r->stage_out = !device_accessible(out);
    r->need = (r->stage_in ? r->in_bytes : 0) + (r->stage_out ? r->out_bytes : 0);
    // multi-pass scratch: the temporaries of one chunk of transforms (the chunking of
    // multi-pass batches bounds them)
    if (!r->P->single) {
        const int fk = kind == KFA_SERVE_R2C ? FK_R2C : kind == KFA_SERVE_C2R ? FK_C2R : FK_C2C;
        const size_t per = multipass_bytes_per_transform(*r->P, fk) + 1024;
        r->need += std::min((size_t)howmany * per, std::max(per, scratch_room()));
    }
    r->t_submit = now_s();
    {
        std::lock_guard<std::mutex> g(S->mu);
        r->id = S->next_id++;
        S->live[r->id] = r;
        S->pending.push_back(r);
        S->submitted++;
        if (S->t_first == 0)
            S->t_first = r->t_submit;
    }
    S->cv_work.notify_all();
    return r->id;
}

long long kfa_server_submit_nd(void *srv, int kind, int ndims, const int *dims, const kfa_cplan *axes,
                               const kfa_rplan *rp, const void *in, void *out, long long howmany)
{
    Server *S = static_cast<Server *>(srv);
    if (!S || !in || !out || howmany <= 0 || ndims <= 0 || !dims || kind < 0 || kind > 3)
        return fail(-2, "server submit: bad arguments");
    const bool real = kind == KFA_SERVE_R2C || kind == KFA_SERVE_C2R;
    auto r = std::make_shared<ServeReq>();
    r->nd = true;
    r->kind = kind;
    r->howmany = howmany;
    r->in = in;
    r->out = out;
    r->dims.assign(dims, dims + ndims);
    long long total = 1;
    for (int k = 0; k < ndims; ++k) {
        if (dims[k] <= 0)
            return fail(-2, "server submit: bad dimension");
        total *= dims[k];
    }
    r->n = (int)std::min<long long>(total, 0x7fffffff);
    int rc;
    bool multipass = false;
    const int nplans = real ? ndims - 1 : ndims;
    r->axes.resize(nplans);
    for (int k = 0; k < nplans; ++k) {
        if ((rc = build_cplan(&axes[k], S->dev, r->axes[k])))
            return rc;
        multipass |= !r->axes[k]->single;
    }
    if (real) {
        if (!rp || (dims[ndims - 1] & 1))
            return fail(-2, "server submit: real N-D requests need an even last dimension");
        if ((rc = build_cplan(&rp->half, S->dev, r->P)) || (rc = build_rtab(rp, S->dev, r->R)))
            return rc;
        multipass |= !r->P->single;
    } else {
        r->P = r->axes[0];
    }
    const size_t cb = cpx_bytes(S->type), sb = cb / 2;
    const long long last = dims[ndims - 1];
    const size_t cpx_all = cb * (size_t)(total / last * (last / 2 + 1)) * (size_t)howmany;
    const size_t real_all = sb * (size_t)total * (size_t)howmany;
    if (!real)
        r->in_bytes = r->out_bytes = cb * (size_t)total * (size_t)howmany;
    else if (kind == KFA_SERVE_R2C) {
        r->in_bytes = real_all;
        r->out_bytes = cpx_all;
    } else {
        r->in_bytes = cpx_all;
        r->out_bytes = real_all;
    }
    r->stage_in = !device_accessible(in);
    r->stage_out = !device_accessible(out);
    r->need = (r->stage_in ? r->in_bytes : 0) + (r->stage_out ? r->out_bytes : 0);
    if (kind == KFA_SERVE_C2R)  // the inverse runs its complex axes into a temporary
        r->need += cpx_all;
    if (multipass)
        r->need += 2 * scratch_room();
    r->t_submit = now_s();
    {
        std::lock_guard<std::mutex> g(S->mu);
        r->id = S->next_id++;
        S->live[r->id] = r;
        S->pending.push_back(r);
        S->submitted++;
        if (S->t_first == 0)
            S->t_first = r->t_submit;
    }
    S->cv_work.notify_all();
    return r->id;
}

int kfa_server_wait(void *srv, long long ticket, double *latency_s)
{
    Server *S = static_cast<Server *>(srv);
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->live.find(ticket);
    if (it == S->live.end())
        return fail(-2, "unknown ticket %lld", ticket);
    auto r = it->second;
    S->cv_done.wait(lk, [&] { return r->state == 2 || r->state < 0; });
    S->live.erase(ticket);
    if (latency_s)
        *latency_s = r->t_done - r->t_submit;
    return r->state == 2 ? 0 : fail(-3, "request %lld failed: %s", ticket, r->err.c_str());
}

int kfa_server_drain(void *srv)
{
    Server *S = static_cast<Server *>(srv);
    std::unique_lock<std::mutex> lk(S->mu);
    S->cv_done.wait(lk, [&] { return S->pending.empty() && S->inflight_launches == 0; });
    return 0;
}

int kfa_server_get_stats(void *srv, kfa_server_stats *st)
{
    Server *S = static_cast<Server *>(srv);
    std::lock_guard<std::mutex> g(S->mu);
    st->submitted = S->submitted;
    st->completed = S->completed;
    st->failed = S->failed;
    st->launches = S->launches;
    st->grouped_requests = S->grouped;
    st->bytes = S->bytes_done;
    st->budget_bytes = (long long)S->budget;
    st->in_use_bytes = (long long)S->in_use;
    std::vector<double> v(S->lat);
    std::sort(v.begin(), v.end());
    auto q = [&](double f) { return v.empty() ? 0.0 : v[std::min(v.size() - 1, (size_t)(f * (v.size() - 1) + 0.5))]; };
    st->p50_s = q(0.50);
    st->p99_s = q(0.99);
    st->max_s = v.empty() ? 0.0 : v.back();
    double sum = 0;
    for (double x : v)
        sum += x;
    st->mean_s = v.empty() ? 0.0 : sum / (double)v.size();
    st->span_s = S->t_last > S->t_first ? S->t_last - S->t_first : 0.0;
    return 0;
}

void kfa_server_reset_stats(void *srv)
{
    Server *S = static_cast<Server *>(srv);
    std::lock_guard<std::mutex> g(S->mu);
    S->lat.clear();
    S->completed = S->failed = S->launches = S->grouped = S->bytes_done = 0;
    S->submitted = 0;
    S->t_first = S->t_last = 0;
}

void kfa_server_destroy(void *srv)
{
    Server *S = static_cast<Server *>(srv);
    if (!S)
        return;
    {
        std::lock_guard<std::mutex> g(S->mu);
        S->stopping = true;
    }
    S->cv_work.notify_all();
    if (S->th.joinable())
        S->th.join();
    for (hipStream_t s : S->streams) {
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
    }
    if (S->plan_stream) {
        (void)hipStreamSynchronize(S->plan_stream);
        (void)hipStreamDestroy(S->plan_stream);
    }
    delete S;
}

}  // extern "C"
