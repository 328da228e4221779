// kfa_conv.hip -- fast FIR filtering by overlap-scrap (k_conv in kfa_blue.hpp).
//
// The filter spectrum (N complex, scaled by 1/N) comes from the host layer; the engine
// keeps a device copy per (device, N, spectrum contents) together with the stage tables
// of the N-point schedule, and runs every complete block of a call in one launch.
#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../kfa_engine.h"
#include "kfa_blue.hpp"
#include "kfa_internal.hpp"

namespace kfa {

struct ConvPlan {
    const ConvVariant *v = nullptr;
    DevBuf mem_f, mem_i, mem_h;
};

static std::mutex g_conv_mu;
static std::map<std::tuple<int, int, int, std::string>, std::shared_ptr<const ConvPlan>> g_conv;

static int conv_plan(int dev, int nfft, int real, const void *H, std::shared_ptr<const ConvPlan> &out)
{
    const std::string key_h((const char *)H, (size_t)nfft * 8);
    std::lock_guard<std::mutex> g(g_conv_mu);
    const auto key = std::make_tuple(dev, nfft, real, key_h);
    auto it = g_conv.find(key);
    if (it != g_conv.end()) {
        out = it->second;
        return 0;
    }
    auto cp = std::make_shared<ConvPlan>();
    for (const ConvVariant &v : conv_registry())
        if (v.N == nfft && v.real == real) {  // the first registered of a length
            cp->v = &v;
            break;
        }
    if (!cp->v)
        return fail(-2, "fast convolution: no kernel for block length %d (powers of two 256..8192)", nfft);
    const std::vector<unsigned char> tf = float_twiddles(nfft, 0, nfft, 1), ti = float_twiddles(nfft, 1, nfft, 1);
    int rc;
    if ((rc = make_stage_table_raw(tf.data(), KFA_F32, nfft, cp->v->radices, cp->v->tw_pad, 0, cp->mem_f)) ||
        (rc = make_stage_table_raw(ti.data(), KFA_F32, nfft, cp->v->radices, cp->v->tw_pad, 0, cp->mem_i)) ||
        (rc = dev_upload(H, (size_t)nfft * 8, cp->mem_h)))
        return rc;
    g_conv[key] = cp;
    out = cp;
    return 0;
}

void conv_release()
{
    std::lock_guard<std::mutex> g(g_conv_mu);
    g_conv.clear();
}

}  // namespace kfa

using namespace kfa;

extern "C" int kfa_fastfir(int type, int nfft, int real, const void *H, const void *in, void *out, long long nblocks,
                           long long ngood, void *stream, int sync)
{
    if (type != KFA_F32)
        return fail(-2, "fast convolution is float only (the reference scales by 1/nfft in float)");
    if (nfft <= 0 || ngood <= 0 || ngood > nfft || nblocks < 0)
        return fail(-2, "fast convolution: bad sizes");
    hipStream_t s;
    int dev;
    int rc = call_begin(stream, &s, &dev);
    if (rc || nblocks == 0)
        return rc;
    std::shared_ptr<const ConvPlan> cp;
    if ((rc = conv_plan(dev, nfft, real, H, cp)))
        return rc;
    const size_t sb = real ? 4 : 8;
    Operand I, O;
    rc = stage_in(in, sb * (size_t)((nblocks - 1) * ngood + nfft), s, I);
    if (rc == 0)
        rc = stage_out(out, sb * (size_t)(nblocks * ngood), s, O);
    if (rc == 0) {
        ConvArgs a{};
        a.in = I.dev;
        a.out = O.dev;
        a.stw_f = cp->mem_f.get();
        a.stw_i = cp->mem_i.get();
        a.H = cp->mem_h.get();
        a.nblocks = nblocks;
        a.nlines = real ? (nblocks + 1) / 2 : nblocks;
        a.ngood = ngood;
        ++g_launches;
        rc = cp->v->launch(a, s) ? fail(-3, "fast convolution launch failed") : 0;
        if (rc == 0)
            rc = finish_out(O, s);
    }
    release(I, s);
    release(O, s);
    return call_end(s, sync || I.staged || O.staged, rc);
}
