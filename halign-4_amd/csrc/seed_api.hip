// seed_api.hip -- host side of minimizer extraction: buffer layout, the choice
// between the wave form and the sequential form (seed.hip), and the C ABI
// halign_extract_minimizers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "host_common.h"

namespace hal {

hipError_t seed_pack(const char *seqs, const int64_t *off, const int64_t *woff, const int64_t *moff, int n,
                     int64_t max_words, uint64_t *codes, uint64_t *nmask, hipStream_t stream);
hipError_t seed_wave(const int64_t *off, const int64_t *woff, const int64_t *moff, const uint64_t *codes,
                     const uint64_t *nmask, const int64_t *hit_off, int n, int k, int w, uint64_t *hx, uint64_t *hy,
                     int64_t *counts, int32_t *status, hipStream_t stream);
hipError_t seed_seq(const char *seqs, const int64_t *off, const int32_t *list, int nlist, int k, int w,
                    int non_canonical, const int64_t *hit_off, int ws_cap, uint64_t *hx, uint64_t *hy,
                    int64_t *counts, int32_t *status, hipStream_t stream);
int seed_max_wave_window();

// Minimizers of n sequences already in ctx->s_seqs (bytes) with offsets off
// (host copy, rebased to 0, n+1 entries, also uploaded to ctx->s_off).
// Hits of sequence s go to ctx->s_hx/s_hy from hit_off[s] (capacity
// max(len-k+1, 0)); counts and status come back to the host.
int seed_device(halign_ctx *ctx, int n, const std::vector<int64_t> &off, int k, int w, int non_canonical,
                std::vector<int64_t> &hit_off, std::vector<int64_t> &counts, std::vector<int32_t> &status)
{
    hit_off.assign(n + 1, 0);
    std::vector<int64_t> woff(n + 1, 0), moff(n + 1, 0);
    int64_t max_words = 1;
    for (int s = 0; s < n; ++s) {
        const int64_t len = off[s + 1] - off[s];
        const int64_t groups = 2 * (len / 64 + 2);   // 32-base groups, an even number with padding
        woff[s + 1] = woff[s] + groups;
        moff[s + 1] = moff[s] + groups / 2;
        max_words = std::max(max_words, groups);
        hit_off[s + 1] = hit_off[s] + std::max<int64_t>(len - k + 1, 0);
    }
    const size_t nb8 = (size_t)(n + 1) * 8;
    HCHECK(ctx->s_woff.ensure(nb8), "alloc seed offsets");
    HCHECK(ctx->s_moff.ensure(nb8), "alloc seed offsets");
    HCHECK(ctx->s_hoff.ensure(nb8), "alloc seed offsets");
    HCHECK(ctx->s_codes.ensure((size_t)std::max<int64_t>(woff[n], 1) * 8), "alloc packed codes");
    HCHECK(ctx->s_nmask.ensure((size_t)std::max<int64_t>(moff[n], 1) * 8), "alloc N masks");
    HCHECK(ctx->s_hx.ensure((size_t)std::max<int64_t>(hit_off[n], 1) * 8), "alloc hits");
    HCHECK(ctx->s_hy.ensure((size_t)std::max<int64_t>(hit_off[n], 1) * 8), "alloc hits");
    HCHECK(ctx->s_cnt.ensure((size_t)std::max(n, 1) * 8), "alloc hit counts");
    HCHECK(ctx->s_st.ensure((size_t)std::max(n, 1) * 4), "alloc seed status");
    HCHECK(hipMemcpyAsync(ctx->s_woff.p, woff.data(), nb8, hipMemcpyHostToDevice, ctx->stream), "upload");
    HCHECK(hipMemcpyAsync(ctx->s_moff.p, moff.data(), nb8, hipMemcpyHostToDevice, ctx->stream), "upload");
    HCHECK(hipMemcpyAsync(ctx->s_hoff.p, hit_off.data(), nb8, hipMemcpyHostToDevice, ctx->stream), "upload");

    std::vector<int32_t> list;
    status.assign(n, 0);
    const bool wave = ctx->seed_policy == 0 && non_canonical && w > 0 && w <= seed_max_wave_window();
    if (wave) {
        HCHECK(seed_pack(ctx->s_seqs.as<char>(), ctx->s_off.as<int64_t>(), ctx->s_woff.as<int64_t>(),
                         ctx->s_moff.as<int64_t>(), n, max_words, ctx->s_codes.as<uint64_t>(),
                         ctx->s_nmask.as<uint64_t>(), ctx->stream),
               "launch pack_seqs");
        HCHECK(seed_wave(ctx->s_off.as<int64_t>(), ctx->s_woff.as<int64_t>(), ctx->s_moff.as<int64_t>(),
                         ctx->s_codes.as<uint64_t>(), ctx->s_nmask.as<uint64_t>(), ctx->s_hoff.as<int64_t>(), n, k, w,
                         ctx->s_hx.as<uint64_t>(), ctx->s_hy.as<uint64_t>(), ctx->s_cnt.as<int64_t>(),
                         ctx->s_st.as<int32_t>(), ctx->stream),
               "launch minimizers_wave");
        HCHECK(hipMemcpyAsync(status.data(), ctx->s_st.p, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream),
               "download seed status");
        HCHECK(hipStreamSynchronize(ctx->stream), "minimizers_wave");
        for (int s = 0; s < n; ++s)
            if (status[s] & 1) list.push_back(s);
    } else {
        for (int s = 0; s < n; ++s) list.push_back(s);
    }
    if (!list.empty()) {
        const int win = std::max(1, std::min(w, 255));
        const int ws_cap = 4 * win + 8;
        const int nl = (int)list.size();
        HCHECK(ctx->s_list.ensure((size_t)nl * 4), "alloc seed list");
        HCHECK(hipMemcpyAsync(ctx->s_list.p, list.data(), (size_t)nl * 4, hipMemcpyHostToDevice, ctx->stream),
               "upload seed list");
        HCHECK(seed_seq(ctx->s_seqs.as<char>(), ctx->s_off.as<int64_t>(), ctx->s_list.as<int32_t>(), nl, k, w,
                        non_canonical, ctx->s_hoff.as<int64_t>(), ws_cap, ctx->s_hx.as<uint64_t>(),
                        ctx->s_hy.as<uint64_t>(), ctx->s_cnt.as<int64_t>(), ctx->s_st.as<int32_t>(), ctx->stream),
               "launch minimizers_seq");
    }
    counts.assign(n, 0);
    std::vector<int32_t> st2(n, 0);
    HCHECK(hipMemcpyAsync(counts.data(), ctx->s_cnt.p, (size_t)n * 8, hipMemcpyDeviceToHost, ctx->stream),
           "download hit counts");
    HCHECK(hipMemcpyAsync(st2.data(), ctx->s_st.p, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream),
           "download seed status");
    HCHECK(hipStreamSynchronize(ctx->stream), "minimizers");
    for (int s = 0; s < n; ++s) status[s] = st2[s] | (wave && (status[s] & 1) ? 1 : 0) | (wave ? 0 : 1);
    return HALIGN_OK;
}

}  // namespace hal

using namespace hal;

extern "C" int halign_set_seed_policy(halign_ctx *ctx, int policy)
{
    if (!ctx || policy < 0 || policy > 1) return HALIGN_E_ARG;
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->seed_policy = policy;
    return HALIGN_OK;
}

extern "C" int halign_extract_minimizers(halign_ctx *ctx, int32_t n, const char *seqs, const int64_t *offsets,
                                         int32_t k, int32_t w, int32_t non_canonical, uint64_t *x_out,
                                         uint64_t *y_out, int64_t cap, int64_t *hit_offsets, int32_t *status_out)
{
    if (!ctx || n < 0 || (n > 0 && (!seqs || !offsets || !hit_offsets))) return HALIGN_E_ARG;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HCHECK(hipSetDevice(ctx->device), "set device");
    hit_offsets[0] = 0;
    if (n == 0) return HALIGN_OK;
    std::vector<int64_t> off(offsets, offsets + n + 1);
    for (auto &o : off) o -= offsets[0];
    const int64_t total = off[n];
    HCHECK(ctx->s_seqs.ensure((size_t)std::max<int64_t>(total, 1)), "alloc sequences");
    HCHECK(ctx->s_off.ensure((size_t)(n + 1) * 8), "alloc offsets");
    HCHECK(hipMemcpyAsync(ctx->s_seqs.p, seqs + offsets[0], (size_t)total, hipMemcpyHostToDevice, ctx->stream),
           "upload sequences");
    HCHECK(hipMemcpyAsync(ctx->s_off.p, off.data(), (size_t)(n + 1) * 8, hipMemcpyHostToDevice, ctx->stream),
           "upload offsets");
    std::vector<int64_t> hoff, counts;
    std::vector<int32_t> st;
    const int rc = seed_device(ctx, n, off, k, w, non_canonical, hoff, counts, st);
    if (rc != HALIGN_OK) return rc;
    int64_t need = 0;
    for (int s = 0; s < n; ++s) need += counts[s];
    if (need > cap) return fail(ctx, HALIGN_E_CAPACITY, "minimizer output buffer too small");
    std::vector<uint64_t> hx((size_t)std::max<int64_t>(hoff[n], 1)), hy(hx.size());
    HCHECK(hipMemcpy(hx.data(), ctx->s_hx.p, (size_t)hoff[n] * 8, hipMemcpyDeviceToHost), "download hits");
    HCHECK(hipMemcpy(hy.data(), ctx->s_hy.p, (size_t)hoff[n] * 8, hipMemcpyDeviceToHost), "download hits");
    int64_t o = 0;
    for (int s = 0; s < n; ++s) {
        hit_offsets[s] = o;
        if (counts[s]) {
            if (x_out) memcpy(x_out + o, hx.data() + hoff[s], (size_t)counts[s] * 8);
            if (y_out) memcpy(y_out + o, hy.data() + hoff[s], (size_t)counts[s] * 8);
        }
        o += counts[s];
        if (status_out) status_out[s] = st[s];
    }
    hit_offsets[n] = o;
    return HALIGN_OK;
}
