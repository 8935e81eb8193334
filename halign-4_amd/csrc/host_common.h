// host_common.h -- host-side pieces shared by the C-ABI implementation files:
// the context (device, stream, events, reusable device buffers), the error
// helper and the HIP status check.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/halign_hip.h"
#include "dp_common.h"

namespace hal {

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    // At least n bytes. A buffer that has to grow grows by half again (up to
    // limit, when given) so that batches of varying size do not reallocate
    // (hipFree + hipMalloc of a large buffer costs far more than a batch).
    hipError_t ensure(size_t n, size_t limit = 0)
    {
        if (n <= cap) return hipSuccess;
        size_t want = cap ? std::max(n, cap + cap / 2) : n;
        if (limit && want > limit) want = std::max(n, limit);
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        hipError_t e = hipMalloc(&p, want);
        if (e != hipSuccess && want > n) {
            (void)hipGetLastError();
            want = n;
            e = hipMalloc(&p, want);
        }
        if (e == hipSuccess) cap = want;
        return e;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    template <class T> T *as() const { return (T *)p; }
};

// Pinned host memory, grown by half again when it has to grow. The CPU's first
// read of freshly allocated pinned pages is an order of magnitude slower than
// later reads, so buffers that receive results chunk after chunk are kept.
struct HostPinned {
    char *p = nullptr;
    size_t cap = 0;
    bool ensure(size_t n)
    {
        if (n <= cap) return true;
        const size_t want = std::max(n, cap + cap / 2);
        release();
        if (hipHostMalloc((void **)&p, want, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            return false;
        }
        cap = want;
        return true;
    }
    void release()
    {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
};

}  // namespace hal

struct halign_ctx {
    int device = 0;
    int num_cu = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // Side streams (default priority): all three run DP buckets beside the
    // main stream; aux[2] also runs the per-query path's small-segment DPs,
    // which come before that batch's large-segment buckets.
    static constexpr int kAux = 3;
    hipStream_t aux[kAux] = {};
    hipEvent_t fork = nullptr, join[kAux] = {}, pfork = nullptr, pjoin = nullptr;
    std::mutex mu;
    std::string err;
    hal::DevBuf seq, seqb, jobs, outs, counters, tb, cig, qparts;
    std::vector<hal::DpJob> hjobs;   // host copy of the last batch's jobs (kept between batches)
    hal::DevBuf crows, ccounts, ccons;   // consensus rows (host input), counts, characters
    // seeding: sequences, offsets, packed codes and N masks, hits, counts, status, workspace
    hal::DevBuf s_seqs, s_off, s_woff, s_moff, s_codes, s_nmask, s_hoff, s_hx, s_hy, s_cnt, s_st, s_list;
    size_t tb_budget = 0;
    int policy = 0;        // 0: auto kernel choice, 1: multi-wave kernels only, 2+i: ring shape i
    // Batched DP launches are captured into a HIP graph and replayed while the
    // launch plan (kernels, grids, buffers, streams) repeats.
    int graphs = 1;
    struct CachedGraph {
        std::vector<int64_t> sig;
        hipGraphExec_t exec;
    };
    std::vector<CachedGraph> gcache;   // most recent last, at most kGraphCache
    static constexpr size_t kGraphCache = 8;
    int64_t graph_captures = 0, graph_replays = 0;
    int seed_policy = 0;   // 0: wave form where it applies, 1: sequential form only
    // mash sketches: sequences, offsets, key ranges, keys, sort output, counts,
    // sort storage, selection outputs
    hal::DevBuf m_seqs, m_off, m_koff, m_keys, m_keys2, m_nkeys, m_tmp, m_sel;
    // keep-length file runs: one at a time per context, with two record slots
    // and the last run's center kept for the next run
    std::mutex msa_mu;
    hal::HostPinned msa_out[2], msa_tmp;
    struct halign_center *msa_center = nullptr;
    std::string msa_key;
    // per-query path: two pinned staging buffers for query uploads (each
    // reused after the event recorded behind its last copy) and one for
    // downloads of the large-segment records
    hal::HostPinned up_pin[2], dn_pin;
    hal::HostPinned out_pin;   // the batch runner's download of its DpOuts
    hipEvent_t up_ev[2] = {};
    // Centers and pair sets live on the context's streams and buffers. They
    // are counted here (under mu); halign_free with some still live only
    // marks the context closing, and the last child's free releases it.
    int children = 0;
    bool closing = false;
    // this device's address of the static scratch of halign_consensus_noalloc
    void *noalloc_scratch = nullptr;
    // planner cost table, one entry per ring shape index (ring == 0: none)
    std::vector<halign_tune_entry> tune;
    bool tuning = false;   // halign_tune's launches: forced shapes run as they are
    // per-bucket timing (halign_set_bucket_timing): an event pair per DP launch
    bool bucket_timing = false;
    std::vector<halign_bucket_stat> last_buckets;
    std::vector<hipEvent_t> bev;   // created on first use, 2 per launch
};

namespace hal {

// Event pair k of the bucket-timing pool (created as needed); false on failure.
inline bool bucket_events(halign_ctx *ctx, size_t k, hipEvent_t &a, hipEvent_t &b)
{
    while (ctx->bev.size() < 2 * (k + 1)) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return false;
        ctx->bev.push_back(e);
    }
    a = ctx->bev[2 * k];
    b = ctx->bev[2 * k + 1];
    return true;
}

// Registers a new center or pair set on ctx (ctx->mu held by the caller).
inline void ctx_attach(halign_ctx *ctx) { ++ctx->children; }
// Releases the context itself: streams, events, buffers, graphs.
void ctx_destroy(halign_ctx *ctx);
// Unregisters a child whose resources are already released (ctx->mu not
// held); releases the context when it was freed and this was its last child.
void ctx_detach(halign_ctx *ctx);

// Host phase timing on stderr when the environment sets HALIGN_TRACE
struct PhaseClock {
    bool on;
    std::chrono::steady_clock::time_point t;
    const char *scope;
    explicit PhaseClock(const char *sc)
        : on(std::getenv("HALIGN_TRACE") != nullptr), t(std::chrono::steady_clock::now()), scope(sc)
    {
    }
    void mark(const char *what)
    {
        if (!on) return;
        const auto n = std::chrono::steady_clock::now();
        fprintf(stderr, "[halign %s] %-28s %9.3f ms\n", scope, what,
                std::chrono::duration<double, std::milli>(n - t).count());
        t = n;
    }
};

inline int fail(halign_ctx *ctx, int code, const char *what, hipError_t e = hipSuccess)
{
    char buf[512];
    if (e != hipSuccess) snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
    else snprintf(buf, sizeof buf, "%s", what);
    if (ctx) ctx->err = buf;
    return code;
}

#define HCHECK(expr, what)                                           \
    do {                                                             \
        hipError_t e_ = (expr);                                      \
        if (e_ != hipSuccess) return fail(ctx, HALIGN_E_HIP, what, e_); \
    } while (0)

}  // namespace hal
