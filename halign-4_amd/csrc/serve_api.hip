// serve_api.hip -- continuous-batching request server for the per-query path
// (RefAligner::globalAlign with one resident center, src/align/ref_aligner.cpp:
// 91-125), the C ABI halign_server_*.
//
// Each worker owns a context (its own streams, buffers and traceback pool) and
// a copy of the center with its minimizer index resident in HBM. A request
// (a group of queries) is cut into chunks of at most chunk_queries queries. A
// worker that becomes free forms its next batch from the waiting chunks,
// taking one chunk from each waiting request in arrival order and going round
// again until the batch is full (max_batch_queries, or the batch's HBM
// footprint estimate reaches its share of the budget): a request that arrives
// while others are on the GPU is admitted at the next batch any worker forms,
// and a small request is not queued behind all of a large one. A request is
// done when its last chunk is. A query's result does not depend on which
// batch carries it.
//
// Device memory: each worker's traceback pool is allocated whole when the
// server is created (its share of the budget), so a batch's whole-pair
// fallbacks (queries with no anchors, aligned as one unbanded pair) always
// find their traceback slots, however the other buffers have grown; the DP
// launches fit their grids to the pool. The workers measure device memory in
// use after each batch, and halign_server_stats reports it beside the
// footprint estimates: resident (contexts, centers, indexes), the reserved
// pools, and what the batches' working buffers took.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "host_common.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    int64_t id = 0;
    int32_t n = 0;
    std::vector<char> seqs;
    std::vector<int64_t> off;   // n + 1 offsets into seqs
    std::vector<int64_t> est;   // footprint estimate per query
    Clock::time_point t_submit, t_start, t_done;
    bool started = false, done = false;
    int status = HALIGN_OK;
    std::string err;
    int32_t next = 0;           // first query not yet handed to a worker
    int32_t chunks_left = 0;    // chunks handed out or waiting, not finished
    std::vector<std::vector<uint32_t>> qcig;   // CIGAR per query
    std::vector<int32_t> stats;                // 4 per query
    int32_t batch_queries = 0;  // largest batch that carried a chunk of it
    std::vector<uint32_t> cigar;   // packed at completion
    std::vector<int64_t> coff;
};

struct Chunk {
    std::shared_ptr<Request> r;
    int32_t a = 0, b = 0;   // queries [a, b) of the request
};

}  // namespace

struct halign_server {
    int32_t rlen = 0;
    halign_server_config cfg{};
    int32_t chunk_queries = 256;
    int64_t batch_budget = 0;   // HBM footprint estimate allowed per batch
    int device = 0;
    size_t free_at_start = 0;   // free device memory before the workers' contexts
    struct Worker {
        halign_ctx *ctx = nullptr;
        halign_center *center = nullptr;
        std::thread th;
    };
    std::vector<Worker> workers;
    std::vector<int64_t> worker_max_est;   // largest footprint estimate of a batch, per worker
    int64_t resident = 0;       // device bytes of the contexts and centers, measured after creation
    int64_t tb_reserved = 0;    // device bytes of the reserved traceback pools
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Request>> pending;   // requests with queries not yet handed out
    std::unordered_map<int64_t, std::shared_ptr<Request>> live;
    int64_t next_id = 1;
    bool stop = false;
    std::string err;
    // measurements: batches run, the largest footprint estimate of a batch,
    // and the most device memory in use seen after a batch
    int64_t batches = 0, max_est = 0, max_used = 0;
};

namespace {

// HBM footprint estimate of one query of length L against a center of length R
// while its batch is on the device: sequence copies and codes, minimizer hits,
// anchors and their sort buffers, chain workspaces, segments and their CIGAR
// words, the capacity-layout and packed outputs (traceback slots come from the
// context's own pool).
int64_t query_footprint(int64_t L, int64_t R) { return 4096 + 200 * L + 12 * (R + L); }

// the next batch: one chunk from each waiting request in arrival order, round
// and round, until the batch is full (s->mu held)
std::vector<Chunk> form_batch(halign_server *s)
{
    std::vector<Chunk> batch;
    int64_t q = 0, bytes = 0;
    bool progress = true;
    while (progress && !s->pending.empty()) {
        progress = false;
        for (auto it = s->pending.begin(); it != s->pending.end();) {
            auto &r = *it;
            const int32_t b = std::min<int32_t>(r->n, r->next + s->chunk_queries);
            int64_t e = 0;
            for (int32_t k = r->next; k < b; ++k) e += r->est[(size_t)k];
            if (!batch.empty() && (q + (b - r->next) > s->cfg.max_batch_queries || bytes + e > s->batch_budget))
                return batch;
            batch.push_back({r, r->next, b});
            q += b - r->next;
            bytes += e;
            r->next = b;
            progress = true;
            if (r->next >= r->n) it = s->pending.erase(it);
            else ++it;
        }
    }
    return batch;
}

void worker_loop(halign_server *s, int wi)
{
    halign_server::Worker &w = s->workers[wi];
    (void)hipSetDevice(s->device);   // hipMemGetInfo below reads the thread's current device
    for (;;) {
        std::vector<Chunk> batch;
        int64_t est = 0;
        bool more = false;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] { return s->stop || !s->pending.empty(); });
            if (s->pending.empty()) return;   // stop requested and nothing left
            batch = form_batch(s);
            const auto now = Clock::now();
            for (auto &c : batch) {
                if (!c.r->started) {
                    c.r->started = true;
                    c.r->t_start = now;
                }
                for (int32_t k = c.a; k < c.b; ++k) est += c.r->est[(size_t)k];
            }
            more = !s->pending.empty();
        }
        if (more) s->cv_work.notify_one();   // more work for another free worker
        int64_t n = 0, bases = 0;
        for (auto &c : batch) {
            n += c.b - c.a;
            bases += c.r->off[c.b] - c.r->off[c.a];
        }
        std::vector<char> blob((size_t)std::max<int64_t>(bases, 1));
        std::vector<int64_t> off((size_t)n + 1);
        {
            int64_t qi = 0, b0 = 0;
            for (auto &c : batch) {
                const int64_t base = c.r->off[c.a], len = c.r->off[c.b] - base;
                if (len) memcpy(blob.data() + b0, c.r->seqs.data() + base, (size_t)len);
                for (int32_t k = c.a; k < c.b; ++k) off[qi++] = b0 + c.r->off[k] - base;
                b0 += len;
            }
            off[n] = b0;
        }
        const int64_t cap = bases + n * ((int64_t)s->rlen + 1) + 1;
        std::unique_ptr<uint32_t[]> cig(new (std::nothrow) uint32_t[(size_t)cap]);
        std::vector<int64_t> coff((size_t)n + 1);
        std::vector<int32_t> stats((size_t)std::max<int64_t>(n, 1) * 4);
        int rc = cig ? HALIGN_OK : HALIGN_E_NOMEM;
        if (rc == HALIGN_OK && n > 0)
            rc = halign_center_align(w.center, (int32_t)n, blob.data(), off.data(), cig.get(), cap, coff.data(),
                                     stats.data(), nullptr);
        if (n == 0) coff[0] = 0;
        const std::string e = rc == HALIGN_OK ? std::string() : std::string(halign_last_error(w.ctx));
        size_t free_now = 0, total = 0;
        (void)hipMemGetInfo(&free_now, &total);
        std::vector<std::shared_ptr<Request>> finished;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            const auto t_done = Clock::now();
            s->batches += 1;
            s->max_est = std::max(s->max_est, est);
            s->worker_max_est[(size_t)wi] = std::max(s->worker_max_est[(size_t)wi], est);
            if (s->free_at_start > free_now) s->max_used = std::max(s->max_used, (int64_t)(s->free_at_start - free_now));
            int64_t qi = 0;
            for (auto &c : batch) {
                Request &r = *c.r;
                if (rc != HALIGN_OK && r.status == HALIGN_OK) {
                    r.status = rc;
                    r.err = e;
                }
                r.batch_queries = std::max<int32_t>(r.batch_queries, (int32_t)n);
                for (int32_t k = c.a; k < c.b; ++k, ++qi) {
                    if (rc != HALIGN_OK) continue;
                    r.qcig[(size_t)k].assign(cig.get() + coff[qi], cig.get() + coff[qi + 1]);
                    for (int t = 0; t < 4; ++t) r.stats[(size_t)k * 4 + t] = stats[(size_t)qi * 4 + t];
                }
                if (--r.chunks_left == 0) {
                    r.t_done = t_done;
                    finished.push_back(c.r);
                }
            }
            for (auto &rp : finished) {
                Request &r = *rp;
                if (r.status == HALIGN_OK) {
                    size_t tot = 0;
                    for (auto &v : r.qcig) tot += v.size();
                    r.cigar.reserve(tot);
                    r.coff.assign((size_t)r.n + 1, 0);
                    for (int32_t k = 0; k < r.n; ++k) {
                        r.cigar.insert(r.cigar.end(), r.qcig[(size_t)k].begin(), r.qcig[(size_t)k].end());
                        r.coff[(size_t)k + 1] = (int64_t)r.cigar.size();
                    }
                }
                std::vector<std::vector<uint32_t>>().swap(r.qcig);
                r.done = true;
            }
        }
        if (!finished.empty()) s->cv_done.notify_all();
    }
}

void destroy(halign_server *s)
{
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto &w : s->workers)
        if (w.th.joinable()) w.th.join();
    for (auto &w : s->workers) {
        if (w.center) halign_center_free(w.center);
        if (w.ctx) halign_free(w.ctx);
    }
    delete s;
}

}  // namespace

extern "C" {

int halign_server_create(int device, const char *center, int32_t len, const halign_server_config *cfg,
                         halign_server **out)
{
    if (!center || len <= 0 || !out) return HALIGN_E_ARG;
    *out = nullptr;
    halign_server_config c{2, 4096, 0, 256, 0};
    if (cfg) c = *cfg;
    if (c.chunk_queries == 0) c.chunk_queries = 256;
    if (c.workers <= 0 || c.workers > 8 || c.max_batch_queries <= 0 || c.hbm_budget < 0 || c.chunk_queries < 0)
        return HALIGN_E_ARG;
    halign_server *s = new (std::nothrow) halign_server();
    if (!s) return HALIGN_E_NOMEM;
    s->rlen = len;
    s->cfg = c;
    s->chunk_queries = std::min(c.chunk_queries, c.max_batch_queries);   // a chunk fits in a batch
    s->device = device;
    if (hipSetDevice(device) != hipSuccess) {
        delete s;
        return HALIGN_E_HIP;
    }
    size_t free_b = 0, total_b = 0;
    (void)hipMemGetInfo(&free_b, &total_b);
    s->free_at_start = free_b;
    // default: half of free HBM for the batches' working sets, 35% for the
    // traceback pools, split between the workers
    const int64_t budget = c.hbm_budget > 0 ? c.hbm_budget : (int64_t)(0.85 * (double)free_b);
    s->batch_budget = (int64_t)(budget * (0.50 / 0.85)) / c.workers;
    const int64_t tb = (int64_t)(budget * (0.35 / 0.85)) / c.workers;
    s->workers.resize((size_t)c.workers);
    s->worker_max_est.assign((size_t)c.workers, 0);
    for (auto &w : s->workers) {
        int rc = halign_init(device, &w.ctx);
        if (rc == HALIGN_OK) rc = halign_set_traceback_budget(w.ctx, tb);
        if (rc == HALIGN_OK) rc = halign_center_create(w.ctx, center, len, &w.center);
        if (rc != HALIGN_OK) {
            destroy(s);
            return rc;
        }
    }
    (void)hipDeviceSynchronize();
    size_t free_c = 0;
    (void)hipMemGetInfo(&free_c, &total_b);
    s->resident = free_b > free_c ? (int64_t)(free_b - free_c) : 0;
    // the reserved traceback pools (no batch on the device yet: the pools
    // are not in use)
    for (auto &w : s->workers) {
        std::lock_guard<std::mutex> lk(w.ctx->mu);
        if (w.ctx->tb.ensure((size_t)tb, (size_t)tb) != hipSuccess) {
            (void)hipGetLastError();
            destroy(s);
            return HALIGN_E_NOMEM;
        }
        s->tb_reserved += (int64_t)w.ctx->tb.cap;
    }
    for (int i = 0; i < c.workers; ++i) s->workers[i].th = std::thread(worker_loop, s, i);
    *out = s;
    return HALIGN_OK;
}

int halign_server_submit(halign_server *s, int32_t n, const char *queries, const int64_t *offsets,
                         int64_t *request_id)
{
    if (!s || n < 0 || !request_id || (n > 0 && (!queries || !offsets))) return HALIGN_E_ARG;
    auto r = std::make_shared<Request>();
    r->n = n;
    r->off.resize((size_t)n + 1);
    r->est.resize((size_t)n);
    if (n > 0) {
        for (int32_t k = 0; k <= n; ++k) r->off[k] = offsets[k] - offsets[0];
        for (int32_t k = 0; k < n; ++k) {
            if (r->off[k + 1] < r->off[k]) return HALIGN_E_ARG;
            r->est[(size_t)k] = query_footprint(r->off[k + 1] - r->off[k], s->rlen);
        }
        r->seqs.assign(queries + offsets[0], queries + offsets[n]);
    } else {
        r->off[0] = 0;
    }
    r->qcig.resize((size_t)n);
    r->stats.assign((size_t)n * 4, 0);
    r->chunks_left = n > 0 ? (n + s->chunk_queries - 1) / s->chunk_queries : 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) return HALIGN_E_ARG;
        r->id = s->next_id++;
        r->t_submit = Clock::now();
        s->live[r->id] = r;
        *request_id = r->id;
        if (n == 0) {   // nothing to align
            r->t_start = r->t_done = r->t_submit;
            r->coff.assign(1, 0);
            r->done = true;
        } else {
            s->pending.push_back(r);
        }
    }
    if (n == 0) s->cv_done.notify_all();
    else s->cv_work.notify_one();
    return HALIGN_OK;
}

int halign_server_wait(halign_server *s, int64_t request_id, int32_t timeout_ms)
{
    if (!s) return HALIGN_E_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->live.find(request_id);
    if (it == s->live.end()) return HALIGN_E_ARG;
    std::shared_ptr<Request> r = it->second;
    auto ready = [&] { return r->done; };
    if (timeout_ms < 0) s->cv_done.wait(lk, ready);
    else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) return HALIGN_PENDING;
    return HALIGN_OK;
}

int halign_server_result(halign_server *s, int64_t request_id, uint32_t *cigar_out, int64_t cap,
                         int64_t *cigar_offsets, int32_t *stats, double *timing, int64_t *n_cigar)
{
    if (!s) return HALIGN_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->live.find(request_id);
    if (it == s->live.end()) return HALIGN_E_ARG;
    Request &r = *it->second;
    if (!r.done) return HALIGN_PENDING;
    if (r.status != HALIGN_OK) {
        s->err = r.err;
        const int rc = r.status;
        s->live.erase(it);
        return rc;
    }
    if (n_cigar) *n_cigar = (int64_t)r.cigar.size();
    if (timing) {
        timing[0] = std::chrono::duration<double, std::milli>(r.t_start - r.t_submit).count();
        timing[1] = std::chrono::duration<double, std::milli>(r.t_done - r.t_submit).count();
        timing[2] = (double)r.batch_queries;
    }
    if (!cigar_out) return HALIGN_OK;   // size query: the request stays
    if ((int64_t)r.cigar.size() > cap) return HALIGN_E_CAPACITY;
    if (!r.cigar.empty()) memcpy(cigar_out, r.cigar.data(), r.cigar.size() * 4);
    if (cigar_offsets)
        for (int32_t k = 0; k <= r.n; ++k) cigar_offsets[k] = r.coff.empty() ? 0 : r.coff[k];
    if (stats && r.n > 0) memcpy(stats, r.stats.data(), (size_t)r.n * 4 * sizeof(int32_t));
    s->live.erase(it);
    return HALIGN_OK;
}

int halign_server_stats(halign_server *s, halign_server_memory *out)
{
    if (!s || !out) return HALIGN_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    out->batches = s->batches;
    out->max_batch_estimate = s->max_est;
    out->sum_worker_estimates = 0;
    for (int64_t e : s->worker_max_est) out->sum_worker_estimates += e;
    out->batch_budget = s->batch_budget;
    out->resident = s->resident;
    out->traceback_reserved = s->tb_reserved;
    out->device_used = s->max_used;
    out->working_measured = std::max<int64_t>(0, s->max_used - s->resident - s->tb_reserved);
    return HALIGN_OK;
}

const char *halign_server_last_error(const halign_server *s) { return s ? s->err.c_str() : "no server"; }

void halign_server_free(halign_server *s)
{
    if (s) destroy(s);
}

}  // extern "C"
