// serve_api.hip -- continuous-batching request server for the per-query path
// (RefAligner::globalAlign with one resident center, src/align/ref_aligner.cpp:
// 91-125), the C ABI halign_server_*.
//
// Each worker owns a context (its own stream and buffers) and a copy of the
// center with its minimizer index resident in HBM. Requests (groups of
// queries) wait in one FIFO; a free worker takes the requests at the head of
// it, as many as fit its batch limits (query count and an HBM footprint
// estimate), runs them as one batch and hands each request its own results.
// With two or more workers, batches formed while another is on the GPU run
// beside it. A query's result does not depend on which batch carries it.
#include <hip/hip_runtime.h>

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

#include "../../include/halign_hip.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    int64_t id = 0;
    int32_t n = 0;
    std::vector<char> seqs;
    std::vector<int64_t> off;   // n + 1 offsets into seqs
    int64_t est_bytes = 0;
    Clock::time_point t_submit, t_start, t_done;
    bool done = false;
    int status = HALIGN_OK;
    std::string err;
    std::vector<uint32_t> cigar;
    std::vector<int64_t> coff;
    std::vector<int32_t> stats;
    int32_t batch_queries = 0;
};

}  // namespace

struct halign_server {
    int32_t rlen = 0;
    halign_server_config cfg{};
    int64_t batch_budget = 0;   // HBM footprint estimate allowed per batch
    struct Worker {
        halign_ctx *ctx = nullptr;
        halign_center *center = nullptr;
        std::thread th;
    };
    std::vector<Worker> workers;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Request>> pending;
    std::unordered_map<int64_t, std::shared_ptr<Request>> live;
    int64_t next_id = 1;
    bool stop = false;
    std::string err;
};

namespace {

// HBM footprint estimate of one query of length L against a center of length R
// while its batch is on the device: sequence copies and codes, minimizer hits,
// anchors and their sort buffers, chain workspaces, segments and their CIGAR
// words, the capacity-layout and packed outputs (traceback slots come from the
// context's own pool).
int64_t query_footprint(int64_t L, int64_t R) { return 4096 + 200 * L + 12 * (R + L); }

void worker_loop(halign_server *s, int wi)
{
    halign_server::Worker &w = s->workers[wi];
    for (;;) {
        std::vector<std::shared_ptr<Request>> batch;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] { return s->stop || !s->pending.empty(); });
            if (s->pending.empty()) return;   // stop requested and nothing left
            int64_t q = 0, bytes = 0;
            while (!s->pending.empty()) {
                const auto &r = s->pending.front();
                if (!batch.empty() &&
                    (q + r->n > s->cfg.max_batch_queries || bytes + r->est_bytes > s->batch_budget))
                    break;
                q += r->n;
                bytes += r->est_bytes;
                batch.push_back(r);
                s->pending.pop_front();
            }
        }
        const auto t_start = Clock::now();
        int64_t n = 0, bases = 0;
        for (auto &r : batch) {
            n += r->n;
            bases += r->off[r->n];
        }
        std::vector<char> blob((size_t)bases);
        std::vector<int64_t> off((size_t)n + 1);
        {
            int64_t qi = 0, b0 = 0;
            for (auto &r : batch) {
                if (r->off[r->n]) memcpy(blob.data() + b0, r->seqs.data(), (size_t)r->off[r->n]);
                for (int32_t k = 0; k < r->n; ++k) off[qi++] = b0 + r->off[k];
                b0 += r->off[r->n];
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
        {
            std::lock_guard<std::mutex> lk(s->mu);
            const auto t_done = Clock::now();
            int64_t qi = 0;
            for (auto &r : batch) {
                r->t_start = t_start;
                r->t_done = t_done;
                r->status = rc;
                r->err = e;
                r->batch_queries = (int32_t)n;
                if (rc == HALIGN_OK) {
                    const int64_t c0 = coff[qi], c1 = coff[qi + r->n];
                    r->cigar.assign(cig.get() + c0, cig.get() + c1);
                    r->coff.resize((size_t)r->n + 1);
                    for (int32_t k = 0; k <= r->n; ++k) r->coff[k] = coff[qi + k] - c0;
                    r->stats.assign(stats.begin() + 4 * qi, stats.begin() + 4 * (qi + r->n));
                }
                qi += r->n;
                r->done = true;
            }
        }
        s->cv_done.notify_all();
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
    halign_server_config c{2, 4096, 0};
    if (cfg) c = *cfg;
    if (c.workers <= 0 || c.workers > 8 || c.max_batch_queries <= 0 || c.hbm_budget < 0) return HALIGN_E_ARG;
    halign_server *s = new (std::nothrow) halign_server();
    if (!s) return HALIGN_E_NOMEM;
    s->rlen = len;
    s->cfg = c;
    if (hipSetDevice(device) != hipSuccess) {
        delete s;
        return HALIGN_E_HIP;
    }
    size_t free_b = 0, total_b = 0;
    (void)hipMemGetInfo(&free_b, &total_b);
    // default: half of free HBM for the batches' working sets, 35% for the
    // traceback pools, split between the workers
    const int64_t budget = c.hbm_budget > 0 ? c.hbm_budget : (int64_t)(0.85 * (double)free_b);
    s->batch_budget = (int64_t)(budget * (0.50 / 0.85)) / c.workers;
    const int64_t tb = (int64_t)(budget * (0.35 / 0.85)) / c.workers;
    s->workers.resize((size_t)c.workers);
    for (auto &w : s->workers) {
        int rc = halign_init(device, &w.ctx);
        if (rc == HALIGN_OK) rc = halign_set_traceback_budget(w.ctx, tb);
        if (rc == HALIGN_OK) rc = halign_center_create(w.ctx, center, len, &w.center);
        if (rc != HALIGN_OK) {
            destroy(s);
            return rc;
        }
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
    int64_t est = 0;
    if (n > 0) {
        for (int32_t k = 0; k <= n; ++k) r->off[k] = offsets[k] - offsets[0];
        for (int32_t k = 0; k < n; ++k) {
            if (r->off[k + 1] < r->off[k]) return HALIGN_E_ARG;
            est += query_footprint(r->off[k + 1] - r->off[k], s->rlen);
        }
        r->seqs.assign(queries + offsets[0], queries + offsets[n]);
    } else {
        r->off[0] = 0;
    }
    r->est_bytes = est;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) return HALIGN_E_ARG;
        r->id = s->next_id++;
        r->t_submit = Clock::now();
        s->live[r->id] = r;
        s->pending.push_back(r);
        *request_id = r->id;
    }
    s->cv_work.notify_one();
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

const char *halign_server_last_error(const halign_server *s) { return s ? s->err.c_str() : "no server"; }

void halign_server_free(halign_server *s)
{
    if (s) destroy(s);
}

}  // extern "C"
