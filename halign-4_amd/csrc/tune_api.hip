// tune_api.hip -- the measured cost table of the DP batch planner (C ABI
// halign_tune, halign_tune_table) and the estimate the planner reads from it.
//
// A ring shape's kernel runs its jobs as persistent workgroups; with the chip
// full, one job takes about anti-diagonals x ns_diag + ns_job, and
// jobs_in_flight jobs progress at once. halign_tune measures ns_diag at up to
// four bands per shape (a quarter, half, three quarters and all of the widest
// band the ring holds) from two job lengths each, on equal-length synthetic
// pairs that fill the chip once; ns_job is what the length fits leave over.
// Every shape is measured, the uniform-ticket forms of the sub-wave shapes
// included (their jobs are equal here, as the forms require); the planner
// chooses among per-group forms and the uniform pass after it converts the
// buckets whose jobs allow. A query-fixed shape (G lanes x S slots) is
// measured on unbanded jobs whose query fills 1..S slots (its cost per
// anti-diagonal steps with the slots its longest query needs), against two
// target lengths. The library starts every context from the table
// measured when it was built (dp_tune_table.h, written by halign.tune).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "../../include/halign_hip.h"
#include "dp_common.h"
#include "dp_host.h"
#include "host_common.h"

namespace {

#include "dp_tune_table.h"

// 64-bit LCG (Knuth's MMIX constants) for the synthetic pairs
struct Lcg {
    uint64_t s;
    uint32_t next()
    {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        return (uint32_t)(s >> 33);
    }
};

bool same_shape(const halign_tune_entry &e, int wi)
{
    return e.waves == hal::dp_wave_shape_waves(wi) && e.lanes == hal::dp_wave_shape_lanes(wi) &&
           e.ring == hal::dp_wave_shape_ring(wi) && e.uniform == hal::dp_wave_shape_form(wi);
}

// One timed launch of n equal jobs (qlen = tlen = len, band w) forced onto ring
// shape wi; the kernel time in ms, or < 0 on failure.
double time_shape(halign_ctx *ctx, int wi, int n, int len, int w, Lcg &rng)
{
    std::vector<uint8_t> tgt((size_t)len), qry;
    for (auto &c : tgt) c = (uint8_t)(rng.next() & 3);
    qry = tgt;
    for (int k = 0; k < len; ++k)   // about 2% substitutions
        if (rng.next() % 50 == 0) qry[(size_t)k] = (uint8_t)((qry[(size_t)k] + 1 + rng.next() % 3) & 3);
    std::vector<halign_dp_task> tasks((size_t)n);
    for (auto &t : tasks) {
        t.qlen = len;
        t.tlen = len;
        t.w = w;
        t.flag = 0x04;   // GENERIC_SC
        t.query_off = 0;
        t.target_off = 0;
    }
    std::vector<halign_dp_result> res((size_t)n);
    const int64_t cap = (int64_t)n * (2 * len + 4);
    std::vector<uint32_t> cig((size_t)cap);
    const int policy = ctx->policy, graphs = ctx->graphs;
    ctx->policy = 2 + wi;
    ctx->graphs = 0;
    ctx->tuning = true;   // a per-group shape stays per-group (no uniform tickets)
    double ms = -1.0;
    const int rc = halign_dp_batch(ctx, n, tasks.data(), qry.data(), tgt.data(), 0, 0, res.data(), cig.data(), cap,
                                   &ms);
    ctx->policy = policy;
    ctx->graphs = graphs;
    ctx->tuning = false;
    return rc == HALIGN_OK ? ms : -1.0;
}

// One timed launch of n equal unbanded jobs (query qlen, target tlen) forced
// onto query-fixed shape wi; the kernel time in ms, or < 0 on failure.
double time_qfix(halign_ctx *ctx, int wi, int n, int qlen, int tlen, Lcg &rng)
{
    std::vector<uint8_t> tgt((size_t)tlen), qry((size_t)qlen);
    for (auto &c : tgt) c = (uint8_t)(rng.next() & 3);
    const int at = (int)(rng.next() % (uint32_t)(tlen - qlen + 1));
    for (int k = 0; k < qlen; ++k) {   // a piece of the target at about 2% substitutions
        qry[(size_t)k] = tgt[(size_t)(at + k)];
        if (rng.next() % 50 == 0) qry[(size_t)k] = (uint8_t)((qry[(size_t)k] + 1 + rng.next() % 3) & 3);
    }
    std::vector<halign_dp_task> tasks((size_t)n);
    for (auto &t : tasks) {
        t.qlen = qlen;
        t.tlen = tlen;
        t.w = -1;
        t.flag = 0x06;   // GENERIC_SC | RIGHT
        t.query_off = 0;
        t.target_off = 0;
    }
    std::vector<halign_dp_result> res((size_t)n);
    const int64_t cap = (int64_t)n * (qlen + tlen + 4);
    std::vector<uint32_t> cig((size_t)cap);
    const int policy = ctx->policy, graphs = ctx->graphs;
    ctx->policy = 3 + hal::dp_wave_num_shapes() + (wi - hal::dp_wave_qfix_index(0));
    ctx->graphs = 0;
    double ms = -1.0;
    const int rc = halign_dp_batch(ctx, n, tasks.data(), qry.data(), tgt.data(), 0, 0, res.data(), cig.data(), cap,
                                   &ms);
    ctx->policy = policy;
    ctx->graphs = graphs;
    return rc == HALIGN_OK ? ms : -1.0;
}

}  // namespace

namespace hal {

void tune_load_default(halign_ctx *ctx)
{
    const int ns = dp_wave_num_shapes() + dp_wave_num_qfix();
    ctx->tune.assign((size_t)ns, halign_tune_entry{});
    for (const auto &e : kTuneDefault)
        for (int wi = 0; wi < ns; ++wi)
            if (same_shape(e, wi)) ctx->tune[(size_t)wi] = e;
}

bool tune_cost(const halign_ctx *ctx, int wi, double max_nc, double max_nd, double sum_nd, double njobs,
               double &crit, double &tp)
{
    if (wi < 0 || (size_t)wi >= ctx->tune.size()) return false;
    const halign_tune_entry &e = ctx->tune[(size_t)wi];
    if (e.ring == 0 || e.jobs_in_flight <= 0 || e.npoints < 1 || e.npoints > 4) return false;
    if (dp_wave_shape_qfix(wi)) {
        // max_nc: query blocks of the longest query; the cost steps with the
        // slots it needs (the first measured point that holds it)
        for (int k = 0; k < e.npoints; ++k) {
            if (max_nc <= e.ncol[k]) {
                crit = max_nd * e.ns_diag[k] + e.ns_job;
                tp = (sum_nd * e.ns_diag[k] + njobs * e.ns_job) / e.jobs_in_flight;
                return true;
            }
        }
        return false;
    }
    // no estimate below the bands measured (a quarter of the widest the ring
    // holds): a narrower band leaves most of the ring idle in ways the table
    // does not see
    if (max_nc < e.ncol[0]) return false;
    // a two-per-SIMD form is measured with every SIMD holding two workgroups:
    // no estimate for a bucket that cannot fill them (fewer jobs, or a
    // traceback budget without a slot for each; a slot is n_col_ x 8 bytes per
    // anti-diagonal of the longest job, rounded as the launch plan rounds it)
    if (dp_wave_shape_dense(wi) &&
        (njobs < e.jobs_in_flight ||
         (double)e.jobs_in_flight * (double)round_slot((size_t)(max_nd * max_nc * 8.0)) > (double)ctx->tb_budget))
        return false;
    double td = e.ns_diag[0];
    for (int k = 1; k < e.npoints; ++k) {
        if (max_nc <= e.ncol[k - 1]) break;
        const double f = std::min(1.0, (max_nc - e.ncol[k - 1]) / (double)(e.ncol[k] - e.ncol[k - 1]));
        td = e.ns_diag[k - 1] + f * (e.ns_diag[k] - e.ns_diag[k - 1]);
    }
    crit = max_nd * td + e.ns_job;
    tp = (sum_nd * td + njobs * e.ns_job) / e.jobs_in_flight;
    return true;
}

}  // namespace hal

using namespace hal;

extern "C" {

int halign_tune(halign_ctx *ctx)
{
    if (!ctx) return HALIGN_E_ARG;
    PhaseClock clk("tune");
    const int nring = dp_wave_num_shapes();
    const int ns = nring + dp_wave_num_qfix();
    std::vector<halign_tune_entry> tab((size_t)ns, halign_tune_entry{});
    Lcg rng{0x7A11E5ull};
    for (int wi = 0; wi < ns; ++wi) {
        int groups = 0;
        {
            std::lock_guard<std::mutex> lk(ctx->mu);
            if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, HALIGN_E_HIP, "set device");
            groups = dp_wave_max_groups(wi, false, ctx->num_cu);
        }
        if (groups <= 0) return fail(ctx, HALIGN_E_HIP, "occupancy query failed");
        const int conc = groups * dp_wave_jobs_per_group(wi);
        const int ring = dp_wave_shape_ring(wi);
        halign_tune_entry &e = tab[(size_t)wi];
        e.waves = dp_wave_shape_waves(wi);
        e.lanes = dp_wave_shape_lanes(wi);
        e.ring = ring;
        e.uniform = dp_wave_shape_form(wi);
        e.jobs_in_flight = conc;
        int cols[4], np = 0;
        double t0sum = 0;
        if (wi >= nring) {
            // query-fixed: queries filling 1..S slots (at most four points),
            // unbanded, targets of len1 and len1 + 1024 bases, as many jobs as
            // the chip keeps in flight (the estimate divides by that count)
            const int lanes = dp_wave_shape_lanes(wi), slots = ring / lanes;
            const int n = conc;
            for (int k = std::max(1, slots - 3); k <= slots; ++k) cols[np++] = k * lanes;
            for (int c = 0; c < np; ++c) {
                const int qlen = 16 * cols[c];
                const int len1 = std::max(1024, qlen), len2 = len1 + 1024;
                double best[2] = {1e30, 1e30};
                for (int rep = 0; rep < 2; ++rep) {
                    const int lens[2] = {len1, len2};
                    for (int l = 0; l < 2; ++l) {
                        const double ms = time_qfix(ctx, wi, n, qlen, lens[l], rng);
                        if (ms < 0) return HALIGN_E_KERNEL;
                        best[l] = std::min(best[l], ms);
                    }
                }
                const double nd1 = (double)qlen + len1 - 1, nd2 = (double)qlen + len2 - 1;
                const double td = std::max(1e-3, (best[1] - best[0]) * 1e6 / (nd2 - nd1));
                e.ncol[c] = cols[c];
                e.ns_diag[c] = td;
                t0sum += std::max(0.0, best[0] * 1e6 - nd1 * td);
            }
            e.npoints = np;
            e.ns_job = t0sum / np;
            if (clk.on)
                fprintf(stderr, "[halign tune] shape %2d query-fixed %d lanes x %d slots: %6d jobs in flight, ns/diag %.1f (%d blocks) .. %.1f (%d blocks), ns/job %.0f\n",
                        wi, lanes, slots, conc, e.ns_diag[0], e.ncol[0], e.ns_diag[np - 1], e.ncol[np - 1], e.ns_job);
            continue;
        }
        // bands of k/4 of the widest the ring holds (at least 2 columns, at
        // most 513: wider bands are taken as 513); band w = 16 c - 17 gives
        // n_col_ = c for jobs of at least w + 1 bases
        for (int k = 1; k <= 4; ++k) {
            const int c = std::min(513, std::max(2, ((ring - 2) * k + 2) / 4));
            if (np == 0 || c > cols[np - 1]) cols[np++] = c;
        }
        for (int c = 0; c < np; ++c) {
            const int w = 16 * cols[c] - 17;
            const int len1 = std::max(512, 2 * (w + 1)), len2 = 2 * len1;
            double best[2] = {1e30, 1e30};
            for (int rep = 0; rep < 2; ++rep) {
                const int lens[2] = {len1, len2};
                for (int l = 0; l < 2; ++l) {
                    const double ms = time_shape(ctx, wi, conc, lens[l], w, rng);
                    if (ms < 0) return HALIGN_E_KERNEL;
                    best[l] = std::min(best[l], ms);
                }
            }
            const double nd1 = 2.0 * len1 - 1, nd2 = 2.0 * len2 - 1;
            const double td = std::max(1e-3, (best[1] - best[0]) * 1e6 / (nd2 - nd1));
            e.ncol[c] = cols[c];
            e.ns_diag[c] = td;
            t0sum += std::max(0.0, best[0] * 1e6 - nd1 * td);
        }
        e.npoints = np;
        e.ns_job = t0sum / np;
        if (clk.on) {
            fprintf(stderr, "[halign tune] shape %2d waves %d lanes %3d ring %4d%s: %6d jobs in flight, ns/diag", wi,
                    e.waves, e.lanes, ring, e.uniform == 1 ? "u" : e.uniform == 2 ? "d" : " ", conc);
            for (int c = 0; c < np; ++c) fprintf(stderr, " %.1f (n_col %d)", e.ns_diag[c], e.ncol[c]);
            fprintf(stderr, ", ns/job %.0f\n", e.ns_job);
        }
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->tune.swap(tab);
    return HALIGN_OK;
}

int halign_tune_table(halign_ctx *ctx, halign_tune_entry *out, int32_t cap, int32_t *n)
{
    if (!ctx || !n || cap < 0 || (cap > 0 && !out)) return HALIGN_E_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    int32_t k = 0;
    for (const auto &e : ctx->tune) {
        if (e.ring == 0) continue;
        if (k < cap) out[k] = e;
        ++k;
    }
    *n = k;
    return HALIGN_OK;
}

}  // extern "C"
