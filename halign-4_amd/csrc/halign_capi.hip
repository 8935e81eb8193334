// halign_capi.hip -- host side of libhalign_hip.so: the C ABI declared in
// include/halign_hip.h, sequence preparation, job bucketing by ring shape,
// traceback-pool sizing and kernel launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/halign_hip.h"
#include "dp_common.h"
#include "dp_host.h"
#include "host_common.h"

namespace hal {
int dp_num_shapes();
int dp_shape_ring(int idx);
int dp_shape_threads(int idx);
int dp_shape_max_groups(int idx, bool stats, int num_cu);
size_t dp_shape_slot_bytes(int idx, int qlen, int tlen, int ncol);
hipError_t dp_launch(int idx, bool stats, int groups, const DpJob *jobs, int njobs, int *counter,
                     uint8_t *tb_pool, size_t tb_slot_bytes, DpOut *outs, hipStream_t stream);
hipError_t consensus_launch(const uint8_t *rows, int64_t row_stride, int64_t aln_len, int64_t row_end, int64_t batch,
                            int num_cu, uint32_t *counts, char *cons, hipStream_t stream);
hipError_t zero_words_launch(uint32_t *p, int64_t n, hipStream_t stream);
}  // namespace hal

using namespace hal;

// Device scratch of halign_consensus_noalloc: the 7 counts per column
// (uint32) and the consensus characters of one call, for alignments of up to
// kNoAllocMaxColumns columns. One copy per GPU, alive until the process ends;
// host code uses the address hipGetSymbolAddress returns for the context's
// device (halign_ctx::noalloc_scratch), never this name.
static constexpr int64_t kNoAllocMaxColumns = (int64_t)1 << 20;
static constexpr size_t kNoAllocScratchBytes = (size_t)kNoAllocMaxColumns * (7 * 4 + 1);
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];



namespace hal {

void spread_target(const uint8_t *codes, int len, uint16_t *even, uint16_t *odd)
{
    // even[-kSeqPad .. len+kSeqPad), odd[k] = even[k+1]
    const int n = len + 2 * kSeqPad;
    for (int k = 0; k < n; ++k) {
        const int t = k - kSeqPad;
        even[k] = (t >= 0 && t < len) ? (uint16_t)((codes[t] << 8) | 0x0C) : (uint16_t)0x000C;
    }
    for (int k = 0; k < n; ++k) odd[k] = (k + 1 < n) ? even[k + 1] : (uint16_t)0x000C;
}

void spread_query_rev(const uint8_t *codes, int len, uint16_t *even, uint16_t *odd)
{
    const int n = len + 2 * kSeqPad;
    for (int k = 0; k < n; ++k) {
        const int t = k - kSeqPad;   // reversed index
        even[k] = (t >= 0 && t < len) ? (uint16_t)(codes[len - 1 - t] << 8) : (uint16_t)0;
    }
    for (int k = 0; k < n; ++k) odd[k] = (k + 1 < n) ? even[k + 1] : (uint16_t)0;
}

void bytes_target(const uint8_t *codes, int len, uint8_t *out)
{
    for (int k = 0; k < len + 2 * kBytePad; ++k) {
        const int t = k - kBytePad;
        out[k] = (t >= 0 && t < len) ? codes[t] : 0;
    }
}

void bytes_query_rev(const uint8_t *codes, int len, uint8_t *out)
{
    for (int k = 0; k < len + 2 * kBytePad; ++k) {
        const int t = k - kBytePad;
        out[k] = (t >= 0 && t < len) ? codes[len - 1 - t] : 0;
    }
}

// Shape ids: [0, dp_num_shapes()) multi-wave kernels (any mode), then
// kWaveBase + i for ring shape i and, after the ring shapes, query-fixed
// shape i (CIGAR mode).
constexpr int kWaveBase = 16;
constexpr int kNumShapeIds = kWaveBase + 48;

// Shape-table lookups the per-job shape choice makes, computed once: the
// query-fixed shapes' capacities (query blocks) and the ring shape of each
// entry of the automatic first-shape list (-1 where the table has none).
namespace {
// A job's first shape, which sorts it into a bucket: the first of this list,
// {waves, lanes per job, slots}, that holds the band. With the automatic
// policy run_specs then gives each bucket the shape the measured cost table
// prefers (plan_bucket_shapes).
constexpr int kAuto[][3] = {{0, 4, 1},  {0, 4, 2},  {0, 16, 1},  {0, 16, 2}, {0, 16, 3}, {1, 64, 1},
                            {1, 64, 2}, {2, 128, 2}, {3, 192, 2}, {4, 256, 2}, {8, 512, 2}};
constexpr int kNumAuto = (int)(sizeof(kAuto) / sizeof(kAuto[0]));
struct ShapeLookup {
    int nq = 0, nw = 0;
    int qcap[32] = {};       // query-fixed shape q holds queries of at most qcap[q] blocks
    int auto_idx[kNumAuto];  // ring shape of kAuto[k]
    int auto_ring[kNumAuto];
};
const ShapeLookup &shape_lookup()
{
    static const ShapeLookup t = [] {
        ShapeLookup x;
        x.nw = dp_wave_num_shapes();
        x.nq = std::min(dp_wave_num_qfix(), 32);
        for (int q = 0; q < x.nq; ++q) x.qcap[q] = dp_wave_shape_ring(dp_wave_qfix_index(q));
        for (int k = 0; k < kNumAuto; ++k) {
            x.auto_idx[k] = -1;
            x.auto_ring[k] = 0;
            for (int i = 0; i < x.nw && x.auto_idx[k] < 0; ++i)
                if (dp_wave_shape_waves(i) == kAuto[k][0] && dp_wave_shape_lanes(i) == kAuto[k][1] &&
                    dp_wave_shape_ring(i) == kAuto[k][1] * kAuto[k][2]) {
                    x.auto_idx[k] = i;
                    x.auto_ring[k] = dp_wave_shape_ring(i);
                }
        }
        return x;
    }();
    return t;
}
}  // namespace

// Query-fixed shape (dp_qfix.hip) for a CIGAR-mode job, -1 if none: the band
// is the whole matrix (w >= max(qlen, tlen) - 1), the query is the short side
// and fits the shape. forced >= 0: only that query-fixed shape; else the
// smallest that holds the query (the most jobs per wave).
static int qfix_pick(int qlen, int tlen, int w, int forced)
{
    if (qlen <= 0 || qlen > tlen || w < std::max(qlen, tlen) - 1) return -1;
    const int nb = (qlen + 15) / 16;
    const ShapeLookup &t = shape_lookup();
    if (forced >= 0) return forced < t.nq && nb <= t.qcap[forced] ? forced : -1;
    int best = -1;
    for (int q = 0; q < t.nq; ++q)
        if (nb <= t.qcap[q] && (best < 0 || t.qcap[q] < t.qcap[best])) best = q;
    return best;
}
constexpr int kDpStreams = 1 + halign_ctx::kAux;   // main + side streams for DP buckets

int pick_shape(int ncol, bool stats, int policy, int qlen, int tlen, int w)
{
    const ShapeLookup &t = shape_lookup();
    const int nw = t.nw;
    if (policy == 2 + nw) return dp_num_shapes() - 1;   // the wide-band kernel only
    if (!stats && policy != 1) {
        // query-fixed shapes: automatically for the jobs they hold, or forced
        // (policy 3 + nw + q) where the job allows
        if (policy == 0 || policy >= 3 + nw) {
            const int q = qfix_pick(qlen, tlen, w, policy == 0 ? -1 : policy - 3 - nw);
            if (q >= 0) return kWaveBase + dp_wave_qfix_index(q);
        }
        if (policy >= 2 && policy - 2 < nw && ncol <= dp_wave_shape_ring(policy - 2) - 2)
            return kWaveBase + policy - 2;
        for (int k = 0; k < kNumAuto; ++k)
            if (t.auto_idx[k] >= 0 && ncol <= t.auto_ring[k] - 2) return kWaveBase + t.auto_idx[k];
    }
    const int ns = dp_num_shapes();
    for (int i = 0; i < ns; ++i)
        if (ncol <= dp_shape_ring(i) - 1) return i;
    return -1;
}

// Batch-level shape choice for the ring-kernel CIGAR buckets (automatic
// policy), from the measured cost table (tune_api.hip). Buckets given the same
// shape (and gap direction) run as one launch. A launch on shape s takes about
// max(crit, tp): crit = its longest job's time on s, tp = all its jobs' time /
// s's jobs in flight. The launches share the main and the three side streams
// (one hardware queue each), longest first on the least-loaded stream, and the
// chip: the batch takes about max(stream makespan, sum of tp). Start each
// bucket on its cheapest shape in tp, then move one bucket at a time to the
// shape (possibly one another bucket already has: a merge) that lowers that
// estimate most, while any move does.
struct BucketCost {
    int bi, rt;
    double max_nd, sum_nd, n;
    int max_nc;
    std::vector<int> shapes;   // candidate ring shapes, by rising tp
    int cur;                   // the bucket's shape (a ring shape index)
};

static double plan_estimate(const halign_ctx *ctx, const std::vector<BucketCost> &bc)
{
    struct Group { int shape, rt, max_nc; double max_nd, sum_nd, n; };
    std::vector<Group> gs;
    for (const auto &c : bc) {
        Group *g = nullptr;
        for (auto &x : gs)
            if (x.shape == c.cur && x.rt == c.rt) g = &x;
        if (!g) {
            gs.push_back({c.cur, c.rt, 0, 0, 0, 0});
            g = &gs.back();
        }
        g->max_nc = std::max(g->max_nc, c.max_nc);
        g->max_nd = std::max(g->max_nd, c.max_nd);
        g->sum_nd += c.sum_nd;
        g->n += c.n;
    }
    std::vector<double> t;
    double sum_tp = 0;
    for (const auto &g : gs) {
        double cr, tp;
        if (!tune_cost(ctx, g.shape, g.max_nc, g.max_nd, g.sum_nd, g.n, cr, tp)) return 1e300;
        t.push_back(std::max(cr, tp));
        sum_tp += tp;
    }
    std::sort(t.begin(), t.end(), std::greater<double>());
    double load[kDpStreams] = {};
    for (double x : t) *std::min_element(load, load + kDpStreams) += x;
    return std::max(*std::max_element(load, load + kDpStreams), sum_tp);
}

static void plan_bucket_shapes(const halign_ctx *ctx, const std::vector<Spec> &specs,
                               const std::vector<std::vector<int>> &buckets, std::vector<int> &shape_of_bucket)
{
    std::vector<BucketCost> bc;
    for (size_t bi = 0; bi < buckets.size(); ++bi) {
        const auto &b = buckets[bi];
        const int shape = (int)bi / 4;
        shape_of_bucket[bi] = shape;
        if (b.empty() || shape < kWaveBase || (bi & 2)) continue;
        // a query-fixed bucket (full band, short query) chooses among the
        // query-fixed shapes that hold its longest query (max_nc: query
        // blocks), a ring bucket among the ring shapes that hold its band
        const bool qf = dp_wave_shape_qfix(shape - kWaveBase) != 0;
        BucketCost c{(int)bi, (int)(bi & 1), 0, 0, (double)b.size(), 0, {}, 0};
        for (int i : b) {
            const Spec &s = specs[i];
            const double nd = (double)s.qlen + s.tlen - 1;
            c.max_nc = std::max(c.max_nc, qf ? (s.qlen + 15) / 16 : n_col(s.qlen, s.tlen, eff_w(s.w, s.qlen, s.tlen)));
            c.max_nd = std::max(c.max_nd, nd);
            c.sum_nd += nd;
        }
        std::vector<std::pair<double, int>> cand;
        for (int i = 0; i < dp_wave_num_shapes() + dp_wave_num_qfix(); ++i) {
            double cr, tp;
            const bool fits = qf ? dp_wave_shape_qfix(i) && c.max_nc <= dp_wave_shape_ring(i)
                                 : !dp_wave_shape_qfix(i) && !dp_wave_shape_uniform(i) &&
                                       c.max_nc <= dp_wave_shape_ring(i) - 2;
            if (fits && tune_cost(ctx, i, c.max_nc, c.max_nd, c.sum_nd, c.n, cr, tp)) cand.push_back({tp, i});
        }
        if (cand.empty()) continue;
        std::sort(cand.begin(), cand.end());
        for (auto &x : cand) c.shapes.push_back(x.second);
        c.cur = c.shapes[0];
        bc.push_back(c);
    }
    double est = plan_estimate(ctx, bc);
    for (int guard = 0; guard < 64 && !bc.empty(); ++guard) {
        double best = est;
        int bb = -1, bs = -1;
        for (size_t k = 0; k < bc.size(); ++k) {
            const int old = bc[k].cur;
            for (int s : bc[k].shapes) {
                if (s == old) continue;
                bc[k].cur = s;
                const double e = plan_estimate(ctx, bc);
                if (e < best * (1.0 - 1e-9)) { best = e; bb = (int)k; bs = s; }
            }
            bc[k].cur = old;
        }
        if (bb < 0) break;
        bc[(size_t)bb].cur = bs;
        est = best;
    }
    for (auto &c : bc) {
        shape_of_bucket[c.bi] = kWaveBase + c.cur;
        if (std::getenv("HALIGN_TRACE")) {
            fprintf(stderr, "[halign plan] bucket %d: %.0f jobs, n_col <= %d, longest %.0f anti-diagonals:", c.bi,
                    c.n, c.max_nc, c.max_nd);
            for (int i : c.shapes) {
                double cr, tp;
                if (!tune_cost(ctx, i, c.max_nc, c.max_nd, c.sum_nd, c.n, cr, tp)) continue;
                fprintf(stderr, " %s%dx%d/%d%s crit %.2f tp %.2f ms", i == c.cur ? "*" : "", dp_wave_shape_waves(i),
                        dp_wave_shape_lanes(i), dp_wave_shape_ring(i), dp_wave_shape_dense(i) ? "d" : "", cr * 1e-6,
                        tp * 1e-6);
            }
            fprintf(stderr, "\n");
        }
    }
    if (std::getenv("HALIGN_TRACE") && !bc.empty())
        fprintf(stderr, "[halign plan] estimate %.2f ms\n", est * 1e-6);
}

// Positions 0..n-1 ordered by falling key, equal keys in position order: a
// counting sort when the keys span little more than n values (lengths), else
// one sort of packed words.
static void falling_order(const std::vector<uint32_t> &key, std::vector<uint32_t> &pos)
{
    const size_t n = key.size();
    pos.resize(n);
    if (n == 0) return;
    const uint32_t mx = *std::max_element(key.begin(), key.end());
    const uint32_t mn = *std::min_element(key.begin(), key.end());
    const size_t span = (size_t)(mx - mn) + 1;
    if (span <= 4 * n + 65536) {
        std::vector<uint32_t> cnt(span + 1, 0);
        for (uint32_t k : key) ++cnt[mx - k + 1];
        for (size_t v = 1; v <= span; ++v) cnt[v] += cnt[v - 1];
        for (size_t i = 0; i < n; ++i) pos[cnt[mx - key[i]]++] = (uint32_t)i;
        return;
    }
    std::vector<uint64_t> pk(n);
    for (size_t i = 0; i < n; ++i) pk[i] = (uint64_t)(mx - key[i]) << 32 | (uint32_t)i;
    std::sort(pk.begin(), pk.end());
    for (size_t i = 0; i < n; ++i) pos[i] = (uint32_t)pk[i];
}

// Uniform tickets: a sub-wave bucket whose jobs share (qlen, tlen, band) in
// groups can run on the uniform form of its shape: the bucket becomes its keys
// in falling anti-diagonal count, each padded with repeats of its last job to
// a whole number of tickets (jobs per wave). The uniform form keeps the band
// bookkeeping in scalars. The bucket takes that form when the measured table
// puts it (padding included) below the per-group form (without table entries:
// when padding adds at most a quarter more jobs); a uniform shape forced by
// the kernel policy always does.
static void form_uniform_tickets(const halign_ctx *ctx, std::vector<Spec> &specs,
                                 std::vector<std::vector<int>> &buckets, bool keep)
{
    const int nb = (int)buckets.size();
    std::vector<std::vector<int>> moved(nb);
    for (int bi = 0; bi < nb; ++bi) {
        auto &b = buckets[bi];
        const int shape = bi / 4;
        if (b.empty() || shape < kWaveBase) continue;
        const int wi = shape - kWaveBase;
        const int G = dp_wave_jobs_per_group(wi);
        const int ui = dp_wave_uniform_of(wi);
        const bool forced = dp_wave_shape_uniform(wi) != 0;   // a uniform shape must get its tickets
        if (G <= 1 || (ui == wi && !forced) || (keep && !forced)) continue;
        // keys (qlen + tlen, qlen, tlen, band), falling; equal keys keep the
        // bucket's order: a counting sort on qlen + tlen, then each run of
        // equal qlen + tlen (already in bucket order) sorted stably by the rest
        std::vector<uint32_t> k1(b.size()), pos;
        for (size_t k = 0; k < b.size(); ++k) k1[k] = (uint32_t)(specs[b[k]].qlen + specs[b[k]].tlen);
        falling_order(k1, pos);
        struct Key { uint64_t a, b; int idx; };
        std::vector<Key> keys(b.size());
        for (size_t k = 0; k < b.size(); ++k) {
            const Spec &x = specs[b[pos[k]]];
            keys[k] = {(uint64_t)(uint32_t)(x.qlen + x.tlen) << 32 | (uint32_t)x.qlen,
                       (uint64_t)(uint32_t)x.tlen << 32 | (uint32_t)eff_w(x.w, x.qlen, x.tlen), b[pos[k]]};
        }
        for (size_t a = 0; a < keys.size();) {
            size_t e = a + 1;
            while (e < keys.size() && keys[e].a >> 32 == keys[a].a >> 32) ++e;
            if (e - a > 1)
                std::stable_sort(keys.begin() + a, keys.begin() + e, [](const Key &x, const Key &y) {
                    return x.a != y.a ? x.a > y.a : x.b > y.b;
                });
            a = e;
        }
        std::vector<int> srt(b.size());
        for (size_t k = 0; k < b.size(); ++k) srt[k] = keys[k].idx;
        std::vector<int> padded;
        padded.reserve(srt.size() + srt.size() / 2);
        for (size_t a = 0; a < srt.size();) {
            size_t e = a;
            while (e < srt.size() && keys[e].a == keys[a].a && keys[e].b == keys[a].b) ++e;
            padded.insert(padded.end(), srt.begin() + a, srt.begin() + e);
            for (size_t c = e - a; c % G; ++c) padded.push_back(srt[e - 1]);
            a = e;
        }
        if (!forced) {
            double max_nd = 0, sum_nd = 0, sum_ndp = 0;
            int max_nc = 0;
            for (int i : b) {
                const Spec &x = specs[i];
                max_nd = std::max(max_nd, (double)x.qlen + x.tlen - 1);
                max_nc = std::max(max_nc, n_col(x.qlen, x.tlen, eff_w(x.w, x.qlen, x.tlen)));
                sum_nd += (double)x.qlen + x.tlen - 1;
            }
            for (int i : padded) sum_ndp += (double)specs[i].qlen + specs[i].tlen - 1;
            double cg, tg, cu, tu;
            if (tune_cost(ctx, wi, max_nc, max_nd, sum_nd, (double)b.size(), cg, tg) &&
                tune_cost(ctx, ui, max_nc, max_nd, sum_ndp, (double)padded.size(), cu, tu)) {
                if (std::max(cu, tu) >= std::max(cg, tg)) continue;
            } else if (padded.size() * 4 > b.size() * 5) {
                continue;
            }
        }
        for (int i : b) specs[i].shape = kWaveBase + ui;
        auto &dst = moved[(kWaveBase + ui) * 4 + (bi & 3)];
        dst.insert(dst.end(), padded.begin(), padded.end());
        b.clear();
    }
    for (int bi = 0; bi < nb; ++bi)
        if (!moved[bi].empty()) buckets[bi].insert(buckets[bi].end(), moved[bi].begin(), moved[bi].end());
}

// Run a list of jobs whose sequences already sit in ctx->seq (spread u16) and
// whose CIGARs go to ctx->cig. outs receives one DpOut per job, host side.
int run_specs(halign_ctx *ctx, std::vector<Spec> &specs, bool plain, std::vector<DpOut> &outs,
              double *kernel_ms)
{
    const size_t n = specs.size();
    outs.assign(n, DpOut{});
    if (kernel_ms) *kernel_ms = 0.0;
    if (n == 0) return HALIGN_OK;
    PhaseClock clk("run_specs");
    const uint16_t *seq = ctx->seq.as<uint16_t>();
    const uint8_t *seqb = ctx->seqb.as<uint8_t>();
    uint32_t *cig = ctx->cig.as<uint32_t>();

    // bucket by (shape, stats, gap direction)
    const int ns = kNumShapeIds;
    if (ns - kWaveBase < dp_wave_num_shapes() + dp_wave_num_qfix())
        return fail(ctx, HALIGN_E_KERNEL, "shape id space smaller than the shape table");
    const int nb = 4 * ns;
    std::vector<std::vector<int>> buckets(nb);
    for (size_t i = 0; i < n; ++i)
        buckets[specs[i].shape * 4 + (specs[i].stats ? 2 : 0) + ((specs[i].flags & kFlagRight) ? 1 : 0)]
            .push_back((int)i);
    clk.mark("bucketing");
    if (ctx->policy == 0 && !plain) {
        std::vector<int> to_shape(nb);
        plan_bucket_shapes(ctx, specs, buckets, to_shape);
        std::vector<std::vector<int>> moved(nb);
        for (int bi = 0; bi < nb; ++bi) {
            auto &b = buckets[bi];
            if (b.empty()) continue;
            const int to = to_shape[bi];
            for (int i : b) {
                specs[i].shape = to;
                // a query-fixed shape's rows hold all its blocks
                if (to >= kWaveBase && dp_wave_shape_qfix(to - kWaveBase))
                    specs[i].tb_bytes = (size_t)(specs[i].qlen + specs[i].tlen - 1) *
                                        (size_t)dp_wave_shape_ring(to - kWaveBase) * 8;
            }
            auto &dst = moved[to * 4 + (bi & 3)];
            dst.insert(dst.end(), b.begin(), b.end());
        }
        buckets.swap(moved);
    }

    clk.mark("shape plan");
    if (!plain) form_uniform_tickets(ctx, specs, buckets, ctx->tuning);
    clk.mark("uniform tickets");
    std::vector<int> order;
    order.reserve(n);
    for (int bi = 0; bi < nb; ++bi) {
        auto &b = buckets[bi];
        // row shapes run several jobs in one wave: neighbours of similar length
        // (longest first) keep the rows of a wave busy for the same time
        // (uniform tickets are already ordered and padded)
        if (!b.empty() && bi / 4 >= kWaveBase && dp_wave_jobs_per_group(bi / 4 - kWaveBase) > 1 &&
            !dp_wave_shape_uniform(bi / 4 - kWaveBase)) {
            // falling qlen + tlen, equal lengths in bucket order
            std::vector<uint32_t> key(b.size()), pos;
            for (size_t k = 0; k < b.size(); ++k) key[k] = (uint32_t)(specs[b[k]].qlen + specs[b[k]].tlen);
            falling_order(key, pos);
            std::vector<int> srt(b.size());
            for (size_t k = 0; k < b.size(); ++k) srt[k] = b[pos[k]];
            b.swap(srt);
        }
        order.insert(order.end(), b.begin(), b.end());
    }
    clk.mark("job order");
    // with uniform tickets a job may appear more than once (its repeats
    // compute and write the same results)
    const size_t nj = order.size();
    HCHECK(ctx->jobs.ensure(nj * sizeof(DpJob)), "alloc jobs");
    HCHECK(ctx->outs.ensure(nj * sizeof(DpOut)), "alloc outs");
    HCHECK(ctx->counters.ensure((size_t)(plain ? nj : nb) * sizeof(int)), "alloc counters");
    if (ctx->hjobs.size() < nj) ctx->hjobs.resize(nj);   // kept: no zero fill of a fresh array per batch
    DpJob *hjobs = ctx->hjobs.data();
    for (size_t k = 0; k < nj; ++k) {
        const Spec &s = specs[order[k]];
        DpJob &j = hjobs[k];
        j.tgt_even = seq + s.tgt_elem;
        j.tgt_odd = seq + s.tgt_odd_elem;
        j.qr_even = seq + s.qry_elem;
        j.qr_odd = seq + s.qry_odd_elem;
        j.tgt_b = seqb + s.tgt_b;
        j.qr_b = seqb + s.qry_b;
        j.tlen = s.tlen; j.qlen = s.qlen; j.w = s.w; j.flags = s.flags;
        j.cigar = cig + s.cig_off;
        j.cig_cap = s.cig_cap;
        j.pad_ = 0;
    }
    clk.mark("jobs");
    HCHECK(hipMemcpyAsync(ctx->jobs.p, hjobs, nj * sizeof(DpJob), hipMemcpyHostToDevice, ctx->stream),
           "upload jobs");
    HCHECK(hipMemsetAsync(ctx->counters.p, 0, ctx->counters.cap, ctx->stream), "clear counters");
    clk.mark("job upload");

    // traceback pool: one slot per resident workgroup (per job group of a
    // sub-wave workgroup)
    size_t need_slot = 0;
    for (auto &s : specs) need_slot = std::max(need_slot, s.tb_bytes);
    if (need_slot > ctx->tb_budget) return fail(ctx, HALIGN_E_NOMEM, "traceback of one job exceeds the device budget");
    // the pool is planned within the budget and within what the device can
    // still give it (the pool it holds plus the free HBM, less 1 GiB for the
    // rest of the call), so contexts that share a device shrink their launches
    // instead of failing an allocation
    size_t budget = ctx->tb_budget;
    {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) == hipSuccess) {
            const size_t keep = (size_t)1 << 30;
            budget = std::min(budget, ctx->tb.cap + (fr > keep ? fr - keep : 0));
            budget = std::max(budget, need_slot);
        }
    }

    HCHECK(hipEventRecord(ctx->ev0, ctx->stream), "event");
    // launches bracketed by bucket-timing events, in event-pair order
    struct Timed { int family, shape, stream, right; size_t pos, njobs; double est; };
    std::vector<Timed> timed;
    std::vector<std::pair<size_t, size_t>> qranges;   // job ranges of the combined query-fixed launch
    DpJob *djobs = ctx->jobs.as<DpJob>();
    DpOut *douts = ctx->outs.as<DpOut>();
    int *ctr = ctx->counters.as<int>();
    if (plain) {
        size_t pos = 0;
        for (int bi = 0; bi < nb; ++bi) {
            auto &b = buckets[bi];
            if (b.empty()) continue;
            const int shape = bi / 4;
            const bool stats = (bi & 2) != 0;
            const bool rt = (bi & 1) != 0;
            size_t slot = 0;
            for (int i : b) slot = std::max(slot, specs[i].tb_bytes);
            slot = (slot + 255) & ~(size_t)255;
            const bool wave = shape >= kWaveBase;
            const int wi = shape - kWaveBase;
            const size_t jpg = wave ? (size_t)dp_wave_jobs_per_group(wi) : 1;
            HCHECK(ctx->tb.ensure(slot * jpg), "alloc traceback");
            for (size_t k = 0; k < b.size(); ++k) {
                hipError_t e = wave ? dp_wave_launch(wi, rt, 1, djobs + pos + k, 1, ctr + pos + k,
                                                     ctx->tb.as<uint8_t>(), slot, douts + pos + k, ctx->stream)
                                    : dp_launch(shape, stats, 1, djobs + pos + k, 1, ctr + pos + k,
                                                ctx->tb.as<uint8_t>(), slot, douts + pos + k, ctx->stream);
                HCHECK(e, "launch dp (plain)");
            }
            pos += b.size();
        }
    } else {
        // One launch per bucket. Buckets are independent (own jobs, outputs,
        // counter and traceback region), so they run side by side on the main
        // stream and the three side streams (one hardware queue each: the
        // runtime's default is four per process), longest estimate first on
        // the least loaded stream: narrow-band buckets are bound by their
        // jobs' anti-diagonal chains, not by the chip, and overlap well.
        // wi: the ring shape launched (the bucket's, or the one-per-SIMD form
        // of a two-per-SIMD shape the budget cannot fill)
        struct Launch { int bi; size_t pos, slot, jpg, off; int groups; double est; int stream; bool prio; int wi; };
        std::vector<QfixPart> qparts;   // parts of the combined query-fixed launch (bi = -1)
        std::vector<Launch> ls;
        size_t pos = 0;
        for (int bi = 0; bi < nb; ++bi) {
            auto &b = buckets[bi];
            if (b.empty()) continue;
            const int shape = bi / 4;
            const bool stats = (bi & 2) != 0;
            const bool rt = (bi & 1) != 0;
            Launch L{bi, pos, 0, 1, 0, 0, 0.0, 0, false, shape - kWaveBase};
            for (int i : b) L.slot = std::max(L.slot, specs[i].tb_bytes);
            L.slot = round_slot(L.slot);
            const bool wave = shape >= kWaveBase;
            const int wi = shape - kWaveBase;
            L.jpg = wave ? (size_t)dp_wave_jobs_per_group(wi) : 1;
            int groups = wave ? dp_wave_max_groups(wi, rt, ctx->num_cu) : dp_shape_max_groups(shape, stats, ctx->num_cu);
            if (groups <= 0) return fail(ctx, HALIGN_E_HIP, "occupancy query failed");
            L.groups = (int)std::min<size_t>(groups, (b.size() + L.jpg - 1) / L.jpg);
            // estimated time (ns): the measured table for ring shapes; for the
            // multi-wave kernels, cells at about 650 cells per ns over the chip
            // (a query-fixed shape's table is by query blocks)
            const bool qf = wave && dp_wave_shape_qfix(wi);
            double max_nd = 0, sum_nd = 0, blocks = 0;
            int max_nc = 0;
            for (int i : b) {
                const Spec &s = specs[i];
                const double nd = (double)s.qlen + s.tlen - 1;
                const int nc = n_col(s.qlen, s.tlen, eff_w(s.w, s.qlen, s.tlen));
                max_nd = std::max(max_nd, nd);
                max_nc = std::max(max_nc, qf ? (s.qlen + 15) / 16 : nc);
                sum_nd += nd;
                blocks += nd * nc;
            }
            double crit = 0, tp = 0;
            const bool tab = wave && tune_cost(ctx, wi, max_nc, max_nd, sum_nd, (double)b.size(), crit, tp);
            L.est = tab ? std::max(crit, tp) : blocks * 16.0 / 650.0;
            // bound by its longest job's anti-diagonal chain rather than by the
            // chip: its waves issue first on the SIMDs they share with other
            // launches' waves (which fill the issue slots they leave)
            L.prio = tab && crit > tp;
            ls.push_back(L);
            pos += b.size();
        }
        // traceback regions: each launch its own; shrink the widest launches to the budget
        auto pool_bytes = [&]() {
            size_t t = 0;
            for (auto &L : ls) t += L.slot * L.jpg * (size_t)L.groups;
            return t;
        };
        size_t min_bytes = 0;
        for (auto &L : ls) min_bytes += L.slot * L.jpg;
        // Tickets (jpg jobs each) are taken one at a time, so n tickets on g
        // groups run in ceil(n / g) rounds; when the traceback budget has cut
        // g below the chip's resident count, the fewest groups with as many
        // rounds leave no partial last round (4,096 config-2 pairs on 1,152
        // groups of a shape that fits two per SIMD: 1,024 groups, one per
        // SIMD, four rounds).
        // A two-per-SIMD launch keeps every SIMD's pair (its rounds are not
        // rebalanced); under the automatic policy one the budget has cut below
        // that runs in the one-per-SIMD form.
        auto balance = [&](Launch &L) {
            const size_t tickets = (buckets[L.bi].size() + L.jpg - 1) / L.jpg;
            if (L.groups <= 0 || tickets == 0) return;
            const bool wv = L.bi / 4 >= kWaveBase;
            if (wv && dp_wave_shape_dense(L.wi)) {
                const int full = dp_wave_max_groups(L.wi, (L.bi & 1) != 0, ctx->num_cu);
                if (ctx->policy != 0 || L.groups >= full || (size_t)L.groups >= tickets) return;
                L.wi = dp_wave_sparse_of(L.wi);
                L.groups = std::min(L.groups, dp_wave_max_groups(L.wi, (L.bi & 1) != 0, ctx->num_cu));
            }
            const size_t rounds = (tickets + (size_t)L.groups - 1) / (size_t)L.groups;
            L.groups = (int)((tickets + rounds - 1) / rounds);
        };
        const bool side = min_bytes <= budget;   // else one launch at a time in one region
        if (side) {
            while (pool_bytes() > budget) {
                Launch *big = nullptr;
                for (auto &L : ls)
                    if (L.groups > 1 && (!big || L.slot * L.jpg * L.groups > big->slot * big->jpg * big->groups))
                        big = &L;
                if (!big) break;
                big->groups = std::max(1, (int)(big->groups * 0.75));
            }
            for (auto &L : ls) balance(L);
            size_t off = 0;
            for (auto &L : ls) {
                L.off = off;
                off += L.slot * L.jpg * (size_t)L.groups;
            }
            clk.mark("launch plan");
            HCHECK(ctx->tb.ensure(std::max<size_t>(off, 1), budget), "alloc traceback pool");
            clk.mark("traceback pool");
            // Query-fixed launches bound by their jobs' chains run as the parts
            // of one combined launch (dp_qfix_combined): side by side rather
            // than one after another on the four streams. Each part keeps its
            // jobs, counter and traceback region; the combined estimate is its
            // longest part's.
            if (!plain) {
                std::vector<Launch> keep, merged;
                for (auto &L : ls) {
                    const int sh = L.bi / 4;
                    if (sh >= kWaveBase && dp_wave_shape_qfix(sh - kWaveBase) && L.prio && !(L.bi & 2))
                        merged.push_back(L);
                    else
                        keep.push_back(L);
                }
                if (merged.size() >= 2) {
                    Launch C{-1, 0, 0, 1, 0, 0, 0.0, 0, true};
                    int first = 0;
                    for (auto &L : merged) {
                        const int wi = L.bi / 4 - kWaveBase;
                        qparts.push_back({djobs + L.pos, douts + L.pos, ctr + L.bi, ctx->tb.as<uint8_t>() + L.off,
                                          (uint64_t)L.slot, (int32_t)buckets[L.bi].size(),
                                          2 * (wi - dp_wave_qfix_index(0)) + (L.bi & 1), first, 0});
                        first += L.groups;
                        C.est = std::max(C.est, L.est);
                    }
                    C.groups = first;
                    keep.push_back(C);
                    for (auto &L : merged) qranges.push_back({L.pos, buckets[L.bi].size()});
                    ls.swap(keep);
                    HCHECK(ctx->qparts.ensure(qparts.size() * sizeof(QfixPart)), "alloc combined parts");
                    HCHECK(hipMemcpyAsync(ctx->qparts.p, qparts.data(), qparts.size() * sizeof(QfixPart),
                                          hipMemcpyHostToDevice, ctx->stream),
                           "upload combined parts");
                }
            }
            std::vector<int> ord(ls.size());
            for (size_t k = 0; k < ls.size(); ++k) ord[k] = (int)k;
            std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return ls[x].est > ls[y].est; });
            double load[kDpStreams] = {};
            for (int k : ord) {
                int sidx = 0;
                for (int q = 1; q < kDpStreams; ++q)
                    if (load[q] < load[sidx]) sidx = q;
                ls[k].stream = sidx;
                load[sidx] += ls[k].est;
            }
        } else {
            for (auto &L : ls) {
                L.groups = (int)std::min<size_t>(L.groups, std::max<size_t>(1, budget / (L.slot * L.jpg)));
                balance(L);
                L.off = 0;
                L.stream = 0;
            }
            size_t mx = 0;
            for (auto &L : ls) mx = std::max(mx, L.slot * L.jpg * (size_t)L.groups);
            HCHECK(ctx->tb.ensure(std::max<size_t>(mx, 1), budget), "alloc traceback pool");
        }
        bool used[kDpStreams] = {true};
        for (auto &L : ls) used[L.stream] = true;
        // the plan's signature: a repeat replays the captured graph
        std::vector<int64_t> sig;
        for (auto &L : ls) {
            if (L.bi < 0) {   // the combined launch: its parts table and where it lives
                const int64_t v[] = {-1, L.groups, (int64_t)ctx->qparts.p, (int64_t)qparts.size(), L.stream};
                sig.insert(sig.end(), v, v + 5);
                for (const QfixPart &q : qparts) {
                    const int64_t w[] = {(int64_t)q.jobs, (int64_t)q.outs, (int64_t)q.counter, (int64_t)q.tb_pool,
                                         (int64_t)q.tb_slot_bytes, q.njobs, q.shape, q.first};
                    sig.insert(sig.end(), w, w + 8);
                }
                continue;
            }
            const int64_t v[] = {L.bi, L.groups, (int64_t)(djobs + L.pos), (int64_t)buckets[L.bi].size(),
                                 (int64_t)(ctr + L.bi), (int64_t)(ctx->tb.as<uint8_t>() + L.off), (int64_t)L.slot,
                                 (int64_t)(douts + L.pos), L.stream, L.prio, L.wi};
            sig.insert(sig.end(), v, v + 11);
        }
        const bool use_graphs = ctx->graphs && !ctx->bucket_timing;
        int hit = -1;
        for (size_t k = 0; use_graphs && k < ctx->gcache.size(); ++k)
            if (ctx->gcache[k].sig == sig) hit = (int)k;
        if (hit >= 0) {
            auto cg = ctx->gcache[hit];
            ctx->gcache.erase(ctx->gcache.begin() + hit);
            ctx->gcache.push_back(cg);
            HCHECK(hipGraphLaunch(cg.exec, ctx->stream), "graph launch");
            ++ctx->graph_replays;
            ls.clear();   // nothing to launch directly
        } else if (use_graphs) {
            // relaxed mode: other threads may allocate while this stream captures
            HCHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed), "begin capture");
        }
        const bool capturing = use_graphs && !ls.empty();
        auto end_capture = [&]() {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(ctx->stream, &g);
            if (g) (void)hipGraphDestroy(g);
        };
#define HCHECK_CAP(expr, what)                                              \
    do {                                                                    \
        hipError_t e_ = (expr);                                             \
        if (e_ != hipSuccess) {                                             \
            if (capturing) end_capture();                                   \
            return fail(ctx, HALIGN_E_HIP, what, e_);                        \
        }                                                                   \
    } while (0)
        bool any_side = false;
        for (int q = 1; q < kDpStreams; ++q) any_side |= used[q];
        if (!ls.empty() && any_side) HCHECK_CAP(hipEventRecord(ctx->fork, ctx->stream), "event");
        for (int q = 1; q < kDpStreams && !ls.empty(); ++q)
            if (used[q]) HCHECK_CAP(hipStreamWaitEvent(ctx->aux[q - 1], ctx->fork, 0), "stream wait");
        for (auto &L : ls) {
            if (L.bi < 0) {   // the combined query-fixed launch
                hipStream_t st = L.stream ? ctx->aux[L.stream - 1] : ctx->stream;
                hipEvent_t ta = nullptr, tz = nullptr;
                if (ctx->bucket_timing) {
                    if (!bucket_events(ctx, timed.size(), ta, tz)) return fail(ctx, HALIGN_E_HIP, "bucket events");
                    HCHECK(hipEventRecord(ta, st), "event");
                    size_t nj = 0;
                    for (const QfixPart &q : qparts) nj += (size_t)q.njobs;
                    timed.push_back({3, -1, L.stream, -1, 0, nj, L.est});
                }
                HCHECK_CAP(dp_qfix_combined_launch(ctx->qparts.as<QfixPart>(), (int)qparts.size(), L.groups, st),
                           "launch dp (combined)");
                if (tz) HCHECK(hipEventRecord(tz, st), "event");
                continue;
            }
            const int shape = L.bi / 4;
            const bool stats = (L.bi & 2) != 0;
            const bool rt = (L.bi & 1) != 0;
            const bool wave = shape >= kWaveBase;
            const int wi = L.wi;
            if (clk.on)
                fprintf(stderr, "[halign run_specs] bucket: %s %d x%d lanes/job %d ring %d, %zu jobs, %d groups, stream %d, est %.3g ms\n",
                        wave ? (dp_wave_shape_uniform(wi) ? "ring(uniform)" : dp_wave_shape_dense(wi) ? "ring(dense)" : "ring")
                             : "multiwave",
                        wave ? dp_wave_shape_waves(wi) : shape,
                        wave ? dp_wave_jobs_per_group(wi) : 1, wave ? dp_wave_shape_lanes(wi) : 0,
                        wave ? dp_wave_shape_ring(wi) : 0, buckets[L.bi].size(), L.groups, L.stream, L.est * 1e-6);
            hipStream_t st = L.stream ? ctx->aux[L.stream - 1] : ctx->stream;
            uint8_t *tb = ctx->tb.as<uint8_t>() + L.off;
            const int nj = (int)buckets[L.bi].size();
            hipEvent_t ta = nullptr, tz = nullptr;
            if (ctx->bucket_timing) {
                if (!bucket_events(ctx, timed.size(), ta, tz)) return fail(ctx, HALIGN_E_HIP, "bucket events");
                HCHECK(hipEventRecord(ta, st), "event");
                timed.push_back({wave ? 0 : 2, wave ? wi : shape, L.stream, rt ? 1 : 0, L.pos, (size_t)nj, L.est});
            }
            hipError_t e = wave ? dp_wave_launch(wi, rt, L.groups, djobs + L.pos, nj, ctr + L.bi, tb, L.slot,
                                                 douts + L.pos, st, nullptr, L.prio)
                                : dp_launch(shape, stats, L.groups, djobs + L.pos, nj, ctr + L.bi, tb, L.slot,
                                            douts + L.pos, st);
            HCHECK_CAP(e, "launch dp");
            if (tz) HCHECK(hipEventRecord(tz, st), "event");
        }
        for (int q = 1; q < kDpStreams && !ls.empty(); ++q) {
            if (!used[q]) continue;
            HCHECK_CAP(hipEventRecord(ctx->join[q - 1], ctx->aux[q - 1]), "event");
            HCHECK_CAP(hipStreamWaitEvent(ctx->stream, ctx->join[q - 1], 0), "stream wait");
        }
#undef HCHECK_CAP
        if (capturing) {
            hipGraph_t g = nullptr;
            HCHECK(hipStreamEndCapture(ctx->stream, &g), "end capture");
            hipGraphExec_t ge = nullptr;
            const hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            HCHECK(ie, "instantiate graph");
            if (ctx->gcache.size() >= halign_ctx::kGraphCache) {
                (void)hipGraphExecDestroy(ctx->gcache.front().exec);
                ctx->gcache.erase(ctx->gcache.begin());
            }
            ctx->gcache.push_back({sig, ge});
            ++ctx->graph_captures;
            HCHECK(hipGraphLaunch(ge, ctx->stream), "graph launch");
        }
    }
    clk.mark("launches");
    HCHECK(hipEventRecord(ctx->ev1, ctx->stream), "event");
    if (!ctx->out_pin.ensure(std::max<size_t>(nj, 1) * sizeof(DpOut))) return fail(ctx, HALIGN_E_NOMEM, "pinned buffer");
    const DpOut *tmp = reinterpret_cast<const DpOut *>(ctx->out_pin.p);
    HCHECK(hipMemcpyAsync(ctx->out_pin.p, ctx->outs.p, nj * sizeof(DpOut), hipMemcpyDeviceToHost, ctx->stream),
           "download outs");
    HCHECK(hipStreamSynchronize(ctx->stream), "dp kernels");
    clk.mark("kernels");
    float ms = 0.f;
    HCHECK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1), "event time");
    if (kernel_ms) *kernel_ms = ms;
    for (size_t k = 0; k < nj; ++k) outs[order[k]] = tmp[k];
    for (size_t k = 0; k < timed.size(); ++k) {
        const Timed &t = timed[k];
        float bms = 0.f;
        HCHECK(hipEventElapsedTime(&bms, ctx->bev[2 * k], ctx->bev[2 * k + 1]), "bucket event time");
        halign_bucket_stat b{t.family, t.shape, t.stream, t.right, (int64_t)t.njobs, 0, bms, t.est * 1e-6};
        if (t.family == 3) {   // the combined launch: the cells of every part
            for (const auto &q : qranges)
                for (size_t j = q.first; j < q.first + q.second; ++j) b.cells += tmp[j].cells;
        } else {
            for (size_t j = t.pos; j < t.pos + t.njobs; ++j) b.cells += tmp[j].cells;
        }
        ctx->last_buckets.push_back(b);
    }
    for (size_t i = 0; i < n; ++i) {
        const int st = outs[i].status;
        if (st != kStOk) {
            char msg[128];
            snprintf(msg, sizeof msg, "dp job %zu returned status %d", i, st);
            return fail(ctx, HALIGN_E_KERNEL, msg);
        }
    }
    return HALIGN_OK;
}

bool fill_spec_shape(Spec &s, bool full_stats, int policy)
{
    const int w = eff_w(s.w, s.qlen, s.tlen);
    const int nc = n_col(s.qlen, s.tlen, w);
    s.stats = full_stats || predict_zdrop(s.qlen, s.tlen, w);
    s.shape = pick_shape(nc, s.stats, policy, s.qlen, s.tlen, w);
    if (s.shape == dp_num_shapes() - 1) s.stats = true;   // the wide-band kernel computes every field
    // query-fixed shapes keep every block of the shape in each anti-diagonal's row
    const int rowblk = s.shape >= kWaveBase && dp_wave_shape_qfix(s.shape - kWaveBase)
                           ? dp_wave_shape_ring(s.shape - kWaveBase)
                           : nc;
    s.tb_bytes = s.shape >= 0 && s.shape < kWaveBase ? dp_shape_slot_bytes(s.shape, s.qlen, s.tlen, nc)
                                                     : (size_t)(s.qlen + s.tlen - 1) * (size_t)rowblk * 8;
    return s.shape >= 0;
}

}  // namespace hal

void hal::ctx_destroy(halign_ctx *ctx)
{
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->seq.release(); ctx->seqb.release(); ctx->jobs.release(); ctx->outs.release();
    ctx->counters.release(); ctx->tb.release(); ctx->cig.release(); ctx->qparts.release();
    ctx->crows.release(); ctx->ccounts.release(); ctx->ccons.release();
    for (hal::DevBuf *b : {&ctx->s_seqs, &ctx->s_off, &ctx->s_woff, &ctx->s_moff, &ctx->s_codes, &ctx->s_nmask,
                           &ctx->s_hoff, &ctx->s_hx, &ctx->s_hy, &ctx->s_cnt, &ctx->s_st, &ctx->s_list,
                           &ctx->m_seqs, &ctx->m_off, &ctx->m_koff, &ctx->m_keys, &ctx->m_keys2, &ctx->m_nkeys,
                           &ctx->m_tmp, &ctx->m_sel})
        b->release();
    for (auto &cg : ctx->gcache) (void)hipGraphExecDestroy(cg.exec);
    for (hipEvent_t e : ctx->bev) (void)hipEventDestroy(e);
    ctx->bev.clear();
    for (auto &b : ctx->msa_out) b.release();
    ctx->msa_tmp.release();
    for (int k = 0; k < 2; ++k) {
        if (ctx->up_ev[k]) {
            (void)hipEventSynchronize(ctx->up_ev[k]);
            (void)hipEventDestroy(ctx->up_ev[k]);
        }
        ctx->up_pin[k].release();
    }
    ctx->dn_pin.release();
    ctx->out_pin.release();
    (void)hipEventDestroy(ctx->ev0);
    (void)hipEventDestroy(ctx->ev1);
    for (int k = 0; k < halign_ctx::kAux; ++k) {
        (void)hipStreamSynchronize(ctx->aux[k]);
        (void)hipStreamDestroy(ctx->aux[k]);
        (void)hipEventDestroy(ctx->join[k]);
    }
    (void)hipEventDestroy(ctx->fork);
    (void)hipEventDestroy(ctx->pfork);
    (void)hipEventDestroy(ctx->pjoin);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

void hal::ctx_detach(halign_ctx *ctx)
{
    bool last = false;
    {
        std::lock_guard<std::mutex> lock(ctx->mu);
        --ctx->children;
        last = ctx->closing && ctx->children == 0;
    }
    if (last) ctx_destroy(ctx);
}

extern "C" {

const char *halign_version(void) { return "halign-mi355x 0.1 (gfx950)"; }

int halign_init(int device, halign_ctx **out)
{
    if (!out) return HALIGN_E_ARG;
    *out = nullptr;
    halign_ctx *ctx = new (std::nothrow) halign_ctx();
    if (!ctx) return HALIGN_E_NOMEM;
    ctx->device = device;
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) { delete ctx; return HALIGN_E_HIP; }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) { delete ctx; return HALIGN_E_HIP; }
    ctx->num_cu = prop.multiProcessorCount;
    bool ok = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreate(&ctx->ev0) == hipSuccess && hipEventCreate(&ctx->ev1) == hipSuccess &&
              hipEventCreateWithFlags(&ctx->fork, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&ctx->pfork, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&ctx->pjoin, hipEventDisableTiming) == hipSuccess;
    for (int k = 0; ok && k < halign_ctx::kAux; ++k)
        ok = hipStreamCreateWithFlags(&ctx->aux[k], hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&ctx->join[k], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        delete ctx;
        return HALIGN_E_HIP;
    }
    // this device's copy of the no-allocation scratch array (device current)
    if (hipGetSymbolAddress(&ctx->noalloc_scratch, HIP_SYMBOL(scratch)) != hipSuccess || !ctx->noalloc_scratch) {
        ctx_destroy(ctx);
        return HALIGN_E_HIP;
    }
    size_t free_b = 0, total_b = 0;
    (void)hipMemGetInfo(&free_b, &total_b);
    // Traceback pool: one slot per resident workgroup. A pair of ~30 kb at
    // band 4,665 needs a 140 MB slot, and the two-per-SIMD form of its shape
    // keeps 2,048 workgroups resident (287 GB), so the pool may take all of
    // the free HBM but a reserve of 4% (at least 8 GiB) for sequences, jobs,
    // CIGARs and the per-query path's buffers. It is allocated as launches
    // need it, not up front.
    {
        const size_t reserve = std::max<size_t>(free_b / 25, (size_t)8 << 30);
        ctx->tb_budget = free_b > 2 * reserve ? free_b - reserve : free_b / 2;
    }
    tune_load_default(ctx);
    *out = ctx;
    return HALIGN_OK;
}

void halign_free(halign_ctx *ctx)
{
    if (!ctx) return;
    if (ctx->msa_center) {   // the keep-length run's cached center is the context's own
        halign_center *mc = ctx->msa_center;
        ctx->msa_center = nullptr;
        halign_center_free(mc);
    }
    bool now = false;
    {
        std::lock_guard<std::mutex> lock(ctx->mu);
        if (ctx->closing) return;   // freed twice while children were live
        ctx->closing = true;
        now = ctx->children == 0;
    }
    if (now) ctx_destroy(ctx);
}

const char *halign_last_error(const halign_ctx *ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int halign_band_info(int32_t qlen, int32_t tlen, int32_t w, int32_t *ncol, int32_t *overflow, int64_t *cells)
{
    if (qlen <= 0 || tlen <= 0) return HALIGN_E_ARG;
    const int we = eff_w(w, qlen, tlen);
    if (ncol) *ncol = n_col(qlen, tlen, we);
    if (overflow) *overflow = predict_zdrop(qlen, tlen, we) ? 1 : 0;
    if (cells) {
        // sum of en0 - st0 + 1 over the anti-diagonals computed before any overflow
        // (src/align/ksw2_extz2_sse.c:107-114)
        int64_t c = 0;
        for (int r = 0; r < qlen + tlen - 1; ++r) {
            const int st0 = std::max(std::max(0, r - qlen + 1), (r - we + 1) >> 1);
            const int en0 = std::min(std::min(tlen - 1, r), (r + we) >> 1);
            if (st0 > en0) break;
            c += en0 - st0 + 1;
        }
        *cells = c;
    }
    return HALIGN_OK;
}

int halign_ring_shapes(int32_t *waves, int32_t *lanes, int32_t *ring, int32_t cap)
{
    const int n = dp_wave_num_shapes();
    for (int i = 0; i < n && i < cap; ++i) {
        if (waves) waves[i] = dp_wave_shape_waves(i);
        if (lanes) lanes[i] = dp_wave_shape_lanes(i);
        if (ring) ring[i] = dp_wave_shape_ring(i);
    }
    return n;
}

int halign_qfix_shapes(int32_t *lanes, int32_t *slots, int32_t cap)
{
    const int n = dp_wave_num_qfix();
    for (int q = 0; q < n && q < cap; ++q) {
        const int idx = dp_wave_qfix_index(q);
        if (lanes) lanes[q] = dp_wave_shape_lanes(idx);
        if (slots) slots[q] = dp_wave_shape_ring(idx) / dp_wave_shape_lanes(idx);
    }
    return n;
}

int halign_set_graph_mode(halign_ctx *ctx, int32_t on)
{
    if (!ctx) return HALIGN_E_ARG;
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->graphs = on ? 1 : 0;
    return HALIGN_OK;
}

int halign_graph_counts(halign_ctx *ctx, int64_t *captures, int64_t *replays)
{
    if (!ctx) return HALIGN_E_ARG;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (captures) *captures = ctx->graph_captures;
    if (replays) *replays = ctx->graph_replays;
    return HALIGN_OK;
}

int halign_plan_shape(halign_ctx *ctx, int32_t n, int32_t qlen, int32_t tlen, int32_t w, int32_t *ring_shape)
{
    if (!ctx || !ring_shape || n <= 0 || qlen <= 0 || tlen <= 0) return HALIGN_E_ARG;
    std::lock_guard<std::mutex> lock(ctx->mu);
    *ring_shape = -1;
    Spec s{};
    s.qlen = qlen;
    s.tlen = tlen;
    s.w = w;
    if (!fill_spec_shape(s, false, ctx->policy)) return HALIGN_OK;
    if (s.shape < kWaveBase) return HALIGN_OK;
    std::vector<Spec> specs((size_t)n, s);
    std::vector<std::vector<int>> buckets(4 * kNumShapeIds);
    for (int i = 0; i < n; ++i) buckets[(size_t)s.shape * 4].push_back(i);
    std::vector<int> to(buckets.size());
    if (ctx->policy == 0) plan_bucket_shapes(ctx, specs, buckets, to);
    else to[(size_t)s.shape * 4] = s.shape;
    *ring_shape = to[(size_t)s.shape * 4] - kWaveBase;
    return HALIGN_OK;
}

int halign_set_bucket_timing(halign_ctx *ctx, int32_t on)
{
    if (!ctx) return HALIGN_E_ARG;
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->bucket_timing = on != 0;
    ctx->last_buckets.clear();
    return HALIGN_OK;
}

int halign_last_buckets(halign_ctx *ctx, halign_bucket_stat *out, int32_t cap, int32_t *n)
{
    if (!ctx || !n || cap < 0 || (cap > 0 && !out)) return HALIGN_E_ARG;
    std::lock_guard<std::mutex> lock(ctx->mu);
    *n = (int32_t)ctx->last_buckets.size();
    if (!out) return HALIGN_OK;   // size query: the records stay
    for (int32_t k = 0; k < *n && k < cap; ++k) out[k] = ctx->last_buckets[(size_t)k];
    ctx->last_buckets.clear();
    return HALIGN_OK;
}

int halign_set_traceback_budget(halign_ctx *ctx, int64_t bytes)
{
    if (!ctx || bytes <= 0) return HALIGN_E_ARG;
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->tb_budget = (size_t)bytes;
    return HALIGN_OK;
}

int halign_set_kernel_policy(halign_ctx *ctx, int policy)
{
    if (!ctx || policy < 0 || policy > 2 + dp_wave_num_shapes() + dp_wave_num_qfix()) return HALIGN_E_ARG;
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->policy = policy;
    return HALIGN_OK;
}

int halign_dp_batch(halign_ctx *ctx, int32_t n, const halign_dp_task *tasks, const uint8_t *query_codes,
                    const uint8_t *target_codes, int32_t full_stats, int32_t plain,
                    halign_dp_result *results, uint32_t *cigar_out, int64_t cigar_cap, double *kernel_ms)
{
    if (!ctx || n < 0 || (n > 0 && (!tasks || !results))) return HALIGN_E_ARG;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HCHECK(hipSetDevice(ctx->device), "set device");
    // spread sequences: per task, target even/odd then query even/odd
    std::vector<Spec> specs(n);
    int64_t elems = 0, cig_total = 0, nbytes = 0;
    for (int i = 0; i < n; ++i) {
        const halign_dp_task &t = tasks[i];
        if (t.qlen < 0 || t.tlen < 0) return fail(ctx, HALIGN_E_ARG, "negative length");
        if (!(t.flag & 0x04) || (t.flag & ~0x06)) return fail(ctx, HALIGN_E_UNSUPPORTED, "flag must be GENERIC_SC [| RIGHT]");
        const int64_t tn = t.tlen + 2 * kSeqPad, qn = t.qlen + 2 * kSeqPad;
        Spec &s = specs[i];
        s.tgt_elem = elems + kSeqPad; s.tgt_odd_elem = elems + tn + kSeqPad;
        s.qry_elem = elems + 2 * tn + kSeqPad; s.qry_odd_elem = elems + 2 * tn + qn + kSeqPad;
        elems += 2 * tn + 2 * qn;
        elems = (elems + 1) & ~(int64_t)1;   // keep every copy 4-byte aligned
        s.tgt_b = nbytes + kBytePad;
        nbytes += t.tlen + 2 * kBytePad;
        s.qry_b = nbytes + kBytePad;
        nbytes += t.qlen + 2 * kBytePad;
        s.tlen = t.tlen; s.qlen = t.qlen; s.w = t.w;
        s.flags = (t.flag & 0x02) ? kFlagRight : 0;
        for (int k = 0; k < t.qlen; ++k) if (query_codes[t.query_off + k] > 3) { s.flags |= kFlagHasN; break; }
        if (!(s.flags & kFlagHasN))
            for (int k = 0; k < t.tlen; ++k) if (target_codes[t.target_off + k] > 3) { s.flags |= kFlagHasN; break; }
        s.cig_off = cig_total;
        s.cig_cap = t.qlen + t.tlen + 4;
        cig_total += s.cig_cap;
        if (t.qlen > 0 && t.tlen > 0 && !fill_spec_shape(s, full_stats != 0, ctx->policy))
            return fail(ctx, HALIGN_E_UNSUPPORTED, "band wider than the largest ring shape");
    }
    std::vector<uint16_t> host((size_t)elems + 2, 0);
    std::vector<uint8_t> hostb((size_t)nbytes + 16, 0);
    for (int i = 0; i < n; ++i) {
        const halign_dp_task &t = tasks[i];
        const Spec &s = specs[i];
        spread_target(target_codes + t.target_off, t.tlen, host.data() + s.tgt_elem - kSeqPad,
                      host.data() + s.tgt_odd_elem - kSeqPad);
        spread_query_rev(query_codes + t.query_off, t.qlen, host.data() + s.qry_elem - kSeqPad,
                         host.data() + s.qry_odd_elem - kSeqPad);
        bytes_target(target_codes + t.target_off, t.tlen, hostb.data() + s.tgt_b - kBytePad);
        bytes_query_rev(query_codes + t.query_off, t.qlen, hostb.data() + s.qry_b - kBytePad);
    }
    HCHECK(ctx->seq.ensure(host.size() * 2), "alloc sequences");
    HCHECK(ctx->seqb.ensure(hostb.size()), "alloc byte sequences");
    HCHECK(hipMemcpyAsync(ctx->seqb.p, hostb.data(), hostb.size(), hipMemcpyHostToDevice, ctx->stream),
           "upload byte sequences");
    HCHECK(ctx->cig.ensure((size_t)std::max<int64_t>(cig_total, 1) * 4), "alloc cigars");
    HCHECK(hipMemcpyAsync(ctx->seq.p, host.data(), host.size() * 2, hipMemcpyHostToDevice, ctx->stream),
           "upload sequences");
    // empty sequences: ksw2 returns immediately with a reset ez (:56-57)
    std::vector<Spec> run;
    std::vector<int> run_idx;
    for (int i = 0; i < n; ++i)
        if (tasks[i].qlen > 0 && tasks[i].tlen > 0) { run.push_back(specs[i]); run_idx.push_back(i); }
    std::vector<DpOut> outs;
    int rc = run_specs(ctx, run, plain != 0, outs, kernel_ms);
    if (rc != HALIGN_OK) return rc;
    std::vector<uint32_t> hcig((size_t)std::max<int64_t>(cig_total, 1));
    HCHECK(hipMemcpy(hcig.data(), ctx->cig.p, (size_t)cig_total * 4, hipMemcpyDeviceToHost), "download cigars");
    int64_t out_pos = 0;
    for (int i = 0; i < n; ++i) {
        halign_dp_result &r = results[i];
        r.max = 0; r.zdropped = 0; r.max_q = r.max_t = r.mqe_t = r.mte_q = -1;
        r.mqe = r.mte = r.score = -0x40000000; r.n_cigar = 0; r.status = 0; r.pad_ = 0; r.cells = 0;
        r.cigar_off = out_pos;
    }
    for (size_t k = 0; k < run.size(); ++k) {
        const int i = run_idx[k];
        const DpOut &o = outs[k];
        halign_dp_result &r = results[i];
        r.max = o.max; r.zdropped = o.zdropped; r.max_q = o.max_q; r.max_t = o.max_t;
        r.mqe = o.mqe; r.mqe_t = o.mqe_t; r.mte = o.mte; r.mte_q = o.mte_q; r.score = o.score;
        r.n_cigar = o.n_cigar; r.status = o.status; r.cells = o.cells;
    }
    for (int i = 0; i < n; ++i) {
        halign_dp_result &r = results[i];
        r.cigar_off = out_pos;
        if (out_pos + r.n_cigar > cigar_cap) return fail(ctx, HALIGN_E_CAPACITY, "cigar_out too small");
        const uint32_t *src = hcig.data() + specs[i].cig_off;
        for (int k = 0; k < r.n_cigar; ++k) cigar_out[out_pos + k] = src[r.n_cigar - 1 - k];
        out_pos += r.n_cigar;
    }
    return HALIGN_OK;
}

int halign_ksw_extz2(halign_ctx *ctx, void *km, int qlen, const uint8_t *query, int tlen, const uint8_t *target,
                     int8_t m, const int8_t *mat, int8_t q, int8_t e, int w, int zdrop, int end_bonus, int flag,
                     halign_extz_t *ez)
{
    (void)km;
    if (!ctx || !ez) return HALIGN_E_ARG;
    static const int8_t dna5[25] = {5, -4, -4, -4, 0, -4, 5, -4, -4, 0, -4, -4, 5, -4, 0,
                                    -4, -4, -4, 5, 0, 0, 0, 0, 0, 0};
    // ksw_reset_extz (include/ksw2.h:184-189)
    ez->max_q = ez->max_t = ez->mqe_t = ez->mte_q = -1;
    ez->max = 0; ez->score = ez->mqe = ez->mte = -0x40000000;
    ez->n_cigar = 0; ez->zdropped = 0; ez->reach_end = 0;
    if (m <= 0 || qlen <= 0 || tlen <= 0) return HALIGN_OK;
    if (m != 5 || !mat || memcmp(mat, dna5, 25) != 0 || q != 6 || e != 2 || zdrop >= 0 || end_bonus != 0 ||
        !(flag & 0x04) || (flag & ~0x06))
        return fail(ctx, HALIGN_E_UNSUPPORTED, "ksw_extz2: only the center-star path configuration is supported");
    halign_dp_task t{qlen, tlen, w, flag, 0, 0};
    halign_dp_result r;
    std::vector<uint32_t> cig((size_t)qlen + tlen + 4);
    int rc = halign_dp_batch(ctx, 1, &t, query, target, 1, 1, &r, cig.data(), (int64_t)cig.size(), nullptr);
    if (rc != HALIGN_OK) return rc;
    ez->max = r.max; ez->zdropped = r.zdropped; ez->max_q = r.max_q; ez->max_t = r.max_t;
    ez->mqe = r.mqe; ez->mqe_t = r.mqe_t; ez->mte = r.mte; ez->mte_q = r.mte_q; ez->score = r.score;
    if (r.n_cigar > ez->m_cigar) {
        // ksw_push_cigar's growth (include/ksw2.h:113-123): 4 when empty, then doubling
        int m = ez->m_cigar;
        while (m < r.n_cigar) m = m ? m << 1 : 4;
        uint32_t *nc = (uint32_t *)realloc(ez->cigar, (size_t)m * 4);
        if (!nc) return fail(ctx, HALIGN_E_NOMEM, "cigar realloc");
        ez->cigar = nc;
        ez->m_cigar = m;
    }
    memcpy(ez->cigar, cig.data(), (size_t)r.n_cigar * 4);
    ez->n_cigar = r.n_cigar;
    return HALIGN_OK;
}

int halign_global_align_batch(halign_ctx *ctx, int32_t n, const char *const *refs, const int32_t *ref_lens,
                              const char *const *queries, const int32_t *query_lens, const int32_t *flags,
                              uint32_t *cigar_out, int64_t cigar_cap, int64_t *offsets, int32_t *scores)
{
    if (!ctx || n < 0 || (n > 0 && (!refs || !ref_lens || !queries || !query_lens || !flags || !offsets)))
        return HALIGN_E_ARG;
    // encode (ScoreChar2Idx) and build tasks for the non-empty pairs
    std::vector<uint8_t> qc, tc;
    std::vector<halign_dp_task> tasks;
    std::vector<int> idx;
    for (int i = 0; i < n; ++i) {
        const int rl = ref_lens[i], ql = query_lens[i];
        if (rl < 0 || ql < 0) return fail(ctx, HALIGN_E_ARG, "negative length");
        if (rl == 0 || ql == 0) continue;
        halign_dp_task t;
        t.qlen = ql; t.tlen = rl; t.w = auto_band(rl, ql); t.flag = flags[i];
        t.query_off = (int64_t)qc.size(); t.target_off = (int64_t)tc.size();
        for (int k = 0; k < ql; ++k) qc.push_back(score_code((unsigned char)queries[i][k]));
        for (int k = 0; k < rl; ++k) tc.push_back(score_code((unsigned char)refs[i][k]));
        tasks.push_back(t);
        idx.push_back(i);
    }
    std::vector<halign_dp_result> res(tasks.size());
    int64_t tot = 0;
    for (auto &t : tasks) tot += t.qlen + t.tlen + 4;
    std::vector<uint32_t> cig((size_t)std::max<int64_t>(tot, 1));
    if (!tasks.empty()) {
        int rc = halign_dp_batch(ctx, (int32_t)tasks.size(), tasks.data(), qc.data(), tc.data(), 0, 0, res.data(),
                                 cig.data(), (int64_t)cig.size(), nullptr);
        if (rc != HALIGN_OK) return rc;
    }
    // assemble in input order; empty-side pairs are pure I or D (src/align/align.cpp:26-34)
    int64_t pos = 0;
    size_t k = 0;
    for (int i = 0; i < n; ++i) {
        offsets[i] = pos;
        const int rl = ref_lens[i], ql = query_lens[i];
        if (rl == 0 || ql == 0) {
            if (scores) scores[i] = -0x40000000;
            if (rl == 0 && ql > 0) {
                if (pos + 1 > cigar_cap) return fail(ctx, HALIGN_E_CAPACITY, "cigar_out too small");
                cigar_out[pos++] = (uint32_t)ql << 4 | 1u;
            } else if (ql == 0 && rl > 0) {
                if (pos + 1 > cigar_cap) return fail(ctx, HALIGN_E_CAPACITY, "cigar_out too small");
                cigar_out[pos++] = (uint32_t)rl << 4 | 2u;
            }
            continue;
        }
        const halign_dp_result &r = res[k++];
        if (scores) scores[i] = r.score;
        if (pos + r.n_cigar > cigar_cap) return fail(ctx, HALIGN_E_CAPACITY, "cigar_out too small");
        memcpy(cigar_out + pos, cig.data() + r.cigar_off, (size_t)r.n_cigar * 4);
        pos += r.n_cigar;
    }
    offsets[n] = pos;
    return HALIGN_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Device-resident pair set (kernel-only DP workload and serving)
// ---------------------------------------------------------------------------
namespace {

// encode ASCII queries [first, first+count) into reversed spread even/odd copies
__device__ inline uint8_t ascii_code(unsigned char c)
{
    return (c == 'A' || c == 'a') ? 0 : (c == 'C' || c == 'c') ? 1 : (c == 'G' || c == 'g') ? 2
         : (c == 'T' || c == 't') ? 3 : 4;
}

__global__ void prep_queries(const char *__restrict__ raw, const int64_t *__restrict__ qoff,
                             const int64_t *__restrict__ region, const int64_t *__restrict__ bregion,
                             int first, int count, uint16_t *__restrict__ seq, uint8_t *__restrict__ seqb)
{
    const int qi = blockIdx.y;
    if (qi >= count) return;
    const int64_t beg = qoff[first + qi], len = qoff[first + qi + 1] - beg;
    const int64_t n = len + 2 * kSeqPad;
    uint16_t *even = seq + region[qi];
    uint16_t *odd = even + n;
    auto code = [&](int64_t t) -> uint8_t {   // reversed query code at index t, 0 outside
        if (t < 0 || t >= len) return 0;
        return ascii_code((unsigned char)raw[beg + len - 1 - t]);
    };
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        even[k] = (uint16_t)(code(k - kSeqPad) << 8);
        odd[k] = (k + 1 < n) ? (uint16_t)(code(k + 1 - kSeqPad) << 8) : (uint16_t)0;
    }
    uint8_t *bq = seqb + bregion[qi];
    const int64_t nb = len + 2 * kBytePad;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nb; k += (int64_t)gridDim.x * blockDim.x)
        bq[k] = code(k - kBytePad);
}

}  // namespace

hipError_t hal::prep_queries_launch(const char *raw, const int64_t *qoff, const int64_t *region,
                                    const int64_t *bregion, int first, int count, int64_t mean_len, uint16_t *seq,
                                    uint8_t *seqb, hipStream_t stream)
{
    if (count <= 0) return hipSuccess;
    // enough workgroups of 256 for a query of the mean length (longer ones loop)
    const int gx = (int)std::min<int64_t>(64, std::max<int64_t>(1, (mean_len + 2 * kSeqPad + 255) / 256));
    hipLaunchKernelGGL(prep_queries, dim3(gx, count), dim3(256), 0, stream, raw, qoff, region, bregion, first, count,
                       seq, seqb);
    return hipGetLastError();
}

struct halign_pairset {
    halign_ctx *ctx = nullptr;
    int32_t clen = 0, n = 0;
    DevBuf center;     // spread even/odd center
    DevBuf centerb;    // center as padded byte codes
    DevBuf raw, qoff;  // ASCII queries and their offsets
    std::vector<int64_t> hoff;
    std::vector<uint8_t> hasn;
    bool center_hasn = false;
    DevBuf work, cig, region, bregion;
    int32_t last_first = 0, last_count = 0;
    std::vector<DpOut> outs;
    std::vector<int64_t> cig_off;
};

extern "C" {

int halign_consensus(halign_ctx *ctx, const char *rows, int64_t nrows, int64_t aln_len, int64_t row_stride,
                     int32_t rows_on_device, int64_t batch, uint64_t seq_limit, uint32_t *counts_out,
                     char *cons_out, uint64_t *num_seqs, double *kernel_ms)
{
    if (!ctx || !rows || nrows <= 0 || aln_len <= 0 || row_stride < aln_len || batch <= 0 || !counts_out ||
        !cons_out)
        return fail(ctx, HALIGN_E_ARG, "consensus: empty or inconsistent input");
    std::lock_guard<std::mutex> lock(ctx->mu);
    HCHECK(hipSetDevice(ctx->device), "set device");
    // rows counted: r % (batch + 1) != batch, the first seq_limit of them
    const int64_t B = batch;
    const int64_t avail = (nrows / (B + 1)) * B + std::min<int64_t>(nrows % (B + 1), B);
    int64_t counted = avail, row_end = nrows;
    if (seq_limit > 0 && (uint64_t)avail > seq_limit) {
        counted = (int64_t)seq_limit;
        const int64_t m = counted - 1;   // index among counted rows of the last one
        row_end = (m / B) * (B + 1) + (m % B) + 1;
    }
    if (counted < 1) counted = 1;        // the first record is always counted
    const int64_t stride4 = (aln_len + 3) & ~(int64_t)3;
    const uint8_t *drows;
    int64_t dstride;
    if (rows_on_device) {
        if (row_stride % 4 != 0 || row_stride < stride4 || ((uintptr_t)rows & 3))
            return fail(ctx, HALIGN_E_ARG, "consensus: device rows need a 4-byte aligned stride >= aln_len rounded to 4");
        drows = (const uint8_t *)rows;
        dstride = row_stride;
    } else {
        HCHECK(ctx->crows.ensure((size_t)(row_end * stride4)), "alloc consensus rows");
        HCHECK(hipMemcpy2DAsync(ctx->crows.p, (size_t)stride4, rows, (size_t)row_stride, (size_t)aln_len,
                                (size_t)row_end, hipMemcpyHostToDevice, ctx->stream),
               "upload consensus rows");
        drows = ctx->crows.as<uint8_t>();
        dstride = stride4;
    }
    HCHECK(ctx->ccounts.ensure((size_t)aln_len * 7 * 4), "alloc counts");
    HCHECK(ctx->ccons.ensure((size_t)aln_len), "alloc consensus");
    HCHECK(hipMemsetAsync(ctx->ccounts.p, 0, (size_t)aln_len * 7 * 4, ctx->stream), "clear counts");
    HCHECK(hipEventRecord(ctx->ev0, ctx->stream), "event");
    HCHECK(consensus_launch(drows, dstride, aln_len, row_end, B, ctx->num_cu, ctx->ccounts.as<uint32_t>(),
                            ctx->ccons.as<char>(), ctx->stream),
           "launch consensus");
    HCHECK(hipEventRecord(ctx->ev1, ctx->stream), "event");
    HCHECK(hipMemcpyAsync(counts_out, ctx->ccounts.p, (size_t)aln_len * 7 * 4, hipMemcpyDeviceToHost, ctx->stream),
           "download counts");
    HCHECK(hipMemcpyAsync(cons_out, ctx->ccons.p, (size_t)aln_len, hipMemcpyDeviceToHost, ctx->stream),
           "download consensus");
    HCHECK(hipStreamSynchronize(ctx->stream), "consensus kernels");
    if (kernel_ms) {
        float ms = 0.f;
        HCHECK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1), "event time");
        *kernel_ms = ms;
    }
    if (num_seqs) *num_seqs = (uint64_t)counted;
    return HALIGN_OK;
}

int halign_consensus_noalloc(halign_ctx *ctx, const char *rows_dev, int64_t nrows, int64_t aln_len,
                             int64_t row_stride, int64_t batch, uint64_t seq_limit, uint32_t *counts_out,
                             char *cons_out, uint64_t *num_seqs)
{
    if (!ctx || !rows_dev || nrows <= 0 || aln_len <= 0 || row_stride < aln_len || batch <= 0 || !counts_out ||
        !cons_out)
        return fail(ctx, HALIGN_E_ARG, "consensus_noalloc: empty or inconsistent input");
    if (aln_len > kNoAllocMaxColumns)
        return fail(ctx, HALIGN_E_CAPACITY, "consensus_noalloc: more columns than the static scratch holds (1,048,576)");
    const int64_t stride4 = (aln_len + 3) & ~(int64_t)3;
    if (row_stride % 4 != 0 || row_stride < stride4 || ((uintptr_t)rows_dev & 3))
        return fail(ctx, HALIGN_E_ARG, "consensus_noalloc: device rows need a 4-byte aligned stride >= aln_len rounded to 4");
    std::lock_guard<std::mutex> lock(ctx->mu);
    HCHECK(hipSetDevice(ctx->device), "set device");
    const int64_t B = batch;
    const int64_t avail = (nrows / (B + 1)) * B + std::min<int64_t>(nrows % (B + 1), B);
    int64_t counted = avail, row_end = nrows;
    if (seq_limit > 0 && (uint64_t)avail > seq_limit) {
        counted = (int64_t)seq_limit;
        const int64_t m = counted - 1;
        row_end = (m / B) * (B + 1) + (m % B) + 1;
    }
    if (counted < 1) counted = 1;
    // counts then characters, both in this device's scratch; the previous
    // call's counts are cleared by a fill kernel
    uint32_t *counts = (uint32_t *)ctx->noalloc_scratch;
    char *cons = (char *)ctx->noalloc_scratch + (size_t)kNoAllocMaxColumns * 7 * 4;
    HCHECK(zero_words_launch(counts, aln_len * 7, ctx->stream), "clear counts");
    HCHECK(consensus_launch((const uint8_t *)rows_dev, row_stride, aln_len, row_end, B, ctx->num_cu, counts, cons,
                            ctx->stream),
           "launch consensus");
    HCHECK(hipMemcpyAsync(counts_out, counts, (size_t)aln_len * 7 * 4, hipMemcpyDeviceToHost, ctx->stream),
           "download counts");
    HCHECK(hipMemcpyAsync(cons_out, cons, (size_t)aln_len, hipMemcpyDeviceToHost, ctx->stream), "download consensus");
    HCHECK(hipStreamSynchronize(ctx->stream), "consensus kernels");
    if (num_seqs) *num_seqs = (uint64_t)counted;
    return HALIGN_OK;
}

int halign_pairset_create(halign_ctx *ctx, const char *center, int32_t clen, int32_t n, const char *queries,
                          const int64_t *query_offsets, halign_pairset **out)
{
    if (!ctx || !center || clen <= 0 || n < 0 || !queries || !query_offsets || !out) return HALIGN_E_ARG;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HCHECK(hipSetDevice(ctx->device), "set device");
    halign_pairset *ps = new (std::nothrow) halign_pairset();
    if (!ps) return HALIGN_E_NOMEM;
    ps->ctx = ctx; ps->clen = clen; ps->n = n;
    std::vector<uint8_t> cc(clen);
    for (int i = 0; i < clen; ++i) {
        cc[i] = score_code((unsigned char)center[i]);
        if (cc[i] > 3) ps->center_hasn = true;
    }
    const int64_t cn = clen + 2 * kSeqPad;
    std::vector<uint16_t> hc((size_t)cn * 2);
    spread_target(cc.data(), clen, hc.data(), hc.data() + cn);
    std::vector<uint8_t> hcb((size_t)clen + 2 * kBytePad);
    bytes_target(cc.data(), clen, hcb.data());
    const int64_t total = query_offsets[n] - query_offsets[0];
    ps->hoff.assign(query_offsets, query_offsets + n + 1);
    for (auto &o : ps->hoff) o -= query_offsets[0];
    ps->hasn.assign(n, 0);
    for (int i = 0; i < n; ++i)
        for (int64_t k = ps->hoff[i]; k < ps->hoff[i + 1]; ++k)
            if (score_code((unsigned char)queries[query_offsets[0] + k]) > 3) { ps->hasn[i] = 1; break; }
    hipError_t e;
    if ((e = ps->center.ensure(hc.size() * 2)) != hipSuccess || (e = ps->centerb.ensure(hcb.size())) != hipSuccess ||
        (e = ps->raw.ensure((size_t)std::max<int64_t>(total, 1))) != hipSuccess ||
        (e = ps->qoff.ensure((size_t)(n + 1) * 8)) != hipSuccess) {
        delete ps;
        return fail(ctx, HALIGN_E_NOMEM, "pairset alloc", e);
    }
    HCHECK(hipMemcpy(ps->center.p, hc.data(), hc.size() * 2, hipMemcpyHostToDevice), "upload center");
    HCHECK(hipMemcpy(ps->centerb.p, hcb.data(), hcb.size(), hipMemcpyHostToDevice), "upload center bytes");
    HCHECK(hipMemcpy(ps->raw.p, queries + query_offsets[0], (size_t)total, hipMemcpyHostToDevice), "upload queries");
    HCHECK(hipMemcpy(ps->qoff.p, ps->hoff.data(), (size_t)(n + 1) * 8, hipMemcpyHostToDevice), "upload offsets");
    ctx_attach(ctx);
    *out = ps;
    return HALIGN_OK;
}

int halign_pairset_run(halign_pairset *ps, int32_t first, int32_t count, int32_t plain, double *kernel_ms,
                       int64_t *cells)
{
    if (!ps || first < 0 || count < 0 || first + count > ps->n) return HALIGN_E_ARG;
    halign_ctx *ctx = ps->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HCHECK(hipSetDevice(ctx->device), "set device");
    // The pair set's sequences live in its own buffers; run_specs reads from
    // ctx->seq, so lay the center and the prepared queries out there.
    const int64_t cn = ps->clen + 2 * kSeqPad;
    std::vector<int64_t> region(count), bregion(count);
    int64_t elems = 2 * cn;
    const int64_t cbn = ps->clen + 2 * kBytePad;
    int64_t nbytes = cbn;
    for (int i = 0; i < count; ++i) {
        region[i] = elems;
        bregion[i] = nbytes;
        const int64_t ql = ps->hoff[first + i + 1] - ps->hoff[first + i];
        elems += 2 * (ql + 2 * kSeqPad);
        elems = (elems + 1) & ~(int64_t)1;
        nbytes += ql + 2 * kBytePad;
    }
    HCHECK(ctx->seq.ensure((size_t)elems * 2 + 16), "alloc sequences");
    HCHECK(ctx->seqb.ensure((size_t)nbytes + 16), "alloc byte sequences");
    HCHECK(ps->region.ensure((size_t)std::max(count, 1) * 8), "alloc regions");
    HCHECK(ps->bregion.ensure((size_t)std::max(count, 1) * 8), "alloc regions");
    HCHECK(hipMemcpyAsync(ctx->seq.p, ps->center.p, (size_t)cn * 2 * 2, hipMemcpyDeviceToDevice, ctx->stream),
           "copy center");
    HCHECK(hipMemcpyAsync(ctx->seqb.p, ps->centerb.p, (size_t)cbn, hipMemcpyDeviceToDevice, ctx->stream),
           "copy center bytes");
    HCHECK(hipMemcpyAsync(ps->region.p, region.data(), (size_t)count * 8, hipMemcpyHostToDevice, ctx->stream),
           "upload regions");
    HCHECK(hipMemcpyAsync(ps->bregion.p, bregion.data(), (size_t)count * 8, hipMemcpyHostToDevice, ctx->stream),
           "upload regions");
    if (count > 0) {
        dim3 grid(64, count);
        hipLaunchKernelGGL(prep_queries, grid, dim3(256), 0, ctx->stream, ps->raw.as<const char>(),
                           ps->qoff.as<const int64_t>(), ps->region.as<const int64_t>(),
                           ps->bregion.as<const int64_t>(), first, count, ctx->seq.as<uint16_t>(),
                           ctx->seqb.as<uint8_t>());
        HCHECK(hipGetLastError(), "launch prep");
    }
    std::vector<Spec> specs(count);
    int64_t cig_total = 0;
    ps->cig_off.assign(count, 0);
    for (int i = 0; i < count; ++i) {
        Spec &s = specs[i];
        const int32_t ql = (int32_t)(ps->hoff[first + i + 1] - ps->hoff[first + i]);
        s.tgt_elem = kSeqPad; s.tgt_odd_elem = cn + kSeqPad;
        s.qry_elem = region[i] + kSeqPad;
        s.qry_odd_elem = region[i] + (ql + 2 * kSeqPad) + kSeqPad;
        s.tgt_b = kBytePad;
        s.qry_b = bregion[i] + kBytePad;
        s.tlen = ps->clen; s.qlen = ql;
        s.w = auto_band(ps->clen, ql);
        s.flags = kFlagRight | ((ps->center_hasn || ps->hasn[first + i]) ? kFlagHasN : 0);
        s.cig_off = cig_total;
        s.cig_cap = ql + ps->clen + 4;
        ps->cig_off[i] = cig_total;
        cig_total += s.cig_cap;
        if (ql == 0) return fail(ctx, HALIGN_E_ARG, "empty query in pair set");
        if (!fill_spec_shape(s, false, ctx->policy)) return fail(ctx, HALIGN_E_UNSUPPORTED, "band too wide");
    }
    HCHECK(ctx->cig.ensure((size_t)std::max<int64_t>(cig_total, 1) * 4), "alloc cigars");
    int rc = run_specs(ctx, specs, plain != 0, ps->outs, kernel_ms);
    if (rc != HALIGN_OK) return rc;
    if (cells) {
        int64_t c = 0;
        for (auto &o : ps->outs) c += o.cells;
        *cells = c;
    }
    ps->last_first = first;
    ps->last_count = count;
    return HALIGN_OK;
}

int halign_pairset_result(halign_pairset *ps, int32_t idx, int32_t *score, uint32_t *cigar, int32_t cigar_cap,
                          int32_t *n_cigar)
{
    if (!ps || idx < ps->last_first || idx >= ps->last_first + ps->last_count) return HALIGN_E_ARG;
    halign_ctx *ctx = ps->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    const int k = idx - ps->last_first;
    const DpOut &o = ps->outs[k];
    if (score) *score = o.score;
    if (n_cigar) *n_cigar = o.n_cigar;
    if (cigar) {
        if (o.n_cigar > cigar_cap) return fail(ctx, HALIGN_E_CAPACITY, "cigar buffer too small");
        std::vector<uint32_t> tmp((size_t)std::max(o.n_cigar, 1));
        HCHECK(hipMemcpy(tmp.data(), ctx->cig.as<uint32_t>() + ps->cig_off[k], (size_t)o.n_cigar * 4,
                         hipMemcpyDeviceToHost),
               "download cigar");
        for (int i = 0; i < o.n_cigar; ++i) cigar[i] = tmp[o.n_cigar - 1 - i];
    }
    return HALIGN_OK;
}

void halign_pairset_free(halign_pairset *ps)
{
    if (!ps) return;
    halign_ctx *ctx = ps->ctx;
    {
        std::lock_guard<std::mutex> lock(ctx->mu);
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        ps->center.release(); ps->centerb.release(); ps->raw.release(); ps->bregion.release(); ps->qoff.release();
        ps->work.release(); ps->cig.release(); ps->region.release();
    }
    delete ps;
    ctx_detach(ctx);   // releases the context if it was freed and this was its last child
}

}  // extern "C"
