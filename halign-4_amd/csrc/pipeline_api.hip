// pipeline_api.hip -- host side of the per-query path against a resident
// center (RefAligner::globalAlign, src/align/ref_aligner.cpp:91-125, with one
// center): minimizers -> collect_anchors -> chainAnchors -> globalAlignMM2,
// all on the device, and the C ABI halign_center_*.
//
// Host round trips per batch are small: per-query anchor totals, per-query
// chain lengths, the size of the segment CIGAR buffer, the records of the
// large segments (aligned by the ring / full-statistics kernels), and the
// result.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <rocprim/rocprim.hpp>
#include <string>
#include <thread>
#include <vector>

#include "dp_host.h"
#include "pipeline_common.h"
#include "pipeline_host.h"

namespace hal {

int seed_device(halign_ctx *ctx, int n, const std::vector<int64_t> &off, int k, int w, int non_canonical,
                std::vector<int64_t> &hit_off, std::vector<int64_t> &counts, std::vector<int32_t> &status);
hipError_t anchors_launch_count(const CenterIndexDev &ci, const uint64_t *qx, const uint64_t *qy,
                                const int64_t *hit_off, const int64_t *hit_cnt, int nq, const SeedFilter &f,
                                int groups, uint32_t *scratch, int32_t *hidx, uint32_t *nanch, int64_t *totals,
                                hipStream_t stream);
hipError_t anchors_launch_write(const CenterIndexDev &ci, const uint64_t *qx, const uint64_t *qy,
                                const int64_t *hit_off, const int64_t *hit_cnt, int nq, int groups,
                                const int32_t *hidx, const uint32_t *nanch, const int64_t *aoff, AnchorRec *out,
                                hipStream_t stream);
hipError_t anchor_keys_launch(const AnchorRec *anchors, int64_t n, uint64_t *keys, uint32_t *vals,
                              hipStream_t stream);
hipError_t chain_launch(const AnchorRec *anchors, const uint32_t *order, const int64_t *aoff, const int64_t *na,
                        int nq, const ChainParamsDev &cp, const float *lut, int32_t *f, int32_t *p, uint32_t *chain_idx,
                        int32_t *chain_cnt, hipStream_t stream);
hipError_t mm2_segments_launch(const AnchorRec *anchors, const uint32_t *chain_idx, const int64_t *aoff,
                               const int32_t *chain_cnt, const int64_t *qlens, uint32_t rlen, const int64_t *soff,
                               int nq, SegRec *segs, int32_t *nseg, int64_t *caps, hipStream_t stream);
hipError_t seg_offsets_launch(SegRec *segs, const int64_t *cig_off, int64_t n, hipStream_t stream);
hipError_t seg_split_launch(SegRec *segs, int64_t n, const uint8_t *center, const int64_t *qrev_off,
                            const int64_t *qlens, const uint8_t *qrev, uint32_t *cigs, int32_t *ncig,
                            int64_t *lane_list, unsigned long long *lane_count, int64_t *host_list,
                            unsigned long long *host_count, hipStream_t stream);
hipError_t lane_jobs_launch(int phase, int64_t nmax, const SegRec *segs, const int64_t *list,
                            const unsigned long long *count, const LaneShapeBounds &b, const uint8_t *center,
                            const uint8_t *qrev, const int64_t *qrev_off, const int64_t *qlens, uint32_t *cigs,
                            int *hist, int *range, int *cursor, DpJob *jobs, int64_t *job_seg, const DpOut *outs,
                            SegRec *wsegs, int32_t *ncig, int *err, hipStream_t stream);
hipError_t project_rows_launch(const uint32_t *cig, const int64_t *coff, const char *qseq, const int64_t *qoff,
                               int nq, char *rows, int64_t stride, int64_t *row_len, int64_t *ws, uint8_t *has_ins,
                               hipStream_t stream);
hipError_t mm2_assemble_launch(const SegRec *segs, const int64_t *soff, const int32_t *nseg, const uint32_t *cigs,
                               const int32_t *ncig, const int64_t *qlens, uint32_t rlen, const int64_t *ooff, int nq,
                               uint32_t *out, int64_t *nout, hipStream_t stream);
hipError_t records_launch(int phase, const uint32_t *cig, const int64_t *coff, const char *qseq, const int64_t *qoff,
                          const char *names, const int64_t *name_off, int nq, int64_t rlen, const int32_t *map,
                          int64_t width, int64_t *row_len, int64_t *ws, uint8_t *has_ins, int64_t *sz,
                          unsigned long long *counts, char *out, hipStream_t stream);

}  // namespace hal

using namespace hal;

struct halign_center {
    halign_ctx *ctx = nullptr;
    int32_t rlen = 0;
    bool hasn = false;
    std::vector<uint16_t> spread;           // center even/odd spread codes (kSeqPad padding)
    std::vector<uint8_t> bytes;             // center byte codes (kBytePad padding)
    std::vector<uint64_t> hx, hy;           // the center's minimizer hits, in the reference's order
    DevBuf sx, sy, dh, dstart, docc, lut;   // sorted hits, distinct-hash table, log2f table
    int nd = 0;
    uint64_t thr = 0;
    size_t scratch_groups = 0;
    // per-batch buffers
    DevBuf hcnt, scratch, hidx, nanch, totals, acount, aoff, anchors, keys, keys2, vals, vals2, sort_tmp;
    DevBuf fw, pw, chain_idx, chain_cnt, qlens, soff, segs, nseg, caps, cig_off, scan_tmp, region, bregion;
    DevBuf qrev_off, ncig, llist, hlist, hcount, hsegs, hvals, ooff, out, nout, ocnt, coff, packed;
    DevBuf rows, rlen_d, rws, hins;   // projected rows, their lengths, op workspace, insertion flags
    DevBuf nm, nmoff, rsz, roff, rcnt, recs;   // keep-length FASTA records: names, sizes, offsets, counts, bytes
    DevBuf cmap;         // column of each center position in its aligned record (cwidth > 0)
    int64_t cwidth = 0;
    DevBuf lsmall, ljobs, ljseg, louts, ltb;   // small segments as ring jobs: counters, jobs, outputs, traceback
    // host lists of the large segments' DP jobs and results, kept between
    // batches (a fresh list per batch paid its page faults every time)
    std::vector<Spec> hspecs;
    std::vector<DpOut> houts;
};

namespace {

__global__ void gather_segs(const SegRec *__restrict__ segs, const int64_t *__restrict__ list, int64_t n,
                            SegRec *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = segs[list[i]];
}

// CIGAR lengths for the output scan; a negative length (a CIGAR that does not
// consume the pair) counts 0 here and fails the call on the host
__global__ void clamp_counts(const int64_t *__restrict__ nout, int n, int64_t *__restrict__ cnt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) cnt[i] = nout[i] > 0 ? nout[i] : 0;
}

// one workgroup per query: its CIGAR from the per-query capacity layout to the packed output
__global__ void compact_cigars(const uint32_t *__restrict__ src, const int64_t *__restrict__ ooff,
                               const int64_t *__restrict__ cnt, const int64_t *__restrict__ coff,
                               uint32_t *__restrict__ dst)
{
    const int q = blockIdx.x;
    const uint32_t *a = src + ooff[q];
    uint32_t *b = dst + coff[q];
    const int64_t m = cnt[q];
    for (int64_t k = threadIdx.x; k < m; k += blockDim.x) b[k] = a[k];
}

__global__ void scatter_ncig(const int64_t *__restrict__ list, const int32_t *__restrict__ vals, int64_t n,
                             int32_t *__restrict__ ncig)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ncig[list[i]] = vals[i];
}

constexpr int kK = 15, kW = 10;   // RefAligner's k-mer size and window (include/config.hpp defaults)

// small segments (kSegLane): the sub-wave ring shapes (lanes per job, slots)
// groups of 4 lanes with 1 and 2 slots, of 16 lanes with 1 and 2 slots; the
// query-fixed shapes (lanes per job, slots) for queries of up to 16, 64 and
// 128 bases
constexpr int kLaneShapeDesc[kLaneShapes][2] = {{4, 1}, {4, 2}, {16, 1}, {16, 2}, {1, 1}, {4, 1}, {4, 2}};
// small-segment counters: bucket ranges, tickets, error flag (the head, read
// back in the diagnostic mode), then the per-bin histogram and cursors
constexpr int kLsRange = 0, kLsTicket = 2 * kLaneBuckets, kLsErr = 3 * kLaneBuckets, kLsHead = 64;
constexpr int kLsHist = kLsHead, kLsCursor = kLsHist + kLaneBuckets * kLaneNd, kLsInts = kLsCursor + kLaneBuckets * kLaneNd;
static_assert(kLsErr < kLsHead, "small-segment counter head");

template <class T>
hipError_t up(DevBuf &b, const std::vector<T> &v, hipStream_t st)
{
    hipError_t e = b.ensure(std::max<size_t>(v.size(), 1) * sizeof(T));
    if (e != hipSuccess || v.empty()) return e;
    return hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st);
}

template <class T>
hipError_t down(std::vector<T> &v, const void *src, size_t n, hipStream_t st)
{
    v.resize(n);
    if (n == 0) return hipSuccess;
    hipError_t e = hipMemcpyAsync(v.data(), src, n * sizeof(T), hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(st);
}

// SeedHitBase order (include/seed.h:68-74): hash, rid, pos, strand
bool hit_less(uint64_t ax, uint64_t ay, uint64_t bx, uint64_t by)
{
    const uint64_t ha = ax >> 8, hb = bx >> 8;
    if (ha != hb) return ha < hb;
    const uint32_t ra = (uint32_t)(ay >> 32) & 0x7fffffffu, rb = (uint32_t)(by >> 32) & 0x7fffffffu;
    if (ra != rb) return ra < rb;
    const uint32_t pa = (uint32_t)ay, pb = (uint32_t)by;
    if (pa != pb) return pa < pb;
    return (ay >> 63) < (by >> 63);
}

struct Batch {
    int n = 0;
    std::vector<int64_t> off, qlen, hit_off, hit_cnt, aoff, na;
    std::vector<int32_t> seed_status, chain_cnt, nseg;
    std::vector<int64_t> soff, ooff, nout;
    int64_t packed = 0;   // total CIGAR words of the batch (packed in c->packed at c->coff)
    int64_t anchors = 0, segs = 0, host_segs = 0, lane_segs = 0, cells_host = 0;
    double ms = 0;
};

// memcpy of n bytes split over four host threads (one copy below 1 MiB, or
// when a thread cannot be started)
static void par_copy(char *d, const char *s, size_t n)
{
    constexpr int kThreads = 4;
    if (n < ((size_t)1 << 20)) {
        std::memcpy(d, s, n);
        return;
    }
    const size_t part = ((n + kThreads - 1) / kThreads + 63) & ~(size_t)63;
    std::thread th[kThreads - 1];
    for (int t = 1; t < kThreads; ++t) {
        const size_t a = std::min(n, (size_t)t * part), e = std::min(n, a + part);
        if (a >= e) break;
        try {
            th[t - 1] = std::thread([=] { std::memcpy(d + a, s + a, e - a); });
        } catch (...) {   // copy the rest here
            std::memcpy(d + a, s + a, n - a);
            break;
        }
    }
    std::memcpy(d, s, std::min(n, part));
    for (auto &x : th)
        if (x.joinable()) x.join();
}

// Host-to-device copy of n pageable bytes on stream st through the context's
// two pinned staging buffers: chunk k is copied into buffer k & 1 by host
// threads while the DMA of chunk k-1 runs, and a buffer is refilled only
// after the event recorded behind its previous copy (of this call or the
// last one) has completed. Small copies go directly.
static int upload_staged(halign_ctx *ctx, void *dst, const char *src, size_t n, hipStream_t st)
{
    constexpr size_t kChunk = (size_t)8 << 20;
    if (n <= ((size_t)4 << 20)) {
        HCHECK(hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, st), "upload");
        return HALIGN_OK;
    }
    for (int k = 0; k < 2; ++k) {
        if (!ctx->up_pin[k].ensure(kChunk)) return fail(ctx, HALIGN_E_NOMEM, "pinned staging buffer");
        if (!ctx->up_ev[k]) HCHECK(hipEventCreateWithFlags(&ctx->up_ev[k], hipEventDisableTiming), "event");
    }
    for (size_t off = 0, k = 0; off < n; off += kChunk, ++k) {
        const int b = (int)(k & 1);
        const size_t len = std::min(kChunk, n - off);
        HCHECK(hipEventSynchronize(ctx->up_ev[b]), "staging buffer free");
        par_copy(ctx->up_pin[b].p, src + off, len);
        HCHECK(hipMemcpyAsync((char *)dst + off, ctx->up_pin[b].p, len, hipMemcpyHostToDevice, st), "upload");
        HCHECK(hipEventRecord(ctx->up_ev[b], st), "event");
    }
    return HALIGN_OK;
}

// Stages: 1 = anchors, 2 = chains, 3 = CIGARs (everything).
int run_batch(halign_center *c, int n, const char *queries, const int64_t *offsets, int stop, Batch &b)
{
    halign_ctx *ctx = c->ctx;
    hipStream_t st = ctx->stream;
    PhaseClock clk("run_batch");
    b.n = n;
    b.off.assign(offsets, offsets + n + 1);
    for (auto &o : b.off) o -= offsets[0];
    b.qlen.resize(n);
    for (int i = 0; i < n; ++i) b.qlen[i] = b.off[i + 1] - b.off[i];
    const int64_t total = b.off[n];
    HCHECK(ctx->s_seqs.ensure((size_t)std::max<int64_t>(total, 1)), "alloc queries");
    HCHECK(ctx->s_off.ensure((size_t)(n + 1) * 8), "alloc offsets");
    if (total) {
        const int urc = upload_staged(ctx, ctx->s_seqs.p, queries + offsets[0], (size_t)total, st);
        if (urc != HALIGN_OK) return urc;
    }
    HCHECK(hipMemcpyAsync(ctx->s_off.p, b.off.data(), (size_t)(n + 1) * 8, hipMemcpyHostToDevice, st),
           "upload offsets");

    // minimizers: extractMinimizer(query, 15, 10, non-canonical)
    int rc = seed_device(ctx, n, b.off, kK, kW, 1, b.hit_off, b.hit_cnt, b.seed_status);
    if (rc != HALIGN_OK) return rc;
    clk.mark("upload + minimizers");
    HCHECK(up(c->hcnt, b.hit_cnt, st), "upload hit counts");

    // anchors: count per query, then write in the reference's order
    const int64_t nhits = b.hit_off[n];
    const int groups = std::max(1, std::min(n, ctx->num_cu * 4));
    if ((size_t)groups > c->scratch_groups) {
        const size_t sz = (size_t)groups * (size_t)std::max(c->nd, 1) * 4;
        HCHECK(c->scratch.ensure(sz), "alloc anchor scratch");
        HCHECK(hipMemsetAsync(c->scratch.p, 0, sz, st), "clear anchor scratch");
        c->scratch_groups = groups;
    }
    HCHECK(c->hidx.ensure((size_t)std::max<int64_t>(nhits, 1) * 4), "alloc");
    HCHECK(c->nanch.ensure((size_t)std::max<int64_t>(nhits, 1) * 4), "alloc");
    HCHECK(c->acount.ensure((size_t)std::max(n, 1) * 8), "alloc");
    const CenterIndexDev ci{c->sx.as<uint64_t>(),     c->sy.as<uint64_t>(),   c->dh.as<uint64_t>(),
                            c->dstart.as<uint32_t>(), c->docc.as<uint32_t>(), c->nd, c->thr};
    const SeedFilter f{0.01, 500};   // SeedFilterParams defaults (include/anchor.h:65-71)
    HCHECK(anchors_launch_count(ci, ctx->s_hx.as<uint64_t>(), ctx->s_hy.as<uint64_t>(), ctx->s_hoff.as<int64_t>(),
                                c->hcnt.as<int64_t>(), n, f, groups, c->scratch.as<uint32_t>(), c->hidx.as<int32_t>(),
                                c->nanch.as<uint32_t>(), c->acount.as<int64_t>(), st),
           "launch anchors_count");
    HCHECK(down(b.na, c->acount.p, n, st), "download anchor counts");
    clk.mark("anchor counts");
    b.aoff.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) b.aoff[i + 1] = b.aoff[i] + b.na[i];
    b.anchors = b.aoff[n];
    const int64_t A = std::max<int64_t>(b.anchors, 1);
    HCHECK(up(c->aoff, b.aoff, st), "upload anchor offsets");
    HCHECK(c->anchors.ensure((size_t)A * sizeof(AnchorRec)), "alloc anchors");
    HCHECK(anchors_launch_write(ci, ctx->s_hx.as<uint64_t>(), ctx->s_hy.as<uint64_t>(), ctx->s_hoff.as<int64_t>(),
                                c->hcnt.as<int64_t>(), n, groups, c->hidx.as<int32_t>(), c->nanch.as<uint32_t>(),
                                c->aoff.as<int64_t>(), c->anchors.as<AnchorRec>(), st),
           "launch anchors_write");
    if (stop == 1) {
        HCHECK(hipStreamSynchronize(st), "anchors");
        return HALIGN_OK;
    }

    // chains: each query's anchors sorted by (pos_ref, pos_qry), then the DP
    HCHECK(c->keys.ensure((size_t)A * 8), "alloc");
    HCHECK(c->keys2.ensure((size_t)A * 8), "alloc");
    HCHECK(c->vals.ensure((size_t)A * 4), "alloc");
    HCHECK(c->vals2.ensure((size_t)A * 4), "alloc");
    HCHECK(anchor_keys_launch(c->anchors.as<AnchorRec>(), b.anchors, c->keys.as<uint64_t>(), c->vals.as<uint32_t>(),
                              st),
           "launch anchor_keys");
    if (b.anchors > 0) {
        size_t tmp = 0;
        HCHECK(rocprim::segmented_radix_sort_pairs(nullptr, tmp, c->keys.as<uint64_t>(), c->keys2.as<uint64_t>(),
                                                   c->vals.as<uint32_t>(), c->vals2.as<uint32_t>(),
                                                   (unsigned)b.anchors, (unsigned)n, c->aoff.as<int64_t>(),
                                                   c->aoff.as<int64_t>() + 1, 0u, 64u, st),
               "sort storage size");
        HCHECK(c->sort_tmp.ensure(std::max<size_t>(tmp, 1)), "alloc sort storage");
        HCHECK(rocprim::segmented_radix_sort_pairs(c->sort_tmp.p, tmp, c->keys.as<uint64_t>(),
                                                   c->keys2.as<uint64_t>(), c->vals.as<uint32_t>(),
                                                   c->vals2.as<uint32_t>(), (unsigned)b.anchors, (unsigned)n,
                                                   c->aoff.as<int64_t>(), c->aoff.as<int64_t>() + 1, 0u, 64u, st),
               "segmented sort");
    }
    HCHECK(c->fw.ensure((size_t)A * 4), "alloc");
    HCHECK(c->pw.ensure((size_t)A * 4), "alloc");
    HCHECK(c->chain_idx.ensure((size_t)A * 4), "alloc");
    HCHECK(c->chain_cnt.ensure((size_t)std::max(n, 1) * 4), "alloc");
    const ChainParamsDev cp{5000, 5000, 500, 25, 5000, 3, 40, 0.01f, 0.01f};   // ChainParams defaults
    HCHECK(chain_launch(c->anchors.as<AnchorRec>(), c->vals2.as<uint32_t>(), c->aoff.as<int64_t>(),
                        c->acount.as<int64_t>(), n, cp, c->lut.as<float>(), c->fw.as<int32_t>(), c->pw.as<int32_t>(),
                        c->chain_idx.as<uint32_t>(), c->chain_cnt.as<int32_t>(), st),
           "launch chain");
    HCHECK(down(b.chain_cnt, c->chain_cnt.p, n, st), "download chain lengths");
    clk.mark("anchors + sort + chains");
    if (stop == 2) return HALIGN_OK;

    // MM2 segments
    b.soff.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) b.soff[i + 1] = b.soff[i] + (b.chain_cnt[i] ? 2 * b.chain_cnt[i] + 1 : 1);
    b.segs = b.soff[n];
    const int64_t S = std::max<int64_t>(b.segs, 1);
    HCHECK(up(c->soff, b.soff, st), "upload segment offsets");
    HCHECK(up(c->qlens, b.qlen, st), "upload query lengths");
    HCHECK(c->segs.ensure((size_t)S * sizeof(SegRec)), "alloc segments");
    HCHECK(c->nseg.ensure((size_t)std::max(n, 1) * 4), "alloc");
    HCHECK(c->caps.ensure((size_t)S * 8), "alloc");
    HCHECK(c->cig_off.ensure((size_t)S * 8), "alloc");
    HCHECK(mm2_segments_launch(c->anchors.as<AnchorRec>(), c->chain_idx.as<uint32_t>(), c->aoff.as<int64_t>(),
                               c->chain_cnt.as<int32_t>(), c->qlens.as<int64_t>(), (uint32_t)c->rlen,
                               c->soff.as<int64_t>(), n, c->segs.as<SegRec>(), c->nseg.as<int32_t>(),
                               c->caps.as<int64_t>(), st),
           "launch mm2_segments");
    {
        size_t tmp = 0;
        HCHECK(rocprim::exclusive_scan(nullptr, tmp, c->caps.as<int64_t>(), c->cig_off.as<int64_t>(), (int64_t)0,
                                       (size_t)b.segs, rocprim::plus<int64_t>(), st),
               "scan storage size");
        HCHECK(c->scan_tmp.ensure(std::max<size_t>(tmp, 1)), "alloc scan storage");
        HCHECK(rocprim::exclusive_scan(c->scan_tmp.p, tmp, c->caps.as<int64_t>(), c->cig_off.as<int64_t>(),
                                       (int64_t)0, (size_t)b.segs, rocprim::plus<int64_t>(), st),
               "scan segment CIGAR offsets");
    }
    int64_t last[2] = {0, 0};
    HCHECK(hipMemcpyAsync(&last[0], c->cig_off.as<int64_t>() + (b.segs - 1), 8, hipMemcpyDeviceToHost, st), "down");
    HCHECK(hipMemcpyAsync(&last[1], c->caps.as<int64_t>() + (b.segs - 1), 8, hipMemcpyDeviceToHost, st), "down");
    HCHECK(hipStreamSynchronize(st), "segments");
    clk.mark("segments");
    const int64_t cig_total = std::max<int64_t>(last[0] + last[1], 1);
    HCHECK(seg_offsets_launch(c->segs.as<SegRec>(), c->cig_off.as<int64_t>(), b.segs, st), "launch seg offsets");

    // sequences in ctx->seq / ctx->seqb as the DP runner expects: center first,
    // then each query's reversed copies
    const int64_t cn = c->rlen + 2 * kSeqPad, cbn = c->rlen + 2 * kBytePad;
    std::vector<int64_t> region(n), bregion(n), qrev_off(n);
    int64_t elems = 2 * cn, nbytes = cbn;
    for (int i = 0; i < n; ++i) {
        region[i] = elems;
        bregion[i] = nbytes;
        qrev_off[i] = nbytes + kBytePad;
        elems += 2 * (b.qlen[i] + 2 * kSeqPad);
        elems = (elems + 1) & ~(int64_t)1;
        nbytes += b.qlen[i] + 2 * kBytePad;
    }
    HCHECK(ctx->seq.ensure((size_t)elems * 2 + 16), "alloc sequences");
    HCHECK(ctx->seqb.ensure((size_t)nbytes + 16), "alloc byte sequences");
    HCHECK(hipMemcpyAsync(ctx->seq.p, c->spread.data(), c->spread.size() * 2, hipMemcpyHostToDevice, st),
           "upload center");
    HCHECK(hipMemcpyAsync(ctx->seqb.p, c->bytes.data(), c->bytes.size(), hipMemcpyHostToDevice, st),
           "upload center bytes");
    HCHECK(up(c->region, region, st), "upload regions");
    HCHECK(up(c->bregion, bregion, st), "upload regions");
    HCHECK(up(c->qrev_off, qrev_off, st), "upload regions");
    HCHECK(prep_queries_launch(ctx->s_seqs.as<char>(), ctx->s_off.as<int64_t>(), c->region.as<int64_t>(),
                               c->bregion.as<int64_t>(), 0, n, n ? b.off[n] / n : 0, ctx->seq.as<uint16_t>(),
                               ctx->seqb.as<uint8_t>(), st),
           "launch prep_queries");

    // small segments: exact ones take nM here, the others one lane each on a
    // side stream beside the large segments' DP (disjoint segments, CIGAR
    // ranges and ncig entries); large ones to the host list
    HCHECK(ctx->cig.ensure((size_t)cig_total * 4), "alloc segment CIGARs");
    HCHECK(c->ncig.ensure((size_t)S * 4), "alloc");
    HCHECK(c->llist.ensure((size_t)S * 8), "alloc");
    HCHECK(c->hlist.ensure((size_t)S * 8), "alloc");
    HCHECK(c->hcount.ensure(16), "alloc");
    HCHECK(hipMemsetAsync(c->hcount.p, 0, 16, st), "clear");
    unsigned long long *lcnt = c->hcount.as<unsigned long long>(), *hcnt = lcnt + 1;
    HCHECK(seg_split_launch(c->segs.as<SegRec>(), b.segs, ctx->seqb.as<uint8_t>() + kBytePad,
                            c->qrev_off.as<int64_t>(), c->qlens.as<int64_t>(), ctx->seqb.as<uint8_t>(),
                            ctx->cig.as<uint32_t>(), c->ncig.as<int32_t>(), c->llist.as<int64_t>(), lcnt,
                            c->hlist.as<int64_t>(), hcnt, st),
           "launch seg_split");
    // small segments: DpJobs built on the device, bucketed by shape and gap
    // direction (longest first within a bucket), run by the sub-wave ring and
    // query-fixed kernels with their job ranges read on the device (the
    // buckets run one after another and share one traceback region: a small
    // segment's traceback is at most kLaneMaxTb / 2 bytes on a ring shape and
    // kLaneNd rows of the shape's blocks on a query-fixed one)
    hipStream_t side = ctx->aux[2];
    HCHECK(hipEventRecord(ctx->pfork, st), "event");
    HCHECK(hipStreamWaitEvent(side, ctx->pfork, 0), "stream wait");
    HCHECK(c->lsmall.ensure(kLsInts * 4), "alloc");
    HCHECK(c->ljobs.ensure((size_t)S * sizeof(DpJob)), "alloc small-segment jobs");
    HCHECK(c->ljseg.ensure((size_t)S * 8), "alloc");
    HCHECK(c->louts.ensure((size_t)S * sizeof(DpOut)), "alloc");
    HCHECK(hipMemsetAsync(c->lsmall.p, 0, kLsInts * 4, side), "clear");
    int *ls = c->lsmall.as<int>();
    LaneShapeBounds lb;
    int lshape[kLaneShapes], lgroups[kLaneShapes][2];
    size_t lslot[kLaneShapes], ltb = 0;
    for (int k = 0; k < kLaneShapes; ++k) {
        const int g = kLaneShapeDesc[k][0], sl = kLaneShapeDesc[k][1];
        lshape[k] = -1;
        if (k < kLaneRing) {
            lshape[k] = dp_wave_find_shape(0, g, sl);
            if (lshape[k] >= 0) lb.ncol_max[k] = dp_wave_shape_ring(lshape[k]) - 2;
            lslot[k] = ((size_t)kLaneMaxTb / 2 + 255) & ~(size_t)255;
        } else {
            for (int q = 0; q < dp_wave_num_qfix(); ++q) {
                const int idx = dp_wave_qfix_index(q);
                if (dp_wave_shape_lanes(idx) == g && dp_wave_shape_ring(idx) == g * sl) lshape[k] = idx;
            }
            lb.qcap[k - kLaneRing] = 16 * g * sl;
            lslot[k] = (size_t)kLaneNd * (size_t)(g * sl) * 8;   // kLaneNd rows of the shape's blocks
        }
        if (lshape[k] < 0) return fail(ctx, HALIGN_E_HIP, "small-segment shape missing");
        for (int rt = 0; rt < 2; ++rt) {
            lgroups[k][rt] = std::max(1, dp_wave_max_groups(lshape[k], rt != 0, ctx->num_cu));
            ltb = std::max(ltb, lslot[k] * (size_t)dp_wave_jobs_per_group(lshape[k]) * (size_t)lgroups[k][rt]);
        }
    }
    HCHECK(c->ltb.ensure(ltb), "alloc small-segment traceback");
    const uint8_t *cbytes = ctx->seqb.as<uint8_t>() + kBytePad;
    HCHECK(lane_jobs_launch(0, S, c->segs.as<SegRec>(), c->llist.as<int64_t>(), lcnt, lb, cbytes,
                            ctx->seqb.as<uint8_t>(), c->qrev_off.as<int64_t>(), c->qlens.as<int64_t>(),
                            ctx->cig.as<uint32_t>(), ls + kLsHist, ls + kLsRange, ls + kLsCursor, c->ljobs.as<DpJob>(),
                            c->ljseg.as<int64_t>(), nullptr, nullptr, nullptr, nullptr, side),
           "launch small-segment jobs");
    // bucket timing: pairs from kSmallEv on of the context's event pool (the
    // batch runner's launches take the pairs from 0)
    constexpr size_t kSmallEv = 200;
    for (int k = 0; k < kLaneShapes; ++k)
        for (int rt = 0; rt < 2; ++rt) {
            const int bk = 2 * k + rt;
            hipEvent_t ta = nullptr, tz = nullptr;
            if (ctx->bucket_timing) {
                if (!bucket_events(ctx, kSmallEv + bk, ta, tz)) return fail(ctx, HALIGN_E_HIP, "bucket events");
                HCHECK(hipEventRecord(ta, side), "event");
            }
            HCHECK(dp_wave_launch(lshape[k], rt != 0, lgroups[k][rt], c->ljobs.as<DpJob>(), 0, ls + kLsTicket + bk,
                                  c->ltb.as<uint8_t>(), lslot[k], c->louts.as<DpOut>(), side, ls + kLsRange + 2 * bk),
                   "launch small-segment DP");
            if (tz) HCHECK(hipEventRecord(tz, side), "event");
        }
    if (ctx->bucket_timing) {   // diagnostic mode: wait here and record the launches
        int hls[kLsHead];
        HCHECK(hipMemcpyAsync(hls, ls, sizeof hls, hipMemcpyDeviceToHost, side), "download ranges");
        HCHECK(hipStreamSynchronize(side), "small-segment DPs");
        int njobs = 0;
        for (int bk = 0; bk < 2 * kLaneShapes; ++bk) njobs = std::max(njobs, hls[kLsRange + 2 * bk] + hls[kLsRange + 2 * bk + 1]);
        std::vector<DpOut> lo((size_t)std::max(njobs, 1));
        if (njobs > 0)
            HCHECK(hipMemcpy(lo.data(), c->louts.p, (size_t)njobs * sizeof(DpOut), hipMemcpyDeviceToHost), "download");
        for (int k = 0; k < kLaneShapes; ++k)
            for (int rt = 0; rt < 2; ++rt) {
                const int bk = 2 * k + rt;
                float bms = 0.f;
                HCHECK(hipEventElapsedTime(&bms, ctx->bev[2 * (kSmallEv + bk)], ctx->bev[2 * (kSmallEv + bk) + 1]),
                       "bucket event time");
                const int first = hls[kLsRange + 2 * bk], cnt = hls[kLsRange + 2 * bk + 1];
                halign_bucket_stat b{1, lshape[k], 3, rt, cnt, 0, bms, 0.0};
                for (int j = first; j < first + cnt; ++j) b.cells += lo[(size_t)j].cells;
                if (cnt > 0) ctx->last_buckets.push_back(b);
            }
    }
    HCHECK(lane_jobs_launch(1, S, c->segs.as<SegRec>(), c->llist.as<int64_t>(), lcnt, lb, cbytes,
                            ctx->seqb.as<uint8_t>(), c->qrev_off.as<int64_t>(), c->qlens.as<int64_t>(),
                            ctx->cig.as<uint32_t>(), nullptr, nullptr, nullptr, nullptr, c->ljseg.as<int64_t>(),
                            c->louts.as<DpOut>(), c->segs.as<SegRec>(), c->ncig.as<int32_t>(), ls + kLsErr, side),
           "launch small-segment results");
    HCHECK(hipEventRecord(ctx->pjoin, side), "event");
    // on every return from here on, later work on the main stream waits for the small-segment DPs
    struct JoinSide {
        halign_ctx *ctx;
        ~JoinSide() { (void)hipStreamWaitEvent(ctx->stream, ctx->pjoin, 0); }
    } join_side{ctx};

    // large segments and band overflows: the ring / full-statistics kernels
    unsigned long long nh = 0;
    HCHECK(hipMemcpyAsync(&nh, hcnt, 8, hipMemcpyDeviceToHost, st), "down");
    HCHECK(hipStreamSynchronize(st), "collect");
    clk.mark("prep + collect");
    b.host_segs = (int64_t)nh;
    if (nh > 0) {
        HCHECK(c->hsegs.ensure((size_t)nh * sizeof(SegRec)), "alloc");
        hipLaunchKernelGGL(gather_segs, dim3((unsigned)((nh + 255) / 256)), dim3(256), 0, st, c->segs.as<SegRec>(),
                           c->hlist.as<int64_t>(), (int64_t)nh, c->hsegs.as<SegRec>());
        HCHECK(hipGetLastError(), "launch gather");
        // through pinned memory: a pageable download here waited for the
        // small-segment DPs on the side stream
        if (!ctx->dn_pin.ensure((size_t)nh * sizeof(SegRec))) return fail(ctx, HALIGN_E_NOMEM, "pinned buffer");
        HCHECK(hipMemcpyAsync(ctx->dn_pin.p, c->hsegs.p, (size_t)nh * sizeof(SegRec), hipMemcpyDeviceToHost, st),
               "download large segments");
        HCHECK(hipStreamSynchronize(st), "download large segments");
        const SegRec *hs = reinterpret_cast<const SegRec *>(ctx->dn_pin.p);
        clk.mark("large segment download");
        // keep the device list order (atomic order) paired with the records
        std::vector<Spec> &specs = c->hspecs;
        specs.resize(nh);
        for (size_t k = 0; k < nh; ++k) {
            const SegRec &sg = hs[k];
            Spec &s = specs[k];
            const int64_t L = b.qlen[sg.query];
            const int64_t rq = L - sg.qe;   // reversed index of the segment's last query base
            s.tgt_elem = kSeqPad + sg.rs;
            s.tgt_odd_elem = cn + kSeqPad + sg.rs;
            s.qry_elem = region[sg.query] + kSeqPad + rq;
            s.qry_odd_elem = region[sg.query] + (L + 2 * kSeqPad) + kSeqPad + rq;
            s.tgt_b = kBytePad + sg.rs;
            s.qry_b = bregion[sg.query] + kBytePad + rq;
            s.tlen = (int32_t)(sg.re - sg.rs);
            s.qlen = (int32_t)(sg.qe - sg.qs);
            s.w = auto_band(s.tlen, s.qlen);
            s.flags = ((sg.flags & kSegRight) ? kFlagRight : 0) | kFlagHasN;
            s.cig_off = sg.cig_off;
            s.cig_cap = s.tlen + s.qlen + 1;
            if (!fill_spec_shape(s, false, ctx->policy)) return fail(ctx, HALIGN_E_UNSUPPORTED, "segment band too wide");
        }
        std::vector<DpOut> &outs = c->houts;
        double ms = 0;
        clk.mark("large segment specs");
        rc = run_specs(ctx, specs, false, outs, &ms);
        clk.mark("large segment DP");
        if (rc != HALIGN_OK) return rc;
        std::vector<int32_t> nv(nh);
        for (size_t k = 0; k < nh; ++k) {
            nv[k] = outs[k].n_cigar;
            b.cells_host += outs[k].cells;
        }
        HCHECK(up(c->hvals, nv, st), "upload CIGAR lengths");
        hipLaunchKernelGGL(scatter_ncig, dim3((unsigned)((nh + 255) / 256)), dim3(256), 0, st, c->hlist.as<int64_t>(),
                           c->hvals.as<int32_t>(), (int64_t)nh, c->ncig.as<int32_t>());
        HCHECK(hipGetLastError(), "launch scatter");
    }
    HCHECK(hipStreamWaitEvent(st, ctx->pjoin, 0), "stream wait");   // the small-segment DPs are done before assembly

    // assembly: appendCigar over the segments of each query
    b.ooff.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) b.ooff[i + 1] = b.ooff[i] + b.qlen[i] + c->rlen + 1;
    HCHECK(up(c->ooff, b.ooff, st), "upload output offsets");
    HCHECK(c->out.ensure((size_t)std::max<int64_t>(b.ooff[n], 1) * 4), "alloc output");
    HCHECK(c->nout.ensure((size_t)std::max(n, 1) * 8), "alloc");
    HCHECK(mm2_assemble_launch(c->segs.as<SegRec>(), c->soff.as<int64_t>(), c->nseg.as<int32_t>(),
                               ctx->cig.as<uint32_t>(), c->ncig.as<int32_t>(), c->qlens.as<int64_t>(),
                               (uint32_t)c->rlen, c->ooff.as<int64_t>(), n, c->out.as<uint32_t>(),
                               c->nout.as<int64_t>(), st),
           "launch assemble");
    // packed output: exclusive scan of the lengths, then one copy per query
    HCHECK(c->ocnt.ensure((size_t)(n + 1) * 8), "alloc");
    HCHECK(c->coff.ensure((size_t)(n + 1) * 8), "alloc");
    hipLaunchKernelGGL(clamp_counts, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, c->nout.as<int64_t>(), n,
                       c->ocnt.as<int64_t>());
    HCHECK(hipGetLastError(), "launch clamp");
    HCHECK(hipMemsetAsync(c->ocnt.as<int64_t>() + n, 0, 8, st), "clear");
    {
        size_t tmp = 0;
        HCHECK(rocprim::exclusive_scan(nullptr, tmp, c->ocnt.as<int64_t>(), c->coff.as<int64_t>(), (int64_t)0,
                                       (size_t)n + 1, rocprim::plus<int64_t>(), st),
               "scan storage size");
        HCHECK(c->scan_tmp.ensure(std::max<size_t>(tmp, 1)), "alloc scan storage");
        HCHECK(rocprim::exclusive_scan(c->scan_tmp.p, tmp, c->ocnt.as<int64_t>(), c->coff.as<int64_t>(), (int64_t)0,
                                       (size_t)n + 1, rocprim::plus<int64_t>(), st),
               "scan CIGAR offsets");
    }
    HCHECK(c->packed.ensure((size_t)std::max<int64_t>(b.ooff[n], 1) * 4), "alloc packed output");
    hipLaunchKernelGGL(compact_cigars, dim3((unsigned)n), dim3(256), 0, st, c->out.as<uint32_t>(),
                       c->ooff.as<int64_t>(), c->ocnt.as<int64_t>(), c->coff.as<int64_t>(), c->packed.as<uint32_t>());
    HCHECK(hipGetLastError(), "launch compact");
    b.nout.resize(n);
    int lane_err = 0;
    HCHECK(hipMemcpyAsync(b.nout.data(), c->nout.p, (size_t)n * 8, hipMemcpyDeviceToHost, st), "download");
    HCHECK(hipMemcpyAsync(&b.packed, c->coff.as<int64_t>() + n, 8, hipMemcpyDeviceToHost, st), "download");
    HCHECK(hipMemcpyAsync(&lane_err, ls + kLsErr, 4, hipMemcpyDeviceToHost, st), "download");
    HCHECK(hipStreamSynchronize(st), "assembly");
    if (lane_err) return fail(ctx, HALIGN_E_KERNEL, "a small-segment DP job ended with a kernel status");
    clk.mark("assembly + compaction");
    return HALIGN_OK;
}

}  // namespace

extern "C" {

int halign_center_create(halign_ctx *ctx, const char *center, int32_t len, halign_center **out)
{
    return halign_center_create_indexed(ctx, center, len, center, len, out);
}

int halign_center_create_indexed(halign_ctx *ctx, const char *center, int32_t len, const char *index_seq,
                                 int32_t index_len, halign_center **out)
{
    if (!ctx || !center || len <= 0 || !out || !index_seq || index_len < 0) return HALIGN_E_ARG;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HCHECK(hipSetDevice(ctx->device), "set device");
    halign_center *c = new (std::nothrow) halign_center();
    if (!c) return HALIGN_E_NOMEM;
    c->ctx = ctx;
    c->rlen = len;
    std::vector<uint8_t> codes(len);
    for (int i = 0; i < len; ++i) {
        codes[i] = score_code((unsigned char)center[i]);
        if (codes[i] > 3) c->hasn = true;
    }
    const int64_t cn = len + 2 * kSeqPad;
    c->spread.resize((size_t)cn * 2);
    spread_target(codes.data(), len, c->spread.data(), c->spread.data() + cn);
    c->bytes.resize((size_t)len + 2 * kBytePad);
    bytes_target(codes.data(), len, c->bytes.data());
    // the center's minimizers on the device (same kernels as the queries)
    // (from index_seq; RefAligner::globalAlign extracts them from the aligned
    // sequence itself when the given ones are empty, ref_aligner.cpp:103-106)
    std::vector<int64_t> hoff, cnt;
    std::vector<int32_t> sst;
    for (int pass = 0; pass < 2; ++pass) {
        const char *src = pass == 0 ? index_seq : center;
        const int32_t n = pass == 0 ? index_len : len;
        std::vector<int64_t> off{0, n};
        if (ctx->s_seqs.ensure((size_t)std::max(n, 1)) != hipSuccess || ctx->s_off.ensure(16) != hipSuccess) {
            delete c;
            return fail(ctx, HALIGN_E_NOMEM, "alloc center");
        }
        if (n) (void)hipMemcpy(ctx->s_seqs.p, src, (size_t)n, hipMemcpyHostToDevice);
        (void)hipMemcpy(ctx->s_off.p, off.data(), 16, hipMemcpyHostToDevice);
        const int rc = seed_device(ctx, 1, off, kK, kW, 1, hoff, cnt, sst);
        if (rc != HALIGN_OK) {
            delete c;
            return rc;
        }
        if (cnt[0] > 0 || index_seq == center) break;
    }
    c->hx.resize(cnt[0]);
    c->hy.resize(cnt[0]);
    if (cnt[0]) {
        (void)hipMemcpy(c->hx.data(), ctx->s_hx.p, (size_t)cnt[0] * 8, hipMemcpyDeviceToHost);
        (void)hipMemcpy(c->hy.data(), ctx->s_hy.p, (size_t)cnt[0] * 8, hipMemcpyDeviceToHost);
    }
    // index: hits sorted by (hash, rid, pos, strand), distinct hashes, threshold
    const size_t nh = c->hx.size();
    std::vector<size_t> ord(nh);
    for (size_t i = 0; i < nh; ++i) ord[i] = i;
    std::sort(ord.begin(), ord.end(),
              [&](size_t a, size_t b) { return hit_less(c->hx[a], c->hy[a], c->hx[b], c->hy[b]); });
    std::vector<uint64_t> sx(nh), sy(nh), dh;
    std::vector<uint32_t> dstart, docc;
    for (size_t i = 0; i < nh; ++i) { sx[i] = c->hx[ord[i]]; sy[i] = c->hy[ord[i]]; }
    for (size_t i = 0; i < nh;) {
        size_t j = i;
        while (j < nh && (sx[j] >> 8) == (sx[i] >> 8)) ++j;
        dh.push_back(sx[i] >> 8);
        dstart.push_back((uint32_t)i);
        docc.push_back((uint32_t)(j - i));
        i = j;
    }
    c->nd = (int)dh.size();
    // compute_ref_occ_threshold (src/seed/anchor.cpp:73-117), defaults f = 2e-4, U = [10, 1e6]
    uint64_t f_cutoff = UINT64_MAX;
    const size_t n_skip = (size_t)std::floor(2e-4 * (double)dh.size());
    if (n_skip > 0) {
        std::vector<uint32_t> tmp(docc);
        std::nth_element(tmp.begin(), tmp.begin() + (n_skip - 1), tmp.end(), std::greater<uint32_t>());
        f_cutoff = tmp[n_skip - 1];
    }
    const uint64_t capped = std::min<uint64_t>(1000000, f_cutoff);
    c->thr = std::max<uint64_t>(10, capped);
    // log2f of the C library for the chain penalty (mg_log2, src/seed/anchor.cpp:227-230)
    std::vector<float> lut(502);
    lut[0] = 0.0f;
    for (int d = 1; d <= 501; ++d) lut[d] = log2f((float)d);
    hipStream_t s = ctx->stream;
    if (up(c->sx, sx, s) != hipSuccess || up(c->sy, sy, s) != hipSuccess || up(c->dh, dh, s) != hipSuccess ||
        up(c->dstart, dstart, s) != hipSuccess || up(c->docc, docc, s) != hipSuccess ||
        up(c->lut, lut, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
        delete c;
        return fail(ctx, HALIGN_E_HIP, "upload center index");
    }
    ctx_attach(ctx);
    *out = c;
    return HALIGN_OK;
}

void halign_center_free(halign_center *c)
{
    if (!c) return;
    halign_ctx *ctx = c->ctx;
    {
    std::lock_guard<std::mutex> lock(ctx->mu);
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (int k = 0; k < halign_ctx::kAux; ++k) (void)hipStreamSynchronize(ctx->aux[k]);
    for (DevBuf *b : {&c->sx, &c->sy, &c->dh, &c->dstart, &c->docc, &c->lut, &c->hcnt, &c->scratch, &c->hidx,
                      &c->nanch, &c->totals, &c->acount, &c->aoff, &c->anchors, &c->keys, &c->keys2, &c->vals,
                      &c->vals2, &c->sort_tmp, &c->fw, &c->pw, &c->chain_idx, &c->chain_cnt, &c->qlens,
                      &c->soff, &c->segs, &c->nseg, &c->caps, &c->cig_off, &c->scan_tmp, &c->region, &c->bregion,
                      &c->qrev_off, &c->ncig, &c->llist, &c->hlist, &c->hcount, &c->hsegs, &c->hvals, &c->ooff,
                      &c->out, &c->nout, &c->ocnt, &c->coff, &c->packed, &c->rows,
                      &c->rlen_d, &c->rws, &c->hins, &c->nm, &c->nmoff, &c->rsz, &c->roff, &c->rcnt,
                      &c->recs, &c->cmap, &c->lsmall, &c->ljobs, &c->ljseg, &c->louts, &c->ltb})
        b->release();
    }
    delete c;
    ctx_detach(ctx);   // releases the context if it was freed and this was its last child
}

int halign_center_anchors(halign_center *c, int32_t n, const char *queries, const int64_t *offsets,
                          void *anchors_out, int64_t cap, int64_t *anchor_offsets)
{
    if (!c || n < 0 || (n > 0 && (!queries || !offsets || !anchor_offsets))) return HALIGN_E_ARG;
    halign_ctx *ctx = c->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HCHECK(hipSetDevice(ctx->device), "set device");
    anchor_offsets[0] = 0;
    if (n == 0) return HALIGN_OK;
    Batch b;
    const int rc = run_batch(c, n, queries, offsets, 1, b);
    if (rc != HALIGN_OK) return rc;
    if (b.anchors > cap) return fail(ctx, HALIGN_E_CAPACITY, "anchor output too small");
    for (int i = 0; i <= n; ++i) anchor_offsets[i] = b.aoff[i];
    if (b.anchors && anchors_out)
        HCHECK(hipMemcpy(anchors_out, c->anchors.p, (size_t)b.anchors * sizeof(AnchorRec), hipMemcpyDeviceToHost),
               "download anchors");
    return HALIGN_OK;
}

int halign_center_chains(halign_center *c, int32_t n, const char *queries, const int64_t *offsets,
                         void *chains_out, int64_t cap, int64_t *chain_offsets)
{
    if (!c || n < 0 || (n > 0 && (!queries || !offsets || !chain_offsets))) return HALIGN_E_ARG;
    halign_ctx *ctx = c->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HCHECK(hipSetDevice(ctx->device), "set device");
    chain_offsets[0] = 0;
    if (n == 0) return HALIGN_OK;
    Batch b;
    const int rc = run_batch(c, n, queries, offsets, 2, b);
    if (rc != HALIGN_OK) return rc;
    std::vector<AnchorRec> an;
    std::vector<uint32_t> ci;
    HCHECK(down(an, c->anchors.p, (size_t)b.anchors, ctx->stream), "download anchors");
    HCHECK(down(ci, c->chain_idx.p, (size_t)b.anchors, ctx->stream), "download chains");
    int64_t o = 0;
    for (int i = 0; i < n; ++i) {
        chain_offsets[i] = o;
        if (o + b.chain_cnt[i] > cap) return fail(ctx, HALIGN_E_CAPACITY, "chain output too small");
        for (int k = 0; k < b.chain_cnt[i]; ++k)
            if (chains_out) ((AnchorRec *)chains_out)[o + k] = an[ci[b.aoff[i] + k]];
        o += b.chain_cnt[i];
    }
    chain_offsets[n] = o;
    return HALIGN_OK;
}

}  // extern "C"

namespace {

// halign_center_align, and with rows_out the projected rows as well
int center_align(halign_center *c, int32_t n, const char *queries, const int64_t *offsets, uint32_t *cigar_out,
                 int64_t cap, int64_t *cigar_offsets, int32_t *stats, double *ms, char *rows_out, int64_t row_stride,
                 int64_t *row_len, uint8_t *has_insertion)
{
    if (!c || n < 0 || (n > 0 && (!queries || !offsets || !cigar_offsets))) return HALIGN_E_ARG;
    if (rows_out && (row_stride <= 0 || !row_len || !has_insertion)) return HALIGN_E_ARG;
    halign_ctx *ctx = c->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HCHECK(hipSetDevice(ctx->device), "set device");
    cigar_offsets[0] = 0;
    if (n == 0) return HALIGN_OK;
    hipEvent_t e0, e1;
    HCHECK(hipEventCreate(&e0), "event");
    HCHECK(hipEventCreate(&e1), "event");
    struct Events {
        hipEvent_t a, b;
        ~Events() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); }
    } evs{e0, e1};
    HCHECK(hipEventRecord(e0, ctx->stream), "event");
    Batch b;
    const int rc = run_batch(c, n, queries, offsets, 3, b);
    if (rc != HALIGN_OK) return rc;
    int64_t o = 0;
    for (int i = 0; i < n; ++i) {
        if (b.nout[i] < 0)
            return fail(ctx, HALIGN_E_KERNEL,
                        b.nout[i] == -2 ? "zero-length op in a segment CIGAR" : "MM2 CIGAR does not consume the pair");
        o += b.nout[i];
    }
    if (o > cap) return fail(ctx, HALIGN_E_CAPACITY, "CIGAR output too small");
    if (rows_out) {
        // rows from the packed CIGARs and the raw queries, still on the device
        HCHECK(c->rows.ensure((size_t)n * (size_t)row_stride), "alloc rows");
        HCHECK(c->rlen_d.ensure((size_t)n * 8), "alloc");
        HCHECK(c->rws.ensure((size_t)std::max<int64_t>(b.packed, 1) * 8), "alloc");
        HCHECK(c->hins.ensure((size_t)n), "alloc");
        HCHECK(project_rows_launch(c->packed.as<uint32_t>(), c->coff.as<int64_t>(), ctx->s_seqs.as<char>(),
                                   ctx->s_off.as<int64_t>(), n, c->rows.as<char>(), row_stride,
                                   c->rlen_d.as<int64_t>(), c->rws.as<int64_t>(), c->hins.as<uint8_t>(), ctx->stream),
               "launch project_rows");
    }
    HCHECK(hipEventRecord(e1, ctx->stream), "event");
    if (cigar_out && b.packed > 0)
        HCHECK(hipMemcpyAsync(cigar_out, c->packed.p, (size_t)b.packed * 4, hipMemcpyDeviceToHost, ctx->stream),
               "download CIGARs");
    if (rows_out) {
        HCHECK(hipMemcpyAsync(rows_out, c->rows.p, (size_t)n * (size_t)row_stride, hipMemcpyDeviceToHost, ctx->stream),
               "download rows");
        HCHECK(hipMemcpyAsync(row_len, c->rlen_d.p, (size_t)n * 8, hipMemcpyDeviceToHost, ctx->stream), "download");
        HCHECK(hipMemcpyAsync(has_insertion, c->hins.p, (size_t)n, hipMemcpyDeviceToHost, ctx->stream), "download");
    }
    HCHECK(hipStreamSynchronize(ctx->stream), "download results");
    o = 0;
    for (int i = 0; i < n; ++i) {
        cigar_offsets[i] = o;
        o += b.nout[i];
        if (stats) {
            stats[4 * i + 0] = b.chain_cnt[i] ? 2 * b.chain_cnt[i] + 1 : 1;
            stats[4 * i + 1] = b.chain_cnt[i];
            stats[4 * i + 2] = b.seed_status[i];
            stats[4 * i + 3] = (int32_t)b.na[i];
        }
    }
    cigar_offsets[n] = o;
    if (ms) {
        float t = 0;
        HCHECK(hipEventElapsedTime(&t, e0, e1), "event time");
        *ms = t;
    }
    return HALIGN_OK;
}

}  // namespace

extern "C" {

int halign_center_align(halign_center *c, int32_t n, const char *queries, const int64_t *offsets,
                        uint32_t *cigar_out, int64_t cap, int64_t *cigar_offsets, int32_t *stats, double *ms)
{
    return center_align(c, n, queries, offsets, cigar_out, cap, cigar_offsets, stats, ms, nullptr, 0, nullptr,
                        nullptr);
}

int halign_center_align_rows(halign_center *c, int32_t n, const char *queries, const int64_t *offsets,
                             uint32_t *cigar_out, int64_t cap, int64_t *cigar_offsets, char *rows_out,
                             int64_t row_stride, int64_t *row_len, uint8_t *has_insertion, int32_t *stats, double *ms)
{
    if (!rows_out) return HALIGN_E_ARG;
    return center_align(c, n, queries, offsets, cigar_out, cap, cigar_offsets, stats, ms, rows_out, row_stride,
                        row_len, has_insertion);
}

int halign_center_align_fasta(halign_center *c, int32_t n, const char *queries, const int64_t *offsets,
                              const char *names, const int64_t *name_offsets, char *out, int64_t cap,
                              int64_t *out_len, int64_t *ins_bytes, int32_t *n_insertion, double *ms)
{
    return hal::center_align_fasta(c, n, queries, offsets, names, name_offsets, out, cap, out_len, ins_bytes,
                                   n_insertion, ms, nullptr);
}

}  // extern "C"

namespace hal {

int center_set_columns(halign_center *c, const int32_t *map, int64_t width)
{
    halign_ctx *ctx = c->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HCHECK(hipSetDevice(ctx->device), "set device");
    if (!map) {
        c->cwidth = 0;
        return HALIGN_OK;
    }
    HCHECK(c->cmap.ensure((size_t)std::max(c->rlen, 1) * 4), "alloc column map");
    HCHECK(hipMemcpy(c->cmap.p, map, (size_t)c->rlen * 4, hipMemcpyHostToDevice), "upload column map");
    c->cwidth = width;
    return HALIGN_OK;
}

int center_align_fasta(halign_center *c, int32_t n, const char *queries, const int64_t *offsets, const char *names,
                       const int64_t *name_offsets, char *out, int64_t cap, int64_t *out_len, int64_t *ins_bytes,
                       int32_t *n_insertion, double *ms, uint8_t *has_ins_out)
{
    if (!c || n < 0 || !out_len || !ins_bytes || !n_insertion) return HALIGN_E_ARG;
    if (n > 0 && (!queries || !offsets || !names || !name_offsets || !out)) return HALIGN_E_ARG;
    halign_ctx *ctx = c->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HCHECK(hipSetDevice(ctx->device), "set device");
    *out_len = *ins_bytes = 0;
    *n_insertion = 0;
    if (n == 0) return HALIGN_OK;
    // every record is '>' name '\n' row '\n' with a row of the center's length
    // (or of its aligned record's); a query named "@..." gets none (its SAM
    // line would be skipped)
    const int64_t width = c->cwidth > 0 ? c->cwidth : c->rlen;
    const int32_t *cmap = c->cwidth > 0 ? c->cmap.as<int32_t>() : nullptr;
    std::vector<int64_t> noff((size_t)n + 1);
    int64_t total = 0;
    for (int i = 0; i <= n; ++i) {
        noff[i] = name_offsets[i] - name_offsets[0];
        if (i > 0) {
            if (noff[i] < noff[i - 1]) return HALIGN_E_ARG;
            if (noff[i] > noff[i - 1] && names[name_offsets[i - 1]] == '@') continue;
            total += noff[i] - noff[i - 1] + width + 3;
        }
    }
    if (total > cap) return fail(ctx, HALIGN_E_CAPACITY, "record output too small");
    hipEvent_t e0, e1;
    HCHECK(hipEventCreate(&e0), "event");
    HCHECK(hipEventCreate(&e1), "event");
    struct Events {
        hipEvent_t a, b;
        ~Events() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); }
    } evs{e0, e1};
    hipStream_t st = ctx->stream;
    HCHECK(hipEventRecord(e0, st), "event");
    Batch b;
    const int rc = run_batch(c, n, queries, offsets, 3, b);
    if (rc != HALIGN_OK) return rc;
    for (int i = 0; i < n; ++i)
        if (b.nout[i] < 0)
            return fail(ctx, HALIGN_E_KERNEL,
                        b.nout[i] == -2 ? "zero-length op in a segment CIGAR" : "MM2 CIGAR does not consume the pair");
    const size_t nn = (size_t)n;
    HCHECK(c->nm.ensure((size_t)std::max<int64_t>(noff[n], 1)), "alloc names");
    HCHECK(c->nmoff.ensure((nn + 1) * 8), "alloc");
    HCHECK(c->rlen_d.ensure(nn * 8), "alloc");
    HCHECK(c->rws.ensure((size_t)std::max<int64_t>(b.packed, 1) * 8), "alloc");
    HCHECK(c->hins.ensure(nn), "alloc");
    HCHECK(c->rsz.ensure((2 * nn + 1) * 8), "alloc");
    HCHECK(c->roff.ensure((2 * nn + 1) * 8), "alloc");
    HCHECK(c->rcnt.ensure(16), "alloc");
    HCHECK(c->recs.ensure((size_t)std::max<int64_t>(total, 1)), "alloc records");
    if (noff[n])
        HCHECK(hipMemcpyAsync(c->nm.p, names + name_offsets[0], (size_t)noff[n], hipMemcpyHostToDevice, st),
               "upload names");
    HCHECK(hipMemcpyAsync(c->nmoff.p, noff.data(), (nn + 1) * 8, hipMemcpyHostToDevice, st), "upload name offsets");
    HCHECK(hipMemsetAsync(c->rcnt.p, 0, 16, st), "clear counts");
    const uint32_t *cig = c->packed.as<uint32_t>();
    const int64_t *coff = c->coff.as<int64_t>();
    for (int phase = 0; phase < 3; ++phase) {
        if (phase == 2) {
            size_t tmp = 0;
            HCHECK(rocprim::exclusive_scan(nullptr, tmp, c->rsz.as<int64_t>(), c->roff.as<int64_t>(), (int64_t)0,
                                           2 * nn + 1, rocprim::plus<int64_t>(), st),
                   "scan storage");
            HCHECK(c->scan_tmp.ensure(std::max<size_t>(tmp, 1)), "alloc scan storage");
            HCHECK(rocprim::exclusive_scan(c->scan_tmp.p, tmp, c->rsz.as<int64_t>(), c->roff.as<int64_t>(),
                                           (int64_t)0, 2 * nn + 1, rocprim::plus<int64_t>(), st),
                   "scan record sizes");
        }
        HCHECK(records_launch(phase, cig, coff, ctx->s_seqs.as<char>(), ctx->s_off.as<int64_t>(), c->nm.as<char>(),
                              c->nmoff.as<int64_t>(), n, c->rlen, cmap, width, c->rlen_d.as<int64_t>(),
                              c->rws.as<int64_t>(),
                              c->hins.as<uint8_t>(), phase == 2 ? c->roff.as<int64_t>() : c->rsz.as<int64_t>(),
                              c->rcnt.as<unsigned long long>(), c->recs.as<char>(), st),
               "launch records");
    }
    HCHECK(hipEventRecord(e1, st), "event");
    unsigned long long cnt[2] = {0, 0};
    int64_t ins = 0;
    if (total) HCHECK(hipMemcpyAsync(out, c->recs.p, (size_t)total, hipMemcpyDeviceToHost, st), "download records");
    HCHECK(hipMemcpyAsync(cnt, c->rcnt.p, 16, hipMemcpyDeviceToHost, st), "download");
    HCHECK(hipMemcpyAsync(&ins, c->roff.as<int64_t>() + n, 8, hipMemcpyDeviceToHost, st), "download");
    if (has_ins_out)
        HCHECK(hipMemcpyAsync(has_ins_out, c->hins.p, nn, hipMemcpyDeviceToHost, st), "download insertion flags");
    HCHECK(hipStreamSynchronize(st), "download records");
    if (cnt[1]) return fail(ctx, HALIGN_E_KERNEL, "a row's length differs from the center's");
    *out_len = total;
    *ins_bytes = ins;
    *n_insertion = (int32_t)cnt[0];
    if (ms) {
        float t = 0;
        HCHECK(hipEventElapsedTime(&t, e0, e1), "event time");
        *ms = t;
    }
    return HALIGN_OK;
}

}  // namespace hal
