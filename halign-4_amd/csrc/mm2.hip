// mm2.hip -- align::globalAlignMM2 (src/align/align.cpp:189-305) on the
// device, for many queries against one center:
//
//  * mm2_segments: one wave per query turns its chain into the 2n+1 segments
//    of the reference (first segment LEFT-aligned, the others RIGHT; bounds
//    clipped to the sequences; each segment starts where the previous one
//    ended) and classifies them: empty, one side empty, small DP, large DP
//    or band overflow. A query without a chain gets one whole-pair segment
//    whose CIGAR is taken as it is (align.cpp:204-206).
//  * seg_split: small segments that are exact matches get nM directly; the
//    others become DpJobs built on the device (lane_jobs_*), longest first in
//    each bucket, for the sub-wave ring kernels (dp_wave.hip) or, when the
//    band is the whole matrix and the query the short side, the query-fixed
//    kernels (dp_qfix.hip), each one globalAlignKSW2 (align.cpp:22-65);
//    large segments and band overflows go to the host's batch runner.
//  * mm2_assemble: one wave per query appends the segment CIGARs with
//    appendCigar's merge (src/align/cigar.cpp:302-329; the ring and
//    full-statistics kernels store a CIGAR last op first), replacing a segment
//    whose CIGAR does not consume it exactly by I(query) D(ref)
//    (align.cpp:243-261).
//  * project_rows: delQueryToRefByCigar rows and hasInsertion; rows_scan,
//    record_sizes, write_records: the keep-length FASTA records.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

#include "dp_common.h"
#include "pipeline_common.h"

namespace hal {

// align::auto_band(qlen = ref length, tlen = query length) (include/align.h:159-172)
__device__ __forceinline__ int auto_band_dev(int ref_len, int query_len)
{
    const int d = ref_len > query_len ? ref_len - query_len : query_len - ref_len;
    const int mx = ref_len > query_len ? ref_len : query_len;
    if ((double)d / (double)mx > 0.5) return -1;
    return 200 + (int)(0.1 * (double)(ref_len + query_len / 2));
}

// band overflow ("zdropped") before the last anti-diagonal
__device__ __forceinline__ bool overflows(int qlen, int tlen, int w)
{
    if (w == 0 && qlen + tlen - 1 >= 2) return true;
    const int r = qlen + tlen - 2;
    return (r - qlen + 1 > ((r + w) >> 1)) || (((r - w + 1) >> 1) > tlen - 1);
}

__device__ __forceinline__ uint32_t classify(uint32_t srl, uint32_t sql)
{
    if (srl == 0 && sql == 0) return kSegEmpty;
    if (srl == 0 || sql == 0) return kSegPure;
    int w = auto_band_dev((int)srl, (int)sql);
    if (w < 0) w = srl > sql ? (int)srl : (int)sql;
    if (overflows((int)sql, (int)srl, w)) return kSegHost;
    int ncol = (int)(srl < sql ? srl : sql);
    ncol = ((ncol < w + 1 ? ncol : w + 1) + 15) / 16 + 1;
    const int64_t tb = (int64_t)(srl + sql - 1) * ncol * 16;
    if (srl <= (uint32_t)kLaneMaxT && tb <= kLaneMaxTb) return kSegLane;
    return kSegHost;
}

__device__ __forceinline__ uint32_t seg_cig_cap(uint32_t cls, uint32_t srl, uint32_t sql)
{
    if (cls == kSegEmpty) return 0;
    if (cls == kSegPure) return 1;
    return srl + sql + 1;
}

// One wave per query (four per workgroup). chain_idx/chain_cnt from
// chain_queries; segments of query q go to segs[soff[q] ..]; nseg[q] receives
// their count; caps[s] the CIGAR capacity of segment s.
//
// The reference appends 2n+1 segments for a chain of n anchors: to the first
// anchor, over each anchor's span, to the next anchor, and to the ends. Each
// starts where the previous one ended, and its end is its target clipped to
// the sequence and not below its start, so the ends are running maxima of
// the clipped targets: lane k takes segment k of a chunk of 64, an inclusive
// max-scan gives its end, the lane before it (or the chunk's carry) its start.
__global__ __launch_bounds__(256) void mm2_segments(const AnchorRec *__restrict__ anchors,
                                                    const uint32_t *__restrict__ chain_idx,
                                                    const int64_t *__restrict__ aoff,
                                                    const int32_t *__restrict__ chain_cnt,
                                                    const int64_t *__restrict__ qlens, uint32_t rlen,
                                                    const int64_t *__restrict__ soff, int nq,
                                                    SegRec *__restrict__ segs, int32_t *__restrict__ nseg,
                                                    int64_t *__restrict__ caps)
{
    const int q = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
    if (q >= nq) return;   // uniform per wave
    const uint32_t qlen = (uint32_t)qlens[q];
    const int n = chain_cnt[q];
    SegRec *out = segs + soff[q];
    int64_t *cap = caps + soff[q];
    auto emit = [&](int k, uint32_t rs, uint32_t re, uint32_t qs, uint32_t qe, uint32_t flags) {
        SegRec sr;
        sr.rs = rs; sr.re = re; sr.qs = qs; sr.qe = qe;
        sr.flags = flags; sr.query = (uint32_t)q; sr.pad = 0; sr.cig_off = 0;
        sr.cls = classify(re - rs, qe - qs);
        out[k] = sr;
        cap[k] = seg_cig_cap(sr.cls, re - rs, qe - qs);
    };
    if (n == 0) {   // the whole pair (align.cpp:204-206)
        if (lane == 0) emit(0, 0, rlen, 0, qlen, kSegRight | kSegRaw);
        nseg[q] = 1;
        return;
    }
    const auto mx = [](int32_t x, int32_t y) { return x > y ? x : y; };
    const uint32_t *ch = chain_idx + aoff[q];
    const int ns = 2 * n + 1;
    int32_t cr = 0, cq = 0;   // end of the segments so far
    for (int k0 = 0; k0 < ns; k0 += 64) {
        const int k = k0 + lane;
        uint64_t tr = rlen, tq = qlen;   // target end; the last segment runs to the ends
        if (k < ns - 1) {
            // k = 0: first anchor's start; k = 2i+1: anchor i's span end; k = 2i+2: anchor i+1's start
            const AnchorRec a = anchors[ch[k >> 1]];
            const uint64_t sp = (k & 1) ? a.span : 0;
            tr = (uint64_t)a.pos_ref + sp;
            tq = (uint64_t)a.pos_qry + sp;
        }
        const int32_t er = (int32_t)(tr < rlen ? tr : rlen), eq = (int32_t)(tq < qlen ? tq : qlen);
        const int32_t re = mx(wave_scan_incl(k < ns ? er : 0, 0, mx), cr);
        const int32_t qe = mx(wave_scan_incl(k < ns ? eq : 0, 0, mx), cq);
        const int32_t rs = wave_shr1_id(re, cr), qs = wave_shr1_id(qe, cq);
        if (k < ns) emit(k, (uint32_t)rs, (uint32_t)re, (uint32_t)qs, (uint32_t)qe, k == 0 ? 0u : kSegRight);
        cr = __builtin_amdgcn_readlane(re, 63);
        cq = __builtin_amdgcn_readlane(qe, 63);
    }
    nseg[q] = ns;   // every lane stores the same value
}

__global__ void seg_set_offsets(SegRec *__restrict__ segs, const int64_t *__restrict__ cig_off, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) segs[i].cig_off = cig_off[i];
}



// One-gap segments: with a full band (auto_band -1) ksw_extz2 finds an
// optimal alignment exactly. For a query part Q of length m shorter than the
// reference part R of length n, every alignment has at least g = n - m
// deleted bases; one run of them with m matching ACGT pairs scores
// 5m - (6 + 2g), and anything else (a mismatch or N pair, a second gap run,
// an insertion) scores less. One run with all m bases matching is
// M^a D^g M^(m-a) with Q[0,a) = R[0,a) and Q[a,m) = R[a+g,n). If Q = R[g,n)
// (Q is R's tail) the second part holds for every a, so the optimal
// alignments are exactly those with a up to the common ACGT prefix of Q and
// R; if Q = R[0,m) (R's head) they are M^(m-b) D^g M^b with b up to the
// common suffix. Among them ksw_extz2 returns the leftmost gap (a = 0,
// resp. the largest b) and, with KSW_EZ_RIGHT, the rightmost one (the
// largest a, resp. b = 0): its tie order puts gaps left or right. With the
// roles swapped (m > n) the same holds with I. Returns the number of ops
// written to op (2 or 3), or 0 when neither case applies.
__device__ __forceinline__ int one_gap(const uint8_t *__restrict__ r, const uint8_t *__restrict__ qr, int n, int m,
                                       bool right, uint32_t (&op)[3])
{
    // r[k]: reference base k; qr[-k]: query base k (the query is stored reversed)
    if (n == m || auto_band_dev(n, m) >= 0) return 0;
    const bool del = m < n;
    const int s = del ? m : n, g = del ? n - m : m - n;   // matched bases, gap length
    const uint32_t gop = (uint32_t)g << 4 | (del ? 2u : 1u);
    int a = -1;   // matched bases before the gap
    bool tail = true;
    for (int k = 0; k < s && tail; ++k) {
        const uint8_t x = del ? r[g + k] : r[k], y = del ? qr[-k] : qr[-(g + k)];
        tail = x < 4 && x == y;
    }
    if (tail) {
        a = 0;
        if (right)
            while (a < s && r[a] < 4 && r[a] == qr[-a]) ++a;   // common prefix of R and Q
    } else {
        bool head = true;
        for (int k = 0; k < s && head; ++k) head = r[k] < 4 && r[k] == qr[-k];
        if (head) {
            a = s;
            if (!right) {
                int b = 0;   // common suffix of R and Q
                while (b < s && r[n - 1 - b] < 4 && r[n - 1 - b] == qr[-(m - 1 - b)]) ++b;
                a = s - b;
            }
        }
    }
    if (a < 0) return 0;
    int k = 0;
    if (a > 0) op[k++] = (uint32_t)a << 4;
    op[k++] = gop;
    if (s - a > 0) op[k++] = (uint32_t)(s - a) << 4;
    return k;
}

// Segments split three ways, one thread per segment: an equal-length small
// segment whose bases are identical ACGT (the k-mer of an anchor, nearly every
// one) has the all-match path as its unique optimum, so ksw_extz2 returns the
// CIGAR nM; a segment whose optimum is one gap (one_gap: typically the ends
// of a query cut from the center, a long reference part against the few query
// bases before the first anchor) gets its two or three ops. Both are written here,
// class kSegExact. Other small segments go to the lane list, large ones to the
// host list (both appended with an atomic).
__global__ __launch_bounds__(256) void seg_split(SegRec *__restrict__ segs, int64_t n,
                                                 const uint8_t *__restrict__ center,
                                                 const int64_t *__restrict__ qrev_off, const int64_t *__restrict__ qlens,
                                                 const uint8_t *__restrict__ qrev, uint32_t *__restrict__ cigs,
                                                 int32_t *__restrict__ ncig, int64_t *__restrict__ lane_list,
                                                 unsigned long long *__restrict__ lane_count,
                                                 int64_t *__restrict__ host_list,
                                                 unsigned long long *__restrict__ host_count)
{
    // list slots: ranks in LDS, one global atomic per list and workgroup
    __shared__ unsigned int lc[2];
    __shared__ unsigned long long lbase[2];
    if (threadIdx.x < 2) lc[threadIdx.x] = 0;
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int which = -1;   // 0: lane list, 1: host list
    unsigned int rank = 0;
    if (i < n) {
        const SegRec sg = segs[i];
        uint32_t op[3];
        int nop = 0;
        if (sg.cls == kSegHost || sg.cls == kSegLane)
            nop = one_gap(center + sg.rs, qrev + qrev_off[sg.query] + (qlens[sg.query] - 1 - sg.qs),
                          (int)(sg.re - sg.rs), (int)(sg.qe - sg.qs), (sg.flags & kSegRight) != 0, op);
        const int srl0 = (int)(sg.re - sg.rs), sql0 = (int)(sg.qe - sg.qs);
        int w0 = auto_band_dev(srl0, sql0);
        if (w0 < 0) w0 = srl0 > sql0 ? srl0 : sql0;
        if (nop > 0) {
            for (int k = 0; k < nop; ++k) cigs[sg.cig_off + k] = op[k];
            ncig[i] = nop;
            segs[i].cls = kSegExact;
        } else if (sg.cls == kSegHost && !(sg.flags & kSegRaw) && overflows(sql0, srl0, w0)) {
            // The band leaves the matrix before the last anti-diagonal: ksw_extz2
            // stops there (zdropped) and backtracks from its maximum, a cell
            // short of the last one, so the CIGAR never consumes the segment and
            // globalAlignMM2 replaces it by I(query) D(ref) whatever its ops
            // (align.cpp:243-261). No DP: an empty CIGAR, which the assembly
            // replaces the same way.
            ncig[i] = 0;
            segs[i].cls = kSegExact;
        } else if (sg.cls == kSegHost) {
            which = 1;
        } else if (sg.cls == kSegLane) {
            const int srl = (int)(sg.re - sg.rs), sql = (int)(sg.qe - sg.qs);
            bool same = srl == sql;
            if (same) {
                const uint8_t *qr = qrev + qrev_off[sg.query] + (qlens[sg.query] - 1 - sg.qs);
                const uint8_t *t = center + sg.rs;
                for (int k = 0; k < srl && same; ++k) same = t[k] < 4 && t[k] == qr[-k];
            }
            if (same) {
                cigs[sg.cig_off] = (uint32_t)srl << 4;   // M
                ncig[i] = 1;
                segs[i].cls = kSegExact;
            } else {
                which = 0;
            }
        }
        if (which >= 0) rank = atomicAdd(&lc[which], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 2 && lc[threadIdx.x])
        lbase[threadIdx.x] = atomicAdd(threadIdx.x ? host_count : lane_count, (unsigned long long)lc[threadIdx.x]);
    __syncthreads();
    if (which == 0) lane_list[lbase[0] + rank] = i;
    if (which == 1) host_list[lbase[1] + rank] = i;
}

// ---- small segments as DP jobs built on the device ----

// bucket (shape * 2 + gap direction) and anti-diagonal count of a small
// segment: a query-fixed shape when the band is the whole matrix and the query
// the short side (as the batch runner's qfix_pick), else the first sub-wave
// ring shape whose n_col_ bound holds the band
__device__ __forceinline__ int lane_bucket(const SegRec &sg, const LaneShapeBounds &b, int &nd)
{
    const int srl = (int)(sg.re - sg.rs), sql = (int)(sg.qe - sg.qs);
    nd = srl + sql - 1;
    const int rt = (sg.flags & kSegRight) ? 1 : 0;
    int w = auto_band_dev(srl, sql);
    if (w < 0) w = srl > sql ? srl : sql;
    if (sql <= srl && w >= srl - 1) {
        for (int q = 0; q < kLaneQfix; ++q)
            if (sql <= b.qcap[q]) return (kLaneRing + q) * 2 + rt;
    }
    int nc = srl < sql ? srl : sql;
    nc = ((nc < w + 1 ? nc : w + 1) + 15) / 16 + 1;
    int sh = 0;
    while (sh < kLaneRing - 1 && nc > b.ncol_max[sh]) ++sh;
    return sh * 2 + rt;
}

// bin of a small segment: its bucket's kLaneNd bins by anti-diagonal count
// (a small segment has 1..kLaneNd anti-diagonals: its traceback is at most
// kLaneMaxTb bytes at 16 per block and n_col_ >= 2)
constexpr int kLaneBins = kLaneBuckets * kLaneNd;
__device__ __forceinline__ int lane_bin(const SegRec &sg, const LaneShapeBounds &b)
{
    int nd = 0;
    const int bk = lane_bucket(sg, b, nd);
    nd = nd < 1 ? 1 : nd > kLaneNd ? kLaneNd : nd;
    return bk * kLaneNd + nd - 1;
}

// per-bin job counts (hist[kLaneBins], cleared before): counted in LDS, one
// global atomic per bin and workgroup (by the thread that opened the bin)
__global__ __launch_bounds__(256) void lane_jobs_count(const SegRec *__restrict__ segs,
                                                       const int64_t *__restrict__ list,
                                                       const unsigned long long *__restrict__ count,
                                                       LaneShapeBounds b, int *__restrict__ hist)
{
    __shared__ int lc[kLaneBins];
    for (int k = (int)threadIdx.x; k < kLaneBins; k += 256) lc[k] = 0;
    __syncthreads();
    const int64_t li = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int bin = -1, rank = 0;
    if (li < (int64_t)*count) {
        bin = lane_bin(segs[list[li]], b);
        rank = atomicAdd(&lc[bin], 1);
    }
    __syncthreads();
    if (bin >= 0 && rank == 0) atomicAdd(&hist[bin], lc[bin]);
}

// range[2k] = first job of bucket k, range[2k+1] = its job count; cursor[bin]
// = the first job of the bin: the bins of a bucket from the most
// anti-diagonals down (one workgroup)
__global__ __launch_bounds__(64) void lane_jobs_offsets(const int *__restrict__ hist, int *__restrict__ range,
                                                        int *__restrict__ cursor)
{
    __shared__ int tot[kLaneBuckets], first[kLaneBuckets];
    const int t = (int)threadIdx.x;
    if (t < kLaneBuckets) {
        int sum = 0;
        for (int d = 0; d < kLaneNd; ++d) sum += hist[t * kLaneNd + d];
        tot[t] = sum;
    }
    __syncthreads();
    if (t == 0) {
        int at = 0;
        for (int k = 0; k < kLaneBuckets; ++k) {
            first[k] = at;
            at += tot[k];
        }
    }
    __syncthreads();
    if (t < kLaneBuckets) {
        range[2 * t] = first[t];
        range[2 * t + 1] = tot[t];
        int at = first[t];
        for (int d = kLaneNd - 1; d >= 0; --d) {
            cursor[t * kLaneNd + d] = at;
            at += hist[t * kLaneNd + d];
        }
    }
}

// one DpJob per small segment at its bin's cursor (byte-code windows of the
// center and of the reversed query, as the host builds them for large
// segments); ranks within the workgroup in LDS, one global atomic per bin
__global__ __launch_bounds__(256) void lane_jobs_place(
    const SegRec *__restrict__ segs, const int64_t *__restrict__ list, const unsigned long long *__restrict__ count,
    LaneShapeBounds b, const uint8_t *__restrict__ center, const uint8_t *__restrict__ qrev,
    const int64_t *__restrict__ qrev_off, const int64_t *__restrict__ qlens, uint32_t *__restrict__ cigs,
    int *__restrict__ cursor, DpJob *__restrict__ jobs, int64_t *__restrict__ job_seg)
{
    __shared__ int lc[kLaneBins], lbase[kLaneBins];
    for (int k = (int)threadIdx.x; k < kLaneBins; k += 256) lc[k] = 0;
    __syncthreads();
    const int64_t li = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = li < (int64_t)*count;
    int64_t i = 0;
    SegRec sg{};
    int bin = 0, rank = 0;
    if (valid) {
        i = list[li];
        sg = segs[i];
        bin = lane_bin(sg, b);
        rank = atomicAdd(&lc[bin], 1);
    }
    __syncthreads();
    if (valid && rank == 0) lbase[bin] = atomicAdd(&cursor[bin], lc[bin]);
    __syncthreads();
    if (!valid) return;
    const int pos = lbase[bin] + rank;
    const int srl = (int)(sg.re - sg.rs), sql = (int)(sg.qe - sg.qs);
    DpJob j;
    j.tgt_even = j.tgt_odd = j.qr_even = j.qr_odd = nullptr;   // spread copies: multi-wave kernel only
    j.tgt_b = center + sg.rs;
    j.qr_b = qrev + qrev_off[sg.query] + (qlens[sg.query] - sg.qe);   // reversed index of base qe-1
    j.tlen = srl;
    j.qlen = sql;
    j.w = auto_band_dev(srl, sql);
    j.flags = ((sg.flags & kSegRight) ? kFlagRight : 0) | kFlagHasN;
    j.cigar = cigs + sg.cig_off;
    j.cig_cap = srl + sql + 1;
    j.pad_ = 0;
    jobs[pos] = j;
    job_seg[pos] = i;
}

// CIGAR lengths back to the segments (class kSegRing: ops stored last first);
// a job that did not end with status ok sets *err
__global__ void lane_jobs_finish(const DpOut *__restrict__ outs, const int64_t *__restrict__ job_seg,
                                 const unsigned long long *__restrict__ count, SegRec *__restrict__ segs,
                                 int32_t *__restrict__ ncig, int *__restrict__ err)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= (int64_t)*count) return;
    const int64_t i = job_seg[j];
    const DpOut o = outs[j];
    if (o.status != kStOk) atomicOr(err, 1);
    ncig[i] = o.n_cigar;
    segs[i].cls = kSegRing;
}

// phase 0: jobs (hist: kLaneBins counters, cleared before; cursor:
// kLaneBins; range: 2 * kLaneBuckets); phase 1: results back to the segments
hipError_t lane_jobs_launch(int phase, int64_t nmax, const SegRec *segs, const int64_t *list,
                            const unsigned long long *count, const LaneShapeBounds &b, const uint8_t *center,
                            const uint8_t *qrev, const int64_t *qrev_off, const int64_t *qlens, uint32_t *cigs,
                            int *hist, int *range, int *cursor, DpJob *jobs, int64_t *job_seg, const DpOut *outs,
                            SegRec *wsegs, int32_t *ncig, int *err, hipStream_t stream)
{
    if (nmax <= 0) return hipSuccess;
    const dim3 grid((unsigned)((nmax + 255) / 256)), blk(256);
    if (phase == 0) {
        hipLaunchKernelGGL(lane_jobs_count, grid, blk, 0, stream, segs, list, count, b, hist);
        hipLaunchKernelGGL(lane_jobs_offsets, dim3(1), dim3(64), 0, stream, hist, range, cursor);
        hipLaunchKernelGGL(lane_jobs_place, grid, blk, 0, stream, segs, list, count, b, center, qrev, qrev_off, qlens,
                           cigs, cursor, jobs, job_seg);
    } else {
        hipLaunchKernelGGL(lane_jobs_finish, grid, blk, 0, stream, outs, job_seg, count, wsegs, ncig, err);
    }
    return hipGetLastError();
}


// CIGAR consumption (src/align/cigar.cpp:332-365): ref = M D N = X, query = M I S = X
__device__ __forceinline__ void consumed(const uint32_t *c, int n, int64_t &rl, int64_t &ql)
{
    rl = ql = 0;
    for (int i = 0; i < n; ++i) {
        const uint32_t op = c[i] & 0xf, len = c[i] >> 4;
        if (op == 0 || op == 2 || op == 3 || op == 7 || op == 8) rl += len;
        if (op == 0 || op == 1 || op == 4 || op == 7 || op == 8) ql += len;
    }
}


// One wave per query (four per workgroup): the final CIGAR at out[ooff[q] ..],
// its length in nout[q] (-1 when the total does not consume both sequences:
// the caller then aligns the whole pair, align.cpp:296-302; -2 when a segment
// CIGAR holds a zero-length op, which no source writes).
//
// The segments go in chunks of 64, lane k holding segment k of the chunk and
// its op list: none (both sides empty), the one op of a one-side-empty
// segment, the forced I(query) D(ref) pair of a CIGAR that does not consume
// its segment, or the CIGAR itself, forward or last op first. The chunk's ops
// then go in tiles of 64, lane t taking op t (its segment found by binary
// search over the op offsets). appendCigar's merge of equal neighbours is a
// run-length encoding of that op stream: an op whose type differs from the
// op before it opens output op (number of such heads so far - 1), and the
// last op of each run writes the run's summed length; the run open at a
// tile's end carries into the next tile, which rewrites its output op.
__global__ __launch_bounds__(256) void mm2_assemble(const SegRec *__restrict__ segs, const int64_t *__restrict__ soff,
                                                    const int32_t *__restrict__ nseg, const uint32_t *__restrict__ cigs,
                                                    const int32_t *__restrict__ ncig, const int64_t *__restrict__ qlens,
                                                    uint32_t rlen, const int64_t *__restrict__ ooff, int nq,
                                                    uint32_t *__restrict__ out, int64_t *__restrict__ nout)
{
    __shared__ int32_t s_incl[4][64], s_nc[4][64];
    __shared__ uint32_t s_v0[4][64], s_v1[4][64], s_mode[4][64];
    __shared__ int64_t s_off[4][64];
    const int wv = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
    const int q = (int)blockIdx.x * 4 + wv;
    if (q >= nq) return;   // uniform per wave
    const auto add = [](int32_t x, int32_t y) { return x + y; };
    const auto mx = [](int32_t x, int32_t y) { return x > y ? x : y; };
    const int64_t qlen = qlens[q], cap = (int64_t)rlen + qlen + 1, sbase = soff[q];
    uint32_t *o = out + ooff[q];
    const SegRec *sg = segs + sbase;
    const int ns = nseg[q];
    int32_t c_type = -1, c_idx = -1, c_len = 0;   // the run open so far
    int64_t rl = 0, ql = 0;                       // this lane's share of the consumption
    bool zero = false;
    for (int k0 = 0; k0 < ns; k0 += 64) {
        const int k = k0 + lane;
        uint32_t mode = 0, v0 = 0, v1 = 0;   // mode 1: ops v0 (v1); 2: CIGAR forward; 3: last op first
        int32_t m = 0, nc = 0;
        int64_t off = 0;
        if (k < ns) {
            const SegRec sr = sg[k];
            const uint32_t srl = sr.re - sr.rs, sql = sr.qe - sr.qs;
            if (sr.cls == kSegPure) {   // globalAlignKSW2 on an empty side (align.cpp:26-34)
                mode = 1;
                m = 1;
                v0 = srl == 0 ? (sql << 4 | 1u) : (srl << 4 | 2u);
            } else if (sr.cls != kSegEmpty) {
                nc = ncig[sbase + k];
                off = sr.cig_off;
                bool forced = false;
                if (!(sr.flags & kSegRaw)) {
                    int64_t crl, cql;
                    consumed(cigs + off, nc, crl, cql);
                    forced = crl != srl || cql != sql;
                }
                if (forced) {   // I(query) D(ref), empty ones left out
                    mode = 1;
                    const uint32_t i_op = sql << 4 | 1u, d_op = srl << 4 | 2u;
                    m = (sql > 0) + (srl > 0);
                    v0 = sql > 0 ? i_op : d_op;
                    v1 = d_op;
                } else {
                    mode = (sr.cls == kSegHost || sr.cls == kSegRing) ? 3u : 2u;
                    m = nc;
                }
            }
        }
        const int32_t incl = wave_scan_incl(m, 0, add);
        const int32_t T = __builtin_amdgcn_readlane(incl, 63);
        s_incl[wv][lane] = incl;
        s_nc[wv][lane] = nc;
        s_v0[wv][lane] = v0;
        s_v1[wv][lane] = v1;
        s_mode[wv][lane] = mode;
        s_off[wv][lane] = off;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        for (int t0 = 0; t0 < T; t0 += 64) {
            const int t = t0 + lane;
            const bool valid = t < T;
            uint32_t v = 0;
            if (valid) {
                int lo = 0, hi = 63;   // first segment whose inclusive op count passes t
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (s_incl[wv][mid] > t) hi = mid;
                    else lo = mid + 1;
                }
                const int i = t - (lo ? s_incl[wv][lo - 1] : 0);
                const uint32_t md = s_mode[wv][lo];
                if (md == 1) v = i ? s_v1[wv][lo] : s_v0[wv][lo];
                else v = cigs[s_off[wv][lo] + (md == 3 ? s_nc[wv][lo] - 1 - i : i)];
            }
            const int32_t ty = valid ? (int32_t)(v & 0xfu) : -2;
            const int32_t ln = valid ? (int32_t)(v >> 4) : 0;
            zero |= valid && ln == 0;
            if (valid) {   // cigar::getRefLength / getQueryLength (src/align/cigar.cpp:332-365)
                if (ty == 0 || ty == 2 || ty == 3 || ty == 7 || ty == 8) rl += ln;
                if (ty == 0 || ty == 1 || ty == 4 || ty == 7 || ty == 8) ql += ln;
            }
            const int32_t head = valid && ty != wave_shr1_id(ty, c_type) ? 1 : 0;
            const int32_t idx = c_idx + wave_scan_incl(head, 0, add);
            const int32_t P = wave_scan_incl(ln, 0, add);
            const int32_t hb = wave_scan_incl(head ? P - ln : INT_MIN, INT_MIN, mx);
            const int32_t run = hb == INT_MIN ? c_len + P : P - hb;
            // the next lane opens a run, or this is the tile's last op
            const int32_t next_head = __builtin_amdgcn_update_dpp(1, head, 0x130, 0xF, 0xF, false);
            if (valid && (next_head || t + 1 == T) && idx < cap) o[idx] = (uint32_t)run << 4 | (uint32_t)ty;
            const int tl = (T - t0 < 64 ? T - t0 : 64) - 1;
            c_type = __builtin_amdgcn_readlane(ty, tl);
            c_idx = __builtin_amdgcn_readlane(idx, tl);
            c_len = __builtin_amdgcn_readlane(run, tl);
        }
        // the next chunk rewrites the LDS tables
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
    for (int d = 32; d > 0; d >>= 1) {
        rl += __shfl_xor(rl, d);
        ql += __shfl_xor(ql, d);
    }
    const bool any_zero = __ballot(zero) != 0;
    const bool raw = ns == 1 && (sg[0].flags & kSegRaw);
    // every lane stores the same value
    nout[q] = any_zero ? -2 : ((!raw && (rl != (int64_t)rlen || ql != qlen)) ? -1 : (int64_t)c_idx + 1);
}

hipError_t mm2_segments_launch(const AnchorRec *anchors, const uint32_t *chain_idx, const int64_t *aoff,
                               const int32_t *chain_cnt, const int64_t *qlens, uint32_t rlen, const int64_t *soff,
                               int nq, SegRec *segs, int32_t *nseg, int64_t *caps, hipStream_t stream)
{
    if (nq <= 0) return hipSuccess;
    hipLaunchKernelGGL(mm2_segments, dim3((nq + 3) / 4), dim3(256), 0, stream, anchors, chain_idx, aoff, chain_cnt,
                       qlens, rlen, soff, nq, segs, nseg, caps);
    return hipGetLastError();
}

hipError_t seg_offsets_launch(SegRec *segs, const int64_t *cig_off, int64_t n, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(seg_set_offsets, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, segs, cig_off, n);
    return hipGetLastError();
}


// cigar::delQueryToRefByCigar (src/align/cigar.cpp:212-299) of one query in
// one workgroup of 256: row_scan runs a block scan over the ops giving each
// op's end in the row and start on the query (w[k] = row end << 32 | query
// start, laid out like the CIGAR) and returns the row length (M/S/=/X and D/N
// advance the row, I/H/P do not); *ins (shared) is set when an I op occurs
// (cigar::hasInsertion, :70-79). row_fill then writes row columns [0, limit):
// every column finds its op by binary search (M/S/=/X copy query bases, D/N
// write '-'). An empty CIGAR leaves the query as it is (:214-216).
__device__ int64_t row_scan(const uint32_t *__restrict__ c, int64_t n, int64_t qlen, int64_t *__restrict__ w,
                            int64_t *sh, int *ins)
{
    int64_t carry = 0;
    for (int64_t k0 = 0; k0 < n; k0 += 256) {
        const int64_t k = k0 + threadIdx.x;
        int64_t v = 0;
        if (k < n) {
            const uint32_t op = c[k] & 0xf, len = c[k] >> 4;
            const bool col = !(op == 1 || op == 5 || op == 6);               // advances the row
            const bool qry = !(op == 2 || op == 3 || op == 5 || op == 6);    // advances the query
            v = ((int64_t)(col ? len : 0) << 32) | (qry ? len : 0);
            if (op == 1) *ins = 1;
        }
        int64_t tot = 0;
        const int64_t incl = block_scan_incl(v, sh, &tot) + carry;
        if (k < n) w[k] = (incl & ~0xffffffffll) | ((incl - v) & 0xffffffffll);
        carry += tot;
    }
    __syncthreads();
    return n == 0 ? qlen : (carry >> 32);
}

__device__ void row_fill(const uint32_t *__restrict__ c, int64_t n, const int64_t *__restrict__ w,
                         const char *__restrict__ qs, int64_t qlen, int64_t limit, char *__restrict__ row,
                         const int32_t *__restrict__ map = nullptr)
{
    if (n == 0) {
        for (int64_t p = threadIdx.x; p < qlen && p < limit; p += 256) row[map ? map[p] : p] = qs[p];
        return;
    }
    for (int64_t p = threadIdx.x; p < limit; p += 256) {
        int64_t lo = 0, hi = n;   // first op whose end in the row is > p
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if ((w[mid] >> 32) > p) hi = mid;
            else lo = mid + 1;
        }
        char ch = '-';
        if (lo < n) {
            const uint32_t op = c[lo] & 0xf, len = c[lo] >> 4;
            if (op != 2 && op != 3) {
                const int64_t start = (w[lo] >> 32) - (int64_t)len;
                ch = qs[(w[lo] & 0xffffffffll) + (p - start)];
            }
        }
        row[map ? map[p] : p] = ch;
    }
}

// rows of a batch: query q's row at rows + q * stride, its length in
// row_len[q] (rows longer than stride are cut at stride), hasInsertion in
// has_ins[q]; ws holds the op scan (one int64 per CIGAR word).
__global__ __launch_bounds__(256) void project_rows(const uint32_t *__restrict__ cig, const int64_t *__restrict__ coff,
                                                    const char *__restrict__ qseq, const int64_t *__restrict__ qoff,
                                                    char *__restrict__ rows, int64_t stride,
                                                    int64_t *__restrict__ row_len, int64_t *__restrict__ ws,
                                                    uint8_t *__restrict__ has_ins)
{
    __shared__ int64_t sh[256];
    __shared__ int ins;
    const int q = blockIdx.x;
    const int64_t n = coff[q + 1] - coff[q];
    const int64_t qlen = qoff[q + 1] - qoff[q];
    if (threadIdx.x == 0) ins = 0;
    __syncthreads();
    const int64_t out_len = row_scan(cig + coff[q], n, qlen, ws + coff[q], sh, &ins);
    if (threadIdx.x == 0) {
        has_ins[q] = (uint8_t)ins;
        row_len[q] = out_len;
    }
    row_fill(cig + coff[q], n, ws + coff[q], qseq + qoff[q], qlen, out_len < stride ? out_len : stride,
             rows + (int64_t)q * stride);
}

// Keep-length FASTA records of RefAligner::mergeAlignedResults for one center
// (src/align/ref_aligner.cpp:676-740, keep_length): record q is
// '>' name '\n' row '\n' (seq_io::SeqWriter::writeFasta with an unlimited line
// width, src/utils/seq_io.cpp:236-268), the rows of queries with an insertion
// first (their SAM files are merged first, :504-535 and :713-717), then the
// others (:719-732), each group in input order. Three passes:
//  * rows_scan: the op scan, row length and insertion flag of every query;
//  * record_sizes: sz[q] (insertion group) and sz[n + q] (the others) hold the
//    record's size in its group and 0 in the other (0 in both for a query
//    whose name starts with '@'), sz[2n] = 0, so that one
//    exclusive scan over the 2n + 1 entries gives every record's byte offset;
//    counts[0] += queries with an insertion, counts[1] += rows whose length
//    is not the center's (the reference throws on those, :590-596, :657-663);
//  * write_records: header, row and newline at the offset. The row has the
//    center's length, or with a column map (several references: the row of
//    the chosen reference spread over the columns of its aligned record,
//    padQueryToRefByCigar with that record's M/D CIGAR, ref_aligner.cpp:
//    190-196, 486-496) `width` bytes, '-' where the aligned record has a gap.
// SamReader::next skips SAM lines that start with '@' (src/utils/seq_io.cpp:
// 433-436), so a query named "@..." never reaches the merged FASTA
__device__ __forceinline__ bool is_dropped(const char *names, const int64_t *name_off, int q)
{
    return name_off[q + 1] > name_off[q] && names[name_off[q]] == '@';
}

__global__ __launch_bounds__(256) void rows_scan(const uint32_t *__restrict__ cig, const int64_t *__restrict__ coff,
                                                 const int64_t *__restrict__ qoff, int64_t *__restrict__ row_len,
                                                 int64_t *__restrict__ ws, uint8_t *__restrict__ has_ins)
{
    __shared__ int64_t sh[256];
    __shared__ int ins;
    const int q = blockIdx.x;
    if (threadIdx.x == 0) ins = 0;
    __syncthreads();
    const int64_t out_len = row_scan(cig + coff[q], coff[q + 1] - coff[q], qoff[q + 1] - qoff[q], ws + coff[q], sh,
                                     &ins);
    if (threadIdx.x == 0) {
        has_ins[q] = (uint8_t)ins;
        row_len[q] = out_len;
    }
}

__global__ __launch_bounds__(256) void record_sizes(const char *__restrict__ names,
                                                    const int64_t *__restrict__ name_off,
                                                    const uint8_t *__restrict__ has_ins,
                                                    const int64_t *__restrict__ row_len, int nq, int64_t rlen,
                                                    int64_t width, int64_t *__restrict__ sz,
                                                    unsigned long long *__restrict__ counts)
{
    __shared__ unsigned int n_ins, n_bad;
    if (threadIdx.x == 0) n_ins = n_bad = 0;
    __syncthreads();
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q < nq) {
        const int64_t nl = name_off[q + 1] - name_off[q];
        const bool keep = !is_dropped(names, name_off, q);
        const int64_t size = keep ? nl + width + 3 : 0;
        const bool ins = has_ins[q] != 0;
        sz[q] = ins ? size : 0;
        sz[nq + q] = ins ? 0 : size;
        if (keep && ins) atomicAdd(&n_ins, 1u);
        if (keep && row_len[q] != rlen) atomicAdd(&n_bad, 1u);
    }
    if (q == 0) sz[2 * (int64_t)nq] = 0;
    __syncthreads();
    if (threadIdx.x == 0) {
        if (n_ins) atomicAdd(&counts[0], (unsigned long long)n_ins);
        if (n_bad) atomicAdd(&counts[1], (unsigned long long)n_bad);
    }
}

__global__ __launch_bounds__(256) void write_records(const uint32_t *__restrict__ cig, const int64_t *__restrict__ coff,
                                                     const int64_t *__restrict__ ws, const char *__restrict__ qseq,
                                                     const int64_t *__restrict__ qoff, const char *__restrict__ names,
                                                     const int64_t *__restrict__ name_off,
                                                     const uint8_t *__restrict__ has_ins,
                                                     const int64_t *__restrict__ rec_off, int nq, int64_t rlen,
                                                     const int32_t *__restrict__ map, int64_t width,
                                                     char *__restrict__ out)
{
    const int q = blockIdx.x;
    if (is_dropped(names, name_off, q)) return;
    char *rec = out + (has_ins[q] ? rec_off[q] : rec_off[nq + q]);
    const int64_t nl = name_off[q + 1] - name_off[q];
    const char *nm = names + name_off[q];
    for (int64_t p = threadIdx.x; p < nl + 2; p += 256) rec[p] = p == 0 ? '>' : (p == nl + 1 ? '\n' : nm[p - 1]);
    char *row = rec + nl + 2;
    if (threadIdx.x == 0) row[width] = '\n';
    const int64_t n = coff[q + 1] - coff[q];
    const int64_t qlen = qoff[q + 1] - qoff[q];
    const int64_t have = n == 0 ? qlen : ws[coff[q] + n - 1] >> 32;
    const int64_t fill = have < rlen ? have : rlen;
    if (map) {
        // padQueryToRefByCigar with the reference's aligned row: its gap
        // columns are '-', its bases take the row's characters
        for (int64_t p = threadIdx.x; p < width; p += 256) row[p] = '-';
        __syncthreads();
    } else {
        // a row shorter than the center (reported through record_sizes) is
        // padded so that every byte of the record is written
        for (int64_t p = fill + threadIdx.x; p < rlen; p += 256) row[p] = '-';
    }
    row_fill(cig + coff[q], n, ws + coff[q], qseq + qoff[q], qlen, fill, row, map);
}

hipError_t project_rows_launch(const uint32_t *cig, const int64_t *coff, const char *qseq, const int64_t *qoff,
                               int nq, char *rows, int64_t stride, int64_t *row_len, int64_t *ws, uint8_t *has_ins,
                               hipStream_t stream)
{
    if (nq <= 0) return hipSuccess;
    hipLaunchKernelGGL(project_rows, dim3((unsigned)nq), dim3(256), 0, stream, cig, coff, qseq, qoff, rows, stride,
                       row_len, ws, has_ins);
    return hipGetLastError();
}

hipError_t records_launch(int phase, const uint32_t *cig, const int64_t *coff, const char *qseq, const int64_t *qoff,
                          const char *names, const int64_t *name_off, int nq, int64_t rlen, const int32_t *map,
                          int64_t width, int64_t *row_len, int64_t *ws, uint8_t *has_ins, int64_t *sz,
                          unsigned long long *counts, char *out, hipStream_t stream)
{
    // phase 0: rows_scan; 1: record_sizes into sz (2n + 1 entries); 2:
    // write_records with sz = the exclusive scan of those sizes
    if (nq <= 0) return hipSuccess;
    if (phase == 0)
        hipLaunchKernelGGL(rows_scan, dim3((unsigned)nq), dim3(256), 0, stream, cig, coff, qoff, row_len, ws, has_ins);
    else if (phase == 1)
        hipLaunchKernelGGL(record_sizes, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, stream, names, name_off,
                           has_ins, row_len, nq, rlen, width, sz, counts);
    else
        hipLaunchKernelGGL(write_records, dim3((unsigned)nq), dim3(256), 0, stream, cig, coff, ws, qseq, qoff, names,
                           name_off, has_ins, sz, nq, rlen, map, width, out);
    return hipGetLastError();
}

hipError_t seg_split_launch(SegRec *segs, int64_t n, const uint8_t *center, const int64_t *qrev_off,
                            const int64_t *qlens, const uint8_t *qrev, uint32_t *cigs, int32_t *ncig,
                            int64_t *lane_list, unsigned long long *lane_count, int64_t *host_list,
                            unsigned long long *host_count, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(seg_split, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, segs, n, center, qrev_off,
                       qlens, qrev, cigs, ncig, lane_list, lane_count, host_list, host_count);
    return hipGetLastError();
}


hipError_t mm2_assemble_launch(const SegRec *segs, const int64_t *soff, const int32_t *nseg, const uint32_t *cigs,
                               const int32_t *ncig, const int64_t *qlens, uint32_t rlen, const int64_t *ooff, int nq,
                               uint32_t *out, int64_t *nout, hipStream_t stream)
{
    if (nq <= 0) return hipSuccess;
    hipLaunchKernelGGL(mm2_assemble, dim3((nq + 3) / 4), dim3(256), 0, stream, segs, soff, nseg, cigs, ncig, qlens,
                       rlen, ooff, nq, out, nout);
    return hipGetLastError();
}

}  // namespace hal
