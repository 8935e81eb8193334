// anchor.hip -- anchors and chains on the device.
//
//  * anchors_count / anchors_write: minimizer::collect_anchors
//    (src/seed/minimizer.cpp:220-356) for many queries against one center
//    index. The center's hits are sorted by (hash, rid, pos, strand) once on
//    the host; the distinct-hash table gives each hash its run [start, occ)
//    and the reference-occurrence threshold (src/seed/anchor.cpp:73-117) is
//    a single number. A workgroup handles one query at a time: it finds each
//    query hit's hash in the table, counts per center hash how often the
//    query holds it (atomics on a per-workgroup scratch array that it
//    restores to zero), applies the query-occurrence and reference-frequency
//    filters, and a block scan places the anchors in the reference's order
//    (query scan order, then sorted reference order).
//
//  * chain_queries: anchor::chainAnchors (src/seed/anchor.cpp:254-461) with
//    one lane per query over its anchors sorted by (pos_ref, pos_qry) (the
//    path has one center and non-canonical hits, so rid_ref = 0 and
//    is_rev = 0 for every anchor). The gap penalty is evaluated with
//    separately rounded float operations in the reference's order and
//    log2f from a table computed by the C library on the host.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

#include "pipeline_common.h"

namespace hal {

__device__ __forceinline__ int find_hash(const uint64_t *__restrict__ dh, int nd, uint64_t h)
{
    int lo = 0, hi = nd - 1;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        const uint64_t v = dh[mid];
        if (v == h) return mid;
        if (v < h) lo = mid + 1;
        else hi = mid - 1;
    }
    return -1;
}

// anchors a query hit expands to (0 if filtered out); d = its center hash index
__device__ __forceinline__ uint32_t hit_anchors(const CenterIndexDev &ci, const SeedFilter &f, int d, uint32_t qocc,
                                                double limit, uint32_t qpos)
{
    if (d < 0) return 0;
    if (f.q_occ_frac > 0.0 && (double)qocc > limit) return 0;   // --q-occ-frac (:306-314)
    const uint32_t rocc = ci.docc[d];
    if ((uint64_t)rocc > ci.thr) {                               // -f/-U with -e sampling (:316-329)
        if (f.sample_every_bp == 0) return 0;
        if ((uint64_t)qpos % f.sample_every_bp != 0) return 0;
    }
    return rocc;
}

// Per query: the center hash index of every hit (binary search), the query's
// own occurrence count of each center hash (block counters, cleared again
// after the query), the anchors each hit expands to and their total.
// MODE 2: hashes and counters in LDS; 1: counters in LDS; 0: counters in the
// block's slice of scratch (global atomics, performed at the L2 of the
// block's XCD: waiting for them before the barrier orders them for the
// block, no cache maintenance needed).
template <int MODE>
__global__ __launch_bounds__(256) void anchors_count(CenterIndexDev ci, const uint64_t *__restrict__ qx,
                                                     const uint64_t *__restrict__ qy,
                                                     const int64_t *__restrict__ hit_off,
                                                     const int64_t *__restrict__ hit_cnt, int nq_all, SeedFilter f,
                                                     uint32_t *__restrict__ scratch, int32_t *__restrict__ hidx,
                                                     uint32_t *__restrict__ nanch, int64_t *__restrict__ totals)
{
    __shared__ int64_t sh[256];
    extern __shared__ uint64_t dyn[];
    const int nd = ci.nd;
    const uint64_t *dh = ci.dh;
    uint32_t *cnt;
    if constexpr (MODE == 2) {
        for (int k = threadIdx.x; k < nd; k += 256) dyn[k] = ci.dh[k];
        dh = dyn;
        cnt = (uint32_t *)(dyn + nd);
    } else if constexpr (MODE == 1) {
        cnt = (uint32_t *)dyn;
    } else {
        cnt = scratch + (size_t)blockIdx.x * (size_t)nd;
    }
    if constexpr (MODE != 0)
        for (int k = threadIdx.x; k < nd; k += 256) cnt[k] = 0;
    __syncthreads();
    for (int q = blockIdx.x; q < nq_all; q += gridDim.x) {
        const int64_t base = hit_off[q], n = hit_cnt[q];
        const double limit = f.q_occ_frac * (double)n;
        for (int64_t i = threadIdx.x; i < n; i += 256) {
            const int d = nd > 0 ? find_hash(dh, nd, qx[base + i] >> 8) : -1;
            hidx[base + i] = d;
            if (d >= 0) atomicAdd(&cnt[d], 1u);
        }
        if constexpr (MODE == 0) __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        int64_t mine = 0;
        for (int64_t i = threadIdx.x; i < n; i += 256) {
            const int d = hidx[base + i];
            uint32_t na = 0;
            if (d >= 0) {
                uint32_t qocc;
                if constexpr (MODE == 0) qocc = __hip_atomic_load(&cnt[d], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                else qocc = cnt[d];
                na = hit_anchors(ci, f, d, qocc, limit, (uint32_t)(qy[base + i] & 0xffffffffu));
            }
            nanch[base + i] = na;
            mine += na;
        }
        __syncthreads();
        for (int64_t i = threadIdx.x; i < n; i += 256) {
            const int d = hidx[base + i];
            if (d >= 0) atomicSub(&cnt[d], 1u);
        }
        int64_t tot = 0;
        if constexpr (MODE == 0) __builtin_amdgcn_s_waitcnt(0);
        (void)block_scan_incl(mine, sh, &tot);   // its barriers order the clears before the next query
        if (threadIdx.x == 0) totals[q] = tot;
    }
}

__global__ __launch_bounds__(256) void anchors_write(CenterIndexDev ci, const uint64_t *__restrict__ qx,
                                                     const uint64_t *__restrict__ qy,
                                                     const int64_t *__restrict__ hit_off,
                                                     const int64_t *__restrict__ hit_cnt, int nq_all,
                                                     const int32_t *__restrict__ hidx,
                                                     const uint32_t *__restrict__ nanch,
                                                     const int64_t *__restrict__ aoff, AnchorRec *__restrict__ out)
{
    __shared__ int64_t sh[256];
    for (int q = blockIdx.x; q < nq_all; q += gridDim.x) {
        const int64_t base = hit_off[q], n = hit_cnt[q];
        int64_t carry = aoff[q];
        for (int64_t c0 = 0; c0 < n; c0 += 256) {
            const int64_t i = c0 + threadIdx.x;
            const uint32_t na = i < n ? nanch[base + i] : 0;
            int64_t chunk = 0;
            const int64_t incl = block_scan_incl(na, sh, &chunk);
            const int64_t at = carry + incl - na;
            if (na) {
                const uint64_t x = qx[base + i], y = qy[base + i];
                const int d = hidx[base + i];
                const uint32_t qrid = (uint32_t)(y >> 32) & 0x7fffffffu, qpos = (uint32_t)(y & 0xffffffffu);
                const uint32_t qstrand = (uint32_t)(y >> 63), qspan = (uint32_t)(x & 0xff);
                const uint32_t s0 = ci.dstart[d];
                for (uint32_t r = 0; r < na; ++r) {
                    const uint64_t rx = ci.sx[s0 + r], ry = ci.sy[s0 + r];
                    const uint32_t rspan = (uint32_t)(rx & 0xff);
                    AnchorRec a;
                    a.hash = x >> 8;
                    a.rid_ref = (uint32_t)(ry >> 32) & 0x7fffffffu;
                    a.pos_ref = (uint32_t)(ry & 0xffffffffu);
                    a.rid_qry = qrid;
                    a.pos_qry = qpos;
                    a.span = rspan < qspan ? rspan : qspan;
                    a.is_rev = ((uint32_t)(ry >> 63) != qstrand) ? 1u : 0u;
                    out[at + r] = a;
                }
            }
            carry += chunk;
        }
    }
}

// Chaining score of predecessor j for anchor i (chainScoreSimple,
// src/seed/anchor.cpp:232-262): fields of j, anchor i in scalars, the log2
// table in LDS. INT_MIN: j cannot precede i. Branch-free: every lane
// evaluates the penalty (the table index clamped) and the tests select.
__device__ __forceinline__ int32_t chain_score(uint32_t i_rr, uint32_t i_pr, uint32_t i_pq, uint32_t j_rr,
                                               uint32_t j_pr, uint32_t j_pq, uint32_t j_span,
                                               const ChainParamsDev &cp, const float *lut)
{
#pragma clang fp contract(off)
    // each operation rounded on its own, no fused multiply-add: the
    // reference's production build has none
    const int32_t dq = (int32_t)i_pq - (int32_t)j_pq;
    const int32_t dr = (int32_t)i_pr - (int32_t)j_pr;
    const int32_t dd = dr > dq ? dr - dq : dq - dr;
    const int32_t dg = dr < dq ? dr : dq;
    const bool ok = i_rr == j_rr && dq > 0 && dq <= cp.max_dist_x && dr > 0 && dr <= cp.max_dist_y && dd <= cp.bw;
    const int32_t q_span = (int32_t)j_span;
    int32_t sc = q_span < dg ? q_span : dg;
    const int32_t di = ok && dd >= 1 ? dd + 1 : 0;   // lut[0] = 0
    const float lin = cp.gap_penalty * (float)dd + cp.skip_penalty * (float)dg;
    const float pen = lin + 0.5f * lut[di];
    if (dd > 0 || dg > q_span) sc -= (int32_t)pen;
    return ok ? sc : INT_MIN;
}

// lane l receives lane l-1 (lane 0 keeps its own value): DPP wave_shr:1
__device__ __forceinline__ int32_t wave_shr1(int32_t v)
{
    return __builtin_amdgcn_update_dpp(v, v, 0x138, 0xF, 0xF, false);
}

constexpr int kChainMaxIter = 8192;   // bound on ChainParams::max_iter (LDS mark bitmap)
constexpr int kChainLut = 502;        // log2 table entries: bw <= 500

// chainAnchors (src/seed/anchor.cpp:264-330), one wave per query. order:
// anchors of query q sorted by (pos_ref, pos_qry), as indices into `anchors`
// at [aoff[q], aoff[q] + na[q]); f, pr: int32 workspaces of the same shape.
// The best chain goes to chain_idx at aoff[q] (anchor indices, in sorted
// order) and its length to chain_cnt[q].
//
// The predecessor scan of anchor i (j = i-1 down to iter_start, stopping after
// max_skip skipped predecessors) runs 64 candidates at a time: lane l takes
// j = jtop - l. A candidate "improves" iff its total beats the running maximum
// of the candidates before it (an exclusive prefix max over lanes); the skip
// counter then runs over the improving / skipped candidates of the chunk in
// order (ballot bit masks, scalar loop), which gives the break position, and
// the last improving candidate before it gives f[i] and pr[i]. The reference's
// mark t[pr[j]] = i, set by every scanned candidate j with a score, decides
// "skipped" for a later candidate: t[j] == i iff some scored candidate in
// (j, i-1] has predecessor j, so the marks of a whole chunk may be set before
// any of its candidates are tested (a candidate below the break marks only
// indices below the break). They live in an LDS bitmap over the scan window.
// The 64 most recent anchors (fields, f, pr) ride in a register ring shifted one
// lane per anchor; older candidates (chunks after the first) come from memory.
__global__ __launch_bounds__(64) void chain_wave(const AnchorRec *__restrict__ anchors,
                                                 const uint32_t *__restrict__ order, const int64_t *__restrict__ aoff,
                                                 const int64_t *__restrict__ na_all, ChainParamsDev cp,
                                                 const float *__restrict__ lut, int32_t *__restrict__ fw,
                                                 int32_t *__restrict__ pw, uint32_t *__restrict__ chain_idx,
                                                 int32_t *__restrict__ chain_cnt)
{
    __shared__ uint32_t marks[kChainMaxIter / 32];
    __shared__ float slut[kChainLut];
    const int q = blockIdx.x;
    const int lane = (int)threadIdx.x;
    for (int k = lane; k < kChainLut; k += 64) slut[k] = lut[k];
    __syncthreads();
    const int64_t base = aoff[q];
    const int n = (int)na_all[q];
    const uint32_t *ord = order + base;
    int32_t *f = fw + base, *pr = pw + base;

    // ring: lane l holds anchor i-1-l at the top of iteration i
    uint32_t r_pr = 0, r_pq = 0, r_sp = 0, r_rr = 0;   // r_rr = rid_ref * 2 + is_rev
    int32_t r_f = 0, r_p = -1;
    // st window: lane l holds anchor st_base + l (pos_ref, rid_ref * 2 + is_rev)
    int st = 0, st_base = 0;
    uint32_t s_pr = 0, s_rr = 0;
    auto load_st = [&]() {
        const int k = st_base + lane;
        if (k < n) {
            const AnchorRec a = anchors[ord[k]];
            s_pr = a.pos_ref;
            s_rr = a.rid_ref * 2u + a.is_rev;
        }
    };
    load_st();
    // ord block: lane l holds ord[ob + l]
    int ob = 0;
    uint32_t o_v = lane < n ? ord[lane] : 0u;
    int32_t max_f_global = 0;
    int best_end = -1;

    for (int i = 0; i < n; ++i) {
        if (i - ob >= 64) {
            ob += 64;
            o_v = ob + lane < n ? ord[ob + lane] : 0u;
        }
        const uint32_t oi = __builtin_amdgcn_readlane(o_v, i - ob);
        const AnchorRec ai = anchors[oi];
        const uint32_t ai_rr = ai.rid_ref * 2u + ai.is_rev;

        // st: first anchor from st on that is in range of i (the reference's while loop)
        for (;;) {
            const int k = st_base + lane;
            const bool cand = k >= st && k < i;
            const bool out = s_rr != ai_rr || (int32_t)(ai.pos_ref - s_pr) > cp.max_dist_x;
            const uint64_t stop = __ballot(cand && !out);
            if (stop) {
                st = st_base + (int)__builtin_ctzll(stop);
                break;
            }
            const int lim = i < st_base + 64 ? i : st_base + 64;
            if (st < lim) st = lim;
            if (st >= i) break;
            st_base = st;
            load_st();
        }
        const int iter_start = (i - st > cp.max_iter) ? i - cp.max_iter : st;

        int32_t max_f = (int32_t)ai.span;
        int max_j = -1, n_skip = 0;
        const int wdw = (i - iter_start + 31) >> 5;   // bitmap words of the window
        for (int d = lane; d < wdw; d += 64) marks[d] = 0u;
        for (int jtop = i - 1; jtop >= iter_start; jtop -= 64) {
            const int j = jtop - lane;
            const bool inwin = j >= iter_start;
            uint32_t a_pr, a_pq, a_sp, a_rr;
            int32_t a_f, a_p;
            if (jtop == i - 1) {
                a_pr = r_pr; a_pq = r_pq; a_sp = r_sp; a_rr = r_rr; a_f = r_f; a_p = r_p;
            } else {
                // values written at least 64 anchors ago; complete this wave's stores first
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                a_pr = a_pq = a_sp = a_rr = 0;
                a_f = 0; a_p = -1;
                if (inwin) {
                    const AnchorRec aj = anchors[ord[j]];
                    a_pr = aj.pos_ref; a_pq = aj.pos_qry; a_sp = aj.span; a_rr = aj.rid_ref * 2u + aj.is_rev;
                    a_f = __hip_atomic_load(f + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    a_p = __hip_atomic_load(pr + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
            const int32_t sc = chain_score(ai_rr, ai.pos_ref, ai.pos_qry, a_rr, a_pr, a_pq, a_sp, cp, slut);
            const bool valid = inwin && sc != INT_MIN;
            const int32_t total = valid ? a_f + sc : INT_MIN;
            if (valid && a_p >= iter_start) {
                const int bt = i - 1 - a_p;
                atomicOr(&marks[bt >> 5], 1u << (bt & 31));
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
            const int bj = inwin ? i - 1 - j : 0;
            const bool marked = inwin && ((marks[bj >> 5] >> (bj & 31)) & 1u);
            // exclusive prefix max of the scored totals over the lanes before
            const auto mx = [](int32_t x, int32_t y) { return x > y ? x : y; };
            const auto mn = [](int32_t x, int32_t y) { return x < y ? x : y; };
            const auto add = [](int32_t x, int32_t y) { return x + y; };
            const int32_t prev = wave_shr1_id(wave_scan_incl(total, INT_MIN, mx), INT_MIN);
            const int32_t before = prev > max_f ? prev : max_f;
            const bool imp = valid && total > before;
            const bool skp = valid && !imp && marked;
            const uint64_t I = __ballot(imp);
            // The skip counter over the candidates in lane order: an improving
            // one lowers it (not below 0), a skipped one raises it and breaks
            // the scan when it passes max_skip. With step d_l = -1 / +1 / 0 and
            // S_l the inclusive sum of the steps, the counter after lane l is
            // S_l - min(-n_skip, min over lanes m <= l of S_m) (the reflected
            // walk n_l = max(n_(l-1) + d_l, 0)); the first skipped lane with a
            // counter above max_skip is the break.
            const int32_t S = wave_scan_incl(imp ? -1 : (skp ? 1 : 0), 0, add);
            const int32_t P = wave_scan_incl(S, INT_MAX, mn);
            const int32_t cnt_l = S - (P < -n_skip ? P : -n_skip);
            const uint64_t B = __ballot(skp && cnt_l > cp.max_skip);
            const int brk = B ? (int)__builtin_ctzll(B) : 64;
            if (!B) n_skip = __builtin_amdgcn_readlane(cnt_l, 63);
            const uint64_t Iv = brk == 64 ? I : (I & ((1ull << brk) - 1ull));
            if (Iv) {
                const int lb = 63 - (int)__builtin_clzll(Iv);
                max_f = __builtin_amdgcn_readlane(total, lb);
                max_j = jtop - lb;
            }
            if (brk < 64) break;
        }
        f[i] = max_f;   // every lane stores the same value
        pr[i] = max_j;
        if (max_f > max_f_global) {
            max_f_global = max_f;
            best_end = i;
        }
        // shift the ring by one lane; anchor i enters at lane 0
        r_pr = (uint32_t)wave_shr1((int32_t)r_pr);
        r_pq = (uint32_t)wave_shr1((int32_t)r_pq);
        r_sp = (uint32_t)wave_shr1((int32_t)r_sp);
        r_rr = (uint32_t)wave_shr1((int32_t)r_rr);
        r_f = wave_shr1(r_f);
        r_p = wave_shr1(r_p);
        const bool l0 = lane == 0;
        r_pr = l0 ? ai.pos_ref : r_pr;
        r_pq = l0 ? ai.pos_qry : r_pq;
        r_sp = l0 ? ai.span : r_sp;
        r_rr = l0 ? ai_rr : r_rr;
        r_f = l0 ? max_f : r_f;
        r_p = l0 ? max_j : r_p;
    }

    // the chain: predecessor links from best_end, read 64 at a time
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    int cnt = 0;
    if (!(max_f_global < cp.min_score || best_end < 0)) {
        int wb = -1;
        int32_t pv = 0;
        for (int cur = best_end; cur >= 0;) {
            if (wb < 0 || cur < wb) {
                wb = cur - 63 > 0 ? cur - 63 : 0;
                pv = wb + lane <= cur ? __hip_atomic_load(pr + wb + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                      : -1;
            }
            ++cnt;
            cur = __builtin_amdgcn_readlane(pv, cur - wb);
        }
        if (cnt < cp.min_cnt) {
            cnt = 0;
        } else {
            wb = -1;
            uint32_t ov = 0;
            int k = cnt - 1;
            for (int cur = best_end; cur >= 0; --k) {
                if (wb < 0 || cur < wb) {
                    wb = cur - 63 > 0 ? cur - 63 : 0;
                    const bool in = wb + lane <= cur;
                    pv = in ? __hip_atomic_load(pr + wb + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : -1;
                    ov = in ? ord[wb + lane] : 0u;
                }
                chain_idx[base + k] = __builtin_amdgcn_readlane(ov, cur - wb);   // every lane, same value
                cur = __builtin_amdgcn_readlane(pv, cur - wb);
            }
        }
    }
    chain_cnt[q] = cnt;
}

// sort keys of the anchors: (pos_ref, pos_qry); values: anchor index
__global__ void anchor_keys(const AnchorRec *__restrict__ anchors, int64_t n, uint64_t *__restrict__ keys,
                            uint32_t *__restrict__ vals)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys[i] = ((uint64_t)anchors[i].pos_ref << 32) | anchors[i].pos_qry;
    vals[i] = (uint32_t)i;
}

constexpr size_t kAnchorLds = 62 * 1024;   // dynamic LDS of anchors_count (64 KiB with its scan array)

hipError_t anchors_launch_count(const CenterIndexDev &ci, const uint64_t *qx, const uint64_t *qy,
                                const int64_t *hit_off, const int64_t *hit_cnt, int nq, const SeedFilter &f,
                                int groups, uint32_t *scratch, int32_t *hidx, uint32_t *nanch, int64_t *totals,
                                hipStream_t stream)
{
    if (nq <= 0) return hipSuccess;
    const size_t nd = (size_t)(ci.nd > 0 ? ci.nd : 0);
    if (nd * 12 <= kAnchorLds)
        hipLaunchKernelGGL(anchors_count<2>, dim3(groups), dim3(256), nd * 12, stream, ci, qx, qy, hit_off, hit_cnt,
                           nq, f, scratch, hidx, nanch, totals);
    else if (nd * 4 <= kAnchorLds)
        hipLaunchKernelGGL(anchors_count<1>, dim3(groups), dim3(256), nd * 4, stream, ci, qx, qy, hit_off, hit_cnt, nq,
                           f, scratch, hidx, nanch, totals);
    else
        hipLaunchKernelGGL(anchors_count<0>, dim3(groups), dim3(256), 0, stream, ci, qx, qy, hit_off, hit_cnt, nq, f,
                           scratch, hidx, nanch, totals);
    return hipGetLastError();
}

hipError_t anchors_launch_write(const CenterIndexDev &ci, const uint64_t *qx, const uint64_t *qy,
                                const int64_t *hit_off, const int64_t *hit_cnt, int nq, int groups,
                                const int32_t *hidx, const uint32_t *nanch, const int64_t *aoff, AnchorRec *out,
                                hipStream_t stream)
{
    if (nq <= 0) return hipSuccess;
    hipLaunchKernelGGL(anchors_write, dim3(groups), dim3(256), 0, stream, ci, qx, qy, hit_off, hit_cnt, nq, hidx,
                       nanch, aoff, out);
    return hipGetLastError();
}

hipError_t anchor_keys_launch(const AnchorRec *anchors, int64_t n, uint64_t *keys, uint32_t *vals,
                              hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(anchor_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, anchors, n, keys, vals);
    return hipGetLastError();
}

hipError_t chain_launch(const AnchorRec *anchors, const uint32_t *order, const int64_t *aoff, const int64_t *na,
                        int nq, const ChainParamsDev &cp, const float *lut, int32_t *f, int32_t *p, uint32_t *chain_idx,
                        int32_t *chain_cnt, hipStream_t stream)
{
    if (nq <= 0) return hipSuccess;
    if (cp.max_iter < 0 || cp.max_iter > kChainMaxIter || cp.bw > kChainLut - 2) return hipErrorInvalidValue;
    hipLaunchKernelGGL(chain_wave, dim3(nq), dim3(64), 0, stream, anchors, order, aoff, na, cp, lut, f, p, chain_idx,
                       chain_cnt);
    return hipGetLastError();
}

}  // namespace hal
