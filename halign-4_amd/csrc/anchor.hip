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

// block-wide inclusive scan of one value per thread (256 threads); *total
// receives the sum over the block
__device__ __forceinline__ int64_t block_scan_incl(int64_t v, int64_t *sh, int64_t *total)
{
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1) {
        const int64_t a = t >= d ? sh[t - d] : 0;
        __syncthreads();
        sh[t] += a;
        __syncthreads();
    }
    const int64_t r = sh[t];
    *total = sh[255];
    __syncthreads();
    return r;
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

__global__ __launch_bounds__(256) void anchors_count(CenterIndexDev ci, const uint64_t *__restrict__ qx,
                                                     const uint64_t *__restrict__ qy,
                                                     const int64_t *__restrict__ hit_off,
                                                     const int64_t *__restrict__ hit_cnt, int nq_all, SeedFilter f,
                                                     uint32_t *__restrict__ scratch, int32_t *__restrict__ hidx,
                                                     uint32_t *__restrict__ nanch, int64_t *__restrict__ totals)
{
    __shared__ int64_t sh[256];
    uint32_t *cnt = scratch + (size_t)blockIdx.x * (size_t)ci.nd;
    for (int q = blockIdx.x; q < nq_all; q += gridDim.x) {
        const int64_t base = hit_off[q], n = hit_cnt[q];
        const double limit = f.q_occ_frac * (double)n;
        for (int64_t i = threadIdx.x; i < n; i += 256) {
            const int d = ci.nd > 0 ? find_hash(ci.dh, ci.nd, qx[base + i] >> 8) : -1;
            hidx[base + i] = d;
            if (d >= 0) atomicAdd(&cnt[d], 1u);
        }
        __threadfence();
        __syncthreads();
        int64_t mine = 0;
        for (int64_t i = threadIdx.x; i < n; i += 256) {
            const int d = hidx[base + i];
            uint32_t na = 0;
            if (d >= 0) {
                const uint32_t qocc = __hip_atomic_load(&cnt[d], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
        (void)block_scan_incl(mine, sh, &tot);
        if (threadIdx.x == 0) totals[q] = tot;
        __threadfence();
        __syncthreads();
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

// chainScoreSimple (src/seed/anchor.cpp:254-301); INT_MIN = no link
__device__ __forceinline__ int32_t chain_score(const AnchorRec &ai, const AnchorRec &aj, const ChainParamsDev &cp,
                                               const float *__restrict__ lut)
{
    if (ai.rid_ref != aj.rid_ref) return INT_MIN;
    if (ai.is_rev != aj.is_rev) return INT_MIN;
    const int32_t dq = (int32_t)ai.pos_qry - (int32_t)aj.pos_qry;
    if (dq <= 0 || dq > cp.max_dist_x) return INT_MIN;
    const int32_t dr = (int32_t)ai.pos_ref - (int32_t)aj.pos_ref;
    if (dr <= 0 || dr > cp.max_dist_y) return INT_MIN;
    const int32_t dd = dr > dq ? dr - dq : dq - dr;
    if (dd > cp.bw) return INT_MIN;
    const int32_t dg = dr < dq ? dr : dq;
    const int32_t q_span = (int32_t)aj.span;
    int32_t sc = q_span < dg ? q_span : dg;
    if (dd > 0 || dg > q_span) {
        // each operation rounded on its own, no contraction: the reference's
        // production build has no FMA
        const float lin = __fadd_rn(__fmul_rn(cp.gap_penalty, (float)dd), __fmul_rn(cp.skip_penalty, (float)dg));
        const float lp = dd >= 1 ? lut[dd + 1] : 0.0f;
        sc -= (int32_t)__fadd_rn(lin, __fmul_rn(0.5f, lp));
    }
    return sc;
}

// One lane per query. order: anchors of query q sorted by (pos_ref, pos_qry)
// as indices into `anchors` at [aoff[q], aoff[q] + na[q]); f, pr, tk: int32
// workspaces of the same shape. The best chain goes to chain_idx at aoff[q]
// (anchor indices, in sorted order) and its length to chain_cnt[q].
__global__ void chain_queries(const AnchorRec *__restrict__ anchors, const uint32_t *__restrict__ order,
                              const int64_t *__restrict__ aoff, const int64_t *__restrict__ na_all, int nq_all,
                              ChainParamsDev cp, const float *__restrict__ lut, int32_t *__restrict__ fw,
                              int32_t *__restrict__ pw, int32_t *__restrict__ tw, uint32_t *__restrict__ chain_idx,
                              int32_t *__restrict__ chain_cnt)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq_all) return;
    const int64_t base = aoff[q], n = na_all[q];
    const uint32_t *ord = order + base;
    int32_t *f = fw + base, *p = pw + base, *t = tw + base;
    for (int64_t i = 0; i < n; ++i) t[i] = 0;
    int64_t st = 0, best_end = -1;
    int32_t max_f_global = 0;
    for (int64_t i = 0; i < n; ++i) {
        const AnchorRec ai = anchors[ord[i]];
        int32_t max_f = (int32_t)ai.span;
        int64_t max_j = -1;
        int32_t n_skip = 0;
        while (st < i) {
            const AnchorRec as = anchors[ord[st]];
            if (as.rid_ref != ai.rid_ref || as.is_rev != ai.is_rev ||
                (int32_t)(ai.pos_ref - as.pos_ref) > cp.max_dist_x)
                ++st;
            else
                break;
        }
        const int64_t iter_start = (i - st > cp.max_iter) ? (i - cp.max_iter) : st;
        for (int64_t j = i - 1; j >= iter_start; --j) {
            const int32_t sc = chain_score(ai, anchors[ord[j]], cp, lut);
            if (sc == INT_MIN) continue;
            const int32_t total = f[j] + sc;
            if (total > max_f) {
                max_f = total;
                max_j = j;
                if (n_skip > 0) --n_skip;
            } else if (t[j] == (int32_t)i) {
                if (++n_skip > cp.max_skip) break;
            }
            if (p[j] >= 0) t[p[j]] = (int32_t)i;
        }
        f[i] = max_f;
        p[i] = (int32_t)max_j;
        if (max_f > max_f_global) {
            max_f_global = max_f;
            best_end = i;
        }
    }
    int32_t cnt = 0;
    if (!(max_f_global < cp.min_score || best_end < 0)) {
        for (int64_t cur = best_end; cur >= 0; cur = p[cur]) ++cnt;
        if (cnt < cp.min_cnt) {
            cnt = 0;
        } else {
            int64_t cur = best_end;
            for (int k = cnt - 1; k >= 0; --k) {
                chain_idx[base + k] = ord[cur];
                cur = p[cur];
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

hipError_t anchors_launch_count(const CenterIndexDev &ci, const uint64_t *qx, const uint64_t *qy,
                                const int64_t *hit_off, const int64_t *hit_cnt, int nq, const SeedFilter &f,
                                int groups, uint32_t *scratch, int32_t *hidx, uint32_t *nanch, int64_t *totals,
                                hipStream_t stream)
{
    if (nq <= 0) return hipSuccess;
    hipLaunchKernelGGL(anchors_count, dim3(groups), dim3(256), 0, stream, ci, qx, qy, hit_off, hit_cnt, nq, f, scratch,
                       hidx, nanch, totals);
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
                        int nq, const ChainParamsDev &cp, const float *lut, int32_t *f, int32_t *p, int32_t *t,
                        uint32_t *chain_idx, int32_t *chain_cnt, hipStream_t stream)
{
    if (nq <= 0) return hipSuccess;
    hipLaunchKernelGGL(chain_queries, dim3((nq + 63) / 64), dim3(64), 0, stream, anchors, order, aoff, na, nq, cp, lut,
                       f, p, t, chain_idx, chain_cnt);
    return hipGetLastError();
}

}  // namespace hal
