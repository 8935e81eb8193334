// dp_host.h -- host-side helpers of the DP path shared by the C-ABI files:
// reference scoring/band rules, the job description and the batched runner.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <vector>

#include "dp_common.h"
#include "host_common.h"

namespace hal {

// ScoreChar2Idx (reference include/align.h:72-89)
inline uint8_t score_code(unsigned char c)
{
    switch (c) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return 4;
    }
}

// align::auto_band(qlen=ref_len, tlen=query_len) (reference include/align.h:159-172)
inline int auto_band(int ref_len, int query_len)
{
    const int d = std::abs(ref_len - query_len);
    const int mx = std::max(ref_len, query_len);
    if ((double)d / (double)mx > 0.5) return -1;
    return 200 + (int)(0.1 * (double)(ref_len + query_len / 2));
}

inline int eff_w(int w, int qlen, int tlen) { return w < 0 ? std::max(qlen, tlen) : w; }

// A launch's traceback slot: its largest job's bytes rounded up to 1/32 of
// their power of two (at least 256 B), so that batches of similar jobs share a
// plan (and its graph)
inline size_t round_slot(size_t bytes)
{
    size_t g = 256;
    while (g * 64 <= bytes) g <<= 1;
    return (bytes + g - 1) / g * g;
}

// ksw2's n_col_ (reference src/align/ksw2_extz2_sse.c:75-76)
inline int n_col(int qlen, int tlen, int w)
{
    int n = std::min(qlen, tlen);
    n = std::min(n, w + 1);
    return (n + 15) / 16 + 1;
}

// Band overflow ("zdropped" break, src/align/ksw2_extz2_sse.c:111-114) happens
// iff some anti-diagonal has st > en. Of the ways st > en can hold, two grow
// monotonically with r (tested on the last anti-diagonal) and one needs w = 0:
// (r-w+1)>>1 > (r+w)>>1 at every odd r, so from r = 1 on. The others cannot
// hold for qlen, tlen >= 1 and w >= 0 (checked exhaustively over small sizes
// by tests/test_host_logic.py).
inline bool predict_zdrop(int qlen, int tlen, int w)
{
    if (w == 0 && qlen + tlen - 1 >= 2) return true;
    const int r = qlen + tlen - 2;
    if (r - qlen + 1 > (r + w) >> 1) return true;
    if ((r - w + 1) >> 1 > tlen - 1) return true;
    return false;
}



// Host-side description of one DP job before its device pointers are known.
struct Spec {
    int64_t tgt_elem, qry_elem;   // element offsets of the even copies (pad included before)
    int64_t tgt_odd_elem, qry_odd_elem;
    int64_t tgt_b, qry_b;         // byte offsets of the 0..4 code copies (pad included before)
    int32_t tlen, qlen, w, flags;
    int64_t cig_off;              // uint32 offset into the device cigar buffer
    int32_t cig_cap;
    int shape;
    bool stats;
    size_t tb_bytes;
};

// Run DP jobs whose sequences sit in ctx->seq (16-bit spread) and ctx->seqb
// (byte codes) and whose CIGARs go to ctx->cig; one DpOut per job.
int run_specs(halign_ctx *ctx, std::vector<Spec> &specs, bool plain, std::vector<DpOut> &outs, double *kernel_ms);
// choose the kernel shape and traceback size of a job; false = band too wide
bool fill_spec_shape(Spec &s, bool full_stats, int policy);
// Reversed query copies for a batch of raw ASCII queries (see halign_capi.hip);
// mean_len sizes the grid (workgroups per query)
hipError_t prep_queries_launch(const char *raw, const int64_t *qoff, const int64_t *region, const int64_t *bregion,
                               int first, int count, int64_t mean_len, uint16_t *seq, uint8_t *seqb,
                               hipStream_t stream);
void spread_target(const uint8_t *codes, int len, uint16_t *even, uint16_t *odd);

// Ring-form kernels (dp_wave.hip): shape table, cost figures, launch
int dp_wave_num_shapes();
int dp_wave_shape_waves(int idx);
int dp_wave_shape_ring(int idx);
int dp_wave_jobs_per_group(int idx);
int dp_wave_shape_lanes(int idx);
int dp_wave_shape_uniform(int idx);   // 1: a sub-wave shape's uniform-ticket form
int dp_wave_uniform_of(int idx);      // the uniform-ticket form of a sub-wave shape (else idx)
int dp_wave_shape_dense(int idx);     // 1: the two-per-SIMD form of a one-wave shape
int dp_wave_shape_form(int idx);      // 0 plain, 1 uniform-ticket form, 2 two-per-SIMD form
int dp_wave_sparse_of(int idx);       // the one-per-SIMD form of a dense shape (else idx)
// query-fixed shapes (dp_qfix.hip) follow the ring shapes: shape index
// dp_wave_qfix_index(q) for q < dp_wave_num_qfix(); their "ring" is G*S blocks
int dp_wave_num_qfix();
int dp_wave_qfix_index(int q);
int dp_wave_shape_qfix(int idx);
int dp_wave_max_groups(int idx, bool rt, int num_cu);
hipError_t dp_wave_launch(int idx, bool rt, int groups, const DpJob *jobs, int njobs, int *counter,
                          uint8_t *tb_pool, size_t tb_slot_bytes, DpOut *outs, hipStream_t stream,
                          const int *range = nullptr, bool prio = false);
// several query-fixed buckets in one launch (dp_qfix.hip); parts in device memory
hipError_t dp_qfix_combined_launch(const QfixPart *parts, int nparts, int groups, hipStream_t stream);
// index of the ring shape with these waves (0: sub-wave), lanes per job and slots; -1 if none
inline int dp_wave_find_shape(int waves, int lanes, int slots)
{
    for (int i = 0; i < dp_wave_num_shapes(); ++i)
        if (dp_wave_shape_waves(i) == waves && dp_wave_shape_lanes(i) == lanes &&
            dp_wave_shape_ring(i) == lanes * slots)
            return i;
    return -1;
}
void bytes_target(const uint8_t *codes, int len, uint8_t *out);

// Planner cost table (tune_api.hip): the table measured when the library was
// built, and a bucket's estimate on ring shape wi (ns): crit = its longest
// job, tp = its share of the chip. false: no measurement for wi.
void tune_load_default(halign_ctx *ctx);
bool tune_cost(const halign_ctx *ctx, int wi, double max_nc, double max_nd, double sum_nd, double njobs,
               double &crit, double &tp);

}  // namespace hal
