// pipeline_common.h -- records shared by the per-query path's kernels and
// their host side: the center index, filter and chain parameters, anchors
// and MM2 segments.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace hal {

// anchor::Anchor (include/anchor.h:39-50), 32 bytes as in the reference's
// layout: hash, rid_ref, pos_ref, rid_qry, pos_qry, span, is_rev.
struct AnchorRec {
    uint64_t hash;
    uint32_t rid_ref, pos_ref, rid_qry, pos_qry, span, is_rev;
};
static_assert(sizeof(AnchorRec) == 32, "anchor record");

// A center's minimizer index on the device: hits sorted by
// (hash, rid, pos, strand) (include/seed.h:68-74), the distinct hashes with
// their first sorted hit and occurrence count, and the reference-side
// occurrence threshold max(u_floor, min(u_ceil, f-cutoff)).
struct CenterIndexDev {
    const uint64_t *sx, *sy;      // sorted hits (MinimizerHit x, y)
    const uint64_t *dh;           // distinct hashes, ascending
    const uint32_t *dstart, *docc;
    int nd;
    uint64_t thr;
};

// SeedFilterParams (include/anchor.h:65-71), the query-side part
struct SeedFilter {
    double q_occ_frac;
    uint64_t sample_every_bp;
};

// ChainParams (include/anchor.h:116-126)
struct ChainParamsDev {
    int32_t max_dist_x, max_dist_y, bw, max_skip, max_iter, min_cnt, min_score;
    float gap_penalty, skip_penalty;
};

// One globalAlignMM2 segment (src/align/align.cpp:221-267): ref [rs, re),
// query [qs, qe) after clipping; flags: gap direction and how it is aligned.
struct SegRec {
    uint32_t rs, re, qs, qe;
    uint32_t flags;   // bit 1: RIGHT (ksw2 flag bit); bit 8: whole pair without a length check
    uint32_t query;   // query index in the batch
    uint32_t cls;     // kSegEmpty / kSegPure / kSegLane / kSegHost / kSegExact / kSegRing
    uint32_t pad;
    int64_t cig_off;  // offset of its CIGAR in the segment CIGAR buffer
};
static_assert(sizeof(SegRec) == 40, "segment record");

constexpr uint32_t kSegRight = 0x02;
constexpr uint32_t kSegRaw = 0x100;   // empty-chain whole pair: CIGAR taken as is

constexpr uint32_t kSegEmpty = 0;     // both sides empty: no CIGAR
constexpr uint32_t kSegPure = 1;      // one side empty: a single I or D
constexpr uint32_t kSegLane = 2;      // small DP (a sub-wave ring job)
constexpr uint32_t kSegHost = 3;      // large DP or band overflow: ring / full-statistics kernels
constexpr uint32_t kSegExact = 4;     // small, equal lengths, identical ACGT bases: CIGAR nM without a DP
constexpr uint32_t kSegRing = 5;      // small DP done by a ring / sub-wave kernel (CIGAR stored last op first)

// small segments as DP jobs built on the device: bucket = shape * 2 + gap
// direction. Shapes 0 .. kLaneRing-1 are sub-wave ring shapes (the first whose
// n_col_ bound holds the band); shapes kLaneRing .. are query-fixed shapes,
// which take the segments whose band is the whole matrix and whose query is
// the short side (the first whose query capacity holds the query). Within a
// bucket the jobs run longest first (a wave runs to its longest job).
constexpr int kLaneRing = 4;
constexpr int kLaneQfix = 3;
constexpr int kLaneShapes = kLaneRing + kLaneQfix;
constexpr int kLaneBuckets = 2 * kLaneShapes;
struct LaneShapeBounds {
    int ncol_max[kLaneRing];   // ring shape k serves n_col_ <= ncol_max[k]
    int qcap[kLaneQfix];       // query-fixed shape q serves queries of <= qcap[q] bases
};

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

// lane l receives lane l-1, lane 0 receives id
__device__ __forceinline__ int32_t wave_shr1_id(int32_t v, int32_t id)
{
    return __builtin_amdgcn_update_dpp(id, v, 0x138, 0xF, 0xF, false);
}

// inclusive scan over the 64 lanes with DPP: row_shr 1, 2, 4, 8 inside each
// row of 16 lanes, then row_bcast:15 into rows 1 and 3 and row_bcast:31 into
// rows 2 and 3; a lane with no source lane takes the identity id
template <class Op>
__device__ __forceinline__ int32_t wave_scan_incl(int32_t v, int32_t id, Op op)
{
    v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x111, 0xF, 0xF, false));
    v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x112, 0xF, 0xF, false));
    v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x114, 0xF, 0xF, false));
    v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x118, 0xF, 0xF, false));
    v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x142, 0xA, 0xF, false));
    v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x143, 0xC, 0xF, false));
    return v;
}

// small-segment limits: target length and traceback bytes (16 per block and
// anti-diagonal; the ring kernels store half of that)
constexpr int kLaneMaxT = 512;
constexpr int kLaneMaxTb = 16384;
// anti-diagonals of a small segment: at most kLaneMaxTb / 16 / n_col_, and
// n_col_ >= 2
constexpr int kLaneNd = kLaneMaxTb / 32;

}  // namespace hal
