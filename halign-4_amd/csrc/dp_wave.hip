// dp_wave.hip -- ring form of the extz2 DP (CIGAR + score mode) for bands up
// to 64*W*S-2 blocks. Same semantics as dp_kernel.hip (which restates
// src/align/ksw2_extz2_sse.c:23-304, SSE2 branch), organised so that as little
// as possible on the per-anti-diagonal critical path waits on memory or LDS:
//
//  * a workgroup of W waves owns the job; block b lives in slot (b mod R)/NL of
//    lane (b mod R) mod NL (NL = 64*W lanes, R = NL*S blocks); neighbour values
//    move with a DPP wave rotate, and only the first lane of each wave takes
//    its neighbour from LDS (one barrier per anti-diagonal when W > 1);
//  * the query window of a block is a 16-byte register window that slides one
//    base per anti-diagonal (v_alignbyte), fed from the left lane's window; only
//    the first block of a band, whose left neighbour is idle, takes its new base
//    from a scalar load issued a diagonal ahead;
//  * target and query are 0..4 byte codes, four cells per VGPR;
//  * no per-block score array is kept: inside the band the score is recomputed
//    every diagonal, right of the band it is the never-written 0 of the
//    reference's calloc'd array, and left of the band (first block only) the
//    frozen last in-band value is kept in one 8-VGPR set per lane;
//  * the H-row edge (score/mte) is tracked in scalars fed by v_readlane (one
//    wave) or by two LDS bytes read one diagonal later (several waves);
//  * the gap direction (RIGHT/LEFT) is a template parameter, so the cell loop
//    has no runtime branch and no register state merges at a join.
#include <atomic>

#include "dp_device.h"

namespace hal {

// Mask of the cells c < k of a block for each register j (cells j and j+8 in
// the low and high halves), k = 0..16: wave-uniform band-edge masks.
struct HalfMaskTable { uint32_t m[17][8]; };
static constexpr HalfMaskTable make_half_masks()
{
    HalfMaskTable t{};
    for (int k = 0; k <= 16; ++k)
        for (int j = 0; j < 8; ++j)
            t.m[k][j] = (j < k ? 0x0000ffffu : 0u) | (j + 8 < k ? 0xffff0000u : 0u);
    return t;
}
__constant__ HalfMaskTable kHalfMask = make_half_masks();

// cell c of a block held as registers j = c & 7, half c >> 3 (x<<8 form)
__device__ __forceinline__ int cell_byte(const uint32_t (&a)[8], int c)
{
    return (int)((pick8(a, c & 7) >> ((c >> 3) * 16 + 8)) & 0xffu);
}
// Byte of cell c (0..15) of slot k of lane `from`, all three wave-uniform: one
// v_readlane of the register that holds the cell, picked by a uniform switch
// (no per-lane select over the block's registers).
template <int S>
__device__ __forceinline__ int uniform_cell_byte(const uint32_t (&A)[S][8], int k, int c, int from)
{
    uint32_t v = 0;
    switch (k * 8 + (c & 7)) {
#define HAL_RL(kk, jj)                                                                       \
    case (kk) * 8 + (jj):                                                                    \
        if constexpr ((kk) < S) v = __builtin_amdgcn_readlane(A[(kk) < S ? (kk) : 0][jj], from); \
        break;
#define HAL_RL8(kk) HAL_RL(kk, 0) HAL_RL(kk, 1) HAL_RL(kk, 2) HAL_RL(kk, 3) HAL_RL(kk, 4) HAL_RL(kk, 5) \
    HAL_RL(kk, 6) HAL_RL(kk, 7)
        HAL_RL8(0) HAL_RL8(1) HAL_RL8(2) HAL_RL8(3) HAL_RL8(4)
#undef HAL_RL8
#undef HAL_RL
    default:
        break;
    }
    static_assert(S <= 5, "uniform_cell_byte covers up to 5 slots");
    return (int)((v >> ((c >> 3) * 16 + 8)) & 0xffu);
}

// lane i receives lane i-1 of its group of SUB lanes (SUB = 16: DPP row_ror:1;
// SUB = 4: DPP quad_perm [3,0,1,2])
template <int SUB>
__device__ __forceinline__ uint32_t group_ror1(uint32_t v)
{
    static_assert(SUB == 4 || SUB == 16, "groups of 4 or 16 lanes");
    return (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, SUB == 16 ? 0x121 : 0x93, 0xF, 0xF, false);
}

// Backtrack of include/ksw2.h:129-161 for a job owned by a group of SUB lanes
// (sub-wave mode). Every lane of the group walks the same path; traceback rows
// are staged 16 anti-diagonals at a time into the group's LDS chunk (16 x 16
// bytes: along 16 anti-diagonals the path moves at most 15 target columns, so two
// blocks cover it), lane t of the group loading anti-diagonals rc - t - SUB*m.
// Cell layout 1 of backtrack_to_cigar; ops go to J.cigar last op first, written
// by every lane of the group with the same value.
template <int SUB>
__device__ __forceinline__ int group_backtrack(const DpJob &J, const uint8_t *__restrict__ tb, size_t rowb, int w,
                                      bool right, int i0, int j0, uint32_t (*chunk_)[4])
{
    uint32_t(*chunk)[4] = chunk_;   // LDS reads and writes once inlined (see backtrack_to_cigar)
    const int t = threadIdx.x & (SUB - 1);
    const int qlen = J.qlen, tlen = J.tlen;
    int nout = 0, i = i0, j = j0, state = 0, rc = -1, bb0 = 0, cur_op = -1, cur_len = 0;
    auto push = [&](int op, int len) {
        if (op == cur_op) {
            cur_len += len;
            return;
        }
        if (cur_len > 0) {
            if (nout < J.cig_cap) J.cigar[nout] = (uint32_t)cur_len << 4 | (uint32_t)cur_op;
            ++nout;
        }
        cur_op = op;
        cur_len = len;
    };
    while (i >= 0 && j >= 0) {
        const int r = i + j;
        if (rc < 0 || r < rc - 15) {
            rc = r;
            bb0 = (i - 15 > 0 ? i - 15 : 0) >> 4;
            // unconditional loads (row and block clamped into the band, as in
            // backtrack_to_cigar) issued together, then the LDS writes
            uint32_t wv[16 / SUB][4];
#pragma unroll
            for (int m = 0; m < 16 / SUB; ++m) {
                const int rr = rc - t - SUB * m > 0 ? rc - t - SUB * m : 0;
                Bounds br{0, 0, 0, 0};
                diag_bounds(rr, qlen, tlen, w, br);
                const int sb = br.st >> 4, eb = br.en >> 4;
                const uint8_t *row = tb + (size_t)rr * rowb;
#pragma unroll
                for (int q2 = 0; q2 < 2; ++q2) {
                    const int bb = min(max(bb0 + q2, sb), eb);
                    const uint64_t v = __hip_atomic_load((const uint64_t *)(row + (size_t)(bb - sb) * 8),
                                                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    wv[m][2 * q2] = (uint32_t)v;
                    wv[m][2 * q2 + 1] = (uint32_t)(v >> 32);
                }
            }
#pragma unroll
            for (int m = 0; m < 16 / SUB; ++m)
#pragma unroll
                for (int q2 = 0; q2 < 4; ++q2) chunk[t + SUB * m][q2] = wv[m][q2];
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
        }
        Bounds b;
        diag_bounds(r, qlen, tlen, w, b);
        int force = -1, d = 0;
        if (i < b.st) force = 2;
        if (i > b.en) force = 1;
        if (force < 0) {
            const int c = i & 15;
            const int dw = (c & 7) >> 2, sh = (((c & 1) << 1) | (c >> 3)) * 8 + (((c & 7) >> 1) & 1) * 4;
            const uint32_t word = chunk[rc - r][2 * ((i >> 4) - bb0) + dw];
            const uint32_t nib = (word >> sh) & 15u;
            uint32_t fA = (nib >> 3) & 1, fB = (nib >> 2) & 1, fX = (nib >> 1) & 1, fY = nib & 1;
            if (right) { fA ^= 1; fB ^= 1; fX ^= 1; fY ^= 1; }
            d = (int)((fB ? 2u : fA) | (fX << 3) | (fY << 4));
        }
        if (state == 0) state = d & 7;
        else if (!((d >> (state + 2)) & 1)) state = 0;
        if (state == 0) state = d & 7;
        if (force >= 0) state = force;
        if (state == 0) { push(0, 1); --i; --j; }
        else if (state == 1) { push(2, 1); --i; }
        else { push(1, 1); --j; }
    }
    if (i >= 0) push(2, i + 1);
    if (j >= 0) push(1, j + 1);
    push(-1, 0);   // flush the last run
    return nout;
}

// PAD (one-wave shapes of four or more slots, the widest bands: traceback
// slots of ~140 MB, so HBM, not the chip, caps the jobs in flight): more than
// 32 KB of LDS keeps at most four workgroups on a CU, one per SIMD. When their
// registers allow two per SIMD and fewer workgroups than that are resident,
// the dispatcher stacks pairs on some SIMDs while others idle. The
// two-per-SIMD form of these shapes (dp_wave_dense) has no pad.
template <int W, int S, bool PAD = (W == 1 && S >= 4)>
struct WaveSmem {
    uint4 xch[2][W][S];   // last lane of each wave: x, v of cell 15, query window dword 3 (w unused)
    union {
        uint32_t chunk[64][10];       // backtrack staging (one job per workgroup)
        uint32_t group[16][16][4];    // backtrack staging per job (sub-wave mode)
        uint8_t spread[PAD ? 33 * 1024 : 4];
    };
    uint8_t logu[2], logv[2];   // H-edge bytes of the previous anti-diagonal (W > 1)
    int job;
};

// Sub-wave mode (SUB = 4 or 16, W = 1): each group of SUB lanes of the wave
// runs its own job, so one wave advances 64/SUB narrow bands at once. The ring
// is the group's SUB lanes x S slots (neighbours by DPP within the group), band
// bounds are per group, the band-edge masks are computed from the bounds, the
// H-edge bytes come from the owning lane by ds_bpermute one anti-diagonal later,
// and the backtrack is group_backtrack.
// UNI (sub-wave mode only): every group of the wave holds a job with the same
// qlen, tlen and band, so the band bounds, masks and ring positions are
// wave-uniform (scalar) as in the one-wave ring; only the sequences, the
// H-row edge and the backtrack stay per group.
template <int W, int S, bool RT, int SUB = 0, bool UNI = false, bool PAD>
__device__ __forceinline__ void dp_wave_job(const DpJob &J, uint8_t *__restrict__ tb, DpOut *__restrict__ out,
                            WaveSmem<W, S, PAD> &sm)
{
    constexpr bool ROW = SUB != 0;
    constexpr bool ROWB = ROW && !UNI;   // per-group band bounds
    static_assert(!ROW || (W == 1 && (SUB == 4 || SUB == 16)), "sub-wave groups of 4 or 16 lanes in one wave");
    static_assert(!UNI || ROW, "uniform bounds are a sub-wave mode");
    constexpr int NL = ROW ? SUB : 64 * W;   // lanes per job
    constexpr int R = NL * S;                // ring size in blocks
    const int L = ROW ? (int)(threadIdx.x & (NL - 1)) : (int)threadIdx.x, lane = L & 63;
    const int wave = ROW ? 0 : __builtin_amdgcn_readfirstlane(L >> 6);
    const int qlen = UNI ? __builtin_amdgcn_readfirstlane(J.qlen) : J.qlen;
    const int tlen = UNI ? __builtin_amdgcn_readfirstlane(J.tlen) : J.tlen;
    int w = UNI ? __builtin_amdgcn_readfirstlane(J.w) : J.w;
    if (w < 0) w = tlen > qlen ? tlen : qlen;
    // UNI: the N correction of the score runs for every group when any job has
    // an N (it subtracts 0 where no N is involved)
    const bool hasn = UNI ? __builtin_amdgcn_ballot_w64((J.flags & kFlagHasN) != 0) != 0
                          : (J.flags & kFlagHasN) != 0;
    const int ndiag = qlen + tlen - 1;
    int ncol = qlen < tlen ? qlen : tlen;
    ncol = ((ncol < w + 1 ? ncol : w + 1) + 15) / 16 + 1;
    const size_t rowb = (size_t)ncol * 8;
    // one spare ring block: a block entering two diagonals ahead must not reuse the
    // slot of the left neighbour of the next diagonal's first block
    if (ncol > R - 2) {
        if (wave == 0) { out->status = kStTooWide; out->n_cigar = 0; }   // same value from every lane
        return;
    }
    const uint8_t *__restrict__ tgt = J.tgt_b;
    const uint8_t *__restrict__ qr = J.qr_b;

    // Block state: register j holds cells j (low half) and j+8 (high half),
    // int8 values in the high byte of each half. T4/Q4: byte codes, cell c in
    // byte c&3 of dword c>>2.
    uint32_t U[S][8], V[S][8], X[S][8], Y[S][8], T4[S][4], Q4[S][4];
    uint32_t FZ[8];   // frozen scores of the first block (left of the band)
#pragma unroll
    for (int k = 0; k < S; ++k) {
#pragma unroll
        for (int j = 0; j < 8; ++j) U[k][j] = V[k][j] = X[k][j] = Y[k][j] = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) T4[k][j] = Q4[k][j] = 0;
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) FZ[j] = 0x10001000u;

    // block 0 for anti-diagonal 0 lives in lane 0, slot 0 (every lane loads the
    // same words and keeps them on lane 0: no lane-0 branch near the loop head)
    {
        uint32_t t0[4], q0[4];
        load16(tgt, t0);
        load16(qr + (qlen - 1), q0);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            T4[0][j] = L == 0 ? t0[j] : 0u;
            Q4[0][j] = L == 0 ? q0[j] : 0u;
        }
    }

    int last_st = -1, last_en = -1, last_bs = 0;
    int32_t He = kNegInf, Hl = kNegInf;
    int last_en0 = -1;
    int32_t score = kNegInf, mte = kNegInf, mte_q = -1;
    int64_t cells = 0;
    bool zd = false;

    // H-row edge (:227-269): H[en0] = H[en0-1] + u - qe, H[en0-1] += v - qe, fed
    // with u at en0 and v at en0-1 (v at 0 when en0 == 0) of anti-diagonal rr.
    auto edge = [&](int rr, int st0, int en0, int en, int ue, int vl) {
        if (rr == 0) {
            He = vl - 2 * kQE;
        } else if (en0 == 0) {
            He = He + vl - kQE;
        } else {
            const int32_t Hp = (en0 == last_en0) ? Hl : He;
            Hl = (en0 - 1 >= st0) ? Hp + vl - kQE : Hp;
            He = Hp + ue - kQE;
        }
        last_en0 = en0;
        if (en0 == tlen - 1 && He > mte) { mte = He; mte_q = rr - en; }
        if (rr == ndiag - 1 && en0 == tlen - 1) score = He;
    };
    int p_st0 = 0, p_en0 = 0, p_en = 0, rlast = 0;   // bounds of the diagonal whose edge is pending
    uint32_t pu = 0, pv = 0;                          // its edge bytes (row mode)

    // Bounds of r and r+1 are known at the top of diagonal r. Loads for a
    // diagonal (new block windows, the first block's new query base) are
    // issued two diagonals ahead so their latency is hidden by a full diagonal.
    Bounds bd, bd1;
    diag_bounds(0, qlen, tlen, w, bd);
    bool ok1 = ndiag > 1 && diag_bounds(1, qlen, tlen, w, bd1);
    uint32_t fe = 0, fe_next = 0;   // first block's new query base for r and r+1
    if (ok1) {
        const int nbe = bd1.en >> 4;
        if (nbe > (bd.en >> 4)) {
            const int nm = nbe % R;
            if (L == nm % NL) {
                const int kn = nm / NL;
#pragma unroll
                for (int k = 0; k < S; ++k) {
                    if (k == kn) {
                        load16(tgt + 16 * nbe, T4[k]);
                        load16(qr + (qlen - 1 + 16 * nbe), Q4[k]);   // window of diagonal 0, slid at 1
                    }
                }
            }
        }
        if ((bd1.st >> 4) == (bd.st >> 4)) fe_next = load_byte_uniform(qr + (qlen - 2 + (bd1.st & ~15)));
    }

    // Ring bookkeeping kept across anti-diagonals: the band's first block bs
    // only grows, by at most one block per anti-diagonal (st = max(0, r-qlen+1,
    // (r-w+1)>>1)), so blk[k] (this lane's block in slot k, within [bs, bs+R))
    // changes only when bs does, and only for the block that left the band: it
    // comes back R blocks on. bsm = bs mod R, the first block's ring position.
    int blk[S];
#pragma unroll
    for (int k = 0; k < S; ++k) blk[k] = NL * k + L;
    int cur_bs = 0, bsm = 0;

    for (int r = 0; r < ndiag; ++r) {
        const int st0 = bd.st0, en0 = bd.en0, st = bd.st, en = bd.en;
        const int bs = st >> 4, be = en >> 4;
        if (bs != cur_bs) {
#pragma unroll
            for (int k = 0; k < S; ++k) blk[k] += blk[k] < bs ? R : 0;
            bsm = bsm + (bs - cur_bs) < R ? bsm + (bs - cur_bs) : bsm + (bs - cur_bs) - R;
            cur_bs = bs;
        }
        // ring positions of the last block (be) and of the block before it
        const int bem = bsm + (be - bs) < R ? bsm + (be - bs) : bsm + (be - bs) - R;
        const int bem1 = bem > 0 ? bem - 1 : R - 1;
        cells += en0 - st0 + 1;
        // band-edge masks (uniform): cells < st0 of the first block, cells <= en0 of the last
        uint32_t leftm[8], lastm[8];
        if constexpr (ROWB) {
            // per-row bounds: half (c < k) is all ones iff c - k < 0 as int16, for
            // the cell pair (j, j+8) of register j
            const uint32_t kl = (uint32_t)(st0 & 15) * 0x00010001u, ke1 = (uint32_t)((en0 & 15) + 1) * 0x00010001u;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const uint32_t cj = (uint32_t)j | (uint32_t)(j + 8) << 16;
                leftm[j] = pk_ashr15(pk_sub(cj, kl));
                lastm[j] = pk_ashr15(pk_sub(cj, ke1));
            }
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                leftm[j] = kHalfMask.m[st0 & 15][j];
                lastm[j] = kHalfMask.m[(en0 & 15) + 1][j];
            }
        }
        bool act[S];
#pragma unroll
        for (int k = 0; k < S; ++k) act[k] = blk[k] <= be;

        // neighbour values of the previous anti-diagonal: x, v of cell 15 of the left
        // block (high half of register 7) and byte 15 of its query window
        uint32_t nx[S], nv[S], nq[S];
        {
            uint32_t rx[S], rv[S], rq[S];
#pragma unroll
            for (int k = 0; k < S; ++k) {
                if constexpr (ROW) {
                    rx[k] = group_ror1<NL>(X[k][7]);
                    rv[k] = group_ror1<NL>(V[k][7]);
                    rq[k] = group_ror1<NL>(Q4[k][3]);
                } else {
                    rx[k] = wave_ror1(X[k][7]);
                    rv[k] = wave_ror1(V[k][7]);
                    rq[k] = wave_ror1(Q4[k][3]);
                }
            }
            if constexpr (W > 1) {
                const int p = r & 1;
                if (lane == 63) {
#pragma unroll
                    for (int k = 0; k < S; ++k) sm.xch[p][wave][k] = make_uint4(X[k][7], V[k][7], Q4[k][3], 0u);
                }
                __syncthreads();
                // lane 0 of wave i > 0: last lane of wave i-1, same slot; lane 0 of
                // wave 0: last lane of wave W-1, previous slot. Every lane reads the
                // (same) words, one 16-byte read per slot, and lane 0 selects them:
                // no lane-0 branch around the reads.
                const int ws = wave ? wave - 1 : W - 1;
#pragma unroll
                for (int k = 0; k < S; ++k) {
                    const int kp = (k + S - 1) % S;
                    uint4 e = sm.xch[p][ws][wave ? k : kp];
                    asm volatile("" : "+v"(e.x), "+v"(e.y), "+v"(e.z));   // keep the read unconditional
                    nx[k] = lane ? rx[k] : e.x;
                    nv[k] = lane ? rv[k] : e.y;
                    nq[k] = lane ? rq[k] : e.z;
                }
                // the edge bytes of diagonal r-1 were stored before this barrier
                if (r > 0) edge(r - 1, p_st0, p_en0, p_en, sm.logu[p ^ 1], sm.logv[p ^ 1]);
            } else {
#pragma unroll
                for (int k = 0; k < S; ++k) {
                    const int kp = (k + S - 1) % S;
                    nx[k] = lane ? rx[k] : rx[kp];
                    nv[k] = lane ? rv[k] : rv[kp];
                    nq[k] = lane ? rq[k] : rq[kp];
                }
                // row mode: the edge bytes of diagonal r-1 (bpermute issued a diagonal ago)
                if constexpr (ROW) {
                    if (r > 0) edge(r - 1, p_st0, p_en0, p_en, (int)pu, (int)pv);
                }
            }
        }
        // slide every query window one base (branch-free): a block loaded ahead
        // of its entry holds the window of the diagonal it was loaded at and
        // slides with the rest, taking the true query bytes from its left
        // neighbour; windows of blocks outside the band are never read
        if (r > 0) {
#pragma unroll
            for (int k = 0; k < S; ++k) {
                const uint32_t in = (blk[k] == last_bs) ? (fe << 24) : nq[k];
                Q4[k][3] = __builtin_amdgcn_alignbyte(Q4[k][3], Q4[k][2], 3);
                Q4[k][2] = __builtin_amdgcn_alignbyte(Q4[k][2], Q4[k][1], 3);
                Q4[k][1] = __builtin_amdgcn_alignbyte(Q4[k][1], Q4[k][0], 3);
                Q4[k][0] = __builtin_amdgcn_alignbyte(Q4[k][0], in, 3);
            }
        }
        // left boundary of the rounded region (:118-122): the first block's left
        // neighbour x, v are those of (r-1, st-1) when that cell was computed,
        // else 0 (v = q on the top row, st = 0 and r > 0). One wave: the lane
        // holding the first block replaces them in that block's slot with
        // selects; several waves: behind a branch on that lane (the form each
        // measured faster in).
        {
            const int ks = bsm / NL;
            const bool keep = st > 0 && st - 1 >= last_st && st - 1 <= last_en;
            const uint32_t v1 = (st == 0 && r) ? (uint32_t)kQ << 24 : 0u;
            if constexpr (W == 1) {
                const bool own = !keep && L == bsm - ks * NL;
#pragma unroll
                for (int k = 0; k < S; ++k) {
                    const bool mk = own && k == ks;
                    nx[k] = mk ? 0u : nx[k];
                    nv[k] = mk ? v1 : nv[k];
                }
            } else if (L == bsm - ks * NL) {
#pragma unroll
                for (int k = 0; k < S; ++k) {
                    const uint32_t m = keep ? 0u : sel_mask(k, ks);
                    nx[k] = nx[k] & ~m;
                    nv[k] = (nv[k] & ~m) | (v1 & m);
                }
            }
        }
        // top boundary y[r]=0, u[r]=q (:123)
        if (en >= r) {
            const int rb = r >> 4, rc = r & 15;
#pragma unroll
            for (int k = 0; k < S; ++k) {
                if (act[k] && blk[k] == rb) {
                    set_cell(U[k], rc, r ? (uint32_t)kQ << 8 : 0u);
                    set_cell(Y[k], rc, 0u);
                }
            }
        }

        // en0 is in block be; cell tv = en0 - 1 in be or, at a block start, be - 1
        const int ke = bem / NL, le = bem - ke * NL;
        const int tv = en0 > 0 ? en0 - 1 : 0;
        const int tvm = (tv >> 4) == be ? bem : bem1;
        const int kv = tvm / NL, lv = tvm - kv * NL;
        uint32_t ue_v = 0, vl_v = 0;

#pragma unroll
        for (int k = 0; k < S; ++k) {
            if (!act[k]) continue;
            const int b = blk[k];
            uint32_t z[8];
            {
                // fresh s+2(q+e) per byte: table[t^q] = [21,12,12,12,16,16,16,16]
                uint32_t zb[4];
#pragma unroll
                for (int d = 0; d < 4; ++d) zb[d] = __builtin_amdgcn_perm(0x10101010u, 0x0C0C0C15u, T4[k][d] ^ Q4[k][d]);
                if (hasn) {   // one branch per block (N:N scores 0)
#pragma unroll
                    for (int d = 0; d < 4; ++d) zb[d] -= ((T4[k][d] & Q4[k][d] & 0x04040404u) >> 2) * 5u;
                }
                // spread: register j <- cells j and j+8, high byte of each half
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    z[j] = __builtin_amdgcn_perm(zb[(j >> 2) + 2], zb[j >> 2],
                                                 0x040C000Cu + (uint32_t)(j & 3) * 0x01000100u);
                // right of the band: the never-written score 0 (s+16 = 16)
                // (several waves: only slot bem / NL can hold block be)
                if ((W == 1 || k == bem / NL) && b == be) {
#pragma unroll
                    for (int j = 0; j < 8; ++j) z[j] = bfi(lastm[j], z[j], 0x10001000u);
                }
                // left of the band (first block): the frozen last in-band score
                if ((W == 1 || k == bsm / NL) && b == bs) {
#pragma unroll
                    for (int j = 0; j < 8; ++j) z[j] = bfi(leftm[j], FZ[j], z[j]);
#pragma unroll
                    for (int j = 0; j < 8; ++j) FZ[j] = z[j];
                }
            }
            uint32_t C[4];
            const uint32_t x7o = X[k][7], v7o = V[k][7];
#pragma unroll
            for (int m = 3; m >= 0; --m) {
                uint32_t fa[2], fb[2], fx[2], fy[2];
#pragma unroll
                for (int h = 1; h >= 0; --h) {
                    const int j = 2 * m + h;   // descending: X[j-1], V[j-1] are still old
                    // previous-diagonal x, v of cells (j-1, j+7)
                    const uint32_t xs = j ? X[k][j - 1] : cat16(x7o, nx[k]);
                    const uint32_t vs = j ? V[k][j - 1] : cat16(v7o, nv[k]);
                    const uint32_t a = pk_add(xs, vs);
                    const uint32_t ut = U[k][j];
                    const uint32_t bb = pk_add(Y[k][j], ut);
                    const uint32_t z0 = z[j];
                    fa[h] = RT ? pk_subsat(a, z0) : pk_subsat(z0, a);
                    uint32_t zz = pk_maxu(z0, a);
                    fb[h] = RT ? pk_subsat(bb, zz) : pk_subsat(zz, bb);
                    zz = pk_maxu(zz, bb);
                    zz = pk_minu(zz, 0x15001500u);
                    // v before u: the old u (ut) dies before the new one is
                    // written, so both update in place (no register copies)
                    V[k][j] = pk_sub(zz, ut);
                    U[k][j] = pk_sub(zz, vs);
                    zz = pk_sub(zz, 0x06000600u);
                    const uint32_t a2 = pk_sub(a, zz), b2 = pk_sub(bb, zz);
                    X[k][j] = pk_maxi(a2, 0u);
                    Y[k][j] = pk_maxi(b2, 0u);
                    fx[h] = RT ? a2 : pk_subsat(0u, a2);
                    fy[h] = RT ? b2 : pk_subsat(0u, b2);
                }
                // 4 flags of cells (2m, 2m+8, 2m+1, 2m+9): low nibbles for even m, high for odd
                C[m] = (m & 1) ? tb_nibbles<true>(fa[0], fa[1], fb[0], fb[1], fx[0], fx[1], fy[0], fy[1])
                               : tb_nibbles<false>(fa[0], fa[1], fb[0], fb[1], fx[0], fx[1], fy[0], fy[1]);
            }
            uint2 o;
            o.x = bfi(0xF0F0F0F0u, C[1], C[0]);
            o.y = bfi(0xF0F0F0F0u, C[3], C[2]);
            // row base uniform, block offset a 32-bit lane value (a slot is < 4 GB):
            // the store takes the scalar base and the 32-bit offset as they are
            *(uint2 *)(tb + (size_t)r * rowb + (uint32_t)((b - bs) * 8)) = o;
            if constexpr (ROW) {   // per-group bounds: the owning lane's own copy
                if (k == ke) ue_v = (uint32_t)cell_byte(U[k], en0 & 15);
                if (k == kv) vl_v = (uint32_t)cell_byte(V[k], tv & 15);
            }
        }

        if constexpr (ROW) {
            const int base = (int)(threadIdx.x & ~(uint32_t)(NL - 1));
            pu = (uint32_t)__builtin_amdgcn_ds_bpermute((base + le) << 2, (int)ue_v);
            pv = (uint32_t)__builtin_amdgcn_ds_bpermute((base + lv) << 2, (int)vl_v);
            p_st0 = st0; p_en0 = en0; p_en = en;
        } else if constexpr (W == 1) {
            edge(r, st0, en0, en, uniform_cell_byte<S>(U, ke, en0 & 15, le), uniform_cell_byte<S>(V, kv, tv & 15, lv));
        } else {
            // read by every wave after the next barrier; the owning wave writes
            if (wave == (le >> 6)) {
                const int b = uniform_cell_byte<S>(U, ke, en0 & 15, le & 63);
                if (lane == 0) sm.logu[r & 1] = (uint8_t)b;
            }
            if (wave == (lv >> 6)) {
                const int b = uniform_cell_byte<S>(V, kv, tv & 15, lv & 63);
                if (lane == 0) sm.logv[r & 1] = (uint8_t)b;
            }
            p_st0 = st0; p_en0 = en0; p_en = en;
        }
        rlast = r;

        last_st = st; last_en = en; last_bs = bs;
        if (r + 1 < ndiag) {
            if (!ok1) { zd = true; break; }
            bd = bd1;
            fe = fe_next;
            fe_next = 0;
            ok1 = r + 2 < ndiag && diag_bounds(r + 2, qlen, tlen, w, bd1);
            if (ok1) {
                // prefetch for r+2 (bd = bounds of r+1, bd1 = bounds of r+2)
                const int nbs = bd1.st >> 4, nbe = bd1.en >> 4;
                if (nbe > (bd.en >> 4)) {   // a block entering the band at r+2
                    const int nm = nbe % R;
                    if (L == nm % NL) {
                        const int kn = nm / NL;
#pragma unroll
                        for (int k = 0; k < S; ++k) {
                            if (k == kn) {
#pragma unroll
                                for (int j = 0; j < 8; ++j) U[k][j] = V[k][j] = X[k][j] = Y[k][j] = 0;
                                load16(tgt + 16 * nbe, T4[k]);
                                // the window of diagonal r: slid at r+1 and r+2
                                load16(qr + (qlen - 1 - r + 16 * nbe), Q4[k]);
                            }
                        }
                    }
                }
                // the first block keeps its place at r+2: its window's new base from memory
                if (nbs == (bd.st >> 4)) fe_next = load_byte_uniform(qr + (qlen - 1 - (r + 2) + 16 * nbs));
            }
        }
    }

    if constexpr (W > 1) {
        // traceback rows of every wave and the last edge bytes become visible
        __threadfence();
        __syncthreads();
        edge(rlast, p_st0, p_en0, p_en, sm.logu[rlast & 1], sm.logv[rlast & 1]);
    }
    if constexpr (ROW) {
        edge(rlast, p_st0, p_en0, p_en, (int)pu, (int)pv);
        // the row's traceback stores complete before its lanes read each other's rows
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    }

    int nout = 0, status = kStOk;
    if (zd) status = kStNeedStats;
    else if constexpr (ROW) {
        nout = group_backtrack<NL>(J, tb, rowb, w, RT, tlen - 1, qlen - 1, sm.group[threadIdx.x / NL]);
        if (nout > J.cig_cap) status = kStCigarOverflow;
    } else if (wave == 0) {
        nout = backtrack_to_cigar(J, tb, rowb, w, RT, tlen - 1, qlen - 1, sm.chunk, 1);
        if (nout > J.cig_cap) status = kStCigarOverflow;
    }
    // every lane of wave 0 stores the same (uniform) values: no lane-0 branch here
    if (wave == 0) {
        out->score = score; out->n_cigar = nout; out->zdropped = zd ? 1 : 0; out->status = status;
        out->max = 0; out->max_t = -1; out->max_q = -1;
        out->mqe = kNegInf; out->mqe_t = -1;
        out->mte = mte; out->mte_q = mte_q; out->pad_ = 0; out->cells = cells;
    }
}

// Persistent loop: each workgroup pulls jobs from a queue and reuses its own
// traceback slot; all jobs of one launch share the gap direction.
// range (optional, on the device): {first job, job count} within jobs/outs,
// for job lists built on the device.
template <int W, int S, bool RT, bool PAD>
__device__ __forceinline__ void wave_persistent(const DpJob *__restrict__ jobs, int njobs, int *__restrict__ counter,
                                                uint8_t *__restrict__ tb_pool, size_t tb_slot_bytes,
                                                DpOut *__restrict__ outs, const int *__restrict__ range, int prio)
{
    if (prio) __builtin_amdgcn_s_setprio(3);
    if (range) {
        jobs += range[0];
        outs += range[0];
        njobs = range[1];
    }
    __shared__ WaveSmem<W, S, PAD> sm;
    uint8_t *tb = tb_pool + (size_t)blockIdx.x * tb_slot_bytes;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (;;) {
        // Job ticket: every lane of wave 0 executes the atomic (lane 0 adds 1,
        // the others 0) and readfirstlane keeps lane 0's ticket, so no lane-0
        // branch sits at the loop's back edge. Behind a lane-0 branch the
        // compiler may fuse the previous job's lane-0 output writes with the
        // next lane-0 atomic into a path that only lane 0 leaves the job loop
        // by, while the other lanes re-run the old job forever (with one wave
        // per workgroup the barrier is only a wave barrier and does not stop
        // this). Several waves share the ticket through LDS; those barriers
        // also keep the next job's traceback writes behind wave 0's backtrack.
        int jid;
        if constexpr (W == 1) {
            jid = __builtin_amdgcn_readfirstlane(atomicAdd(counter, threadIdx.x == 0 ? 1 : 0));
        } else {
            if (wave == 0) sm.job = __builtin_amdgcn_readfirstlane(atomicAdd(counter, threadIdx.x == 0 ? 1 : 0));
            __syncthreads();
            jid = __builtin_amdgcn_readfirstlane(sm.job);
            __syncthreads();
        }
        if (jid >= njobs) break;
        const DpJob J = jobs[jid];
        dp_wave_job<W, S, RT>(J, tb, &outs[jid], sm);
    }
}

template <int W, int S, bool RT>
__global__ __launch_bounds__(64 * W) void dp_wave_persistent(const DpJob *__restrict__ jobs, int njobs,
                                                             int *__restrict__ counter,
                                                             uint8_t *__restrict__ tb_pool, size_t tb_slot_bytes,
                                                             DpOut *__restrict__ outs, const int *__restrict__ range,
                                                             int prio)
{
    wave_persistent<W, S, RT, (W == 1 && S >= 4)>(jobs, njobs, counter, tb_pool, tb_slot_bytes, outs, range, prio);
}

// Two-per-SIMD form of a one-wave shape: registers capped at 256 so that two
// workgroups share a SIMD, and no LDS pad. One wave per SIMD issues one
// instruction at a time, so the scalar bookkeeping and branches of an
// anti-diagonal stall its VALU; a second wave fills those issue slots (1x5 on
// jobs of band 4,665: 1.45x the cells per second of the one-per-SIMD form with
// every SIMD holding two). It pays only when the traceback budget holds two
// slots per SIMD: with fewer workgroups the dispatcher stacks pairs on some
// SIMDs while others idle.
template <int W, int S, bool RT>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(2))) void dp_wave_dense(
    const DpJob *__restrict__ jobs, int njobs, int *__restrict__ counter, uint8_t *__restrict__ tb_pool,
    size_t tb_slot_bytes, DpOut *__restrict__ outs, const int *__restrict__ range, int prio)
{
    wave_persistent<W, S, RT, false>(jobs, njobs, counter, tb_pool, tb_slot_bytes, outs, range, prio);
}

// Sub-wave persistent kernel: one wave, G = 64/SUB jobs per ticket (group g
// of SUB lanes takes job G*t + g); groups past the last job repeat the last
// job, writing the same values to the same outputs. Group g of workgroup b owns
// traceback slot G*b + g.
template <int SUB, int S, bool RT>
__global__ __launch_bounds__(64) void dp_sub_persistent(const DpJob *__restrict__ jobs, int njobs,
                                                        int *__restrict__ counter, uint8_t *__restrict__ tb_pool,
                                                        size_t tb_slot_bytes, DpOut *__restrict__ outs,
                                                        const int *__restrict__ range, int prio)
{
    if (prio) __builtin_amdgcn_s_setprio(3);
    constexpr int G = 64 / SUB;
    if (range) {
        jobs += range[0];
        outs += range[0];
        njobs = range[1];
    }
    __shared__ WaveSmem<1, S> sm;
    const int g = (int)(threadIdx.x / SUB);
    uint8_t *tb = tb_pool + ((size_t)blockIdx.x * G + (size_t)g) * tb_slot_bytes;
    for (;;) {
        // ticket as in dp_wave_persistent: all lanes run the atomic, lane 0 adds 1
        const int t = __builtin_amdgcn_readfirstlane(atomicAdd(counter, threadIdx.x == 0 ? 1 : 0));
        if (G * t >= njobs) break;
        const int jid = G * t + g < njobs ? G * t + g : njobs - 1;
        const DpJob J = jobs[jid];
        dp_wave_job<1, S, RT, SUB>(J, tb, &outs[jid], sm);
    }
}

// Sub-wave persistent kernel over uniform tickets: the G jobs of every ticket
// share qlen, tlen and band (the host pads each key's jobs to a multiple of G
// with repeats of its last job), so the band bookkeeping runs in scalars.
template <int SUB, int S, bool RT>
__global__ __launch_bounds__(64) void dp_usub_persistent(const DpJob *__restrict__ jobs, int njobs,
                                                         int *__restrict__ counter, uint8_t *__restrict__ tb_pool,
                                                         size_t tb_slot_bytes, DpOut *__restrict__ outs,
                                                         const int *__restrict__ range, int prio)
{
    if (prio) __builtin_amdgcn_s_setprio(3);
    constexpr int G = 64 / SUB;
    if (range) {
        jobs += range[0];
        outs += range[0];
        njobs = range[1];
    }
    __shared__ WaveSmem<1, S> sm;
    const int g = (int)(threadIdx.x / SUB);
    uint8_t *tb = tb_pool + ((size_t)blockIdx.x * G + (size_t)g) * tb_slot_bytes;
    for (;;) {
        const int t = __builtin_amdgcn_readfirstlane(atomicAdd(counter, threadIdx.x == 0 ? 1 : 0));
        if (G * t >= njobs) break;
        const int jid = G * t + g < njobs ? G * t + g : njobs - 1;
        const DpJob J = jobs[jid];
        dp_wave_job<1, S, RT, SUB, true>(J, tb, &outs[jid], sm);
    }
}

// Ring shapes: (W waves, S slots per lane) serve n_col_ <= 64*W*S - 2;
// sub-wave shapes (SUB lanes x S slots per job, 64/SUB jobs per wave) serve
// n_col_ <= SUB*S - 2; query-fixed shapes (G lanes x S slots per job, 64/G
// jobs per wave, dp_qfix.hip) serve full-band jobs of at most 16*G*S query
// bases. The query-fixed shapes follow the ring shapes in the index space.
// Dense shapes: the two-per-SIMD form (dp_wave_dense) of a ring shape, listed
// after the ring shapes with the same (waves, slots).
#define HAL_WAVE_SHAPES(X) \
    X(1, 1) X(1, 2) X(2, 1) X(1, 3) X(3, 1) X(2, 2) X(4, 1) X(5, 1) X(3, 2) X(6, 1) X(4, 2) X(8, 1) X(8, 2) X(16, 1) \
    X(2, 3) X(1, 5) X(2, 4)
#define HAL_DENSE_SHAPES(X) X(1, 5)
#define HAL_SUB_SHAPES(X) X(16, 1) X(16, 2) X(16, 3) X(4, 1) X(4, 2)
#define HAL_QFIX_SHAPES(X) X(1, 1) X(4, 1) X(4, 2) X(16, 1) X(16, 2) X(16, 3) X(64, 1) X(64, 2)

const void *dp_qfix_kernel(int g, int s, bool rt);   // dp_qfix.hip

// waves = 0: sub-wave shape with groups of sub lanes; uni: its uniform-ticket
// form; dense: the two-per-SIMD form of a ring shape; waves = -1: query-fixed
// shape with groups of sub lanes
struct ShapeDesc { int waves, sub, slots, uni, dense; };
static const ShapeDesc kShapes[] = {
#define X(w, s) {w, 0, s, 0, 0},
    HAL_WAVE_SHAPES(X)
#undef X
#define X(w, s) {w, 0, s, 0, 1},
    HAL_DENSE_SHAPES(X)
#undef X
#define X(g, s) {0, g, s, 0, 0},
    HAL_SUB_SHAPES(X)
#undef X
#define X(g, s) {0, g, s, 1, 0},
    HAL_SUB_SHAPES(X)
#undef X
#define X(g, s) {-1, g, s, 0, 0},
    HAL_QFIX_SHAPES(X)
#undef X
};
constexpr int kNumShapes = sizeof(kShapes) / sizeof(kShapes[0]);
constexpr int kNumQfix = 0
#define X(g, s) +1
    HAL_QFIX_SHAPES(X)
#undef X
    ;
constexpr int kNumRing = kNumShapes - kNumQfix;

static const void *wave_kernel(int idx, bool rt)
{
    static const void *const tab[][2] = {
#define X(w, s) {(const void *)&dp_wave_persistent<w, s, false>, (const void *)&dp_wave_persistent<w, s, true>},
        HAL_WAVE_SHAPES(X)
#undef X
#define X(w, s) {(const void *)&dp_wave_dense<w, s, false>, (const void *)&dp_wave_dense<w, s, true>},
        HAL_DENSE_SHAPES(X)
#undef X
#define X(g, s) {(const void *)&dp_sub_persistent<g, s, false>, (const void *)&dp_sub_persistent<g, s, true>},
        HAL_SUB_SHAPES(X)
#undef X
#define X(g, s) {(const void *)&dp_usub_persistent<g, s, false>, (const void *)&dp_usub_persistent<g, s, true>},
        HAL_SUB_SHAPES(X)
#undef X
    };
    if (idx >= kNumRing) return dp_qfix_kernel(kShapes[idx].sub, kShapes[idx].slots, rt);
    return tab[idx][rt ? 1 : 0];
}

static int block_threads(int idx) { return kShapes[idx].waves > 0 ? 64 * kShapes[idx].waves : 64; }

int dp_wave_num_shapes() { return kNumRing; }
int dp_wave_num_qfix() { return kNumQfix; }
int dp_wave_qfix_index(int q) { return kNumRing + q; }
int dp_wave_shape_qfix(int idx) { return kShapes[idx].waves < 0; }
int dp_wave_shape_waves(int idx) { return kShapes[idx].waves; }
int dp_wave_shape_lanes(int idx) { return kShapes[idx].waves > 0 ? 64 * kShapes[idx].waves : kShapes[idx].sub; }
int dp_wave_shape_ring(int idx) { return dp_wave_shape_lanes(idx) * kShapes[idx].slots; }
int dp_wave_jobs_per_group(int idx) { return kShapes[idx].waves > 0 ? 1 : 64 / kShapes[idx].sub; }
int dp_wave_shape_uniform(int idx) { return kShapes[idx].uni; }
int dp_wave_shape_dense(int idx) { return kShapes[idx].dense; }
// 0 plain, 1 uniform-ticket form, 2 two-per-SIMD form
int dp_wave_shape_form(int idx) { return kShapes[idx].uni ? 1 : kShapes[idx].dense ? 2 : 0; }
// the one-per-SIMD form of dense shape idx (itself for any other shape)
int dp_wave_sparse_of(int idx)
{
    const ShapeDesc &d = kShapes[idx];
    if (!d.dense) return idx;
    for (int i = 0; i < kNumRing; ++i)
        if (kShapes[i].waves == d.waves && !kShapes[i].dense && !kShapes[i].uni && kShapes[i].slots == d.slots) return i;
    return idx;
}
// the uniform-ticket form of sub-wave shape idx (itself if it has none)
int dp_wave_uniform_of(int idx)
{
    const ShapeDesc &d = kShapes[idx];
    if (d.waves || d.uni) return idx;
    for (int i = 0; i < kNumShapes; ++i)
        if (!kShapes[i].waves && kShapes[i].uni && kShapes[i].sub == d.sub && kShapes[i].slots == d.slots) return i;
    return idx;
}

// Resident workgroups per CU of each kernel: the occupancy query costs
// microseconds and every planned launch asks, so each answer is kept (the
// device is gfx950 throughout; 0 = not asked yet).
int dp_wave_max_groups(int idx, bool rt, int num_cu)
{
    static std::atomic<int> cache[kNumShapes][2];
    int per_cu = cache[idx][rt ? 1 : 0].load(std::memory_order_relaxed);
    if (per_cu <= 0) {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, wave_kernel(idx, rt), block_threads(idx), 0) !=
                hipSuccess ||
            per_cu <= 0)
            return 0;
        cache[idx][rt ? 1 : 0].store(per_cu, std::memory_order_relaxed);
    }
    return per_cu * num_cu;
}

// tb_pool holds groups * dp_wave_jobs_per_group(idx) slots of tb_slot_bytes;
// range (device, optional) overrides jobs/outs offset and njobs; prio: the
// waves raise their issue priority (s_setprio), for a launch bound by its
// jobs' anti-diagonal chains that shares the chip with other launches
hipError_t dp_wave_launch(int idx, bool rt, int groups, const DpJob *jobs, int njobs, int *counter,
                          uint8_t *tb_pool, size_t tb_slot_bytes, DpOut *outs, hipStream_t stream, const int *range,
                          bool prio)
{
    int pr = prio ? 1 : 0;
    void *args[] = {(void *)&jobs, (void *)&njobs, (void *)&counter, (void *)&tb_pool,
                    (void *)&tb_slot_bytes, (void *)&outs, (void *)&range, (void *)&pr};
    return hipLaunchKernel(wave_kernel(idx, rt), dim3(groups), dim3(block_threads(idx)), args, 0, stream);
}

}  // namespace hal
