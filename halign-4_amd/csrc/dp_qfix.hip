// dp_qfix.hip -- the extz2 DP (CIGAR + score mode) for jobs whose band is the
// whole matrix and whose query is the short side: the anchor-less whole pairs
// of the per-query path (a query of a few hundred bases against the center,
// auto band -1) and the end segments that pair a few query bases with a long
// reference part. Same cells, values and traceback flags as
// src/align/ksw2_extz2_sse.c:23-304 (restated in dp_kernel.hip and dp_wave.hip),
// organised along the other axis:
//
//  * cells of an anti-diagonal are indexed by their query position i (target
//    t = r - i), 16 per block, and the blocks never move: block b (query
//    bases 16b..16b+15) is slot b / G of lane b % G of its job's group of G
//    lanes. A full band needs no ring, no entering or leaving blocks and no
//    band-edge masks: every in-band cell reads only in-band cells of the
//    previous anti-diagonal or the two fixed boundaries (u = q, y = 0 above
//    the first query base; x = 0, v = q left of the first target base);
//  * in this layout x and v come from the same query position of the previous
//    anti-diagonal and u and y from the one before it (the target layout has
//    it the other way round), so the neighbour exchange moves u, y and the
//    target window; the query window of a block is loaded once per job;
//  * the row edge of the H matrix that ksw2 follows for score and mte
//    (:224-262, en0 = min(r, tlen - 1)) is query cell 0 while r < tlen, a
//    fixed register of the group's first lane, and walks down the query only
//    for the last qlen anti-diagonals;
//  * G = 1 runs one job per lane (64 per wave) for queries of at most 16 bases.
//
// The traceback record of a block is the 8-byte nibble record of dp_wave.hip
// (cell c in dword (c&7)>>2, byte ((c&1)<<1)|(c>>3), nibble ((c&7)>>1)&1); row
// r of a job holds all G*S blocks, and the rows of the workgroup's 64/G jobs
// are interleaved (row r of job g at (r * 64/G + g) * rowb in the workgroup's
// region), so one store instruction of the wave writes one contiguous piece
// of memory whatever G is (64 one-lane jobs: 512 bytes) instead of 64/G
// scattered pieces in as many pages. qfix_backtrack walks ksw_backtrack's state
// machine (include/ksw2.h:129-161) over it.
#include "dp_device.h"

namespace hal {

// lane l of a group of G lanes receives lane l-1 (lane 0 receives lane G-1)
template <int G>
__device__ __forceinline__ uint32_t qf_ror1(uint32_t v)
{
    static_assert(G == 1 || G == 4 || G == 16 || G == 64, "groups of 1, 4, 16 or 64 lanes");
    if constexpr (G == 64) return wave_ror1(v);
    else if constexpr (G == 16) return (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x121, 0xF, 0xF, false);
    else if constexpr (G == 4) return (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x93, 0xF, 0xF, false);
    else return v;
}

// byte of cell c of slot k in this lane's registers (k, c per lane)
template <int S>
__device__ __forceinline__ uint32_t lane_cell_byte(const uint32_t (&A)[S][8], int k, int c)
{
    uint32_t v = 0;
#pragma unroll
    for (int kk = 0; kk < S; ++kk)
#pragma unroll
        for (int j = 0; j < 8; ++j) v = (kk == k && j == (c & 7)) ? A[kk][j] : v;
    return (v >> ((c >> 3) * 16 + 8)) & 0xffu;
}

template <int G>
struct QfixStage {
    // anti-diagonals staged per chunk: along ROWS anti-diagonals the path
    // moves at most ROWS - 1 query bases, so with ROWS = 16 two blocks cover
    // it; a one-lane job has a single block
    static constexpr int ROWS = G == 1 ? 32 : 16;
    static constexpr int BW = G == 1 ? 2 : 4;   // dwords per staged row
};

template <int G>
struct QfixSmem {
    uint32_t chunk[64 / G][QfixStage<G>::ROWS][QfixStage<G>::BW];
};

// Backtrack of include/ksw2.h:129-161 (is_rot=1, is_rev=0, no forced states:
// the band is the whole matrix) from (target tlen-1, query qlen-1), run by
// every lane of the job's group on the same path; tb is the job's row 0 and
// rstride the distance between its rows. Ops go to J.cigar last op first;
// returns the op count (nothing past J.cig_cap is written).
template <int G>
__device__ __forceinline__ int qfix_backtrack(const DpJob &J, const uint8_t *__restrict__ tb, size_t rstride, int nblk_row,
                                     bool right, uint32_t (*chunk_)[QfixStage<G>::BW])
{
    constexpr int ROWS = QfixStage<G>::ROWS, BW = QfixStage<G>::BW;
    constexpr int PER = G >= ROWS ? 1 : ROWS / G;   // rows each lane stages
    uint32_t(*chunk)[BW] = chunk_;   // LDS reads and writes once inlined (see backtrack_to_cigar)
    const int t = (int)(threadIdx.x & (G - 1));
    int nout = 0, i = J.tlen - 1, j = J.qlen - 1, state = 0, rc = -1, bb0 = 0, cur_op = -1, cur_len = 0;
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
        if (rc < 0 || r < rc - (ROWS - 1)) {
            rc = r;
            bb0 = BW == 2 ? 0 : (j - (ROWS - 1) > 0 ? j - (ROWS - 1) : 0) >> 4;
            // unconditional loads (row and block clamped: rows before 0 and
            // blocks past the row are never read by the walk) issued together,
            // then the LDS writes
            uint32_t wv[PER][BW];
#pragma unroll
            for (int m = 0; m < PER; ++m) {
                const int rr = rc - t - G * m > 0 ? rc - t - G * m : 0;
                const uint8_t *row = tb + (size_t)rr * rstride;
#pragma unroll
                for (int q2 = 0; q2 < BW / 2; ++q2) {
                    const int bb = min(bb0 + q2, nblk_row - 1);
                    const uint64_t v = __hip_atomic_load((const uint64_t *)(row + (size_t)bb * 8), __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT);
                    wv[m][2 * q2] = (uint32_t)v;
                    wv[m][2 * q2 + 1] = (uint32_t)(v >> 32);
                }
            }
#pragma unroll
            for (int m = 0; m < PER; ++m) {
                const int d = t + G * m;
                if (d < ROWS) {
#pragma unroll
                    for (int q2 = 0; q2 < BW; ++q2) chunk[d][q2] = wv[m][q2];
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
        }
        const int c = j & 15;
        const int dw = (c & 7) >> 2, sh = (((c & 1) << 1) | (c >> 3)) * 8 + (((c & 7) >> 1) & 1) * 4;
        const uint32_t word = chunk[rc - r][2 * ((j >> 4) - bb0) + dw];
        const uint32_t nib = (word >> sh) & 15u;
        uint32_t fA = (nib >> 3) & 1, fB = (nib >> 2) & 1, fX = (nib >> 1) & 1, fY = nib & 1;
        if (right) { fA ^= 1; fB ^= 1; fX ^= 1; fY ^= 1; }
        const int d = (int)((fB ? 2u : fA) | (fX << 3) | (fY << 4));
        if (state == 0) state = d & 7;
        else if (!((d >> (state + 2)) & 1)) state = 0;
        if (state == 0) state = d & 7;
        if (state == 0) { push(0, 1); --i; --j; }
        else if (state == 1) { push(2, 1); --i; }
        else { push(1, 1); --j; }
    }
    if (i >= 0) push(2, i + 1);
    if (j >= 0) push(1, j + 1);
    push(-1, 0);   // flush the last run
    return nout;
}

// One job per group of G lanes (64 / G jobs per wave, all of them started
// together and run to the longest one's last anti-diagonal; rows past a
// shorter job's end land inside the workgroup's region, which holds the
// longest job of the launch for every group, and are never read). tb: this
// job's row 0; its rows are 64/G * G*S*8 bytes apart.
template <int G, int S, bool RT>
__device__ __forceinline__ void qfix_job(const DpJob &J, uint8_t *__restrict__ tb, DpOut *__restrict__ out, QfixSmem<G> &sm)
{
    static_assert(G > 1 || S == 1, "one-lane jobs hold a single block");
    constexpr int NB = G * S;                   // blocks per traceback row
    constexpr uint32_t rowb = (uint32_t)NB * 8;
    constexpr size_t rstride = (size_t)(64 / G) * rowb;
    const int lane = (int)threadIdx.x, l = lane & (G - 1), gbase = lane & ~(G - 1);
    const int qlen = J.qlen, tlen = J.tlen, ndiag = qlen + tlen - 1;
    const int nblk = (qlen + 15) >> 4;
    // wave-wide: the longest job, the shortest target, the longest query
    int ndmax = ndiag, tmin = tlen, qmax = qlen;
    if constexpr (G < 64) {
#pragma unroll
        for (int o = G; o < 64; o <<= 1) {
            ndmax = max(ndmax, __shfl_xor(ndmax, o));
            tmin = min(tmin, __shfl_xor(tmin, o));
            qmax = max(qmax, __shfl_xor(qmax, o));
        }
    }
    ndmax = __builtin_amdgcn_readfirstlane(ndmax);
    tmin = __builtin_amdgcn_readfirstlane(tmin);
    qmax = __builtin_amdgcn_readfirstlane(qmax);
    // the N correction of the score runs for every job when any job has an N
    // (it subtracts 0 where no N is involved)
    const bool hasn = __builtin_amdgcn_ballot_w64((J.flags & kFlagHasN) != 0) != 0;

    uint32_t U[S][8], V[S][8], X[S][8], Y[S][8], T4[S][4], Q4[S][4];
#pragma unroll
    for (int k = 0; k < S; ++k) {
#pragma unroll
        for (int j = 0; j < 8; ++j) U[k][j] = V[k][j] = X[k][j] = Y[k][j] = 0;
        // query bases 16b..16b+15 of block b, forward, from the reversed copy
        // (blocks past the query read its last block: their cells are never in band)
        const int b = min(k * G + l, nblk - 1);
        uint32_t o[4];
        load16(J.qr_b + (qlen - 16 - 16 * b), o);
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            Q4[k][d] = __builtin_bswap32(o[3 - d]);
            T4[k][d] = 0;
        }
    }
    // target bytes entering the first block: tw = target[r0..r0+15] for the
    // current 16 anti-diagonals; the next 16 are fetched 16 anti-diagonals
    // ahead as the five aligned words that hold them (tr; tsh = byte offset)
    // and aligned only when they are needed, so the load is not waited on
    const uint8_t *__restrict__ tgt = J.tgt_b;
    uint32_t tw[4], tr[5];
    load16(tgt, tw);
    auto fetch = [&](int at) {
        const uintptr_t a = (uintptr_t)(tgt + at);
        const uint32_t *d = (const uint32_t *)(a & ~(uintptr_t)3);
#pragma unroll
        for (int i = 0; i < 5; ++i) tr[i] = d[i];
        return (uint32_t)(a & 3);
    };
    uint32_t tsh = fetch(16);

    int32_t He = kNegInf, Hl = kNegInf, score = kNegInf, mte = kNegInf, mte_q = -1;
    int last_en0 = -1;
    uint8_t *__restrict__ trow = tb + (size_t)(l * 8);

    for (int r = 0; r < ndmax; ++r, trow += rstride) {
        // target[r] for cell 0 of the first block
        if ((r & 15) == 0 && r) {
#pragma unroll
            for (int d = 0; d < 4; ++d) tw[d] = __builtin_amdgcn_alignbyte(tr[d + 1], tr[d], tsh);
            tsh = fetch(min(r + 16, tlen + 32));
        }
        const int rs = r & 15;
        const uint32_t tdw = (rs & 8) ? ((rs & 4) ? tw[3] : tw[2]) : ((rs & 4) ? tw[1] : tw[0]);
        const uint32_t tin = tdw << ((3 - (rs & 3)) * 8);

        // previous anti-diagonal's u, y of query cell 15 of the block before
        // (high half of register 7) and its target byte; the first block takes
        // the boundary above the first query base and the new target byte
        uint32_t nu[S], ny[S], nt[S];
        {
            const uint32_t bu = r ? (uint32_t)kQ << 24 : 0u;
            if constexpr (G == 1) {
                nu[0] = bu;
                ny[0] = 0u;
                nt[0] = tin;
            } else {
                uint32_t ru[S], ry[S], rt[S];
#pragma unroll
                for (int k = 0; k < S; ++k) {
                    ru[k] = qf_ror1<G>(U[k][7]);
                    ry[k] = qf_ror1<G>(Y[k][7]);
                    rt[k] = qf_ror1<G>(T4[k][3]);
                }
#pragma unroll
                for (int k = 0; k < S; ++k) {
                    nu[k] = l ? ru[k] : (k ? ru[k ? k - 1 : 0] : bu);
                    ny[k] = l ? ry[k] : (k ? ry[k ? k - 1 : 0] : 0u);
                    nt[k] = l ? rt[k] : (k ? rt[k ? k - 1 : 0] : tin);
                }
            }
        }
        // the cell entering the band at query position r (target 0) reads the
        // boundary left of the first target base: x = 0, v = q (0 at r = 0)
        if (r < 16 * NB) {
            const int bq = r >> 4, ks = bq / G, c = r & 15;
            if (l == bq - ks * G) {
#pragma unroll
                for (int k = 0; k < S; ++k) {
                    if (k == ks) {
                        set_cell(X[k], c, 0u);
                        set_cell(V[k], c, r ? (uint32_t)kQ << 8 : 0u);
                    }
                }
            }
        }

#pragma unroll
        for (int k = 0; k < S; ++k) {
            if (k * G * 16 >= qmax) break;   // no job of the wave reaches this slot
            // slide the target window one base: cell c takes cell c-1's byte
            T4[k][3] = __builtin_amdgcn_alignbyte(T4[k][3], T4[k][2], 3);
            T4[k][2] = __builtin_amdgcn_alignbyte(T4[k][2], T4[k][1], 3);
            T4[k][1] = __builtin_amdgcn_alignbyte(T4[k][1], T4[k][0], 3);
            T4[k][0] = __builtin_amdgcn_alignbyte(T4[k][0], nt[k], 3);
            uint32_t z[8];
            {
                // s + 2(q+e) per byte: table[t^q] = [21,12,12,12,16,16,16,16]
                uint32_t zb[4];
#pragma unroll
                for (int d = 0; d < 4; ++d) zb[d] = __builtin_amdgcn_perm(0x10101010u, 0x0C0C0C15u, T4[k][d] ^ Q4[k][d]);
                if (hasn) {   // N:N scores 0
#pragma unroll
                    for (int d = 0; d < 4; ++d) zb[d] -= ((T4[k][d] & Q4[k][d] & 0x04040404u) >> 2) * 5u;
                }
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    z[j] = __builtin_amdgcn_perm(zb[(j >> 2) + 2], zb[j >> 2],
                                                 0x040C000Cu + (uint32_t)(j & 3) * 0x01000100u);
            }
            uint32_t C[4];
            const uint32_t u7o = U[k][7], y7o = Y[k][7];
#pragma unroll
            for (int m = 3; m >= 0; --m) {
                uint32_t fa[2], fb[2], fx[2], fy[2];
#pragma unroll
                for (int h = 1; h >= 0; --h) {
                    const int j = 2 * m + h;   // descending: U[j-1], Y[j-1] are still old
                    // previous anti-diagonal: x, v of the same cells, u, y of cells (j-1, j+7)
                    const uint32_t xs = X[k][j], vs = V[k][j];
                    const uint32_t us = j ? U[k][j - 1] : cat16(u7o, nu[k]);
                    const uint32_t ys = j ? Y[k][j - 1] : cat16(y7o, ny[k]);
                    const uint32_t a = pk_add(xs, vs);
                    const uint32_t bb = pk_add(ys, us);
                    const uint32_t z0 = z[j];
                    fa[h] = RT ? pk_subsat(a, z0) : pk_subsat(z0, a);
                    uint32_t zz = pk_maxu(z0, a);
                    fb[h] = RT ? pk_subsat(bb, zz) : pk_subsat(zz, bb);
                    zz = pk_maxu(zz, bb);
                    zz = pk_minu(zz, 0x15001500u);
                    U[k][j] = pk_sub(zz, vs);
                    V[k][j] = pk_sub(zz, us);
                    zz = pk_sub(zz, 0x06000600u);
                    const uint32_t a2 = pk_sub(a, zz), b2 = pk_sub(bb, zz);
                    X[k][j] = pk_maxi(a2, 0u);
                    Y[k][j] = pk_maxi(b2, 0u);
                    fx[h] = RT ? a2 : pk_subsat(0u, a2);
                    fy[h] = RT ? b2 : pk_subsat(0u, b2);
                }
                C[m] = (m & 1) ? tb_nibbles<true>(fa[0], fa[1], fb[0], fb[1], fx[0], fx[1], fy[0], fy[1])
                               : tb_nibbles<false>(fa[0], fa[1], fb[0], fb[1], fx[0], fx[1], fy[0], fy[1]);
            }
            uint2 o;
            o.x = bfi(0xF0F0F0F0u, C[1], C[0]);
            o.y = bfi(0xF0F0F0F0u, C[3], C[2]);
            *(uint2 *)(trow + (size_t)(k * G * 8)) = o;
        }

        // H-row edge (:224-262): en0 = min(r, tlen-1); H[en0] takes u at en0
        // and H[en0-1] v at en0-1, i.e. query cells imin = r - en0 and imin+1
        // (cell imin when en0 = 0). While r < tlen that is cells 0 and 1 of the
        // first block (this lane's own registers in the group's first lane).
        {
            const int en0 = r < tlen - 1 ? r : tlen - 1;
            const int st0 = r - qlen + 1 > 0 ? r - qlen + 1 : 0;
            const int imin = r - en0;
            int ue = (int)((U[0][0] >> 8) & 0xffu);
            int vl = (int)(((en0 > 0 ? V[0][1] : V[0][0]) >> 8) & 0xffu);
            if (r >= tmin) {   // some job's edge has left the first query cell
                const int iu = imin, iv = en0 > 0 ? imin + 1 : imin;
                const int bu = iu >> 4, bv = iv >> 4;
                int du = (int)lane_cell_byte<S>(U, bu / G, iu & 15);
                int dv = (int)lane_cell_byte<S>(V, bv / G, iv & 15);
                if constexpr (G > 1) {
                    du = __builtin_amdgcn_ds_bpermute((gbase + (bu & (G - 1))) << 2, du);
                    dv = __builtin_amdgcn_ds_bpermute((gbase + (bv & (G - 1))) << 2, dv);
                }
                if (imin > 0) {
                    ue = du;
                    vl = dv;
                }
            }
            if (r == 0) {
                He = vl - 2 * kQE;
            } else if (en0 == 0) {
                He = He + vl - kQE;
            } else {
                const int32_t Hp = (en0 == last_en0) ? Hl : He;
                Hl = (en0 - 1 >= st0) ? Hp + vl - kQE : Hp;
                He = Hp + ue - kQE;
            }
            last_en0 = en0;
            if (r < ndiag && en0 == tlen - 1) {
                if (He > mte) {
                    mte = He;
                    mte_q = r - ((en0 & ~15) + 15);
                }
                if (r == ndiag - 1) score = He;
            }
        }
    }

    // the rows are complete before the group's lanes read each other's
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    if constexpr (G > 1) {   // the group's first lane followed the edge
        score = __builtin_amdgcn_ds_bpermute(gbase << 2, score);
        mte = __builtin_amdgcn_ds_bpermute(gbase << 2, mte);
        mte_q = __builtin_amdgcn_ds_bpermute(gbase << 2, mte_q);
    }
    const int nout = qfix_backtrack<G>(J, tb, rstride, NB, RT, sm.chunk[lane / G]);
    // every lane of the group stores the same values
    out->score = score; out->n_cigar = nout; out->zdropped = 0;
    out->status = nout > J.cig_cap ? kStCigarOverflow : kStOk;
    out->max = 0; out->max_t = -1; out->max_q = -1;
    out->mqe = kNegInf; out->mqe_t = -1;
    out->mte = mte; out->mte_q = mte_q; out->pad_ = 0;
    out->cells = (int64_t)qlen * tlen;   // every cell of the matrix is in band
}

// Persistent loop of one workgroup: 64/G jobs per ticket (group g of G lanes
// takes job (64/G)*t + g; groups past the last job repeat it, writing the
// same values to the same outputs). Workgroup wg owns the 64/G traceback slots
// from (64/G)*wg on as one region with the groups' rows interleaved. All jobs
// share the gap direction.
template <int G, int S, bool RT>
__device__ __forceinline__ void qfix_run(const DpJob *__restrict__ jobs, int njobs, int *__restrict__ counter,
                                         uint8_t *__restrict__ tb_pool, size_t tb_slot_bytes,
                                         DpOut *__restrict__ outs, int wg, QfixSmem<G> &sm)
{
    constexpr int NG = 64 / G;
    const int g = (int)(threadIdx.x / G);
    uint8_t *tb = tb_pool + (size_t)wg * NG * tb_slot_bytes + (size_t)g * (G * S * 8);
    for (;;) {
        // ticket as in dp_wave_persistent: all lanes run the atomic, lane 0 adds 1
        const int t = __builtin_amdgcn_readfirstlane(atomicAdd(counter, threadIdx.x == 0 ? 1 : 0));
        if (NG * t >= njobs) break;
        const int jid = NG * t + g < njobs ? NG * t + g : njobs - 1;
        const DpJob J = jobs[jid];
        qfix_job<G, S, RT>(J, tb, &outs[jid], sm);
    }
}

template <int G, int S, bool RT>
__global__ __launch_bounds__(64) void dp_qfix_persistent(const DpJob *__restrict__ jobs, int njobs,
                                                         int *__restrict__ counter, uint8_t *__restrict__ tb_pool,
                                                         size_t tb_slot_bytes, DpOut *__restrict__ outs,
                                                         const int *__restrict__ range, int prio)
{
    if (prio) __builtin_amdgcn_s_setprio(3);
    if (range) {
        jobs += range[0];
        outs += range[0];
        njobs = range[1];
    }
    __shared__ QfixSmem<G> sm;
    qfix_run<G, S, RT>(jobs, njobs, counter, tb_pool, tb_slot_bytes, outs, (int)blockIdx.x, sm);
}

// Several query-fixed buckets bound by their jobs' anti-diagonal chains in one
// launch: each workgroup serves the part whose workgroup range holds it, with
// that part's shape, so the chains run side by side instead of one after
// another on a stream (a process has four hardware queues). The waves raise
// their issue priority as the chain-bound single launches do.
__global__ __launch_bounds__(64) void dp_qfix_combined(const QfixPart *__restrict__ parts, int nparts)
{
    __builtin_amdgcn_s_setprio(3);
    __shared__ union {
        QfixSmem<1> g1;
        QfixSmem<4> g4;
        QfixSmem<16> g16;
        QfixSmem<64> g64;
    } sm;
    int p = 0;
    for (int k = 1; k < nparts; ++k)
        if ((int)blockIdx.x >= parts[k].first) p = k;
    const QfixPart P = parts[p];
    const int wg = (int)blockIdx.x - P.first;
#define HAL_QP(idx, gg, ss, mem)                                                                            \
    case 2 * (idx): qfix_run<gg, ss, false>(P.jobs, P.njobs, P.counter, P.tb_pool, P.tb_slot_bytes, P.outs, wg, sm.mem); break; \
    case 2 * (idx) + 1: qfix_run<gg, ss, true>(P.jobs, P.njobs, P.counter, P.tb_pool, P.tb_slot_bytes, P.outs, wg, sm.mem); break;
    switch (P.shape) {
        HAL_QP(0, 1, 1, g1) HAL_QP(1, 4, 1, g4) HAL_QP(2, 4, 2, g4) HAL_QP(3, 16, 1, g16) HAL_QP(4, 16, 2, g16)
        HAL_QP(5, 16, 3, g16) HAL_QP(6, 64, 1, g64) HAL_QP(7, 64, 2, g64)
    default:
        break;
    }
#undef HAL_QP
}

hipError_t dp_qfix_combined_launch(const QfixPart *parts, int nparts, int groups, hipStream_t stream)
{
    void *args[] = {(void *)&parts, (void *)&nparts};
    return hipLaunchKernel((const void *)&dp_qfix_combined, dim3(groups), dim3(64), args, 0, stream);
}

const void *dp_qfix_kernel(int g, int s, bool rt)
{
#define HAL_QF(gg, ss)                                                                              \
    if (g == (gg) && s == (ss))                                                                     \
        return rt ? (const void *)&dp_qfix_persistent<gg, ss, true> : (const void *)&dp_qfix_persistent<gg, ss, false>;
    HAL_QF(1, 1) HAL_QF(4, 1) HAL_QF(4, 2) HAL_QF(16, 1) HAL_QF(16, 2) HAL_QF(16, 3) HAL_QF(64, 1) HAL_QF(64, 2)
#undef HAL_QF
    return nullptr;
}

}  // namespace hal
