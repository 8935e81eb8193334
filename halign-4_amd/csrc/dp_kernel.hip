// dp_kernel.hip -- banded affine-gap global DP with ksw2 extz2 semantics on
// CDNA4 (gfx950), plus device backtrack to a CIGAR.
//
// Restates src/align/ksw2_extz2_sse.c:23-304 (SSE2 branch; production build
// has no -march) and ksw_backtrack include/ksw2.h:129-161 for the path's
// fixed parameters: dna5 matrix (+5/-4/0 for N), q=6, e=2, zdrop<0,
// end_bonus=0, flags GENERIC_SC with or without RIGHT.
//
// Work decomposition. One workgroup of W waves owns one DP job. The
// reference walks anti-diagonals r and, on each, a region of target columns
// [st, en] rounded out to 16-cell blocks; cells outside the exact band read
// whatever stale values the persistent per-column arrays hold. Here every lane
// owns 16-cell blocks: block b lives in slot (b mod R) / (64W) of lane
// (b mod R) mod (64W), R = 64*W*S, so the active blocks of a diagonal are a
// contiguous run of the ring and the ring slides with the band without any
// data movement. A lane's block state (u, v, x, y, score, target, query
// window) stays in VGPRs for the whole life of the block, which reproduces
// the reference's persistent arrays, stale margin reads included.
//
// Arithmetic. Each 32-bit VGPR holds two cells as 16-bit halves, the int8
// value in the HIGH byte and 0 in the low byte ("x<<8"). With that layout the
// packed 16-bit VALU ops give the reference's 8-bit lane semantics exactly:
// v_pk_add/sub_u16 wrap modulo 256 in the high byte, v_pk_max/min_u16 are the
// unsigned byte max/min, v_pk_max_i16 the signed one, and a saturating
// v_pk_sub_i16 puts the signed compare result in bit 15. No masking or
// sign-extension op is needed anywhere in the recurrence.
//
// Traceback: 4 bits per cell (two state-compare bits and the two gap
// continuation bits), 8 bytes per block per anti-diagonal, rows of
// n_col_*8 bytes (n_col_ exactly as the reference sizes its matrix).
#include <atomic>
#include <climits>

#include "dp_device.h"

namespace hal {

template <int W, int S>
struct Smem {
    uint32_t xch[2][W][S][3];
    int32_t logH[64][2];
    uint32_t chunk[64][10];
    int32_t red[64 * W][3];
    uint8_t logu[64], logv[64];
    int job;
};

// One DP job on one workgroup. Every thread keeps the uniform (per-diagonal)
// state redundantly; all branches on it are uniform.
template <int W, int S, bool STATS>
__device__ void dp_job(const DpJob &J, uint8_t *__restrict__ tb, DpOut *__restrict__ out,
                       Smem<W, S> &sm)
{
    constexpr int NL = 64 * W;
    constexpr int R = NL * S;
    const int L = threadIdx.x, lane = L & 63;
    const int wave = __builtin_amdgcn_readfirstlane(L >> 6);
    const int qlen = J.qlen, tlen = J.tlen;
    int w = J.w;
    if (w < 0) w = tlen > qlen ? tlen : qlen;
    const bool right = (J.flags & kFlagRight) != 0;
    const bool hasn = (J.flags & kFlagHasN) != 0;
    const int ndiag = qlen + tlen - 1;
    int ncol = qlen < tlen ? qlen : tlen;
    ncol = ((ncol < w + 1 ? ncol : w + 1) + 15) / 16 + 1;
    const size_t rowb = (size_t)ncol * 8;
    if (ncol > R - 1) {
        if (wave == 0) { out->status = kStTooWide; out->n_cigar = 0; }   // same value from every lane
        return;
    }

    uint32_t U[S][8], V[S][8], X[S][8], Y[S][8], Z[S][8], T[S][8], Q[S][8];
    int32_t HC[STATS ? S : 1][16];
#pragma unroll
    for (int k = 0; k < S; ++k) {
#pragma unroll
        for (int j = 0; j < 8; ++j) { U[k][j] = V[k][j] = X[k][j] = Y[k][j] = 0; Z[k][j] = 0x10001000u; T[k][j] = Q[k][j] = 0; }
        if constexpr (STATS) {
#pragma unroll
            for (int c = 0; c < 16; ++c) HC[k][c] = kNegInf;
        }
    }
    int32_t bH = 0, bR = -1, bK = 0;   // STATS: best (H, r, key) of this lane, valid iff bH > 0

    // uniform state
    int last_st = -1, last_en = -1;
    int32_t He = kNegInf, Hl = kNegInf;
    int last_en0 = -1;
    int32_t score = kNegInf, mte = kNegInf, mte_q = -1, mqe = kNegInf, mqe_t = -1;
    int64_t cells = 0;
    bool zd = false;
    int rlast = 0;                        // last computed anti-diagonal
    int p_st0 = 0, p_en0 = 0, p_en = 0;   // its bounds (its edge log is pending)

    // consume the edge log of anti-diagonal rr (src/align/ksw2_extz2_sse.c:227-269)
    auto consume = [&](int rr, int st0, int en0, int en) {
        const int ue = sm.logu[rr & 63], vl = sm.logv[rr & 63];
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
        if constexpr (STATS) {
            if (rr - st0 == qlen - 1 && sm.logH[rr & 63][1] > mqe) { mqe = sm.logH[rr & 63][1]; mqe_t = st0; }
        }
        if (rr == ndiag - 1 && en0 == tlen - 1) score = He;
    };

    // element pointers (as dwords) for 16-cell windows starting at element e
    auto tgt_ptr = [&](int e) -> const uint32_t * {
        return (const uint32_t *)((e & 1) ? (J.tgt_odd + (e - 1)) : (J.tgt_even + e));
    };
    auto qry_ptr = [&](int e) -> const uint32_t * {
        return (const uint32_t *)((e & 1) ? (J.qr_odd + (e - 1)) : (J.qr_even + e));
    };

    // load query windows (and set up new blocks) for anti-diagonal r
    auto prefetch = [&](int r, const Bounds &bd, int prev_be) {
        const int bs = bd.st >> 4, be = bd.en >> 4, bsm = bs % R;
#pragma unroll
        for (int k = 0; k < S; ++k) {
            int d = NL * k + L - bsm;
            if (d < 0) d += R;
            const int b = bs + d;
            if (b <= be) {
                if (b > prev_be) {   // first activation: calloc state, load target block
                    const uint32_t *tp = tgt_ptr(16 * b);
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        U[k][j] = V[k][j] = X[k][j] = Y[k][j] = 0;
                        Z[k][j] = 0x10001000u;
                        T[k][j] = tp[j];
                    }
                }
                const uint32_t *qp = qry_ptr(qlen - 1 - r + 16 * b);
#pragma unroll
                for (int j = 0; j < 8; ++j) Q[k][j] = qp[j];
            }
        }
    };

    Bounds bd;
    diag_bounds(0, qlen, tlen, w, bd);
    prefetch(0, bd, -1);

    for (int r = 0; r < ndiag; ++r) {
        const int st0 = bd.st0, en0 = bd.en0, st = bd.st, en = bd.en;
        const int bs = st >> 4, be = en >> 4, bsm = bs % R;
        cells += en0 - st0 + 1;
        int blk[S];
        bool act[S];
#pragma unroll
        for (int k = 0; k < S; ++k) {
            int d = NL * k + L - bsm;
            if (d < 0) d += R;
            blk[k] = bs + d;
            act[k] = blk[k] <= be;
        }

        // top boundary y[r]=0, u[r]=q (:123)
        if (en >= r) {
            const int rb = r >> 4, rc = r & 15;
#pragma unroll
            for (int k = 0; k < S; ++k) {
                if (act[k] && blk[k] == rb) {
                    set_half(U[k], rc, r ? (uint32_t)kQ << 8 : 0u);
                    set_half(Y[k], rc, 0u);
                }
            }
        }

        // neighbour exchange: x[t-1], v[t-1] (and H) of the previous diagonal
        uint32_t nx[S], nv[S];
        int32_t nh[S];
        {
            uint32_t rx[S], rv[S];
            int32_t rh[S];
            // lane i receives lane i-1, lane 0 lane 63 (DPP rotate, no LDS round trip)
#pragma unroll
            for (int k = 0; k < S; ++k) {
                rx[k] = wave_ror1(X[k][7]);
                rv[k] = wave_ror1(V[k][7]);
                if constexpr (STATS) rh[k] = (int32_t)wave_ror1((uint32_t)HC[k][15]); else rh[k] = 0;
            }
            if constexpr (W > 1) {
                if (lane == 63) {
#pragma unroll
                    for (int k = 0; k < S; ++k) {
                        sm.xch[r & 1][wave][k][0] = X[k][7];
                        sm.xch[r & 1][wave][k][1] = V[k][7];
                        if constexpr (STATS) sm.xch[r & 1][wave][k][2] = (uint32_t)HC[k][15];
                    }
                }
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < S; ++k) { nx[k] = rx[k]; nv[k] = rv[k]; nh[k] = rh[k]; }
            if (lane == 0) {
                if constexpr (W == 1) {
#pragma unroll
                    for (int k = 0; k < S; ++k) {
                        const int kp = (k + S - 1) % S;
                        nx[k] = rx[kp]; nv[k] = rv[kp]; nh[k] = rh[kp];
                    }
                } else {
#pragma unroll
                    for (int k = 0; k < S; ++k) {
                        const int ws = wave ? wave - 1 : W - 1;
                        const int ks = wave ? k : (k + S - 1) % S;
                        nx[k] = sm.xch[r & 1][ws][ks][0];
                        nv[k] = sm.xch[r & 1][ws][ks][1];
                        nh[k] = (int32_t)sm.xch[r & 1][ws][ks][2];
                    }
                }
            }
        }
        // the log of diagonal r-1 is complete once every wave passed the barrier
        if (r > 0) consume(r - 1, p_st0, p_en0, p_en);

        // left boundary of the rounded region (:118-122)
        if (L == bsm % NL) {
            const int ks = bsm / NL;
            const bool keep = st > 0 && st - 1 >= last_st && st - 1 <= last_en;
            const uint32_t v1 = (st == 0 && r) ? (uint32_t)kQ << 24 : 0u;
#pragma unroll
            for (int k = 0; k < S; ++k) {
                const uint32_t m = keep ? 0u : sel_mask(k, ks);
                nx[k] = nx[k] & ~m;
                nv[k] = (nv[k] & ~m) | (v1 & m);
            }
        }

        // cell recurrence (:26-47, :197-222) over all active blocks
#pragma unroll
        for (int k = 0; k < S; ++k) {
            if (!act[k]) continue;
            const int b = blk[k];
            const int lo = st0 - 16 * b, hi = en0 - 16 * b;
            // score fill only inside [st0, en0] (:142-143): elsewhere the old value stays
            {
                uint32_t zf[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const uint32_t sel = T[k][j] ^ Q[k][j];
                    uint32_t z = __builtin_amdgcn_perm(0x10101010u, 0x0C0C0C15u, sel);
                    if (hasn) z -= ((T[k][j] & Q[k][j] & 0x04000400u) >> 10) * 0x500u;
                    zf[j] = z;
                }
                if (lo <= 0 && hi >= 15) {
#pragma unroll
                    for (int j = 0; j < 8; ++j) Z[k][j] = zf[j];
                } else {
                    const int l0 = lo < 0 ? 0 : lo, h0 = hi > 15 ? 15 : hi;
                    const uint32_t m16 = (h0 >= l0) ? (((2u << h0) - 1u) & ~((1u << l0) - 1u)) : 0u;
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        const uint32_t m = ((m16 >> (2 * j)) & 1u ? 0x0000ffffu : 0u) |
                                           ((m16 >> (2 * j + 1)) & 1u ? 0xffff0000u : 0u);
                        Z[k][j] = bfi(m, zf[j], Z[k][j]);
                    }
                }
            }
            uint32_t C[4];
            auto cells_rl = [&](auto right_tag) {
                constexpr bool RT = decltype(right_tag)::value;
#pragma unroll
                for (int m = 3; m >= 0; --m) {
                    uint32_t fa[2], fb[2], fx[2], fy[2];
#pragma unroll
                    for (int h = 1; h >= 0; --h) {
                        const int j = 2 * m + h;
                        const uint32_t xlo = j ? X[k][j - 1] : nx[k];
                        const uint32_t vlo = j ? V[k][j - 1] : nv[k];
                        const uint32_t xs = cat16(X[k][j], xlo);
                        const uint32_t vs = cat16(V[k][j], vlo);
                        const uint32_t a = pk_add(xs, vs);
                        const uint32_t ut = U[k][j];
                        const uint32_t bb = pk_add(Y[k][j], ut);
                        const uint32_t z0 = Z[k][j];
                        fa[h] = RT ? pk_subsat(a, z0) : pk_subsat(z0, a);
                        uint32_t z = pk_maxu(z0, a);
                        fb[h] = RT ? pk_subsat(bb, z) : pk_subsat(z, bb);
                        z = pk_maxu(z, bb);
                        z = pk_minu(z, 0x15001500u);
                        U[k][j] = pk_sub(z, vs);
                        V[k][j] = pk_sub(z, ut);
                        z = pk_sub(z, 0x06000600u);
                        const uint32_t a2 = pk_sub(a, z), b2 = pk_sub(bb, z);
                        X[k][j] = pk_maxi(a2, 0u);
                        Y[k][j] = pk_maxi(b2, 0u);
                        fx[h] = RT ? a2 : pk_subsat(0u, a2);
                        fy[h] = RT ? b2 : pk_subsat(0u, b2);
                    }
                    // 4 flags of cells (2m, 2m+8, 2m+1, 2m+9): low nibbles for even m, high for odd
                    C[m] = (m & 1) ? tb_nibbles<true>(fa[0], fa[1], fb[0], fb[1], fx[0], fx[1], fy[0], fy[1])
                                   : tb_nibbles<false>(fa[0], fa[1], fb[0], fb[1], fx[0], fx[1], fy[0], fy[1]);
                }
            };
            if (right) cells_rl(std::integral_constant<bool, true>{});
            else cells_rl(std::integral_constant<bool, false>{});
            uint2 o;
            o.x = bfi(0xF0F0F0F0u, C[1], C[0]);
            o.y = bfi(0xF0F0F0F0u, C[3], C[2]);
            *(uint2 *)(tb + (size_t)r * rowb + (size_t)(b - bs) * 8) = o;

            // H bookkeeping
            if (b == (en0 >> 4)) sm.logu[r & 63] = (uint8_t)byte_of(U[k], en0 & 15);
            {
                const int tv = en0 > 0 ? en0 - 1 : 0;
                if (b == (tv >> 4)) sm.logv[r & 63] = (uint8_t)byte_of(V[k], tv & 15);
            }
            if constexpr (STATS) {
                // exact H row (:227-261): fix H[en0] from the pre-update H[en0-1], then add
                // v-qe over [st0, en0)
                // in the lane that owns en0, ce = en0 & 15 (wave-uniform): the edge cell's
                // neighbours are read and written through uniform switches
                const int ce = en0 - 16 * b, cu = en0 & 15;
                int32_t hfix = 0;
                const bool own_e = (ce >= 0 && ce < 16);
                if (r == 0) hfix = byte_of(V[k], 0) - 2 * kQE;
                else if (en0 == 0) hfix = HC[k][0] + byte_of(V[k], 0) - kQE;
                else hfix = (cu > 0 ? pick16u(HC[k], cu - 1) : nh[k]) + byte_of_u(U[k], cu) - kQE;
                // branch-free: cell c is row column t = 16b + c, in range iff lo <= c < ce
                const int lo = st0 - 16 * b;
#pragma unroll
                for (int c = 0; c < 16; ++c) {
                    const int32_t dv = (int32_t)((V[k][c >> 1] >> ((c & 1) * 16 + 8)) & 0xffu) - kQE;
                    HC[k][c] += (c >= lo && c < ce) ? dv : 0;
                }
                put16u(HC[k], cu, hfix, own_e);
                // per-diagonal max with the reference's tie order (:228-260), folded into a
                // running lexicographic (H desc, r asc, key asc) best: the block's largest H
                // over [st0, en0], then the smallest key among the cells that hold it (the
                // key is computed for those cells only, usually one)
                int32_t hb = INT_MIN;
#pragma unroll
                for (int c = 0; c < 16; ++c) hb = max(hb, (c >= lo && c <= ce) ? HC[k][c] : INT_MIN);
                uint32_t eq = 0;
#pragma unroll
                for (int c = 0; c < 16; ++c) eq |= (uint32_t)(c >= lo && c <= ce && HC[k][c] == hb) << c;
                if (eq && (hb > bH || (hb == bH && bR == r))) {
                    const int en1 = st0 + ((en0 - st0) >> 2 << 2);
                    int kb = INT_MAX;
                    for (uint32_t m = eq; m; m &= m - 1) {
                        const int t = 16 * b + __builtin_ctz(m);
                        const int key = (t == en0) ? 0
                                        : (t < en1) ? 1 + (((t - st0) & 3) << 20) + ((t - st0) >> 2)
                                                    : 1 + (4 << 20) + (t - en1);
                        kb = key < kb ? key : kb;
                    }
                    if (hb > bH || kb < bK) { bH = hb; bR = r; bK = kb; }
                }
                const int cs = st0 - 16 * b;
                if (r - st0 == qlen - 1 && cs >= 0 && cs < 16) sm.logH[r & 63][1] = pick16u(HC[k], st0 & 15);
            }
        }

        p_st0 = st0; p_en0 = en0; p_en = en; rlast = r;
        last_st = st; last_en = en;
        if (r + 1 < ndiag) {
            if (!diag_bounds(r + 1, qlen, tlen, w, bd)) { zd = true; break; }
            prefetch(r + 1, bd, be);
        }
    }
    __threadfence();
    __syncthreads();
    consume(rlast, p_st0, p_en0, p_en);

    // STATS: reduce the per-lane best (H desc, r asc, key asc) over the workgroup
    int32_t emax = 0, emax_t = -1, emax_q = -1;
    if constexpr (STATS) {
        sm.red[L][0] = bH; sm.red[L][1] = bR; sm.red[L][2] = bK;
        __syncthreads();
        int32_t gH = 0, gR = -1, gK = 0;
        for (int i = 0; i < NL; ++i) {
            const int32_t h = sm.red[i][0], rr = sm.red[i][1], kk = sm.red[i][2];
            if (rr < 0 || h <= 0) continue;
            if (h > gH || (h == gH && (rr < gR || (rr == gR && kk < gK)))) { gH = h; gR = rr; gK = kk; }
        }
        if (gR >= 0 && gH > 0) {
            Bounds b2;
            diag_bounds(gR, qlen, tlen, w, b2);
            const int en1 = b2.st0 + ((b2.en0 - b2.st0) >> 2 << 2);
            int t;
            if (gK == 0) t = b2.en0;
            else if (gK < 1 + (4 << 20)) t = b2.st0 + 4 * ((gK - 1) & 0xfffff) + ((gK - 1) >> 20);
            else t = en1 + (gK - 1 - (4 << 20));
            emax = gH; emax_t = t; emax_q = gR - t;
        }
        __syncthreads();
    }

    // backtrack (include/ksw2.h:129-161) on wave 0, traceback rows staged in LDS
    int i0 = -1, j0 = -1;
    if (!zd) { i0 = tlen - 1; j0 = qlen - 1; }
    else if (STATS && emax_t >= 0 && emax_q >= 0) { i0 = emax_t; j0 = emax_q; }
    int nout = 0;
    int status = kStOk;
    if (zd && !STATS) status = kStNeedStats;
    if (wave == 0 && i0 >= 0 && status == kStOk) {
        nout = backtrack_to_cigar(J, tb, rowb, w, right, i0, j0, sm.chunk, 0);
        if (nout > J.cig_cap) status = kStCigarOverflow;
    }
    // every lane of wave 0 stores the same (uniform) values: no lane-0 branch
    // next to the job loop's back edge
    if (wave == 0) {
        out->score = score; out->n_cigar = nout; out->zdropped = zd ? 1 : 0; out->status = status;
        out->max = (uint32_t)emax; out->max_t = emax_t; out->max_q = emax_q;
        out->mqe = STATS ? mqe : kNegInf; out->mqe_t = STATS ? mqe_t : -1;
        out->mte = mte; out->mte_q = mte_q; out->pad_ = 0; out->cells = cells;
    }
    __syncthreads();
}

// Wide bands (any n_col_): the same recurrence and statistics as dp_job, with
// the band's blocks of an anti-diagonal processed in passes of one block per
// lane (pass p holds blocks bs + 64W*p + L). Block state (u, v, x, y, score,
// 16 H values) lives between diagonals in a ring of n_col_ + 2 records of 56
// words in global memory, after the job's traceback rows: block b in record
// b mod (n_col_ + 2), so a block entering the band never overwrites one still
// in it. Passes run highest first, so the first lane of pass p reads its left
// neighbour (the last block of pass p - 1) as the previous diagonal left it;
// within a pass neighbours come by DPP and LDS as in dp_job.
constexpr int kWideWords = 56;   // U V X Y Z (8 words each) + 16 H

template <int W>
__device__ void dp_job_wide(const DpJob &J, uint8_t *__restrict__ tb, DpOut *__restrict__ out, Smem<W, 1> &sm)
{
    constexpr int NL = 64 * W;
    const int L = threadIdx.x, lane = L & 63;
    const int wave = __builtin_amdgcn_readfirstlane(L >> 6);
    const int qlen = J.qlen, tlen = J.tlen;
    int w = J.w;
    if (w < 0) w = tlen > qlen ? tlen : qlen;
    const bool right = (J.flags & kFlagRight) != 0;
    const bool hasn = (J.flags & kFlagHasN) != 0;
    const int ndiag = qlen + tlen - 1;
    int ncol = qlen < tlen ? qlen : tlen;
    ncol = ((ncol < w + 1 ? ncol : w + 1) + 15) / 16 + 1;
    const size_t rowb = (size_t)ncol * 8;
    const int RG = ncol + 2;
    uint32_t *__restrict__ ring =
        (uint32_t *)(((uintptr_t)(tb + (size_t)ndiag * rowb) + 15) & ~(uintptr_t)15);

    int32_t bH = 0, bR = -1, bK = 0;   // best (H, r, key) of this lane, valid iff bH > 0
    int last_st = -1, last_en = -1, prev_be = -1;
    int32_t He = kNegInf, Hl = kNegInf;
    int last_en0 = -1;
    int32_t score = kNegInf, mte = kNegInf, mte_q = -1, mqe = kNegInf, mqe_t = -1;
    int64_t cells = 0;
    bool zd = false;
    int rlast = 0, p_st0 = 0, p_en0 = 0, p_en = 0, pc = 0;

    auto consume = [&](int rr, int st0, int en0, int en) {
        const int ue = sm.logu[rr & 63], vl = sm.logv[rr & 63];
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
        if (rr - st0 == qlen - 1 && sm.logH[rr & 63][1] > mqe) { mqe = sm.logH[rr & 63][1]; mqe_t = st0; }
        if (rr == ndiag - 1 && en0 == tlen - 1) score = He;
    };
    auto tgt_ptr = [&](int e) -> const uint32_t * {
        return (const uint32_t *)((e & 1) ? (J.tgt_odd + (e - 1)) : (J.tgt_even + e));
    };
    auto qry_ptr = [&](int e) -> const uint32_t * {
        return (const uint32_t *)((e & 1) ? (J.qr_odd + (e - 1)) : (J.qr_even + e));
    };

    Bounds bd;
    diag_bounds(0, qlen, tlen, w, bd);
    for (int r = 0; r < ndiag; ++r) {
        const int st0 = bd.st0, en0 = bd.en0, st = bd.st, en = bd.en;
        const int bs = st >> 4, be = en >> 4;
        const int npass = (be - bs + NL) / NL;
        cells += en0 - st0 + 1;
        for (int p = npass - 1; p >= 0; --p, ++pc) {
            const int b = bs + NL * p + L;
            const bool act = b <= be;
            const bool fresh = b > prev_be;   // first activation: the reference's calloc'd zeros
            uint32_t U[8], V[8], X[8], Y[8], Z[8], T[8], Q[8];
            int32_t HC[16];
            uint32_t *rec = ring + (size_t)(b % RG) * kWideWords;
            if (act && !fresh) {
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    U[j] = rec[j]; V[j] = rec[8 + j]; X[j] = rec[16 + j]; Y[j] = rec[24 + j]; Z[j] = rec[32 + j];
                }
#pragma unroll
                for (int c = 0; c < 16; ++c) HC[c] = (int32_t)rec[40 + c];
            } else {
#pragma unroll
                for (int j = 0; j < 8; ++j) { U[j] = V[j] = X[j] = Y[j] = 0; Z[j] = 0x10001000u; }
#pragma unroll
                for (int c = 0; c < 16; ++c) HC[c] = kNegInf;
            }
            if (act) {
                const uint32_t *tp = tgt_ptr(16 * b);
                const uint32_t *qp = qry_ptr(qlen - 1 - r + 16 * b);
#pragma unroll
                for (int j = 0; j < 8; ++j) { T[j] = tp[j]; Q[j] = qp[j]; }
            } else {
#pragma unroll
                for (int j = 0; j < 8; ++j) T[j] = Q[j] = 0;
            }
            // top boundary y[r]=0, u[r]=q (:123)
            if (act && en >= r && b == (r >> 4)) {
                set_half(U, r & 15, r ? (uint32_t)kQ << 8 : 0u);
                set_half(Y, r & 15, 0u);
            }
            // neighbour exchange: x[t-1], v[t-1], H of the previous diagonal
            uint32_t nx = wave_ror1(X[7]), nv = wave_ror1(V[7]);
            int32_t nh = (int32_t)wave_ror1((uint32_t)HC[15]);
            if constexpr (W > 1) {
                if (lane == 63) {
                    sm.xch[pc & 1][wave][0][0] = X[7];
                    sm.xch[pc & 1][wave][0][1] = V[7];
                    sm.xch[pc & 1][wave][0][2] = (uint32_t)HC[15];
                }
            }
            __syncthreads();
            if constexpr (W > 1) {
                if (lane == 0 && wave > 0) {
                    nx = sm.xch[pc & 1][wave - 1][0][0];
                    nv = sm.xch[pc & 1][wave - 1][0][1];
                    nh = (int32_t)sm.xch[pc & 1][wave - 1][0][2];
                }
            }
            if (L == 0) {   // left neighbour: last block of the next (lower) pass, still as diagonal r-1 left it
                const int bl = b - 1;
                if (bl >= 0 && bl <= prev_be && bl >= bs - 1) {
                    const uint32_t *lr = ring + (size_t)(bl % RG) * kWideWords;
                    nx = lr[16 + 7]; nv = lr[8 + 7]; nh = (int32_t)lr[40 + 15];
                } else {
                    nx = 0u; nv = 0u; nh = kNegInf;
                }
            }
            // the log of diagonal r-1 is complete once every wave passed the first barrier of r
            if (r > 0 && p == npass - 1) consume(r - 1, p_st0, p_en0, p_en);
            // left boundary of the rounded region (:118-122)
            if (L == 0 && p == 0) {
                const bool keep = st > 0 && st - 1 >= last_st && st - 1 <= last_en;
                const uint32_t v1 = (st == 0 && r) ? (uint32_t)kQ << 24 : 0u;
                if (!keep) { nx = 0u; nv = v1; }
            }
            if (act) {
                const int lo = st0 - 16 * b, hi = en0 - 16 * b;
                {
                    uint32_t zf[8];
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        const uint32_t sel = T[j] ^ Q[j];
                        uint32_t z = __builtin_amdgcn_perm(0x10101010u, 0x0C0C0C15u, sel);
                        if (hasn) z -= ((T[j] & Q[j] & 0x04000400u) >> 10) * 0x500u;
                        zf[j] = z;
                    }
                    if (lo <= 0 && hi >= 15) {
#pragma unroll
                        for (int j = 0; j < 8; ++j) Z[j] = zf[j];
                    } else {
                        const int l0 = lo < 0 ? 0 : lo, h0 = hi > 15 ? 15 : hi;
                        const uint32_t m16 = (h0 >= l0) ? (((2u << h0) - 1u) & ~((1u << l0) - 1u)) : 0u;
#pragma unroll
                        for (int j = 0; j < 8; ++j) {
                            const uint32_t m = ((m16 >> (2 * j)) & 1u ? 0x0000ffffu : 0u) |
                                               ((m16 >> (2 * j + 1)) & 1u ? 0xffff0000u : 0u);
                            Z[j] = bfi(m, zf[j], Z[j]);
                        }
                    }
                }
                uint32_t C[4];
                auto cells_rl = [&](auto right_tag) {
                    constexpr bool RT = decltype(right_tag)::value;
#pragma unroll
                    for (int m = 3; m >= 0; --m) {
                        uint32_t fa[2], fb[2], fx[2], fy[2];
#pragma unroll
                        for (int h = 1; h >= 0; --h) {
                            const int j = 2 * m + h;
                            const uint32_t xlo = j ? X[j - 1] : nx;
                            const uint32_t vlo = j ? V[j - 1] : nv;
                            const uint32_t xs = cat16(X[j], xlo);
                            const uint32_t vs = cat16(V[j], vlo);
                            const uint32_t a = pk_add(xs, vs);
                            const uint32_t ut = U[j];
                            const uint32_t bb = pk_add(Y[j], ut);
                            const uint32_t z0 = Z[j];
                            fa[h] = RT ? pk_subsat(a, z0) : pk_subsat(z0, a);
                            uint32_t z = pk_maxu(z0, a);
                            fb[h] = RT ? pk_subsat(bb, z) : pk_subsat(z, bb);
                            z = pk_maxu(z, bb);
                            z = pk_minu(z, 0x15001500u);
                            U[j] = pk_sub(z, vs);
                            V[j] = pk_sub(z, ut);
                            z = pk_sub(z, 0x06000600u);
                            const uint32_t a2 = pk_sub(a, z), b2 = pk_sub(bb, z);
                            X[j] = pk_maxi(a2, 0u);
                            Y[j] = pk_maxi(b2, 0u);
                            fx[h] = RT ? a2 : pk_subsat(0u, a2);
                            fy[h] = RT ? b2 : pk_subsat(0u, b2);
                        }
                        C[m] = (m & 1) ? tb_nibbles<true>(fa[0], fa[1], fb[0], fb[1], fx[0], fx[1], fy[0], fy[1])
                                       : tb_nibbles<false>(fa[0], fa[1], fb[0], fb[1], fx[0], fx[1], fy[0], fy[1]);
                    }
                };
                if (right) cells_rl(std::integral_constant<bool, true>{});
                else cells_rl(std::integral_constant<bool, false>{});
                uint2 o;
                o.x = bfi(0xF0F0F0F0u, C[1], C[0]);
                o.y = bfi(0xF0F0F0F0u, C[3], C[2]);
                *(uint2 *)(tb + (size_t)r * rowb + (size_t)(b - bs) * 8) = o;

                if (b == (en0 >> 4)) sm.logu[r & 63] = (uint8_t)byte_of(U, en0 & 15);
                {
                    const int tv = en0 > 0 ? en0 - 1 : 0;
                    if (b == (tv >> 4)) sm.logv[r & 63] = (uint8_t)byte_of(V, tv & 15);
                }
                // exact H row and the per-diagonal maximum, as dp_job with statistics
                {
                    const int ce = en0 - 16 * b, cu = en0 & 15;
                    int32_t hfix = 0;
                    const bool own_e = (ce >= 0 && ce < 16);
                    if (r == 0) hfix = byte_of(V, 0) - 2 * kQE;
                    else if (en0 == 0) hfix = HC[0] + byte_of(V, 0) - kQE;
                    else hfix = (cu > 0 ? pick16u(HC, cu - 1) : nh) + byte_of_u(U, cu) - kQE;
#pragma unroll
                    for (int c = 0; c < 16; ++c) {
                        const int32_t dv = (int32_t)((V[c >> 1] >> ((c & 1) * 16 + 8)) & 0xffu) - kQE;
                        HC[c] += (c >= lo && c < ce) ? dv : 0;
                    }
                    put16u(HC, cu, hfix, own_e);
                    int32_t hb = INT_MIN;
#pragma unroll
                    for (int c = 0; c < 16; ++c) hb = max(hb, (c >= lo && c <= ce) ? HC[c] : INT_MIN);
                    uint32_t eq = 0;
#pragma unroll
                    for (int c = 0; c < 16; ++c) eq |= (uint32_t)(c >= lo && c <= ce && HC[c] == hb) << c;
                    if (eq && (hb > bH || (hb == bH && bR == r))) {
                        const int en1 = st0 + ((en0 - st0) >> 2 << 2);
                        int kb = INT_MAX;
                        for (uint32_t m = eq; m; m &= m - 1) {
                            const int t = 16 * b + __builtin_ctz(m);
                            const int key = (t == en0) ? 0
                                            : (t < en1) ? 1 + (((t - st0) & 3) << 20) + ((t - st0) >> 2)
                                                        : 1 + (4 << 20) + (t - en1);
                            kb = key < kb ? key : kb;
                        }
                        if (hb > bH || kb < bK) { bH = hb; bR = r; bK = kb; }
                    }
                    const int cs = st0 - 16 * b;
                    if (r - st0 == qlen - 1 && cs >= 0 && cs < 16) sm.logH[r & 63][1] = pick16u(HC, st0 & 15);
                }
                // state back to the ring (the lower pass reads only its own records
                // and this pass's first block's left neighbour, which is not written here)
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    rec[j] = U[j]; rec[8 + j] = V[j]; rec[16 + j] = X[j]; rec[24 + j] = Y[j]; rec[32 + j] = Z[j];
                }
#pragma unroll
                for (int c = 0; c < 16; ++c) rec[40 + c] = (uint32_t)HC[c];
            }
        }
        p_st0 = st0; p_en0 = en0; p_en = en; rlast = r;
        last_st = st; last_en = en; prev_be = be;
        // the ring records of this diagonal are read by other lanes in the next one
        __threadfence_block();
        __syncthreads();
        if (r + 1 < ndiag && !diag_bounds(r + 1, qlen, tlen, w, bd)) { zd = true; break; }
    }
    __threadfence();
    __syncthreads();
    consume(rlast, p_st0, p_en0, p_en);

    int32_t emax = 0, emax_t = -1, emax_q = -1;
    {
        sm.red[L][0] = bH; sm.red[L][1] = bR; sm.red[L][2] = bK;
        __syncthreads();
        int32_t gH = 0, gR = -1, gK = 0;
        for (int i = 0; i < NL; ++i) {
            const int32_t h = sm.red[i][0], rr = sm.red[i][1], kk = sm.red[i][2];
            if (rr < 0 || h <= 0) continue;
            if (h > gH || (h == gH && (rr < gR || (rr == gR && kk < gK)))) { gH = h; gR = rr; gK = kk; }
        }
        if (gR >= 0 && gH > 0) {
            Bounds b2;
            diag_bounds(gR, qlen, tlen, w, b2);
            const int en1 = b2.st0 + ((b2.en0 - b2.st0) >> 2 << 2);
            int t;
            if (gK == 0) t = b2.en0;
            else if (gK < 1 + (4 << 20)) t = b2.st0 + 4 * ((gK - 1) & 0xfffff) + ((gK - 1) >> 20);
            else t = en1 + (gK - 1 - (4 << 20));
            emax = gH; emax_t = t; emax_q = gR - t;
        }
        __syncthreads();
    }
    int i0 = -1, j0 = -1;
    if (!zd) { i0 = tlen - 1; j0 = qlen - 1; }
    else if (emax_t >= 0 && emax_q >= 0) { i0 = emax_t; j0 = emax_q; }
    int nout = 0;
    int status = kStOk;
    if (wave == 0 && i0 >= 0) {
        nout = backtrack_to_cigar(J, tb, rowb, w, right, i0, j0, sm.chunk, 0);
        if (nout > J.cig_cap) status = kStCigarOverflow;
    }
    if (wave == 0) {
        out->score = score; out->n_cigar = nout; out->zdropped = zd ? 1 : 0; out->status = status;
        out->max = (uint32_t)emax; out->max_t = emax_t; out->max_q = emax_q;
        out->mqe = mqe; out->mqe_t = mqe_t;
        out->mte = mte; out->mte_q = mte_q; out->pad_ = 0; out->cells = cells;
    }
    __syncthreads();
}

template <int W>
__global__ __launch_bounds__(64 * W) void dp_wide_persistent(const DpJob *__restrict__ jobs, int njobs,
                                                             int *__restrict__ counter, uint8_t *__restrict__ tb_pool,
                                                             size_t tb_slot_bytes, DpOut *__restrict__ outs)
{
    __shared__ Smem<W, 1> sm;
    uint8_t *tb = tb_pool + (size_t)blockIdx.x * tb_slot_bytes;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (;;) {
        if (wave == 0) sm.job = __builtin_amdgcn_readfirstlane(atomicAdd(counter, threadIdx.x == 0 ? 1 : 0));
        __syncthreads();
        const int jid = __builtin_amdgcn_readfirstlane(sm.job);
        __syncthreads();
        if (jid >= njobs) break;
        const DpJob J = jobs[jid];
        dp_job_wide<W>(J, tb, &outs[jid], sm);
    }
}

// Persistent kernel: each workgroup pulls jobs from a queue and reuses its own
// traceback slot. With njobs == 1 and one workgroup it is the plain
// one-pair-per-launch form.
template <int W, int S, bool STATS>
__global__ __launch_bounds__(64 * W) void dp_persistent(const DpJob *__restrict__ jobs, int njobs,
                                                        int *__restrict__ counter, uint8_t *__restrict__ tb_pool,
                                                        size_t tb_slot_bytes, DpOut *__restrict__ outs)
{
    __shared__ Smem<W, S> sm;
    uint8_t *tb = tb_pool + (size_t)blockIdx.x * tb_slot_bytes;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (;;) {
        // Job ticket from an atomic that every lane of wave 0 executes (lane 0
        // adds 1), shared through LDS and read back as a scalar: no lane-0
        // branch at the loop's back edge, and the loop exit is uniform.
        if (wave == 0) sm.job = __builtin_amdgcn_readfirstlane(atomicAdd(counter, threadIdx.x == 0 ? 1 : 0));
        __syncthreads();
        const int jid = __builtin_amdgcn_readfirstlane(sm.job);
        __syncthreads();
        if (jid >= njobs) break;
        const DpJob J = jobs[jid];
        dp_job<W, S, STATS>(J, tb, &outs[jid], sm);
    }
}

// (W, S) shapes: ring of R = 64*W*S blocks serves bands with n_col_ <= R-1.
struct Shape { int W, S; };
static const Shape kShapes[] = {{1, 1}, {1, 2}, {1, 3}, {2, 2}, {2, 3}, {4, 2}, {4, 3}, {8, 3}};
constexpr int kNumShapes = sizeof(kShapes) / sizeof(kShapes[0]);

template <int W, int S, bool STATS>
static const void *kernel_ptr() { return (const void *)&dp_persistent<W, S, STATS>; }

static const void *shape_kernel(int idx, bool stats)
{
#define HAL_SHAPE(i, w, s) case i: return stats ? kernel_ptr<w, s, true>() : kernel_ptr<w, s, false>();
    switch (idx) {
        HAL_SHAPE(0, 1, 1) HAL_SHAPE(1, 1, 2) HAL_SHAPE(2, 1, 3) HAL_SHAPE(3, 2, 2)
        HAL_SHAPE(4, 2, 3) HAL_SHAPE(5, 4, 2) HAL_SHAPE(6, 4, 3) HAL_SHAPE(7, 8, 3)
    }
#undef HAL_SHAPE
    return nullptr;
}

// Shape kNumShapes is the wide-band kernel (any n_col_, always with statistics).
constexpr int kWideWaves = 8;

int dp_num_shapes() { return kNumShapes + 1; }
int dp_shape_ring(int idx) { return idx == kNumShapes ? (1 << 30) : 64 * kShapes[idx].W * kShapes[idx].S; }
int dp_shape_threads(int idx) { return 64 * (idx == kNumShapes ? kWideWaves : kShapes[idx].W); }
// traceback slot bytes of a job on shape idx: its traceback rows, and for the
// wide-band kernel the ring of block states behind them
size_t dp_shape_slot_bytes(int idx, int qlen, int tlen, int ncol)
{
    const size_t tbb = (size_t)(qlen + tlen - 1) * (size_t)ncol * 8;
    if (idx != kNumShapes) return tbb;
    return tbb + 16 + (size_t)(ncol + 2) * kWideWords * 4;
}

static const void *any_kernel(int idx, bool stats)
{
    return idx == kNumShapes ? (const void *)&dp_wide_persistent<kWideWaves> : shape_kernel(idx, stats);
}

// resident workgroups per CU of each kernel, asked once (0 = not asked yet)
int dp_shape_max_groups(int idx, bool stats, int num_cu)
{
    static std::atomic<int> cache[kNumShapes + 1][2];
    int per_cu = cache[idx][stats ? 1 : 0].load(std::memory_order_relaxed);
    if (per_cu <= 0) {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, any_kernel(idx, stats), dp_shape_threads(idx), 0) !=
                hipSuccess ||
            per_cu <= 0)
            return 0;
        cache[idx][stats ? 1 : 0].store(per_cu, std::memory_order_relaxed);
    }
    return per_cu * num_cu;
}

hipError_t dp_launch(int idx, bool stats, int groups, const DpJob *jobs, int njobs, int *counter,
                     uint8_t *tb_pool, size_t tb_slot_bytes, DpOut *outs, hipStream_t stream)
{
    void *args[] = {(void *)&jobs, (void *)&njobs, (void *)&counter, (void *)&tb_pool,
                    (void *)&tb_slot_bytes, (void *)&outs};
    return hipLaunchKernel(any_kernel(idx, stats), dim3(groups), dim3(dp_shape_threads(idx)), args, 0, stream);
}

}  // namespace hal
