/*
 * halign_oracle.c -- CPU restatement of HAlign-4's center-star alignment path.
 *
 * TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg as the checker. The product (libhalign_hip.so)
 * never links or calls this file.
 *
 * Each function cites the reference file:line it restates. Written from the
 * reference's documented behaviour as plain scalar C: the reference's SSE2
 * vector code is restated cell by cell (8-bit wrapping arithmetic, signed
 * 8-bit compares, unsigned 8-bit max/min exactly where the SSE2 path uses
 * them), so the stale values the vector code reads at 16-cell block margins
 * are reproduced.
 */
#include "halign_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* scoring constants: dna5_simd_mat include/align.h:109-116,           */
/* KSW2AlignConfig include/align.h:126-135                              */
/* ------------------------------------------------------------------ */
static const int8_t k_mat[25] = {
     5, -4, -4, -4, 0,
    -4,  5, -4, -4, 0,
    -4, -4,  5, -4, 0,
    -4, -4, -4,  5, 0,
     0,  0,  0,  0, 0,
};
enum { K_Q = 6, K_E = 2, K_M = 5 };

/* ScoreChar2Idx include/align.h:72-89: A/C/G/T (either case) -> 0..3, else 4 */
static inline uint8_t score_code(unsigned char c)
{
    switch (c) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return 4;
    }
}

/* 8-bit helpers mirroring _mm_*_epi8 / _mm_*_epu8 lane semantics */
static inline uint8_t w8(int v) { return (uint8_t)(v & 0xff); }
static inline int s8(uint8_t v) { return (int)(int8_t)v; }
static inline uint8_t maxu8(uint8_t a, uint8_t b) { return a > b ? a : b; }
static inline uint8_t minu8(uint8_t a, uint8_t b) { return a < b ? a : b; }

/* ksw_push_cigar include/ksw2.h:113-123 (growable buffer) */
typedef struct { uint32_t *c; int n, m; } cigbuf_t;
static void push_cigar(cigbuf_t *cb, uint32_t op, int len)
{
    if (cb->n == 0 || op != (cb->c[cb->n - 1] & 0xf)) {
        if (cb->n == cb->m) {
            cb->m = cb->m ? cb->m << 1 : 16;
            cb->c = (uint32_t *)realloc(cb->c, (size_t)cb->m * 4);
        }
        cb->c[cb->n++] = (uint32_t)len << 4 | op;
    } else {
        cb->c[cb->n - 1] += (uint32_t)len << 4;
    }
}

/* ksw_backtrack, is_rot=1, is_rev=0, min_intron_len=0: include/ksw2.h:129-161 */
static void backtrack(const uint8_t *p, const int *off, const int *off_end,
                      size_t n_col, int i0, int j0, cigbuf_t *cb)
{
    int i = i0, j = j0, state = 0;
    while (i >= 0 && j >= 0) {
        int r = i + j, force_state = -1, tmp;
        if (i < off[r]) force_state = 2;
        if (i > off_end[r]) force_state = 1;
        tmp = force_state < 0 ? p[(size_t)r * n_col + (size_t)(i - off[r])] : 0;
        if (state == 0) state = tmp & 7;
        else if (!(tmp >> (state + 2) & 1)) state = 0;
        if (state == 0) state = tmp & 7;
        if (force_state >= 0) state = force_state;
        if (state == 0) { push_cigar(cb, 0, 1); --i; --j; }
        else if (state == 1) { push_cigar(cb, 2, 1); --i; }
        else { push_cigar(cb, 1, 1); --j; }
    }
    if (i >= 0) push_cigar(cb, 2, i + 1);
    if (j >= 0) push_cigar(cb, 1, j + 1);
    for (int k = 0; k < cb->n >> 1; ++k) {
        uint32_t t = cb->c[k];
        cb->c[k] = cb->c[cb->n - 1 - k];
        cb->c[cb->n - 1 - k] = t;
    }
}

/* ksw_extz2_sse src/align/ksw2_extz2_sse.c:23-304, SSE2 branch, exact max
 * (no KSW_EZ_APPROX_MAX), with CIGAR, zdrop=-1, end_bonus=0, no EXTZ_ONLY. */
int or_ksw_extz2(int qlen, const uint8_t *query, int tlen, const uint8_t *target,
                 int w, int flag, or_ez_t *ez, uint32_t *cigar, int cigar_cap,
                 int64_t *cells)
{
    /* ksw_reset_extz include/ksw2.h:184-189 */
    ez->max_q = ez->max_t = ez->mqe_t = ez->mte_q = -1;
    ez->max = 0; ez->score = ez->mqe = ez->mte = OR_KSW_NEG_INF;
    ez->n_cigar = 0; ez->zdropped = 0;
    if (cells) *cells = 0;
    if (qlen <= 0 || tlen <= 0) return 0;

    const int q = K_Q, e = K_E, qe = q + e;
    const uint8_t qe2 = (uint8_t)(2 * qe), max_sc = (uint8_t)(k_mat[0] + 2 * qe);
    const int right = (flag & OR_KSW_RIGHT) != 0;
    if (w < 0) w = tlen > qlen ? tlen : qlen;
    const int wl = w, wr = w;
    const int tlen_ = (tlen + 15) / 16;
    int n_col_ = qlen < tlen ? qlen : tlen;
    n_col_ = ((n_col_ < w + 1 ? n_col_ : w + 1) + 15) / 16 + 1;
    const size_t ncol = (size_t)n_col_ * 16;
    const int ndiag = qlen + tlen - 1;

    size_t asz = (size_t)tlen_ * 16;
    uint8_t *u = (uint8_t *)calloc(asz * 5, 1);
    uint8_t *v = u + asz, *x = v + asz, *y = x + asz, *s = y + asz;
    int32_t *H = (int32_t *)malloc(asz * 4);
    for (size_t t = 0; t < asz; ++t) H[t] = OR_KSW_NEG_INF;
    uint8_t *p = (uint8_t *)malloc((size_t)ndiag * ncol + 16);
    int *off = (int *)malloc((size_t)ndiag * 2 * sizeof(int));
    int *off_end = off + ndiag;
    int64_t ncells = 0;

    int last_st = -1, last_en = -1;
    for (int r = 0; r < ndiag; ++r) {
        int st = 0, en = tlen - 1, st0, en0;
        uint8_t x1, v1;
        if (st < r - qlen + 1) st = r - qlen + 1;
        if (en > r) en = r;
        if (st < (r - wr + 1) >> 1) st = (r - wr + 1) >> 1;
        if (en > (r + wl) >> 1) en = (r + wl) >> 1;
        if (st > en) { ez->zdropped = 1; break; }
        st0 = st; en0 = en;
        ncells += en0 - st0 + 1;
        st = st / 16 * 16; en = (en + 16) / 16 * 16 - 1;
        if (st > 0) {
            if (st - 1 >= last_st && st - 1 <= last_en) x1 = x[st - 1], v1 = v[st - 1];
            else x1 = v1 = 0;
        } else {
            x1 = 0; v1 = r ? (uint8_t)q : 0;
        }
        if (en >= r) { y[r] = 0; u[r] = r ? (uint8_t)q : 0; }
        /* GENERIC_SC score fill, only over [st0, en0] (:142-143) */
        for (int t = st0; t <= en0; ++t)
            s[t] = (uint8_t)k_mat[target[t] * K_M + query[r - t]];
        off[r] = st; off_end[r] = en;
        uint8_t *pr = p + (size_t)r * ncol;
        uint8_t xprev = x1, vprev = v1;  /* x[t-1], v[t-1] of diagonal r-1 */
        for (int t = st; t <= en; ++t) {
            uint8_t xt1 = xprev, vt1 = vprev;
            xprev = x[t]; vprev = v[t];
            uint8_t z = w8(s[t] + qe2);
            uint8_t a = w8(xt1 + vt1);
            uint8_t ut = u[t];
            uint8_t b = w8(y[t] + ut);
            uint8_t d;
            if (right) {
                d = s8(z) > s8(a) ? 0 : 1;
                if (!(s8(z) > 0)) z = 0;
                z = maxu8(z, a);
                d = s8(z) > s8(b) ? d : 2;
            } else {
                d = s8(a) > s8(z) ? 1 : 0;
                if (!(s8(z) > 0)) z = 0;
                z = maxu8(z, a);
                d = s8(b) > s8(z) ? 2 : d;
            }
            z = maxu8(z, b);
            z = minu8(z, max_sc);
            u[t] = w8(z - vt1);
            v[t] = w8(z - ut);
            z = w8(z - q);
            a = w8(a - z);
            b = w8(b - z);
            if (right) {
                x[t] = s8(a) < 0 ? 0 : a; d |= s8(a) < 0 ? 0 : 0x08;
                y[t] = s8(b) < 0 ? 0 : b; d |= s8(b) < 0 ? 0 : 0x10;
            } else {
                x[t] = s8(a) > 0 ? a : 0; d |= s8(a) > 0 ? 0x08 : 0;
                y[t] = s8(b) > 0 ? b : 0; d |= s8(b) > 0 ? 0x10 : 0;
            }
            pr[t - st] = d;
        }
        /* exact H row and per-diagonal max (:224-269) */
        int32_t max_H, max_t;
        if (r > 0) {
            int en1 = st0 + (en0 - st0) / 4 * 4, t;
            max_H = H[en0] = en0 > 0 ? H[en0 - 1] + u[en0] - qe : H[en0] + v[en0] - qe;
            max_t = en0;
            int32_t HH[4], tt[4];
            for (int k = 0; k < 4; ++k) HH[k] = max_H, tt[k] = max_t;
            for (t = st0; t < en1; t += 4) {
                for (int k = 0; k < 4; ++k) {
                    H[t + k] += (int32_t)v[t + k] - qe;
                    if (H[t + k] > HH[k]) HH[k] = H[t + k], tt[k] = t;
                }
            }
            for (int k = 0; k < 4; ++k)
                if (max_H < HH[k]) max_H = HH[k], max_t = tt[k] + k;
            for (; t < en0; ++t) {
                H[t] += (int32_t)v[t] - qe;
                if (H[t] > max_H) max_H = H[t], max_t = t;
            }
        } else {
            H[0] = v[0] - qe - qe; max_H = H[0]; max_t = 0;
        }
        if (en0 == tlen - 1 && H[en0] > ez->mte) ez->mte = H[en0], ez->mte_q = r - en;
        if (r - st0 == qlen - 1 && H[st0] > ez->mqe) ez->mqe = H[st0], ez->mqe_t = st0;
        /* ksw_apply_zdrop include/ksw2.h:191-207 with zdrop < 0 */
        if (max_H > (int32_t)ez->max) {
            ez->max = (uint32_t)max_H; ez->max_t = max_t; ez->max_q = r - max_t;
        }
        if (r == qlen + tlen - 2 && en0 == tlen - 1) ez->score = H[tlen - 1];
        last_st = st; last_en = en;
    }
    if (cells) *cells = ncells;

    cigbuf_t cb = {0, 0, 0};
    if (!ez->zdropped)
        backtrack(p, off, off_end, ncol, tlen - 1, qlen - 1, &cb);
    else if (ez->max_t >= 0 && ez->max_q >= 0)
        backtrack(p, off, off_end, ncol, ez->max_t, ez->max_q, &cb);
    free(u); free(H); free(p); free(off);
    ez->n_cigar = cb.n;
    int ret = cb.n;
    if (cigar) {
        if (cb.n > cigar_cap) ret = -1;
        else memcpy(cigar, cb.c, (size_t)cb.n * 4);
    } else if (cigar_cap >= 0 && cb.n > cigar_cap) {
        ret = -1;
    }
    free(cb.c);
    return ret;
}

int or_auto_band(int qlen, int tlen)
{
    /* include/align.h:159-172; called as auto_band(ref.size(), query.size()) */
    int d = qlen - tlen; if (d < 0) d = -d;
    int mx = qlen > tlen ? qlen : tlen;
    if ((double)d / (double)mx > 0.5) return -1;
    return 200 + (int)(0.1 * (double)(qlen + tlen / 2));
}

int or_global_align_ksw2(const char *ref, int rlen, const char *qry, int qlen,
                         int flag, uint32_t *cigar, int cigar_cap, int64_t *cells)
{
    if (cells) *cells = 0;
    if (rlen == 0 || qlen == 0) {   /* src/align/align.cpp:26-34 */
        if (rlen == 0 && qlen > 0) {
            if (cigar_cap < 1) return -1;
            cigar[0] = (uint32_t)qlen << 4 | 1;
            return 1;
        }
        if (qlen == 0 && rlen > 0) {
            if (cigar_cap < 1) return -1;
            cigar[0] = (uint32_t)rlen << 4 | 2;
            return 1;
        }
        return 0;
    }
    uint8_t *re = (uint8_t *)malloc((size_t)rlen + (size_t)qlen);
    uint8_t *qe = re + rlen;
    for (int i = 0; i < rlen; ++i) re[i] = score_code((unsigned char)ref[i]);
    for (int i = 0; i < qlen; ++i) qe[i] = score_code((unsigned char)qry[i]);
    int w = or_auto_band(rlen, qlen);
    or_ez_t ez;
    int n = or_ksw_extz2(qlen, qe, rlen, re, w, flag, &ez, cigar, cigar_cap, cells);
    free(re);
    return n;
}

/* ------------------------------------------------------------------ */
/* minimizers                                                           */
/* ------------------------------------------------------------------ */
/* nt4_table include/seed.h:157-190: ACGT/U either case -> 0..3, U->3 */
static inline uint8_t nt4(unsigned char c)
{
    switch (c) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': case 'U': case 'u': return 3;
    default: return 4;
    }
}

/* splitmix64 src/seed/minimizer.cpp:16-22 */
static inline uint64_t splitmix64(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

int or_extract_minimizer(const char *seq, int n_, int k_, int w_, int non_canonical,
                         uint64_t *x_out, uint64_t *y_out, int cap, int *ub)
{
    if (ub) *ub = 0;
    const uint32_t n = (uint32_t)n_, k = (uint32_t)k_, w = (uint32_t)w_;
    if (k == 0 || w == 0 || n < k) return 0;
    if (k > 31) return 0;
    if (w >= 256) return 0;
    const uint32_t total_kmer = n - k + 1;
    const uint32_t win = w < total_kmer ? w : total_kmer;
    if (win == 0) return 0;
    const uint64_t mask = (1ULL << (2 * k)) - 1ULL;
    const uint64_t shift = 2ULL * (k - 1);
    uint64_t fwd = 0, rev = 0;
    uint32_t valid = 0;
    /* RingMinQueue (:35-97). The reference buffer has `win` slots; indices at
     * or past win are out of bounds there. We keep a larger buffer and flag. */
    const uint32_t cap_ = win;
    const uint32_t big = 4 * win + 8;
    uint64_t *qh = (uint64_t *)malloc(big * sizeof(uint64_t));
    uint32_t *qp = (uint32_t *)malloc(big * sizeof(uint32_t));
    uint32_t head = 0, size = 0;
    uint64_t last_h = 0; uint32_t last_p = 0; int has_last = 0;
    int cnt = 0, overflow = 0;
#define RQ_IDX(off_) ({ uint32_t i_ = head + (off_); if (i_ >= cap_) i_ -= cap_; i_; })
#define RQ_CHECK(i_) do { if ((i_) >= cap_ && ub) *ub |= 1; if ((i_) >= big) { overflow = 1; } } while (0)
    for (uint32_t i = 0; i < n && !overflow; ++i) {
        const uint8_t c = nt4((unsigned char)seq[i]);
        if (c >= 4) {
            fwd = rev = 0; valid = 0; head = 0; size = 0; has_last = 0;
            continue;
        }
        fwd = ((fwd << 2) | c) & mask;
        rev = (rev >> 2) | ((uint64_t)(3U ^ c) << shift);
        if (valid < k) ++valid;
        if (valid < k) continue;
        const uint32_t pos = i + 1 - k;
        const uint64_t code = non_canonical ? fwd : (fwd < rev ? fwd : rev);
        const uint64_t h56 = splitmix64(code) >> 8;
        /* push */
        for (;;) {
            if (!size) break;
            uint32_t bi = RQ_IDX(size - 1);
            RQ_CHECK(bi);
            if (overflow) break;
            if (qh[bi] >= h56) --size; else break;
        }
        if (overflow) break;
        {
            uint32_t wi = RQ_IDX(size);
            RQ_CHECK(wi);
            if (overflow) break;
            if (size >= cap_ && ub) *ub |= 4;   /* full queue: front slot overwritten */
            qh[wi] = h56; qp[wi] = pos; ++size;
        }
        if (pos + 1 < win) continue;
        const uint32_t win_start = pos + 1 - win;
        while (size) {   /* popExpired */
            RQ_CHECK(head);
            if (overflow) break;
            if (qp[head] < win_start) { head = RQ_IDX(1); --size; }
            else break;
        }
        if (overflow) break;
        if (size) {
            RQ_CHECK(head);
            const uint64_t ch = qh[head];
            const uint32_t cp = qp[head];
            if (!has_last || ch != last_h || cp != last_p) {
                if (cnt >= cap) { free(qh); free(qp); return -1; }
                const uint32_t rid_strand = non_canonical ? 0x80000000U
                                                         : (fwd <= rev ? 0x80000000U : 0U);
                x_out[cnt] = (ch << 8) | (uint64_t)(uint8_t)k;
                y_out[cnt] = ((uint64_t)rid_strand << 32) | (uint64_t)cp;
                ++cnt;
                last_h = ch; last_p = cp; has_last = 1;
            }
        }
    }
#undef RQ_IDX
#undef RQ_CHECK
    free(qh); free(qp);
    if (overflow && ub) *ub |= 2;
    return cnt;
}

/* ------------------------------------------------------------------ */
/* anchors                                                              */
/* ------------------------------------------------------------------ */
typedef struct { uint64_t h; uint32_t rid, pos, strand, span; } hit_t;

static hit_t unpack_hit(uint64_t x, uint64_t y)
{
    hit_t h;
    h.h = x >> 8; h.span = (uint32_t)(x & 0xff);
    h.pos = (uint32_t)(y & 0xffffffffULL);
    h.rid = (uint32_t)(y >> 32) & 0x7fffffffU;
    h.strand = (uint32_t)(y >> 63);
    return h;
}

/* SeedHitBase::operator< include/seed.h:68-74 */
static int cmp_hit(const void *pa, const void *pb)
{
    const hit_t *a = (const hit_t *)pa, *b = (const hit_t *)pb;
    if (a->h != b->h) return a->h < b->h ? -1 : 1;
    if (a->rid != b->rid) return a->rid < b->rid ? -1 : 1;
    if (a->pos != b->pos) return a->pos < b->pos ? -1 : 1;
    if (a->strand != b->strand) return a->strand < b->strand ? -1 : 1;
    return 0;
}
static int cmp_u64(const void *pa, const void *pb)
{
    uint64_t a = *(const uint64_t *)pa, b = *(const uint64_t *)pb;
    return a < b ? -1 : a > b ? 1 : 0;
}
static int cmp_size_desc(const void *pa, const void *pb)
{
    uint64_t a = *(const uint64_t *)pa, b = *(const uint64_t *)pb;
    return a > b ? -1 : a < b ? 1 : 0;
}

int or_collect_anchors(const uint64_t *ref_x, const uint64_t *ref_y, int nref,
                       const uint64_t *qry_x, const uint64_t *qry_y, int nqry,
                       or_anchor_t *out, int cap)
{
    /* SeedFilterParams defaults, include/anchor.h:65-71 */
    const or_seed_params_t p = {2e-4, 10, 1000000, 0.01, 500};
    return or_collect_anchors_ex(ref_x, ref_y, nref, qry_x, qry_y, nqry, &p, out, cap);
}

int or_collect_anchors_ex(const uint64_t *ref_x, const uint64_t *ref_y, int nref,
                          const uint64_t *qry_x, const uint64_t *qry_y, int nqry,
                          const or_seed_params_t *p, or_anchor_t *out, int cap)
{
    if (nref == 0 || nqry == 0) return 0;
    hit_t *sr = (hit_t *)malloc((size_t)nref * sizeof(hit_t));
    for (int i = 0; i < nref; ++i) sr[i] = unpack_hit(ref_x[i], ref_y[i]);
    qsort(sr, (size_t)nref, sizeof(hit_t), cmp_hit);
    /* distinct hash intervals (:249-265) */
    uint64_t *dh = (uint64_t *)malloc((size_t)nref * 8);
    uint64_t *dstart = (uint64_t *)malloc((size_t)nref * 8);
    uint64_t *docc = (uint64_t *)malloc((size_t)nref * 8);
    int nd = 0;
    for (int i = 0; i < nref;) {
        int j = i;
        while (j < nref && sr[j].h == sr[i].h) ++j;
        dh[nd] = sr[i].h; dstart[nd] = (uint64_t)i; docc[nd] = (uint64_t)(j - i); ++nd;
        i = j;
    }
    /* compute_ref_occ_threshold src/seed/anchor.cpp:73-117 */
    uint64_t f_cutoff = UINT64_MAX;
    if (p->f_top_frac >= 1.0) {
        f_cutoff = nd > 0 ? 1 : UINT64_MAX;
    } else if (p->f_top_frac > 0.0) {
        size_t n_skip = (size_t)floor(p->f_top_frac * (double)nd);
        if (n_skip > 0) {
            uint64_t *tmp = (uint64_t *)malloc((size_t)nd * 8);
            memcpy(tmp, docc, (size_t)nd * 8);
            qsort(tmp, (size_t)nd, 8, cmp_size_desc);
            f_cutoff = tmp[n_skip - 1];
            free(tmp);
        }
    }
    uint64_t capped = f_cutoff < p->u_ceil ? f_cutoff : p->u_ceil;
    uint64_t thr = capped > p->u_floor ? capped : p->u_floor;
    /* query hash occurrence (:280-284) */
    uint64_t *qh = (uint64_t *)malloc((size_t)nqry * 8);
    for (int i = 0; i < nqry; ++i) qh[i] = qry_x[i] >> 8;
    qsort(qh, (size_t)nqry, 8, cmp_u64);
    const double q_occ_limit = p->q_occ_frac * (double)nqry;   /* used only when q_occ_frac > 0 */
    int cnt = 0, ok = 1;
    for (int iq = 0; iq < nqry && ok; ++iq) {
        hit_t qhit = unpack_hit(qry_x[iq], qry_y[iq]);
        /* hash_index.find: binary search in distinct list */
        int lo = 0, hi = nd - 1, found = -1;
        while (lo <= hi) {
            int mid = (lo + hi) >> 1;
            if (dh[mid] == qhit.h) { found = mid; break; }
            if (dh[mid] < qhit.h) lo = mid + 1; else hi = mid - 1;
        }
        if (found < 0) continue;
        /* qocc: count of this hash in the query */
        size_t a = 0, b = (size_t)nqry;
        while (a < b) { size_t m = (a + b) >> 1; if (qh[m] < qhit.h) a = m + 1; else b = m; }
        size_t c0 = a; b = (size_t)nqry;
        while (a < b) { size_t m = (a + b) >> 1; if (qh[m] <= qhit.h) a = m + 1; else b = m; }
        uint64_t qocc = (uint64_t)(a - c0);
        if (p->q_occ_frac > 0.0 && (double)qocc > q_occ_limit) continue;
        uint64_t ref_occ = docc[found];
        if (ref_occ > thr) {
            if (p->sample_every_bp == 0) continue;
            if (qhit.pos % p->sample_every_bp != 0) continue;
        }
        for (uint64_t i = dstart[found]; i < dstart[found] + ref_occ; ++i) {
            if (cnt >= cap) { ok = 0; break; }
            or_anchor_t an;
            an.hash = qhit.h;
            an.rid_ref = sr[i].rid; an.pos_ref = sr[i].pos;
            an.rid_qry = qhit.rid; an.pos_qry = qhit.pos;
            an.span = sr[i].span < qhit.span ? sr[i].span : qhit.span;
            an.is_rev = sr[i].strand != qhit.strand;
            out[cnt++] = an;
        }
    }
    free(sr); free(dh); free(dstart); free(docc); free(qh);
    return ok ? cnt : -1;
}

/* ------------------------------------------------------------------ */
/* chaining                                                             */
/* ------------------------------------------------------------------ */
void or_fill_log2_lut(float *lut)
{
    lut[0] = 0.0f;
    for (int d = 1; d <= 501; ++d) lut[d] = log2f((float)d);
}

static int cmp_anchor_chain(const void *pa, const void *pb)
{
    const or_anchor_t *a = (const or_anchor_t *)pa, *b = (const or_anchor_t *)pb;
    if (a->rid_ref != b->rid_ref) return a->rid_ref < b->rid_ref ? -1 : 1;
    if (a->is_rev != b->is_rev) return a->is_rev < b->is_rev ? -1 : 1;
    if (a->pos_ref != b->pos_ref) return a->pos_ref < b->pos_ref ? -1 : 1;
    if (a->pos_qry != b->pos_qry) return a->pos_qry < b->pos_qry ? -1 : 1;
    return 0;
}

/* chainScoreSimple src/seed/anchor.cpp:254-301 with ChainParams defaults
 * include/anchor.h:116-126. Float ops in program order, no contraction. */
static int32_t chain_score(const or_anchor_t *ai, const or_anchor_t *aj, const or_chain_params_t *cp,
                           const float *lut)
{
    if (ai->rid_ref != aj->rid_ref) return INT32_MIN;
    if (ai->is_rev != aj->is_rev) return INT32_MIN;
    const int32_t dq = (int32_t)ai->pos_qry - (int32_t)aj->pos_qry;
    if (dq <= 0 || dq > cp->max_dist_x) return INT32_MIN;
    const int32_t dr = (int32_t)ai->pos_ref - (int32_t)aj->pos_ref;
    if (dr <= 0 || dr > cp->max_dist_y) return INT32_MIN;
    const int32_t dd = dr > dq ? dr - dq : dq - dr;
    if (dd > cp->bw) return INT32_MIN;
    const int32_t dg = dr < dq ? dr : dq;
    const int32_t q_span = (int32_t)aj->span;
    int32_t sc = q_span < dg ? q_span : dg;
    if (dd > 0 || dg > q_span) {
        /* each float operation rounded separately, no contraction (the
         * reference's production build has no FMA) */
        volatile float t1 = cp->gap_penalty * (float)dd;
        volatile float t2 = cp->skip_penalty * (float)dg;
        volatile float lin_pen = t1 + t2;
        const float log_pen = dd >= 1 ? lut[dd + 1] : 0.0f;
        volatile float half = 0.5f * log_pen;
        volatile float tot = lin_pen + half;
        sc -= (int32_t)tot;
    }
    return sc;
}

int or_chain_anchors(or_anchor_t *anchors, int n, const float *lut, or_anchor_t *out)
{
    /* ChainParams defaults, include/anchor.h:116-126 */
    const or_chain_params_t cp = {5000, 5000, 500, 25, 5000, 3, 40, 0.01f, 0.01f};
    return or_chain_anchors_ex(anchors, n, &cp, lut, out);
}

int or_chain_anchors_ex(or_anchor_t *anchors, int n, const or_chain_params_t *cp, const float *lut,
                        or_anchor_t *out)
{
    if (n == 0) return 0;
    if (cp->bw > 500) return -1;   /* lut covers dd + 1 <= 501 */
    qsort(anchors, (size_t)n, sizeof(or_anchor_t), cmp_anchor_chain);
    int32_t *f = (int32_t *)malloc((size_t)n * 4);
    int64_t *p = (int64_t *)calloc((size_t)n, 8);
    int32_t *t = (int32_t *)calloc((size_t)n, 4);
    int64_t st = 0, best_end = -1;
    int32_t max_f_global = 0;
    for (int64_t i = 0; i < n; ++i) {
        const or_anchor_t *ai = &anchors[i];
        int32_t max_f = (int32_t)ai->span;
        int64_t max_j = -1;
        int32_t n_skip = 0;
        while (st < i) {
            const or_anchor_t *as = &anchors[st];
            if (as->rid_ref != ai->rid_ref || as->is_rev != ai->is_rev ||
                (int32_t)(ai->pos_ref - as->pos_ref) > cp->max_dist_x) ++st;
            else break;
        }
        int64_t iter_start = (i - st > cp->max_iter) ? (i - cp->max_iter) : st;
        for (int64_t j = i - 1; j >= iter_start; --j) {
            const int32_t sc = chain_score(ai, &anchors[j], cp, lut);
            if (sc == INT32_MIN) continue;
            const int32_t total = f[j] + sc;
            if (total > max_f) {
                max_f = total; max_j = j;
                if (n_skip > 0) --n_skip;
            } else if (t[j] == (int32_t)i) {
                if (++n_skip > cp->max_skip) break;
            }
            if (p[j] >= 0) t[p[j]] = (int32_t)i;
        }
        f[i] = max_f; p[i] = max_j;
        if (max_f > max_f_global) { max_f_global = max_f; best_end = i; }
    }
    int cnt = 0;
    if (!(max_f_global < cp->min_score || best_end < 0)) {
        int64_t cur = best_end;
        while (cur >= 0) { ++cnt; cur = p[cur]; }
        if (cnt < cp->min_cnt) cnt = 0;
        else {
            cur = best_end;
            for (int k = cnt - 1; k >= 0; --k) { out[k] = anchors[cur]; cur = p[cur]; }
        }
    }
    free(f); free(p); free(t);
    return cnt;
}

/* ------------------------------------------------------------------ */
/* CIGAR helpers: src/align/cigar.cpp                                   */
/* ------------------------------------------------------------------ */
static int64_t cig_ref_len(const uint32_t *c, int n)
{
    int64_t s = 0;   /* M D N = X consume ref (:332-347) */
    for (int i = 0; i < n; ++i) {
        uint32_t op = c[i] & 0xf;
        if (op == 0 || op == 2 || op == 3 || op == 7 || op == 8) s += c[i] >> 4;
    }
    return s;
}
static int64_t cig_qry_len(const uint32_t *c, int n)
{
    int64_t s = 0;   /* M I S = X consume query (:350-365) */
    for (int i = 0; i < n; ++i) {
        uint32_t op = c[i] & 0xf;
        if (op == 0 || op == 1 || op == 4 || op == 7 || op == 8) s += c[i] >> 4;
    }
    return s;
}
/* appendCigar (:302-329) */
static void cig_append(cigbuf_t *cb, const uint32_t *c, int n)
{
    for (int i = 0; i < n; ++i) {
        uint32_t len = c[i] >> 4, op = c[i] & 0xf;
        if (len == 0) continue;
        if (cb->n > 0 && (cb->c[cb->n - 1] & 0xf) == op) cb->c[cb->n - 1] += len << 4;
        else {
            if (cb->n == cb->m) {
                cb->m = cb->m ? cb->m << 1 : 64;
                cb->c = (uint32_t *)realloc(cb->c, (size_t)cb->m * 4);
            }
            cb->c[cb->n++] = c[i];
        }
    }
}

int or_has_insertion(const uint32_t *c, int n)
{
    for (int i = 0; i < n; ++i) if ((c[i] & 0xf) == 1) return 1;
    return 0;
}

int or_cigar_to_string(const uint32_t *c, int n, char *out, int cap)
{
    static const char ops[] = "MIDNSHP=X";
    int len = 0;
    for (int i = 0; i < n; ++i) {
        char buf[16];
        int m = 0; uint32_t l = c[i] >> 4;
        do { buf[m++] = (char)('0' + l % 10); l /= 10; } while (l);
        if (len + m + 1 > cap) return -1;
        while (m) out[len++] = buf[--m];
        out[len++] = ops[(c[i] & 0xf) < 9 ? (c[i] & 0xf) : 0];
    }
    if (len < cap) out[len] = 0;
    return len;
}

int or_del_query_to_ref(const char *qry, int qlen, const uint32_t *c, int n,
                        char *out, int cap)
{
    int64_t out_len = 0, consume = 0;
    for (int i = 0; i < n; ++i) {
        uint32_t op = c[i] & 0xf, l = c[i] >> 4;
        if (op == 1) consume += l;
        else if (op == 5 || op == 6) continue;
        else if (op == 2 || op == 3) out_len += l;
        else { out_len += l; consume += l; }
    }
    if (n == 0) {   /* empty CIGAR: query returned unchanged (:214-216) */
        if (qlen > cap) return -1;
        memcpy(out, qry, (size_t)qlen);
        return qlen;
    }
    if (consume != qlen || out_len > cap) return -1;
    int64_t w = 0, r = 0;
    for (int i = 0; i < n; ++i) {
        uint32_t op = c[i] & 0xf, l = c[i] >> 4;
        if (op == 1) { r += l; continue; }
        if (op == 5 || op == 6) continue;
        if (op == 2 || op == 3) { memset(out + w, '-', l); w += l; continue; }
        memcpy(out + w, qry + r, l); w += l; r += l;
    }
    return (int)out_len;
}

/* ------------------------------------------------------------------ */
/* globalAlignMM2 src/align/align.cpp:189-305                           */
/* ------------------------------------------------------------------ */
static int cmp_anchor_q(const void *pa, const void *pb)
{
    const or_anchor_t *a = (const or_anchor_t *)pa, *b = (const or_anchor_t *)pb;
    if (a->pos_qry != b->pos_qry) return a->pos_qry < b->pos_qry ? -1 : 1;
    if (a->pos_ref != b->pos_ref) return a->pos_ref < b->pos_ref ? -1 : 1;
    return 0;
}

typedef struct {
    const char *ref, *qry; size_t rlen, qlen, ref_pos, qry_pos;
    cigbuf_t res; int64_t cells; int32_t nseg, nforced; uint32_t *tmp; int tmpcap;
} mm2_t;

static void mm2_segment(mm2_t *m, size_t rs, size_t re, size_t qs, size_t qe, int flag)
{
    if (rs > m->rlen) rs = m->rlen;
    if (re > m->rlen) re = m->rlen;
    if (qs > m->qlen) qs = m->qlen;
    if (qe > m->qlen) qe = m->qlen;
    if (re < rs) re = rs;
    if (qe < qs) qe = qs;
    const int srl = (int)(re - rs), sql = (int)(qe - qs);
    int64_t cells = 0;
    int need = srl + sql + 4;
    if (need > m->tmpcap) {
        m->tmpcap = need;
        m->tmp = (uint32_t *)realloc(m->tmp, (size_t)need * 4);
    }
    int nc = or_global_align_ksw2(m->ref + rs, srl, m->qry + qs, sql, flag, m->tmp, m->tmpcap, &cells);
    m->cells += cells;
    m->nseg++;
    if (cig_ref_len(m->tmp, nc) != srl || cig_qry_len(m->tmp, nc) != sql) {
        uint32_t forced[2]; int nf = 0;
        if (sql > 0) forced[nf++] = (uint32_t)sql << 4 | 1;
        if (srl > 0) forced[nf++] = (uint32_t)srl << 4 | 2;
        cig_append(&m->res, forced, nf);
        m->nforced++;
        m->ref_pos = re; m->qry_pos = qe;
        return;
    }
    cig_append(&m->res, m->tmp, nc);
    m->ref_pos = rs + (size_t)srl;
    m->qry_pos = qs + (size_t)sql;
}

static int emit_cigar(const uint32_t *c, int n, uint32_t *cigar, int cap)
{
    if (n > cap) return -1;
    memcpy(cigar, c, (size_t)n * 4);
    return n;
}

int or_global_align_mm2(const char *ref, int rlen, const char *qry, int qlen,
                        const or_anchor_t *anchors, int nanchors, const float *lut,
                        uint32_t *cigar, int cigar_cap, int64_t *cells, int32_t *stats)
{
    or_anchor_t *sorted = (or_anchor_t *)malloc(((size_t)nanchors + 1) * sizeof(or_anchor_t));
    or_anchor_t *chain = (or_anchor_t *)malloc(((size_t)nanchors + 1) * sizeof(or_anchor_t));
    if (nanchors) memcpy(sorted, anchors, (size_t)nanchors * sizeof(or_anchor_t));
    int nch = or_chain_anchors(sorted, nanchors, lut, chain);
    if (stats) { stats[0] = stats[1] = stats[2] = 0; stats[3] = nch; }
    if (cells) *cells = 0;
    int ret;
    if (nch == 0) {
        free(sorted); free(chain);
        if (stats) stats[2] = 1;
        int64_t c = 0;
        ret = or_global_align_ksw2(ref, rlen, qry, qlen, OR_KSW_GENERIC_SC | OR_KSW_RIGHT,
                                   cigar, cigar_cap, &c);
        if (cells) *cells = c;
        return ret;
    }
    qsort(chain, (size_t)nch, sizeof(or_anchor_t), cmp_anchor_q);
    mm2_t m;
    memset(&m, 0, sizeof m);
    m.ref = ref; m.qry = qry; m.rlen = (size_t)rlen; m.qlen = (size_t)qlen;
    const int RF = OR_KSW_GENERIC_SC | OR_KSW_RIGHT, LF = OR_KSW_GENERIC_SC;
    mm2_segment(&m, m.ref_pos, chain[0].pos_ref, m.qry_pos, chain[0].pos_qry, LF);
    for (int i = 0; i < nch; ++i) {
        const or_anchor_t *a = &chain[i];
        mm2_segment(&m, m.ref_pos, (size_t)a->pos_ref + a->span, m.qry_pos,
                    (size_t)a->pos_qry + a->span, RF);
        if (i + 1 < nch)
            mm2_segment(&m, m.ref_pos, chain[i + 1].pos_ref, m.qry_pos, chain[i + 1].pos_qry, RF);
    }
    mm2_segment(&m, m.ref_pos, m.rlen, m.qry_pos, m.qlen, RF);
    if (stats) { stats[0] = m.nseg; stats[1] = m.nforced; }
    if (cig_ref_len(m.res.c, m.res.n) != rlen || cig_qry_len(m.res.c, m.res.n) != qlen) {
        if (stats) stats[2] = 1;
        int64_t c = 0;
        ret = or_global_align_ksw2(ref, rlen, qry, qlen, RF, cigar, cigar_cap, &c);
        if (cells) *cells = m.cells + c;
    } else {
        ret = emit_cigar(m.res.c, m.res.n, cigar, cigar_cap);
        if (cells) *cells = m.cells;
    }
    free(m.res.c); free(m.tmp); free(sorted); free(chain);
    return ret;
}

int or_ref_global_align(const char *ref, int rlen, const char *qry, int qlen,
                        const uint64_t *ref_x, const uint64_t *ref_y, int nref,
                        const float *lut, uint32_t *cigar, int cigar_cap,
                        int64_t *cells, int32_t *stats, int *ub)
{
    int ub1 = 0, ub2 = 0;
    uint64_t *rx = 0, *ry = 0;
    if (nref == 0) {
        int cap = rlen + 1;
        rx = (uint64_t *)malloc((size_t)cap * 16); ry = rx + cap;
        nref = or_extract_minimizer(ref, rlen, 15, 10, 1, rx, ry, cap, &ub1);
        ref_x = rx; ref_y = ry;
    }
    int qcap = qlen + 1;
    uint64_t *qx = (uint64_t *)malloc((size_t)qcap * 16), *qy = qx + qcap;
    int nq = or_extract_minimizer(qry, qlen, 15, 10, 1, qx, qy, qcap, &ub2);
    /* anchors: each query hit expands to at most its ref occurrence count */
    int acap = nq * 16 + 16, na;
    or_anchor_t *an = 0;
    for (;;) {
        an = (or_anchor_t *)realloc(an, (size_t)acap * sizeof(or_anchor_t));
        na = or_collect_anchors(ref_x, ref_y, nref, qx, qy, nq, an, acap);
        if (na >= 0) break;
        acap *= 4;
    }
    int ret = or_global_align_mm2(ref, rlen, qry, qlen, an, na, lut, cigar, cigar_cap, cells, stats);
    if (ub) *ub = ub1 | ub2;
    free(an); free(qx); free(rx);
    return ret;
}

/* ------------------------------------------------------------------ */
/* consensus counting                                                   */
/* ------------------------------------------------------------------ */
static inline int base_idx(unsigned char c)
{
    switch (c) {   /* k_base_map include/consensus.h:126-145 */
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    case 'U': case 'u': return 4;
    case '-': case '.': return 6;
    default: return 5;
    }
}

int64_t or_consensus(const char *rows, int64_t nrows, int64_t aln_len,
                     int64_t batch_size, uint64_t seq_limit,
                     uint32_t *counts, char *cons_out)
{
    if (nrows <= 0 || aln_len <= 0 || batch_size <= 0) return -1;
    memset(counts, 0, (size_t)aln_len * 7 * 4);
    /* Reader cursor model of :620-685: the first record starts batch 1;
     * after each processed batch one record is read and discarded. */
    int64_t next = 0;          /* index of the next record the reader yields */
    uint64_t num_seqs = 0;
    int64_t in_batch = 0;
    /* first record */
    {
        const char *row = rows;
        for (int64_t i = 0; i < aln_len; ++i) counts[i * 7 + base_idx((unsigned char)row[i])]++;
        next = 1; num_seqs = 1; in_batch = 1;
    }
    while (seq_limit == 0 || num_seqs < seq_limit) {
        while (in_batch < batch_size && (seq_limit == 0 || num_seqs < seq_limit)) {
            if (next >= nrows) break;
            const char *row = rows + next * aln_len;
            for (int64_t i = 0; i < aln_len; ++i) counts[i * 7 + base_idx((unsigned char)row[i])]++;
            ++next; ++in_batch; ++num_seqs;
        }
        in_batch = 0;              /* batch processed and cleared */
        if (next >= nrows) break;  /* reader.next(rec) at EOF */
        ++next;                    /* this record is read and never counted */
    }
    for (int64_t i = 0; i < aln_len; ++i) {
        const uint32_t *sc = counts + i * 7;
        uint32_t best = sc[0]; char ch = 'A';
        if (sc[1] > best) best = sc[1], ch = 'C';
        if (sc[2] > best) best = sc[2], ch = 'G';
        if (sc[3] > best) best = sc[3], ch = 'T';
        if (sc[4] > best) best = sc[4], ch = 'U';
        cons_out[i] = ch;
    }
    return (int64_t)num_seqs;
}

/* ------------------------------------------------------------------ */
/* FASTA/FASTQ records: kseq_read (include/kseq.h) as KseqReader uses it */
/* ------------------------------------------------------------------ */
typedef struct {
    const unsigned char *b; int64_t n, i;
} fx_src_t;

typedef struct {
    char *p; int64_t l, m;
} fx_str_t;

static void fx_put(fx_str_t *s, char c)
{
    if (s->l + 1 > s->m) {
        s->m = s->m ? 2 * s->m : 64;
        s->p = (char *)realloc(s->p, (size_t)s->m);
    }
    s->p[s->l++] = c;
}

static int fx_space(int c)
{
    return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r';
}

/* ks_getuntil2: bytes up to the delimiter (a whitespace byte when space != 0,
 * else '\n'; consumed, returned in *dret, 0 at the end of the data) are
 * appended to dst. Returns -1, reading nothing and trimming nothing, when the
 * data was already exhausted. Otherwise a line read (space == 0) leaves dst
 * one '\r' shorter when dst, counted whole, is longer than one byte and ends
 * in '\r'. */
static int fx_until(fx_src_t *s, int space, fx_str_t *dst, int *dret)
{
    if (dret) *dret = 0;
    if (s->i >= s->n) return -1;
    while (s->i < s->n) {
        const int c = s->b[s->i++];
        if (space ? fx_space(c) : c == '\n') {
            if (dret) *dret = c;
            break;
        }
        fx_put(dst, (char)c);
    }
    if (!space && dst->l > 1 && dst->p[dst->l - 1] == '\r') --dst->l;
    return 0;
}

static int fx_getc(fx_src_t *s) { return s->i < s->n ? s->b[s->i++] : -1; }

/* kseq_read: 1 = a record in id/desc/seq, 0 = end of data, -2 = malformed
 * FASTQ. *hdr carries the header byte already consumed ('>' / '@', 0 = none). */
static int fx_read(fx_src_t *s, int *hdr, fx_str_t *id, fx_str_t *desc, fx_str_t *seq, fx_str_t *qual)
{
    int c;
    if (*hdr == 0) {
        while ((c = fx_getc(s)) >= 0 && c != '>' && c != '@') {}
        if (c < 0) return 0;
        *hdr = c;
    }
    id->l = desc->l = seq->l = qual->l = 0;
    if (fx_until(s, 1, id, &c) < 0) return 0;
    if (c != '\n') fx_until(s, 0, desc, NULL);
    while ((c = fx_getc(s)) >= 0 && c != '>' && c != '+' && c != '@') {
        if (c == '\n') continue;
        fx_put(seq, (char)c);
        fx_until(s, 0, seq, NULL);
    }
    if (c == '>' || c == '@') *hdr = c;
    if (c != '+') return 1;
    while ((c = fx_getc(s)) >= 0 && c != '\n') {}
    if (c < 0) return -2;
    while (fx_until(s, 0, qual, NULL) >= 0 && qual->l < seq->l) {}
    *hdr = 0;
    return seq->l == qual->l ? 1 : -2;
}

static int fx_emit(const fx_str_t *v, char *out, int64_t *off, int k, int64_t cap)
{
    if (off[k] + v->l > cap) return -1;
    if (v->l) memcpy(out + off[k], v->p, (size_t)v->l);
    off[k + 1] = off[k] + v->l;
    return 0;
}

int or_fastx_parse(const char *buf, int64_t len, int max_records,
                   char *id_out, int64_t *id_off, char *desc_out, int64_t *desc_off,
                   char *seq_out, int64_t *seq_off, int64_t cap)
{
    fx_src_t s = {(const unsigned char *)buf, len, 0};
    fx_str_t id = {0}, desc = {0}, seq = {0}, qual = {0};
    int hdr = 0, n = 0, rc = 0;
    id_off[0] = desc_off[0] = seq_off[0] = 0;
    while (n < max_records) {
        rc = fx_read(&s, &hdr, &id, &desc, &seq, &qual);
        if (rc <= 0) break;
        if (fx_emit(&id, id_out, id_off, n, cap) || fx_emit(&desc, desc_out, desc_off, n, cap) ||
            fx_emit(&seq, seq_out, seq_off, n, cap)) {
            rc = -1;
            break;
        }
        ++n;
        rc = 0;
    }
    free(id.p); free(desc.p); free(seq.p); free(qual.p);
    return rc < 0 ? rc : n;
}

/* ------------------------------------------------------------------ */
/* keep-length merge: RefAligner::mergeAlignedResults, one center       */
/* ------------------------------------------------------------------ */
/* seq_io::cleanSequence's table (include/utils.h:161-181) */
static char clean_char(unsigned char c)
{
    switch (c) {
    case 'A': case 'a': return 'A';
    case 'C': case 'c': return 'C';
    case 'G': case 'g': return 'G';
    case 'T': case 't': return 'T';
    case 'U': case 'u': return 'U';
    case '-': return '-';
    default: return 'N';
    }
}

int64_t or_keep_length_fasta(const char *center_id, int id_len, const char *center_desc, int desc_len,
                             const char *center_seq, int64_t center_len, int n, const char *queries,
                             const int64_t *qoff, const char *names, const int64_t *noff,
                             const uint32_t *cigars, const int64_t *coff, char *out, int64_t cap)
{
    int64_t o = 0, expect = -1;
#define OR_PUT(src, len_)                                          \
    do {                                                           \
        if (o + (int64_t)(len_) > cap) return -1;                  \
        if ((len_) > 0) memcpy(out + o, (src), (size_t)(len_));    \
        o += (int64_t)(len_);                                      \
    } while (0)
    OR_PUT(">", 1);
    OR_PUT(center_id, id_len);
    if (desc_len > 0) {
        OR_PUT(" ", 1);
        OR_PUT(center_desc, desc_len);
    }
    OR_PUT("\n", 1);
    if (o + center_len + 1 > cap) return -1;
    for (int64_t i = 0; i < center_len; ++i) out[o + i] = clean_char((unsigned char)center_seq[i]);
    o += center_len;
    OR_PUT("\n", 1);   /* writeFasta: an empty sequence is an empty line */
    /* insertion group first, then the others */
    for (int pass = 0; pass < 2; ++pass) {
        for (int q = 0; q < n; ++q) {
            const uint32_t *c = cigars + coff[q];
            const int nc = (int)(coff[q + 1] - coff[q]);
            if (or_has_insertion(c, nc) != (pass == 0)) continue;
            const int64_t nl = noff[q + 1] - noff[q];
            /* SamReader::next skips lines starting with '@' (src/utils/seq_io.cpp:
               433-436): a query named "@..." does not come back from its SAM */
            if (nl > 0 && names[noff[q]] == '@') continue;
            OR_PUT(">", 1);
            OR_PUT(names + noff[q], nl);
            OR_PUT("\n", 1);
            if (o >= cap) return -1;
            const int rl = or_del_query_to_ref(queries + qoff[q], (int)(qoff[q + 1] - qoff[q]), c, nc, out + o,
                                               (int)((cap - o) < 0x7fffffff ? cap - o : 0x7fffffff));
            if (rl < 0) return -1;
            if (expect < 0) expect = rl;
            else if (rl != expect) return -1;
            o += rl;
            OR_PUT("\n", 1);
        }
    }
#undef OR_PUT
    return o;
}

/* ------------------------------------------------------------------ */
/* mash sketches: src/mash/mash.cpp, src/mash/hash.cpp                 */
/* ------------------------------------------------------------------ */
static uint64_t rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }

uint64_t or_hash2bit(uint64_t code, uint32_t seed)
{
    /* the 8 input bytes are the code's big-endian bytes; the first and the
       last four read little-endian */
    const uint32_t in1 = __builtin_bswap32((uint32_t)(code >> 32));
    const uint32_t in2 = __builtin_bswap32((uint32_t)code);
    uint64_t s = seed;
    s ^= (uint64_t)__builtin_bswap32((uint32_t)s) << 32;
    /* default secret, bytes 8..15 and 16..23 read little-endian */
    const uint64_t bitflip = (0x1cad21f72c81017cULL ^ 0xdb979083e96dd4deULL) - s;
    uint64_t h = ((uint64_t)in2 + ((uint64_t)in1 << 32)) ^ bitflip;
    const uint64_t mx2 = 0x9fb21c651e98df25ULL;
    h ^= rotl64(h, 49) ^ rotl64(h, 24);
    h *= mx2;
    h ^= (h >> 35) + 8;
    h *= mx2;
    return h ^ (h >> 28);
}

int or_mash_sketch(const char *seq, int64_t n, int k, int sketch_size, int noncanonical, uint32_t seed,
                   uint64_t *out, int cap)
{
    if (k == 0 || sketch_size <= 0 || n < k || k > 32) return 0;
    /* (1 << 2k) - 1 with the shift count taken mod 64 as the reference's x86
       build computes it: k = 32 shifts by 64 there (undefined in C++) and the
       mask comes out 0, so every k-mer codes as 0 */
    const uint64_t mask = (1ULL << ((2 * k) & 63)) - 1ULL;
    const int shift = 2 * (k - 1);
    uint64_t *all = (uint64_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(uint64_t));
    int64_t m = 0;
    uint64_t fwd = 0, rev = 0;
    int valid = 0;
    for (int64_t i = 0; i < n; ++i) {
        int c;
        switch (seq[i]) {
        case 'A': case 'a': c = 0; break;
        case 'C': case 'c': c = 1; break;
        case 'G': case 'g': c = 2; break;
        case 'T': case 't': case 'U': case 'u': c = 3; break;
        default: c = 4;
        }
        if (c == 4) {
            fwd = rev = 0;
            valid = 0;
            continue;
        }
        fwd = ((fwd << 2) | (uint64_t)c) & mask;
        rev = (rev >> 2) | ((uint64_t)(3 ^ c) << shift);
        if (valid < k) ++valid;
        if (valid < k) continue;
        const uint64_t code = noncanonical ? fwd : (fwd < rev ? fwd : rev);
        all[m++] = or_hash2bit(code, seed);
    }
    qsort(all, (size_t)m, sizeof(uint64_t), cmp_u64);
    int cnt = 0;
    for (int64_t i = 0; i < m && cnt < sketch_size; ++i) {
        if (i > 0 && all[i] == all[i - 1]) continue;
        if (cnt >= cap) { free(all); return -1; }
        out[cnt++] = all[i];
    }
    free(all);
    return cnt;
}

double or_mash_jaccard(const uint64_t *a, int na, const uint64_t *b, int nb)
{
    if (na == 0 && nb == 0) return 1.0;
    if (na == 0 || nb == 0) return 0.0;
    int i = 0, j = 0, inter = 0;
    while (i < na && j < nb) {
        if (a[i] < b[j]) ++i;
        else if (b[j] < a[i]) ++j;
        else { ++inter; ++i; ++j; }
    }
    return (double)inter / (double)(na < nb ? na : nb);
}
