/*
 * synth.c -- seeded synthetic workloads for the five benchmark
 * configurations (center sequence + mutated queries). Host-only helper used
 * by bench.py and the tests; identical output on every machine.
 *
 * PRNG: xoshiro256** seeded through splitmix64. Base = top 2 bits of a draw.
 * Mutation model, applied per source position: with p_sub substitute one of
 * the 3 other bases uniformly; else with p_indel an indel event, insertion or
 * deletion with equal odds, of length l with P(l) = 2^-l (or, with
 * probability p_long, l uniform in [long_min, long_max]); an insertion emits l
 * random bases before the current base, a deletion skips l bases.
 * Query i of configuration c is generated from seed (0xC0+c) << 32 | i.
 * Integer arithmetic and exact double comparisons only: the same bytes on
 * every machine (tests/test_synth.py holds checksums).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "synth_tables.h"

typedef struct { uint64_t s[4]; } rng_t;

static uint64_t sm64(uint64_t *x)
{
    uint64_t z = (*x += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
static void rng_seed(rng_t *r, uint64_t seed)
{
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) r->s[i] = sm64(&x);
}
static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static uint64_t next(rng_t *r)
{
    uint64_t *s = r->s;
    const uint64_t res = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
    return res;
}
static double unif(rng_t *r) { return (double)(next(r) >> 11) * (1.0 / 9007199254740992.0); }
static const char B[4] = {'A', 'C', 'G', 'T'};
static int bidx(char c) { return c == 'A' ? 0 : c == 'C' ? 1 : c == 'G' ? 2 : 3; }

typedef struct { char *p; int64_t n, cap; int overflow; } buf_t;
static inline void put(buf_t *b, char c)
{
    if (b->n < b->cap) b->p[b->n] = c; else b->overflow = 1;
    b->n++;
}

static void mutate(rng_t *r, const char *src, int64_t len, double p_sub, double p_indel,
                   double p_long, int long_min, int long_max, buf_t *out)
{
    int64_t i = 0;
    while (i < len) {
        const double u = unif(r);
        if (u < p_sub) {
            put(out, B[(bidx(src[i]) + 1 + (int)(next(r) % 3)) & 3]);
            ++i;
        } else if (u < p_sub + p_indel) {
            int64_t l;
            if (p_long > 0 && unif(r) < p_long) l = long_min + (int64_t)(next(r) % (uint64_t)(long_max - long_min + 1));
            else { l = 1; while (next(r) >> 63) ++l; }
            if (next(r) >> 63) {
                for (int64_t k = 0; k < l; ++k) put(out, B[next(r) >> 62]);
                put(out, src[i]);
                ++i;
            } else {
                i += l;
            }
        } else {
            put(out, src[i]);
            ++i;
        }
    }
}

/* center lengths of configurations 1..5 */
static const int64_t kCenterLen[6] = {0, 16569, 29903, 20000, 10000, 29903};

int64_t synth_center_len(int cfg) { return (cfg >= 1 && cfg <= 5) ? kCenterLen[cfg] : -1; }

int64_t synth_center(int cfg, char *out, int64_t cap)
{
    if (cfg < 1 || cfg > 5) return -1;
    const int64_t n = kCenterLen[cfg];
    if (cap < n) return -n;
    rng_t r;
    rng_seed(&r, 0xC0u + (uint64_t)cfg);
    for (int64_t i = 0; i < n; ++i) out[i] = B[next(&r) >> 62];
    return n;
}

/* lognormal(median 1500, sigma 0.8) clamped to [200, 20000]: the committed
 * quantile table, entry picked by the top 12 bits of one draw (integers only,
 * so the lengths do not depend on the machine's libm) */
static int64_t lognormal_len(rng_t *r) { return kLognormalLen[next(r) >> 52]; }

/* one query of configuration cfg into out; returns its length (or -needed) */
static int64_t one_query(int cfg, int64_t idx, const char *center, char *tmp, int64_t tmpcap,
                         char *out, int64_t cap)
{
    rng_t r;
    rng_seed(&r, ((uint64_t)(0xC0u + cfg) << 32) | (uint64_t)idx);
    const int64_t cl = kCenterLen[cfg];
    buf_t o = {out, 0, cap, 0};
    if (cfg == 1) {
        mutate(&r, center, cl, 0.01, 0.002, 0, 0, 0, &o);
    } else if (cfg == 2) {
        buf_t m = {tmp, 0, tmpcap, 0};
        mutate(&r, center, cl, 0.005, 0.001, 0, 0, 0, &m);
        if (m.overflow) return -1;
        const int64_t L = 29000 + (int64_t)(next(&r) % 1001);
        if (L <= m.n) {
            const int64_t st = (int64_t)(next(&r) % (uint64_t)(m.n - L + 1));
            for (int64_t k = 0; k < L; ++k) put(&o, tmp[st + k]);
        } else {
            const int64_t extra = L - m.n, left = extra / 2;
            for (int64_t k = 0; k < left; ++k) put(&o, B[next(&r) >> 62]);
            for (int64_t k = 0; k < m.n; ++k) put(&o, tmp[k]);
            for (int64_t k = 0; k < extra - left; ++k) put(&o, B[next(&r) >> 62]);
        }
    } else if (cfg == 3) {
        const int64_t L = lognormal_len(&r);
        const int64_t st = (int64_t)(next(&r) % (uint64_t)(cl - L + 1));
        mutate(&r, center + st, L, 0.01, 0.002, 0, 0, 0, &o);
    } else if (cfg == 4) {
        mutate(&r, center, cl, 0.05, 0.01, 1.0 / 50.0, 50, 500, &o);
    } else if (cfg == 5) {
        const int64_t L = 1 + (int64_t)(next(&r) % 30000);
        if (L <= cl) {
            const int64_t st = (int64_t)(next(&r) % (uint64_t)(cl - L + 1));
            mutate(&r, center + st, L, 0.008, 0.002, 0, 0, 0, &o);
        } else {
            if (tmpcap < L) return -1;
            memcpy(tmp, center, (size_t)cl);
            for (int64_t k = cl; k < L; ++k) tmp[k] = B[next(&r) >> 62];
            mutate(&r, tmp, L, 0.008, 0.002, 0, 0, 0, &o);
        }
    } else {
        return -1;
    }
    return o.overflow ? -o.n : o.n;
}

/* Queries [first, first+count) of configuration cfg, concatenated into out
 * with offsets[count+1]. Returns total bytes, or a negative value if out is
 * too small (then -returned is a safe capacity). Uses OpenMP when built with it. */
int64_t synth_queries(int cfg, int64_t first, int64_t count, char *out, int64_t cap, int64_t *offsets)
{
    if (cfg < 1 || cfg > 5 || count < 0) return 0;
    const int64_t cl = kCenterLen[cfg];
    char *center = (char *)malloc((size_t)cl);
    synth_center(cfg, center, cl);
    const int64_t per = 3 * 30000 + 4096;   /* upper bound of one query of any config */
    int64_t *lens = (int64_t *)calloc((size_t)count + 1, sizeof(int64_t));
    char *work = (char *)malloc((size_t)per);
    /* pass 1: lengths (sequential generation is cheap relative to the copies) */
    int bad = 0;
#pragma omp parallel
    {
        char *tmp = (char *)malloc((size_t)per), *q = (char *)malloc((size_t)per);
#pragma omp for schedule(dynamic, 64)
        for (int64_t i = 0; i < count; ++i) {
            const int64_t n = one_query(cfg, first + i, center, tmp, per, q, per);
            if (n < 0) bad = 1;
            lens[i] = n;
        }
        free(tmp); free(q);
    }
    free(work);
    if (bad) { free(center); free(lens); return 0; }
    int64_t tot = 0;
    for (int64_t i = 0; i < count; ++i) { offsets[i] = tot; tot += lens[i]; }
    offsets[count] = tot;
    if (tot > cap) { free(center); free(lens); return -tot; }
#pragma omp parallel
    {
        char *tmp = (char *)malloc((size_t)per);
#pragma omp for schedule(dynamic, 64)
        for (int64_t i = 0; i < count; ++i)
            one_query(cfg, first + i, center, tmp, per, out + offsets[i], lens[i]);
        free(tmp);
    }
    free(center); free(lens);
    return tot;
}
