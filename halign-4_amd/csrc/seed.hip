// seed.hip -- minimizer extraction (minimizer::extractMinimizer,
// src/seed/minimizer.cpp:103-177; RingMinQueue :35-97; splitmix64 :16-22;
// nt4_table include/seed.h:157-190; MinimizerHit include/seed.h:206-265).
//
// Two device forms with identical results:
//
//  * minimizers_wave: one wave per sequence, 64 k-mer starts per step. A
//    sequence is read as 2-bit codes packed big-endian into 64-bit words plus
//    a bitmask of non-ACGTU bases (pack_seqs), so a k-mer code is two shifts.
//    The reference's ring queue keeps, at each k-mer, the rightmost minimum of
//    the last `win` hashes of the current N-free run; here that is a range
//    minimum over two power-of-two blocks of a sparse table kept in LDS
//    (levels 1..32, 128-position rings that reach into the previous step).
//    Emission is "valid, past the first window, and (first in the run or a
//    different (hash, pos) than the step before)", compacted with ballot/mbcnt.
//
//  * minimizers_seq: one lane per sequence, the reference's queue step by step
//    (a ring of `win` slots indexed as the reference indexes it, in a larger
//    per-lane LDS workspace so that the reference's out-of-range slots are
//    modelled instead of faulting); minimizers_seqw, the same steps with one
//    wave per sequence for windows of up to 14 (hashes of 64 positions at a
//    time by the lanes, the workspace one entry per lane).
//
// The queue differs from a plain windowed minimum in one case: when it holds
// `win` entries (the last win hashes strictly increasing) and a larger hash
// arrives, the push lands on the front slot. minimizers_wave detects a run of
// win + 1 strictly increasing hashes, stops, and marks the sequence for
// minimizers_seq. Canonical k-mers and windows above 32 always take
// minimizers_seq.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

namespace hal {

constexpr int kSeedWaves = 4;        // waves per workgroup in minimizers_wave
constexpr int kMaxWaveWin = 32;      // largest window of the wave form
constexpr int kRing = 128;           // LDS ring of positions per sparse-table level
constexpr int kLevels = 6;           // block lengths 1, 2, 4, 8, 16, 32

// status bits per sequence
constexpr int kSeedNeedSeq = 1;      // wave form stopped: run minimizers_seq
constexpr int kSeedUb = 2;           // reference indexes past its ring buffer here
constexpr int kSeedOverflow = 4;     // even the larger workspace overflowed (output truncated)

__device__ __forceinline__ uint32_t nt4(uint32_t c)
{
    switch (c) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': case 'U': case 'u': return 3;
    default: return 4;
    }
}

__device__ __forceinline__ uint64_t splitmix64(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

// One thread per 32 bases: word w of a sequence holds bases 32w..32w+31, base
// 32w+j in bits 2(31-j)+1..2(31-j) (big-endian, so a window read off the top of
// the bits is the reference's rolling code with the oldest base highest); mask
// word m has bit j set when base 64m+j is not A/C/G/T/U. A sequence of len
// bases owns 2*(len/64+2) code words and len/64+2 mask words, the tail zero.
__global__ void pack_seqs(const char *__restrict__ seqs, const int64_t *__restrict__ off,
                          const int64_t *__restrict__ woff, const int64_t *__restrict__ moff, int n,
                          uint64_t *__restrict__ codes, uint64_t *__restrict__ nmask)
{
    const int s = blockIdx.y * gridDim.z + blockIdx.z;
    if (s >= n) return;
    const int64_t len = off[s + 1] - off[s];
    const char *q = seqs + off[s];
    const int64_t nw = 2 * (len / 64 + 2);   // 32-base groups incl. padding (an even count)
    for (int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; w < nw; w += (int64_t)gridDim.x * blockDim.x) {
        uint64_t cw = 0;
        uint32_t nb = 0;
        for (int j = 0; j < 32; ++j) {
            const int64_t b = 32 * w + j;
            uint32_t c = 0, isn = 0;
            if (b < len) {
                c = nt4((uint8_t)q[b]);
                isn = c >> 2;
                c &= 3;
            }
            cw |= (uint64_t)c << (2 * (31 - j));
            nb |= isn << j;
        }
        codes[woff[s] + w] = cw;
        // two 32-base words make one 64-bit mask word
        uint32_t *m32 = (uint32_t *)(nmask + moff[s]);
        m32[w] = nb;
    }
}

struct Node {
    uint64_t h;
    uint32_t p;
};

// later block first: on equal hashes the later (rightmost) position wins
__device__ __forceinline__ Node pick(Node later, Node earlier) { return later.h <= earlier.h ? later : earlier; }

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ __launch_bounds__(64 * kSeedWaves) void minimizers_wave(
    const int64_t *__restrict__ off, const int64_t *__restrict__ woff, const int64_t *__restrict__ moff,
    const uint64_t *__restrict__ codes, const uint64_t *__restrict__ nmask, const int64_t *__restrict__ hit_off,
    int n, int k, int w, uint64_t *__restrict__ hx, uint64_t *__restrict__ hy, int64_t *__restrict__ counts,
    int32_t *__restrict__ status)
{
    __shared__ uint64_t sh_h[kSeedWaves][kLevels][kRing];
    __shared__ uint32_t sh_p[kSeedWaves][kLevels][kRing];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int s = blockIdx.x * kSeedWaves + wv;
    if (s >= n) return;   // whole wave
    uint64_t(*H)[kRing] = sh_h[wv];
    uint32_t(*P)[kRing] = sh_p[wv];

    const int64_t len = off[s + 1] - off[s];
    int64_t nout = 0;
    if (k <= 0 || w <= 0 || len < k || k > 31 || w >= 256) {
        if (lane == 0) { counts[s] = 0; status[s] = 0; }
        return;
    }
    const int64_t nk = len - k + 1;                  // k-mer starts
    const int win = (int)(w < nk ? w : nk);
    const uint64_t *cw = codes + woff[s];
    const uint64_t *nm = nmask + moff[s];
    uint64_t *ox = hx + hit_off[s], *oy = hy + hit_off[s];
    const uint64_t kmask = ((uint64_t)1 << (2 * k)) - 1;
    int levels = 0;                                  // sparse-table levels needed: 2^levels <= win
    while ((2 << levels) <= win) ++levels;

    // carries from the previous step (values of position base-1)
    bool c_valid = false, c_emitting = false;
    Node c_front{0, 0};
    uint64_t c_h = 0;
    int64_t c_seg = 0;      // start of the N-free run containing base-1
    int64_t c_run = 0;      // strictly increasing steps ending at base-1
    bool quirk = false;

    for (int64_t base = 0; base < nk; base += 64) {
        const int64_t p = base + lane;
        const bool in = p < nk;
        bool valid = false;
        uint64_t h = ~(uint64_t)0;
        if (in) {
            const int64_t mw = p >> 6;
            const int mb = (int)(p & 63);
            uint64_t m = nm[mw] >> mb;
            if (mb) m |= nm[mw + 1] << (64 - mb);
            valid = (m & ((((uint64_t)1) << k) - 1)) == 0;
            const int64_t ww = p >> 5;
            const int ob = (int)(p & 31) * 2;
            uint64_t v = cw[ww] << ob;
            if (ob) v |= cw[ww + 1] >> (64 - ob);
            const uint64_t code = (v >> (64 - 2 * k)) & kmask;
            if (valid) h = splitmix64(code) >> 8;
        }
        // previous position's values (lane - 1, or the carry for lane 0)
        const uint64_t vmask = __ballot(valid);
        const bool prev_valid = lane ? ((vmask >> (lane - 1)) & 1) : c_valid;
        const uint64_t h_prev_l = __shfl_up(h, 1);
        const uint64_t h_prev = lane ? h_prev_l : c_h;

        // start of the N-free run containing p
        const uint64_t starts = __ballot(valid && !prev_valid);
        const uint64_t upto = lane == 63 ? ~(uint64_t)0 : (((uint64_t)2 << lane) - 1);
        const uint64_t sm = starts & upto;
        const int64_t seg = sm ? base + (63 - __builtin_clzll(sm)) : c_seg;

        // queue front-overwrite case: win + 1 strictly increasing hashes in one run
        const bool inc = valid && prev_valid && h > h_prev;
        const uint64_t breaks = ~__ballot(inc) & upto;
        const int64_t run = breaks ? (int64_t)(lane - (63 - __builtin_clzll(breaks))) : c_run + lane + 1;
        if (__ballot(valid && run >= win)) {
            quirk = true;
            break;
        }

        // sparse table: level j at position q = rightmost minimum of [q-2^j+1, q]
        const int slot = (int)(p & (kRing - 1));
        H[0][slot] = h;
        P[0][slot] = (uint32_t)p;
        for (int j = 1; j <= levels; ++j) {
            wave_sync();
            const int o = (int)((p - (1 << (j - 1))) & (kRing - 1));
            const Node a{H[j - 1][slot], P[j - 1][slot]};
            const Node b{H[j - 1][o], P[j - 1][o]};
            const Node r = pick(a, b);
            wave_sync();
            H[j][slot] = r.h;
            P[j][slot] = r.p;
        }
        wave_sync();
        // front = rightmost minimum over [max(p-win+1, seg), p]
        Node front{0, 0};
        if (valid) {
            int64_t a = p - win + 1;
            if (a < seg) a = seg;
            const int L = (int)(p - a + 1);
            const int J = 31 - __builtin_clz((unsigned)L);
            const int s2 = (int)((a + (1 << J) - 1) & (kRing - 1));
            front = pick(Node{H[J][slot], P[J][slot]}, Node{H[J][s2], P[J][s2]});
        }
        wave_sync();

        const bool emitting = valid && p + 1 >= win;
        // every shuffle runs with all lanes active: ds_bpermute reads 0 from a lane
        // that is masked off, so a shuffle inside `lane ? ... : ...` would lose lane 0
        const int em_l = __shfl_up((int)emitting, 1);
        const bool prev_emitting = lane ? em_l != 0 : c_emitting;
        const uint64_t fh_l = __shfl_up(front.h, 1);
        const uint32_t fp_l = __shfl_up(front.p, 1);
        const uint64_t fh_prev = lane ? fh_l : c_front.h;
        const uint32_t fp_prev = lane ? fp_l : c_front.p;
        const bool emit = emitting && (!prev_emitting || front.h != fh_prev || front.p != fp_prev);
        const uint64_t em = __ballot(emit);
        if (emit) {
            const int64_t at = nout + __builtin_popcountll(em & (((uint64_t)1 << lane) - 1));
            ox[at] = (front.h << 8) | (uint64_t)(uint8_t)k;
            oy[at] = ((uint64_t)0x80000000u << 32) | (uint64_t)front.p;   // non-canonical: strand set
        }
        nout += __builtin_popcountll(em);

        // carries for the next step: position base+63
        c_valid = __shfl(valid ? 1 : 0, 63) != 0;
        c_emitting = __shfl(emitting ? 1 : 0, 63) != 0;
        c_h = __shfl(h, 63);
        c_front.h = __shfl(front.h, 63);
        c_front.p = __shfl(front.p, 63);
        c_seg = __shfl(seg, 63);
        c_run = __shfl(run, 63);
    }
    if (lane == 0) {
        counts[s] = quirk ? 0 : nout;
        status[s] = quirk ? kSeedNeedSeq : 0;
    }
}

// One lane per listed sequence: the reference's loop (minimizer.cpp:118-170)
// with its RingMinQueue. The queue lives in LDS, ws_cap (4*win+8) entries per
// lane (the dynamic shared memory holds blockDim.x of them): every queue
// access is on the sequence's dependent chain, so its latency is the kernel's.
__global__ void minimizers_seq(const char *__restrict__ seqs, const int64_t *__restrict__ off,
                               const int32_t *__restrict__ list, int nlist, int k_, int w_, int non_canonical,
                               const int64_t *__restrict__ hit_off, int ws_cap, uint64_t *__restrict__ hx,
                               uint64_t *__restrict__ hy, int64_t *__restrict__ counts, int32_t *__restrict__ status)
{
    extern __shared__ uint64_t seq_lds[];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nlist) return;
    const int s = list[t];
    const char *seq = seqs + off[s];
    const uint32_t n = (uint32_t)(off[s + 1] - off[s]), k = (uint32_t)k_, w = (uint32_t)w_;
    int st = 0;
    int64_t cnt = 0;
    if (k == 0 || w == 0 || n < k || k > 31 || w >= 256) {
        counts[s] = 0;
        status[s] = 0;
        return;
    }
    const uint32_t total = n - k + 1;
    const uint32_t win = w < total ? w : total;
    const uint64_t mask = ((uint64_t)1 << (2 * k)) - 1;
    const uint64_t shift = 2ull * (k - 1);
    uint64_t *qh = seq_lds + (int64_t)threadIdx.x * ws_cap;
    uint32_t *qp = (uint32_t *)(seq_lds + (int64_t)blockDim.x * ws_cap) + (int64_t)threadIdx.x * ws_cap;
    const uint32_t big = 4 * win + 8 <= (uint32_t)ws_cap ? 4 * win + 8 : (uint32_t)ws_cap;
    uint64_t *ox = hx + hit_off[s], *oy = hy + hit_off[s];
    uint64_t fwd = 0, rev = 0;
    uint32_t valid = 0, head = 0, size = 0, last_p = 0;
    uint64_t last_h = 0;
    bool has_last = false, over = false;
    // the reference's index: one wrap by `win` (minimizer.cpp:38-41)
    auto idx = [&](uint32_t o) {
        uint32_t i = head + o;
        if (i >= win) i -= win;
        return i;
    };
    auto check = [&](uint32_t i) {
        if (i >= win) st |= kSeedUb;
        if (i >= big) over = true;
    };
    for (uint32_t i = 0; i < n && !over; ++i) {
        const uint32_t c = nt4((uint8_t)seq[i]);
        if (c >= 4) {
            fwd = rev = 0;
            valid = 0;
            head = size = 0;
            has_last = false;
            continue;
        }
        fwd = ((fwd << 2) | c) & mask;
        rev = (rev >> 2) | ((uint64_t)(3u ^ c) << shift);
        if (valid < k) ++valid;
        if (valid < k) continue;
        const uint32_t pos = i + 1 - k;
        const uint64_t code = non_canonical ? fwd : (fwd < rev ? fwd : rev);
        const uint64_t h56 = splitmix64(code) >> 8;
        while (size) {   // pop_back while back.h >= h56
            const uint32_t bi = idx(size - 1);
            check(bi);
            if (over) break;
            if (qh[bi] >= h56) --size;
            else break;
        }
        if (over) break;
        {
            const uint32_t wi = idx(size);
            check(wi);
            if (over) break;
            qh[wi] = h56;
            qp[wi] = pos;
            ++size;
        }
        if (pos + 1 < win) continue;
        const uint32_t win_start = pos + 1 - win;
        while (size) {   // popExpired
            check(head);
            if (over) break;
            if (qp[head] < win_start) {
                head = idx(1);
                --size;
            } else {
                break;
            }
        }
        if (over) break;
        if (size) {
            check(head);
            const uint64_t ch = qh[head];
            const uint32_t cp = qp[head];
            if (!has_last || ch != last_h || cp != last_p) {
                const uint32_t strand = non_canonical ? 0x80000000u : (fwd <= rev ? 0x80000000u : 0u);
                ox[cnt] = (ch << 8) | (uint64_t)(uint8_t)k;
                oy[cnt] = ((uint64_t)strand << 32) | (uint64_t)cp;
                ++cnt;
                last_h = ch;
                last_p = cp;
                has_last = true;
            }
        }
    }
    if (over) st |= kSeedOverflow;
    counts[s] = cnt;
    status[s] = st;
}

// The same loop with one wave per listed sequence, for queue workspaces of at
// most 64 entries (w <= 14): for 64 positions at a time the lanes classify
// the base, build the k-mer codes from the bases staged in LDS and hash them;
// then the whole wave steps the reference's queue through those positions, the
// workspace held one entry per lane (entry e in lane e, read and written with
// readlane / writelane at wave-uniform indices), so the dependent chain is only
// the queue's own steps.
__global__ __launch_bounds__(64) void minimizers_seqw(const char *__restrict__ seqs, const int64_t *__restrict__ off,
                                                      const int32_t *__restrict__ list, int nlist, int k_, int w_,
                                                      int non_canonical, const int64_t *__restrict__ hit_off,
                                                      int ws_cap, uint64_t *__restrict__ hx, uint64_t *__restrict__ hy,
                                                      int64_t *__restrict__ counts, int32_t *__restrict__ status)
{
    __shared__ uint8_t code[128];   // base classes of positions b-64 .. b+63
    const int t = (int)blockIdx.x;
    if (t >= nlist) return;
    const int lane = (int)threadIdx.x;
    const int s = list[t];
    const char *seq = seqs + off[s];
    const uint32_t n = (uint32_t)(off[s + 1] - off[s]), k = (uint32_t)k_, w = (uint32_t)w_;
    if (k == 0 || w == 0 || n < k || k > 31 || w >= 256) {
        if (lane == 0) {
            counts[s] = 0;
            status[s] = 0;
        }
        return;
    }
    const uint32_t total = n - k + 1;
    const uint32_t win = w < total ? w : total;
    const uint64_t mask = ((uint64_t)1 << (2 * k)) - 1;
    const uint32_t shift = 2 * (k - 1);
    const uint32_t big = 4 * win + 8 <= (uint32_t)ws_cap ? 4 * win + 8 : (uint32_t)ws_cap;   // <= 64
    uint64_t *ox = hx + hit_off[s], *oy = hy + hit_off[s];
    uint32_t qlo = 0, qhi = 0, qpos = 0;   // queue workspace entry `lane`: hash, position
    uint32_t head = 0, size = 0, last_p = 0;
    uint64_t last_h = 0;
    bool has_last = false, over = false;
    int st = 0;
    int64_t cnt = 0;
    auto idx = [&](uint32_t o) {   // the reference's index: one wrap by `win`
        uint32_t i = head + o;
        if (i >= win) i -= win;
        return i;
    };
    auto check = [&](uint32_t i) {
        if (i >= win) st |= kSeedUb;
        if (i >= big) over = true;
    };
    auto rl64 = [](uint32_t lo, uint32_t hi, uint32_t l) {
        return (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)hi, (int)l) << 32 |
               (uint32_t)__builtin_amdgcn_readlane((int)lo, (int)l);
    };
    code[64 + lane] = 4;
    for (uint32_t b = 0; b < n && !over; b += 64) {
        // classes of positions b-64 .. b+63 (4: not ACGTU, or before the sequence)
        __syncthreads();
        code[lane] = code[64 + lane];
        const uint32_t i = b + (uint32_t)lane;
        code[64 + lane] = (uint8_t)(i < n ? nt4((uint8_t)seq[i]) : 4u);
        __syncthreads();
        // 0: a base outside ACGTU (the queue restarts), 1: no k-mer ends here,
        // 2: the k-mer of bases i-k+1 .. i, with its hash and strand
        uint32_t cls = code[64 + lane] >= 4 ? 0u : 1u, strand = 0, hlo = 0, hhi = 0;
        if (cls && i + 1 >= k) {
            uint64_t f = 0, r = 0;
            bool ok = true;
#pragma unroll
            for (int d = 0; d < 31; ++d) {
                if (d < (int)k) {
                    const uint32_t c = code[64 + lane - (int)k + 1 + d];
                    ok = ok && c < 4;
                    f = ((f << 2) | (c & 3)) & mask;
                    r = (r >> 2) | ((uint64_t)(3u ^ (c & 3)) << shift);
                }
            }
            if (ok) {
                cls = 2;
                const uint64_t h56 = splitmix64(non_canonical ? f : (f < r ? f : r)) >> 8;
                hlo = (uint32_t)h56;
                hhi = (uint32_t)(h56 >> 32);
                strand = (non_canonical || f <= r) ? 1u : 0u;
            }
        }
        const uint32_t jn = n - b < 64 ? n - b : 64;
        for (uint32_t j = 0; j < jn; ++j) {
            const uint32_t cj = (uint32_t)__builtin_amdgcn_readlane((int)cls, (int)j);
            if (cj == 0) {
                head = size = 0;
                has_last = false;
                continue;
            }
            if (cj == 1) continue;
            const uint64_t h = rl64(hlo, hhi, j);
            const uint32_t pos = b + j + 1 - k;
            while (size) {   // pop_back while back.h >= h
                const uint32_t bi = idx(size - 1);
                check(bi);
                if (over) break;
                if (rl64(qlo, qhi, bi) >= h) --size;
                else break;
            }
            if (over) break;
            {
                const uint32_t wi = idx(size);
                check(wi);
                if (over) break;
                qlo = lane == (int)wi ? (uint32_t)h : qlo;
                qhi = lane == (int)wi ? (uint32_t)(h >> 32) : qhi;
                qpos = lane == (int)wi ? pos : qpos;
                ++size;
            }
            if (pos + 1 < win) continue;
            const uint32_t win_start = pos + 1 - win;
            while (size) {   // popExpired
                check(head);
                if (over) break;
                if ((uint32_t)__builtin_amdgcn_readlane((int)qpos, (int)head) < win_start) {
                    head = idx(1);
                    --size;
                } else {
                    break;
                }
            }
            if (over) break;
            if (size) {
                check(head);
                if (over) break;
                const uint64_t ch = rl64(qlo, qhi, head);
                const uint32_t cp = (uint32_t)__builtin_amdgcn_readlane((int)qpos, (int)head);
                if (!has_last || ch != last_h || cp != last_p) {
                    const uint32_t sj = (uint32_t)__builtin_amdgcn_readlane((int)strand, (int)j);
                    if (lane == 0) {
                        ox[cnt] = (ch << 8) | (uint64_t)(uint8_t)k;
                        oy[cnt] = ((uint64_t)(sj ? 0x80000000u : 0u) << 32) | (uint64_t)cp;
                    }
                    ++cnt;
                    last_h = ch;
                    last_p = cp;
                    has_last = true;
                }
            }
        }
    }
    if (lane == 0) {
        counts[s] = cnt;
        status[s] = st | (over ? kSeedOverflow : 0);
    }
}

hipError_t seed_pack(const char *seqs, const int64_t *off, const int64_t *woff, const int64_t *moff, int n,
                     int64_t max_words, uint64_t *codes, uint64_t *nmask, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    const int gx = (int)((max_words + 255) / 256 < 64 ? (max_words + 255) / 256 : 64);
    const int gz = n < 1024 ? n : 1024;
    const int gy = (n + gz - 1) / gz;
    hipLaunchKernelGGL(pack_seqs, dim3(gx > 0 ? gx : 1, gy, gz), dim3(256), 0, stream, seqs, off, woff, moff, n,
                       codes, nmask);
    return hipGetLastError();
}

hipError_t seed_wave(const int64_t *off, const int64_t *woff, const int64_t *moff, const uint64_t *codes,
                     const uint64_t *nmask, const int64_t *hit_off, int n, int k, int w, uint64_t *hx, uint64_t *hy,
                     int64_t *counts, int32_t *status, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(minimizers_wave, dim3((n + kSeedWaves - 1) / kSeedWaves), dim3(64 * kSeedWaves), 0, stream,
                       off, woff, moff, codes, nmask, hit_off, n, k, w, hx, hy, counts, status);
    return hipGetLastError();
}

hipError_t seed_seq(const char *seqs, const int64_t *off, const int32_t *list, int nlist, int k, int w,
                    int non_canonical, const int64_t *hit_off, int ws_cap, uint64_t *hx, uint64_t *hy,
                    int64_t *counts, int32_t *status, hipStream_t stream)
{
    if (nlist <= 0) return hipSuccess;
    if (ws_cap <= 64) {   // the workspace fits one entry per lane: one wave per sequence
        hipLaunchKernelGGL(minimizers_seqw, dim3(nlist), dim3(64), 0, stream, seqs, off, list, nlist, k, w,
                           non_canonical, hit_off, ws_cap, hx, hy, counts, status);
        return hipGetLastError();
    }
    // 12 bytes of queue per entry and lane; lanes per workgroup so that a
    // workgroup takes at most 48 KiB of LDS (64 lanes for the path's w = 10)
    const size_t per_lane = (size_t)ws_cap * 12;
    int lanes = (int)std::min<size_t>(64, std::max<size_t>(1, (48u << 10) / per_lane));
    const size_t lds = (size_t)lanes * per_lane;
    hipLaunchKernelGGL(minimizers_seq, dim3((nlist + lanes - 1) / lanes), dim3(lanes), lds, stream, seqs, off, list,
                       nlist, k, w, non_canonical, hit_off, ws_cap, hx, hy, counts, status);
    return hipGetLastError();
}

int seed_max_wave_window() { return kMaxWaveWin; }

}  // namespace hal
