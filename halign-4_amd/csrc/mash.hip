// mash.hip -- Mash-style bottom-k sketches and the Jaccard choice of a
// reference (reference src/mash/mash.cpp:32-164, src/mash/hash.cpp; used by
// RefAligner::alignOneQueryToRef when the center file holds several
// sequences, src/align/ref_aligner.cpp:217-251).
//
//  * mash_kmers: one workgroup per sequence; every position's k-mer is read
//    whole (k <= 32 bytes), valid when no base of it is outside ACGTU, hashed
//    (XXH3-64 of the big-endian 8-byte code) and appended to the sequence's
//    key range (order does not matter: the range is sorted next);
//  * a rocPRIM segmented radix sort of the keys;
//  * mash_take: one wave per sequence keeps the first sketch_size distinct
//    keys, which is the reference's bottom-k heap with its seen-set;
//  * mash_select: one workgroup per query counts, for every reference, the
//    query hashes found in the reference sketch (binary search) and keeps the
//    first reference with the largest |a n b| / min(|a|, |b|), compared
//    exactly by cross-multiplication (1 for two empty sketches, 0 for one).
#include <hip/hip_runtime.h>

#include <cstdint>

namespace hal {

__device__ __forceinline__ uint64_t rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }

// XXH3-64 with `seed` of 8 input bytes that hold `code` big-endian: the
// 4..8-byte path of the vendored xxHash (default secret bytes 8..23)
__device__ __forceinline__ uint64_t hash2bit(uint64_t code, uint64_t seed_mix)
{
    const uint32_t in1 = __builtin_bswap32((uint32_t)(code >> 32));
    const uint32_t in2 = __builtin_bswap32((uint32_t)code);
    const uint64_t bitflip = (0x1cad21f72c81017cull ^ 0xdb979083e96dd4deull) - seed_mix;
    uint64_t h = ((uint64_t)in2 + ((uint64_t)in1 << 32)) ^ bitflip;
    const uint64_t mx2 = 0x9fb21c651e98df25ull;
    h ^= rotl64(h, 49) ^ rotl64(h, 24);
    h *= mx2;
    h ^= (h >> 35) + 8;
    h *= mx2;
    return h ^ (h >> 28);
}

// mash's nt4_table (include/mash.h): A C G T/U in either case -> 0..3, else 4
__device__ __forceinline__ int mash_code(unsigned char c)
{
    switch (c) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': case 'U': case 'u': return 3;
    default: return 4;
    }
}

__global__ __launch_bounds__(256) void mash_kmers(const char *__restrict__ seqs, const int64_t *__restrict__ off,
                                                  const int64_t *__restrict__ key_off, int k, int noncanonical,
                                                  uint64_t seed_mix, uint64_t *__restrict__ keys,
                                                  int64_t *__restrict__ nkeys)
{
    __shared__ unsigned int cnt;
    const int s = blockIdx.x;
    const char *q = seqs + off[s];
    const int64_t L = off[s + 1] - off[s];
    uint64_t *out = keys + key_off[s];
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    // (1 << 2k) - 1 with the shift count mod 64, as the reference's x86 build
    // computes it (k = 32: a zero mask, every k-mer codes as 0)
    const uint64_t mask = (1ull << ((2 * k) & 63)) - 1ull;
    for (int64_t i0 = 0; i0 + k <= L; i0 += 256) {
        const int64_t i = i0 + threadIdx.x;   // k-mer [i, i + k)
        bool valid = i + k <= L;
        uint64_t fwd = 0, rev = 0;
        if (valid) {
            for (int j = 0; j < k; ++j) {
                const int c = mash_code((unsigned char)q[i + j]);
                if (c > 3) {
                    valid = false;
                    break;
                }
                fwd = (fwd << 2) | (uint64_t)c;
                rev |= (uint64_t)(3 ^ c) << (2 * j);   // the reference's rolling form: base j of the window at bits 2j
            }
        }
        fwd &= mask;
        const uint64_t code = noncanonical ? fwd : (fwd < rev ? fwd : rev);
        // wave-aggregated append: one LDS atomic per wave
        const uint64_t bal = __ballot(valid);
        const int lane = threadIdx.x & 63;
        unsigned int base = 0;
        if (lane == 0 && bal) base = atomicAdd(&cnt, (unsigned int)__popcll(bal));
        base = __shfl(base, 0);
        if (valid) {
            const unsigned int r = (unsigned int)__popcll(bal & ((1ull << lane) - 1ull));
            out[base + r] = hash2bit(code, seed_mix);
        }
    }
    __syncthreads();
    // the rest of the range sorts last
    const int64_t cap = key_off[s + 1] - key_off[s];
    for (int64_t p = cnt + threadIdx.x; p < cap; p += 256) out[p] = ~0ull;
    if (threadIdx.x == 0) nkeys[s] = cnt;
}

// first `size` distinct keys of each sorted range; sketch s at out[s * size ..]
__global__ __launch_bounds__(64) void mash_take(const uint64_t *__restrict__ keys, const int64_t *__restrict__ key_off,
                                                const int64_t *__restrict__ nkeys, int size,
                                                uint64_t *__restrict__ out, int32_t *__restrict__ count)
{
    const int s = blockIdx.x;
    const int lane = threadIdx.x;
    const uint64_t *a = keys + key_off[s];
    const int64_t n = nkeys[s];
    uint64_t *o = out + (int64_t)s * size;
    int have = 0;
    for (int64_t i0 = 0; i0 < n && have < size; i0 += 64) {
        const int64_t i = i0 + lane;
        bool d = false;
        uint64_t v = 0;
        if (i < n) {
            v = a[i];
            d = i == 0 || a[i - 1] != v;
        }
        const uint64_t bal = __ballot(d);
        const int r = have + __popcll(bal & ((1ull << lane) - 1ull));
        if (d && r < size) o[r] = v;
        have += __popcll(bal);
    }
    if (lane == 0) count[s] = have < size ? have : size;
}

__global__ __launch_bounds__(256) void mash_select(const uint64_t *__restrict__ qsk, const int32_t *__restrict__ qcnt,
                                                   int qstride, const uint64_t *__restrict__ rsk,
                                                   const int32_t *__restrict__ rcnt, int rstride, int nr,
                                                   int32_t *__restrict__ best, int32_t *__restrict__ inter_out,
                                                   int32_t *__restrict__ den_out)
{
    __shared__ int part[4];
    const int q = blockIdx.x;
    const uint64_t *a = qsk + (int64_t)q * qstride;
    const int na = qcnt[q];
    int bi = -1, bn = -1, bd = 1;   // best as bn / bd, start below 0
    for (int r = 0; r < nr; ++r) {
        const uint64_t *b = rsk + (int64_t)r * rstride;
        const int nb = rcnt[r];
        int num, den;
        if (na == 0 && nb == 0) {
            num = 1;
            den = 1;
        } else if (na == 0 || nb == 0) {
            num = 0;
            den = 1;
        } else {
            int hits = 0;
            for (int i = threadIdx.x; i < na; i += 256) {
                const uint64_t v = a[i];
                int lo = 0, hi = nb;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (b[mid] < v) lo = mid + 1;
                    else hi = mid;
                }
                hits += lo < nb && b[lo] == v;
            }
            for (int o = 32; o > 0; o >>= 1) hits += __shfl_xor(hits, o);
            if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = hits;
            __syncthreads();
            num = part[0] + part[1] + part[2] + part[3];
            __syncthreads();
            den = na < nb ? na : nb;
        }
        if ((int64_t)num * bd > (int64_t)bn * den) {   // strictly larger, first one kept
            bi = r;
            bn = num;
            bd = den;
        }
    }
    if (threadIdx.x == 0) {
        best[q] = bi;
        inter_out[q] = bn;
        den_out[q] = bd;
    }
}

hipError_t mash_launch(int phase, int n, const char *seqs, const int64_t *off, const int64_t *key_off, int k,
                       int noncanonical, uint64_t seed_mix, uint64_t *keys, int64_t *nkeys, int size, uint64_t *out,
                       int32_t *count, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    if (phase == 0)
        hipLaunchKernelGGL(mash_kmers, dim3((unsigned)n), dim3(256), 0, stream, seqs, off, key_off, k, noncanonical,
                           seed_mix, keys, nkeys);
    else
        hipLaunchKernelGGL(mash_take, dim3((unsigned)n), dim3(64), 0, stream, keys, key_off, nkeys, size, out, count);
    return hipGetLastError();
}

hipError_t mash_select_launch(int nq, const uint64_t *qsk, const int32_t *qcnt, int qstride, const uint64_t *rsk,
                              const int32_t *rcnt, int rstride, int nr, int32_t *best, int32_t *inter, int32_t *den,
                              hipStream_t stream)
{
    if (nq <= 0) return hipSuccess;
    hipLaunchKernelGGL(mash_select, dim3((unsigned)nq), dim3(256), 0, stream, qsk, qcnt, qstride, rsk, rcnt, rstride,
                       nr, best, inter, den);
    return hipGetLastError();
}

}  // namespace hal
