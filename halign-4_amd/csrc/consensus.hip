// consensus.hip -- per-column base counts and the consensus character over the
// rows of an alignment (consensus::generateConsensusSequence,
// src/consensus/consensus.cpp:611-703; SiteCount include/consensus.h:78-99;
// k_base_map include/consensus.h:126-145; pickConsensusChar
// src/consensus/consensus.cpp:18-35).
//
// Layout: rows are stored row-major with a row stride that is a multiple of 4
// bytes. A lane owns 4 adjacent columns and walks a chunk of rows with one
// dword load per row (a wave reads 256 contiguous bytes of a row), eight rows'
// loads in flight at a time. The class
// of each byte comes from a 256-entry table in LDS; the 7 class counts of a
// column are 16-bit fields of two 64-bit accumulators (classes 0-3 and 4-6),
// so a chunk holds at most 65535 rows, and are added to the global 32-bit
// counts with one atomic per (column, class) per chunk.
//
// Row selection is the reference's reader model: rows are read in batches of
// `batch` records and one record after each full batch is read and dropped
// (row r is skipped iff r % (batch + 1) == batch); counting stops after
// seq_limit counted rows (0 = no limit). The host turns seq_limit into the
// row bound row_end.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace hal {

constexpr int kCountThreads = 256;
constexpr int kMaxChunkRows = 65535;

// k_base_map: A C G T U (either case) -> 0..4, '-' and '.' -> 6, anything else -> 5 (N)
__device__ __forceinline__ uint8_t base_class(uint32_t c)
{
    switch (c) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    case 'U': case 'u': return 4;
    case '-': case '.': return 6;
    default: return 5;
    }
}

__global__ __launch_bounds__(kCountThreads) void count_columns(const uint8_t *__restrict__ rows, int64_t row_stride,
                                                              int64_t aln_len, int64_t row_begin_all,
                                                              int64_t row_end, int64_t rows_per_chunk, int64_t batch,
                                                              uint32_t *__restrict__ counts)
{
    __shared__ uint8_t cls[256];
    cls[threadIdx.x] = base_class(threadIdx.x);
    __syncthreads();

    const int64_t col0 = 4 * ((int64_t)blockIdx.x * kCountThreads + threadIdx.x);
    const int64_t r0 = row_begin_all + (int64_t)blockIdx.y * rows_per_chunk;
    int64_t r1 = r0 + rows_per_chunk;
    if (r1 > row_end) r1 = row_end;
    if (col0 >= aln_len || r0 >= r1) return;

    uint64_t lo[4] = {0, 0, 0, 0}, hi[4] = {0, 0, 0, 0};
    const uint8_t *p = rows + r0 * row_stride + col0;
    int64_t phase = r0 % (batch + 1);   // position inside the read-batch cycle (the same in every lane)
    // 8 rows per step: their loads are independent and in flight together
    constexpr int kU = 8;
    for (int64_t r = r0; r < r1; r += kU, p += kU * row_stride) {
        uint32_t v[kU];
        uint32_t use = 0;
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            // the record read after a full batch is dropped
            const bool cnt = r + u < r1 && phase != batch;
            phase = phase == batch ? 0 : phase + 1;
            v[u] = cnt ? *(const uint32_t *)(p + u * row_stride) : 0u;
            use |= (uint32_t)cnt << u;
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            if (!((use >> u) & 1u)) continue;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const uint32_t k = cls[(v[u] >> (8 * b)) & 0xffu];
                lo[b] += k < 4 ? (uint64_t)1 << (16 * k) : 0;
                hi[b] += k >= 4 ? (uint64_t)1 << (16 * (k - 4)) : 0;
            }
        }
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        if (col0 + b >= aln_len) break;
        uint32_t *c = counts + (col0 + b) * 7;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t n = (uint32_t)(lo[b] >> (16 * k)) & 0xffffu;
            if (n) atomicAdd(c + k, n);
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const uint32_t n = (uint32_t)(hi[b] >> (16 * k)) & 0xffffu;
            if (n) atomicAdd(c + 4 + k, n);
        }
    }
}

// pickConsensusChar: the largest of A, C, G, T, U, earlier wins ties; all zero -> 'A'
__global__ void pick_consensus(const uint32_t *__restrict__ counts, int64_t aln_len, char *__restrict__ cons)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= aln_len) return;
    const uint32_t *s = counts + i * 7;
    uint32_t best = s[0];
    char ch = 'A';
    if (s[1] > best) { best = s[1]; ch = 'C'; }
    if (s[2] > best) { best = s[2]; ch = 'G'; }
    if (s[3] > best) { best = s[3]; ch = 'T'; }
    if (s[4] > best) { best = s[4]; ch = 'U'; }
    cons[i] = ch;
}

__global__ void zero_words(uint32_t *__restrict__ p, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

// Clears n 32-bit words with a fill kernel (no memset node in a captured graph).
hipError_t zero_words_launch(uint32_t *p, int64_t n, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    int64_t g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(zero_words, dim3((unsigned)g), dim3(256), 0, stream, p, n);
    return hipGetLastError();
}

// Launch both kernels on `stream`. counts must be zeroed by the caller.
hipError_t consensus_launch(const uint8_t *rows, int64_t row_stride, int64_t aln_len, int64_t row_end, int64_t batch,
                            int num_cu, uint32_t *counts, char *cons, hipStream_t stream)
{
    const int64_t ncol4 = (aln_len + 3) / 4;
    const int64_t gx = (ncol4 + kCountThreads - 1) / kCountThreads;
    // enough row chunks for ~8 workgroups per CU, each at most kMaxChunkRows rows
    int64_t want = (8LL * num_cu + gx - 1) / gx;
    if (want < 1) want = 1;
    int64_t rpc = (row_end + want - 1) / want;
    if (rpc < 64) rpc = 64;
    if (rpc > kMaxChunkRows) rpc = kMaxChunkRows;
    const int64_t gy = (row_end + rpc - 1) / rpc;
    if (gy > 65535) return hipErrorInvalidValue;   // more than 4.29e9 rows
    if (row_end > 0) {
        hipLaunchKernelGGL(count_columns, dim3((unsigned)gx, (unsigned)gy), dim3(kCountThreads), 0, stream, rows,
                           row_stride, aln_len, (int64_t)0, row_end, rpc, batch, counts);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(pick_consensus, dim3((unsigned)((aln_len + 255) / 256)), dim3(256), 0, stream, counts, aln_len,
                       cons);
    return hipGetLastError();
}

}  // namespace hal
