// consensus.hip -- per-column base counts and the consensus character over the
// rows of an alignment (consensus::generateConsensusSequence,
// src/consensus/consensus.cpp:611-703; SiteCount include/consensus.h:78-99;
// k_base_map include/consensus.h:126-145; pickConsensusChar
// src/consensus/consensus.cpp:18-35).
//
// Layout: rows are stored row-major with a row stride that is a multiple of 4
// bytes (count_columns16 below). Row selection is the reference's reader
// model: rows are read in batches of `batch` records and one record after each
// full batch is read and dropped (row r is skipped iff r % (batch + 1) ==
// batch); counting stops after seq_limit counted rows (0 = no limit). The
// host turns seq_limit into the row bound row_end.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace hal {

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

// count_columns16: a workgroup of 4 waves owns 1,024 adjacent columns of a
// range of rows; lane l of every wave reads bytes [16l, 16l + 16) of the
// workgroup's columns of a row with one 16-byte load, and wave w takes the
// w-th quarter of the range. A byte's class adds 1 << 4k (k = class) to its
// column's accumulator, read from a 256-entry table in LDS, so one dword
// holds all 7 counts in 4-bit fields (8 rows' loads are in flight at a time); every 15 counted rows the fields move to
// 16-bit halves of four dwords (classes 0|4, 1|5, 2|6, 3), which hold up to
// 65,535 rows. The four waves add their counts in LDS and the workgroup adds
// them to the global counts with one atomic per (column, class). The rows the
// reference's reader drops (r % (batch + 1) == batch) are skipped by the whole
// wave.
constexpr int kCols16 = 1024;             // columns per workgroup (64 lanes x 16 bytes)
constexpr int kWaves16 = 4;
constexpr int kMaxWaveRows16 = 65535;     // rows per wave (16-bit counters)

template <bool A16>
__global__ __launch_bounds__(64 * kWaves16) void count_columns16(const uint8_t *__restrict__ rows, int64_t row_stride,
                                                                 int64_t aln_len, int64_t row_end,
                                                                 int64_t rows_per_group, int64_t batch,
                                                                 uint32_t *__restrict__ counts)
{
    __shared__ uint32_t inc[256];
    __shared__ uint32_t part[7][kCols16];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    inc[tid] = 1u << (4 * base_class((uint32_t)tid));
    for (int i = tid; i < 7 * kCols16; i += 64 * kWaves16) (&part[0][0])[i] = 0u;
    __syncthreads();

    const int64_t col0 = (int64_t)blockIdx.x * kCols16 + 16 * lane;
    const int64_t g0 = (int64_t)blockIdx.y * rows_per_group;
    const int64_t per_wave = (rows_per_group + kWaves16 - 1) / kWaves16;
    int64_t r0 = g0 + wave * per_wave, r1 = r0 + per_wave;
    if (r1 > g0 + rows_per_group) r1 = g0 + rows_per_group;
    if (r1 > row_end) r1 = row_end;
    // this lane's bytes of a row: dword d is read from the row at col0 + 4d
    // when it lies inside the row, else from the row's start (those columns are
    // past aln_len and never reported), so every load is unconditional; with
    // a stride that is a multiple of 16 the 16 bytes are all inside or all out
    const int64_t lane_off = col0 + 16 <= row_stride ? col0 : 0;
    int32_t dw_off[4];
#pragma unroll
    for (int d = 0; d < 4; ++d) dw_off[d] = (int32_t)(col0 + 4 * d + 4 <= row_stride ? col0 + 4 * d : 0);
    uint32_t acc[16], c04[16], c15[16], c26[16], c3[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) acc[j] = c04[j] = c15[j] = c26[j] = c3[j] = 0u;
    auto flush = [&]() {
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            c04[j] += acc[j] & 0x000F000Fu;
            c15[j] += (acc[j] >> 4) & 0x000F000Fu;
            c26[j] += (acc[j] >> 8) & 0x000F000Fu;
            c3[j] += (acc[j] >> 12) & 0xFu;
            acc[j] = 0u;
        }
    };
    int64_t phase = r0 % (batch + 1);   // position in the read-batch cycle (uniform)
    int pending = 0;                    // counted rows in acc
    auto load_row = [&](const uint8_t *row, uint32_t (&v)[4]) {
        if (A16) {
            const uint4 x = *(const uint4 *)(row + lane_off);
            v[0] = x.x; v[1] = x.y; v[2] = x.z; v[3] = x.w;
        } else {
#pragma unroll
            for (int d = 0; d < 4; ++d) v[d] = *(const uint32_t *)(row + dw_off[d]);
        }
    };
    // 8 rows per step, loaded unconditionally (rows past the range repeat the
    // range's last row and are not counted), so their loads are in flight together
    constexpr int kU = 8;
    for (int64_t r = r0; r < r1; r += kU) {
        uint32_t v[kU][4];
        uint32_t use = 0;   // counted rows of this step (uniform)
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const bool cnt = r + u < r1 && phase != batch;
            phase = phase == batch ? 0 : phase + 1;
            use |= (uint32_t)cnt << u;
            const int64_t rr = r + u < r1 ? r + u : r1 - 1;
            load_row(rows + rr * row_stride, v[u]);
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            if (!((use >> u) & 1u)) continue;
#pragma unroll
            for (int j = 0; j < 16; ++j) acc[j] += inc[(v[u][j >> 2] >> (8 * (j & 3))) & 0xffu];
            if (++pending == 15) {
                flush();
                pending = 0;
            }
        }
    }
    flush();
    // the workgroup's sum in LDS, then one global atomic per (column, class)
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const int c = 16 * lane + j;
        const uint32_t k[7] = {c04[j] & 0xffffu, c15[j] & 0xffffu, c26[j] & 0xffffu, c3[j] & 0xffffu,
                               c04[j] >> 16,     c15[j] >> 16,     c26[j] >> 16};
#pragma unroll
        for (int q = 0; q < 7; ++q)
            if (k[q]) atomicAdd(&part[q][c], k[q]);
    }
    __syncthreads();
    for (int i = tid; i < 7 * kCols16; i += 64 * kWaves16) {
        const int q = i / kCols16, c = i - q * kCols16;
        const int64_t col = (int64_t)blockIdx.x * kCols16 + c;
        const uint32_t n = part[q][c];
        if (col < aln_len && n) atomicAdd(counts + col * 7 + q, n);
    }
}

// Launch both kernels on `stream`. counts must be zeroed by the caller.
hipError_t consensus_launch(const uint8_t *rows, int64_t row_stride, int64_t aln_len, int64_t row_end, int64_t batch,
                            int num_cu, uint32_t *counts, char *cons, hipStream_t stream)
{
    if (row_end > 0) {
        const int64_t gx = (aln_len + kCols16 - 1) / kCols16;
        // about 4 workgroups (16 waves) per CU in all, each at most 4 x 65,535 rows
        int64_t want = (4LL * num_cu + gx - 1) / gx;
        if (want < 1) want = 1;
        int64_t rpg = (row_end + want - 1) / want;
        if (rpg < 64) rpg = 64;
        if (rpg > (int64_t)kWaves16 * kMaxWaveRows16) rpg = (int64_t)kWaves16 * kMaxWaveRows16;
        const int64_t gy = (row_end + rpg - 1) / rpg;
        if (gy > 65535) return hipErrorInvalidValue;
        const bool a16 = row_stride % 16 == 0 && ((uintptr_t)rows & 15) == 0;
        if (a16)
            hipLaunchKernelGGL(count_columns16<true>, dim3((unsigned)gx, (unsigned)gy), dim3(64 * kWaves16), 0, stream,
                               rows, row_stride, aln_len, row_end, rpg, batch, counts);
        else
            hipLaunchKernelGGL(count_columns16<false>, dim3((unsigned)gx, (unsigned)gy), dim3(64 * kWaves16), 0,
                               stream, rows, row_stride, aln_len, row_end, rpg, batch, counts);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(pick_consensus, dim3((unsigned)((aln_len + 255) / 256)), dim3(256), 0, stream, counts, aln_len,
                       cons);
    return hipGetLastError();
}

}  // namespace hal
