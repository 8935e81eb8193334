// dp_device.h -- device helpers shared by the DP kernels: packed 16-bit
// ops on int8-in-high-byte cells, band bounds, uniform-index register access,
// and the on-device backtrack.
#pragma once
#include <hip/hip_runtime.h>

#include "dp_common.h"

namespace hal {

constexpr int32_t kNegInf = -0x40000000;   // KSW_NEG_INF include/ksw2.h:6
constexpr int kQ = 6, kQE = 8;             // q, q+e of KSW2AlignConfig include/align.h:126-135

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
typedef short s16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t pk_add(uint32_t a, uint32_t b)
{
    return __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2, a) + __builtin_bit_cast(u16x2, b));
}
__device__ __forceinline__ uint32_t pk_sub(uint32_t a, uint32_t b)
{
    return __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2, a) - __builtin_bit_cast(u16x2, b));
}
__device__ __forceinline__ uint32_t pk_maxu(uint32_t a, uint32_t b)
{
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(u16x2, a),
                                                                  __builtin_bit_cast(u16x2, b)));
}
__device__ __forceinline__ uint32_t pk_minu(uint32_t a, uint32_t b)
{
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_min(__builtin_bit_cast(u16x2, a),
                                                                  __builtin_bit_cast(u16x2, b)));
}
__device__ __forceinline__ uint32_t pk_maxi(uint32_t a, uint32_t b)
{
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(s16x2, a),
                                                                  __builtin_bit_cast(s16x2, b)));
}
// signed saturating a - b per half: bit 15 (31) is set iff a < b (as int8 in the high byte)
__device__ __forceinline__ uint32_t pk_subsat(uint32_t a, uint32_t b)
{
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_sub_sat(__builtin_bit_cast(s16x2, a),
                                                                      __builtin_bit_cast(s16x2, b)));
}
// each 16-bit half: all ones if negative as int16, else 0
__device__ __forceinline__ uint32_t pk_ashr15(uint32_t a)
{
    return __builtin_bit_cast(uint32_t, __builtin_bit_cast(s16x2, a) >> (s16x2){15, 15});
}
__device__ __forceinline__ uint32_t bfi(uint32_t m, uint32_t a, uint32_t b) { return (a & m) | (b & ~m); }
// v in a scalar register the compiler cannot see through: a chain of bfi with
// constant masks is otherwise rewritten as and / or3 (five operations per
// three merges instead of three)
__device__ __forceinline__ uint32_t sreg(uint32_t v)
{
    asm("" : "+s"(v));
    return v;
}
// Traceback nibbles of the four cells of registers (2m, 2m+1): the flag words
// of register 2m in f?0 and of 2m+1 in f?1, each flag in bit 15 of its half.
// v_perm selectors 8-11 replicate bit 15 / 31 of the second source and 15 / 31
// of the first into a whole byte, so byte i of each gathered word is 0x00 or
// 0xFF for the cell (2m, 2m+8, 2m+1, 2m+9)[i], and one bfi per flag places it
// (the masks are set where they are used: held across the loop they cost
// scalar registers the loop needs).
// HI: flags (a, b, x, y) in bits (7, 6, 5, 4) of each byte; else in (3, 2, 1, 0).
// The other nibble of each byte is filler for the caller's final bfi.
template <bool HI>
__device__ __forceinline__ uint32_t tb_nibbles(uint32_t fa0, uint32_t fa1, uint32_t fb0, uint32_t fb1,
                                               uint32_t fx0, uint32_t fx1, uint32_t fy0, uint32_t fy1)
{
    const uint32_t ga = __builtin_amdgcn_perm(fa1, fa0, 0x0B0A0908u);
    const uint32_t gb = __builtin_amdgcn_perm(fb1, fb0, 0x0B0A0908u);
    const uint32_t gx = __builtin_amdgcn_perm(fx1, fx0, 0x0B0A0908u);
    const uint32_t gy = __builtin_amdgcn_perm(fy1, fy0, 0x0B0A0908u);
    constexpr uint32_t s = HI ? 4u : 0u;
    return bfi(sreg(0x08080808u << s), ga, bfi(sreg(0x04040404u << s), gb, bfi(sreg(0x02020202u << s), gx, gy)));
}
// cells (2j-1, 2j) from pairs hi=(2j,2j+1), lo=(2j-2,2j-1)
__device__ __forceinline__ uint32_t cat16(uint32_t hi, uint32_t lo) { return __builtin_amdgcn_alignbit(hi, lo, 16); }

// lane i receives lane i-1 (lane 0 receives lane 63): DPP wave_ror:1
__device__ __forceinline__ uint32_t wave_ror1(uint32_t v)
{
    // every lane is written (a full-wave rotate), so no old value is needed
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x13C, 0xF, 0xF, false);
}

// 16 bytes starting at an arbitrary byte address (aligned dword loads + v_alignbyte)
__device__ __forceinline__ void load16(const uint8_t *p, uint32_t (&o)[4])
{
    const uintptr_t a = (uintptr_t)p;
    const uint32_t *d = (const uint32_t *)(a & ~(uintptr_t)3);
    const uint32_t sh = (uint32_t)(a & 3);
    uint32_t wv[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) wv[i] = d[i];
#pragma unroll
    for (int i = 0; i < 4; ++i) o[i] = __builtin_amdgcn_alignbyte(wv[i + 1], wv[i], sh);
}

__device__ __forceinline__ uint32_t load_byte_uniform(const uint8_t *p)
{
    const uintptr_t a = (uintptr_t)p;
    const uint32_t wd = *(const uint32_t *)(a & ~(uintptr_t)3);
    return (wd >> (8 * (uint32_t)(a & 3))) & 0xffu;
}

struct Bounds { int st0, en0, st, en; };

// Band of anti-diagonal r: src/align/ksw2_extz2_sse.c:107-116
__device__ __forceinline__ bool diag_bounds(int r, int qlen, int tlen, int w, Bounds &b)
{
    int st = 0, en = tlen - 1;
    if (st < r - qlen + 1) st = r - qlen + 1;
    if (en > r) en = r;
    if (st < (r - w + 1) >> 1) st = (r - w + 1) >> 1;
    if (en > (r + w) >> 1) en = (r + w) >> 1;
    if (st > en) return false;
    b.st0 = st; b.en0 = en; b.st = st & ~15; b.en = (en & ~15) + 15;
    return true;
}

// Register-array access with a wave-uniform runtime index. Written as mask
// arithmetic over every element so the compiler keeps the arrays in VGPRs
// (an "if (idx == i) a[i] = v" pattern becomes an indexed scratch access).
__device__ __forceinline__ uint32_t sel_mask(int a, int b) { return 0u - (uint32_t)(a == b); }
__device__ __forceinline__ uint32_t pick8(const uint32_t (&a)[8], int j)
{
    uint32_t v = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) v |= a[i] & sel_mask(j, i);
    return v;
}
// cell c (0..15) of a block held as registers j = c & 7, half c >> 3 (x<<8
// form) set to v16 in every register-half that holds it (c wave-uniform)
__device__ __forceinline__ void set_cell(uint32_t (&a)[8], int c, uint32_t v16)
{
    const uint32_t hm = (c >> 3) ? 0xffff0000u : 0x0000ffffu;
    const uint32_t vv = v16 | (v16 << 16);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const uint32_t m = hm & sel_mask(c & 7, i);
        a[i] = (a[i] & ~m) | (vv & m);
    }
}

__device__ __forceinline__ int32_t pick16(const int32_t (&a)[16], int j)
{
    uint32_t v = 0;
#pragma unroll
    for (int i = 0; i < 16; ++i) v |= (uint32_t)a[i] & sel_mask(j, i);
    return (int32_t)v;
}
__device__ __forceinline__ void put16(int32_t (&a)[16], int j, int32_t val)
{
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const uint32_t m = sel_mask(j, i);
        a[i] = (int32_t)(((uint32_t)a[i] & ~m) | ((uint32_t)val & m));
    }
}
// Register-array access with a wave-uniform index through a uniform switch
// (one move or select instead of a select over every element).
__device__ __forceinline__ int32_t pick16u(const int32_t (&a)[16], int j)
{
    int32_t v = 0;
    switch (j) {
#define HAL_C(i) case i: v = a[i]; break;
        HAL_C(0) HAL_C(1) HAL_C(2) HAL_C(3) HAL_C(4) HAL_C(5) HAL_C(6) HAL_C(7)
        HAL_C(8) HAL_C(9) HAL_C(10) HAL_C(11) HAL_C(12) HAL_C(13) HAL_C(14) HAL_C(15)
#undef HAL_C
    default: break;
    }
    return v;
}
// a[j] = val in the lanes where `on` holds (j wave-uniform)
__device__ __forceinline__ void put16u(int32_t (&a)[16], int j, int32_t val, bool on)
{
    switch (j) {
#define HAL_C(i) case i: a[i] = on ? val : a[i]; break;
        HAL_C(0) HAL_C(1) HAL_C(2) HAL_C(3) HAL_C(4) HAL_C(5) HAL_C(6) HAL_C(7)
        HAL_C(8) HAL_C(9) HAL_C(10) HAL_C(11) HAL_C(12) HAL_C(13) HAL_C(14) HAL_C(15)
#undef HAL_C
    default: break;
    }
}
// byte of cell c (cells 2j, 2j+1 in register j, x<<8 halves), c wave-uniform
__device__ __forceinline__ int byte_of_u(const uint32_t (&a)[8], int c)
{
    uint32_t w = 0;
    switch (c >> 1) {
#define HAL_C(i) case i: w = a[i]; break;
        HAL_C(0) HAL_C(1) HAL_C(2) HAL_C(3) HAL_C(4) HAL_C(5) HAL_C(6) HAL_C(7)
#undef HAL_C
    default: break;
    }
    return (int)((w >> ((c & 1) * 16 + 8)) & 0xffu);
}

// unsigned byte value of cell c (0..15) of a block stored as x<<8 halves
__device__ __forceinline__ int byte_of(const uint32_t (&a)[8], int c)
{
    return (int)((pick8(a, c >> 1) >> ((c & 1) * 16 + 8)) & 0xffu);
}
__device__ __forceinline__ void set_half(uint32_t (&a)[8], int c, uint32_t v16)
{
    const uint32_t hm = (c & 1) ? 0xffff0000u : 0x0000ffffu;
    const uint32_t vv = v16 | (v16 << 16);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const uint32_t m = hm & sel_mask(c >> 1, i);
        a[i] = (a[i] & ~m) | (vv & m);
    }
}


// Backtrack of include/ksw2.h:129-161 (is_rot=1, is_rev=0, min_intron_len=0)
// from cell (target i0, query j0), run by one wave. Traceback rows are staged
// 64 anti-diagonals at a time into LDS (chunk, 64 x 40 bytes): along 64
// anti-diagonals the path moves at most 63 target columns, so 5 blocks per row
// cover it. Emits CIGAR ops in backtrack (reversed) order into J.cigar; returns
// the op count (which may exceed J.cig_cap, then nothing past the cap is written).
//
// Cell layouts of a 16-cell block inside its 8-byte traceback record:
//  layout 0 (multi-wave kernel, registers hold cells 2j, 2j+1): cell c is in
//    dword c>>3, byte c&3, nibble (c>>2)&1;
//  layout 1 (single-wave kernel, registers hold cells j, j+8): cell c is in
//    dword (c&7)>>2, byte ((c&1)<<1)|(c>>3), nibble ((c&7)>>1)&1.
// Nibble bits: 3 = state compare a, 2 = state compare b, 1 = x continuation,
// 0 = y continuation, stored as sign bits (inverted for RIGHT alignment).
__device__ inline int backtrack_to_cigar(const DpJob &J, const uint8_t *__restrict__ tb, size_t rowb, int w,
                                         bool right, int i0, int j0, uint32_t (*chunk_)[10], int layout)
{
    volatile uint32_t(*chunk)[10] = chunk_;
    const int lane = threadIdx.x & 63;
    const int qlen = J.qlen, tlen = J.tlen;
    int nout = 0;
    int i = i0, j = j0, state = 0;
    int rc = -1, bb0 = 0;
    int cur_op = -1, cur_len = 0;
    auto flush = [&]() {
        if (cur_len > 0) {
            if (lane == 0 && nout < J.cig_cap) J.cigar[nout] = (uint32_t)cur_len << 4 | (uint32_t)cur_op;
            ++nout;
        }
    };
    auto push = [&](int op, int len) {
        if (op == cur_op) cur_len += len;
        else { flush(); cur_op = op; cur_len = len; }
    };
    while (i >= 0 && j >= 0) {
        const int r = i + j;
        if (rc < 0 || r < rc - 63) {
            rc = r;
            bb0 = (i - 63 > 0 ? i - 63 : 0) >> 4;
            const int rr = rc - lane;
            uint32_t wv[10];
#pragma unroll
            for (int q2 = 0; q2 < 10; ++q2) wv[q2] = 0;
            Bounds br;
            if (rr >= 0 && diag_bounds(rr, qlen, tlen, w, br)) {
                const int sb = br.st >> 4, eb = br.en >> 4;
                const uint8_t *row = tb + (size_t)rr * rowb;
#pragma unroll
                for (int q2 = 0; q2 < 5; ++q2) {
                    const int bb = bb0 + q2;
                    if (bb >= sb && bb <= eb) {
                        const uint64_t v = __hip_atomic_load((const uint64_t *)(row + (size_t)(bb - sb) * 8),
                                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        wv[2 * q2] = (uint32_t)v;
                        wv[2 * q2 + 1] = (uint32_t)(v >> 32);
                    }
                }
            }
#pragma unroll
            for (int q2 = 0; q2 < 10; ++q2) chunk[lane][q2] = wv[q2];
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
        }
        Bounds b3;
        diag_bounds(r, qlen, tlen, w, b3);
        int force = -1, d = 0;
        if (i < b3.st) force = 2;
        if (i > b3.en) force = 1;
        if (force < 0) {
            const int c = i & 15;
            int dw, sh;
            if (layout == 0) { dw = c >> 3; sh = (c & 3) * 8 + ((c >> 2) & 1) * 4; }
            else { dw = (c & 7) >> 2; sh = (((c & 1) << 1) | (c >> 3)) * 8 + (((c & 7) >> 1) & 1) * 4; }
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
    flush();
    return nout;
}

}  // namespace hal
