// dp_common.h -- shared definitions between the host C-ABI layer and the DP
// kernels of libhalign_hip. Plain C layout: no torch or STL types.
#pragma once
#include <stdint.h>

namespace hal {

// Padding (in 16-bit elements) around every spread sequence so that the
// 16-cell windows the kernels read at band margins stay inside the buffer.
constexpr int kSeqPad = 32;
// Padding (bytes) around the 0..4 byte-code copies: a 16-byte window is read as
// five aligned dwords, i.e. up to 19 bytes past its start and 3 before.
constexpr int kBytePad = 64;

// Job flags
constexpr int kFlagRight = 0x02;   // KSW_EZ_RIGHT (include/ksw2.h:9)
constexpr int kFlagHasN = 0x100;   // a base outside ACGT occurs in either sequence

// Status codes written by the kernels
constexpr int kStOk = 0;
constexpr int kStCigarOverflow = 1;   // CIGAR did not fit in cig_cap
constexpr int kStNeedStats = 2;       // band overflow hit in CIGAR-only mode
constexpr int kStTooWide = 3;         // band wider than the ring of this kernel

// One DP job = one ksw_extz2_sse call (src/align/ksw2_extz2_sse.c:23).
// Sequences are "spread": one 16-bit element per base with the 0..4 code in
// the high byte. Target elements carry 0x0C in the low byte (a v_perm
// zero-select), query elements carry 0. Each sequence is given twice: an
// even-aligned copy and a copy shifted by one element, so every 16-element
// window the kernel loads is 4-byte aligned.
struct DpJob {
    const uint16_t *tgt_even;   // tgt_even[k] = target element k (k may be < 0 down to -kSeqPad)
    const uint16_t *tgt_odd;    // tgt_odd[k]  = target element k+1
    const uint16_t *qr_even;    // reversed query: qr_even[k] = query[qlen-1-k]
    const uint16_t *qr_odd;     // qr_odd[k] = qr_even[k+1]
    const uint8_t *tgt_b;       // target as 0..4 byte codes, tgt_b[k] for k in [-kBytePad, tlen+kBytePad)
    const uint8_t *qr_b;        // reversed query as byte codes, same padding
    int32_t tlen, qlen, w, flags;
    uint32_t *cigar;            // output, ops in backtrack (reversed) order
    int32_t cig_cap;
    int32_t pad_;
};

// Result of one job: the ksw_extz_t fields (include/ksw2.h:33-42) plus
// bookkeeping. max/max_t/max_q/mqe/mqe_t are exact only in STATS mode.
struct DpOut {
    int32_t score, n_cigar, zdropped, status;
    uint32_t max;
    int32_t max_t, max_q, mqe, mqe_t, mte, mte_q, pad_;
    int64_t cells;   // sum over computed anti-diagonals of (en0 - st0 + 1)
};

// One part of a combined query-fixed launch: a bucket that runs beside the
// others of the launch in one kernel (workgroups [first, first + groups) of
// the grid serve it). shape: query-fixed shape index * 2 + gap direction
// (1 = RIGHT).
struct QfixPart {
    const DpJob *jobs;
    DpOut *outs;
    int *counter;
    uint8_t *tb_pool;
    uint64_t tb_slot_bytes;
    int32_t njobs, shape, first, pad_;
};

}  // namespace hal
