// mash_host.h -- device-resident sketches for the multi-reference run (the
// caller holds the context's lock).
#pragma once
#include <cstdint>
#include <vector>

#include "host_common.h"

namespace hal {

// mash::sketchFromSequence for n host sequences: sketch i at sk[i * size ..],
// its length in cnt[i] (int32)
int mash_sketch_to_device(halign_ctx *ctx, int n, const char *seqs, const int64_t *offsets, int k, int size,
                          bool noncanonical, uint32_t seed, DevBuf &sk, DevBuf &cnt);

// per query the first reference with the largest Jaccard, as num / den
int mash_select_device(halign_ctx *ctx, int nq, const DevBuf &qsk, const DevBuf &qcnt, int qstride, int nr,
                       const DevBuf &rsk, const DevBuf &rcnt, int rstride, std::vector<int32_t> &best,
                       std::vector<int32_t> &num, std::vector<int32_t> &den);

}  // namespace hal
