// pipeline_host.h -- internal entries of the per-query path used by the
// keep-length file run (the centers live in pipeline_api.hip).
#pragma once
#include <cstdint>

#include "../../include/halign_hip.h"

namespace hal {

// Column map of the center's positions in its aligned record (width columns);
// map = nullptr removes it. With a map, records carry width-byte rows.
int center_set_columns(halign_center *c, const int32_t *map, int64_t width);

// halign_center_align_fasta, plus each query's insertion flag in
// has_ins_out (optional, n bytes)
int center_align_fasta(halign_center *c, int32_t n, const char *queries, const int64_t *offsets, const char *names,
                       const int64_t *name_offsets, char *out, int64_t cap, int64_t *out_len, int64_t *ins_bytes,
                       int32_t *n_insertion, double *ms, uint8_t *has_ins_out);

}  // namespace hal
