"""halign: host-side mirror of HAlign-4's align/seed/anchor/consensus path on
MI355X. The compute runs in libhalign_hip.so (HIP, gfx950) through its C ABI.
"""
from ._lib import (ANCHOR_DTYPE, DNA5_MAT, DP_RESULT_DTYPE, DP_TASK_DTYPE, Center, Context, ExtzT,  # noqa: F401
                   HalignError, PairSet, RefSet, Server, align_sam, free_extz,
                   LIB_PATH, consensus_file, lib, msa_auto_center, msa_keep_length, qfix_policy, qfix_shapes, read_fastx,
                   ring_policy, ring_shapes, shape_names, topk_longest, wide_policy)

FLAG_RIGHT = 0x06   # KSW_EZ_GENERIC_SC | KSW_EZ_RIGHT (include/align.h:134)
FLAG_LEFT = 0x04    # KSW_EZ_GENERIC_SC, first MM2 segment (src/align/align.cpp:194-195)
