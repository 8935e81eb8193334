/*
 * halign_oracle.h -- CPU restatement of HAlign-4's per-query center-star
 * alignment path, used ONLY as a test checker (tests/, smoke(), bench.py's
 * cpu_baseline leg). Nothing in the product library links or calls it.
 *
 * Every function restates a reference function; the citation is on the
 * declaration. Parity of each function is pinned in tests/ against the
 * reference compiled from its own sources (oracle/_ref) and the committed
 * golden fixtures in tests/golden/.
 */
#ifndef HALIGN_ORACLE_H
#define HALIGN_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ksw2 flag bits used on this path (include/ksw2.h:8-14) */
#define OR_KSW_RIGHT 0x02
#define OR_KSW_GENERIC_SC 0x04
#define OR_KSW_NEG_INF (-0x40000000)

/* Result fields of ksw_extz_t (include/ksw2.h:33-42). */
typedef struct {
    uint32_t max;
    int32_t zdropped;
    int32_t max_q, max_t;
    int32_t mqe, mqe_t;
    int32_t mte, mte_q;
    int32_t score;
    int32_t n_cigar;
} or_ez_t;

/* Anchor (include/anchor.h:39-50), laid out as 32 plain bytes. */
typedef struct {
    uint64_t hash;
    uint32_t rid_ref, pos_ref, rid_qry, pos_qry, span, is_rev;
} or_anchor_t;

/*
 * ksw_extz2_sse, production (SSE2) semantics, global mode, dna5 matrix
 * +5/-4/0, q=6, e=2, zdrop disabled, end_bonus 0.
 * src/align/ksw2_extz2_sse.c:23-304 ; ksw_backtrack include/ksw2.h:129-161.
 * query/target are 0..4 codes. Returns n_cigar, or -1 if cigar_cap is too
 * small (cigar may then be NULL to only get the count in ez->n_cigar).
 * If cells != NULL it receives sum_r (en0-st0+1) over computed diagonals.
 */
int or_ksw_extz2(int qlen, const uint8_t *query, int tlen, const uint8_t *target,
                 int w, int flag, or_ez_t *ez, uint32_t *cigar, int cigar_cap,
                 int64_t *cells);

/* align::auto_band(qlen=ref_len, tlen=query_len) as called from
 * globalAlignKSW2 (include/align.h:159-172, src/align/align.cpp:45). */
int or_auto_band(int ref_len, int query_len);

/* align::globalAlignKSW2(ref, query, cfg) src/align/align.cpp:22-65 with
 * cfg.flag = flag. ASCII input. Returns n_cigar or -1 (cap too small). */
int or_global_align_ksw2(const char *ref, int rlen, const char *qry, int qlen,
                         int flag, uint32_t *cigar, int cigar_cap, int64_t *cells);

/* minimizer::extractMinimizer src/seed/minimizer.cpp:103-177 (RingMinQueue
 * :35-97 emulated exactly, including the overwrite of the front slot when the
 * queue is full). *ub is set to 1 when the reference would index past its
 * ring buffer (undefined behaviour there; outputs are then the "larger
 * buffer" continuation), bit 1 if even the larger buffer overflowed (output
 * truncated), bit 2 when a push into a full queue overwrote the front slot
 * (defined behaviour, reproduced). Returns the hit count or -1 (cap). Hits are
 * (x,y) pairs as MinimizerHit (include/seed.h:206-265). */
int or_extract_minimizer(const char *seq, int n, int k, int w, int non_canonical,
                         uint64_t *x_out, uint64_t *y_out, int cap, int *ub);

/* SeedFilterParams (include/anchor.h:65-71) */
typedef struct {
    double f_top_frac;
    uint64_t u_floor, u_ceil;
    double q_occ_frac;
    uint64_t sample_every_bp;
} or_seed_params_t;

/* minimizer::collect_anchors src/seed/minimizer.cpp:220-356 with
 * default_mm2_params() (include/anchor.h:65-75). Returns count or -1. */
int or_collect_anchors(const uint64_t *ref_x, const uint64_t *ref_y, int nref,
                       const uint64_t *qry_x, const uint64_t *qry_y, int nqry,
                       or_anchor_t *out, int cap);
/* the same with explicit SeedFilterParams (the reference tests' form) */
int or_collect_anchors_ex(const uint64_t *ref_x, const uint64_t *ref_y, int nref,
                          const uint64_t *qry_x, const uint64_t *qry_y, int nqry,
                          const or_seed_params_t *p, or_anchor_t *out, int cap);

/* anchor::chainAnchors src/seed/anchor.cpp:332-461 (sorts `anchors` in place).
 * log2_lut[d] must hold log2f((float)d) for d in [0, 501]. Returns chain
 * length (0 = no chain). */
int or_chain_anchors(or_anchor_t *anchors, int n, const float *log2_lut,
                     or_anchor_t *out);
/* ChainParams (include/anchor.h:116-126) */
typedef struct {
    int32_t max_dist_x, max_dist_y, bw, max_skip, max_iter, min_cnt, min_score;
    float gap_penalty, skip_penalty;
} or_chain_params_t;
/* the same with explicit ChainParams (the reference tests' form); bw <= 500 */
int or_chain_anchors_ex(or_anchor_t *anchors, int n, const or_chain_params_t *cp, const float *log2_lut,
                        or_anchor_t *out);

/* align::globalAlignMM2 src/align/align.cpp:189-305. Returns n_cigar or -1.
 * stats (optional, 4 ints): segments run, forced I+D segments, whole-pair
 * fallbacks (empty chain or final mismatch), chain length. */
int or_global_align_mm2(const char *ref, int rlen, const char *qry, int qlen,
                        const or_anchor_t *anchors, int nanchors,
                        const float *log2_lut, uint32_t *cigar, int cigar_cap,
                        int64_t *cells, int32_t *stats);

/* RefAligner::globalAlign src/align/ref_aligner.cpp:91-125 with k=15, w=10,
 * non_canonical=true; ref hits computed here when nref == 0. */
int or_ref_global_align(const char *ref, int rlen, const char *qry, int qlen,
                        const uint64_t *ref_x, const uint64_t *ref_y, int nref,
                        const float *log2_lut, uint32_t *cigar, int cigar_cap,
                        int64_t *cells, int32_t *stats, int *ub);

/* cigar::delQueryToRefByCigar src/align/cigar.cpp:212-299. Returns the row
 * length, or -1 if cap is too small / the CIGAR does not consume qlen. */
int or_del_query_to_ref(const char *qry, int qlen, const uint32_t *cigar, int n,
                        char *out, int cap);

/* cigar::hasInsertion src/align/cigar.cpp:70-79 */
int or_has_insertion(const uint32_t *cigar, int n);

/* cigar::cigarToString src/align/cigar.cpp:82-97. Returns length or -1. */
int or_cigar_to_string(const uint32_t *cigar, int n, char *out, int cap);

/* consensus::generateConsensusSequence counting part
 * src/consensus/consensus.cpp:611-703 with processBatchParallelWithLocals
 * :304-411, pickConsensusChar :18-35, k_base_map include/consensus.h:126-145.
 * rows: nrows rows of aln_len bytes. counts: aln_len*7 u32 (A,C,G,T,U,N,-).
 * Returns number of rows counted, or -1 on error. */
int64_t or_consensus(const char *rows, int64_t nrows, int64_t aln_len,
                     int64_t batch_size, uint64_t seq_limit,
                     uint32_t *counts, char *cons_out);

/* seq_io::KseqReader::next over the vendored kseq_read (src/utils/seq_io.cpp:
 * 131-155, include/kseq.h kseq_read / ks_getuntil2), on an in-memory file of
 * len bytes (already decompressed). Writes up to max_records records: ids,
 * comments and sequences back to back into id_out / desc_out / seq_out
 * (capacity cap bytes each) with n+1 offsets each. Returns the number of
 * records, or -2 for a malformed FASTQ record (kseq's -2; the reference
 * throws), or -1 when an output is too small. */
int or_fastx_parse(const char *buf, int64_t len, int max_records,
                   char *id_out, int64_t *id_off, char *desc_out, int64_t *desc_off,
                   char *seq_out, int64_t *seq_off, int64_t cap);

/* The final keep-length FASTA of RefAligner::mergeAlignedResults for one
 * center and the copy MSA command (src/align/ref_aligner.cpp:676-740):
 * the center record with seq_io::cleanSequence applied (include/utils.h:
 * 161-197; writeConsensusAndReferences :541-561), then per query '>' name
 * '\n' row '\n' (seq_io::SeqWriter::writeFasta, src/utils/seq_io.cpp:236-268,
 * width unlimited) with the row from delQueryToRefByCigar (the insertion SAMs
 * through mergeConsensusAndSamToFasta :139-209) or padQueryToRefByCigar (the
 * other SAMs, convertSamToFastaRecord :467-498; the same row for a CIGAR
 * without I), the queries whose CIGAR has an insertion first (:713-732),
 * in input order with one thread; a query whose name starts with '@' is
 * dropped (SamReader::next skips such lines, src/utils/seq_io.cpp:433-436).
 * center_desc may be empty (no ' ' then).
 * Returns the bytes written, or -1 (capacity, or a row whose length differs
 * from the first row's: the reference throws, :590-596, :657-663). */
int64_t or_keep_length_fasta(const char *center_id, int id_len, const char *center_desc, int desc_len,
                             const char *center_seq, int64_t center_len, int n, const char *queries,
                             const int64_t *qoff, const char *names, const int64_t *noff,
                             const uint32_t *cigars, const int64_t *coff, char *out, int64_t cap);

/* getHash2bit (src/mash/hash.cpp): XXH3-64 with the given seed of the
 * 8-byte big-endian code (the vendored xxHash's 4..8-byte path). */
uint64_t or_hash2bit(uint64_t code, uint32_t seed);

/* mash::sketchFromSequence (src/mash/mash.cpp:32-146): the sketch_size
 * smallest distinct hashes of the valid k-mers (nt4 codes, N resets, forward
 * code or min(forward, reverse complement)), ascending. Returns their count or
 * -1 when cap is too small. */
int or_mash_sketch(const char *seq, int64_t n, int k, int sketch_size, int noncanonical, uint32_t seed,
                   uint64_t *out, int cap);

/* mash::jaccard (src/mash/mash.cpp:151-164): |a n b| / min(|a|, |b|), 1 for
 * two empty sketches, 0 for one. */
double or_mash_jaccard(const uint64_t *a, int na, const uint64_t *b, int nb);

/* log2f table used by or_chain_anchors: lut[d] = log2f(d), d in [0, 501]. */
void or_fill_log2_lut(float *lut);

#ifdef __cplusplus
}
#endif
#endif
