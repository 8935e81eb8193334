/*
 * halign_hip.h -- C ABI of libhalign_hip.so, the MI355X (gfx950) implementation
 * of HAlign-4's per-query center-star alignment path.
 *
 * Plain C: pointers and sizes only. Every entry point returns an int status
 * (HALIGN_OK == 0, negative on error); no exception crosses this boundary.
 * A context owns one HIP device and stream and is safe to share between
 * threads (calls on one context are serialised).
 */
#ifndef HALIGN_HIP_H
#define HALIGN_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HALIGN_OK 0
#define HALIGN_E_ARG (-1)          /* invalid argument */
#define HALIGN_E_UNSUPPORTED (-2)  /* parameters outside the path's fixed configuration */
#define HALIGN_E_HIP (-3)          /* HIP runtime error (see halign_last_error) */
#define HALIGN_E_NOMEM (-4)        /* device or host allocation failed */
#define HALIGN_E_CAPACITY (-5)     /* caller-provided output buffer too small */
#define HALIGN_E_KERNEL (-6)       /* a kernel reported an internal status */
#define HALIGN_PENDING 1           /* not an error: the request is still in flight */

typedef struct halign_ctx halign_ctx;

/* ksw_extz_t field layout (reference include/ksw2.h:33-42). cigar is
 * malloc()ed by halign_ksw_extz2 and must be released with free(), exactly as
 * the reference caller does (src/align/align.cpp:58-63). */
typedef struct {
    uint32_t max : 31, zdropped : 1;
    int max_q, max_t;
    int mqe, mqe_t;
    int mte, mte_q;
    int score;
    int m_cigar, n_cigar;
    int reach_end;
    uint32_t *cigar;
} halign_extz_t;

/* One DP task for halign_dp_batch: sequences as 0..4 codes (A,C,G,T,other),
 * offsets into the query/target code arrays passed alongside. */
typedef struct {
    int32_t qlen, tlen;
    int32_t w;      /* band width; < 0 = unbanded (ksw2 semantics) */
    int32_t flag;   /* KSW_EZ_* bits; GENERIC_SC (0x04) required, RIGHT (0x02) optional */
    int64_t query_off, target_off;
} halign_dp_task;

typedef struct {
    uint32_t max;
    int32_t zdropped, max_q, max_t, mqe, mqe_t, mte, mte_q, score, n_cigar;
    int32_t status, pad_;
    int64_t cells;      /* sum over computed anti-diagonals of the exact band width */
    int64_t cigar_off;  /* offset of this task's CIGAR in the caller's cigar buffer */
} halign_dp_result;

/* Context ------------------------------------------------------------------ */
int halign_init(int device, halign_ctx **out);
/* Frees the context. Centers and pair sets made on it may be freed before or
 * after it: while any is live the context's streams and buffers stay
 * allocated for them (they remain usable) and are released when the last one
 * is freed. The ctx handle itself must not be passed to any call after this. */
void halign_free(halign_ctx *ctx);
const char *halign_last_error(const halign_ctx *ctx);
/* library version string */
const char *halign_version(void);
/* Kernel family choice: 0 = automatic (a query-fixed shape for full-band
 * jobs whose query is the short side, else the ring-form kernel where the band
 * fits, multi-wave otherwise, the wide-band kernel beyond the widest
 * multi-wave ring), 1 = multi-wave kernels only, 2 + i = ring shape i
 * wherever the band fits it (automatic otherwise), 2 + R (R = number of ring
 * shapes) = the wide-band kernel for every job, 3 + R + q = query-fixed shape
 * q (halign_qfix_shapes) for the jobs it holds (automatic otherwise). Results
 * are identical;
 * used to test every family and to tune the shape choice. No reference
 * counterpart (the reference has one CPU kernel). */
int halign_set_kernel_policy(halign_ctx *ctx, int policy);
/* Band geometry of ksw_extz2_sse for (qlen, tlen, w) (reference
 * src/align/ksw2_extz2_sse.c:75-76 n_col_, :107-114 band and overflow): the
 * 16-cell column count n_col_, whether the band overflows ("zdropped"), and
 * the in-band cells computed. Host-only, no GPU needed. No reference
 * counterpart (the reference computes these inline). */
int halign_band_info(int32_t qlen, int32_t tlen, int32_t w, int32_t *ncol, int32_t *overflow,
                     int64_t *cells);
/* Ring shapes of the ring-form kernel: returns their number and fills up to
 * cap entries of waves per workgroup (0 for a sub-wave shape, whose wave runs
 * 64/lanes jobs at once), lanes per job and ring size in 16-cell blocks (a
 * shape serves bands of n_col_ <= ring - 2 blocks). Each sub-wave shape is
 * listed twice: first its per-group form, then its uniform-ticket form (all
 * jobs of a wave share qlen, tlen and band; the library groups the jobs and
 * pads each group with repeats). A ring shape listed a second time with the
 * same waves, lanes and ring is its two-per-SIMD form (registers for two
 * workgroups per SIMD; the planner takes it only for buckets that fill every
 * SIMD with two and whose slots the traceback budget holds). No reference
 * counterpart. */
int halign_ring_shapes(int32_t *waves, int32_t *lanes, int32_t *ring, int32_t cap);
/* Query-fixed shapes: returns their number and fills up to cap entries of
 * lanes per job (1, 4, 16 or 64; a wave runs 64/lanes jobs) and slots per
 * lane. Shape q serves CIGAR-mode jobs whose band is the whole matrix (w < 0
 * or w >= max(qlen, tlen) - 1) and whose query is the shorter side with at
 * most 16 * lanes * slots bases: the anchor-less whole pairs and the long
 * end segments of the per-query path. No reference counterpart. */
int halign_qfix_shapes(int32_t *lanes, int32_t *slots, int32_t cap);

/* Drop-in for ksw_extz2_sse (reference include/ksw2.h:64-65, called at
 * src/align/align.cpp:49-55): the plain one-pair-per-launch GPU form with every
 * ksw_extz_t field exact. Supports the path's configuration only: m=5, the
 * dna5 matrix (include/align.h:109-116), q=6, e=2, zdrop<0, end_bonus=0, flag
 * GENERIC_SC with or without RIGHT; anything else -> HALIGN_E_UNSUPPORTED.
 * km is ignored (kalloc pool in the reference). As ksw2: ez is reset first
 * (ksw_reset_extz, include/ksw2.h:184-189), m <= 0, qlen <= 0 or tlen <= 0
 * return with the reset ez (src/align/ksw2_extz2_sse.c:56-57), and ez->cigar
 * grows by realloc as ksw_push_cigar grows it (m_cigar 4 when empty, then
 * doubling; include/ksw2.h:113-123), so one ez can be reused across calls and
 * its cigar is released with free(). Any band width and length: bands wider
 * than the widest ring kernel run on the wide-band kernel. */
int halign_ksw_extz2(halign_ctx *ctx, void *km, int qlen, const uint8_t *query, int tlen,
                     const uint8_t *target, int8_t m, const int8_t *mat, int8_t q, int8_t e,
                     int w, int zdrop, int end_bonus, int flag, halign_extz_t *ez);

/* Many ksw_extz2_sse calls in one submission (batched persistent form; or the
 * plain form, one launch per task, when plain != 0). full_stats != 0 makes
 * max/max_t/max_q/mqe/mqe_t exact for every task (otherwise they are only
 * computed where the band overflows, which is where the reference uses them).
 * CIGARs are written back to back into cigar_out (capacity cigar_cap uint32),
 * each result holds its offset and length. kernel_ms (optional) receives the
 * summed DP kernel time measured with HIP events. */
int halign_dp_batch(halign_ctx *ctx, int32_t n, const halign_dp_task *tasks,
                    const uint8_t *query_codes, const uint8_t *target_codes, int32_t full_stats,
                    int32_t plain, halign_dp_result *results, uint32_t *cigar_out,
                    int64_t cigar_cap, double *kernel_ms);

/* align::globalAlignKSW2(ref, query, cfg) for n ASCII pairs (reference
 * src/align/align.cpp:22-65, cfg.flag = flags[i]): empty-side handling,
 * ScoreChar2Idx encoding, auto_band, DP, CIGAR. Pair i's CIGAR is written at
 * cigar_out[offsets[i] .. offsets[i+1]); offsets has n+1 entries. */
int halign_global_align_batch(halign_ctx *ctx, int32_t n, const char *const *refs,
                              const int32_t *ref_lens, const char *const *queries,
                              const int32_t *query_lens, const int32_t *flags,
                              uint32_t *cigar_out, int64_t cigar_cap, int64_t *offsets,
                              int32_t *scores);

/* Minimizer kernel choice: 0 = the wave form wherever it applies (the
 * sequential form takes the rest), 1 = the sequential form only. Results are
 * identical; used to test both forms. No reference counterpart. */
int halign_set_seed_policy(halign_ctx *ctx, int policy);

/* minimizer::extractMinimizer (reference src/seed/minimizer.cpp:103-177,
 * declared include/seed.h:283-286) for n sequences: sequence s is
 * seqs[offsets[s] .. offsets[s+1]) (ASCII). Hits are MinimizerHit records
 * (include/seed.h:206-265): x = hash56 << 8 | span, y = (rid | strand<<31) << 32
 * | pos, rid 0; sequence s's hits are x_out/y_out[hit_offsets[s] ..
 * hit_offsets[s+1]) (hit_offsets has n+1 entries; cap = capacity of the
 * outputs). status_out (optional, n entries): bit 0 = computed by the exact
 * sequential form (the reference's queue overwrote its front slot, canonical
 * k-mers, or w > 32), bit 1 = the reference indexes past its ring buffer on
 * this input (undefined behaviour there; the output follows a larger buffer),
 * bit 2 = even the larger buffer overflowed (output truncated). k, w guards as
 * the reference: k = 0, w = 0, len < k, k > 31 or w >= 256 give no hits. */
int halign_extract_minimizers(halign_ctx *ctx, int32_t n, const char *seqs, const int64_t *offsets, int32_t k,
                              int32_t w, int32_t non_canonical, uint64_t *x_out, uint64_t *y_out, int64_t cap,
                              int64_t *hit_offsets, int32_t *status_out);

/* A center kept resident for the per-query path (RefAligner with one
 * center, reference src/align/ref_aligner.cpp:40-41,91-125): its sequence,
 * its minimizers (extractMinimizer(center, 15, 10, non-canonical)) and their
 * index (hits sorted by hash, rid, pos, strand; distinct hashes with their
 * runs; the reference-occurrence threshold of compute_ref_occ_threshold,
 * src/seed/anchor.cpp:111-117). */
typedef struct halign_center halign_center;
int halign_center_create(halign_ctx *ctx, const char *center, int32_t len, halign_center **out);
/* As halign_center_create, aligning to `center` with the minimizer index built
 * from index_seq: RefAligner keeps the minimizers of the center file's record
 * (ref_aligner.cpp:35-44) but aligns to the consensus of that file (:47-60,
 * :243-251), which differs from the record where it has N or lower case.
 * When index_seq has no minimizers the center's own are used (globalAlign
 * extracts them on the spot then, :103-106). */
int halign_center_create_indexed(halign_ctx *ctx, const char *center, int32_t len, const char *index_seq,
                                 int32_t index_len, halign_center **out);
/* Frees the center (before or after its context, see halign_free). */
void halign_center_free(halign_center *c);
/* minimizer::collect_anchors(center hits, query hits) (src/seed/minimizer.cpp:
 * 220-356) for n ASCII queries (queries[offsets[i] .. offsets[i+1])), default
 * SeedFilterParams. anchors_out: 32-byte anchor::Anchor records
 * (include/anchor.h:39-50: hash u64, rid_ref, pos_ref, rid_qry, pos_qry, span,
 * is_rev u32), query i's at anchor_offsets[i] .. anchor_offsets[i+1]. */
int halign_center_anchors(halign_center *c, int32_t n, const char *queries, const int64_t *offsets,
                          void *anchors_out, int64_t cap, int64_t *anchor_offsets);
/* anchor::chainAnchors on those anchors (src/seed/anchor.cpp:332-461, default
 * ChainParams): the best chain of each query, in position order. */
int halign_center_chains(halign_center *c, int32_t n, const char *queries, const int64_t *offsets,
                         void *chains_out, int64_t cap, int64_t *chain_offsets);
/* RefAligner::globalAlign(center, query) (src/align/ref_aligner.cpp:91-125 =
 * extractMinimizer -> collect_anchors -> globalAlignMM2, src/align/align.cpp:
 * 189-305) for n queries: query i's CIGAR at cigar_out[cigar_offsets[i] ..
 * cigar_offsets[i+1]). stats (optional, 4 per query): segments, chain length,
 * minimizer status bits, anchors. ms (optional): device time of the batch. */
int halign_center_align(halign_center *c, int32_t n, const char *queries, const int64_t *offsets,
                        uint32_t *cigar_out, int64_t cap, int64_t *cigar_offsets, int32_t *stats, double *ms);
/* halign_center_align plus, per query, cigar::delQueryToRefByCigar (src/align/
 * cigar.cpp:212-299: the keep-length row, center columns; rows_out holds n
 * rows row_stride bytes apart, row_len[i] its length, a row longer than
 * row_stride is cut) and cigar::hasInsertion (:70-79, the SAM routing of
 * RefAligner::alignOneQueryToRef, src/align/ref_aligner.cpp:212-283) in
 * has_insertion[i]. The rows are made on the device from the CIGARs and the
 * raw query bytes. */
int halign_center_align_rows(halign_center *c, int32_t n, const char *queries, const int64_t *offsets,
                             uint32_t *cigar_out, int64_t cap, int64_t *cigar_offsets, char *rows_out,
                             int64_t row_stride, int64_t *row_len, uint8_t *has_insertion, int32_t *stats,
                             double *ms);

/* The keep-length MSA records of one batch (RefAligner::mergeAlignedResults
 * with keep_length and one center, src/align/ref_aligner.cpp:676-740: rows of
 * queries whose CIGAR has an insertion come from the merged insertion SAMs,
 * :504-535 and :715-717, then the others, :719-732; every row is the query
 * projected by delQueryToRefByCigar / padQueryToRefByCigar, src/align/
 * cigar.cpp:150-299, i.e. the center's length; records as seq_io::SeqWriter::
 * writeFasta with an unlimited line width writes them, src/utils/seq_io.cpp:
 * 236-268). Query i (queries[offsets[i] .. offsets[i+1])) with name
 * names[name_offsets[i] .. name_offsets[i+1]) becomes '>' name '\n' row '\n'
 * (none for a name starting with '@': the reference's SamReader skips such
 * lines, src/utils/seq_io.cpp:433-436); records of queries with an insertion
 * come first, each group in input order:
 * out[0 .. *ins_bytes) holds the *n_insertion records of the first group,
 * out[*ins_bytes .. *out_len) the others. The rows and records are made on the
 * device. cap >= sum over the kept queries of (name length + center length +
 * 3). */
int halign_center_align_fasta(halign_center *c, int32_t n, const char *queries, const int64_t *offsets,
                              const char *names, const int64_t *name_offsets, char *out, int64_t cap,
                              int64_t *out_len, int64_t *ins_bytes, int32_t *n_insertion, double *ms);

#define HALIGN_E_IO (-7)   /* file open/read/write error or malformed FASTA/FASTQ */

/* seq_io::KseqReader (src/utils/seq_io.cpp:67-155 over the vendored
 * include/kseq.h): FASTA and FASTQ records from a gzip or plain file, with
 * kseq's rules for names (up to the first whitespace), comments (the rest of
 * the header line), sequence lines (concatenated, empty lines skipped, one
 * trailing '\r' per line dropped) and quality lines (read, then dropped). A
 * FASTQ record whose quality is missing or of another length is an error
 * (HALIGN_E_IO), as kseq_read's -2 makes the reference throw. */
typedef struct halign_fastx halign_fastx;
int halign_fastx_open(const char *path, halign_fastx **out);
/* Up to max_records next records (fewer at the end of the file; *n = 0 at
 * its end). The views stay valid until the next call: record i's sequence is
 * seqs[seq_offsets[i] .. seq_offsets[i+1]), its name ids[id_offsets[i] ..
 * id_offsets[i+1]), its comment descs[desc_offsets[i] .. desc_offsets[i+1]). */
int halign_fastx_next(halign_fastx *f, int32_t max_records, int32_t *n, const char **seqs,
                      const int64_t **seq_offsets, const char **ids, const int64_t **id_offsets,
                      const char **descs, const int64_t **desc_offsets);
const char *halign_fastx_error(const halign_fastx *f);
void halign_fastx_close(halign_fastx *f);

/* halign4 -c center.fa --keep-length with one center and the copy MSA
 * command (src/halign4.cpp:96-147): the RefAligner constructor's center
 * (consensus of the center file, ref_aligner.cpp:20-70), alignQueryToRef over
 * chunks of `batch` queries (:286-394, batch <= 0: 25600) and
 * mergeAlignedResults (:676-740) into output_path: the center records cleaned
 * (seq_io::cleanSequence, include/utils.h:161-197), then the keep-length rows
 * as halign_center_align_fasta orders them over the whole input (the
 * reference's order with one thread). Reading the next chunk, aligning one on
 * the GPU and writing the previous one's records overlap. A center file with
 * more than one record runs the multi-reference path below with the copy
 * command.
 * Runs on one context take turns; the context keeps two pinned host buffers
 * of one chunk's records and the last run's center with its index until
 * halign_free, so later runs against the same center start warm. The rows of
 * queries without an insertion are spooled to output_path + ".rest.tmp"
 * (removed at the end).
 * stats (optional): {records written, insertion records, bytes written,
 * seconds total, seconds of GPU calls, seconds the GPU waited for input}. */
typedef struct {
    int64_t records, insertion_records, bytes;
    double seconds, gpu_seconds, input_wait_seconds;
} halign_msa_stats;
int halign_msa_keep_length(halign_ctx *ctx, const char *center_path, const char *input_path,
                           const char *output_path, int32_t batch, halign_msa_stats *stats);
/* The same with a center file of several reference sequences (RefAligner
 * with more than one reference and keep_length, ref_aligner.cpp:228-266):
 * each query is aligned to the reference whose sketch has the largest
 * Jaccard with its own (halign_mash_best), and its row is spread over the
 * columns of that reference's record in the aligned file (what the MSA
 * command made of the center file; aligned_path NULL = the copy command, the
 * center file itself), '-' in its gap columns (padQueryToRefByCigar with the
 * record's M/D CIGAR, :400-459, :486-496). The output starts with the aligned
 * file's records, cleaned. Aligned records must all have one width and hold
 * their reference's bases (the reference's final length check and CIGAR
 * lengths); otherwise HALIGN_E_ARG. With one record this is
 * halign_msa_keep_length (aligned_path then gives the file whose consensus
 * is the center). */
int halign_msa_keep_length_refs(halign_ctx *ctx, const char *center_path, const char *aligned_path,
                                const char *input_path, const char *output_path, int32_t batch,
                                halign_msa_stats *stats);

/* RefAligner over its whole reference set (src/align/ref_aligner.cpp:20-73,
 * alignOneQueryToRef :212-283): nref reference records (refs[r], ref_lens[r]
 * bytes, as read from the center file) and the consensus of the MSA command's
 * output for them (the sequence RefAligner names "consensus"). One reference:
 * queries are aligned to the consensus, seeded with the record's minimizers.
 * Several: each query goes to the first reference with the largest sketch
 * Jaccard (k 15, 2,000 hashes, seed 42), aligned to and seeded from it; that
 * is the result with keep_length, otherwise the query is aligned again to the
 * consensus seeded with the consensus's own minimizers (consensus_minimizer,
 * :71-72) and that second alignment is the result (:268-282). */
typedef struct halign_refset halign_refset;
int halign_refset_create(halign_ctx *ctx, int32_t nref, const char *const *refs, const int32_t *ref_lens,
                         const char *consensus, int32_t cons_len, halign_refset **out);
/* For n queries (queries[offsets[i] .. offsets[i+1])): best_ref[i] (optional)
 * the reference chosen (0 with one reference), the first alignment's CIGAR at
 * init_cigar[init_off[i] .. init_off[i+1]) (to that reference, or to the
 * consensus with one reference), the result's CIGAR at final_cigar[final_off
 * [i] .. final_off[i+1]), has_ins[i] (optional) = cigar::hasInsertion of the
 * result (the SAM file it goes to), final_ref[i] (optional) = the reference
 * the result is against, -1 for the consensus. Offsets have n+1 entries. */
int halign_refset_align(halign_refset *rs, int32_t n, const char *queries, const int64_t *offsets,
                        int32_t keep_length, int32_t *best_ref, uint32_t *init_cigar, int64_t init_cap,
                        int64_t *init_off, uint32_t *final_cigar, int64_t final_cap, int64_t *final_off,
                        uint8_t *has_ins, int32_t *final_ref);
void halign_refset_free(halign_refset *rs);
/* RefAligner::alignQueryToRef with one thread (src/align/ref_aligner.cpp:
 * 286-394): the SAM files of thread 0, sam_path (queries whose result has no
 * insertion) and insertion_sam_path (the others), each with the header line
 * "@HD VN:1.6 SO:unknown" and one line per query in input order as
 * writeSamRecord writes it (:128-136, cigar::cigarToString src/align/
 * cigar.cpp:82-97, seq_io::makeSamRecord / SeqWriter::writeSam src/utils/
 * seq_io.cpp:289-360): qname, 0, the reference name ("consensus" or the
 * chosen record's name), 1, 60, the CIGAR, *, 0, 0, the query as read, *.
 * The consensus is that of aligned_path, the MSA command's output for the
 * center file (NULL: the copy command, the center file itself). stats
 * (optional): records, insertion records, bytes written, seconds, seconds in
 * the GPU calls. */
int halign_align_sam(halign_ctx *ctx, const char *center_path, const char *aligned_path, const char *input_path,
                     const char *sam_path, const char *insertion_sam_path, int32_t keep_length, int32_t batch,
                     halign_msa_stats *stats);

/* preprocessInputFasta's consensus input (src/utils/preprocess.cpp:104-160
 * with TopKLongestSelector, src/consensus/selector.cpp:48-215): the k longest
 * records of input_path after seq_io::cleanSequence, longest first and the
 * earlier one first among equal lengths, written as FASTA with 80-column
 * lines; n_records (optional) = records read. Host only; ctx (optional)
 * receives the error message. */
int halign_topk_longest(halign_ctx *ctx, const char *input_path, int32_t k, const char *output_path,
                        int64_t *n_records);
/* consensus::generateConsensusSequence on a file (src/consensus/consensus.cpp:
 * 611-703): the reference's reading loop (batches of `batch` records with the
 * record after each batch dropped, at most seq_limit kept, 0 = all, every kept
 * record of the first one's length), the counts on the GPU (halign_consensus),
 * out_fasta (">consensus", 80-column lines) and out_json (the reference's
 * JSON: num_seqs, aln_len, counts A C G T U N -). cons_out (optional, cap
 * bytes) receives the consensus, cons_len its length. An empty file, an empty
 * first record or a length mismatch: HALIGN_E_ARG (the reference throws). */
int halign_consensus_file(halign_ctx *ctx, const char *aligned_path, const char *out_fasta, const char *out_json,
                          uint64_t seq_limit, int64_t batch, char *cons_out, int64_t cap, int64_t *cons_len);
/* halign4 --keep-length without a center (src/halign4.cpp:87, 119-147): the
 * cons_n longest cleaned records (halign_topk_longest), their alignment by the
 * MSA command (aligned_topk_path: its output; NULL: the copy command, the
 * records as they are, which must then be of one length), its consensus with
 * seq_limit cons_n, and the keep-length run against that consensus
 * (halign_msa_keep_length). Intermediate files sit next to output_path and
 * are removed. */
int halign_msa_auto_center(halign_ctx *ctx, const char *input_path, const char *aligned_topk_path,
                           const char *output_path, int32_t cons_n, int32_t batch, halign_msa_stats *stats);

/* mash::sketchFromSequence (src/mash/mash.cpp:32-146, getHash2bit src/mash/
 * hash.cpp) for n sequences (seqs[offsets[i] .. offsets[i+1]), ASCII): the
 * sketch_size smallest distinct XXH3-64 hashes (seeded) of the k-mers without
 * a non-ACGTU base, forward codes (non_canonical != 0) or canonical ones,
 * ascending; sequence i's at hashes_out[sketch_offsets[i] .. sketch_offsets
 * [i+1]). RefAligner uses k = 15, sketch_size = 2000, seed 42, non-canonical
 * (ref_aligner.cpp:35-44, 217-222). k = 0, k > 32, sketch_size = 0 or a
 * sequence shorter than k give an empty sketch, as in the reference. */
int halign_mash_sketch(halign_ctx *ctx, int32_t n, const char *seqs, const int64_t *offsets, int32_t k,
                       int32_t sketch_size, int32_t non_canonical, uint32_t seed, uint64_t *hashes_out,
                       int64_t cap, int64_t *sketch_offsets);
/* RefAligner::alignOneQueryToRef's choice among several references
 * (ref_aligner.cpp:228-244 with mash::jaccard, src/mash/mash.cpp:151-164):
 * for each query sketch the first reference sketch with the largest
 * |a n b| / min(|a|, |b|) (1 for two empty sketches, 0 for one) in best[i],
 * that value in jaccard[i] (optional). Sketches are sorted unique hash lists
 * as halign_mash_sketch returns them. */
int halign_mash_best(halign_ctx *ctx, int32_t nq, const uint64_t *q_hashes, const int64_t *q_offsets, int32_t nr,
                     const uint64_t *r_hashes, const int64_t *r_offsets, int32_t *best, double *jaccard);

/* Column counts and consensus of an alignment (the counting core of
 * consensus::generateConsensusSequence, reference src/consensus/consensus.cpp:
 * 611-703, pickConsensusChar :18-35, SiteCount include/consensus.h:78-99).
 * rows: nrows rows of aln_len characters, row_stride bytes apart, on the host
 * (rows_on_device = 0) or on the device (rows_on_device != 0: 4-byte aligned,
 * row_stride a multiple of 4 and >= aln_len rounded up to 4). Rows are taken
 * as the reference reads them: in batches of `batch` records (4096 there) with
 * the record after each full batch dropped, and at most seq_limit counted
 * rows (0 = all). counts_out: aln_len x 7 uint32 (A C G T U N -), cons_out:
 * aln_len characters (no terminator), num_seqs: rows counted. nrows <= 0 is
 * HALIGN_E_ARG (the reference throws on empty input). */
int halign_consensus(halign_ctx *ctx, const char *rows, int64_t nrows, int64_t aln_len, int64_t row_stride,
                     int32_t rows_on_device, int64_t batch, uint64_t seq_limit, uint32_t *counts_out,
                     char *cons_out, uint64_t *num_seqs, double *kernel_ms);

/* halign_consensus for rows already on the device (rows_on_device = 1), with
 * no device or host allocation: no counterpart in the reference; it is for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). The counts and consensus
 * characters are built in a static device array of this library (one per
 * GPU, kept until the process ends) and copied to counts_out / cons_out on
 * the host, so aln_len is at most 1,048,576 columns (HALIGN_E_CAPACITY
 * beyond). Calls on one context are serialised; calls on two contexts of the
 * same GPU that overlap race on that array. Results equal halign_consensus's. */
int halign_consensus_noalloc(halign_ctx *ctx, const char *rows_dev, int64_t nrows, int64_t aln_len,
                             int64_t row_stride, int64_t batch, uint64_t seq_limit, uint32_t *counts_out,
                             char *cons_out, uint64_t *num_seqs);

/* Device-resident workload: one center (target) and n queries kept in HBM,
 * aligned pairwise as whole pairs (globalAlignKSW2 semantics). Used for the
 * kernel-only DP throughput measurement and by the serving loop. */
typedef struct halign_pairset halign_pairset;
int halign_pairset_create(halign_ctx *ctx, const char *center, int32_t center_len, int32_t n,
                          const char *queries, const int64_t *query_offsets,
                          halign_pairset **out);
/* Align queries [first, first+count). plain != 0: one pair per launch.
 * kernel_ms: summed time of the DP kernels (HIP events on the library's
 * stream); cells: algorithmic DP cells of the range. */
int halign_pairset_run(halign_pairset *ps, int32_t first, int32_t count, int32_t plain,
                       double *kernel_ms, int64_t *cells);
/* Score and CIGAR of query idx from the last run that covered it. */
int halign_pairset_result(halign_pairset *ps, int32_t idx, int32_t *score, uint32_t *cigar,
                          int32_t cigar_cap, int32_t *n_cigar);
/* Frees the pair set (before or after its context, see halign_free). */
void halign_pairset_free(halign_pairset *ps);

/* on != 0 (the default): the batched DP launches of a call (one persistent
 * kernel per length/shape bucket, on the main and side streams) are captured
 * into a HIP graph, which later calls with the same launch plan replay; 0:
 * direct launches. Results are identical. No reference counterpart. */
int halign_set_graph_mode(halign_ctx *ctx, int32_t on);
/* Graphs captured and graph replays so far on this context. No reference
 * counterpart. */
int halign_graph_counts(halign_ctx *ctx, int64_t *captures, int64_t *replays);

/* Measured cost table of the DP batch planner (the autotuner). For each ring
 * shape (halign_ring_shapes), one job's time on a full chip is about
 * anti-diagonals x ns_diag + ns_job, with ns_diag piecewise linear in the
 * band's 16-cell column count n_col_ between the measured bands (constant
 * beyond the last; a shape is no candidate for bands narrower than the
 * first); the planner (automatic kernel policy) gives
 * each bucket of jobs the shape whose estimate max(longest job, all jobs /
 * jobs_in_flight) is lowest, and spreads the buckets over the streams by the
 * same estimates. The band itself stays the reference's fixed rule
 * (include/align.h:159-172). Field uniform: 1 for the uniform-ticket form of
 * a sub-wave shape, 2 for the two-per-SIMD form of a ring shape. No reference
 * counterpart. */
typedef struct {
    int32_t waves, lanes, ring, uniform;  /* the shape, as halign_ring_shapes lists it */
    int32_t jobs_in_flight;               /* jobs resident at once at full occupancy */
    int32_t npoints;                      /* bands measured (1..4) */
    int32_t ncol[4];                      /* the bands, in 16-cell columns, rising */
    double ns_diag[4];                    /* ns per anti-diagonal of one job at each band */
    double ns_job;                        /* fixed ns per job (ticket, set-up, backtrack start) */
} halign_tune_entry;
/* Measures every ring shape on this context's device (equal-length
 * synthetic jobs filling the chip once, two lengths at up to four bands per
 * shape, about a minute) and makes the result the context's planner table. A new
 * context starts from the table measured when the library was built
 * (halign.tune writes it). Not to be called while other calls use the
 * context. */
int halign_tune(halign_ctx *ctx);
/* The context's planner table: returns HALIGN_OK, fills up to cap entries
 * and sets *n to the number of entries. */
int halign_tune_table(halign_ctx *ctx, halign_tune_entry *out, int32_t cap, int32_t *n);
/* The ring shape (halign_ring_shapes index) the planner gives a batch of n
 * equal jobs (qlen, tlen, band w; w < 0: unbanded), R + q for query-fixed
 * shape q (R = number of ring shapes), or -1 when the job goes to another
 * kernel family (the multi-wave or wide-band kernels). */
int halign_plan_shape(halign_ctx *ctx, int32_t n, int32_t qlen, int32_t tlen, int32_t w, int32_t *ring_shape);

/* Per-bucket DP statistics (a diagnostic mode): with bucket timing on, every
 * DP launch of the context (the planned buckets of the batch runner and the
 * per-query path's small-segment launches) is bracketed by HIP events on its
 * stream, and graph replay is skipped; halign_last_buckets returns the launches
 * recorded since its previous call and forgets them (out NULL: only *n, the
 * records stay). family: 0 a planned
 * ring-kernel bucket, 1 a small-segment launch of the per-query path, 2 a
 * multi-wave or wide-band bucket, 3 a combined launch of several query-fixed
 * buckets bound by their anti-diagonal chains (shape -1, right -1; jobs and
 * cells of all its parts); shape: the ring shape index
 * (halign_ring_shapes, R + q for query-fixed shape q) for 0 and 1. cells: in-band cells computed. No
 * reference counterpart. */
typedef struct {
    int32_t family, shape, stream, right;
    int64_t jobs, cells;
    double kernel_ms;   /* this launch alone, from its events */
    double est_ms;      /* the planner's estimate (0 where it has none) */
} halign_bucket_stat;
int halign_set_bucket_timing(halign_ctx *ctx, int32_t on);
int halign_last_buckets(halign_ctx *ctx, halign_bucket_stat *out, int32_t cap, int32_t *n);

/* Device memory the context may take for traceback slots (default: the HBM
 * free at halign_init but a reserve of 4%, at least 8 GiB; the pool is
 * allocated as launches need it). Each batch plans within the smaller of this
 * budget and the pool already held plus the HBM free then, less 1 GiB, so
 * memory taken elsewhere after halign_init shrinks the launches instead of
 * failing them. No reference counterpart. */
int halign_set_traceback_budget(halign_ctx *ctx, int64_t bytes);

/* Continuous-batching server for RefAligner::globalAlign against one center
 * (src/align/ref_aligner.cpp:91-125; the reference's own request loop is the
 * OpenMP chunk loop of alignQueryToRef, ref_aligner.cpp:348-386). `workers`
 * contexts each keep the center and its minimizer index resident. Requests
 * (groups of ASCII queries) are cut into chunks of chunk_queries queries; a
 * worker that becomes free forms its next batch from one chunk of each
 * waiting request in arrival order, round after round, until max_batch_queries
 * or its share of the HBM footprint budget (hbm_budget bytes over all
 * workers, 0 = 85% of free HBM) is reached. So a request that arrives while
 * others are on the GPU is admitted at the next batch a worker forms, batches
 * overlap on the device, and a small request does not wait for the whole of a
 * large one. A request is done when its last chunk is. A query's CIGAR is the
 * same whichever batch carries it. Each worker's traceback pool (35/85 of the
 * budget over the workers) is allocated whole at create, a reserved pool for
 * the DP of every batch, whole-pair fallbacks included; the batches' other
 * buffers are admitted against the remaining 50/85. Thread-safe. */
typedef struct halign_server halign_server;
typedef struct {
    int32_t workers;            /* concurrent batches (1..8), default 2 */
    int32_t max_batch_queries;  /* queries per batch, default 4096 */
    int64_t hbm_budget;         /* bytes, 0 = 85% of free HBM */
    int32_t chunk_queries;      /* queries per chunk of a request, 0 = 256;
                                   at most max_batch_queries is used */
    int32_t pad_;
} halign_server_config;
int halign_server_create(int device, const char *center, int32_t len, const halign_server_config *cfg,
                         halign_server **out);
/* Copies queries[offsets[0] .. offsets[n]) into the server; *request_id names
 * the request in later calls. */
int halign_server_submit(halign_server *s, int32_t n, const char *queries, const int64_t *offsets,
                         int64_t *request_id);
/* HALIGN_OK once the request is done, HALIGN_PENDING if timeout_ms (>= 0)
 * passes first; timeout_ms < 0 waits without a limit. */
int halign_server_wait(halign_server *s, int64_t request_id, int32_t timeout_ms);
/* Results of a done request: CIGARs as halign_center_align, stats (4 per
 * query), timing[3] = {ms from submit to the start of its first batch, ms
 * from submit to done, queries in the largest batch that carried a chunk of
 * it}, n_cigar = total CIGAR words. cigar_out NULL only
 * reports sizes and keeps the request; otherwise the request is released. A
 * failed request returns its error code (and is released). */
int halign_server_result(halign_server *s, int64_t request_id, uint32_t *cigar_out, int64_t cap,
                         int64_t *cigar_offsets, int32_t *stats, double *timing, int64_t *n_cigar);
/* Device memory of a server, estimated and measured (bytes). Measured use is
 * free device memory at halign_server_create minus free memory after a
 * batch, all workers together, the most seen so far. */
typedef struct {
    int64_t batches;               /* batches run so far */
    int64_t max_batch_estimate;    /* largest footprint estimate of one batch */
    int64_t sum_worker_estimates;  /* sum over workers of the largest estimate of a batch it ran */
    int64_t batch_budget;          /* estimate allowed per batch (per worker) */
    int64_t resident;              /* contexts, centers and indexes (measured after create) */
    int64_t traceback_reserved;    /* traceback pools, allocated whole at create */
    int64_t device_used;           /* most measured device use after a batch */
    int64_t working_measured;      /* device_used - resident - traceback_reserved */
} halign_server_memory;
/* No reference counterpart. */
int halign_server_stats(halign_server *s, halign_server_memory *out);
const char *halign_server_last_error(const halign_server *s);
void halign_server_free(halign_server *s);

#ifdef __cplusplus
}
#endif
#endif
