// ref_harness.cpp -- C-ABI glue over the REAL reference functions, compiled
// together with the reference's own sources by oracle/Makefile into
// oracle/_ref/libhalign_ref.so (this container only; the reference never
// travels). Used to pin the oracle restatement and to write golden fixtures.
// Signatures mirror oracle/halign_oracle.h so tests can compare the two.
#include "align.h"
#include "anchor.h"
#include "consensus.h"
#include "ksw2.h"
#include "seed.h"
#include "utils.h"
#include "config.hpp"
#include "mash.h"
#include "preprocess.h"

#include <cstring>
#include <filesystem>
#include <string>
#include <vector>

namespace {
struct RefEz {
    uint32_t max;
    int32_t zdropped;
    int32_t max_q, max_t, mqe, mqe_t, mte, mte_q, score, n_cigar;
};
struct RefAnchor {
    uint64_t hash;
    uint32_t rid_ref, pos_ref, rid_qry, pos_qry, span, is_rev;
};
anchor::Anchor to_anchor(const RefAnchor& a)
{
    anchor::Anchor r;
    r.hash = a.hash; r.rid_ref = a.rid_ref; r.pos_ref = a.pos_ref;
    r.rid_qry = a.rid_qry; r.pos_qry = a.pos_qry; r.span = a.span; r.is_rev = a.is_rev != 0;
    return r;
}
RefAnchor from_anchor(const anchor::Anchor& a)
{
    RefAnchor r;
    r.hash = a.hash; r.rid_ref = a.rid_ref; r.pos_ref = a.pos_ref;
    r.rid_qry = a.rid_qry; r.pos_qry = a.pos_qry; r.span = a.span; r.is_rev = a.is_rev ? 1u : 0u;
    return r;
}
int copy_cigar(const cigar::Cigar_t& c, uint32_t* out, int cap)
{
    if ((int)c.size() > cap) return -1;
    for (size_t i = 0; i < c.size(); ++i) out[i] = c[i];
    return (int)c.size();
}
minimizer::MinimizerHits to_hits(const uint64_t* x, const uint64_t* y, int n)
{
    minimizer::MinimizerHits h;
    h.reserve((size_t)n);
    for (int i = 0; i < n; ++i) h.emplace_back(x[i], y[i]);
    return h;
}
}  // namespace

extern "C" {

int ref_ksw_extz2(int qlen, const uint8_t* query, int tlen, const uint8_t* target, int w,
                  int flag, RefEz* out, uint32_t* cigar, int cap)
{
    ksw_extz_t ez{};
    ksw_extz2_sse(0, qlen, query, tlen, target, 5, align::dna5_simd_mat, 6, 2, w, -1, 0, flag, &ez);
    out->max = ez.max; out->zdropped = ez.zdropped;
    out->max_q = ez.max_q; out->max_t = ez.max_t; out->mqe = ez.mqe; out->mqe_t = ez.mqe_t;
    out->mte = ez.mte; out->mte_q = ez.mte_q; out->score = ez.score; out->n_cigar = ez.n_cigar;
    int ret = ez.n_cigar;
    if (ez.n_cigar > cap) ret = -1;
    else for (int i = 0; i < ez.n_cigar; ++i) cigar[i] = ez.cigar[i];
    free(ez.cigar);
    return ret;
}

int ref_global_align_ksw2(const char* ref, int rlen, const char* qry, int qlen, int flag,
                          uint32_t* cigar, int cap)
{
    align::KSW2AlignConfig cfg;
    cfg.flag = flag;
    return copy_cigar(align::globalAlignKSW2(std::string(ref, (size_t)rlen),
                                             std::string(qry, (size_t)qlen), cfg),
                      cigar, cap);
}

int ref_extract_minimizer(const char* seq, int n, int k, int w, int non_canonical,
                          uint64_t* x, uint64_t* y, int cap)
{
    auto h = minimizer::extractMinimizer(std::string(seq, (size_t)n), (size_t)k, (size_t)w,
                                         non_canonical != 0);
    if ((int)h.size() > cap) return -1;
    for (size_t i = 0; i < h.size(); ++i) x[i] = h[i].x, y[i] = h[i].y;
    return (int)h.size();
}

int ref_collect_anchors(const uint64_t* rx, const uint64_t* ry, int nref, const uint64_t* qx,
                        const uint64_t* qy, int nqry, RefAnchor* out, int cap)
{
    auto a = minimizer::collect_anchors(to_hits(rx, ry, nref), to_hits(qx, qy, nqry));
    if ((int)a.size() > cap) return -1;
    for (size_t i = 0; i < a.size(); ++i) out[i] = from_anchor(a[i]);
    return (int)a.size();
}

int ref_chain_anchors(RefAnchor* anchors, int n, RefAnchor* out)
{
    anchor::Anchors a;
    for (int i = 0; i < n; ++i) a.push_back(to_anchor(anchors[i]));
    auto c = anchor::chainAnchors(a);
    for (int i = 0; i < n; ++i) anchors[i] = from_anchor(a[(size_t)i]);
    for (size_t i = 0; i < c.size(); ++i) out[i] = from_anchor(c[i]);
    return (int)c.size();
}

int ref_global_align_mm2(const char* ref, int rlen, const char* qry, int qlen,
                         const RefAnchor* anchors, int n, uint32_t* cigar, int cap)
{
    anchor::Anchors a;
    for (int i = 0; i < n; ++i) a.push_back(to_anchor(anchors[i]));
    return copy_cigar(align::globalAlignMM2(std::string(ref, (size_t)rlen),
                                            std::string(qry, (size_t)qlen), a),
                      cigar, cap);
}

// Body of RefAligner::globalAlign (src/align/ref_aligner.cpp:91-125) with the
// production parameters k=15, w=10, non_canonical=true.
int ref_ref_global_align(const char* ref, int rlen, const char* qry, int qlen, uint32_t* cigar,
                         int cap)
{
    const std::string r(ref, (size_t)rlen), q(qry, (size_t)qlen);
    auto rm = minimizer::extractMinimizer(r, 15, 10, true);
    auto qm = minimizer::extractMinimizer(q, 15, 10, true);
    auto anchors = minimizer::collect_anchors(rm, qm);
    return copy_cigar(align::globalAlignMM2(r, q, anchors), cigar, cap);
}

// The same with the center's minimizers computed once by the caller, as
// RefAligner does for its centers (src/align/ref_aligner.cpp:40-41,224-225).
int ref_ref_global_align_pre(const char* ref, int rlen, const uint64_t* rx, const uint64_t* ry, int nref,
                             const char* qry, int qlen, uint32_t* cigar, int cap)
{
    const std::string r(ref, (size_t)rlen), q(qry, (size_t)qlen);
    minimizer::MinimizerHits rm;
    rm.reserve((size_t)nref);
    for (int i = 0; i < nref; ++i) rm.emplace_back(rx[i], ry[i]);
    auto qm = minimizer::extractMinimizer(q, 15, 10, true);
    auto anchors = minimizer::collect_anchors(rm, qm);
    return copy_cigar(align::globalAlignMM2(r, q, anchors), cigar, cap);
}

// consensus::generateConsensusSequence on a FASTA file; the per-column counts
// land in out_json (cereal JSON) for the caller to read.
int ref_consensus(const char* aligned_fasta, const char* out_fasta, const char* out_json,
                  uint64_t seq_limit, int threads, int batch, char* cons_out, int cap)
{
    std::string c = consensus::generateConsensusSequence(aligned_fasta, out_fasta, out_json,
                                                         seq_limit, threads, (size_t)batch);
    if ((int)c.size() > cap) return -1;
    std::memcpy(cons_out, c.data(), c.size());
    return (int)c.size();
}

// seq_io::KseqReader over a file: records' ids, comments and sequences back
// to back with n+1 offsets each (capacity cap bytes per output). Returns the
// record count, -1 when an output is too small, -2 when the reader throws.
int ref_read_fastx(const char* path, int max_records, char* id_out, int64_t* id_off, char* desc_out,
                   int64_t* desc_off, char* seq_out, int64_t* seq_off, int64_t cap)
{
    try {
        seq_io::KseqReader rd(path);
        seq_io::SeqRecord rec;
        int n = 0;
        id_off[0] = desc_off[0] = seq_off[0] = 0;
        while (n < max_records && rd.next(rec)) {
            const std::string* f[3] = {&rec.id, &rec.desc, &rec.seq};
            char* o[3] = {id_out, desc_out, seq_out};
            int64_t* off[3] = {id_off, desc_off, seq_off};
            for (int k = 0; k < 3; ++k) {
                if (off[k][n] + (int64_t)f[k]->size() > cap) return -1;
                std::memcpy(o[k] + off[k][n], f[k]->data(), f[k]->size());
                off[k][n + 1] = off[k][n] + (int64_t)f[k]->size();
            }
            ++n;
        }
        return n;
    } catch (...) {
        return -2;
    }
}

// mash::sketchFromSequence (src/mash/mash.cpp:32-146): the sorted hashes.
// Returns their count, or -1 when cap is too small.
int ref_mash_sketch(const char* seq, int n, int k, int sketch_size, int noncanonical, int seed, uint64_t* out,
                    int cap)
{
    const mash::Sketch sk = mash::sketchFromSequence(std::string(seq, (size_t)n), (size_t)k, (size_t)sketch_size,
                                                     noncanonical != 0, seed);
    if ((int)sk.hashes.size() > cap) return -1;
    for (size_t i = 0; i < sk.hashes.size(); ++i) out[i] = sk.hashes[i];
    return (int)sk.hashes.size();
}

// getHash2bit (src/mash/hash.cpp): XXH3-64 of the big-endian 8-byte code
uint64_t ref_hash2bit(uint64_t code, uint32_t seed) { return getHash2bit(code, seed); }

// mash::jaccard (src/mash/mash.cpp:151-164) of two sorted unique hash lists
double ref_mash_jaccard(const uint64_t* a, int na, const uint64_t* b, int nb, int k)
{
    mash::Sketch x, y;
    x.k = y.k = (size_t)k;
    x.hashes.assign(a, a + na);
    y.hashes.assign(b, b + nb);
    return mash::jaccard(x, y);
}

// preprocessInputFasta (src/utils/preprocess.cpp:19-160) into work_dir: the
// Top-K consensus input lands in work_dir/data/clean_data/consensus_unaligned.fasta.
// Returns the records read, or -2 when the reference throws.
int64_t ref_preprocess(const char* input_path, const char* work_dir, int cons_n)
{
    try {
        return (int64_t)preprocessInputFasta(input_path, work_dir, cons_n);
    } catch (...) {
        return -2;
    }
}

}  // extern "C"
