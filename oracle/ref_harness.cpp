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

// halign4 -c center_path --keep-length with the copy MSA command and one
// thread, as the reference runs it (src/halign4.cpp:96-147): the RefAligner
// constructor (src/align/ref_aligner.cpp:20-70: copy the center file,
// consensus of the copy = the center aligned to; the minimizers of the file's
// record), alignQueryToRef (:286-394) with alignOneQueryToRef's routing
// (:212-283) into thread 0's two SAM files, then mergeAlignedResults
// (:676-740) with processInsertionSequences / mergeConsensusAndSamToFasta
// (:504-535, :139-209), parseAlignedReferencesToCigar (:400-459),
// writeConsensusAndReferences (:541-561), writeInsertionSequences
// (:567-605) and processSamFileBatch / convertSamToFastaRecord (:467-498,
// :612-672). RefAligner itself is not compiled here (it needs mash and the
// external-tool runner); this glue calls the same reference functions in the
// same order. work_dir must exist. Returns 0, or -2 when the reference throws.
int ref_keep_length_msa_refs(const char* center_path, const char* aligned_path, const char* input_path,
                             const char* work_dir, const char* out_path)
{
    namespace fs = std::filesystem;
    try {
        const fs::path wd(work_dir);
        // constructor: the center file's records and minimizers
        std::vector<seq_io::SeqRecord> refs;
        std::vector<minimizer::MinimizerHits> ref_mz;
        {
            seq_io::KseqReader rd(center_path);
            seq_io::SeqRecord rec;
            while (rd.next(rec)) {
                ref_mz.push_back(minimizer::extractMinimizer(rec.seq, 15, 10, true));
                refs.push_back(rec);
            }
        }
        if (refs.empty()) return -3;
        std::vector<mash::Sketch> ref_sk;
        for (const auto& r : refs) ref_sk.push_back(mash::sketchFromSequence(r.seq, 15, 2000, true, 42));
        const fs::path cons_aligned = wd / "cons_aligned.fasta";
        // the MSA command's output: a copy ("cp {input} {output}") or the given aligned file
        file_io::copyFile(aligned_path ? aligned_path : center_path, cons_aligned);
        seq_io::SeqRecord consensus_seq;
        consensus_seq.id = "consensus";
        consensus_seq.seq = consensus::generateConsensusSequence(cons_aligned, wd / "cons.fasta", wd / "cons.json",
                                                                 0, 1, 4096);
        // alignQueryToRef, one thread
        const fs::path sam = wd / "t0.sam", sam_ins = wd / "t0_insertion.sam";
        {
            auto out = seq_io::SeqWriter::Sam(sam);
            auto out_ins = seq_io::SeqWriter::Sam(sam_ins);
            out.writeSamHeader("@HD\tVN:1.6\tSO:unknown");
            out_ins.writeSamHeader("@HD\tVN:1.6\tSO:unknown");
            seq_io::KseqReader rd(input_path);
            seq_io::SeqRecord q;
            while (rd.next(q)) {
                // alignOneQueryToRef (:212-283) with keep_length: the reference sequence
                // with the largest Jaccard (first strict maximum), or the consensus
                // when there is one center
                const mash::Sketch qsk = mash::sketchFromSequence(q.seq, 15, 2000, true, 42);
                auto qm = minimizer::extractMinimizer(q.seq, 15, 10, true);
                size_t best = 0;
                const seq_io::SeqRecord* best_ref = &consensus_seq;
                if (refs.size() > 1) {
                    double bj = -1.0;
                    for (size_t r = 0; r < ref_sk.size(); ++r) {
                        const double j = mash::jaccard(qsk, ref_sk[r]);
                        if (j > bj) {
                            bj = j;
                            best = r;
                        }
                    }
                    best_ref = &refs[best];
                }
                const minimizer::MinimizerHits* rm = &ref_mz[best];
                minimizer::MinimizerHits rtmp;
                if (rm->empty()) {
                    rtmp = minimizer::extractMinimizer(best_ref->seq, 15, 10, true);
                    rm = &rtmp;
                }
                const auto anchors = minimizer::collect_anchors(*rm, qm);
                const cigar::Cigar_t cg = align::globalAlignMM2(best_ref->seq, q.seq, anchors);
                const auto rec = seq_io::makeSamRecord(q, best_ref->id, cigar::cigarToString(cg), 1, 60, 0);
                if (cigar::hasInsertion(cg)) out_ins.writeSam(rec);
                else out.writeSam(rec);
            }
            out.flush();
            out_ins.flush();
        }
        // mergeAlignedResults
        auto parse_aligned = [](const fs::path& p, std::unordered_map<std::string, cigar::Cigar_t>& m) {
            seq_io::KseqReader rd(p);
            seq_io::SeqRecord rec;
            while (rd.next(rec)) {
                cigar::Cigar_t c;
                char cur = 0;
                uint32_t len = 0;
                for (char b : rec.seq) {
                    const char op = b == '-' ? 'D' : 'M';
                    if (op == cur) ++len;
                    else {
                        if (cur && len) c.push_back(cigar::cigarToInt(cur, len));
                        cur = op;
                        len = 1;
                    }
                }
                if (cur && len) c.push_back(cigar::cigarToInt(cur, len));
                m[rec.id] = c;
            }
        };
        std::unordered_map<std::string, cigar::Cigar_t> ref_map, ins_map;
        parse_aligned(cons_aligned, ref_map);
        seq_io::SeqWriter final_writer(out_path, U_MAX);
        // processInsertionSequences -> mergeConsensusAndSamToFasta(keep = true, width 80), then copy
        const fs::path ins_fa = wd / "all_insertion.fasta", ins_aligned = wd / "aligned_insertion.fasta";
        {
            seq_io::SeqWriter w(ins_fa, 80);
            w.writeFasta(consensus_seq);
            seq_io::SamReader r(sam_ins);
            seq_io::SamRecord s;
            while (r.next(s)) {
                seq_io::SeqRecord f = seq_io::samRecordToSeqRecord(s, false);
                if (!s.cigar.empty()) {
                    cigar::Cigar_t ops = cigar::stringToCigar(s.cigar);
                    cigar::delQueryToRefByCigar(f.seq, ops);
                    if (s.rname != "consensus") cigar::padQueryToRefByCigar(f.seq, ref_map.at(s.rname));
                }
                w.writeFasta(f);
            }
            w.flush();
        }
        file_io::copyFile(ins_fa, ins_aligned);
        parse_aligned(ins_aligned, ins_map);
        ref_map[consensus_seq.id] = ins_map[consensus_seq.id];
        // writeConsensusAndReferences
        {
            seq_io::KseqReader rd(cons_aligned);
            seq_io::SeqRecord rec;
            while (rd.next(rec)) {
                seq_io::cleanSequence(rec);
                final_writer.writeFasta(rec);
            }
        }
        // writeInsertionSequences
        size_t expect = 0;
        bool have = false;
        {
            seq_io::KseqReader rd(ins_aligned);
            seq_io::SeqRecord rec;
            bool first = true;
            while (rd.next(rec)) {
                if (first) { first = false; continue; }
                if (!have) { expect = rec.seq.size(); have = true; }
                else if (rec.seq.size() != expect) return -4;
                final_writer.writeFasta(rec);
            }
        }
        // processSamFileBatch over thread 0's SAM file
        {
            seq_io::SamReader r(sam);
            seq_io::SamRecord s;
            while (r.next(s)) {
                seq_io::SeqRecord f = seq_io::samRecordToSeqRecord(s, false);
                if (!s.cigar.empty() && s.cigar != "*") cigar::padQueryToRefByCigar(f.seq, cigar::stringToCigar(s.cigar));
                cigar::padQueryToRefByCigar(f.seq, ref_map.at(s.rname));
                if (!have) { expect = f.seq.size(); have = true; }
                else if (f.seq.size() != expect) return -4;
                final_writer.writeFasta(f);
            }
        }
        final_writer.flush();
        return 0;
    } catch (...) {
        return -2;
    }
}

int ref_keep_length_msa(const char* center_path, const char* input_path, const char* work_dir,
                        const char* out_path)
{
    return ref_keep_length_msa_refs(center_path, nullptr, input_path, work_dir, out_path);
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

// halign4 --keep-length without -c (src/halign4.cpp:87, 119-147) with the
// copy MSA command (or the given aligned Top-K file) and one thread:
// preprocessInputFasta, the consensus of the aligned Top-K with seq_limit
// cons_n, then RefAligner on the consensus file as ref_keep_length_msa_refs
// runs it. Returns 0, or -2 when the reference throws.
int ref_auto_center_keep_length(const char* input_path, const char* aligned_topk, const char* work_dir,
                                const char* out_path, int cons_n)
{
    namespace fs = std::filesystem;
    try {
        const fs::path wd(work_dir);
        preprocessInputFasta(input_path, work_dir, cons_n);
        const fs::path clean = wd / WORKDIR_DATA / DATA_CLEAN;
        const fs::path unaligned = clean / CLEAN_CONS_UNALIGNED, aligned = clean / CLEAN_CONS_ALIGNED;
        file_io::copyFile(aligned_topk ? fs::path(aligned_topk) : unaligned, aligned);
        consensus::generateConsensusSequence(aligned, clean / CLEAN_CONS_FASTA, clean / CLEAN_CONS_JSON,
                                             (uint64_t)cons_n, 1, 4096);
        const fs::path wd2 = wd / "ref_aligner";
        fs::create_directories(wd2);
        const std::string cons_fa = (clean / CLEAN_CONS_FASTA).string();
        return ref_keep_length_msa_refs(cons_fa.c_str(), nullptr, input_path, wd2.string().c_str(), out_path);
    } catch (...) {
        return -2;
    }
}

}  // extern "C"
