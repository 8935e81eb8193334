// refset_api.hip -- RefAligner's per-query routine against its whole
// reference set (src/align/ref_aligner.cpp:20-73 constructor, :212-283
// alignOneQueryToRef), and alignQueryToRef's per-thread SAM files with one
// thread (:286-394, writeSamRecord :128-136).
//
// With one reference the query is aligned to the consensus of the center
// file, seeded with the record's minimizers. With several, it goes to the
// reference with the largest sketch Jaccard (first maximum), aligned to and
// seeded from that reference; with keep_length that alignment is the result,
// otherwise the query is aligned again to the consensus, seeded with the
// consensus's own minimizers, and that second CIGAR is the result. The SAM
// group is cigar::hasInsertion of the result. Every alignment runs through the
// resident centers of the per-query path (pipeline_api.hip) on the GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <unistd.h>
#include <vector>

#include "fastx.h"
#include "mash_host.h"
#include "pipeline_host.h"

using namespace hal;

namespace {

constexpr int kSketchK = 15, kSketchSize = 2000;   // RefAligner's kmer_size, sketch_size (include/config.hpp)
constexpr uint32_t kSketchSeed = 42;               // RefAligner's random_seed (include/align.h)

bool has_insertion(const uint32_t *c, int64_t n)
{
    for (int64_t k = 0; k < n; ++k)
        if ((c[k] & 0xf) == 1) return true;
    return false;
}

}  // namespace

struct halign_refset {
    halign_ctx *ctx = nullptr;
    int nref = 0;
    std::vector<halign_center *> refs;   // several references: each aligned to and seeded from itself
    halign_center *cons = nullptr;       // the consensus (seeded from the single record, or from itself)
    int32_t cons_len = 0;
    std::vector<int32_t> ref_len;
    DevBuf rsk, rcnt, qsk, qcnt;         // reference and query sketches
};

namespace {

// one center over a subset of the queries: CIGARs into per-query vectors
int align_subset(halign_center *c, int32_t rlen, const char *queries, const int64_t *offsets,
                 const std::vector<int32_t> &idx, std::vector<std::vector<uint32_t>> &out)
{
    if (idx.empty()) return HALIGN_OK;
    std::string blob;
    std::vector<int64_t> off{0};
    int64_t cap = 1;
    for (int32_t i : idx) {
        const int64_t a = offsets[i], b = offsets[i + 1];
        blob.append(queries + a, (size_t)(b - a));
        off.push_back((int64_t)blob.size());
        cap += (b - a) + rlen + 1;
    }
    std::vector<uint32_t> cig((size_t)cap);
    std::vector<int64_t> coff(idx.size() + 1);
    const int rc = halign_center_align(c, (int32_t)idx.size(), blob.data(), off.data(), cig.data(), cap, coff.data(),
                                       nullptr, nullptr);
    if (rc != HALIGN_OK) return rc;
    for (size_t j = 0; j < idx.size(); ++j)
        out[(size_t)idx[j]].assign(cig.begin() + coff[j], cig.begin() + coff[j + 1]);
    return HALIGN_OK;
}

int put_cigars(halign_ctx *ctx, const std::vector<std::vector<uint32_t>> &v, uint32_t *out, int64_t cap, int64_t *off)
{
    int64_t pos = 0;
    off[0] = 0;
    for (size_t i = 0; i < v.size(); ++i) {
        if (pos + (int64_t)v[i].size() > cap) return fail(ctx, HALIGN_E_CAPACITY, "refset: CIGAR output too small");
        std::memcpy(out + pos, v[i].data(), v[i].size() * 4);
        pos += (int64_t)v[i].size();
        off[i + 1] = pos;
    }
    return HALIGN_OK;
}

}  // namespace

extern "C" {

int halign_refset_create(halign_ctx *ctx, int32_t nref, const char *const *refs, const int32_t *ref_lens,
                         const char *consensus, int32_t cons_len, halign_refset **out)
{
    if (!ctx || nref <= 0 || !refs || !ref_lens || !consensus || cons_len <= 0 || !out) return HALIGN_E_ARG;
    for (int r = 0; r < nref; ++r)
        if (!refs[r] || ref_lens[r] < 0) return fail(ctx, HALIGN_E_ARG, "refset: bad reference");
    *out = nullptr;
    auto *rs = new (std::nothrow) halign_refset();
    if (!rs) return HALIGN_E_NOMEM;
    rs->ctx = ctx;
    rs->nref = nref;
    rs->cons_len = cons_len;
    rs->ref_len.assign(ref_lens, ref_lens + nref);
    int rc = HALIGN_OK;
    if (nref == 1) {
        // aligned to the consensus, seeded with the record's minimizers (:40-41, :224-251)
        rc = halign_center_create_indexed(ctx, consensus, cons_len, refs[0], ref_lens[0], &rs->cons);
    } else {
        for (int r = 0; r < nref && rc == HALIGN_OK; ++r) {
            if (ref_lens[r] == 0) { rc = fail(ctx, HALIGN_E_ARG, "refset: a reference sequence is empty"); break; }
            halign_center *c = nullptr;
            rc = halign_center_create(ctx, refs[r], ref_lens[r], &c);
            if (rc == HALIGN_OK) rs->refs.push_back(c);
        }
        // the consensus seeded with its own minimizers (consensus_minimizer, :71-72)
        if (rc == HALIGN_OK) rc = halign_center_create(ctx, consensus, cons_len, &rs->cons);
        if (rc == HALIGN_OK) {
            std::vector<int64_t> off(1, 0);
            std::string blob;
            for (int r = 0; r < nref; ++r) {
                blob.append(refs[r], (size_t)ref_lens[r]);
                off.push_back((int64_t)blob.size());
            }
            std::lock_guard<std::mutex> lock(ctx->mu);
            rc = mash_sketch_to_device(ctx, nref, blob.data(), off.data(), kSketchK, kSketchSize, true, kSketchSeed,
                                       rs->rsk, rs->rcnt);
        }
    }
    if (rc != HALIGN_OK) {
        halign_refset_free(rs);
        return rc;
    }
    *out = rs;
    return HALIGN_OK;
}

void halign_refset_free(halign_refset *rs)
{
    if (!rs) return;
    for (auto *c : rs->refs) halign_center_free(c);
    if (rs->cons) halign_center_free(rs->cons);
    {
        std::lock_guard<std::mutex> lock(rs->ctx->mu);
        rs->rsk.release(); rs->rcnt.release(); rs->qsk.release(); rs->qcnt.release();
    }
    delete rs;
}

int halign_refset_align(halign_refset *rs, int32_t n, const char *queries, const int64_t *offsets,
                        int32_t keep_length, int32_t *best_ref, uint32_t *init_cigar, int64_t init_cap,
                        int64_t *init_off, uint32_t *final_cigar, int64_t final_cap, int64_t *final_off,
                        uint8_t *has_ins, int32_t *final_ref)
{
    if (!rs || n < 0 || (n > 0 && (!queries || !offsets)) || !init_off || !final_off || !init_cigar || !final_cigar)
        return HALIGN_E_ARG;
    halign_ctx *ctx = rs->ctx;
    std::vector<std::vector<uint32_t>> init((size_t)n), fin((size_t)n);
    std::vector<int32_t> best((size_t)n, 0);
    int rc = HALIGN_OK;
    if (n > 0 && rs->nref == 1) {
        std::vector<int32_t> all((size_t)n);
        for (int32_t i = 0; i < n; ++i) all[(size_t)i] = i;
        rc = align_subset(rs->cons, rs->cons_len, queries, offsets, all, init);
        if (rc != HALIGN_OK) return rc;
        fin = init;
    } else if (n > 0) {
        // the reference with the largest Jaccard (:228-244)
        std::vector<int64_t> off(offsets, offsets + n + 1);
        for (auto &o : off) o -= offsets[0];
        std::vector<int32_t> num, den;
        {
            std::lock_guard<std::mutex> lock(ctx->mu);
            rc = mash_sketch_to_device(ctx, n, queries + offsets[0], off.data(), kSketchK, kSketchSize, true,
                                       kSketchSeed, rs->qsk, rs->qcnt);
            if (rc == HALIGN_OK)
                rc = mash_select_device(ctx, n, rs->qsk, rs->qcnt, kSketchSize, rs->nref, rs->rsk, rs->rcnt,
                                        kSketchSize, best, num, den);
        }
        if (rc != HALIGN_OK) return rc;
        std::vector<std::vector<int32_t>> members((size_t)rs->nref);
        for (int32_t i = 0; i < n; ++i) members[(size_t)best[(size_t)i]].push_back(i);
        for (int r = 0; r < rs->nref && rc == HALIGN_OK; ++r)
            rc = align_subset(rs->refs[(size_t)r], rs->ref_len[(size_t)r], queries, offsets, members[(size_t)r], init);
        if (rc != HALIGN_OK) return rc;
        if (keep_length) {
            fin = init;
        } else {
            // the second alignment, to the consensus (:268-282)
            std::vector<int32_t> all((size_t)n);
            for (int32_t i = 0; i < n; ++i) all[(size_t)i] = i;
            rc = align_subset(rs->cons, rs->cons_len, queries, offsets, all, fin);
            if (rc != HALIGN_OK) return rc;
        }
    }
    rc = put_cigars(ctx, init, init_cigar, init_cap, init_off);
    if (rc == HALIGN_OK) rc = put_cigars(ctx, fin, final_cigar, final_cap, final_off);
    if (rc != HALIGN_OK) return rc;
    for (int32_t i = 0; i < n; ++i) {
        if (best_ref) best_ref[i] = best[(size_t)i];
        if (has_ins) has_ins[i] = has_insertion(fin[(size_t)i].data(), (int64_t)fin[(size_t)i].size()) ? 1 : 0;
        if (final_ref) final_ref[i] = (rs->nref > 1 && keep_length) ? best[(size_t)i] : -1;
    }
    return HALIGN_OK;
}

}  // extern "C"

namespace {

// cigar::cigarToString (src/align/cigar.cpp:82-97): "" for an empty CIGAR
void append_cigar_string(std::string &o, const uint32_t *c, int64_t n)
{
    static const char kOps[] = "MIDNSHP=X";
    char buf[16];
    for (int64_t k = 0; k < n; ++k) {
        const int len = snprintf(buf, sizeof buf, "%u", c[k] >> 4);
        o.append(buf, (size_t)len);
        const uint32_t op = c[k] & 0xf;
        o += op < 9 ? kOps[op] : '?';
    }
}

// seq_io::SeqWriter::writeSam of makeSamRecord(q, rname, cigar, 1, 60, 0)
// (src/utils/seq_io.cpp:289-360): the reader never fills SeqRecord::qual,
// so QUAL keeps SamRecord's "*"
void append_sam_line(std::string &o, const char *qname, size_t qn, const std::string &rname, const uint32_t *c,
                     int64_t nc, const char *seq, size_t sn)
{
    o.append(qname, qn);
    o += "\t0\t";
    o += rname;
    o += "\t1\t60\t";
    append_cigar_string(o, c, nc);
    o += "\t*\t0\t0\t";
    o.append(seq, sn);
    o += "\t*\n";
}

bool write_all(FILE *f, const std::string &s) { return s.empty() || fwrite(s.data(), 1, s.size(), f) == s.size(); }

}  // namespace

extern "C" int halign_align_sam(halign_ctx *ctx, const char *center_path, const char *aligned_path,
                                const char *input_path, const char *sam_path, const char *insertion_sam_path,
                                int32_t keep_length, int32_t batch, halign_msa_stats *stats)
{
    if (!ctx || !center_path || !input_path || !sam_path || !insertion_sam_path) return HALIGN_E_ARG;
    const auto t0 = std::chrono::steady_clock::now();
    if (batch <= 0) batch = 25600;
    // the center file's records (RefAligner's constructor, :35-44)
    RecordChunk refs;
    {
        FastxReader rd;
        std::string e, id, desc, seq;
        if (!rd.open(center_path, e)) return fail(ctx, HALIGN_E_IO, e.c_str());
        int r;
        while ((r = rd.next(id, desc, seq, e)) > 0) refs.add(id, desc, seq);
        if (r < 0) return fail(ctx, HALIGN_E_IO, e.c_str());
    }
    if (refs.n == 0) return fail(ctx, HALIGN_E_ARG, "center file has no record");
    // the consensus of the MSA command's output for the center file (:47-60;
    // aligned_path NULL: the copy command, the center file itself)
    std::string cons;
    {
        const std::string fa = std::string(sam_path) + ".cons.fa.tmp", js = std::string(sam_path) + ".cons.json.tmp";
        const char *apath = aligned_path ? aligned_path : center_path;
        FILE *probe = fopen(apath, "rb");
        if (!probe) return fail(ctx, HALIGN_E_IO, (std::string("cannot open ") + apath).c_str());
        fseek(probe, 0, SEEK_END);
        const long sz = ftell(probe);
        fclose(probe);
        std::vector<char> buf((size_t)std::max<long>(sz, 1) + 1);
        int64_t clen = 0;
        const int rc = halign_consensus_file(ctx, apath, fa.c_str(), js.c_str(), 0, 4096, buf.data(),
                                             (int64_t)buf.size(), &clen);
        unlink(fa.c_str());
        unlink(js.c_str());
        if (rc != HALIGN_OK) return rc;
        cons.assign(buf.data(), (size_t)clen);
    }
    std::vector<const char *> rptr((size_t)refs.n);
    std::vector<int32_t> rlen((size_t)refs.n);
    std::vector<std::string> rid((size_t)refs.n);
    for (int r = 0; r < refs.n; ++r) {
        rptr[(size_t)r] = refs.seqs.data() + refs.seq_off[r];
        rlen[(size_t)r] = (int32_t)(refs.seq_off[r + 1] - refs.seq_off[r]);
        rid[(size_t)r].assign(refs.ids, (size_t)refs.id_off[r], (size_t)(refs.id_off[r + 1] - refs.id_off[r]));
    }
    halign_refset *rs = nullptr;
    int rc = halign_refset_create(ctx, refs.n, rptr.data(), rlen.data(), cons.data(), (int32_t)cons.size(), &rs);
    if (rc != HALIGN_OK) return rc;
    struct Guard {
        halign_refset *rs;
        FILE *a = nullptr, *b = nullptr;
        ~Guard()
        {
            if (a) fclose(a);
            if (b) fclose(b);
            halign_refset_free(rs);
        }
    } g{rs};
    g.a = fopen(sam_path, "wb");
    g.b = fopen(insertion_sam_path, "wb");
    if (!g.a || !g.b) return fail(ctx, HALIGN_E_IO, "cannot create the SAM files");
    const std::string header = "@HD\tVN:1.6\tSO:unknown\n";   // writeSamHeader in alignQueryToRef (:318-322)
    if (!write_all(g.a, header) || !write_all(g.b, header)) return fail(ctx, HALIGN_E_IO, "SAM write failed");
    FastxReader rd;
    std::string e, id, desc, seq;
    if (!rd.open(input_path, e)) return fail(ctx, HALIGN_E_IO, e.c_str());
    halign_msa_stats st{};
    double gpu_s = 0;
    const std::string consensus_name = "consensus";   // consensus_seq.id (:62)
    for (;;) {
        RecordChunk ch;
        int r = 1;
        while (ch.n < batch && (r = rd.next(id, desc, seq, e)) > 0) ch.add(id, desc, seq);
        if (r < 0) return fail(ctx, HALIGN_E_IO, e.c_str());
        if (ch.n == 0) break;
        int64_t cap = 1;
        int32_t maxr = (int32_t)cons.size();
        for (int32_t x : rlen) maxr = std::max(maxr, x);
        for (int32_t i = 0; i < ch.n; ++i) cap += ch.seq_off[i + 1] - ch.seq_off[i] + maxr + 1;
        std::vector<uint32_t> ic((size_t)cap), fc((size_t)cap);
        std::vector<int64_t> io((size_t)ch.n + 1), fo((size_t)ch.n + 1);
        std::vector<uint8_t> hi((size_t)ch.n);
        std::vector<int32_t> fr((size_t)ch.n);
        const auto g0 = std::chrono::steady_clock::now();
        rc = halign_refset_align(rs, ch.n, ch.seqs.data(), ch.seq_off.data(), keep_length, nullptr, ic.data(), cap,
                                 io.data(), fc.data(), cap, fo.data(), hi.data(), fr.data());
        gpu_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - g0).count();
        if (rc != HALIGN_OK) return rc;
        std::string plain, ins;
        for (int32_t i = 0; i < ch.n; ++i) {
            std::string &o = hi[(size_t)i] ? ins : plain;
            append_sam_line(o, ch.ids.data() + ch.id_off[i], (size_t)(ch.id_off[i + 1] - ch.id_off[i]),
                            fr[(size_t)i] < 0 ? consensus_name : rid[(size_t)fr[(size_t)i]], fc.data() + fo[i],
                            fo[i + 1] - fo[i], ch.seqs.data() + ch.seq_off[i],
                            (size_t)(ch.seq_off[i + 1] - ch.seq_off[i]));
            st.records += 1;
            st.insertion_records += hi[(size_t)i] ? 1 : 0;
        }
        st.bytes += (int64_t)(plain.size() + ins.size());
        if (!write_all(g.a, plain) || !write_all(g.b, ins)) return fail(ctx, HALIGN_E_IO, "SAM write failed");
    }
    if (fflush(g.a) != 0 || fflush(g.b) != 0) return fail(ctx, HALIGN_E_IO, "SAM write failed");
    st.gpu_seconds = gpu_s;
    st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (stats) *stats = st;
    return HALIGN_OK;
}
