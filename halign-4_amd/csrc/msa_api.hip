// msa_api.hip -- halign_msa_keep_length: the keep-length run of halign4 with
// one center, file to file (reference src/halign4.cpp:96-147 with -c and
// --keep-length; RefAligner::RefAligner src/align/ref_aligner.cpp:20-70,
// alignQueryToRef :286-394, mergeAlignedResults :676-740).
//
// Three stages overlap: a reader thread parses the next chunk of queries
// (FastxReader), the calling thread runs a chunk through the GPU
// (halign_center_align_fasta: seeding to CIGARs to keep-length records, all
// on the device, the records landing in pinned host memory), and a writer
// thread appends the previous chunk's records to the output. The reference
// writes the rows of queries with an insertion before all others (it merges
// the insertion SAM files first); the writer puts those straight into the
// output and spools the others to a temporary file next to it, appended at
// the end. Writes are positioned (pwrite) and split over several threads, so
// that copying into the page cache is not one core's memcpy.
#include <hip/hip_runtime.h>

#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "fastx.h"
#include "host_common.h"
#include "mash_host.h"
#include "pipeline_host.h"

using namespace hal;

namespace {

using Clock = std::chrono::steady_clock;

double secs(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double>(b - a).count(); }

// The references of a run. One center: RefAligner aligns to the consensus of
// the (aligned) center file and seeds with the record's minimizers. Several:
// each query goes to the reference with the largest sketch Jaccard, aligned
// to that reference's own sequence, its row spread over the columns of the
// reference's aligned record (all records one width).
struct Refs {
    bool multi = false;
    halign_center *single = nullptr;         // kept by the context
    std::vector<halign_center *> centers;    // several references, owned
    DevBuf rsk, rcnt, qsk, qcnt;             // reference and query sketches
    int64_t width = 0;                       // row length of every record
    ~Refs()
    {
        for (auto *c : centers) halign_center_free(c);
        rsk.release();
        rcnt.release();
        qsk.release();
        qcnt.release();
    }
};

constexpr int kSketchK = 15, kSketchSize = 2000;   // RefAligner's kmer_size, sketch_size (include/config.hpp)
constexpr uint32_t kSketchSeed = 42;               // RefAligner's random_seed (include/align.h)

bool dropped_name(const RecordChunk &ch, int32_t i)
{
    return ch.id_off[i + 1] > ch.id_off[i] && ch.ids[(size_t)ch.id_off[i]] == '@';
}

// One chunk's records ready for the writer: out.p[0 .. ins) first group,
// out.p[ins .. len) the rest.
struct Slot {
    HostPinned *out = nullptr;
    int64_t len = 0, ins = 0;
    bool full = false;
};

struct Shared {
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::unique_ptr<RecordChunk>> ready;   // parsed chunks
    bool reader_done = false;
    std::string err;
    int err_code = HALIGN_OK;
    bool stop = false;
    void set_error(int code, const std::string &e)
    {
        std::lock_guard<std::mutex> lk(mu);
        if (err_code == HALIGN_OK) {
            err_code = code;
            err = e;
        }
        stop = true;
        cv.notify_all();
    }
};

// n bytes at file offset off, in pieces of at least 4 MiB on up to 8 threads
bool write_at(int fd, const char *p, size_t n, int64_t off)
{
    auto one = [fd](const char *q, size_t m, int64_t o) {
        while (m > 0) {
            const ssize_t w = pwrite(fd, q, m, (off_t)o);
            if (w < 0 && errno == EINTR) continue;
            if (w <= 0) return false;
            q += w;
            m -= (size_t)w;
            o += w;
        }
        return true;
    };
    constexpr size_t kPiece = 4 << 20;
    const int parts = (int)std::min<size_t>(8, (n + kPiece - 1) / kPiece);
    if (parts <= 1) return one(p, n, off);
    const size_t step = (n + parts - 1) / parts;
    std::vector<std::thread> th;
    std::vector<char> ok((size_t)parts, 0);
    for (int k = 0; k < parts; ++k) {
        const size_t a = (size_t)k * step, b = std::min(n, a + step);
        th.emplace_back([&, k, a, b] { ok[(size_t)k] = one(p + a, b - a, off + (int64_t)a); });
    }
    for (auto &t : th) t.join();
    for (char v : ok)
        if (!v) return false;
    return true;
}

// bytes [0, n) of file `src` to file `dst` at offset off
bool append_file(int src, int dst, int64_t n, int64_t off)
{
    loff_t si = 0, di = (loff_t)off;
    while (n > 0) {
        const ssize_t c = copy_file_range(src, &si, dst, &di, (size_t)n, 0);
        if (c > 0) {
            n -= c;
            continue;
        }
        if (c < 0 && errno == EINTR) continue;
        break;   // not supported between these files: copy through memory
    }
    std::vector<char> buf(8 << 20);
    while (n > 0) {
        const ssize_t r = pread(src, buf.data(), (size_t)std::min<int64_t>(n, (int64_t)buf.size()), (off_t)si);
        if (r <= 0) return false;
        if (!write_at(dst, buf.data(), (size_t)r, di)) return false;
        si += r;
        di += r;
        n -= r;
    }
    return true;
}

}  // namespace

namespace {

// a chunk through the GPU into slot memory: records of kept queries,
// insertion group first, each group in input order
int align_chunk(halign_ctx *ctx, Refs &R, const RecordChunk &ch, HostPinned &out, int64_t &len, int64_t &ins,
                int32_t &n_ins, int32_t &kept)
{
    int64_t need = 0;
    kept = 0;
    for (int32_t i = 0; i < ch.n; ++i) {
        if (dropped_name(ch, i)) continue;   // no record (its SAM line is skipped)
        need += ch.id_off[i + 1] - ch.id_off[i] + R.width + 3;
        ++kept;
    }
    if (!out.ensure((size_t)std::max<int64_t>(need, 1))) return fail(ctx, HALIGN_E_NOMEM, "pinned host buffer");
    if (!R.multi)
        return center_align_fasta(R.single, ch.n, ch.seqs.data(), ch.seq_off.data(), ch.ids.data(),
                                  ch.id_off.data(), out.p, need, &len, &ins, &n_ins, nullptr, nullptr);
    // several references: the choice per query (alignOneQueryToRef :228-244)
    std::vector<int32_t> best, num, den;
    {
        std::lock_guard<std::mutex> lock(ctx->mu);
        int rc = mash_sketch_to_device(ctx, ch.n, ch.seqs.data(), ch.seq_off.data(), kSketchK, kSketchSize, true,
                                       kSketchSeed, R.qsk, R.qcnt);
        if (rc == HALIGN_OK)
            rc = mash_select_device(ctx, ch.n, R.qsk, R.qcnt, kSketchSize, (int)R.centers.size(), R.rsk, R.rcnt,
                                    kSketchSize, best, num, den);
        if (rc != HALIGN_OK) return rc;
    }
    const int nr = (int)R.centers.size();
    std::vector<std::vector<int32_t>> members((size_t)nr);
    for (int32_t i = 0; i < ch.n; ++i) members[(size_t)best[i]].push_back(i);
    if (!ctx->msa_tmp.ensure((size_t)std::max<int64_t>(need, 1))) return fail(ctx, HALIGN_E_NOMEM, "pinned host buffer");
    std::vector<int64_t> src((size_t)ch.n, 0);
    std::vector<uint8_t> hins((size_t)ch.n, 0);
    int64_t toff = 0;
    n_ins = 0;
    for (int r = 0; r < nr; ++r) {
        const auto &m = members[(size_t)r];
        if (m.empty()) continue;
        std::string blob, names;
        std::vector<int64_t> off{0}, noff{0};
        for (int32_t i : m) {
            blob.append(ch.seqs, (size_t)ch.seq_off[i], (size_t)(ch.seq_off[i + 1] - ch.seq_off[i]));
            names.append(ch.ids, (size_t)ch.id_off[i], (size_t)(ch.id_off[i + 1] - ch.id_off[i]));
            off.push_back((int64_t)blob.size());
            noff.push_back((int64_t)names.size());
        }
        std::vector<uint8_t> flags(m.size());
        int64_t l = 0, ib = 0;
        int32_t ni = 0;
        const int rc = center_align_fasta(R.centers[(size_t)r], (int32_t)m.size(), blob.data(), off.data(),
                                          names.data(), noff.data(), ctx->msa_tmp.p + toff, need - toff, &l, &ib, &ni,
                                          nullptr, flags.data());
        if (rc != HALIGN_OK) return rc;
        // where each member's record landed in this reference's output
        int64_t a = toff, b = toff + ib;
        for (size_t j = 0; j < m.size(); ++j) {
            const int32_t i = m[j];
            hins[(size_t)i] = flags[j];
            if (dropped_name(ch, i)) continue;
            const int64_t sz = ch.id_off[i + 1] - ch.id_off[i] + R.width + 3;
            if (flags[j]) {
                src[(size_t)i] = a;
                a += sz;
            } else {
                src[(size_t)i] = b;
                b += sz;
            }
        }
        toff += l;
        n_ins += ni;
    }
    // the chunk's records in input order, insertion group first
    int64_t d = 0;
    for (int pass = 0; pass < 2; ++pass) {
        for (int32_t i = 0; i < ch.n; ++i) {
            if (dropped_name(ch, i) || (hins[(size_t)i] != 0) != (pass == 0)) continue;
            const int64_t sz = ch.id_off[i + 1] - ch.id_off[i] + R.width + 3;
            std::memcpy(out.p + d, ctx->msa_tmp.p + src[(size_t)i], (size_t)sz);
            d += sz;
        }
        if (pass == 0) ins = d;
    }
    len = d;
    return HALIGN_OK;
}

}  // namespace

extern "C" int halign_msa_keep_length(halign_ctx *ctx, const char *center_path, const char *input_path,
                                      const char *output_path, int32_t batch, halign_msa_stats *stats)
{
    return halign_msa_keep_length_refs(ctx, center_path, nullptr, input_path, output_path, batch, stats);
}

extern "C" int halign_msa_keep_length_refs(halign_ctx *ctx, const char *center_path, const char *aligned_path,
                                           const char *input_path, const char *output_path, int32_t batch,
                                           halign_msa_stats *stats)
{
    if (!ctx || !center_path || !input_path || !output_path) return HALIGN_E_ARG;
    std::lock_guard<std::mutex> run_lock(ctx->msa_mu);
    const auto t0 = Clock::now();
    PhaseClock sclk("msa setup");
    if (batch <= 0) batch = 25600;   // alignQueryToRef's batch in halign4 (include/align.h:432-436)

    // the center file's records, and the aligned records the MSA command made
    // of them (the copy command: the same file)
    auto read_all = [&](const char *path, RecordChunk &rc) -> int {
        FastxReader rd;
        std::string e, id, desc, seq;
        if (!rd.open(path, e)) return fail(ctx, HALIGN_E_IO, e.c_str());
        int r;
        while ((r = rd.next(id, desc, seq, e)) > 0) rc.add(id, desc, seq);
        if (r < 0) return fail(ctx, HALIGN_E_IO, e.c_str());
        return HALIGN_OK;
    };
    RecordChunk refs, aligned;
    int rc = read_all(center_path, refs);
    if (rc != HALIGN_OK) return rc;
    if (refs.n == 0) return fail(ctx, HALIGN_E_ARG, "center file has no record");
    rc = read_all(aligned_path ? aligned_path : center_path, aligned);
    if (rc != HALIGN_OK) return rc;
    if (aligned.n != refs.n) return fail(ctx, HALIGN_E_ARG, "aligned file and center file differ in record count");
    auto seq_of = [](const RecordChunk &c, int i) {
        return std::string(c.seqs, (size_t)c.seq_off[i], (size_t)(c.seq_off[i + 1] - c.seq_off[i]));
    };
    auto id_of = [](const RecordChunk &c, int i) {
        return std::string(c.ids, (size_t)c.id_off[i], (size_t)(c.id_off[i + 1] - c.id_off[i]));
    };
    Refs R;
    if (refs.n == 1) {
        // the consensus of the aligned file is the center aligned to (:47-60)
        const int64_t alen = aligned.seq_off[1];
        if (alen <= 0) return fail(ctx, HALIGN_E_ARG, "center sequence is empty");
        std::string cons((size_t)alen, 'A');
        std::vector<uint32_t> counts((size_t)alen * 7);
        uint64_t nseq = 0;
        rc = halign_consensus(ctx, aligned.seqs.data(), 1, alen, alen, 0, 4096, 0, counts.data(), &cons[0], &nseq,
                              nullptr);
        if (rc != HALIGN_OK) return rc;
        // the context keeps the last run's center (and its batch buffers) for
        // the next run against the same center file
        const std::string cseq = seq_of(refs, 0);
        const std::string key = cons + '\n' + cseq;
        if (!ctx->msa_center || ctx->msa_key != key) {
            if (ctx->msa_center) halign_center_free(ctx->msa_center);
            ctx->msa_center = nullptr;
            ctx->msa_key.clear();
            rc = halign_center_create_indexed(ctx, cons.data(), (int32_t)cons.size(), cseq.data(),
                                              (int32_t)cseq.size(), &ctx->msa_center);
            if (rc != HALIGN_OK) return rc;
            ctx->msa_key = key;
        }
        R.single = ctx->msa_center;
        R.width = (int64_t)cons.size();
    } else {
        // several references: one center each, a column map from its aligned
        // record (looked up by name, the last record of a name winning, as the
        // reference's map does), sketches resident
        R.multi = true;
        R.width = aligned.seq_off[1] - aligned.seq_off[0];
        for (int i = 1; i < aligned.n; ++i)
            if (aligned.seq_off[i + 1] - aligned.seq_off[i] != R.width)
                return fail(ctx, HALIGN_E_ARG, "aligned references differ in length");
        for (int r = 0; r < refs.n; ++r) {
            const std::string seq = seq_of(refs, r), id = id_of(refs, r);
            int a = -1;
            for (int j = 0; j < aligned.n; ++j)
                if (id_of(aligned, j) == id) a = j;
            if (a < 0) return fail(ctx, HALIGN_E_ARG, ("reference '" + id + "' not found in the aligned file").c_str());
            if (seq.empty()) return fail(ctx, HALIGN_E_ARG, "a reference sequence is empty");
            std::vector<int32_t> map;
            const std::string arow = seq_of(aligned, a);
            for (int64_t c = 0; c < (int64_t)arow.size(); ++c)
                if (arow[(size_t)c] != '-') map.push_back((int32_t)c);
            if (map.size() != seq.size())
                return fail(ctx, HALIGN_E_ARG, ("aligned record of '" + id + "' does not hold its bases").c_str());
            halign_center *c = nullptr;
            rc = halign_center_create_indexed(ctx, seq.data(), (int32_t)seq.size(), seq.data(), (int32_t)seq.size(),
                                              &c);
            if (rc != HALIGN_OK) return rc;
            R.centers.push_back(c);
            rc = center_set_columns(c, map.data(), R.width);
            if (rc != HALIGN_OK) return rc;
        }
        std::lock_guard<std::mutex> lock(ctx->mu);
        rc = mash_sketch_to_device(ctx, refs.n, refs.seqs.data(), refs.seq_off.data(), kSketchK, kSketchSize, true,
                                   kSketchSeed, R.rsk, R.rcnt);
        if (rc != HALIGN_OK) return rc;
    }
    sclk.mark("centers and indexes");

    const int out = open(output_path, O_WRONLY | O_CREAT | O_TRUNC, 0644);
    if (out < 0) return fail(ctx, HALIGN_E_IO, (std::string("cannot create ") + output_path).c_str());
    const std::string rest_path = std::string(output_path) + ".rest.tmp";
    const int rest = open(rest_path.c_str(), O_RDWR | O_CREAT | O_TRUNC, 0600);
    if (rest < 0) {
        close(out);
        return fail(ctx, HALIGN_E_IO, (std::string("cannot create ") + rest_path).c_str());
    }
    struct Files {
        int a, b;
        std::string bp;
        ~Files()
        {
            if (a >= 0) close(a);
            close(b);
            unlink(bp.c_str());
        }
    } files{out, rest, rest_path};
    sclk.mark("output files opened");

    // writeConsensusAndReferences (:541-561): the aligned file's records, cleaned
    int64_t out_off = 0, rest_off = 0;
    {
        std::string recs;
        for (int i = 0; i < aligned.n; ++i) {
            recs += '>';
            recs.append(aligned.ids, (size_t)aligned.id_off[i], (size_t)(aligned.id_off[i + 1] - aligned.id_off[i]));
            if (aligned.desc_off[i + 1] > aligned.desc_off[i]) {
                recs += ' ';
                recs.append(aligned.descs, (size_t)aligned.desc_off[i],
                            (size_t)(aligned.desc_off[i + 1] - aligned.desc_off[i]));
            }
            recs += '\n';
            for (int64_t p = aligned.seq_off[i]; p < aligned.seq_off[i + 1]; ++p)
                recs += clean_base((unsigned char)aligned.seqs[(size_t)p]);
            recs += '\n';
        }
        if (!write_at(out, recs.data(), recs.size(), 0)) return fail(ctx, HALIGN_E_IO, "write failed");
        out_off = (int64_t)recs.size();
    }

    Shared sh;
    halign_msa_stats st{};
    st.records = aligned.n;
    st.bytes = 0;

    // reader: chunks of `batch` records, at most two parsed ahead
    std::thread reader([&] {
        FastxReader rd;
        std::string err;
        if (!rd.open(input_path, err)) {
            sh.set_error(HALIGN_E_IO, err);
            std::lock_guard<std::mutex> lk(sh.mu);
            sh.reader_done = true;
            sh.cv.notify_all();
            return;
        }
        std::string id, desc, seq;
        for (;;) {
            auto ch = std::make_unique<RecordChunk>();
            int r = 1;
            while (ch->n < batch && (r = rd.next(id, desc, seq, err)) > 0) ch->add(id, desc, seq);
            if (r < 0) {
                sh.set_error(HALIGN_E_IO, err);
                break;
            }
            std::unique_lock<std::mutex> lk(sh.mu);
            if (ch->n > 0) {
                sh.cv.wait(lk, [&] { return sh.stop || sh.ready.size() < 2; });
                if (sh.stop) break;
                sh.ready.push_back(std::move(ch));
                sh.cv.notify_all();
            }
            if (r == 0) break;
        }
        std::lock_guard<std::mutex> lk(sh.mu);
        sh.reader_done = true;
        sh.cv.notify_all();
    });

    // writer: slots in the order the GPU loop fills them; slot k's first
    // group to the output, the rest to the spool
    Slot slots[2];
    slots[0].out = &ctx->msa_out[0];
    slots[1].out = &ctx->msa_out[1];
    bool producer_done = false;
    std::thread writer([&] {
        for (int k = 0;; k ^= 1) {
            {
                std::unique_lock<std::mutex> lk(sh.mu);
                sh.cv.wait(lk, [&] { return slots[k].full || producer_done || sh.stop; });
                if (sh.stop || !slots[k].full) return;
            }
            Slot &s = slots[k];
            PhaseClock wclk("msa writer");
            const bool ok = write_at(out, s.out->p, (size_t)s.ins, out_off) &&
                            write_at(rest, s.out->p + s.ins, (size_t)(s.len - s.ins), rest_off);
            wclk.mark("chunk written");
            out_off += s.ins;
            rest_off += s.len - s.ins;
            if (!ok) {
                sh.set_error(HALIGN_E_IO, "write failed");
                return;
            }
            std::lock_guard<std::mutex> lk(sh.mu);
            s.full = false;
            sh.cv.notify_all();
        }
    });

    double gpu_s = 0, wait_s = 0;
    for (int k = 0;; k ^= 1) {
        std::unique_ptr<RecordChunk> ch;
        {
            const auto w0 = Clock::now();
            std::unique_lock<std::mutex> lk(sh.mu);
            sh.cv.wait(lk, [&] { return sh.stop || !sh.ready.empty() || sh.reader_done; });
            wait_s += secs(w0, Clock::now());
            if (sh.stop || sh.ready.empty()) break;   // an error, or the reader is done
            ch = std::move(sh.ready.front());
            sh.ready.pop_front();
            sh.cv.notify_all();
            sh.cv.wait(lk, [&] { return sh.stop || !slots[k].full; });   // the writer is done with slot k
            if (sh.stop) break;
        }
        Slot &s = slots[k];
        int32_t n_ins = 0, kept = 0;
        const auto g0 = Clock::now();
        PhaseClock gclk("msa gpu");
        rc = align_chunk(ctx, R, *ch, *s.out, s.len, s.ins, n_ins, kept);
        gpu_s += secs(g0, Clock::now());
        gclk.mark("chunk aligned");
        if (rc != HALIGN_OK) {
            sh.set_error(rc, ctx->err);
            break;
        }
        st.records += kept;
        st.insertion_records += n_ins;
        std::lock_guard<std::mutex> lk(sh.mu);
        s.full = true;
        sh.cv.notify_all();
    }
    {
        std::lock_guard<std::mutex> lk(sh.mu);
        producer_done = true;
        sh.cv.notify_all();
    }
    writer.join();
    {
        std::lock_guard<std::mutex> lk(sh.mu);
        sh.stop = true;   // ends a reader still waiting to hand over a chunk
        sh.cv.notify_all();
    }
    reader.join();
    if (sh.err_code != HALIGN_OK) return fail(ctx, sh.err_code, sh.err.c_str());

    // mergeAlignedResults' order: the other group after the insertion group
    PhaseClock aclk("msa");
    aclk.mark("all chunks written");
    if (!append_file(rest, out, rest_off, out_off)) return fail(ctx, HALIGN_E_IO, "appending the spool file failed");
    aclk.mark("spool appended");
    if (close(out) != 0) {
        files.a = -1;
        return fail(ctx, HALIGN_E_IO, "closing the output failed");
    }
    files.a = -1;
    st.bytes = out_off + rest_off;
    st.seconds = secs(t0, Clock::now());
    st.gpu_seconds = gpu_s;
    st.input_wait_seconds = wait_s;
    if (stats) *stats = st;
    return HALIGN_OK;
}

// ---------------------------------------------------------------------------
// Auto-center pieces: Top-K longest records (preprocessInputFasta with
// TopKLongestSelector), the consensus of an aligned file with its FASTA and
// JSON outputs (generateConsensusSequence), and the keep-length run around
// them.

namespace {

// seq_io::SeqWriter::writeFasta with line width w (src/utils/seq_io.cpp:236-268)
void append_fasta(std::string &o, const std::string &id, const std::string &desc, const std::string &seq, size_t w)
{
    o += '>';
    o += id;
    if (!desc.empty()) {
        o += ' ';
        o += desc;
    }
    o += '\n';
    if (seq.empty()) {
        o += '\n';
        return;
    }
    for (size_t i = 0; i < seq.size(); i += w) {
        o.append(seq, i, std::min(w, seq.size() - i));
        o += '\n';
    }
}

bool write_file(const char *path, const std::string &data)
{
    const int fd = open(path, O_WRONLY | O_CREAT | O_TRUNC, 0644);
    if (fd < 0) return false;
    const bool ok = write_at(fd, data.data(), data.size(), 0);
    return close(fd) == 0 && ok;
}

}  // namespace

extern "C" int halign_topk_longest(halign_ctx *ctx, const char *input_path, int32_t k, const char *output_path,
                                   int64_t *n_records)
{
    if (!input_path || !output_path || k < 0) return HALIGN_E_ARG;   // ctx optional: host only
    FastxReader rd;
    std::string e;
    if (!rd.open(input_path, e)) return fail(ctx, HALIGN_E_IO, e.c_str());
    // the K longest cleaned records, the earlier one first among equal lengths
    struct Item {
        size_t len;
        uint64_t order;
        std::string id, desc, seq;
    };
    auto worse = [](const Item &a, const Item &b) {   // a ranks below b
        return a.len != b.len ? a.len < b.len : a.order > b.order;
    };
    // heap with the lowest-ranked kept record on top
    auto heap_cmp = [&](const Item &a, const Item &b) { return worse(b, a); };
    std::vector<Item> heap;
    std::string id, desc, seq;
    uint64_t order = 0;
    int r;
    while ((r = rd.next(id, desc, seq, e)) > 0) {
        for (auto &ch : seq) ch = clean_base((unsigned char)ch);   // cleanSequence
        Item it{seq.size(), order++, id, desc, seq};
        if (k == 0) continue;
        if ((int32_t)heap.size() < k) {
            heap.push_back(std::move(it));
            std::push_heap(heap.begin(), heap.end(), heap_cmp);
        } else if (worse(heap.front(), it)) {
            std::pop_heap(heap.begin(), heap.end(), heap_cmp);
            heap.back() = std::move(it);
            std::push_heap(heap.begin(), heap.end(), heap_cmp);
        }
    }
    if (r < 0) return fail(ctx, HALIGN_E_IO, e.c_str());
    std::sort(heap.begin(), heap.end(), [&](const Item &a, const Item &b) { return worse(b, a); });
    std::string o;
    for (const auto &it : heap) append_fasta(o, it.id, it.desc, it.seq, 80);
    if (!write_file(output_path, o)) return fail(ctx, HALIGN_E_IO, "write failed");
    if (n_records) *n_records = (int64_t)order;
    return HALIGN_OK;
}

extern "C" int halign_consensus_file(halign_ctx *ctx, const char *aligned_path, const char *out_fasta,
                                     const char *out_json, uint64_t seq_limit, int64_t batch, char *cons_out,
                                     int64_t cap, int64_t *cons_len)
{
    if (!ctx || !aligned_path || !out_fasta || !out_json || batch <= 0) return HALIGN_E_ARG;
    FastxReader rd;
    std::string e, id, desc, seq;
    if (!rd.open(aligned_path, e)) return fail(ctx, HALIGN_E_IO, e.c_str());
    int r = rd.next(id, desc, seq, e);
    if (r < 0) return fail(ctx, HALIGN_E_IO, e.c_str());
    if (r == 0) return fail(ctx, HALIGN_E_ARG, "aligned fasta is empty");
    const size_t aln = seq.size();
    if (aln == 0) return fail(ctx, HALIGN_E_ARG, "first sequence length is 0");
    // the reference's reading loop: batches of `batch` records, the record
    // read after each batch dropped, at most seq_limit records kept
    std::string rows = seq;
    uint64_t kept = 1;
    size_t in_batch = 1;
    for (;;) {
        if (seq_limit != 0 && kept >= seq_limit) break;
        while (in_batch < (size_t)batch && (seq_limit == 0 || kept < seq_limit)) {
            r = rd.next(id, desc, seq, e);
            if (r < 0) return fail(ctx, HALIGN_E_IO, e.c_str());
            if (r == 0) break;
            if (seq.size() != aln) return fail(ctx, HALIGN_E_ARG, "alignment length mismatch when reading");
            rows += seq;
            ++kept;
            ++in_batch;
        }
        in_batch = 0;
        r = rd.next(id, desc, seq, e);   // read and dropped
        if (r < 0) return fail(ctx, HALIGN_E_IO, e.c_str());
        if (r == 0) break;
    }
    std::vector<uint32_t> counts(aln * 7);
    std::string cons(aln, 'A');
    uint64_t nseq = 0;
    // every row read here is counted: no further batch drop on the device
    const int rc = halign_consensus(ctx, rows.data(), (int64_t)kept, (int64_t)aln, (int64_t)aln, 0,
                                    (int64_t)kept + 1, 0, counts.data(), &cons[0], &nseq, nullptr);
    if (rc != HALIGN_OK) return rc;
    // writeConsensusFasta and writeCountsJson (src/consensus/consensus.cpp:38-68)
    std::string fa = ">consensus\n";
    for (size_t i = 0; i < aln; i += 80) {
        fa.append(cons, i, std::min<size_t>(80, aln - i));
        fa += '\n';
    }
    std::string js;
    js.reserve(aln * 160 + 128);
    js += "{\n    \"consensus\": {\n        \"num_seqs\": " + std::to_string(kept) +
          ",\n        \"aln_len\": " + std::to_string(aln) + ",\n        \"counts\": [\n";
    static const char *const kKeys[7] = {"A", "C", "G", "T", "U", "N", "-"};
    for (size_t i = 0; i < aln; ++i) {
        js += "            {\n";
        for (int c = 0; c < 7; ++c) {
            js += "                \"";
            js += kKeys[c];
            js += "\": " + std::to_string(counts[i * 7 + c]);
            js += c < 6 ? ",\n" : "\n";
        }
        js += i + 1 < aln ? "            },\n" : "            }\n";
    }
    js += "        ]\n    }\n}";
    if (!write_file(out_fasta, fa) || !write_file(out_json, js)) return fail(ctx, HALIGN_E_IO, "write failed");
    if (cons_len) *cons_len = (int64_t)aln;
    if (cons_out) {
        if ((int64_t)aln > cap) return fail(ctx, HALIGN_E_CAPACITY, "consensus output too small");
        std::memcpy(cons_out, cons.data(), aln);
    }
    return HALIGN_OK;
}

extern "C" int halign_msa_auto_center(halign_ctx *ctx, const char *input_path, const char *aligned_topk_path,
                                      const char *output_path, int32_t cons_n, int32_t batch,
                                      halign_msa_stats *stats)
{
    if (!ctx || !input_path || !output_path || cons_n <= 0) return HALIGN_E_ARG;
    const std::string base = output_path;
    const std::string topk = base + ".topk.fa", cons = base + ".consensus.fa", json = base + ".consensus.json";
    struct Cleanup {
        std::vector<std::string> f;
        ~Cleanup()
        {
            for (auto &x : f) unlink(x.c_str());
        }
    } cl{{topk, cons, json}};
    // preprocessInputFasta's Top-K file, the MSA command's output for it (the
    // copy command: the file itself), its consensus (seq_limit cons_n), then
    // the keep-length run against the consensus record (src/halign4.cpp:
    // 87, 119-147)
    int rc = halign_topk_longest(ctx, input_path, cons_n, topk.c_str(), nullptr);
    if (rc != HALIGN_OK) return rc;
    rc = halign_consensus_file(ctx, aligned_topk_path ? aligned_topk_path : topk.c_str(), cons.c_str(), json.c_str(),
                               (uint64_t)cons_n, 4096, nullptr, 0, nullptr);
    if (rc != HALIGN_OK) return rc;
    return halign_msa_keep_length(ctx, cons.c_str(), input_path, output_path, batch, stats);
}
