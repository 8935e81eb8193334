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

using namespace hal;

namespace {

using Clock = std::chrono::steady_clock;

double secs(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double>(b - a).count(); }

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

extern "C" int halign_msa_keep_length(halign_ctx *ctx, const char *center_path, const char *input_path,
                                      const char *output_path, int32_t batch, halign_msa_stats *stats)
{
    if (!ctx || !center_path || !input_path || !output_path) return HALIGN_E_ARG;
    std::lock_guard<std::mutex> run_lock(ctx->msa_mu);
    const auto t0 = Clock::now();
    PhaseClock sclk("msa setup");
    if (batch <= 0) batch = 25600;   // alignQueryToRef's batch in halign4 (include/align.h:432-436)

    // the center file: one record (more is the multi-reference path)
    std::string cid, cdesc, cseq, e;
    {
        FastxReader rd;
        if (!rd.open(center_path, e)) return fail(ctx, HALIGN_E_IO, e.c_str());
        std::string id2, desc2, seq2;
        const int r = rd.next(cid, cdesc, cseq, e);
        if (r < 0) return fail(ctx, HALIGN_E_IO, e.c_str());
        if (r == 0) return fail(ctx, HALIGN_E_ARG, "center file has no record");
        const int r2 = rd.next(id2, desc2, seq2, e);
        if (r2 < 0) return fail(ctx, HALIGN_E_IO, e.c_str());
        if (r2 > 0) return fail(ctx, HALIGN_E_UNSUPPORTED, "center file has more than one record");
    }
    if (cseq.empty()) return fail(ctx, HALIGN_E_ARG, "center sequence is empty");
    // RefAligner aligns to the consensus of the (copied) center file
    std::string cons(cseq.size(), 'A');
    {
        std::vector<uint32_t> counts(cseq.size() * 7);
        uint64_t nseq = 0;
        const int rc = halign_consensus(ctx, cseq.data(), 1, (int64_t)cseq.size(), (int64_t)cseq.size(), 0, 4096, 0,
                                        counts.data(), &cons[0], &nseq, nullptr);
        if (rc != HALIGN_OK) return rc;
    }
    // the context keeps the last run's center (and its batch buffers) for the
    // next run against the same center file
    int rc = HALIGN_OK;
    const std::string key = cons + '\n' + cseq;
    if (!ctx->msa_center || ctx->msa_key != key) {
        if (ctx->msa_center) halign_center_free(ctx->msa_center);
        ctx->msa_center = nullptr;
        ctx->msa_key.clear();
        rc = halign_center_create_indexed(ctx, cons.data(), (int32_t)cons.size(), cseq.data(), (int32_t)cseq.size(),
                                          &ctx->msa_center);
        if (rc != HALIGN_OK) return rc;
        ctx->msa_key = key;
    }
    halign_center *center = ctx->msa_center;
    sclk.mark("center and index");

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

    // writeConsensusAndReferences (:541-561): the center record, cleaned
    int64_t out_off = 0, rest_off = 0;
    {
        std::string rec = ">" + cid;
        if (!cdesc.empty()) rec += " " + cdesc;
        rec += "\n";
        const size_t h = rec.size();
        rec.resize(h + cseq.size());
        for (size_t i = 0; i < cseq.size(); ++i) rec[h + i] = clean_base((unsigned char)cseq[i]);
        rec += "\n";
        if (!write_at(out, rec.data(), rec.size(), 0)) return fail(ctx, HALIGN_E_IO, "write failed");
        out_off = (int64_t)rec.size();
    }

    Shared sh;
    halign_msa_stats st{};
    st.records = 1;
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
        int64_t need = 0;
        int32_t kept = 0;
        for (int32_t i = 0; i < ch->n; ++i) {
            const int64_t nl = ch->id_off[i + 1] - ch->id_off[i];
            if (nl > 0 && ch->ids[(size_t)ch->id_off[i]] == '@') continue;   // no record (its SAM line is skipped)
            need += nl + (int64_t)cons.size() + 3;
            ++kept;
        }
        Slot &s = slots[k];
        if (!s.out->ensure((size_t)std::max<int64_t>(need, 1))) {
            sh.set_error(HALIGN_E_NOMEM, "pinned host buffer");
            break;
        }
        int32_t n_ins = 0;
        const auto g0 = Clock::now();
        PhaseClock gclk("msa gpu");
        rc = halign_center_align_fasta(center, ch->n, ch->seqs.data(), ch->seq_off.data(), ch->ids.data(),
                                       ch->id_off.data(), s.out->p, need, &s.len, &s.ins, &n_ins, nullptr);
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
