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
// the insertion SAM files first); the writer appends those straight to the
// output and spools the others to a temporary file next to it, appended at
// the end.
#include <hip/hip_runtime.h>

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

struct PinnedBuf {
    char *p = nullptr;
    size_t cap = 0;
    ~PinnedBuf()
    {
        if (p) (void)hipHostFree(p);
    }
    bool ensure(size_t n)
    {
        if (n <= cap) return true;
        const size_t want = std::max(n, cap + cap / 2);
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
        if (hipHostMalloc((void **)&p, want, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            return false;
        }
        cap = want;
        return true;
    }
};

// One chunk's records ready for the writer: out.p[0 .. ins) first group,
// out.p[ins .. len) the rest.
struct Slot {
    PinnedBuf out;
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

bool write_all(FILE *f, const char *p, size_t n) { return n == 0 || fwrite(p, 1, n, f) == n; }

}  // namespace

extern "C" int halign_msa_keep_length(halign_ctx *ctx, const char *center_path, const char *input_path,
                                      const char *output_path, int32_t batch, halign_msa_stats *stats)
{
    if (!ctx || !center_path || !input_path || !output_path) return HALIGN_E_ARG;
    const auto t0 = Clock::now();
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
    halign_center *center = nullptr;
    int rc = halign_center_create_indexed(ctx, cons.data(), (int32_t)cons.size(), cseq.data(), (int32_t)cseq.size(),
                                          &center);
    if (rc != HALIGN_OK) return rc;
    struct CenterGuard {
        halign_center *c;
        ~CenterGuard() { halign_center_free(c); }
    } cg{center};

    FILE *out = fopen(output_path, "wb");
    if (!out) return fail(ctx, HALIGN_E_IO, (std::string("cannot create ") + output_path).c_str());
    const std::string rest_path = std::string(output_path) + ".rest.tmp";
    FILE *rest = fopen(rest_path.c_str(), "w+b");
    if (!rest) {
        fclose(out);
        return fail(ctx, HALIGN_E_IO, (std::string("cannot create ") + rest_path).c_str());
    }
    struct Files {
        FILE *a, *b;
        std::string bp;
        ~Files()
        {
            if (a) fclose(a);
            if (b) fclose(b);
            std::remove(bp.c_str());
        }
    } files{out, rest, rest_path};
    setvbuf(out, nullptr, _IOFBF, 8 << 20);
    setvbuf(rest, nullptr, _IOFBF, 8 << 20);

    // writeConsensusAndReferences (:541-561): the center record, cleaned
    {
        std::string rec = ">" + cid;
        if (!cdesc.empty()) rec += " " + cdesc;
        rec += "\n";
        const size_t h = rec.size();
        rec.resize(h + cseq.size());
        for (size_t i = 0; i < cseq.size(); ++i) rec[h + i] = clean_base((unsigned char)cseq[i]);
        rec += "\n";
        if (!write_all(out, rec.data(), rec.size())) return fail(ctx, HALIGN_E_IO, "write failed");
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
    bool producer_done = false;
    std::thread writer([&] {
        for (int k = 0;; k ^= 1) {
            {
                std::unique_lock<std::mutex> lk(sh.mu);
                sh.cv.wait(lk, [&] { return slots[k].full || producer_done || sh.stop; });
                if (sh.stop || !slots[k].full) return;
            }
            Slot &s = slots[k];
            const bool ok = write_all(out, s.out.p, (size_t)s.ins) &&
                            write_all(rest, s.out.p + s.ins, (size_t)(s.len - s.ins));
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
        if (!s.out.ensure((size_t)std::max<int64_t>(need, 1))) {
            sh.set_error(HALIGN_E_NOMEM, "pinned host buffer");
            break;
        }
        int32_t n_ins = 0;
        const auto g0 = Clock::now();
        rc = halign_center_align_fasta(center, ch->n, ch->seqs.data(), ch->seq_off.data(), ch->ids.data(),
                                       ch->id_off.data(), s.out.p, need, &s.len, &s.ins, &n_ins, nullptr);
        gpu_s += secs(g0, Clock::now());
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
    if (fflush(rest) != 0 || fseek(rest, 0, SEEK_SET) != 0) return fail(ctx, HALIGN_E_IO, "spool file");
    std::vector<char> buf(8 << 20);
    for (;;) {
        const size_t r = fread(buf.data(), 1, buf.size(), rest);
        if (r == 0) break;
        if (!write_all(out, buf.data(), r)) return fail(ctx, HALIGN_E_IO, "write failed");
    }
    if (ferror(rest)) return fail(ctx, HALIGN_E_IO, "spool file read");
    if (fflush(out) != 0) return fail(ctx, HALIGN_E_IO, "write failed");
    st.bytes = (int64_t)ftell(out);
    st.seconds = secs(t0, Clock::now());
    st.gpu_seconds = gpu_s;
    st.input_wait_seconds = wait_s;
    if (stats) *stats = st;
    return HALIGN_OK;
}
