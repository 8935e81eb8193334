// fastx.cpp -- FastxReader (the record rules are listed in fastx.h) and the
// C ABI halign_fastx_*.
#include "fastx.h"

#include <zlib.h>

#include <cerrno>
#include <cstring>
#include <new>

#include "../../include/halign_hip.h"

namespace hal {

namespace {
constexpr size_t kBuf = 1 << 20;

bool is_ws(int c) { return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r'; }
}  // namespace

FastxReader::~FastxReader()
{
    if (gz_) gzclose((gzFile)gz_);
}

bool FastxReader::open(const char *path, std::string &err)
{
    gz_ = gzopen(path, "rb");
    if (!gz_) {
        err = std::string("cannot open ") + path + ": " + std::strerror(errno);
        return false;
    }
    gzbuffer((gzFile)gz_, 1 << 20);
    buf_.resize(kBuf);
    return true;
}

bool FastxReader::fill()
{
    if (eof_ || bad_) return false;
    const int r = gzread((gzFile)gz_, buf_.data(), (unsigned)buf_.size());
    if (r < 0) {
        bad_ = true;
        return false;
    }
    if (r == 0) {
        eof_ = true;
        return false;
    }
    pos_ = 0;
    len_ = (size_t)r;
    return true;
}

int FastxReader::get()
{
    if (pos_ >= len_ && !fill()) return bad_ ? -2 : -1;
    return buf_[pos_++];
}

// Appends the bytes up to the delimiter (whitespace when space, else '\n';
// consumed, stored in *delim, 0 at the end) to dst. Returns -1 without
// touching dst when nothing is left; -2 on a read error; otherwise 0, after
// dropping one trailing '\r' from a line read when dst is longer than one byte.
int FastxReader::until(bool space, std::string &dst, int *delim)
{
    if (delim) *delim = 0;
    if (pos_ >= len_ && !fill()) return bad_ ? -2 : -1;
    for (;;) {
        size_t i = pos_;
        if (space) {
            while (i < len_ && !is_ws(buf_[i])) ++i;
        } else {
            const void *nl = std::memchr(buf_.data() + pos_, '\n', len_ - pos_);
            i = nl ? (size_t)((const unsigned char *)nl - buf_.data()) : len_;
        }
        dst.append((const char *)buf_.data() + pos_, i - pos_);
        if (i < len_) {
            if (delim) *delim = buf_[i];
            pos_ = i + 1;
            break;
        }
        pos_ = len_;
        if (!fill()) {
            if (bad_) return -2;
            break;
        }
    }
    if (!space && dst.size() > 1 && dst.back() == '\r') dst.pop_back();
    return 0;
}

int FastxReader::next(std::string &id, std::string &desc, std::string &seq, std::string &err)
{
    int c;
    if (hdr_ == 0) {
        while ((c = get()) >= 0 && c != '>' && c != '@') {}
        if (c == -2) return err = "read error", -1;
        if (c < 0) return 0;
        hdr_ = c;
    }
    id.clear();
    desc.clear();
    seq.clear();
    qual_.clear();
    int r = until(true, id, &c);
    if (r == -2) return err = "read error", -1;
    if (r < 0) return 0;
    if (c != '\n' && until(false, desc, nullptr) == -2) return err = "read error", -1;
    while ((c = get()) >= 0 && c != '>' && c != '+' && c != '@') {
        if (c == '\n') continue;
        seq.push_back((char)c);
        if (until(false, seq, nullptr) == -2) return err = "read error", -1;
    }
    if (c == -2) return err = "read error", -1;
    if (c == '>' || c == '@') hdr_ = c;
    if (c != '+') return 1;
    while ((c = get()) >= 0 && c != '\n') {}
    if (c < 0) return err = "record " + id + ": no quality string", -1;
    while ((r = until(false, qual_, nullptr)) >= 0 && qual_.size() < seq.size()) {}
    if (r == -2) return err = "read error", -1;
    hdr_ = 0;
    if (qual_.size() != seq.size()) return err = "record " + id + ": quality length differs from the sequence", -1;
    return 1;
}

void RecordChunk::clear()
{
    n = 0;
    seqs.clear();
    ids.clear();
    descs.clear();
    seq_off.assign(1, 0);
    id_off.assign(1, 0);
    desc_off.assign(1, 0);
}

void RecordChunk::add(const std::string &id, const std::string &desc, const std::string &seq)
{
    ++n;
    seqs += seq;
    ids += id;
    descs += desc;
    seq_off.push_back((int64_t)seqs.size());
    id_off.push_back((int64_t)ids.size());
    desc_off.push_back((int64_t)descs.size());
}

char clean_base(unsigned char c)
{
    switch (c) {
    case 'A': case 'a': return 'A';
    case 'C': case 'c': return 'C';
    case 'G': case 'g': return 'G';
    case 'T': case 't': return 'T';
    case 'U': case 'u': return 'U';
    case '-': return '-';
    default: return 'N';
    }
}

}  // namespace hal

struct halign_fastx {
    hal::FastxReader rd;
    hal::RecordChunk chunk;
    std::string err, id, desc, seq;
    bool failed = false;
};

extern "C" {

int halign_fastx_open(const char *path, halign_fastx **out)
{
    if (!path || !out) return HALIGN_E_ARG;
    *out = nullptr;
    halign_fastx *f = new (std::nothrow) halign_fastx();
    if (!f) return HALIGN_E_NOMEM;
    if (!f->rd.open(path, f->err)) {
        delete f;
        return HALIGN_E_IO;
    }
    *out = f;
    return HALIGN_OK;
}

int halign_fastx_next(halign_fastx *f, int32_t max_records, int32_t *n, const char **seqs,
                      const int64_t **seq_offsets, const char **ids, const int64_t **id_offsets,
                      const char **descs, const int64_t **desc_offsets)
{
    if (!f || max_records <= 0 || !n) return HALIGN_E_ARG;
    *n = 0;
    if (f->failed) return HALIGN_E_IO;
    f->chunk.clear();
    while (f->chunk.n < max_records) {
        const int r = f->rd.next(f->id, f->desc, f->seq, f->err);
        if (r < 0) {
            f->failed = true;
            return HALIGN_E_IO;
        }
        if (r == 0) break;
        f->chunk.add(f->id, f->desc, f->seq);
    }
    *n = f->chunk.n;
    if (seqs) *seqs = f->chunk.seqs.data();
    if (seq_offsets) *seq_offsets = f->chunk.seq_off.data();
    if (ids) *ids = f->chunk.ids.data();
    if (id_offsets) *id_offsets = f->chunk.id_off.data();
    if (descs) *descs = f->chunk.descs.data();
    if (desc_offsets) *desc_offsets = f->chunk.desc_off.data();
    return HALIGN_OK;
}

const char *halign_fastx_error(const halign_fastx *f) { return f ? f->err.c_str() : "no reader"; }

void halign_fastx_close(halign_fastx *f) { delete f; }

}  // extern "C"
