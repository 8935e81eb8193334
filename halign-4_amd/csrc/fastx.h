// fastx.h -- FASTA/FASTQ input and FASTA output of the keep-length run.
//
// FastxReader reads records with the rules of the vendored kseq_read that
// seq_io::KseqReader wraps (reference src/utils/seq_io.cpp:67-155,
// include/kseq.h), from a gzip or plain file (zlib reads both):
//  * records start at a '>' or '@' byte; bytes before the first are skipped;
//  * the name runs to the first whitespace byte (space, \t, \n, \v, \f, \r);
//    unless that byte was '\n', the comment is the rest of the line;
//  * sequence lines follow until a line starts with '>', '@' or '+'; empty
//    lines are skipped, the other lines are appended whole; after each line
//    one trailing '\r' is dropped when the sequence so far is longer than one
//    byte (a line read that finds the data already exhausted drops nothing);
//  * after a '+' line, quality lines are read until they are at least as long
//    as the sequence; a missing or different-length quality is an error;
//  * a comment line loses one trailing '\r' when it is longer than one byte.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace hal {

class FastxReader {
public:
    FastxReader() = default;
    ~FastxReader();
    FastxReader(const FastxReader &) = delete;
    FastxReader &operator=(const FastxReader &) = delete;

    // false with err set when the file cannot be opened
    bool open(const char *path, std::string &err);
    // 1: a record in id / desc / seq; 0: end of the file; -1: error (err)
    int next(std::string &id, std::string &desc, std::string &seq, std::string &err);

private:
    int get();                                  // next byte, -1 at the end, -2 on a read error
    int until(bool space, std::string &dst, int *delim);   // -1: nothing left to read
    bool fill();

    void *gz_ = nullptr;
    std::vector<unsigned char> buf_;
    size_t pos_ = 0, len_ = 0;
    bool eof_ = false, bad_ = false;
    int hdr_ = 0;          // header byte already consumed for the next record
    std::string qual_;
};

// A chunk of records packed for the device: sequences and names back to back
// with n + 1 offsets each.
struct RecordChunk {
    int32_t n = 0;
    std::string seqs, ids;
    std::vector<int64_t> seq_off{0}, id_off{0};
    std::string descs;
    std::vector<int64_t> desc_off{0};
    void clear();
    void add(const std::string &id, const std::string &desc, const std::string &seq);
};

// seq_io::cleanSequence (include/utils.h:161-197): A/C/G/T/U in either case to
// upper case, '-' kept, every other byte to 'N'
char clean_base(unsigned char c);

}  // namespace hal
