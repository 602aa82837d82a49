// Read ingest: FASTQ records located in a file held in memory, the same way on
// the host and on the device (the device parser runs fq_record per record over
// a newline table it builds itself).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "sa_types.hpp"

namespace sa {

// One FASTQ record of a file held in memory (byte offsets into the file).
struct FqLoc {
    uint64_t name_off, comment_off, seq_off, qual_off;
    uint32_t name_len, comment_len, seq_len, qual_len;
    uint8_t end;          // 0 none, 1 or 2: the name ended in "/1" or "/2" (stripped)
    uint8_t has_comment;
    uint8_t err;          // FQ_OK or the first check the record fails
    uint8_t first;        // first byte of the record (the error message of FQ_NO_AT names it)
    uint32_t pad;
};
static_assert(sizeof(FqLoc) == 56, "FqLoc layout");
enum : uint8_t { FQ_OK = 0, FQ_NO_AT = 1, FQ_EMPTY_NAME = 2, FQ_LENGTHS = 3 };

// src/io/mod.rs:32-38 trims with u8::is_ascii_whitespace (no vertical tab)
SA_HD bool ascii_ws(uint8_t c) { return c == ' ' || c == '\t' || c == '\n' || c == '\f' || c == '\r'; }

// A line as BufRead::read_line / read_until(b'\n') return it: [s, e), e past the
// '\n' when the line has one; absent lines (past the end of the file) are empty.
struct FqLine {
    uint64_t s, e;
    bool present;
};

// The record whose four lines are l[0..3] (src/io/fastq.rs:22-95): the name
// line must start with '@'; the header after it is trimmed at its end and split
// at the first space or tab into name and comment (src/io/mod.rs:32-38); an
// empty name is an error; a trailing "/1" or "/2" is cut off the name; the
// sequence and quality lines lose their last byte (the '\n', or the last
// character of an unterminated last line) and then one '\r'; their lengths must
// agree. The '+' line is not read.
SA_HD FqLoc fq_record(const uint8_t* p, const FqLine* l) {
    FqLoc r{};
    const uint64_t hs = l[0].s, he = l[0].e;
    r.first = he > hs ? p[hs] : 0;
    if (he == hs || p[hs] != '@') {
        r.err = FQ_NO_AT;
        return r;
    }
    uint64_t te = he;
    while (te > hs + 1 && ascii_ws(p[te - 1])) --te;
    uint64_t sep = hs + 1;
    while (sep < te && p[sep] != ' ' && p[sep] != '\t') ++sep;
    r.name_off = hs + 1;
    r.name_len = (uint32_t)(sep - (hs + 1));
    if (sep < te) {
        r.has_comment = 1;
        r.comment_off = sep + 1;
        r.comment_len = (uint32_t)(te - (sep + 1));
    }
    if (r.name_len == 0) {
        r.err = FQ_EMPTY_NAME;
        return r;
    }
    if (r.name_len >= 2 && p[r.name_off + r.name_len - 2] == '/' &&
        (p[r.name_off + r.name_len - 1] == '1' || p[r.name_off + r.name_len - 1] == '2')) {
        r.end = p[r.name_off + r.name_len - 1] == '1' ? 1 : 2;
        r.name_len -= 2;
    }
    auto content = [&](const FqLine& x, uint64_t& off, uint32_t& len) {
        off = x.s;
        uint64_t e = x.e > x.s ? x.e - 1 : x.s;
        if (e > x.s && p[e - 1] == '\r') --e;
        len = (uint32_t)(e - x.s);
    };
    content(l[1], r.seq_off, r.seq_len);
    content(l[3], r.qual_off, r.qual_len);
    if (r.seq_len != r.qual_len) r.err = FQ_LENGTHS;
    return r;
}

// Line k of a file of n bytes whose newline positions are nl[0..n_nl).
SA_HD FqLine fq_line(const uint64_t* nl, uint64_t n_nl, uint64_t n, uint64_t k) {
    if (k > n_nl) return FqLine{n, n, false};
    const uint64_t s = k == 0 ? 0 : nl[k - 1] + 1;
    if (k < n_nl) return FqLine{s, nl[k] + 1, true};
    return s < n ? FqLine{s, n, true} : FqLine{n, n, false};
}

// Records in a file of n bytes with n_nl newlines: a record starts at every
// fourth line that is present.
SA_HD uint64_t fq_records(uint64_t n_nl, uint64_t n, bool last_is_newline) {
    const uint64_t lines = n_nl + (n > 0 && !last_is_newline ? 1 : 0);
    return (lines + 3) / 4;
}

SA_HD uint8_t ascii_upper(uint8_t c) { return c >= 'a' && c <= 'z' ? (uint8_t)(c - 32) : c; }

// Error text of a record that fails a check (the wording of the host reader).
std::string fq_error(const FqLoc& r, const std::string& name);

// Parses FASTQ files held in host memory on `device` (r2 may be null): the
// records' locations per file, and their uppercase sequences resident on the
// device in the layout a read set is staged in (reads of both files interleaved,
// read x at d_seq + off[x], 16 zero bytes after the last base). *d_seq is
// allocated with device_alloc. Returns the device milliseconds (copy of the files
// in through the last kernel).
double ingest_fastq(int device, const std::string& r1, const std::string* r2, std::vector<FqLoc>& loc1,
                    std::vector<FqLoc>& loc2, void** d_seq, std::vector<uint64_t>& off);

}  // namespace sa
