// Reference/read input and host-side index construction for the drop-in.
#include <algorithm>
#include <array>
#include <cstdio>
#include <memory>
#include <atomic>
#include <cmath>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <thread>

#include <zlib.h>

#include "sa_host.hpp"
#include "sa_ingest.hpp"

namespace sa {

namespace {
void rstrip(std::string& s) {
    while (!s.empty() && ascii_ws((uint8_t)s.back())) s.pop_back();
}

// src/io/mod.rs:32-38
void split_header(const std::string& h0, std::string& name, std::string& comment, bool& has_comment) {
    std::string h = h0;
    rstrip(h);
    size_t p = h.find_first_of(" \t");
    if (p == std::string::npos) { name = h; comment.clear(); has_comment = false; }
    else { name = h.substr(0, p); comment = h.substr(p + 1); has_comment = true; }
}

template <typename F>
void parallel_for(int threads, size_t n, F f) {
    if (threads <= 1 || n <= 1) {
        for (size_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::vector<std::thread> th;
    std::atomic<size_t> next{0};
    for (int t = 0; t < threads; ++t)
        th.emplace_back([&] {
            for (size_t i; (i = next.fetch_add(1)) < n;) f(i);
        });
    for (auto& t : th) t.join();
}
}  // namespace

// src/io/fasta.rs:116-165 and src/packed_seq.rs:56-92 (wildcards -> position hash).
// The file is read in blocks; each line is split at '\n', trailing whitespace
// is stripped and the bases are packed 2 bits each, 32 to a word.
PackedRef read_fasta(const std::string& path) {
    std::FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("cannot open " + path);
    std::unique_ptr<std::FILE, int (*)(std::FILE*)> close(f, &std::fclose);
    static const auto code = [] {
        std::array<uint8_t, 256> t;
        t.fill(4);
        t['A'] = t['a'] = 0;
        t['C'] = t['c'] = 1;
        t['G'] = t['g'] = 2;
        t['T'] = t['t'] = t['U'] = t['u'] = 3;
        return t;
    }();
    PackedRef r;
    r.word_off.push_back(0);
    uint64_t cur_len = 0, word = 0;
    bool have = false;
    auto flush_word = [&] {
        if (cur_len & 31) r.words.push_back(word);
        word = 0;
    };
    auto finish = [&] {
        if (!have) return;
        flush_word();
        r.len.push_back(cur_len);
        r.word_off.push_back(r.words.size());
    };
    auto line = [&](const char* p, size_t n) {
        if (n && p[0] == '>') {
            finish();
            std::string name, comment;
            bool hc;
            split_header(std::string(p + 1, n - 1), name, comment, hc);
            r.names.push_back(name);
            cur_len = 0;
            word = 0;
            have = true;
            return;
        }
        if (!have) return;
        while (n && ascii_ws((uint8_t)p[n - 1])) --n;
        for (size_t i = 0; i < n; ++i) {
            uint64_t b = code[(unsigned char)p[i]];
            if (b == 4) {
                uint64_t h = cur_len * 0x9e3779b97f4a7c15ULL;
                b = (h ^ (h >> 30)) & 3;
            }
            word |= b << ((cur_len & 31) * 2);
            if ((++cur_len & 31) == 0) {
                r.words.push_back(word);
                word = 0;
            }
        }
    };
    std::vector<char> buf(std::size_t(64) << 20);
    size_t kept = 0;  // bytes of an unfinished line at the front of buf
    for (;;) {
        if (kept == buf.size()) buf.resize(buf.size() * 2);
        size_t got = std::fread(buf.data() + kept, 1, buf.size() - kept, f);
        size_t end = kept + got;
        size_t at = 0;
        for (const char* nl; (nl = (const char*)std::memchr(buf.data() + at, '\n', end - at));) {
            line(buf.data() + at, (size_t)(nl - buf.data()) - at);
            at = (size_t)(nl - buf.data()) + 1;
        }
        if (got == 0) {
            if (at < end) line(buf.data() + at, end - at);
            break;
        }
        std::memmove(buf.data(), buf.data() + at, end - at);
        kept = end - at;
    }
    if (std::ferror(f)) throw std::runtime_error("cannot read " + path);
    finish();
    if (r.names.empty()) throw std::runtime_error("No sequences found in the reference FASTA");
    return r;
}

// src/seeding/parameters.rs:116-159, 181-221
SeedParams seed_params_for_read_length(uint64_t read_length) {
    struct Row { uint64_t canon, thr; int k, s, wmin, wmax; };
    static const Row rows[] = {
        {50, 70, 18, 14, 1, 4},   {75, 90, 20, 16, 1, 6},    {100, 110, 20, 16, 2, 6},
        {125, 135, 20, 16, 3, 8}, {150, 175, 20, 16, 5, 11}, {250, 375, 22, 18, 6, 16},
        {400, ~0ull, 23, 17, 5, 15},
    };
    const Row* r = &rows[6];
    for (const Row& x : rows)
        if (read_length <= x.thr) { r = &x; break; }
    SeedParams p{};
    p.k = r->k;
    p.s = r->s;
    p.t = (r->k - r->s) / 2 + 1;
    p.w_min = r->wmin;
    p.w_max = r->wmax;
    int64_t md = (int64_t)r->canon - 70;
    if (md < 0) md = 0;
    md = std::max<int64_t>(md, r->k);
    md = std::min<int64_t>(md, 255);
    p.max_dist = (int32_t)md;
    p.q = 255;
    p.main_mask = ~0ULL << (9 + 17);
    p.partial_pos = __builtin_ctzll(p.main_mask) - 1;
    p.fwd_main_mask = p.main_mask | (1ULL << p.partial_pos);
    return p;
}

uint32_t profile_for_read_length(uint64_t read_length) {
    static const uint64_t thr[] = {70, 90, 110, 135, 175, 375};
    uint32_t i = 0;
    while (i < 6 && read_length > thr[i]) ++i;
    return i + 1;
}

int pick_bits(const PackedRef& ref, const SeedParams& p) {
    uint64_t total = 0;
    for (uint64_t l : ref.len) total += l;
    uint64_t est = total / (uint64_t)(p.k - p.s + 1) + 1;
    int64_t b = (int64_t)std::log2((double)est);
    b = std::min<int64_t>(std::max<int64_t>(b, 9), 32);
    return (int)b - 1;
}

// src/indexer.rs:18-170, 186-253
void build_index(const PackedRef& ref, SeedParams p, int bits, double filter_fraction, int threads, HostIndex& out) {
    const uint32_t nc = (uint32_t)ref.names.size();
    p.bits = bits;
    std::vector<std::vector<RefEntry>> per(nc);
    parallel_for(threads, nc, [&](size_t c) {
        uint64_t L = ref.len[c];
        if (L < (uint64_t)p.w_max) return;
        if (L >= (1ull << 31)) throw std::runtime_error("contigs longer than 2^31 bp are not supported");
        const uint64_t* w = ref.words.data() + ref.word_off[c];
        std::vector<Syncmer> sy;
        sy.reserve(L / 4 + 16);
        syncmer_scan((uint32_t)L, p, [&](uint32_t i) { return (uint32_t)(w[i >> 5] >> ((i & 31) * 2)) & 3; },
                     [&](uint64_t h, bool fwd, uint32_t pos) { sy.push_back(Syncmer{h, (pos << 1) | (fwd ? 1u : 0u), 0}); });
        std::vector<RefEntry>& v = per[c];
        v.resize(sy.size());
        const uint32_t n = (uint32_t)sy.size();
        for (uint32_t a = 0; a < n; ++a) {
            uint32_t b = pick_strobe2(a, n, p, [&](uint32_t i) { return sy[i]; });
            uint64_t h = randstrobe_hash(sy[a].hash, sy[b].hash, sy[a].pos_fwd & 1, sy[b].pos_fwd & 1, p);
            uint32_t p1 = sy[a].pos_fwd >> 1, p2 = sy[b].pos_fwd >> 1;
            v[a] = RefEntry{(h & REF_HASH_MASK) | (uint64_t)((p2 - p1) & 0xFF), p1, (uint32_t)c};
        }
    });
    // Sort: scatter every contig's entries into 4096 partitions by the top 12
    // hash bits (the leading part of the sort key), then sort the partitions
    // independently. The key (hash_offset, position, ref_index) is a total order.
    constexpr int PB = 12;
    constexpr size_t NP = size_t(1) << PB;
    auto part = [](const RefEntry& e) { return (size_t)(e.hash_offset >> (64 - PB)); };
    std::vector<std::vector<uint64_t>> cnt(nc, std::vector<uint64_t>(NP, 0));
    parallel_for(threads, nc, [&](size_t c) {
        for (const RefEntry& e : per[c]) ++cnt[c][part(e)];
    });
    std::vector<uint64_t> pstart(NP + 1, 0);
    size_t total = 0;
    for (size_t q = 0; q < NP; ++q) {
        pstart[q] = total;
        for (uint32_t c = 0; c < nc; ++c) {
            uint64_t n = cnt[c][q];
            cnt[c][q] = total;  // this contig's first slot in partition q
            total += n;
        }
    }
    pstart[NP] = total;
    auto& rs = out.entries;
    rs.clear();
    rs.resize(total);
    parallel_for(threads, nc, [&](size_t c) {
        std::vector<uint64_t>& at = cnt[c];
        for (const RefEntry& e : per[c]) rs[at[part(e)]++] = e;
        std::vector<RefEntry>().swap(per[c]);
    });
    auto less = [](const RefEntry& a, const RefEntry& b) {
        if (a.hash_offset != b.hash_offset) return a.hash_offset < b.hash_offset;
        if (a.position != b.position) return a.position < b.position;
        return a.ref_index < b.ref_index;
    };
    parallel_for(threads, NP, [&](size_t q) { std::sort(rs.begin() + pstart[q], rs.begin() + pstart[q + 1], less); });
    // bucket directory exactly as src/indexer.rs:77-129 builds it
    std::vector<uint64_t>& bs = out.bucket_starts;
    bs.clear();
    bs.reserve((1ull << bits) + 1);
    std::vector<uint32_t> counts;
    uint64_t unique = rs.empty() ? 0 : 1;
    uint64_t prev = rs.empty() ? 0 : (rs[0].hash_offset & REF_HASH_MASK);
    uint64_t count = 1;
    if (!rs.empty()) bs.push_back(0);
    for (size_t pos = 1; pos < rs.size(); ++pos) {
        uint64_t cur = rs[pos].hash_offset & REF_HASH_MASK;
        if (cur == prev) { ++count; continue; }
        ++unique;
        if (count != 1) counts.push_back((uint32_t)std::min<uint64_t>(count, 0xFFFFFFFFu));
        count = 1;
        uint64_t top = cur >> (64 - bits);
        while (bs.size() <= top) bs.push_back(pos);
        prev = cur;
    }
    if (!rs.empty() && count != 1) counts.push_back((uint32_t)std::min<uint64_t>(count, 0xFFFFFFFFu));
    while (bs.size() < (1ull << bits) + 1) bs.push_back(rs.size());
    std::sort(counts.begin(), counts.end(), std::greater<uint32_t>());
    uint64_t index_cutoff = (uint64_t)((double)unique * filter_fraction);
    uint64_t fc = index_cutoff < counts.size() ? counts[index_cutoff] : (counts.empty() ? 30 : counts.back());
    p.filter_cutoff = std::min<uint64_t>(std::max<uint64_t>(fc, 30), 100);
    p.n_entries = rs.size();
    out.p = p;
}

// The whole file; gzip-compressed (also several members one after another)
// when the name ends in ".gz" (src/io/xopen.rs:9-20).
bool read_file_bytes(const std::string& path, std::string& out, std::string& err) {
    out.clear();
    if (path.size() > 3 && path.compare(path.size() - 3, 3, ".gz") == 0) {
        gzFile g = gzopen(path.c_str(), "rb");
        if (!g) { err = "cannot open " + path; return false; }
        gzbuffer(g, 1 << 20);
        std::vector<char> buf(1 << 22);
        for (;;) {
            const int k = gzread(g, buf.data(), (unsigned)buf.size());
            if (k < 0) {
                int code = 0;
                err = std::string("gzip error in ") + path + ": " + gzerror(g, &code);
                gzclose(g);
                return false;
            }
            if (k == 0) break;
            out.append(buf.data(), (size_t)k);
        }
        gzclose(g);
        return true;
    }
    std::FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) { err = "cannot open " + path; return false; }
    std::fseek(f, 0, SEEK_END);
    const long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (n > 0) {
        out.resize((size_t)n);
        const size_t got = std::fread(&out[0], 1, (size_t)n, f);
        out.resize(got);
    }
    std::fclose(f);
    return true;
}

std::string fq_error(const FqLoc& r, const std::string& name) {
    switch (r.err) {
        case FQ_NO_AT: return "FASTQ record must start with '@'";
        case FQ_EMPTY_NAME: return "FASTQ record identifier is empty";
        case FQ_LENGTHS: return "sequence and quality lengths differ in record " + name;
        default: return "";
    }
}

// Records from their locations in the file's bytes (uppercase sequences);
// fails with the first record's error, if any.
bool records_from_locs(const std::string& raw, const std::vector<FqLoc>& locs, std::vector<Record>& out,
                       std::string& err, int threads) {
    for (const FqLoc& r : locs)
        if (r.err) {
            err = fq_error(r, raw.substr(r.name_off, r.name_len));
            return false;
        }
    const size_t base = out.size();
    out.resize(base + locs.size());
    const uint8_t* p = (const uint8_t*)raw.data();
    parallel_for(threads, (locs.size() + 4095) / 4096, [&](size_t blk) {
        const size_t e = std::min(locs.size(), (blk + 1) * 4096);
        for (size_t i = blk * 4096; i < e; ++i) {
            const FqLoc& l = locs[i];
            Record& r = out[base + i];
            r.name.assign(raw, l.name_off, l.name_len);
            r.has_comment = l.has_comment != 0;
            if (r.has_comment) r.comment.assign(raw, l.comment_off, l.comment_len);
            r.seq.resize(l.seq_len);
            for (uint32_t x = 0; x < l.seq_len; ++x) r.seq[x] = (char)ascii_upper(p[l.seq_off + x]);
            r.qual.assign(raw, l.qual_off, l.qual_len);
            r.has_qual = true;
            r.end = l.end;
        }
    });
    return true;
}

// FASTQ records (src/io/fastq.rs:22-95, fq_record) or FASTA reads
// (src/io/fasta.rs:60-113; src/io/reads.rs:121-135 picks FASTA when the first
// byte is '>') of a file held in memory.
bool parse_records(const std::string& raw, std::vector<Record>& out, std::string& err, int threads) {
    const uint8_t* p = (const uint8_t*)raw.data();
    const uint64_t n = raw.size();
    if (n > 0 && p[0] == '>') {
        Record cur;
        bool have = false;
        uint64_t s = 0;
        while (s < n) {
            const void* q = std::memchr(p + s, '\n', n - s);
            const uint64_t e = q ? (uint64_t)((const uint8_t*)q - p) + 1 : n;
            if (p[s] == '>') {
                if (have) out.push_back(std::move(cur));
                cur = Record();
                split_header(std::string((const char*)p + s + 1, e - s - 1), cur.name, cur.comment, cur.has_comment);
                cur.has_qual = false;
                have = true;
            } else if (have) {
                uint64_t t = e;
                while (t > s && ascii_ws(p[t - 1])) --t;
                for (uint64_t x = s; x < t; ++x) cur.seq.push_back((char)ascii_upper(p[x]));
            }
            s = e;
        }
        if (have) out.push_back(std::move(cur));
        return true;
    }
    // newline table, then each record from its four lines
    std::vector<uint64_t> nl;
    nl.reserve(n / 64 + 4);
    for (uint64_t s = 0; s < n;) {
        const void* q = std::memchr(p + s, '\n', n - s);
        if (!q) break;
        nl.push_back((uint64_t)((const uint8_t*)q - p));
        s = nl.back() + 1;
    }
    const uint64_t nrec = fq_records(nl.size(), n, n > 0 && p[n - 1] == '\n');
    std::vector<FqLoc> locs(nrec);
    parallel_for(threads, (nrec + 4095) / 4096, [&](size_t blk) {
        const uint64_t e = std::min<uint64_t>(nrec, (blk + 1) * 4096);
        for (uint64_t r = blk * 4096; r < e; ++r) {
            FqLine l[4];
            for (int j = 0; j < 4; ++j) l[j] = fq_line(nl.data(), nl.size(), n, 4 * r + (uint64_t)j);
            locs[r] = fq_record(p, l);
        }
    });
    return records_from_locs(raw, locs, out, err, threads);
}

bool read_records(const std::string& path, std::vector<Record>& out, std::string& err) {
    std::string raw;
    if (!read_file_bytes(path, raw, err)) return false;
    return parse_records(raw, out, err, (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
}

// src/main.rs:1001-1012
uint64_t estimate_read_length(const std::vector<Record>& recs) {
    uint64_t s = 0, n = 0;
    for (size_t i = 0; i < recs.size() && i < 500; ++i) { s += recs[i].seq.size(); ++n; }
    return n ? s / n : 0;
}

std::string sam_header(const PackedRef& ref, const Options& o) {
    std::string h = "@HD\tVN:1.6\tSO:unsorted\n";
    for (size_t i = 0; i < ref.names.size(); ++i) h += "@SQ\tSN:" + ref.names[i] + "\tLN:" + std::to_string(ref.len[i]) + "\n";
    if (o.has_rg) {
        h += "@RG\tID:" + o.rg_id;
        for (const auto& f : o.rg_fields) h += "\t" + f;
        h += "\n";
    }
    return h;
}

}  // namespace sa
