// Oracle: seeding, packed reference, index build and lookups.
// TEST INFRASTRUCTURE ONLY (see oracle.hpp).
#include "oracle.hpp"

#include <parallel/algorithm>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <stdexcept>

namespace orc {

// src/seeding/hash.rs:3-27: XXH64 of the 8 little-endian bytes of `input`, seed 0.
uint64_t xxh64(uint64_t input) {
    const uint64_t P1 = 0x9E3779B185EBCA87ULL, P2 = 0xC2B2AE3D27D4EB4FULL, P3 = 0x165667B19E3779F9ULL,
                   P4 = 0x85EBCA77C2B2AE63ULL, P5 = 0x27D4EB2F165667C5ULL;
    auto rotl = [](uint64_t x, int r) { return (x << r) | (x >> (64 - r)); };
    uint64_t acc = P5 + 8;
    acc ^= rotl(input * P2, 31) * P1;
    acc = rotl(acc, 27) * P1 + P4;
    acc ^= acc >> 33;
    acc *= P2;
    acc ^= acc >> 29;
    acc *= P3;
    acc ^= acc >> 32;
    return acc;
}

// src/seeding/parameters.rs:116-159 (read-length profile table) and 181-221.
SeedingParams seeding_params_for_read_length(size_t read_length) {
    struct Row { int canon; size_t thr; int k, s, wmin, wmax; };
    static const Row rows[] = {
        {50, 70, 18, 14, 1, 4},   {75, 90, 20, 16, 1, 6},    {100, 110, 20, 16, 2, 6},
        {125, 135, 20, 16, 3, 8}, {150, 175, 20, 16, 5, 11}, {250, 375, 22, 18, 6, 16},
        {400, SIZE_MAX, 23, 17, 5, 15},
    };
    const Row* r = nullptr;
    for (const Row& x : rows)
        if (read_length <= x.thr) { r = &x; break; }
    // The profile is keyed by the canonical length; parameters come from the row
    // whose threshold covers that canonical length (the same row).
    SeedingParams p;
    p.profile = r->canon;
    p.syncmer = {r->k, r->s, (r->k - r->s) / 2 + 1};
    long md = (long)r->canon - 70;
    if (md < 0) md = 0;
    md = std::max<long>(md, r->k);
    md = std::min<long>(md, 255);
    const int aux_len = 17;
    uint64_t main_mask = ~0ULL << (9 + aux_len);
    p.randstrobe.w_min = r->wmin;
    p.randstrobe.w_max = r->wmax;
    p.randstrobe.q = (1ULL << 8) - 1;
    p.randstrobe.max_dist = (int)md;
    p.randstrobe.main_hash_mask = main_mask;
    p.randstrobe.partial_orientation_pos = __builtin_ctzll(main_mask) - 1;
    p.randstrobe.forward_main_hash_mask = main_mask | (1ULL << p.randstrobe.partial_orientation_pos);
    return p;
}

// src/seeding/syncmers.rs:141-237. `codes` holds 0..3 for ACGT and 4 for anything
// else; a 4 restarts the rolling state.
std::vector<Syncmer> syncmers_codes(const uint8_t* codes, size_t len, const SyncmerParams& p) {
    std::vector<Syncmer> out;
    const int k = p.k, s = p.s, t = p.t;
    const uint64_t kmask = (k == 32) ? ~0ULL : ((1ULL << (2 * k)) - 1);
    const uint64_t smask = (1ULL << (2 * s)) - 1;
    const int kshift = (k - 1) * 2, sshift = (s - 1) * 2;
    uint64_t xk0 = 0, xk1 = 0, xs0 = 0, xs1 = 0;
    size_t l = 0;
    std::vector<uint64_t> qs;  // window of s-mer hashes (front at index 0)
    qs.reserve(64);
    uint64_t qs_min = UINT64_MAX;
    const size_t wlen = (size_t)(k - s + 1);
    for (size_t i = 0; i < len; ++i) {
        uint8_t c = codes[i];
        if (c < 4) {
            xk0 = ((xk0 << 2) | c) & kmask;
            xk1 = (xk1 >> 2) | ((uint64_t)(3 - c) << kshift);
            xs0 = ((xs0 << 2) | c) & smask;
            xs1 = (xs1 >> 2) | ((uint64_t)(3 - c) << sshift);
            ++l;
            if (l < (size_t)s) continue;
            uint64_t hs = xxh64(std::min(xs0, xs1));
            qs.push_back(hs);
            if (qs.size() < wlen) continue;
            if (qs.size() == wlen) {
                for (uint64_t v : qs) qs_min = std::min(qs_min, v);
            } else {
                uint64_t front = qs.front();
                qs.erase(qs.begin());
                if (front == qs_min) {
                    qs_min = UINT64_MAX;
                    for (uint64_t v : qs) qs_min = std::min(qs_min, v);
                } else if (hs < qs_min) {
                    qs_min = hs;
                }
            }
            if (qs[t - 1] == qs_min) {
                uint64_t yk = std::min(xk0, xk1);
                out.push_back({xxh64(yk), xk0 <= xk1, i + 1 - k});
            }
        } else {
            qs_min = UINT64_MAX;
            l = 0;
            xs0 = xs1 = xk0 = xk1 = 0;
            qs.clear();
        }
    }
    return out;
}

// src/seeding/syncmers.rs:14-40 NUCLEOTIDES: ACGTU (any case) -> 0..3, else 4.
static uint8_t nt_code(uint8_t c) {
    switch (c) {
        case 'A': case 'a': return 0;
        case 'C': case 'c': return 1;
        case 'G': case 'g': return 2;
        case 'T': case 't': case 'U': case 'u': return 3;
        default: return 4;
    }
}

std::vector<Syncmer> syncmers_ascii(const uint8_t* seq, size_t len, const SyncmerParams& p) {
    std::vector<uint8_t> codes(len);
    for (size_t i = 0; i < len; ++i) codes[i] = nt_code(seq[i]);
    return syncmers_codes(codes.data(), len, p);
}

// src/seeding/strobes.rs:112-124. Bit 8 (strobe-2 orientation) is cleared by the final mask.
uint64_t randstrobe_hash(uint64_t h1, uint64_t h2, uint64_t f1, uint64_t f2, const RandstrobeParams& p) {
    return ((h1 & p.main_hash_mask) | (h2 & ~p.forward_main_hash_mask) |
            (f1 << p.partial_orientation_pos) | (f2 << 8)) &
           (0xFFFFFFFFFFFFFF00ULL << 1);
}

// src/seeding/strobes.rs:145-176 (strict '<' keeps the first minimum)
std::vector<Randstrobe> randstrobes_from_syncmers(const std::vector<Syncmer>& sy, const RandstrobeParams& p) {
    std::vector<Randstrobe> out;
    out.reserve(sy.size());
    for (size_t a = 0; a < sy.size(); ++a) {
        size_t avail = std::min(sy.size() - a, (size_t)p.w_max + 1);  // deque holds <= w_max+1
        const Syncmer& s1 = sy[a];
        size_t max_pos = s1.position + (size_t)p.max_dist;
        uint64_t min_val = UINT64_MAX;
        const Syncmer* s2 = &sy[a];
        for (size_t i = (size_t)p.w_min; i < avail; ++i) {
            const Syncmer& c = sy[a + i];
            if (c.position > max_pos) break;
            uint64_t ones = (uint64_t)__builtin_popcountll((s1.hash ^ c.hash) & p.q);
            if (ones < min_val) { min_val = ones; s2 = &c; }
        }
        Randstrobe r;
        r.hash = randstrobe_hash(s1.hash, s2->hash, s1.is_forward, s2->is_forward, p);
        r.hash_revcomp = randstrobe_hash(s2->hash, s1.hash, s1.is_forward ^ 1, s2->is_forward ^ 1, p);
        r.strobe1_pos = s1.position;
        r.strobe2_pos = s2->position;
        out.push_back(r);
    }
    return out;
}

// src/seeding/mod.rs:19-75
void randstrobes_query(const uint8_t* seq, size_t len, const SeedingParams& p,
                       std::vector<QueryRandstrobe> out[2]) {
    out[0].clear();
    out[1].clear();
    if (len < (size_t)p.randstrobe.w_max) return;
    std::vector<Syncmer> sy = syncmers_ascii(seq, len, p.syncmer);
    for (const Randstrobe& r : randstrobes_from_syncmers(sy, p.randstrobe))
        out[0].push_back({r.hash, r.hash_revcomp, r.strobe1_pos, r.strobe2_pos + (size_t)p.syncmer.k});
    std::reverse(sy.begin(), sy.end());
    for (Syncmer& s : sy) {
        s.position = len - s.position - (size_t)p.syncmer.k;
        s.is_forward = !s.is_forward;
    }
    for (const Randstrobe& r : randstrobes_from_syncmers(sy, p.randstrobe))
        out[1].push_back({r.hash, r.hash_revcomp, r.strobe1_pos, r.strobe2_pos + (size_t)p.syncmer.k});
}

// src/packed_seq.rs:56-92
void PackedSeq::push(uint8_t c) {
    uint64_t b;
    switch (c) {
        case 'A': case 'a': b = 0; break;
        case 'C': case 'c': b = 1; break;
        case 'G': case 'g': b = 2; break;
        case 'T': case 't': case 'U': case 'u': b = 3; break;
        default: {
            uint64_t h = (uint64_t)len * 0x9e3779b97f4a7c15ULL;
            b = (h ^ (h >> 30)) & 3;
        }
    }
    size_t bit = (len & 31) * 2;
    if (bit == 0) data.push_back(b);
    else data.back() |= b << bit;
    ++len;
}

std::string PackedSeq::decode(size_t start, size_t end) const {
    std::string s;
    s.resize(end > start ? end - start : 0);
    for (size_t i = start; i < end; ++i) s[i - start] = decode_at(i);
    return s;
}

// u8::is_ascii_whitespace, which trim_ascii_end uses (no vertical tab)
static bool ascii_ws(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\f'; }

static void trim_end(std::string& s) {
    while (!s.empty() && ascii_ws(s.back())) s.pop_back();
}

// src/io/mod.rs:32-38 split_header
static void split_header(const std::string& h0, std::string& name, std::string& comment, bool& has_comment) {
    std::string h = h0;
    trim_end(h);
    size_t p = h.find_first_of(" \t");
    if (p == std::string::npos) { name = h; comment.clear(); has_comment = false; }
    else { name = h.substr(0, p); comment = h.substr(p + 1); has_comment = true; }
}

// src/io/fasta.rs:116-165 (no gzip support in the oracle)
std::vector<RefSeq> read_fasta(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::vector<RefSeq> refs;
    std::string line;
    bool have = false;
    while (std::getline(f, line)) {
        if (!line.empty() && line[0] == '>') {
            std::string name, comment;
            bool hc;
            split_header(line.substr(1), name, comment, hc);
            refs.push_back(RefSeq{name, PackedSeq{}});
            have = true;
        } else if (have) {
            trim_end(line);
            for (unsigned char c : line) refs.back().seq.push(c);
        }
    }
    return refs;
}

// BufRead::read_line / read_until(b'\n') over a file held in memory: the next
// line including its '\n' (if it has one); empty at the end of the file.
struct LineReader {
    std::string data;
    size_t at = 0;
    std::string next() {
        if (at >= data.size()) return std::string();
        size_t e = data.find('\n', at);
        e = e == std::string::npos ? data.size() : e + 1;
        std::string l = data.substr(at, e - at);
        at = e;
        return l;
    }
};

// src/io/fastq.rs:22-95 and src/io/fasta.rs:67-114 (reads given as FASTA;
// src/io/reads.rs:121-135 picks it when the first byte is '>')
bool read_reads(const std::string& path, std::vector<Record>& out, std::string& err) {
    std::ifstream f(path, std::ios::binary);
    if (!f) { err = "cannot open " + path; return false; }
    LineReader rd;
    rd.data.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    if (!rd.data.empty() && rd.data[0] == '>') {
        Record cur;
        bool have = false;
        for (std::string line = rd.next(); !line.empty(); line = rd.next()) {
            if (line[0] == '>') {
                if (have) out.push_back(cur);
                cur = Record();
                split_header(line.substr(1), cur.name, cur.comment, cur.has_comment);
                cur.has_qual = false;
                have = true;
            } else if (have) {
                trim_end(line);
                for (char& c : line) c = (c >= 'a' && c <= 'z') ? (char)(c - 32) : c;
                cur.seq += line;
            }
        }
        if (have) out.push_back(cur);
        return true;
    }
    // sequence and quality lines lose their last byte (the '\n', or the last
    // character of an unterminated last line), then one '\r'
    auto content = [](std::string l) {
        if (!l.empty()) l.pop_back();
        if (!l.empty() && l.back() == '\r') l.pop_back();
        return l;
    };
    for (std::string line = rd.next(); !line.empty(); line = rd.next()) {
        if (line[0] != '@') { err = "Record must start with '@'"; return false; }
        Record r;
        split_header(line.substr(1), r.name, r.comment, r.has_comment);
        if (r.name.empty()) { err = "Record identifier is empty"; return false; }
        if (r.name.size() >= 2 && r.name.compare(r.name.size() - 2, 2, "/1") == 0) { r.name.resize(r.name.size() - 2); r.end = 1; }
        else if (r.name.size() >= 2 && r.name.compare(r.name.size() - 2, 2, "/2") == 0) { r.name.resize(r.name.size() - 2); r.end = 2; }
        std::string seq = content(rd.next());
        rd.next();  // the '+' line
        std::string qual = content(rd.next());
        if (seq.size() != qual.size()) { err = "sequence/quality length mismatch in " + r.name; return false; }
        for (char& c : seq) c = (c >= 'a' && c <= 'z') ? (char)(c - 32) : c;
        r.seq = seq;
        r.qual = qual;
        out.push_back(r);
    }
    return true;
}

// ---------------------------------------------------------------- index ---

// src/partition.rs:27-86: lower bound (first element for which pred is false).
template <typename Pred>
static size_t partition_point(const RefRandstrobe* s, size_t n, Pred pred) {
    size_t lo = 0, hi = n;
    while (lo < hi) {
        size_t mid = lo + (hi - lo) / 2;
        if (pred(s[mid])) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// src/index.rs:159-197
size_t Index::get_masked_from(uint64_t hash, uint64_t mask, size_t start) const {
    uint64_t masked = hash & mask;
    size_t top = (size_t)(hash >> (64 - bits));
    size_t ps = (start == SIZE_MAX) ? (size_t)bucket_starts[top] : start;
    size_t pe = (size_t)bucket_starts[top + 1];
    if (ps >= pe) return SIZE_MAX;
    size_t n = pe - ps;
    if (n < 4) {
        for (size_t i = ps; i < pe; ++i) {
            uint64_t h = randstrobes[i].hash() & mask;
            if (h == masked) return i;
            if (h > masked) return SIZE_MAX;
        }
        return SIZE_MAX;
    }
    size_t pos = partition_point(&randstrobes[ps], n, [&](const RefRandstrobe& r) { return (r.hash() & mask) < masked; });
    if (pos < n && (randstrobes[ps + pos].hash() & mask) == masked) return ps + pos;
    return SIZE_MAX;
}

// src/index.rs:272-295
size_t Index::get_count(size_t position, uint64_t mask) const {
    uint64_t key = randstrobes[position].hash();
    uint64_t mk = key & mask;
    size_t top = (size_t)(key >> (64 - bits));
    size_t pe = (size_t)bucket_starts[top + 1];
    if (pe - position < 8) {
        size_t count = 1;
        for (size_t i = position + 1; i < pe; ++i) {
            if ((randstrobes[i].hash() & mask) == mk) ++count; else break;
        }
        return count;
    }
    return partition_point(&randstrobes[position], pe - position,
                           [&](const RefRandstrobe& r) { return (r.hash() & mask) <= mk; });
}

// src/index.rs:254-266
size_t Index::get_count_full(size_t position, uint64_t hash_revcomp) const {
    size_t rc = get_full_forward(hash_revcomp);
    size_t rev = rc == SIZE_MAX ? 0 : get_count(rc, 0xFFFFFFFFFFFFFF00ULL);
    return rev + get_count(position, 0xFFFFFFFFFFFFFF00ULL);
}

// src/index.rs:298-307 (note: guard uses filter_cutoff, index uses cutoff)
bool Index::is_too_frequent_forward(size_t position, size_t cutoff) const {
    if (position + filter_cutoff < randstrobes.size())
        return randstrobes[position].hash() == randstrobes[position + cutoff].hash();
    return false;
}

// src/index.rs:309-323
bool Index::is_too_frequent(size_t position, size_t cutoff, uint64_t hash_revcomp) const {
    if (is_too_frequent_forward(position, cutoff)) return true;
    size_t rc = get_full_forward(hash_revcomp);
    if (rc != SIZE_MAX) {
        if (is_too_frequent_forward(rc, cutoff)) return true;
        size_t count = get_count(position, 0xFFFFFFFFFFFFFF00ULL) + get_count(rc, 0xFFFFFFFFFFFFFF00ULL);
        return count > cutoff;
    }
    return false;
}

// src/index.rs:325-338
bool Index::is_too_frequent_partial(size_t position, size_t cutoff) const {
    if (position + cutoff < randstrobes.size()) {
        uint64_t m = params.randstrobe.main_hash_mask;
        return (randstrobes[position].hash() & m) == (randstrobes[position + cutoff].hash() & m);
    }
    return false;
}

// src/indexer.rs:294-300
int pick_bits(const std::vector<RefSeq>& refs, const SyncmerParams& p) {
    size_t total = 0;
    for (const auto& r : refs) total += r.seq.len;
    size_t est = total / (size_t)(p.k - p.s + 1) + 1;
    long b = (long)std::log2((double)est);
    b = std::min<long>(std::max<long>(b, 9), 32);
    return (int)b - 1;
}

// src/indexer.rs:18-170 and 222-253
Index make_index(const std::vector<RefSeq>& refs, const SeedingParams& p, int bits, double filter_fraction) {
    Index idx;
    idx.params = p;
    idx.bits = bits;
    std::vector<RefRandstrobe>& rs = idx.randstrobes;
    // contigs are independent (src/indexer.rs:65-69 generates them in parallel
    // too); the sort key (hash_offset, position, ref_index) is unique per entry,
    // so the sorted table does not depend on the order entries were produced in
    std::vector<std::vector<RefRandstrobe>> per(refs.size());
#pragma omp parallel for schedule(dynamic, 1)
    for (size_t ri = 0; ri < refs.size(); ++ri) {
        const PackedSeq& seq = refs[ri].seq;
        if (seq.len < (size_t)p.randstrobe.w_max) continue;
        std::vector<uint8_t> codes(seq.len);
        for (size_t i = 0; i < seq.len; ++i) codes[i] = seq.bits(i);
        std::vector<Syncmer> sy = syncmers_codes(codes.data(), seq.len, p.syncmer);
        for (const Randstrobe& r : randstrobes_from_syncmers(sy, p.randstrobe)) {
            uint64_t off = r.strobe2_pos - r.strobe1_pos;
            per[ri].push_back({(r.hash & 0xFFFFFFFFFFFFFF00ULL) | (off & 0xFF), (uint32_t)r.strobe1_pos, (uint32_t)ri});
        }
    }
    size_t total = 0;
    for (const auto& v : per) total += v.size();
    rs.reserve(total);
    for (auto& v : per) {
        rs.insert(rs.end(), v.begin(), v.end());
        std::vector<RefRandstrobe>().swap(v);
    }
    __gnu_parallel::sort(rs.begin(), rs.end(), [](const RefRandstrobe& a, const RefRandstrobe& b) {
        if (a.hash_offset != b.hash_offset) return a.hash_offset < b.hash_offset;
        if (a.position != b.position) return a.position < b.position;
        return a.ref_index < b.ref_index;
    });
    std::vector<uint64_t>& bs = idx.bucket_starts;
    bs.reserve(((size_t)1 << bits) + 1);
    std::vector<size_t> counts;
    uint64_t unique = rs.empty() ? 0 : 1;
    uint64_t prev = rs.empty() ? 0 : rs[0].hash();
    size_t count = 1;
    if (!rs.empty()) bs.push_back(0);
    for (size_t pos = 1; pos < rs.size(); ++pos) {
        uint64_t cur = rs[pos].hash();
        if (cur == prev) { ++count; continue; }
        ++unique;
        if (count != 1) counts.push_back(count);
        count = 1;
        uint64_t top = cur >> (64 - bits);
        while (bs.size() <= top) bs.push_back(pos);
        prev = cur;
    }
    if (!rs.empty() && count != 1) counts.push_back(count);
    while (bs.size() < ((size_t)1 << bits) + 1) bs.push_back(rs.size());
    std::sort(counts.begin(), counts.end(), std::greater<size_t>());
    size_t index_cutoff = (size_t)((double)unique * filter_fraction);
    size_t fc = index_cutoff < counts.size() ? counts[index_cutoff] : (counts.empty() ? 30 : counts.back());
    idx.filter_cutoff = std::min<size_t>(std::max<size_t>(fc, 30), 100);
    return idx;
}

}  // namespace orc
