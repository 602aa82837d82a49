// Read-level decisions of src/mapper.rs and src/maponly.rs, replayed on the
// host over device results.
//
// The device computes NAMs and every base-level extension. The decisions that
// depend on chunk-sequential state -- the per-chunk random generator (fastrand
// Rng, src/main.rs:878) and insert-size estimate (src/main.rs:879) -- are taken
// here, in input order, one host thread per chunk. An extension the replay needs
// but the device has not computed yet is a "miss": the chunk stops at that
// record, the remaining records are replayed in probe mode to predict their
// misses too, all misses of all chunks go to the device as one batch, and the
// chunk resumes. Output is only written by a replay without misses, so it is
// identical to a sequential run.
#include <algorithm>
#include <sched.h>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <array>
#include <immintrin.h>
#include <deque>
#include <exception>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <stdexcept>
#include <thread>
#include <unordered_map>

#include "sa_host.hpp"

namespace sa {

namespace {

const uint16_t PAIRED = 1, PROPER_PAIR = 2, UNMAP = 4, MUNMAP = 8, REVERSE = 0x10, MREVERSE = 0x20, READ1 = 0x40,
               READ2 = 0x80, SECONDARY = 0x100;  // src/io/sam.rs:7-18

// ------------------------------------------------------------------ rng ---
// fastrand 2.3.0 Rng: wyrand (v4.2 constants) with Lemire bounded sampling.
struct Rng {
    uint64_t s = 0;
    uint64_t u64() {
        s += 0x2d358dccaa6c78a5ULL;
        unsigned __int128 t = (unsigned __int128)s * (unsigned __int128)(s ^ 0x8bb84b93962eacc9ULL);
        return (uint64_t)t ^ (uint64_t)(t >> 64);
    }
    uint32_t mod32(uint32_t n) {
        uint32_t r = (uint32_t)u64();
        uint64_t m = (uint64_t)r * n;
        uint32_t hi = (uint32_t)(m >> 32), lo = (uint32_t)m;
        if (lo < n) {
            uint32_t t = (uint32_t)(-n) % n;
            while (lo < t) {
                r = (uint32_t)u64();
                m = (uint64_t)r * n;
                hi = (uint32_t)(m >> 32);
                lo = (uint32_t)m;
            }
        }
        return hi;
    }
    uint64_t mod64(uint64_t n) {
        uint64_t r = u64();
        unsigned __int128 m = (unsigned __int128)r * n;
        uint64_t hi = (uint64_t)(m >> 64), lo = (uint64_t)m;
        if (lo < n) {
            uint64_t t = (uint64_t)(-n) % n;
            while (lo < t) {
                r = u64();
                m = (unsigned __int128)r * n;
                hi = (uint64_t)(m >> 64);
                lo = (uint64_t)m;
            }
        }
        return hi;
    }
};

// src/insertsize.rs:3-50
struct InsertSize {
    uint64_t sample_size = 0;
    float mu = 300.0f, sigma = 100.0f, variance = 10000.0f, sse = 10000.0f;
    void update(uint64_t x0) {
        if (x0 >= 2000) return;
        if (sample_size > 0) {
            float x = (float)x0;
            float e = x - mu;
            mu += e / (float)sample_size;
            sse += e * (x - mu);
            if (sample_size > 1) variance = sse / ((float)sample_size - 1.0f);
            else variance = sse;
            sigma = std::sqrt(variance);
        }
        sample_size += 1;
    }
};

// src/math.rs:1-7
float normal_pdf(float x, float mu, float sigma) {
    float a = (x - mu) / sigma;
    return (float)0.3989422804014327 / sigma * std::exp(-0.5f * a * a);
}

struct Nam {
    uint32_t nam_id, ref_id, ref_start, ref_end, query_start, query_end, matching_bases, n_anchors;
    float score;
    bool is_revcomp;
    uint8_t cons, rev;  // is_consistent as is; reverse_nam_if_needed outcome (0 as is, 1 flip, 2 none)
    uint64_t anchor_off;
    uint32_t projected_ref_start() const { return ref_start > query_start ? ref_start - query_start : 0; }
};

Nam to_nam(const NamRec& r) {
    return Nam{r.nam_id, r.ref_id, r.ref_start, r.ref_end, r.query_start, r.query_end, r.matching_bases, r.n_anchors,
               r.score, (r.flags & 1) != 0, (uint8_t)((r.flags >> 1) & 1), (uint8_t)((r.flags >> 2) & 3), r.anchor_off};
}

// src/nam.rs:94-135 using the device's consistency verdict
bool reverse_nam_if_needed(Nam& n, uint32_t L) {
    if (n.rev == 0) return true;
    if (n.rev == 1) {
        n.is_revcomp = !n.is_revcomp;
        uint32_t qs = L - n.query_end, qe = L - n.query_start;
        n.query_start = qs;
        n.query_end = qe;
        n.cons = 1;
        n.rev = 0;
        return true;
    }
    return false;
}

struct Alignment {
    uint32_t ref_id = 0, ref_start = 0;
    const uint32_t* cigar = nullptr;  // BAM ops (=, X, I, D)
    uint32_t n_cigar = 0;
    uint32_t edit_distance = 0, soft_clip_left = 0, soft_clip_right = 0, score = 0, length = 0;
    bool is_revcomp = false, gapped = false;
    uint32_t global_ed() const { return edit_distance + soft_clip_left + soft_clip_right; }
};

struct Details {
    ReadNamInfo nam{};
    uint32_t inconsistent_nams = 0, mate_rescue = 0, tried_alignment = 0, gapped = 0, best_alignments = 0;
};

// ------------------------------------------------------------ cache key ---
struct Key {
    uint32_t kind, read, flags, ref_id, a, b, c, d;
    uint64_t e;
    bool operator==(const Key& o) const { return std::memcmp(this, &o, sizeof(Key)) == 0; }
};
struct KeyHash {
    size_t operator()(const Key& k) const {
        const uint64_t* w = (const uint64_t*)&k;
        uint64_t h = 0x9e3779b97f4a7c15ULL;
        for (size_t i = 0; i < sizeof(Key) / 8; ++i) h = (h ^ w[i]) * 0xff51afd7ed558ccdULL, h ^= h >> 29;
        return (size_t)h;
    }
};
static_assert(sizeof(Key) == 40, "Key packing");

// Open-addressing table keyed by Key (linear probing, power-of-two capacity, load
// at most 1/2): the per-chunk extension cache and pending set take ~10^4 inserts
// per chunk and round, which a node-per-entry map pays for with an allocation each.
template <typename V>
class KeyTable {
public:
    V* find(const Key& k) {
        if (n_ == 0) return nullptr;
        for (size_t i = KeyHash{}(k) & mask_;; i = (i + 1) & mask_) {
            if (!used_[i]) return nullptr;
            if (slots_[i].k == k) return &slots_[i].v;
        }
    }
    size_t count(const Key& k) { return find(k) ? 1 : 0; }
    // inserts unless present; true if inserted
    bool emplace(const Key& k, const V& v) {
        if (2 * (n_ + 1) > cap_) grow(2 * (n_ + 1));
        for (size_t i = KeyHash{}(k) & mask_;; i = (i + 1) & mask_) {
            if (!used_[i]) {
                used_[i] = 1;
                slots_[i].k = k;
                slots_[i].v = v;
                ++n_;
                return true;
            }
            if (slots_[i].k == k) return false;
        }
    }
    void reserve(size_t n) {
        if (2 * n > cap_) grow(2 * n);
    }
    void clear() {
        if (n_) std::memset(used_.get(), 0, cap_);
        n_ = 0;
    }
    void swap(KeyTable& o) {
        slots_.swap(o.slots_);
        used_.swap(o.used_);
        std::swap(n_, o.n_);
        std::swap(cap_, o.cap_);
        std::swap(mask_, o.mask_);
    }
    size_t size() const { return n_; }

private:
    struct Slot {
        Key k;
        V v;
    };
    void grow(size_t want) {
        size_t c = 16;
        while (c < want) c <<= 1;
        std::unique_ptr<Slot[]> os(new Slot[c]);  // left uninitialised: `used_` says which are live
        std::unique_ptr<uint8_t[]> ou(new uint8_t[c]());
        os.swap(slots_);
        ou.swap(used_);
        const size_t oc = cap_;
        cap_ = c;
        mask_ = c - 1;
        n_ = 0;
        for (size_t i = 0; i < oc; ++i)
            if (ou[i]) emplace(os[i].k, os[i].v);
    }
    std::unique_ptr<Slot[]> slots_;
    std::unique_ptr<uint8_t[]> used_;
    size_t n_ = 0, cap_ = 0, mask_ = 0;
};


struct CachedResult {
    ExtResult r;
    uint64_t cigar_off;  // into Chunk::cigar_pool
};
// A result a decision found: the result and its CIGAR words (r null: a miss)
struct ExtHit {
    const ExtResult* r = nullptr;
    const uint32_t* cigar = nullptr;
};

// Chunk output buffers kept for reuse: a fresh multi-megabyte string per chunk
// and batch costs a page fault per 4 KiB the records are written into (about
// 180,000 per million pairs of SAM); a recycled one is already mapped.
class OutPool {
public:
    std::string take(size_t cap) {
        std::string s;
        {
            std::lock_guard<std::mutex> lk(mu_);
            size_t best = free_.size();
            for (size_t i = 0; i < free_.size(); ++i)
                if (free_[i].capacity() >= cap && (best == free_.size() || free_[i].capacity() < free_[best].capacity()))
                    best = i;
            if (best == free_.size() && !free_.empty()) best = free_.size() - 1;
            if (best < free_.size()) {
                s.swap(free_[best]);
                held_ -= s.capacity();
                free_.erase(free_.begin() + (ptrdiff_t)best);
            }
        }
        s.clear();
        s.reserve(cap);
        return s;
    }
    void give(std::string& s) {
        std::string t;
        t.swap(s);
        t.clear();
        std::lock_guard<std::mutex> lk(mu_);
        if (held_ + t.capacity() > kMaxHeld || free_.size() >= 256) return;  // freed when t goes
        held_ += t.capacity();
        free_.push_back(std::move(t));
    }

private:
    static constexpr size_t kMaxHeld = 4ull << 30;
    std::mutex mu_;
    std::vector<std::string> free_;
    size_t held_ = 0;
};
OutPool g_out_pool;

// The replay state a record starts from: the chunk's generator and insert-size
// estimate (bit for bit).
bool same_state(const Rng& a, const InsertSize& ia, const Rng& b, const InsertSize& ib) {
    static_assert(sizeof(InsertSize) == 24, "InsertSize: u64 then four f32, no padding");
    return a.s == b.s && std::memcmp(&ia, &ib, sizeof(InsertSize)) == 0;
}

// A record decided as a probe without a miss: its output and end state are
// exact if the record really starts from the state it was probed with (its
// decisions read only that state and cached results, which never change).
struct Tentative {
    uint64_t r;            // record (relative to the chunk)
    Rng rng0, rng1;        // state it was probed from, and its state after
    InsertSize isz0, isz1;
    uint64_t t0, t1;       // its output text in the pass's tentative buffer
};

// Per-chunk state of the replay.
struct Chunk {
    uint64_t first, count;  // records (pairs) [first, first+count) of the caller's read set
    const Record* rec1 = nullptr;  // the chunk's first record (and its mate)
    const Record* rec2 = nullptr;
    uint64_t pos = 0;  // index of the chunk's first record (pair) within its device batch
    uint64_t next = 0;       // next record to commit (relative)
    Rng rng;
    InsertSize isize;
    std::string out;
    std::vector<double> abundances;
    KeyTable<CachedResult> cache;
    std::vector<uint32_t> cigar_pool;  // CIGARs of the cached results (grows only between passes)
    // the fused pass's results for the chunk's reads, looked up where they lie
    // (its requests are sorted by read): those of read spec_r0 + x are
    // spec_req[spec_at[x] .. spec_at[x + 1]) with their results and CIGARs
    const ExtRequest* spec_req = nullptr;
    const ExtResult* spec_res = nullptr;
    const uint64_t* spec_cigoff = nullptr;
    const uint32_t* spec_cig = nullptr;
    uint64_t spec_r0 = 0;
    std::vector<uint32_t> spec_at;
    std::vector<uint8_t> spec_seen;  // looked up by a decision (counted once)
    std::vector<Key> pending;
    KeyTable<char> pending_set;
    uint64_t spec_used = 0;  // fused-pass results the decisions looked up
    // probes of the last pass without a miss (in record order) and their text;
    // the current pass writes the next ones
    std::vector<Tentative> tent_prev, tent_cur;
    std::string text_prev, text_cur;
    // replay cost: records decided and committed, records decided as probes,
    // probes taken over unchanged, and the nanoseconds spent on each and on
    // writing output text
    uint64_t n_commit = 0, n_probe = 0, n_accept = 0, ns_commit = 0, ns_probe = 0, ns_write = 0;
    bool done() const { return next >= count; }
};

struct Ctx {
    const Options& o;
    const PackedRef& ref;
    const NamBatch& nb;
    const Record* r1;  // records of the current chunk (indexed from its first record)
    const Record* r2;
    bool paired;
    uint64_t rel0;  // batch index of the chunk's first record
    uint32_t k;
    Chunk* ch;
    bool missed;   // a lookup of this record had no result yet (it is now pending)
    bool abandon;  // the record stopped early: its state changes are discarded
    Rng* rng;
    InsertSize* isize;
};

// Look up an extension result: a seed extension first among the fused pass's
// results for its read (a few requests, read in order as the records are), then
// in the chunk's cache of requested ones. A missing one is registered as
// pending and reported as no alignment; the record then runs on as a probe that
// collects its other requests, and is not committed in this pass. (A fused-pass
// result that failed is not taken: the decision requests it again.)
ExtHit lookup(Ctx& c, const Key& key) {
    Chunk& ch = *c.ch;
    if (key.kind == EXT_SEED && ch.spec_req && key.read >= ch.spec_r0 && key.read - ch.spec_r0 + 1 < ch.spec_at.size()) {
        const uint64_t x = key.read - ch.spec_r0;
        for (uint32_t y = ch.spec_at[x]; y < ch.spec_at[x + 1]; ++y) {
            const ExtRequest& q = ch.spec_req[y];
            const uint64_t e = (q.flags & 1) ? 0 : (q.anchor_off | (uint64_t)q.n_anchors << 48);
            if (q.flags == key.flags && q.ref_id == key.ref_id && q.ref_start == key.a && q.ref_end == key.b &&
                q.query_start == key.c && q.query_end == key.d && e == key.e && ch.spec_res[y].status == 0) {
                if (!ch.spec_seen[y]) {
                    ch.spec_seen[y] = 1;
                    ch.spec_used++;
                }
                return ExtHit{&ch.spec_res[y], ch.spec_cig + ch.spec_cigoff[y]};
            }
        }
    }
    if (const CachedResult* hit = ch.cache.find(key)) return ExtHit{&hit->r, ch.cigar_pool.data() + hit->cigar_off};
    if (ch.pending_set.emplace(key, 1)) ch.pending.push_back(key);
    c.missed = true;
    return ExtHit{};
}

// src/mapper.rs:532-620 (device request kind EXT_SEED)
std::optional<Alignment> extend_seed(Ctx& c, const Nam& nam, uint32_t read_id, uint32_t L, bool consistent,
                                     bool use_ssw) {
    Key key{EXT_SEED, read_id, (use_ssw ? 1u : 0u) | (consistent ? 2u : 0u) | (nam.is_revcomp ? 4u : 0u), nam.ref_id,
            nam.ref_start, nam.ref_end, nam.query_start, nam.query_end, use_ssw ? 0 : nam.anchor_off};
    if (!use_ssw) key.e |= (uint64_t)nam.n_anchors << 48;
    const ExtHit h = lookup(c, key);
    if (!h.r || !h.r->found) return std::nullopt;
    const ExtResult& r = *h.r;
    Alignment a;
    a.ref_id = nam.ref_id;
    a.ref_start = r.ref_start;
    a.cigar = h.cigar;
    a.n_cigar = r.n_cigar;
    a.edit_distance = r.edit_distance;
    a.soft_clip_left = r.query_start;
    a.soft_clip_right = L - r.query_end;
    a.score = r.score;
    a.length = r.length;
    a.is_revcomp = nam.is_revcomp;
    a.gapped = r.gapped != 0;
    return a;
}

// src/mapper.rs:1094-1153 (device request kind EXT_RESCUE)
std::optional<Alignment> rescue_align(Ctx& c, const Nam& mate, uint32_t read_id, uint32_t L, float mu, float sigma) {
    float w = mu + 5.0f * sigma;
    uint64_t wi = w > 0 ? (uint64_t)w : 0;
    uint64_t rs, re;
    bool rc;
    if (mate.is_revcomp) {
        rc = false;
        uint64_t p = mate.projected_ref_start();
        rs = p > wi ? p - wi : 0;
        re = p + L / 2;
    } else {
        rc = true;
        uint64_t b = (uint64_t)mate.ref_end + L - mate.query_end;
        rs = b > L / 2 ? b - L / 2 : 0;
        re = b + wi;
    }
    uint64_t rl = c.ref.len[mate.ref_id];
    rs = std::min(rs, rl);
    re = std::min(re, rl);
    if (re < rs + c.k) return std::nullopt;
    Key key{EXT_RESCUE, read_id, rc ? 4u : 0u, mate.ref_id, (uint32_t)rs, (uint32_t)re, 0, 0, 0};
    const ExtHit h = lookup(c, key);
    if (!h.r || !h.r->found) return std::nullopt;
    const ExtResult& r = *h.r;
    Alignment a;
    a.ref_id = mate.ref_id;
    a.ref_start = r.ref_start;
    a.cigar = h.cigar;
    a.n_cigar = r.n_cigar;
    a.edit_distance = r.edit_distance;
    a.soft_clip_left = r.query_start;
    a.soft_clip_right = L - r.query_end;
    a.score = r.score;
    a.length = r.length;
    a.is_revcomp = !mate.is_revcomp;
    a.gapped = true;
    return a;
}

// ----------------------------------------------------------------- output ---
void put_u(std::string& s, uint64_t v) {
    char b[24];
    int n = 0;
    do { b[n++] = (char)('0' + v % 10); v /= 10; } while (v);
    while (n) s += b[--n];
}

// Decimal digits of 0..99, two characters each.
struct DigitPairs {
    char d[200];
    DigitPairs() {
        for (int i = 0; i < 100; ++i) {
            d[2 * i] = (char)('0' + i / 10);
            d[2 * i + 1] = (char)('0' + i % 10);
        }
    }
};
const DigitPairs kPairs;

// comp_ascii as a table
struct CompTable {
    char t[256];
    CompTable() {
        for (int c = 0; c < 256; ++c) t[c] = (char)comp_ascii((uint8_t)c);
    }
};
const CompTable kComp;

// Reverse complement of n bytes (comp_ascii of each, last first) and the
// reverse of n bytes, 16 at a time with SSSE3 byte shuffles: a block whose
// bytes are all A, C, G, T, U or N (checked by mapping each byte's low nibble
// back to the letter it must be; the unused nibbles map to a byte with another
// nibble, so they never pass) is complemented by a low-nibble lookup; any other
// block takes the table.
__attribute__((target("ssse3"))) void revcomp_ssse3(char* dst, const char* src, size_t n) {
    const __m128i rev = _mm_setr_epi8(15, 14, 13, 12, 11, 10, 9, 8, 7, 6, 5, 4, 3, 2, 1, 0);
    // low nibble -> the letter with that nibble, and its complement
    const __m128i letter = _mm_setr_epi8(1, 'A', 0, 'C', 'T', 'U', 0, 'G', 0, 0, 0, 0, 0, 0, 'N', 0);
    const __m128i comp = _mm_setr_epi8(0, 'T', 0, 'G', 'A', 'A', 0, 'C', 0, 0, 0, 0, 0, 0, 'N', 0);
    const __m128i lo = _mm_set1_epi8(0x0F);
    size_t x = 0;
    for (; x + 16 <= n; x += 16) {
        const __m128i v = _mm_loadu_si128((const __m128i*)(src + n - 16 - x));
        const __m128i nib = _mm_and_si128(v, lo);
        const __m128i ok = _mm_cmpeq_epi8(_mm_shuffle_epi8(letter, nib), v);
        if (_mm_movemask_epi8(ok) == 0xFFFF) {
            _mm_storeu_si128((__m128i*)(dst + x), _mm_shuffle_epi8(_mm_shuffle_epi8(comp, nib), rev));
        } else {
            for (size_t y = 0; y < 16; ++y) dst[x + y] = kComp.t[(uint8_t)src[n - 1 - x - y]];
        }
    }
    for (; x < n; ++x) dst[x] = kComp.t[(uint8_t)src[n - 1 - x]];
}

__attribute__((target("ssse3"))) void reverse_ssse3(char* dst, const char* src, size_t n) {
    const __m128i rev = _mm_setr_epi8(15, 14, 13, 12, 11, 10, 9, 8, 7, 6, 5, 4, 3, 2, 1, 0);
    size_t x = 0;
    for (; x + 16 <= n; x += 16)
        _mm_storeu_si128((__m128i*)(dst + x),
                         _mm_shuffle_epi8(_mm_loadu_si128((const __m128i*)(src + n - 16 - x)), rev));
    for (; x < n; ++x) dst[x] = src[n - 1 - x];
}

const bool kSsse3 = __builtin_cpu_supports("ssse3");

void revcomp_into(char* dst, const char* src, size_t n) {
    if (kSsse3) return revcomp_ssse3(dst, src, n);
    for (size_t x = 0; x < n; ++x) dst[x] = kComp.t[(uint8_t)src[n - 1 - x]];
}

void reverse_into(char* dst, const char* src, size_t n) {
    if (kSsse3) return reverse_ssse3(dst, src, n);
    for (size_t x = 0; x < n; ++x) dst[x] = src[n - 1 - x];
}

// Raw output cursor for SAM records: the caller reserves enough room.
struct Out {
    char* p;
    void str(const std::string& x) { std::memcpy(p, x.data(), x.size()); p += x.size(); }
    void chr(char c) { *p++ = c; }
    void lit(const char* x, size_t n) { std::memcpy(p, x, n); p += n; }
    // decimal, two digits per step from the pair table
    void u(uint64_t v) {
        char b[24];
        int n = 24;
        while (v >= 100) {
            const uint64_t q = v / 100;
            const uint32_t r = (uint32_t)(v - 100 * q);
            n -= 2;
            std::memcpy(b + n, kPairs.d + 2 * r, 2);
            v = q;
        }
        if (v >= 10) {
            n -= 2;
            std::memcpy(b + n, kPairs.d + 2 * v, 2);
        } else {
            b[--n] = (char)('0' + v);
        }
        std::memcpy(p, b + n, (size_t)(24 - n));
        p += 24 - n;
    }
    void i(int64_t v) {
        if (v < 0) { *p++ = '-'; u((uint64_t)(-v)); }
        else u((uint64_t)v);
    }
};

// SAM record fields (src/io/sam.rs:21-43)
struct Sam {
    const Record* rec;
    uint16_t flags = 0;
    int64_t rname = -1;  // reference id
    int64_t pos = -1;
    uint8_t mapq = 0;
    const Alignment* al = nullptr;  // CIGAR source (mapped)
    bool rnext_eq = false;
    int64_t rnext = -1;
    int64_t mate_pos = -1;
    bool has_tlen = false;
    int32_t tlen = 0;
    bool revseq = false;
    const Details* det = nullptr;
    bool is_mapped() const { return (flags & UNMAP) == 0; }
};

void write_cigar(Out& s, const Alignment& a, bool eqx) {
    static const char L[] = "MIDNSHP=X";
    uint32_t pend_op = 99, pend_len = 0;
    auto push = [&](uint32_t op, uint32_t len) {
        if (len == 0) return;
        if (!eqx && (op == 7 || op == 8)) op = 0;
        if (op == pend_op) { pend_len += len; return; }
        if (pend_len) { s.u(pend_len); s.chr(L[pend_op]); }
        pend_op = op;
        pend_len = len;
    };
    push(4, a.soft_clip_left);
    for (uint32_t i = 0; i < a.n_cigar; ++i) push(a.cigar[i] & 0xF, a.cigar[i] >> 4);
    push(4, a.soft_clip_right);
    if (pend_len) { s.u(pend_len); s.chr(L[pend_op]); }
}

// src/io/sam.rs:53-138
void write_sam(std::string& out, const Sam& r, const Ctx& c) {
    const std::string& seq = r.rec->seq;
    const std::string& q = r.rec->qual;
    // upper bound of the record's length
    size_t bound = r.rec->name.size() + seq.size() + q.size() + r.rec->comment.size() + c.o.rg_id.size() + 400 +
                   (r.al ? 12 * (r.al->n_cigar + 2) : 0);
    if (r.rname >= 0) bound += c.ref.names[(size_t)r.rname].size();
    if (r.rnext >= 0) bound += c.ref.names[(size_t)r.rnext].size();
    const size_t old = out.size();
    out.resize(old + bound);
    Out s{&out[old]};
    s.str(r.rec->name); s.chr('\t');
    s.u(r.flags); s.chr('\t');
    if (r.rname >= 0) s.str(c.ref.names[(size_t)r.rname]); else s.chr('*');
    s.chr('\t');
    s.u(r.pos >= 0 ? (uint64_t)r.pos + 1 : 0); s.chr('\t');
    s.u(r.mapq); s.chr('\t');
    if (r.al) write_cigar(s, *r.al, c.o.eqx); else s.chr('*');
    s.chr('\t');
    if (r.rnext_eq) s.chr('=');
    else if (r.rnext >= 0) s.str(c.ref.names[(size_t)r.rnext]);
    else s.chr('*');
    s.chr('\t');
    s.u(r.mate_pos >= 0 ? (uint64_t)r.mate_pos + 1 : 0); s.chr('\t');
    s.i(r.has_tlen ? r.tlen : 0); s.chr('\t');
    const bool sec = (r.flags & SECONDARY) != 0;
    if (sec || seq.empty()) s.chr('*');
    else if (r.revseq) {
        revcomp_into(s.p, seq.data(), seq.size());
        s.p += seq.size();
    } else s.str(seq);
    s.chr('\t');
    if (sec || !r.rec->has_qual || q.empty()) s.chr('*');
    else if (r.revseq) {
        reverse_into(s.p, q.data(), q.size());
        s.p += q.size();
    } else s.str(q);
    if (r.al) { s.lit("\tNM:i:", 6); s.u(r.al->edit_distance); s.lit("\tAS:i:", 6); s.u(r.al->score); }
    if (c.o.details && r.det) {
        s.lit("\tna:i:", 6); s.u(r.det->nam.n_nams);
        s.lit("\tal:i:", 6); s.u(r.det->tried_alignment);
        s.lit("\tga:i:", 6); s.u(r.det->gapped);
        s.lit("\tX0:i:", 6); s.u(r.det->best_alignments);
        if (r.flags & PAIRED) { s.lit("\tmr:i:", 6); s.u(r.det->mate_rescue); }
    }
    if (c.o.has_rg) { s.lit("\tRG:Z:", 6); s.str(c.o.rg_id); }
    if (c.o.fastq_comments && r.rec->has_comment) { s.chr('\t'); s.str(r.rec->comment); }
    s.chr('\n');
    out.resize((size_t)(s.p - out.data()));
}

// src/mapper.rs:115-243
Sam make_unmapped(const Record& rec, const Details* d) {
    Sam r;
    r.rec = &rec;
    r.flags = UNMAP;
    r.det = d;
    return r;
}
Sam make_mapped(const Alignment& a, const Record& rec, uint8_t mapq, bool secondary, const Details* d) {
    Sam r;
    r.rec = &rec;
    if (a.is_revcomp) r.flags |= REVERSE;
    if (secondary) { mapq = 0; r.flags |= SECONDARY; }
    r.revseq = a.is_revcomp;
    r.al = &a;
    r.rname = a.ref_id;
    r.pos = a.ref_start;
    r.mapq = mapq;
    r.det = d;
    return r;
}

// src/mapper.rs:245-345
void make_paired(const Alignment* al[2], const Record* recs[2], const uint8_t mq[2], const Details det[2],
                 bool secondary, bool proper, std::vector<Sam>& v) {
    size_t b = v.size();
    for (int i = 0; i < 2; ++i)
        v.push_back(al[i] ? make_mapped(*al[i], *recs[i], mq[i], secondary, &det[i]) : make_unmapped(*recs[i], &det[i]));
    v[b].flags |= READ1;
    v[b + 1].flags |= READ2;
    for (int i = 0; i < 2; ++i) {
        Sam& s = v[b + i];
        s.flags |= PAIRED;
        if (proper) s.flags |= PROPER_PAIR;
        const Alignment* mate = al[1 - i];
        if (mate) {
            if (mate->is_revcomp) s.flags |= MREVERSE;
            s.mate_pos = mate->ref_start;
            s.rnext = mate->ref_id;
            if (al[i]) {
                const Alignment* th = al[i];
                if (th->ref_id == mate->ref_id) {
                    s.rnext_eq = true;
                    int64_t tl;
                    if (mate->ref_start > th->ref_start) tl = (int64_t)(mate->ref_start - th->ref_start + mate->length);
                    else if (th->ref_start > mate->ref_start) tl = -(int64_t)(th->ref_start - mate->ref_start + th->length);
                    else if (i == 0) tl = -(int64_t)mate->length;
                    else tl = (int64_t)th->length;
                    s.has_tlen = true;
                    s.tlen = (int32_t)tl;
                }
            } else {
                s.rnext_eq = true;
                s.rname = mate->ref_id;
                s.pos = mate->ref_start;
            }
        } else {
            s.flags |= MUNMAP;
            if (al[i]) { s.rnext_eq = true; s.mate_pos = al[i]->ref_start; }
        }
    }
    if (secondary) {
        if (!v[b + 1].is_mapped()) v.pop_back();
        else if (!v[b].is_mapped()) { v[b] = v[b + 1]; v.pop_back(); }
    }
}

// src/mapper.rs:1302-1313
uint8_t mapping_quality(const std::vector<Nam>& nams) {
    if (nams.size() <= 1) return 60;
    float s1 = nams[0].score, s2 = nams[1].score;
    float mm = (float)std::min<uint32_t>(nams[0].n_anchors, 10) / 10.0f;
    float u = 40.0f * (1.0f - s2 / s1) * mm * std::log(s1);
    u = std::min(u, 60.0f);
    if (!(u > 0.0f)) return 0;
    return (uint8_t)u;
}

// src/nam.rs:168-188 + src/shuffle.rs:7-25 (the device already did the stable sort)
void shuffle_best(std::vector<Nam>& nams, Rng& rng) {
    if (nams.empty()) return;
    size_t nb = 0;
    while (nb < nams.size() && nams[nb].score == nams[0].score) ++nb;
    if (nb > 1)
        for (size_t i = 1; i < nb; ++i) std::swap(nams[i], nams[rng.mod64(i + 1)]);
}

void nams_of(const Ctx& c, uint64_t read_id, std::vector<Nam>& v) {
    v.clear();
    v.reserve(c.nb.nam_off[read_id + 1] - c.nb.nam_off[read_id]);
    for (uint64_t x = c.nb.nam_off[read_id]; x < c.nb.nam_off[read_id + 1]; ++x) v.push_back(to_nam(c.nb.nams[x]));
}

// ----------------------------------------------------------------- SE ---
// src/mapper.rs:349-528
void align_single_end(Ctx& c, uint64_t rec_idx, std::vector<Sam>& out, std::deque<Alignment>& store,
                      std::deque<std::array<Details, 2>>& dstore) {
    const Record& rec = c.r1[rec_idx];
    const uint32_t read_id = (uint32_t)(c.rel0 + rec_idx);
    const uint32_t L = (uint32_t)rec.seq.size();
    // per-thread vectors reused from record to record (a record's call does not nest)
    thread_local std::vector<Nam> nams;
    nams_of(c, read_id, nams);
    shuffle_best(nams, *c.rng);
    dstore.emplace_back();
    Details& det = dstore.back()[0];
    det.nam = c.nb.info[read_id];
    if (nams.empty()) { out.push_back(make_unmapped(rec, &det)); return; }
    thread_local std::vector<Alignment> alignments;
    alignments.clear();
    float nam_max = nams[0].score;
    uint32_t best_ed = 0xFFFFFFFFu, best_score = 0, second = 0, n_best = 0;
    size_t best_index = 0;
    std::optional<Alignment> best;
    for (size_t tries = 0; tries < nams.size() && tries < c.o.max_tries; ++tries) {
        Nam& nam = nams[tries];
        float dropoff = nam.score / nam_max;
        if ((tries > 1 && best_ed == 0) || dropoff < c.o.dropoff) break;
        if (!nam.cons) { det.inconsistent_nams++; continue; }
        auto al = extend_seed(c, nam, read_id, L, true, c.o.use_ssw);
        if (!al) continue;
        det.tried_alignment++;
        det.gapped += al->gapped ? 1 : 0;
        if (al->score >= best_score) {
            second = best_score;
            bool upd = false;
            if (al->score > best_score) { n_best = 1; upd = true; }
            else {
                n_best++;
                if (c.rng->mod32(n_best) == 0) upd = true;
            }
            if (upd) {
                best_score = al->score;
                best = *al;
                best_index = alignments.size();
                if (c.o.max_secondary == 0) best_ed = al->global_ed();
            }
        } else if (al->score > second) {
            second = al->score;
        }
        if (c.o.max_secondary > 0) alignments.push_back(*al);
    }
    if (!best) { out.push_back(make_unmapped(rec, &det)); return; }
    uint32_t num = 60 * (best_score - second);
    uint8_t mapq = (uint8_t)((num + best_score - 1) / best_score);
    store.push_back(*best);
    out.push_back(make_mapped(store.back(), rec, mapq, false, &det));
    if (c.o.max_secondary > 0) {
        alignments[best_index] = alignments.back();
        alignments.pop_back();
        std::stable_sort(alignments.begin(), alignments.end(),
                         [](const Alignment& a, const Alignment& b) { return a.score > b.score; });
        for (size_t i = 0; i < alignments.size() && i < c.o.max_secondary; ++i) {
            const Alignment& a = alignments[i];
            uint32_t diff = a.score > best_score ? a.score - best_score : 0;
            if (diff > 2 * (uint32_t)c.o.scores.mismatch + (uint32_t)c.o.scores.gap_open) break;  // saturating_sub never exceeds
            store.push_back(a);
            out.push_back(make_mapped(store.back(), rec, mapq, true, &det));
        }
    }
}

// ----------------------------------------------------------------- PE ---
struct ScoredPair {
    double score;
    std::optional<Alignment> a1, a2;
};

// A stable sort by score, largest first (as std::stable_sort with a > b), in
// place for the short lists of one read or pair; longer lists take
// std::stable_sort (which allocates a buffer).
template <typename T>
void sort_by_score_desc(std::vector<T>& v) {
    if (v.size() > 32) {
        std::stable_sort(v.begin(), v.end(), [](const T& a, const T& b) { return a.score > b.score; });
        return;
    }
    for (size_t i = 1; i < v.size(); ++i) {
        T x = std::move(v[i]);
        size_t j = i;
        for (; j > 0 && v[j - 1].score < x.score; --j) v[j] = std::move(v[j - 1]);
        v[j] = std::move(x);
    }
}

// src/mapper.rs:1018-1091
void rescue_read(Ctx& c, uint32_t read2, uint32_t L2, uint32_t read1, uint32_t L1, std::vector<Nam>& nams1,
                 Details det[2], float mu, float sigma, std::vector<ScoredPair>& pairs) {
    float max1 = nams1[0].score;
    // per-thread buffers reused from pair to pair (the call does not nest)
    thread_local std::vector<Alignment> al1;
    thread_local std::vector<std::optional<Alignment>> al2;
    al1.clear();
    al2.clear();
    for (size_t t = 0; t < nams1.size() && t < c.o.max_tries; ++t) {
        Nam& nam = nams1[t];
        if (nam.score / max1 < c.o.dropoff) break;
        bool cons = reverse_nam_if_needed(nam, L1);
        det[0].inconsistent_nams += !cons;
        auto a = extend_seed(c, nam, read1, L1, cons, true);
        if (a) {
            det[0].gapped += a->gapped;
            al1.push_back(*a);
            det[0].tried_alignment++;
            auto b = rescue_align(c, nam, read2, L2, mu, sigma);
            if (b) det[1].mate_rescue++;
            al2.push_back(b);
        }
    }
    pairs.clear();
    for (const Alignment& a1 : al1) {
        for (const auto& a2 : al2) {
            if (a2) {
                uint32_t dist = a1.ref_start > a2->ref_start ? a1.ref_start - a2->ref_start : a2->ref_start - a1.ref_start;
                float score = (float)(a1.score + a2->score);
                if ((a1.is_revcomp ^ a2->is_revcomp) && (float)dist < mu + 4.0f * sigma)
                    score += std::log(normal_pdf((float)dist, mu, sigma));
                else
                    score -= 10.0f;
                pairs.push_back({(double)score, a1, *a2});
            } else {
                pairs.push_back({(double)a1.score - 10.0, a1, std::nullopt});
            }
        }
    }
}

// src/mapper.rs:1184-1198
bool is_proper_pair(const Alignment* a1, const Alignment* a2, float mu, float sigma) {
    if (!a1 || !a2) return false;
    int64_t dist = (int64_t)a2->ref_start - (int64_t)a1->ref_start;
    bool same = a1->ref_id == a2->ref_id;
    bool r1r2 = !a1->is_revcomp && a2->is_revcomp && dist >= 0;
    bool r2r1 = !a2->is_revcomp && a1->is_revcomp && dist <= 0;
    float lim = mu + sigma * 6.0f;
    uint64_t limi = lim > 0 ? (uint64_t)lim : 0;
    return same && (uint64_t)(dist < 0 ? -dist : dist) <= limi && (r1r2 || r2r1);
}

// src/mapper.rs:1200-1220
bool is_proper_nam_pair(const Nam& n1, const Nam& n2, float mu, float sigma) {
    if (n1.ref_id != n2.ref_id || n1.is_revcomp == n2.is_revcomp) return false;
    uint32_t a = n1.projected_ref_start(), b = n2.projected_ref_start();
    bool r1r2 = n2.is_revcomp && a <= b && (float)(b - a) < mu + 10.0f * sigma;
    bool r2r1 = n1.is_revcomp && b <= a && (float)(a - b) < mu + 10.0f * sigma;
    return r1r2 || r2r1;
}

struct NamPair { float score; const Nam* n1; const Nam* n2; };

// src/mapper.rs:1224-1299
void best_scoring_nam_pairs(const std::vector<Nam>& n1, const std::vector<Nam>& n2, float mu, float sigma,
                            std::vector<NamPair>& pairs) {
    pairs.clear();
    if (n1.empty() && n2.empty()) return;
    thread_local std::vector<uint32_t> added1, added2;
    added1.clear();
    added2.clear();
    size_t best_joint = 0;
    for (size_t a = 0; a < n1.size() && a < 1000; ++a) {
        for (size_t b = 0; b < n2.size() && b < 1000; ++b) {
            size_t joint = (size_t)n1[a].n_anchors + n2[b].n_anchors;
            if (joint < best_joint / 2) break;
            if (is_proper_nam_pair(n1[a], n2[b], mu, sigma)) {
                pairs.push_back({n1[a].score + n2[b].score, &n1[a], &n2[b]});
                added1.push_back(n1[a].nam_id);
                added2.push_back(n2[b].nam_id);
                best_joint = std::max(joint, best_joint);
            }
        }
    }
    auto has = [](const std::vector<uint32_t>& v, uint32_t x) { return std::find(v.begin(), v.end(), x) != v.end(); };
    if (!n1.empty()) {
        size_t bh = best_joint > 0 ? best_joint : n1[0].n_anchors;
        for (const Nam& n : n1) {
            if (n.n_anchors < bh / 2) break;
            if (has(added1, n.nam_id)) continue;
            pairs.push_back({n.score, &n, nullptr});
        }
    }
    if (!n2.empty()) {
        size_t bh = best_joint > 0 ? best_joint : n2[0].n_anchors;
        for (const Nam& n : n2) {
            if (n.n_anchors < bh / 2) break;
            if (has(added2, n.nam_id)) continue;
            pairs.push_back({n.score, nullptr, &n});
        }
    }
    sort_by_score_desc(pairs);
}

// src/mapper.rs:1456-1470
float top_dropoff(const std::vector<Nam>& nams) {
    const Nam& n = nams[0];
    if (n.n_anchors <= 2) return 1.0f;
    if (nams.size() > 1) return (float)nams[1].n_anchors / (float)n.n_anchors;
    return 0.0f;
}

struct PairedResult {
    bool proper = false;
    Alignment p1, p2;
    std::vector<ScoredPair>* pairs = nullptr;  // the caller's buffer
};

// src/mapper.rs:778-1014
// The scored pairs go to `pairs` (cleared here; a per-thread buffer of the caller).
PairedResult extend_paired_seeds(Ctx& c, std::vector<Nam> nams[2], const uint32_t rid[2], const uint32_t L[2],
                                 Details det[2], float mu, float sigma, std::vector<ScoredPair>& pairs) {
    PairedResult res;
    res.pairs = &pairs;
    pairs.clear();
    const float dropoff = c.o.dropoff;
    if (nams[0].empty() && nams[1].empty()) return res;
    if (!nams[0].empty() && nams[1].empty()) {
        rescue_read(c, rid[1], L[1], rid[0], L[0], nams[0], det, mu, sigma, pairs);
        return res;
    }
    if (nams[0].empty() && !nams[1].empty()) {
        Details sw[2] = {det[1], det[0]};
        rescue_read(c, rid[0], L[0], rid[1], L[1], nams[1], sw, mu, sigma, pairs);
        det[0] = sw[1];
        det[1] = sw[0];
        for (auto& p : pairs) std::swap(p.a1, p.a2);
        return res;
    }
    if (top_dropoff(nams[0]) < dropoff && top_dropoff(nams[1]) < dropoff &&
        is_proper_nam_pair(nams[0][0], nams[1][0], mu, sigma)) {
        const Nam& m1 = nams[0][0];
        const Nam& m2 = nams[1][0];
        bool c1 = m1.cons, c2 = m2.cons;
        det[0].inconsistent_nams += !c1;
        det[1].inconsistent_nams += !c2;
        // both requests are registered before a new miss stops the record
        size_t before = c.ch->pending.size();
        auto a1 = extend_seed(c, m1, rid[0], L[0], c1, true);
        auto a2 = extend_seed(c, m2, rid[1], L[1], c2, true);
        if (c.ch->pending.size() != before) {
            c.abandon = true;
            return res;
        }
        if (a1 && a2) {
            det[0].tried_alignment++;
            det[0].gapped += a1->gapped;
            det[1].tried_alignment++;
            det[1].gapped += a2->gapped;
            res.proper = true;
            res.p1 = *a1;
            res.p2 = *a2;
            return res;
        }
    }
    // extension results by NAM id, per read (a few NAMs each: a list, searched
    // linearly, in per-thread buffers; the call does not nest)
    thread_local std::vector<std::pair<uint32_t, std::optional<Alignment>>> cache[2];
    auto cached = [&](int i, uint32_t id) -> std::optional<Alignment>* {
        for (auto& e : cache[i])
            if (e.first == id) return &e.second;
        return nullptr;
    };
    std::optional<Alignment> indv[2];
    for (int i = 0; i < 2; ++i) {
        cache[i].clear();
        bool cons = reverse_nam_if_needed(nams[i][0], L[i]);
        det[i].inconsistent_nams += !cons;
        indv[i] = extend_seed(c, nams[i][0], rid[i], L[i], cons, true);
        det[i].tried_alignment++;
        det[i].gapped += indv[i] ? indv[i]->gapped : 0;
        cache[i].emplace_back(nams[i][0].nam_id, indv[i]);
    }
    thread_local std::vector<NamPair> np;
    best_scoring_nam_pairs(nams[0], nams[1], mu, sigma, np);
    std::vector<ScoredPair>& out = pairs;
    float max_score = np[0].score;
    for (const NamPair& p : np) {
        float sd = p.score / max_score;
        if (out.size() >= c.o.max_tries || sd < dropoff) break;
        std::optional<Alignment> als[2];
        const Nam* ns[2] = {p.n1, p.n2};
        for (int i = 0; i < 2; ++i) {
            std::optional<Alignment> al;
            if (ns[i]) {
                Nam th = *ns[i];
                if (std::optional<Alignment>* hit = cached(i, th.nam_id)) {
                    al = *hit;
                } else {
                    bool cons = reverse_nam_if_needed(th, L[i]);
                    det[i].inconsistent_nams += !cons;
                    al = extend_seed(c, th, rid[i], L[i], cons, true);
                    det[i].tried_alignment++;
                    det[i].gapped += al ? al->gapped : 0;
                    cache[i].emplace_back(th.nam_id, al);
                }
            } else {
                Nam other = *ns[1 - i];
                det[1 - i].inconsistent_nams += !reverse_nam_if_needed(other, L[1 - i]);
                al = rescue_align(c, other, rid[i], L[i], mu, sigma);
                if (al) { det[i].mate_rescue++; det[i].tried_alignment++; }
            }
            if ((al ? al->score : 0) > (indv[i] ? indv[i]->score : 0)) indv[i] = al;
            als[i] = al;
        }
        if (!als[0] || !als[1]) continue;
        const Alignment& a1 = *als[0];
        const Alignment& a2 = *als[1];
        float lim = mu + 10.0f * sigma;
        uint64_t limi = lim > 0 ? (uint64_t)lim : 0;
        bool r1r2 = a2.is_revcomp && !a1.is_revcomp && a1.ref_start <= a2.ref_start && (uint64_t)(a2.ref_start - a1.ref_start) < limi;
        bool r2r1 = a1.is_revcomp && !a2.is_revcomp && a2.ref_start <= a1.ref_start && (uint64_t)(a1.ref_start - a2.ref_start) < limi;
        double cs;
        if (r1r2 || r2r1) {
            uint32_t x = a1.ref_start > a2.ref_start ? a1.ref_start - a2.ref_start : a2.ref_start - a1.ref_start;
            cs = (double)a1.score + (double)a2.score + std::max(-20.0 + 0.001, (double)std::log(normal_pdf((float)x, mu, sigma)));
        } else {
            cs = (double)a1.score + (double)a2.score - 20.0;
        }
        out.push_back({cs, a1, a2});
    }
    if (indv[0] && indv[1]) out.push_back({(double)indv[0]->score + (double)indv[1]->score - 20.0, indv[0], indv[1]});
    return res;
}

// src/mapper.rs:1343-1373
void dedup_pairs(std::vector<ScoredPair>& v) {
    if (v.size() < 2) return;
    auto same = [](const std::optional<Alignment>& a, const std::optional<Alignment>& b) {
        if (!a && !b) return true;
        if (!a || !b) return false;
        return a->ref_start == b->ref_start && a->ref_id == b->ref_id;
    };
    size_t k = 0, j = 1;
    for (size_t i = 1; i < v.size(); ++i)
        if (!same(v[k].a1, v[i].a1) || !same(v[k].a2, v[i].a2)) { v[j] = v[i]; ++j; k = i; }
    v.resize(j);
}

// src/mapper.rs:1436-1453
uint8_t alignment_quality(const std::vector<ScoredPair>& v) {
    if (v.size() <= 1) return 60;
    float s1 = (float)v[0].score, s2 = (float)v[1].score;
    if (s1 > 0 && s2 > 0) {
        float d = std::min(s1 - s2, 60.0f);
        if (!(d > 0)) return 0;
        return (uint8_t)d;
    }
    if (s1 > 0 && s2 <= 0) return 60;
    return 1;
}

// src/mapper.rs:624-768
void align_paired_end(Ctx& c, uint64_t idx, std::vector<Sam>& out, std::deque<Alignment>& store,
                      std::deque<std::array<Details, 2>>& dstore) {
    const Record* recs[2] = {&c.r1[idx], &c.r2[idx]};
    const uint64_t rel = c.rel0 + idx;
    const uint32_t rid[2] = {(uint32_t)(2 * rel), (uint32_t)(2 * rel + 1)};
    const uint32_t L[2] = {(uint32_t)recs[0]->seq.size(), (uint32_t)recs[1]->seq.size()};
    dstore.emplace_back();
    Details* det = dstore.back().data();
    thread_local std::vector<Nam> nams[2];
    for (int i = 0; i < 2; ++i) {
        nams_of(c, rid[i], nams[i]);
        shuffle_best(nams[i], *c.rng);
        det[i].nam = c.nb.info[rid[i]];
    }
    InsertSize& isz = *c.isize;
    float mu = isz.mu, sigma = isz.sigma;
    thread_local std::vector<ScoredPair> pair_buf;
    PairedResult pr = extend_paired_seeds(c, nams, rid, L, det, mu, sigma, pair_buf);
    if (c.abandon) return;
    if (pr.proper) {
        bool ps = is_proper_pair(&pr.p1, &pr.p2, mu, sigma);
        if (ps && isz.sample_size < 400 && pr.p1.edit_distance + pr.p2.edit_distance < 3) {
            uint32_t d = pr.p1.ref_start > pr.p2.ref_start ? pr.p1.ref_start - pr.p2.ref_start : pr.p2.ref_start - pr.p1.ref_start;
            isz.update(d);
        }
        uint8_t mq[2] = {mapping_quality(nams[0]), mapping_quality(nams[1])};
        det[0].best_alignments = 1;
        det[1].best_alignments = 1;
        store.push_back(pr.p1);
        store.push_back(pr.p2);
        const Alignment* al[2] = {&store[store.size() - 2], &store[store.size() - 1]};
        make_paired(al, recs, mq, det, false, ps, out);
        return;
    }
    std::vector<ScoredPair>& v = *pr.pairs;
    sort_by_score_desc(v);
    dedup_pairs(v);
    size_t nb = 0;
    while (nb < v.size() && v[nb].score == v[0].score) ++nb;
    det[0].best_alignments = (uint32_t)nb;
    det[1].best_alignments = (uint32_t)nb;
    if (nb > 1) std::swap(v[0], v[(size_t)c.rng->mod64(nb)]);
    uint8_t mq[2];
    if (v.size() == 1 && !is_proper_pair(v[0].a1 ? &*v[0].a1 : nullptr, v[0].a2 ? &*v[0].a2 : nullptr, mu, sigma)) {
        mq[0] = mapping_quality(nams[0]);
        mq[1] = mapping_quality(nams[1]);
    } else {
        mq[0] = mq[1] = alignment_quality(v);
    }
    if (v.empty()) {
        Sam a = make_unmapped(*recs[0], &det[0]), b = make_unmapped(*recs[1], &det[1]);
        a.flags |= PAIRED | READ1 | MUNMAP;
        b.flags |= PAIRED | READ2 | MUNMAP;
        out.push_back(a);
        out.push_back(b);
        return;
    }
    const double dropoff = (double)(uint8_t)(2 * c.o.scores.mismatch + c.o.scores.gap_open);  // u8 sum, src/mapper.rs:746
    size_t limit = c.o.max_secondary == 0 ? 1 : c.o.max_secondary;
    bool secondary = false;
    double smax = v[0].score;
    for (size_t i = 0; i < v.size() && i < limit; ++i) {
        if (c.o.max_secondary != 0 && !(smax - v[i].score < dropoff)) break;
        const Alignment* al[2] = {nullptr, nullptr};
        if (v[i].a1) { store.push_back(*v[i].a1); al[0] = &store.back(); }
        if (v[i].a2) { store.push_back(*v[i].a2); al[1] = &store.back(); }
        bool ps = is_proper_pair(al[0], al[1], mu, sigma);
        uint8_t emq[2] = {secondary ? (uint8_t)0 : mq[0], secondary ? (uint8_t)0 : mq[1]};
        make_paired(al, recs, emq, det, secondary, ps, out);
        secondary = true;
    }
}

// -------------------------------------------------------------- maponly ---
// src/io/paf.rs:37-60
void write_paf(std::string& s, const Nam& n, const Record& rec, int end, const PackedRef& ref, int mapq) {
    s += rec.name;
    if (end == 1) s += "/1";
    else if (end == 2) s += "/2";
    s += '\t'; put_u(s, rec.seq.size());
    s += '\t'; put_u(s, n.query_start);
    s += '\t'; put_u(s, n.query_end);
    s += '\t'; s += n.is_revcomp ? '-' : '+';
    s += '\t'; s += ref.names[n.ref_id];
    s += '\t'; put_u(s, ref.len[n.ref_id]);
    s += '\t'; put_u(s, n.ref_start);
    s += '\t'; put_u(s, n.ref_end);
    s += '\t'; put_u(s, n.matching_bases);
    s += '\t'; put_u(s, n.ref_end - n.ref_start);
    s += '\t'; put_u(s, (uint64_t)mapq);
    s += '\n';
}

struct NamPairO { Nam n1, n2; double score; };

// src/maponly.rs:380-458
void find_pairs(const std::vector<Nam*>& fwd, const std::vector<Nam*>& rev, float mu, float sigma, bool swap,
                std::vector<NamPairO>& out) {
    float md = std::ceil(mu + 10.0f * sigma);
    uint64_t max_dist = md > 0 ? (uint64_t)md : 0;
    size_t rp = 0;
    bool have_last = false;
    size_t last_id = 0;
    double last_score = 0;
    for (const Nam* f : fwd) {
        while (rp < rev.size() && (rev[rp]->ref_id < f->ref_id ||
                                   (rev[rp]->ref_id == f->ref_id && rev[rp]->projected_ref_start() < f->projected_ref_start())))
            ++rp;
        if (rp == rev.size()) break;
        if (rev[rp]->ref_id > f->ref_id) continue;
        bool have = false;
        size_t best_id = 0;
        double best_score = 0;
        for (size_t i = rp; i < rev.size() && rev[i]->ref_id == f->ref_id &&
                            (uint64_t)(rev[i]->projected_ref_start() - f->projected_ref_start()) <= max_dist;
             ++i) {
            const Nam* r = rev[i];
            uint32_t x = f->ref_start > r->ref_start ? f->ref_start - r->ref_start : r->ref_start - f->ref_start;
            double s = (double)f->score + (double)r->score +
                       std::max(0.001, (double)std::log(normal_pdf((float)x, mu, sigma) + 1.0f));
            if (!have || s > best_score) { have = true; best_id = i; best_score = s; }
        }
        if (!have) continue;
        if (have_last && last_id == best_id) {
            if (best_score <= last_score) continue;
            out.pop_back();
        }
        if (swap) out.push_back({*rev[best_id], *f, best_score});
        else out.push_back({*f, *rev[best_id], best_score});
        have_last = true;
        last_id = best_id;
        last_score = best_score;
        rp = best_id;
    }
}

// src/maponly.rs:314-375 (tie order of Rust's sort_unstable: parity unpinned)
std::vector<NamPairO> get_nam_pairs(std::vector<Nam>& n1, std::vector<Nam>& n2, float mu, float sigma, bool both1, bool both2) {
    std::vector<NamPairO> res;
    if (n1.empty() || n2.empty()) return res;
    auto split = [](std::vector<Nam>& v, bool both) -> size_t {
        if (both) {
            size_t left = 0, right = v.size();
            while (left < right) {
                if (v[left].is_revcomp) { --right; std::swap(v[left], v[right]); }
                else ++left;
            }
            return left;
        }
        return v[0].is_revcomp ? 0 : v.size();
    };
    size_t e1 = split(n1, both1), e2 = split(n2, both2);
    auto key_sort = [](std::vector<Nam>& v, size_t a, size_t b) {
        std::sort(v.begin() + a, v.begin() + b, [](const Nam& x, const Nam& y) {
            if (x.ref_id != y.ref_id) return x.ref_id < y.ref_id;
            return x.projected_ref_start() < y.projected_ref_start();
        });
    };
    auto run = [&](size_t fa, size_t fb, std::vector<Nam>& fv, size_t ra, size_t rb, std::vector<Nam>& rv, bool swap) {
        if (fa == fb || ra == rb) return;
        key_sort(fv, fa, fb);
        key_sort(rv, ra, rb);
        std::vector<Nam*> f, r;
        for (size_t i = fa; i < fb; ++i) f.push_back(&fv[i]);
        for (size_t i = ra; i < rb; ++i) r.push_back(&rv[i]);
        find_pairs(f, r, mu, sigma, swap, res);
    };
    run(0, e1, n1, e2, n2.size(), n2, false);
    run(0, e2, n2, e1, n1.size(), n1, true);
    std::stable_sort(res.begin(), res.end(), [](const NamPairO& a, const NamPairO& b) { return a.score > b.score; });
    return res;
}

void map_only(Ctx& c, uint64_t idx, std::string& s, std::vector<double>& ab) {
    const PackedRef& ref = c.ref;
    if (!c.paired) {
        const Record& r1 = c.r1[idx];
        uint32_t id = (uint32_t)(c.rel0 + idx);
        thread_local std::vector<Nam> nams;
        nams_of(c, id, nams);
        shuffle_best(nams, *c.rng);
        if (c.o.mode == 1) {
            if (!nams.empty()) write_paf(s, nams[0], r1, 0, ref, mapping_quality(nams));
        } else {
            size_t nb = 0;
            while (nb < nams.size() && nams[nb].score == nams[0].score) ++nb;
            double w = (double)r1.seq.size() / (double)nb;
            for (size_t i = 0; i < nb; ++i) ab[nams[i].ref_id] += w;
        }
        return;
    }
    const Record& r1 = c.r1[idx];
    const Record& r2 = c.r2[idx];
    uint64_t rel = c.rel0 + idx;
    uint32_t id1 = (uint32_t)(2 * rel), id2 = (uint32_t)(2 * rel + 1);
    thread_local std::vector<Nam> n1, n2;
    nams_of(c, id1, n1);
    nams_of(c, id2, n2);
    if (n1.empty() && n2.empty()) return;
    InsertSize& isz = *c.isize;
    std::vector<NamPairO> np = get_nam_pairs(n1, n2, isz.mu, isz.sigma, c.nb.info[id1].both_orientations != 0,
                                             c.nb.info[id2].both_orientations != 0);
    // sort_nams on the partitioned vectors: stable by score, then shuffle the best
    auto sort_shuffle = [&](std::vector<Nam>& v) {
        std::stable_sort(v.begin(), v.end(), [](const Nam& a, const Nam& b) { return a.score > b.score; });
        shuffle_best(v, *c.rng);
    };
    sort_shuffle(n1);
    sort_shuffle(n2);
    double indiv = (n1.empty() ? 0.0 : (double)n1[0].score) + (n2.empty() ? 0.0 : (double)n2[0].score);
    bool pair = !np.empty() && np[0].score >= indiv / 2.0;
    if (pair && isz.sample_size < 400) {
        const Nam& a = np[0].n1;
        const Nam& b = np[0].n2;
        isz.update(a.ref_start > b.ref_start ? a.ref_start - b.ref_start : b.ref_start - a.ref_start);
    }
    if (c.o.mode == 1) {
        if (pair) {
            write_paf(s, np[0].n1, r1, 1, ref, 255);
            write_paf(s, np[0].n2, r2, 2, ref, 255);
        } else {
            if (!n1.empty()) write_paf(s, n1[0], r1, 1, ref, 255);
            if (!n2.empty()) write_paf(s, n2[0], r2, 2, ref, 255);
        }
    } else if (pair) {
        size_t nb = 0;
        while (nb < np.size() && np[nb].score == np[0].score) ++nb;
        double w1 = (double)r1.seq.size() / (double)nb, w2 = (double)r2.seq.size() / (double)nb;
        for (size_t i = 0; i < nb; ++i) { ab[np[i].n1.ref_id] += w1; ab[np[i].n2.ref_id] += w2; }
    } else {
        for (int e = 0; e < 2; ++e) {
            const std::vector<Nam>& v = e == 0 ? n1 : n2;
            size_t Lr = e == 0 ? r1.seq.size() : r2.seq.size();
            size_t nb = 0;
            while (nb < v.size() && v[nb].score == v[0].score) ++nb;
            double w = (double)Lr / (double)nb;
            for (size_t i = 0; i < nb; ++i) ab[v[i].ref_id] += w;
        }
    }
}

// One pass over a chunk: commit records until the first miss, then probe the rest.
// Records that miss leave the state as their probe run made it, except records
// that stop early, whose state changes are undone. (Control flow uses flags, not
// exceptions: throwing from many threads serialises on the unwinder.)
void run_chunk(Ctx base, Chunk& ch) {
    std::vector<Sam> recs;
    std::deque<Alignment> store;
    std::deque<std::array<Details, 2>> dstore;
    Ctx c = base;
    c.ch = &ch;
    c.r1 = ch.rec1;
    c.r2 = ch.rec2;
    c.rel0 = ch.pos;
    Rng rng = ch.rng;
    InsertSize isz = ch.isize;
    bool probing = false;
    using clk = std::chrono::steady_clock;
    auto t_rec = clk::now();
    ch.tent_cur.clear();
    ch.text_cur.clear();
    size_t tp = 0;  // next of the last pass's probes to check
    for (uint64_t r = ch.next; r < ch.count; ++r) {
        const uint64_t idx = r;
        if (!probing && c.o.mode == 0) {
            // a record the last pass probed from the state it now starts from:
            // take its output and end state instead of deciding it again
            while (tp < ch.tent_prev.size() && ch.tent_prev[tp].r < r) ++tp;
            if (tp < ch.tent_prev.size() && ch.tent_prev[tp].r == r &&
                same_state(rng, isz, ch.tent_prev[tp].rng0, ch.tent_prev[tp].isz0)) {
                const Tentative& t = ch.tent_prev[tp];
                ch.out.append(ch.text_prev, t.t0, t.t1 - t.t0);
                rng = t.rng1;
                isz = t.isz1;
                ch.next = r + 1;
                ch.rng = rng;
                ch.isize = isz;
                ch.n_accept++;
                t_rec = clk::now();
                continue;
            }
        }
        const Rng rng0 = rng;
        const InsertSize isz0 = isz;
        c.rng = &rng;
        c.isize = &isz;
        c.missed = false;
        c.abandon = false;
        recs.clear();
        store.clear();
        dstore.clear();
        if (c.o.mode == 0) {
            if (c.paired) {
                align_paired_end(c, idx, recs, store, dstore);
                if (!c.abandon && c.o.only_mapped && !recs[0].is_mapped() && !recs[1].is_mapped()) recs.clear();
            } else {
                align_single_end(c, idx, recs, store, dstore);
                if (c.o.only_mapped && !recs[0].is_mapped()) recs.clear();
            }
            if (!c.missed && !c.abandon) {
                const auto w0 = clk::now();
                if (!probing) {
                    for (const Sam& s : recs) write_sam(ch.out, s, c);
                } else {
                    // a probe without a miss: kept for the next pass
                    const uint64_t t0 = ch.text_cur.size();
                    for (const Sam& s : recs) write_sam(ch.text_cur, s, c);
                    ch.tent_cur.push_back(Tentative{r, rng0, rng, isz0, isz, t0, ch.text_cur.size()});
                }
                ch.ns_write += (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(clk::now() - w0).count();
            }
        } else {
            // mapping-only modes do no extension, so they never miss
            map_only(c, idx, ch.out, ch.abundances);
        }
        if (c.abandon) {
            rng = rng0;
            isz = isz0;
        }
        if (c.missed || c.abandon) probing = true;
        if (!probing) {
            ch.next = r + 1;
            ch.rng = rng;
            ch.isize = isz;
        }
        const auto t1 = clk::now();
        const uint64_t ns = (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(t1 - t_rec).count();
        t_rec = t1;
        if (probing) {
            ch.n_probe++;
            ch.ns_probe += ns;
        } else {
            ch.n_commit++;
            ch.ns_commit += ns;
        }
    }
    ch.tent_prev.swap(ch.tent_cur);
    ch.text_prev.swap(ch.text_cur);
}

// CPUs this process may run on (its affinity mask; hardware_concurrency counts
// every CPU of the machine, also those a container or scheduler withholds).
unsigned usable_cpus() {
    cpu_set_t set;
    CPU_ZERO(&set);
    if (sched_getaffinity(0, sizeof set, &set) == 0) {
        const int n = CPU_COUNT(&set);
        if (n > 0) return (unsigned)n;
    }
    return std::max(1u, std::thread::hardware_concurrency());
}

// Worker threads shared by every parallel_for of the process (the engines' host
// passes run concurrently and draw from one pool instead of each starting its own
// threads per call). A job takes at most `threads` - 1 helpers besides its caller.
class HostPool {
public:
    struct Job {
        std::function<void(size_t)> f;
        size_t n = 0;
        int helpers = 0;  // helpers still allowed to join
        std::atomic<size_t> next{0}, done{0};
        std::mutex m;
        std::condition_variable cv;
        void run() {
            size_t k = 0;
            for (size_t i; (i = next.fetch_add(1)) < n; ++k) f(i);
            if (k && done.fetch_add(k) + k == n) {
                std::lock_guard<std::mutex> lk(m);
                cv.notify_all();
            }
        }
    };
    static HostPool& get() {
        static HostPool p;
        return p;
    }
    void submit(const std::shared_ptr<Job>& j) {
        {
            std::lock_guard<std::mutex> lk(m_);
            q_.push_back(j);
        }
        cv_.notify_all();
    }

private:
    HostPool() {
        const unsigned n = std::max(2u, usable_cpus());
        for (unsigned t = 0; t < n; ++t) th_.emplace_back([this] { loop(); });
    }
    ~HostPool() {
        {
            std::lock_guard<std::mutex> lk(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : th_) t.join();
    }
    void loop() {
        std::unique_lock<std::mutex> lk(m_);
        for (;;) {
            cv_.wait(lk, [&] { return stop_ || !q_.empty(); });
            if (stop_) return;
            std::shared_ptr<Job> j = q_.front();
            if (--j->helpers <= 0 || j->next.load() >= j->n) q_.pop_front();
            lk.unlock();
            j->run();
            lk.lock();
        }
    }
    std::mutex m_;
    std::condition_variable cv_;
    std::deque<std::shared_ptr<HostPool::Job>> q_;
    std::vector<std::thread> th_;
    bool stop_ = false;
};

template <typename F>
void parallel_for(int threads, size_t n, F f) {
    if (threads <= 1 || n <= 1) {
        for (size_t i = 0; i < n; ++i) f(i);
        return;
    }
    auto j = std::make_shared<HostPool::Job>();
    j->f = [&f](size_t i) { f(i); };
    j->n = n;
    j->helpers = std::min<int>(threads, (int)n) - 1;
    HostPool::get().submit(j);
    j->run();
    std::unique_lock<std::mutex> lk(j->m);
    j->cv.wait(lk, [&] { return j->done.load() == j->n; });
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Splits records [0, n) into chunks (src/main.rs:591-604): records until the
// cumulative bases exceed chunk_size, the crossing record included.
std::vector<Chunk> make_chunks(const Record* r1, const Record* r2, uint64_t n, uint64_t chunk_size) {
    std::vector<Chunk> chunks;
    uint64_t nt = 0, first = 0;
    for (uint64_t i = 0; i < n; ++i) {
        nt += r1[i].seq.size() + (r2 ? r2[i].seq.size() : 0);
        if (nt > chunk_size || i + 1 == n) {
            Chunk c;
            c.first = first;
            c.count = i + 1 - first;
            c.rec1 = r1 + first;
            c.rec2 = r2 ? r2 + first : nullptr;
            chunks.push_back(std::move(c));
            first = i + 1;
            nt = 0;
        }
    }
    return chunks;
}

uint64_t chunk_bases(const Chunk& c) {
    uint64_t b = 0;
    for (uint64_t i = 0; i < c.count; ++i) b += c.rec1[i].seq.size() + (c.rec2 ? c.rec2[i].seq.size() : 0);
    return b;
}

// Where the reads of a device batch come from: staged from the chunks' host
// records (d_base null), or already in device memory, each chunk's reads
// (records in order, paired: r1 then r2 of each pair) at d_base + seg[chunk].
struct BatchReads {
    const uint8_t* d_base = nullptr;
    std::vector<uint64_t> seg;
};

const ChainParams kChain{50, 0.1f, 0.05f, 0.7f, -1, 0.01f, 10.0f};

// One device batch of whole chunks (of one or several read sets) on one engine:
// reads to the device, NAMs, then replay passes and extension rounds until every
// chunk is committed. Each chunk keeps its own output, RNG and insert-size state,
// so a chunk's result does not depend on the batch it ran in.
void run_batch(Engine& eng, const PackedRef& ref, const Options& o, bool paired, Chunk* const* chs, size_t nch,
               const BatchReads& br, RunStats& st, bool keep) {
    double t0 = now_ms();
    std::vector<uint64_t> off(1, 0), seg_len(nch, 0);
    uint64_t nrec = 0;
    for (size_t x = 0; x < nch; ++x) {
        Chunk& ch = *chs[x];
        ch.pos = nrec;
        nrec += ch.count;
        const uint64_t s0 = off.back();
        for (uint64_t i = 0; i < ch.count; ++i) {
            off.push_back(off.back() + ch.rec1[i].seq.size());
            if (paired) off.push_back(off.back() + ch.rec2[i].seq.size());
        }
        seg_len[x] = off.back() - s0;
    }
    const uint32_t n_reads = (uint32_t)(off.size() - 1);
    if (br.d_base) {
        eng.gather_reads(br.d_base, br.seg.data(), seg_len.data(), (uint32_t)nch, off.data(), n_reads);
    } else {
        std::vector<uint8_t> seq(off.back());
        uint64_t at = 0;
        for (size_t x = 0; x < nch; ++x) {
            const Chunk& ch = *chs[x];
            for (uint64_t i = 0; i < ch.count; ++i) {
                std::memcpy(seq.data() + at, ch.rec1[i].seq.data(), ch.rec1[i].seq.size());
                at += ch.rec1[i].seq.size();
                if (paired) {
                    std::memcpy(seq.data() + at, ch.rec2[i].seq.data(), ch.rec2[i].seq.size());
                    at += ch.rec2[i].seq.size();
                }
            }
        }
        eng.upload_reads(seq.data(), off.data(), n_reads);
    }
    double t1 = now_ms();
    const MapParams mp{o.mcs, o.rescue_distance, o.lookup_lanes, o.seed_lanes};
    NamBatch nb;
    SpecBatch spec;
    // the fused pass also computes the extensions the decisions are expected to
    // need (alignment mode only: the other modes extend nothing)
    const SpecParams sp{paired ? 1u : 0u, o.use_ssw ? 1u : 0u, (uint32_t)std::min<uint64_t>(o.max_tries, 1u << 20),
                        o.dropoff, o.spec_per_read};
    const bool fused = o.fused && o.mode == 0 && eng.map_fused(mp, kChain, sp, o.scores, nb, spec);
    if (!fused) eng.find_nams(mp, kChain, nb);
    st.graph_replays += spec.replayed ? 1 : 0;
    st.fused_overflows += spec.overflow ? 1 : 0;
    double t2 = now_ms();
    st.upload_ms += t1 - t0;
    st.nam_ms += t2 - t1;
    st.nams += nb.nams.size();
    const uint32_t k = (uint32_t)eng.params().k;
    Ctx base{o, ref, nb, nullptr, nullptr, paired, 0, k, nullptr, false, false, nullptr, nullptr};
    for (size_t x = 0; x < nch; ++x) {
        chs[x]->abundances.assign(ref.names.size(), 0.0);
        chs[x]->out = g_out_pool.take(chs[x]->count * (paired ? 1400 : 700));
        chs[x]->spec_used = 0;
        chs[x]->n_commit = chs[x]->n_probe = chs[x]->n_accept = chs[x]->ns_commit = chs[x]->ns_probe =
            chs[x]->ns_write = 0;
        chs[x]->tent_prev.clear();
        chs[x]->tent_cur.clear();
        if (o.mode == 0) {
            chs[x]->text_prev = g_out_pool.take(chs[x]->count * (paired ? 1400 : 700));
            chs[x]->text_cur = g_out_pool.take(chs[x]->count * (paired ? 1400 : 700));
        }
    }
    if (fused) {
        st.fused_batches++;
        st.spec_requests += spec.req.size();
        // each chunk's view of the fused pass's results (they stay in the
        // engine's staging memory until its next batch): the requests of each of
        // its reads, found by one walk over the requests, which are sorted by read
        const uint64_t per = paired ? 2 : 1;
        parallel_for(o.threads, nch, [&](size_t x) {
            Chunk& ch = *chs[x];
            const uint64_t r0 = per * ch.pos, r1 = per * (ch.pos + ch.count);
            auto by_read = [](const ExtRequest& q, uint64_t r) { return q.read_id < r; };
            const ExtRequest* end = spec.req.end();
            const ExtRequest* lo = std::lower_bound((const ExtRequest*)spec.req.begin(), end, r0, by_read);
            const ExtRequest* hi = std::lower_bound(lo, end, r1, by_read);
            const uint64_t y0 = (uint64_t)(lo - spec.req.begin());
            ch.spec_req = lo;
            ch.spec_res = spec.res.begin() + y0;
            ch.spec_cigoff = spec.cigar_off.begin() + y0;
            ch.spec_cig = spec.cigars.begin();
            ch.spec_r0 = r0;
            ch.spec_at.assign(r1 - r0 + 1, 0);
            uint32_t y = 0;
            for (uint64_t rd = 0; rd <= r1 - r0; ++rd) {
                while (lo + y < hi && lo[y].read_id < r0 + rd) ++y;
                ch.spec_at[rd] = y;
            }
            ch.spec_seen.assign((size_t)(hi - lo), 0);
        });
    }
    for (int round = 0;; ++round) {
        double h0 = now_ms();
        parallel_for(o.threads, nch, [&](size_t x) {
            Chunk& ch = *chs[x];
            if (!ch.done() || round == 0) run_chunk(base, ch);
        });
        double h1 = now_ms();
        st.host_ms += h1 - h0;
        // requests of all chunks, chunk by chunk (built in parallel)
        std::vector<uint64_t> roff(nch + 1, 0);
        for (size_t x = 0; x < nch; ++x) roff[x + 1] = roff[x] + chs[x]->pending.size();
        const uint64_t nreq = roff[nch];
        bool all_done = true;
        for (size_t x = 0; x < nch; ++x) all_done &= chs[x]->done();
        if (nreq == 0) {
            if (!all_done) throw std::runtime_error("replay stalled without pending extensions");
            break;
        }
        std::vector<ExtRequest> req(nreq);
        parallel_for(o.threads, nch, [&](size_t x) {
            const Chunk& ch = *chs[x];
            ExtRequest* out_r = req.data() + roff[x];
            for (size_t y = 0; y < ch.pending.size(); ++y) {
                const Key& kk = ch.pending[y];
                ExtRequest r{};
                r.kind = kk.kind;
                r.read_id = kk.read;
                r.flags = kk.flags;
                r.ref_id = kk.ref_id;
                if (kk.kind == EXT_SEED) {
                    r.ref_start = kk.a; r.ref_end = kk.b; r.query_start = kk.c; r.query_end = kk.d;
                    r.anchor_off = kk.e & ((1ull << 48) - 1);
                    r.n_anchors = (uint32_t)(kk.e >> 48);
                } else {
                    r.win_start = kk.a; r.win_end = kk.b;
                }
                out_r[y] = r;
            }
        });
        st.rounds++;
        st.requests += nreq;
        std::vector<ExtResult> res;
        std::vector<uint32_t> cig;
        std::vector<uint64_t> coff;
        double g0 = now_ms();
        st.host_ms += g0 - h1;
        eng.extend(req.data(), (uint32_t)nreq, o.scores, res, cig, coff);
        double g1 = now_ms();
        st.extend_ms += g1 - g0;
        // results into the chunks' caches (each chunk owns its cache and CIGAR pool)
        parallel_for(o.threads, nch, [&](size_t x) {
            Chunk& ch = *chs[x];
            const uint64_t a = roff[x], e = roff[x + 1];
            ch.cache.reserve(ch.cache.size() + (e - a));
            ch.cigar_pool.reserve(ch.cigar_pool.size() + (coff[e] - coff[a]));
            for (uint64_t y = a; y < e; ++y) {
                CachedResult cr;
                cr.r = res[y];
                cr.cigar_off = ch.cigar_pool.size();
                ch.cigar_pool.insert(ch.cigar_pool.end(), cig.begin() + (ptrdiff_t)coff[y],
                                     cig.begin() + (ptrdiff_t)coff[y + 1]);
                ch.cache.emplace(ch.pending[y - a], cr);
            }
            ch.pending.clear();
            ch.pending_set.clear();
        });
        st.host_ms += now_ms() - g1;
    }
    for (size_t x = 0; x < nch; ++x) {
        const Chunk& ch = *chs[x];
        st.spec_used += ch.spec_used;
        st.replay_committed += ch.n_commit;
        st.replay_accepted += ch.n_accept;
        st.replay_probed += ch.n_probe;
        st.replay_commit_ms += ch.ns_commit * 1e-6;
        st.replay_probe_ms += ch.ns_probe * 1e-6;
        st.replay_write_ms += ch.ns_write * 1e-6;
    }
    // the batch's caches are no longer needed; output text stays until it is joined
    parallel_for(o.threads, nch, [&](size_t x) {
        Chunk& ch = *chs[x];
        decltype(ch.cache)().swap(ch.cache);
        std::vector<uint32_t>().swap(ch.cigar_pool);
        ch.spec_req = nullptr;
        std::vector<uint32_t>().swap(ch.spec_at);
        std::vector<uint8_t>().swap(ch.spec_seen);
        g_out_pool.give(ch.text_prev);
        g_out_pool.give(ch.text_cur);
        std::vector<Tentative>().swap(ch.tent_prev);
        std::vector<Tentative>().swap(ch.tent_cur);
        if (!keep) g_out_pool.give(ch.out);
    });
}

// --aemb table (src/main.rs output_abundances): per contig, the summed weights
// divided by the contig length.
void abundance_table(const PackedRef& ref, const std::vector<Chunk>& chunks, std::string& out) {
    std::vector<double> abundances(ref.names.size(), 0.0);
    for (const Chunk& c : chunks)
        for (size_t i = 0; i < abundances.size() && i < c.abundances.size(); ++i) abundances[i] += c.abundances[i];
    char buf[64];
    for (size_t i = 0; i < ref.names.size(); ++i) {
        snprintf(buf, sizeof buf, "\t%.6f\n", abundances[i] / (double)ref.len[i]);
        out += ref.names[i];
        out += buf;
    }
}

void add_stats(RunStats& to, const RunStats& b) {
    to.upload_ms += b.upload_ms;
    to.nam_ms += b.nam_ms;
    to.extend_ms += b.extend_ms;
    to.host_ms += b.host_ms;
    to.requests += b.requests;
    to.rounds += b.rounds;
    to.nams += b.nams;
    to.fused_batches += b.fused_batches;
    to.spec_requests += b.spec_requests;
    to.spec_used += b.spec_used;
    to.graph_replays += b.graph_replays;
    to.fused_overflows += b.fused_overflows;
    to.replay_committed += b.replay_committed;
    to.replay_accepted += b.replay_accepted;
    to.replay_probed += b.replay_probed;
    to.replay_commit_ms += b.replay_commit_ms;
    to.replay_probe_ms += b.replay_probe_ms;
    to.replay_write_ms += b.replay_write_ms;
}

}  // namespace

void map_records(const std::vector<Engine*>& engines, const PackedRef& ref, const std::vector<Record>& r1,
                 const std::vector<Record>& r2, const Options& o, std::string* out, RunStats* stats,
                 const DeviceReads* dreads) {
    if (engines.empty()) throw std::invalid_argument("map_records: no engine");
    const double t_start = now_ms();
    const bool paired = !r2.empty();
    const uint64_t n = paired ? std::min(r1.size(), r2.size()) : r1.size();
    std::vector<Chunk> chunks = make_chunks(r1.data(), paired ? r2.data() : nullptr, n, o.chunk_size);
    // device batches of whole chunks: at most batch_bases each, and as many as a
    // multiple of the engine count, so that the engines finish their last batches
    // together instead of one engine running the remainder alone
    std::vector<uint64_t> cbases(chunks.size());
    uint64_t total_bases = 0;
    for (size_t c = 0; c < chunks.size(); ++c) total_bases += cbases[c] = chunk_bases(chunks[c]);
    const uint64_t ne = engines.size();
    uint64_t nb = std::max<uint64_t>(1, (total_bases + o.batch_bases - 1) / std::max<uint64_t>(1, o.batch_bases));
    nb = (nb + ne - 1) / ne * ne;
    const uint64_t target = std::max<uint64_t>(1, (total_bases + nb - 1) / nb);
    std::vector<std::pair<size_t, size_t>> batches;
    for (size_t cb = 0; cb < chunks.size();) {
        size_t ce = cb;
        uint64_t bases = 0;
        while (ce < chunks.size() && (ce == cb || bases < target)) bases += cbases[ce++];
        batches.push_back({cb, ce});
        cb = ce;
    }
    std::vector<RunStats> bst(batches.size());
    const bool keep = out != nullptr || o.mode == 2;
    auto one = [&](Engine& eng, size_t b) {
        const size_t cb = batches[b].first, ce = batches[b].second;
        std::vector<Chunk*> chs;
        BatchReads br;
        for (size_t c = cb; c < ce; ++c) {
            chs.push_back(&chunks[c]);
            if (dreads) br.seg.push_back(dreads->off[paired ? 2 * chunks[c].first : chunks[c].first]);
        }
        if (dreads) br.d_base = dreads->d_seq;
        run_batch(eng, ref, o, paired, chs.data(), chs.size(), br, bst[b], keep);
    };
    // batches are dealt to the engines; each engine works on its own stream, so one
    // batch's device stages overlap another's host replay
    if (engines.size() == 1 || batches.size() <= 1) {
        try {
            for (size_t b = 0; b < batches.size(); ++b) one(*engines[0], b);
        } catch (...) {
            engines[0]->drain();
            throw;
        }
    } else {
        std::atomic<size_t> next{0};
        std::vector<std::thread> th;
        std::vector<std::exception_ptr> err(engines.size());
        for (size_t e = 0; e < std::min(engines.size(), batches.size()); ++e)
            th.emplace_back([&, e] {
                try {
                    for (size_t b; (b = next.fetch_add(1)) < batches.size();) one(*engines[e], b);
                } catch (...) {
                    err[e] = std::current_exception();
                    next.store(batches.size());
                    engines[e]->drain();
                }
            });
        for (auto& t : th) t.join();
        for (auto& e : err)
            if (e) std::rethrow_exception(e);
    }
    // output in input order (copied in parallel)
    if (out && o.mode != 2) {
        std::vector<uint64_t> pos(chunks.size() + 1, out->size());
        for (size_t c = 0; c < chunks.size(); ++c) pos[c + 1] = pos[c] + chunks[c].out.size();
        out->resize(pos[chunks.size()]);
        parallel_for(o.threads, chunks.size(), [&](size_t x) {
            Chunk& ch = chunks[x];
            if (!ch.out.empty()) std::memcpy(&(*out)[pos[x]], ch.out.data(), ch.out.size());
            g_out_pool.give(ch.out);
        });
    }
    if (o.mode == 2 && out) abundance_table(ref, chunks, *out);
    if (stats) {
        for (const RunStats& b : bst) add_stats(*stats, b);
        stats->records += n;
        stats->total_ms += now_ms() - t_start;
    }
}

void map_records(Engine& eng, const PackedRef& ref, const std::vector<Record>& r1, const std::vector<Record>& r2,
                 const Options& o, std::string& out, RunStats* stats) {
    map_records(std::vector<Engine*>{&eng}, ref, r1, r2, o, &out, stats, nullptr);
}

// ------------------------------------------------------------- serving ---

struct Server::Impl {
    struct Req {
        uint64_t id = 0;
        std::vector<Record> r1, r2;
        bool paired = false;
        std::vector<Chunk> chunks;
        std::vector<uint64_t> chunk_b;    // bases per chunk
        std::vector<uint64_t> chunk_seg;  // offset of each chunk's reads in `packed`, then in the arena
        std::vector<uint8_t> packed;      // the reads' bases, packed by the submitting thread
        size_t next_chunk = 0;            // next chunk to hand to a batch
        size_t left = 0;                  // chunks not finished
        uint64_t bytes = 0;               // bases of all reads
        uint64_t dev = 0, dev_size = 0;   // arena block
        double t_submit = 0, t_admit = 0, t_done = 0;
        bool done = false, failed = false;
        bool claimed = false;  // a wait call holds the ticket
        std::string out, err;
    };

    std::vector<Engine*> engines;
    const PackedRef& ref;
    Options o;
    int device;
    uint8_t* arena = nullptr;
    uint64_t arena_bytes = 0;
    std::map<uint64_t, uint64_t> free_blocks;  // offset -> size
    mutable std::mutex m;
    std::condition_variable cv_admit, cv_work, cv_done;
    std::deque<Req*> waiting;  // submitted, reads not yet on the device
    std::deque<Req*> ready;    // admitted, with chunks not yet in a batch
    uint64_t ready_bases = 0;  // bases of the chunks in `ready` not yet in a batch
    std::unordered_map<uint64_t, std::unique_ptr<Req>> reqs;
    uint64_t next_id = 1;
    bool stop = false;
    int waiters = 0;  // threads inside Server::wait
    Server::Stats st{};
    std::thread admitter;
    std::vector<std::thread> workers;

    Impl(std::vector<Engine*> e, const PackedRef& r, const Options& opt, int dev, uint64_t bytes)
        : engines(std::move(e)), ref(r), o(opt), device(dev), arena_bytes(bytes) {}

    static uint64_t round_up(uint64_t x) { return (x + 255) & ~(uint64_t)255; }

    // first fit in the arena; false when no free block is large enough
    bool alloc(uint64_t size, uint64_t& off) {
        for (auto it = free_blocks.begin(); it != free_blocks.end(); ++it) {
            if (it->second < size) continue;
            off = it->first;
            const uint64_t rest = it->second - size;
            free_blocks.erase(it);
            if (rest) free_blocks[off + size] = rest;
            return true;
        }
        return false;
    }
    void release(uint64_t off, uint64_t size) {
        auto it = free_blocks.emplace(off, size).first;
        auto nx = std::next(it);
        if (nx != free_blocks.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_blocks.erase(nx);
        }
        if (it != free_blocks.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_blocks.erase(it);
            }
        }
    }
    uint64_t arena_used() const {
        uint64_t f = 0;
        for (const auto& b : free_blocks) f += b.second;
        return arena_bytes - f;
    }

    // Admits requests in arrival order while their reads fit in the arena, and
    // copies their reads to the device (outside the lock, on its own stream).
    void admit_loop() {
        std::unique_lock<std::mutex> lk(m);
        for (;;) {
            uint64_t off = 0;
            cv_admit.wait(lk, [&] {
                return stop || (!waiting.empty() && alloc(round_up(waiting.front()->bytes + 16), off));
            });
            if (stop) return;
            Req* r = waiting.front();
            waiting.pop_front();
            r->dev = off;
            r->dev_size = round_up(r->bytes + 16);
            st.arena_peak = std::max(st.arena_peak, arena_used());
            lk.unlock();
            std::string err;
            try {
                device_copy_in(device, arena + r->dev, r->packed.data(), r->bytes);
                for (uint64_t& sgo : r->chunk_seg) sgo += r->dev;
                std::vector<uint8_t>().swap(r->packed);
            } catch (const std::exception& e) {
                err = e.what();
            }
            lk.lock();
            r->t_admit = now_ms();
            if (!err.empty()) {
                r->failed = true;
                r->err = err;
                finish_locked(r);
            } else {
                ready.push_back(r);
                ready_bases += r->bytes;
                cv_work.notify_all();
            }
        }
    }

    void finish_locked(Req* r) {
        release(r->dev, r->dev_size);
        r->dev_size = 0;
        r->done = true;
        r->t_done = now_ms();
        st.completed++;
        cv_done.notify_all();
        cv_admit.notify_all();
    }

    // One engine: forms device batches from the oldest admitted chunks of any
    // requests (same pairedness) and runs them. A batch takes batch_bases when the
    // queue is short (latency) and up to 8x that when a backlog waits (each
    // batch's fixed cost of passes and waits is then spread over more reads).
    void work_loop(size_t e) {
        std::unique_lock<std::mutex> lk(m);
        for (;;) {
            cv_work.wait(lk, [&] { return stop || !ready.empty(); });
            if (stop) return;
            std::vector<std::pair<Req*, size_t>> take;
            uint64_t bases = 0;
            const bool paired = ready.front()->paired;
            const uint64_t share = ready_bases / engines.size();
            const uint64_t limit = std::min<uint64_t>(std::max<uint64_t>(share, o.batch_bases), 8 * o.batch_bases);
            while (!ready.empty() && ready.front()->paired == paired && (take.empty() || bases < limit)) {
                Req* r = ready.front();
                const size_t c = r->next_chunk++;
                take.push_back({r, c});
                bases += r->chunk_b[c];
                ready_bases -= r->chunk_b[c];
                if (r->next_chunk == r->chunks.size()) ready.pop_front();
            }
            size_t nreq = 0;
            for (size_t i = 0; i < take.size(); ++i) nreq += i == 0 || take[i].first != take[i - 1].first;
            st.batches++;
            st.batch_requests += nreq;
            st.max_requests_per_batch = std::max<uint64_t>(st.max_requests_per_batch, nreq);
            lk.unlock();
            std::vector<Chunk*> chs;
            BatchReads br;
            br.d_base = arena;
            for (const auto& t : take) {
                chs.push_back(&t.first->chunks[t.second]);
                br.seg.push_back(t.first->chunk_seg[t.second]);
            }
            std::string err;
            RunStats rs;
            const double t0 = now_ms();
            try {
                run_batch(*engines[e], ref, o, paired, chs.data(), chs.size(), br, rs, true);
            } catch (const std::exception& ex) {
                err = ex.what();
            }
            const double t1 = now_ms();
            // requests whose last chunk this was: output joined outside the lock
            std::vector<Req*> fin;
            lk.lock();
            st.busy_ms += t1 - t0;
            for (const auto& t : take) {
                Req* r = t.first;
                if (!err.empty() && !r->failed) {
                    r->failed = true;
                    r->err = err;
                }
                if (--r->left == 0) fin.push_back(r);
            }
            lk.unlock();
            for (Req* r : fin) {
                if (r->failed) continue;
                if (o.mode == 2) {
                    abundance_table(ref, r->chunks, r->out);
                } else {
                    uint64_t total = 0;
                    for (const Chunk& ch : r->chunks) total += ch.out.size();
                    r->out.reserve(total);
                    for (Chunk& ch : r->chunks) {
                        r->out += ch.out;
                        std::string().swap(ch.out);
                    }
                }
            }
            lk.lock();
            for (Req* r : fin) finish_locked(r);
        }
    }
};

Server::Server(std::vector<Engine*> engines, const PackedRef& ref, const Options& o, int device,
               uint64_t arena_bytes)
    : impl_(std::make_unique<Impl>(std::move(engines), ref, o, device, arena_bytes)) {
    if (impl_->engines.empty()) throw std::invalid_argument("Server: no engine");
    if (arena_bytes < 4096) throw std::invalid_argument("Server: arena smaller than 4 KiB");
    impl_->arena = (uint8_t*)device_alloc(device, arena_bytes);
    impl_->free_blocks[0] = arena_bytes;
    impl_->st.arena_bytes = arena_bytes;
    impl_->admitter = std::thread([this] { impl_->admit_loop(); });
    for (size_t e = 0; e < impl_->engines.size(); ++e) impl_->workers.emplace_back([this, e] { impl_->work_loop(e); });
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lk(impl_->m);
        impl_->stop = true;
    }
    impl_->cv_admit.notify_all();
    impl_->cv_work.notify_all();
    impl_->admitter.join();
    for (auto& t : impl_->workers) t.join();
    {
        // the workers are gone: every request still unfinished fails, and the
        // threads blocked in wait return before the server's state goes away
        std::unique_lock<std::mutex> lk(impl_->m);
        for (auto& kv : impl_->reqs) {
            Impl::Req* r = kv.second.get();
            if (r->done) continue;
            r->failed = r->done = true;
            r->err = "server destroyed before the request completed";
        }
        impl_->cv_done.notify_all();
        impl_->cv_done.wait(lk, [&] { return impl_->waiters == 0; });
    }
    device_free(impl_->device, impl_->arena);
}

double steady_ms() { return now_ms(); }

uint64_t Server::submit(std::vector<Record> r1, std::vector<Record> r2, double t_arrival) {
    const double t_in = t_arrival >= 0 ? t_arrival : now_ms();
    auto r = std::make_unique<Impl::Req>();
    r->paired = !r2.empty();
    if (r->paired && r1.size() != r2.size()) throw std::invalid_argument("submit: mates differ in count");
    r->r1 = std::move(r1);
    r->r2 = std::move(r2);
    const uint64_t n = r->r1.size();
    r->chunks = make_chunks(r->r1.data(), r->paired ? r->r2.data() : nullptr, n, impl_->o.chunk_size);
    for (const Chunk& c : r->chunks) {
        r->chunk_b.push_back(chunk_bases(c));
        r->bytes += r->chunk_b.back();
    }
    // the reads packed here, by the submitting thread, so that admission is one copy
    r->packed.resize(r->bytes);
    {
        uint64_t at = 0;
        for (const Chunk& ch : r->chunks) {
            r->chunk_seg.push_back(at);
            for (uint64_t i = 0; i < ch.count; ++i) {
                std::memcpy(r->packed.data() + at, ch.rec1[i].seq.data(), ch.rec1[i].seq.size());
                at += ch.rec1[i].seq.size();
                if (r->paired) {
                    std::memcpy(r->packed.data() + at, ch.rec2[i].seq.data(), ch.rec2[i].seq.size());
                    at += ch.rec2[i].seq.size();
                }
            }
        }
    }
    if (Impl::round_up(r->bytes + 16) > impl_->arena_bytes)
        throw std::invalid_argument("submit: request reads (" + std::to_string(r->bytes) +
                                    " bases) exceed the device arena");
    r->left = r->chunks.size();
    std::lock_guard<std::mutex> lk(impl_->m);
    r->id = impl_->next_id++;
    r->t_submit = t_in;
    Impl::Req* p = r.get();
    impl_->reqs.emplace(p->id, std::move(r));
    impl_->st.submitted++;
    if (p->chunks.empty()) {
        // no records: nothing to map (an abundance table still lists every contig)
        if (impl_->o.mode == 2) abundance_table(impl_->ref, p->chunks, p->out);
        p->t_admit = p->t_submit;
        p->dev_size = 0;
        p->done = true;
        p->t_done = p->t_submit;
        impl_->st.completed++;
        impl_->cv_done.notify_all();
    } else {
        impl_->waiting.push_back(p);
        impl_->cv_admit.notify_all();
    }
    return p->id;
}

Server::WaitStatus Server::wait(uint64_t ticket, std::string& out, RequestTimes& t, std::string& err) {
    std::unique_lock<std::mutex> lk(impl_->m);
    auto it = impl_->reqs.find(ticket);
    if (it == impl_->reqs.end()) {
        err = "unknown ticket " + std::to_string(ticket);
        return WAIT_BAD_TICKET;
    }
    Impl::Req* r = it->second.get();
    if (r->claimed) {
        err = "ticket " + std::to_string(ticket) + " is already being waited for";
        return WAIT_BAD_TICKET;
    }
    r->claimed = true;
    impl_->waiters++;
    impl_->cv_done.wait(lk, [&] { return r->done; });
    t.admit_ms = r->t_admit - r->t_submit;
    t.done_ms = r->t_done - r->t_submit;
    const bool ok = !r->failed;
    if (ok) out.swap(r->out);
    else err = r->err;
    impl_->reqs.erase(ticket);  // by key: submits during the wait may have rehashed the table
    impl_->waiters--;
    if (impl_->stop) impl_->cv_done.notify_all();  // the destructor waits for the last waiter
    return ok ? WAIT_OK : WAIT_FAILED;
}

Server::Stats Server::stats() const {
    std::lock_guard<std::mutex> lk(impl_->m);
    return impl_->st;
}

}  // namespace sa
