// C entry points into the oracle for the Python test-suite (ctypes).
// TEST INFRASTRUCTURE ONLY (see oracle.hpp).
#include "oracle.hpp"

#include <cstring>
#include <string>
#include <vector>

using namespace orc;

namespace {
void put_cigar(const Cigar& c, char* buf, size_t cap) {
    std::string s = c.str();
    if (cap == 0) return;
    size_t n = std::min(cap - 1, s.size());
    std::memcpy(buf, s.data(), n);
    buf[n] = 0;
}
Scores scores_of(const int* sc) {
    Scores s;
    s.match = sc[0]; s.mismatch = sc[1]; s.gap_open = sc[2]; s.gap_extend = sc[3]; s.end_bonus = sc[4];
    return s;
}
}  // namespace

extern "C" {

uint64_t orc_xxh64(uint64_t x) { return xxh64(x); }

// Fills params as 8 int64: k s t w_min w_max max_dist q main_hash_mask
void orc_seeding_params(uint64_t read_length, int64_t* out) {
    SeedingParams p = seeding_params_for_read_length(read_length);
    out[0] = p.syncmer.k; out[1] = p.syncmer.s; out[2] = p.syncmer.t;
    out[3] = p.randstrobe.w_min; out[4] = p.randstrobe.w_max; out[5] = p.randstrobe.max_dist;
    out[6] = (int64_t)p.randstrobe.q; out[7] = (int64_t)p.randstrobe.main_hash_mask;
}

// Syncmers of an ASCII sequence for (k, s): writes hash, is_forward, position triples.
int64_t orc_syncmers(const uint8_t* seq, uint64_t len, int k, int s, uint64_t* hash, uint8_t* fwd, uint64_t* pos,
                     int64_t cap) {
    SyncmerParams p{k, s, (k - s) / 2 + 1};
    auto v = syncmers_ascii(seq, len, p);
    int64_t n = (int64_t)v.size();
    for (int64_t i = 0; i < n && i < cap; ++i) { hash[i] = v[i].hash; fwd[i] = v[i].is_forward; pos[i] = v[i].position; }
    return n;
}

// Query randstrobes of a read under the profile for `read_length`.
// out arrays (cap entries each, per orientation o at offset o*cap): hash, hash_rc, start, end.
void orc_randstrobes_query(const uint8_t* seq, uint64_t len, uint64_t read_length, uint64_t* hash, uint64_t* hash_rc,
                           uint64_t* start, uint64_t* end, int64_t cap, int64_t* counts) {
    SeedingParams p = seeding_params_for_read_length(read_length);
    std::vector<QueryRandstrobe> q[2];
    randstrobes_query(seq, len, p, q);
    for (int o = 0; o < 2; ++o) {
        counts[o] = (int64_t)q[o].size();
        for (int64_t i = 0; i < counts[o] && i < cap; ++i) {
            hash[o * cap + i] = q[o][i].hash;
            hash_rc[o * cap + i] = q[o][i].hash_revcomp;
            start[o * cap + i] = q[o][i].start;
            end[o * cap + i] = q[o][i].end;
        }
    }
}

// Chaining: anchors as triples (ref_id, ref_start, query_start); returns best score.
float orc_chain_best(const uint64_t* triples, int64_t n, uint64_t k, uint64_t read_len) {
    std::vector<Anchor> a;
    for (int64_t i = 0; i < n; ++i) a.push_back({triples[3 * i], triples[3 * i + 1], triples[3 * i + 2]});
    ChainingParams cp;
    return collinear_chaining_best(a, k, read_len, cp);
}

float orc_chain_score(uint64_t dq, uint64_t dr, uint64_t k) {
    ChainingParams cp;
    return chaining_score(dq, dr, k, cp);
}

// Hamming alignment; returns score, fills segment and cigar.
int orc_hamming_align(const uint8_t* q, const uint8_t* r, uint64_t n, const int* sc, uint64_t* qs, uint64_t* qe,
                      uint64_t* ed, char* cigar, uint64_t cap) {
    auto a = hamming_align(q, r, n, scores_of(sc));
    *qs = a->query_start; *qe = a->query_end; *ed = a->edit_distance;
    put_cigar(a->cigar, cigar, cap);
    return (int)a->score;
}

void orc_highest_scoring_segment(const uint8_t* q, const uint8_t* r, uint64_t n, int match, int mismatch,
                                 int end_bonus, uint64_t* s, uint64_t* e, int* score) {
    size_t a, b;
    highest_scoring_segment(q, r, n, match, mismatch, end_bonus, a, b, *score);
    *s = a; *e = b;
}

// SSW aligner (src/aligner.rs:80-151). Returns 0 when no alignment; else 1.
// out: score, query_start, query_end, ref_start, ref_end, edit_distance
int orc_ssw_align(const uint8_t* q, uint64_t ql, const uint8_t* r, uint64_t rl, const int* sc, int64_t* out,
                  char* cigar, uint64_t cap) {
    auto a = aligner_align(q, ql, r, rl, scores_of(sc));
    if (!a) return 0;
    out[0] = a->score; out[1] = (int64_t)a->query_start; out[2] = (int64_t)a->query_end;
    out[3] = (int64_t)a->ref_start; out[4] = (int64_t)a->ref_end; out[5] = (int64_t)a->edit_distance;
    put_cigar(a->cigar, cigar, cap);
    return 1;
}

// Raw striped SSW on translated codes. out: score1, ref_begin1, ref_end1, read_begin1, read_end1, flag, ncigar
void orc_ssw_raw(const int8_t* read, int read_len, const int8_t* ref, int ref_len, const int8_t* mat, int n, int go,
                 int ge, int64_t* out, uint32_t* cigar, int64_t cap) {
    SswResult s = ssw_align(read, read_len, ref, ref_len, mat, n, go, ge, std::max(read_len / 2, 15));
    out[0] = s.score1; out[1] = s.ref_begin1; out[2] = s.ref_end1; out[3] = s.read_begin1; out[4] = s.read_end1;
    out[5] = s.flag; out[6] = (int64_t)s.cigar.size();
    for (int64_t i = 0; i < (int64_t)s.cigar.size() && i < cap; ++i) cigar[i] = s.cigar[i];
}

// Piecewise pieces. mode: 0 global, 1 xdrop GlobalStartLocalEnd, 2 xdrop LocalStartGlobalEnd.
// out: score, query_start, query_end, ref_start, ref_end
void orc_pw_align(const uint8_t* q, uint64_t ql, const uint8_t* r, uint64_t rl, const int* sc, int mode, int xdrop,
                  int64_t* out, char* cigar, uint64_t cap) {
    PwResult p = mode == 0 ? pw_global_alignment(q, ql, r, rl, scores_of(sc))
                           : pw_xdrop_alignment(q, ql, r, rl, scores_of(sc), mode == 1, xdrop);
    out[0] = p.score; out[1] = (int64_t)p.query_start; out[2] = (int64_t)p.query_end;
    out[3] = (int64_t)p.ref_start; out[4] = (int64_t)p.ref_end;
    put_cigar(p.cigar, cigar, cap);
}

// extend_piecewise with anchors given as (ref_start, query_start) pairs, last anchor first.
int orc_extend_piecewise(const uint8_t* q, uint64_t ql, const uint8_t* r, uint64_t rl, const uint64_t* pairs,
                         int64_t n, uint64_t padding, const int* sc, uint64_t k, int xdrop, int64_t* out, char* cigar,
                         uint64_t cap) {
    std::vector<Anchor> a;
    for (int64_t i = 0; i < n; ++i) a.push_back({0, pairs[2 * i], pairs[2 * i + 1]});
    auto res = extend_piecewise(q, ql, r, rl, a, padding, scores_of(sc), k, xdrop);
    if (!res) return 0;
    out[0] = res->score; out[1] = (int64_t)res->query_start; out[2] = (int64_t)res->query_end;
    out[3] = (int64_t)res->ref_start; out[4] = (int64_t)res->ref_end; out[5] = (int64_t)res->edit_distance;
    put_cigar(res->cigar, cigar, cap);
    return 1;
}

// remove_spurious_anchors over (ref_start, query_start) pairs in place; returns the new count.
int64_t orc_remove_spurious_anchors(uint64_t* pairs, int64_t n) {
    std::vector<Anchor> a;
    for (int64_t i = 0; i < n; ++i) a.push_back({0, pairs[2 * i], pairs[2 * i + 1]});
    remove_spurious_anchors(a);
    for (size_t i = 0; i < a.size(); ++i) { pairs[2 * i] = a[i].ref_start; pairs[2 * i + 1] = a[i].query_start; }
    return (int64_t)a.size();
}

// fastrand stream for the RNG tests: n outputs of gen_u64 from seed.
void orc_rng_u64(uint64_t seed, int64_t n, uint64_t* out) {
    Rng r(seed);
    for (int64_t i = 0; i < n; ++i) out[i] = r.gen_u64();
}

}  // extern "C"

namespace {
// The index built from `fasta` for `read_length`, kept for later calls with the
// same arguments.
struct CachedIndex {
    std::string path;
    uint64_t rl = 0;
    std::vector<RefSeq> refs;
    Index index;
};
CachedIndex& cached_index(const char* fasta, uint64_t read_length) {
    static CachedIndex c;
    if (c.path != fasta || c.rl != read_length) {
        SeedingParams sp = seeding_params_for_read_length(read_length);
        c.refs = read_fasta(fasta);
        c.index = make_index(c.refs, sp, pick_bits(c.refs, sp.syncmer), 0.0002);
        c.path = fasta;
        c.rl = read_length;
    }
    return c;
}
}  // namespace

extern "C" {

// find_hits (src/hit.rs:270-333) of both orientations of one read against the
// index built from `fasta` for `read_length`. Orientation o writes counts[o] hits
// of 7 words (query_start, query_end, position, hash, hash_revcomp, is_partial,
// is_filtered) at out + o * cap * 7 and its 7 detail counters at details + o * 7.
// Returns -1 if an orientation has more than cap hits.
int64_t orc_find_hits(const char* fasta, uint64_t read_length, const uint8_t* seq, uint64_t len, int mcs,
                      uint64_t rescue_distance, uint64_t cap, uint64_t* out, uint64_t* counts, uint64_t* details) {
    const Index& index = cached_index(fasta, read_length).index;
    std::vector<QueryRandstrobe> qr[2];
    randstrobes_query(seq, len, index.params, qr);
    for (int o = 0; o < 2; ++o) {
        HitsDetails d;
        std::vector<Hit> hits = find_hits(qr[o], index, (Mcs)mcs, index.filter_cutoff, rescue_distance, d);
        if (hits.size() > cap) return -1;
        counts[o] = hits.size();
        uint64_t* w = out + (uint64_t)o * cap * 7;
        for (const Hit& h : hits) {
            *w++ = h.query_start; *w++ = h.query_end; *w++ = h.position; *w++ = h.hash; *w++ = h.hash_revcomp;
            *w++ = h.is_partial; *w++ = h.is_filtered;
        }
        uint64_t* dd = details + o * 7;
        dd[0] = d.full_not_found; dd[1] = d.full_filtered; dd[2] = d.full_found; dd[3] = d.partial_not_found;
        dd[4] = d.partial_filtered; dd[5] = d.partial_found; dd[6] = d.rescued;
    }
    return 0;
}

// speculative_extensions of one single-end read (index as orc_find_hits, default
// scores unless sc is given): per extension 10 words (nam_index, found, gapped,
// score, ref_start, length, query_start, query_end, edit_distance, cigar length)
// at out + 10 i and its CIGAR text in cigars, each after the previous one. Returns
// the number of extensions, or -1 if cap or cigar_cap is too small.
int64_t orc_speculative_extensions(const char* fasta, uint64_t read_length, const char* seq, int use_ssw,
                                   uint64_t spec, const int* sc, uint64_t cap, uint64_t* out, char* cigars,
                                   uint64_t cigar_cap) {
    CachedIndex& ci = cached_index(fasta, read_length);
    Mapper m;
    m.index = &ci.index;
    m.refs = &ci.refs;
    m.mp.use_ssw = use_ssw != 0;
    if (sc) m.sc = scores_of(sc);
    std::vector<SpecExtension> ext = speculative_extensions(m, seq, spec);
    if (ext.size() > cap) return -1;
    uint64_t at = 0;
    for (size_t i = 0; i < ext.size(); ++i) {
        const SpecExtension& e = ext[i];
        uint64_t* w = out + 10 * i;
        w[0] = e.nam_index; w[1] = e.found; w[2] = e.gapped; w[3] = e.score; w[4] = e.ref_start; w[5] = e.length;
        w[6] = e.query_start; w[7] = e.query_end; w[8] = e.edit_distance; w[9] = e.cigar.size();
        if (at + e.cigar.size() > cigar_cap) return -1;
        std::memcpy(cigars + at, e.cigar.data(), e.cigar.size());
        at += e.cigar.size();
    }
    return (int64_t)ext.size();
}

}  // extern "C"
