// CPU oracle for the strobealign read-mapping hot path.
//
// TEST INFRASTRUCTURE ONLY. This is a plain, sequential C++ restatement of the
// reference (strobealign 0.18.0-alpha, Rust) used to check the HIP product.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// build, load or run it; the product library never links it.
//
// Every function cites the reference file:line whose behaviour it restates.
// Parity status of each part is listed in the oracle's Makefile header and
// in the test suite; functions whose behaviour the reference's own fixtures
// and known-answer tests cannot pin say "parity unpinned" next to them.
#pragma once
#include <chrono>
#include <atomic>
#include <cstdint>
#include <cstddef>
#include <string>
#include <vector>
#include <optional>

namespace orc {

// Per-thread stage clocks in seconds, the timers of src/main.rs:821-846:
// randstrobes, find_hits (with the anchor expansion, as chainer.rs:222-224 counts
// it), chaining (anchor sort, collinear chaining, chain extraction), sort_nams, and
// extend (extend_seed and rescue_align).
struct StageTimes {
    double randstrobes = 0, find_hits = 0, chaining = 0, sort_nams = 0, extend = 0;
};
StageTimes& stage_times();  // the calling thread's
struct StageTimer {
    double& acc;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    explicit StageTimer(double& a) : acc(a) {}
    ~StageTimer() { acc += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

// ---------------------------------------------------------------- seeding ---
// src/seeding/hash.rs:11-27
uint64_t xxh64(uint64_t input);

struct SyncmerParams { int k, s, t; };
struct RandstrobeParams {
    int w_min, w_max;
    uint64_t q;
    int max_dist;
    uint64_t main_hash_mask;
    uint64_t forward_main_hash_mask;
    int partial_orientation_pos;
};
struct SeedingParams {
    int profile;  // 0 noisy, 50,75,100,125,150,250,400 = canonical read length
    SyncmerParams syncmer;
    RandstrobeParams randstrobe;
};
// src/seeding/parameters.rs:181-221 (SeedingParameters::new(read_length))
SeedingParams seeding_params_for_read_length(size_t read_length);

struct Syncmer { uint64_t hash; bool is_forward; size_t position; };
struct Randstrobe { uint64_t hash, hash_revcomp; size_t strobe1_pos, strobe2_pos; };
struct QueryRandstrobe { uint64_t hash, hash_revcomp; size_t start, end; };

// src/seeding/syncmers.rs:171-237 over ASCII (N resets) or 2-bit codes
std::vector<Syncmer> syncmers_ascii(const uint8_t* seq, size_t len, const SyncmerParams& p);
std::vector<Syncmer> syncmers_codes(const uint8_t* codes, size_t len, const SyncmerParams& p);
// src/seeding/strobes.rs:145-176
std::vector<Randstrobe> randstrobes_from_syncmers(const std::vector<Syncmer>& s, const RandstrobeParams& p);
// src/seeding/strobes.rs:112-124
uint64_t randstrobe_hash(uint64_t h1, uint64_t h2, uint64_t fwd1, uint64_t fwd2, const RandstrobeParams& p);
// src/seeding/mod.rs:19-75
void randstrobes_query(const uint8_t* seq, size_t len, const SeedingParams& p,
                       std::vector<QueryRandstrobe> out[2]);

// ----------------------------------------------------------- references ---
// src/packed_seq.rs:12-126
struct PackedSeq {
    std::vector<uint64_t> data;
    size_t len = 0;
    void push(uint8_t c);
    uint8_t bits(size_t i) const { return (data[i >> 5] >> ((i & 31) * 2)) & 3; }
    uint8_t decode_at(size_t i) const { return "ACGT"[bits(i)]; }
    std::string decode(size_t start, size_t end) const;
};
struct RefSeq { std::string name; PackedSeq seq; };
// src/io/fasta.rs:116-165
std::vector<RefSeq> read_fasta(const std::string& path);

// ---------------------------------------------------------------- index ---
// src/index.rs:15-65
struct RefRandstrobe {
    uint64_t hash_offset;
    uint32_t position;
    uint32_t ref_index;
    uint64_t hash() const { return hash_offset & 0xFFFFFFFFFFFFFF00ULL; }
    size_t strobe2_offset() const { return hash_offset & 0xFF; }
};
static_assert(sizeof(RefRandstrobe) == 16, "RefRandstrobe layout");

struct Index {
    SeedingParams params;
    int bits = 0;
    size_t filter_cutoff = 0;
    std::vector<RefRandstrobe> randstrobes;
    std::vector<uint64_t> bucket_starts;
    int k() const { return params.syncmer.k; }
    size_t size() const { return randstrobes.size(); }
    // src/index.rs:159-197 (returns SIZE_MAX when absent)
    size_t get_masked_from(uint64_t hash, uint64_t mask, size_t start /*SIZE_MAX = bucket start*/) const;
    size_t get_full_forward(uint64_t h) const { return get_masked_from(h, 0xFFFFFFFFFFFFFF00ULL, SIZE_MAX); }
    size_t get_partial(uint64_t h) const { return get_masked_from(h, params.randstrobe.main_hash_mask, SIZE_MAX); }
    // src/index.rs:272-295
    size_t get_count(size_t position, uint64_t mask) const;
    size_t get_count_full(size_t position, uint64_t hash_revcomp) const;
    // src/index.rs:298-338
    bool is_too_frequent_forward(size_t position, size_t cutoff) const;
    bool is_too_frequent(size_t position, size_t cutoff, uint64_t hash_revcomp) const;
    bool is_too_frequent_partial(size_t position, size_t cutoff) const;
};
// src/indexer.rs:294-300
int pick_bits(const std::vector<RefSeq>& refs, const SyncmerParams& p);
// src/indexer.rs:18-170
Index make_index(const std::vector<RefSeq>& refs, const SeedingParams& p, int bits, double filter_fraction);

// ----------------------------------------------------------- hits/chain ---
enum class Mcs { Always, Rescue, Off, FirstStrobe };  // src/mcsstrategy.rs:4-19

struct Hit {
    size_t query_start, query_end, position;
    uint64_t hash, hash_revcomp;
    bool is_partial, is_filtered;
};
struct HitsDetails {
    size_t full_not_found = 0, full_filtered = 0, full_found = 0;
    size_t partial_not_found = 0, partial_filtered = 0, partial_found = 0, rescued = 0;
};
// src/hit.rs:270-333
std::vector<Hit> find_hits(const std::vector<QueryRandstrobe>& qr, const Index& index, Mcs mcs,
                           size_t filter_cutoff, size_t rescue_distance, HitsDetails& details);

struct Anchor {
    size_t ref_id, ref_start, query_start;
    bool operator<(const Anchor& o) const {
        if (ref_id != o.ref_id) return ref_id < o.ref_id;
        if (ref_start != o.ref_start) return ref_start < o.ref_start;
        return query_start < o.query_start;
    }
    bool operator==(const Anchor& o) const {
        return ref_id == o.ref_id && ref_start == o.ref_start && query_start == o.query_start;
    }
};

struct Nam {
    size_t nam_id, ref_start, ref_end, query_start, query_end, matching_bases, ref_id;
    float score;
    bool is_revcomp;
    std::vector<Anchor> anchors;
    size_t projected_ref_start() const { return ref_start > query_start ? ref_start - query_start : 0; }
};

struct ChainingParams {
    size_t max_lookback = 50;
    float diag_diff_penalty = 0.1f, gap_length_penalty = 0.05f, valid_score_threshold = 0.7f;
    long max_ref_gap = -1;  // -1: read length
    float matches_weight = 0.01f, max_diagonal_ratio = 10.0f;
};

struct NamDetails {
    HitsDetails hits;
    size_t n_reads = 0, n_randstrobes = 0, n_anchors = 0, n_nams = 0;
    bool both_orientations = false;
};

// src/chainer.rs:182-267 + src/nam.rs:138-166
std::vector<Nam> get_nams_by_chaining(const uint8_t* seq, size_t len, const Index& index,
                                      const ChainingParams& cp, size_t rescue_distance, Mcs mcs,
                                      NamDetails& details);
// chaining pieces exposed for known-answer tests (src/chainer.rs:85-180, 286-302)
float chaining_score(size_t dq, size_t dr, size_t k, const ChainingParams& p);
float collinear_chaining_best(const std::vector<Anchor>& anchors, size_t k, size_t read_len,
                              const ChainingParams& p);

// ------------------------------------------------------------------ rng ---
// fastrand 2.3.0 Rng (wyrand + Lemire bounded sampling). Parity unpinned: the
// crate is not vendored; restated from the published wyrand generator.
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t gen_u64();
    uint32_t gen_mod_u32(uint32_t n);
    uint64_t gen_mod_u64(uint64_t n);
};

// src/nam.rs:168-188 + src/shuffle.rs:7-25
void sort_nams(std::vector<Nam>& nams, Rng& rng);
// bounded draws with n >= 2 made so far by every Rng of the process
std::atomic<uint64_t>& rng_draws();

// ---------------------------------------------------------------- cigar ---
// src/cigar.rs:4-243. op codes: 0 M 1 I 2 D 3 N 4 S 5 H 6 P 7 = 8 X
struct Cigar {
    std::vector<std::pair<uint8_t, uint32_t>> ops;
    void push(uint8_t op, size_t len);
    void extend(const Cigar& o) { for (auto& p : o.ops) push(p.first, p.second); }
    Cigar reversed() const { Cigar c; c.ops.assign(ops.rbegin(), ops.rend()); return c; }
    Cigar with_m() const;
    Cigar with_eqx(const uint8_t* q, const uint8_t* r) const;
    size_t edit_distance() const;
    std::string str() const;
};

// -------------------------------------------------------------- aligner ---
struct Scores { int match = 2, mismatch = 8, gap_open = 12, gap_extend = 1, end_bonus = 10; };

struct AlignmentInfo {
    Cigar cigar;
    size_t edit_distance = 0, ref_start = 0, ref_end = 0, query_start = 0, query_end = 0;
    uint32_t score = 0;
};

// Result of one SSW call, same fields as ext/ssw/ssw.h s_align (score1,
// ref_begin1, ref_end1, read_begin1, read_end1, flag) plus the BAM cigar.
struct SswResult {
    uint16_t score1 = 0, score2 = 0;
    int32_t ref_begin1 = -1, ref_end1 = -1, read_begin1 = -1, read_end1 = -1, ref_end2 = -1;
    uint16_t flag = 0;
    std::vector<uint32_t> cigar;
};
// ext/ssw/ssw.c:789-922 restated (striped u8/i16 lanes emulated), flag semantics as used by
// src/ssw.rs:156-191 (flag 0x0f, filters 0, filterd INT_MAX)
SswResult ssw_align(const int8_t* read, int read_len, const int8_t* ref, int ref_len,
                    const int8_t* mat, int n, int gap_open, int gap_extend, int mask_len);
// src/ssw.rs:156-191 + src/aligner.rs:80-151
std::optional<AlignmentInfo> aligner_align(const uint8_t* query, size_t qlen, const uint8_t* ref,
                                           size_t rlen, const Scores& sc);
// src/aligner.rs:166-259
std::optional<AlignmentInfo> hamming_align(const uint8_t* q, const uint8_t* r, size_t len, const Scores& sc);
void highest_scoring_segment(const uint8_t* q, const uint8_t* r, size_t n, int match, int mismatch,
                             int end_bonus, size_t& start, size_t& end, int& score);
AlignmentInfo hamming_align_global(const uint8_t* q, const uint8_t* r, size_t n, int match, int mismatch);

// src/piecewisealigner.rs. block-aligner 0.5.1 is restated as exact DP with a
// fixed traceback preference: parity unpinned beyond the reference's KATs.
struct PwResult { int score = 0; size_t query_start = 0, query_end = 0, ref_start = 0, ref_end = 0; Cigar cigar; };
PwResult pw_global_alignment(const uint8_t* q, size_t ql, const uint8_t* r, size_t rl, const Scores& sc);
PwResult pw_xdrop_alignment(const uint8_t* q, size_t ql, const uint8_t* r, size_t rl, const Scores& sc,
                            bool global_start_local_end, int xdrop);
std::optional<AlignmentInfo> extend_piecewise(const uint8_t* q, size_t ql, const uint8_t* r, size_t rl,
                                              const std::vector<Anchor>& anchors, size_t padding,
                                              const Scores& sc, size_t k, int xdrop);
void remove_spurious_anchors(std::vector<Anchor>& anchors);

// --------------------------------------------------------------- mapper ---
struct Record {
    std::string name, comment;
    bool has_comment = false;
    std::string seq, qual;
    bool has_qual = true;
    int end = 0;  // 0 none, 1, 2 (src/io/record.rs:12-23)
};

struct MappingParams {
    size_t max_secondary = 0;
    float dropoff_threshold = 0.5f;
    size_t rescue_distance = 100;
    size_t max_tries = 20;
    Mcs mcs = Mcs::Always;
    bool output_unmapped = true;
    bool use_ssw = false;
};

struct SamOpts {
    bool eqx = false, details = false, fastq_comments = false;
    bool has_rg = false;
    std::string rg_id;
};

struct Mapper {
    const Index* index;
    const std::vector<RefSeq>* refs;
    MappingParams mp;
    ChainingParams cp;
    Scores sc;
    SamOpts so;
    int xdrop = 500;
    int mode = 0;  // 0 SAM, 1 PAF, 2 abundances
    std::vector<double> abundances;
    // src/main.rs:873-988: one chunk with fresh Rng(0) and insert-size state
    void map_chunk(const std::vector<std::pair<Record, std::optional<Record>>>& chunk, std::string& out);
};

// Test hook for the batch drop-in: the NAMs of one single-end read after the
// stable score sort (src/nam.rs:170, no shuffle) and extend_seed
// (src/mapper.rs:532-620) of the first `spec` of them that are consistent, in
// NAM order (the try loop's order without its stopping rules).
struct SpecExtension {
    size_t nam_index = 0;
    bool found = false, gapped = false;
    uint32_t score = 0;
    size_t ref_start = 0, length = 0, query_start = 0, query_end = 0, edit_distance = 0;
    std::string cigar;
};
std::vector<SpecExtension> speculative_extensions(const Mapper& m, const std::string& seq, size_t spec);

// src/io/sam.rs:177-201
std::string sam_header(const std::vector<RefSeq>& refs, const SamOpts& so,
                       const std::vector<std::string>& rg_fields);
// src/main.rs:1001-1012 estimate_read_length
size_t estimate_read_length(const std::vector<Record>& recs);

// src/io/fastq.rs:22-95 (also FASTA reads, src/io/reads.rs:121-135)
bool read_reads(const std::string& path, std::vector<Record>& out, std::string& err);

}  // namespace orc
