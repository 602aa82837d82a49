// Plain-old-data layouts shared by the HIP kernels, the host driver and the
// C ABI (include/sa_hip.h). Everything here is trivially copyable so it can be
// memcpy'd between host and device unchanged.
#pragma once
#include <cstddef>
#include <cstdint>

#if defined(__HIPCC__)
#define SA_HD __host__ __device__ inline
#define SA_UNROLL _Pragma("unroll")
#else
#define SA_HD inline
#define SA_UNROLL
#endif

namespace sa {

constexpr uint64_t REF_HASH_MASK = 0xFFFFFFFFFFFFFF00ULL;  // src/index.rs:33
constexpr uint32_t NONE32 = 0xFFFFFFFFu;
constexpr uint64_t NONE64 = ~0ULL;

// Seeding + index parameters (src/seeding/parameters.rs:163-176, src/index.rs:68-87).
struct SeedParams {
    int32_t k, s, t;
    int32_t w_min, w_max, max_dist;
    uint64_t q;
    uint64_t main_mask;       // main_hash_mask
    uint64_t fwd_main_mask;   // forward_main_hash_mask
    int32_t partial_pos;      // partial_orientation_pos
    int32_t bits;             // bucket bits
    uint64_t filter_cutoff;
    uint64_t n_entries;       // randstrobes.len()
};

// 16-byte index entry, same layout as RefRandstrobe (src/index.rs:15-30).
struct RefEntry {
    uint64_t hash_offset;
    uint32_t position;
    uint32_t ref_index;
};
static_assert(sizeof(RefEntry) == 16, "RefEntry must match RefRandstrobe");

// Query randstrobe (src/seeding/mod.rs:10-16).
struct QRand {
    uint64_t hash, hash_rc;
    uint32_t start, end;
};

// Hit (src/hit.rs:19-28) plus the anchor run the chainer will expand.
struct Hit {
    uint64_t hash, hash_rc;
    uint64_t position;
    uint64_t run_start;   // first entry of the run that becomes anchors (NONE64: none)
    uint32_t run_len;
    uint32_t query_start, query_end;
    uint32_t flags;       // bit0 partial, bit1 filtered
};

// HitsDetails (src/hit.rs:32-52), per read orientation.
struct HitCounts {
    uint32_t full_not_found, full_filtered, full_found;
    uint32_t partial_not_found, partial_filtered, partial_found;
    uint32_t rescued;
    uint32_t n_hits;
};

// Chaining parameters (src/chainer.rs:22-44), Scores (src/aligner.rs:11-30).
struct ChainParams {
    uint32_t max_lookback;
    float diag_diff_penalty, gap_length_penalty, valid_score_threshold;
    int64_t max_ref_gap;  // <0: read length
    float matches_weight, max_diagonal_ratio;
};

// Which extensions the fused batch pass computes ahead of the read-level
// decisions (speculatively: the decisions take what they need and request the
// rest). Single-end: extend_seed of each consistent NAM among the first
// max_tries (all NAMs tied at the top score when there are more of them) down
// to the score dropoff (src/mapper.rs:387-406). Paired-end: extend_seed of each
// mate's top-scoring NAMs, as the proper-pair and individual paths of
// src/mapper.rs:808-866 call it. At most max_per_read per read.
struct SpecParams {
    uint32_t paired;
    uint32_t use_ssw;  // single-end: SSW instead of piecewise extension (--ssw)
    uint32_t max_tries;
    float dropoff;
    uint32_t max_per_read;
};

struct MapParams {
    uint32_t mcs;  // 0 Always, 1 Rescue, 2 Off, 3 FirstStrobe (src/mcsstrategy.rs:4-19)
    uint32_t rescue_distance;
    // lanes per read orientation in the lookup stage: 1 (a thread runs the
    // orientation's lookups in turn) or 8, 16, 64 (a lane per randstrobe); the
    // hits are the same either way
    uint32_t lookup_lanes = 16;
    // lanes per read in seeding: 1 (a thread scans the read) or 64 (a wave, a
    // lane per position and then per syncmer); the randstrobes are the same
    uint32_t seed_lanes = 64;
};

// NAM (src/nam.rs:18-31). Chain anchors are stored separately as
// (ref_start, query_start) pairs, last anchor first, at `anchor_off`.
struct NamRec {
    uint32_t nam_id, ref_id;
    uint32_t ref_start, ref_end;
    uint32_t query_start, query_end;
    uint32_t matching_bases, n_anchors;
    float score;
    uint32_t flags;  // bit0 is_revcomp; bit1 is_consistent(as is); bits2..3 reverse_nam_if_needed: 0 as-is ok, 1 flipped ok, 2 inconsistent
    uint64_t anchor_off;
};
static_assert(sizeof(NamRec) == 48, "NamRec layout");

// Per-read NAM-finding details (src/details.rs:8-31 NamDetails without timers).
struct ReadNamInfo {
    HitCounts hits[2];
    uint32_t n_randstrobes, n_anchors, n_nams;
    uint32_t both_orientations;
};

// Extension request / result (src/mapper.rs:532-620 extend_seed, 1094-1153 rescue_align).
// The mapping path issues EXT_SEED and EXT_RESCUE. The other kinds are the
// aligner's own entry points on an explicit reference window [win_start,
// win_end) of contig ref_id (the per-stage drop-in sa_extend_batch):
//   EXT_HAMMING   hamming_align (src/aligner.rs:223-259), None when the lengths differ;
//   EXT_ALIGN_PW  Aligner::align_piecewise (src/aligner.rs:153-163): extend_piecewise
//                 with the caller's anchors (anchor_off, n_anchors: (ref_start,
//                 query_start) pairs, last anchor first, contig coordinates), no
//                 pruning; ref_start holds the padding;
//   EXT_PW_GLOBAL global_alignment (src/piecewisealigner.rs:84-120);
//   EXT_PW_XDROP  xdrop_alignment (src/piecewisealigner.rs:161-229), GlobalStartLocalEnd,
//                 or LocalStartGlobalEnd when flags bit0 is set.
// The last two report the i32 score (score field, two's complement) with the
// end coordinates of the block aligner, found set.
enum ExtKind : uint32_t {
    EXT_SEED = 0,
    EXT_RESCUE = 1,
    EXT_HAMMING = 2,
    EXT_ALIGN_PW = 3,
    EXT_PW_GLOBAL = 4,
    EXT_PW_XDROP = 5
};

struct ExtRequest {
    uint32_t kind;
    uint32_t read_id;       // read whose sequence is aligned
    uint32_t flags;         // SEED: bit0 use_ssw, bit1 consistent, bit2 query is reverse complement
                            // RESCUE: bit2 query is reverse complement
    uint32_t ref_id;
    // SEED: the NAM as the caller sees it (possibly flipped by reverse_nam_if_needed)
    uint32_t ref_start, ref_end, query_start, query_end;
    uint64_t anchor_off;    // piecewise: chain anchors on the device
    uint32_t n_anchors;
    // RESCUE: reference window [win_start, win_end)
    uint32_t win_start, win_end;
    uint32_t cigar_cap;     // u32 slots reserved for this request's CIGAR
    uint64_t cigar_off;
};

// A gapped alignment the prep kernel hands on: SSW over a reference window, or
// piecewise extension (src/piecewisealigner.rs:396-510) over a decoded window.
struct ExtJob {
    uint32_t req;      // request index (query and result slot)
    uint32_t ref_id;
    uint64_t ws;       // window start (absolute)
    uint32_t wlen;     // window length
    uint32_t flags;    // SSW_WORD_FIRST: the byte pass is known to saturate
};
// An SSW job whose local score is at least 255 - bias (the byte pass's
// saturation point) starts with the 16-bit pass: the byte pass would only find
// that it saturates.
constexpr uint32_t SSW_WORD_FIRST = 1;

struct ExtResult {
    uint32_t found;        // 0: None
    uint32_t gapped;
    uint32_t score;
    uint32_t ref_start;    // absolute reference start (result_ref_start)
    uint32_t length;       // reference span
    uint32_t query_start, query_end;
    uint32_t edit_distance;
    uint32_t n_cigar;      // ops written (BAM encoding len<<4|op); > cigar_cap means overflow
    uint32_t status;       // 0 ok, 1 cigar overflow, 2 scratch overflow
};

}  // namespace sa
