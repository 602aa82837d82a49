/*
 * sa_hip.h -- C ABI of the MI355X (gfx950) drop-in for strobealign's read-mapping
 * path: seeding -> index lookup -> chaining/NAMs -> extension -> SAM/PAF.
 *
 * Conventions (mirroring the one native boundary of the reference, the SSW FFI
 * declared in src/ssw/raw.rs:26-51 and linked by build.rs:2):
 *   - every function returns 0 on success and a negative code on error; the text
 *     of the last error on the calling thread is available via sa_last_error();
 *   - the library allocates handles and result buffers; the caller releases them
 *     with the matching *_destroy / sa_free function;
 *   - no C++ exception crosses this boundary.
 * Threading: handles may be used from one thread at a time; different contexts
 * (one per device) may be used concurrently.
 */
#ifndef SA_HIP_H
#define SA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sa_ctx sa_ctx;         /* one device: stream, buffers */
typedef struct sa_index sa_index;     /* device-resident reference + randstrobe index */

#define SA_OK 0
#define SA_ERR_ARG -1
#define SA_ERR_IO -2
#define SA_ERR_DEVICE -3
#define SA_ERR_INTERNAL -4

/* Text of the last error on this thread (empty if none). Returns its length. */
size_t sa_last_error(char* buf, size_t cap);

/* Library version string, e.g. "0.1.0 (gfx950)". */
const char* sa_version(void);

/* Device context. Replaces nothing in the reference (which has no device). */
int sa_ctx_create(int device, sa_ctx** out);
void sa_ctx_destroy(sa_ctx* ctx);

/* Seeding parameters (src/seeding/parameters.rs:163-176; RandstrobeParameters,
 * src/seeding/strobes.rs:8-23), plus the bucket bits and filter cutoff of the
 * index (src/index.rs:68-87). */
typedef struct sa_seeding_params {
    int32_t k, s, t;
    int32_t w_min, w_max, max_dist;
    uint64_t q;
    uint64_t main_hash_mask;
    uint64_t forward_main_hash_mask;
    int32_t partial_orientation_pos;
    int32_t bits;
    uint64_t filter_cutoff;
} sa_seeding_params;

/* SeedingParameters::new(read_length) (src/seeding/parameters.rs:181-221). */
int sa_seeding_params_for_read_length(uint64_t read_length, sa_seeding_params* out);

/* Loads a FASTA reference (src/io/fasta.rs:116-165) and builds the randstrobe
 * index for the profile of `read_length` (src/indexer.rs:18-170, make_index;
 * bits from pick_bits, src/indexer.rs:294-300, when bits <= 0), then uploads
 * reference and index to the context's device. */
int sa_index_build(sa_ctx* ctx, const char* fasta_path, uint64_t read_length, int bits, int threads,
                   sa_index** out);

/* Uploads an existing index: `entries` are n 16-byte RefRandstrobe records
 * (src/index.rs:15-30, sorted as make_index sorts them), `bucket_starts` has
 * 2^bits + 1 entries (src/index.rs:76-83), and the packed reference uses the
 * PackedSeq layout (src/packed_seq.rs:1-15) with each contig starting on a
 * 64-bit word boundary. These are the in-memory and .sti layouts of the
 * reference (src/index.rs:374-409), so a .sti file can be streamed in. */
int sa_index_upload(sa_ctx* ctx, const sa_seeding_params* params, const void* entries, uint64_t n,
                    const uint64_t* bucket_starts, const uint64_t* packed_ref_words, uint64_t n_words,
                    const uint64_t* contig_word_offsets, const uint64_t* contig_lengths,
                    const char* const* contig_names, uint32_t n_contigs, sa_index** out);
void sa_index_destroy(sa_index* index);
/* Writes the index as a .sti file (StrobemerIndex::write, src/index.rs:374-409):
 * "STI\1", format version 7, an 8-byte reserved chunk (length, contents), then
 * filter cutoff, bits, profile number (of read_length's profile), k, s, w_min,
 * w_max, q and max_dist as u32, main_hash_mask as u64, the 16-byte entries and
 * the bucket starts each after its u64 length; native byte order. */
int sa_index_save_sti(const sa_index* index, uint64_t read_length, const char* path);
/* Loads a .sti file (read_index, src/index.rs:412-495) with its FASTA reference:
 * the header is checked against the parameters for read_length and `bits`
 * (<= 0: pick_bits of the reference) as the reference checks it, and the entries
 * and bucket starts stream from the file through pinned staging memory straight
 * into device memory (no host copy of the index). */
int sa_index_load_sti(sa_ctx* ctx, const char* fasta_path, const char* sti_path, uint64_t read_length, int bits,
                      sa_index** out);
/* sa_index_build with make_index (src/indexer.rs:18-170) run on the device:
   syncmers per 4096-position chunk, randstrobes per syncmer, two stable radix
   sorts, bucket directory and filter cutoff from device passes. Same entries,
   bucket starts and parameters as the host build; up to 2^31 - 1 entries. */
int sa_index_build_device(sa_ctx* ctx, const char* fasta_path, uint64_t read_length, int bits, sa_index** out);
/* Copies an index back to the host: n_entries 16-byte entries (RefRandstrobe
   layout, src/index.rs:24-30) and 2^bits + 1 bucket starts (src/index.rs:93-97). */
int sa_index_download(const sa_index* index, void* entries, uint64_t* bucket_starts);
int sa_index_info(const sa_index* index, sa_seeding_params* params, uint64_t* n_entries, uint32_t* n_contigs);

/* Mapping options: the flags of src/main.rs:61-283 that the mapping path uses. */
typedef struct sa_map_options {
    int32_t mode;            /* 0 SAM, 1 PAF (-x), 2 abundances (--aemb) */
    int32_t eqx;             /* --eqx */
    int32_t details;         /* --details */
    int32_t fastq_comments;  /* -C */
    int32_t only_mapped;     /* -U */
    int32_t use_ssw;         /* --ssw */
    const char* rg_id;       /* --rg-id (NULL: none) */
    const char* const* rg;   /* --rg fields */
    int32_t n_rg;
    uint32_t max_secondary;  /* -N */
    float dropoff;           /* -S (0.5) */
    uint32_t rescue_distance;/* -R (100) */
    uint32_t max_tries;      /* -M (20) */
    uint32_t mcs;            /* --mcs: 0 always, 1 rescue, 2 off, 3 first-strobe */
    uint64_t chunk_size;     /* --chunk-size (1000000) */
    int32_t threads;         /* host threads for the decision replay and output */
    uint64_t batch_bases;    /* bases per device batch (0: default) */
    int32_t slots;           /* batches in flight on the device at once (0: default 4) */
    /* Alignment scores (src/aligner.rs:11-30; -A/-B/-O/-E/-L and --xdrop of
     * src/main.rs:246-272): match score, mismatch / gap-open / gap-extend
     * penalties (0..127 each), soft-clipping end bonus, piecewise x-drop. */
    int32_t match_score, mismatch_penalty, gap_open_penalty, gap_extend_penalty;
    int32_t end_bonus, xdrop;
    /* Lanes per read orientation in the lookup stage (find_hits): 1 (one thread
     * makes the orientation's lookups in turn) or 8, 16, 64 (a lane per
     * randstrobe); 0: default 16. The hits never depend on it, only the speed. */
    int32_t lookup_lanes;
    /* Batch pass: 0 the plain per-batch stages (seeding to NAMs, then the
     * decisions request each extension they need); 1 the fused pass, which runs
     * seeding, lookups, chaining, NAMs and the extensions the decisions are
     * expected to need (each single-end read's leading NAMs, each mate's top
     * NAMs) as one recorded graph replayed per batch, with sizes kept on the
     * device; the decisions then request only what it did not foresee. The
     * output is identical either way. */
    int32_t fused;
    /* Lanes per read in seeding (randstrobes of the query): 1 (a thread scans
     * the read) or 64 (a wave: a lane per position for the syncmers, then a
     * lane per syncmer for the strobes; reads above 512 bp take the thread
     * form); 0: default 64. The randstrobes never depend on it. */
    int32_t seed_lanes;
} sa_map_options;

/* Fills `opt` with the reference defaults. */
void sa_map_options_default(sa_map_options* opt);

/* Maps FASTQ/FASTA files (reads2 may be NULL for single-end) and returns the
 * complete output (SAM with header but without @PG, PAF, or the abundance
 * table) in *out, which the caller releases with sa_free. Replaces the per-chunk
 * Mapper::map_chunk loop of src/main.rs:873-988 with its calls into
 * align_single_end_read / align_paired_end_read (src/mapper.rs:349, 624) and
 * map_*_end_read / abundances_* (src/maponly.rs:18, 57, 114, 198). Reads of
 * at most 512 bp: a batch with a longer read fails (non-zero return, the
 * reason in sa_last_error); so do sa_map_records, sa_map_reads and the
 * serving loop's requests. */
int sa_map_files(sa_index* index, const sa_map_options* opt, const char* reads1, const char* reads2, char** out,
                 size_t* out_len);

/* Same for records held in memory (uppercase or lowercase ASCII sequences,
 * qualities may be NULL). r2 arrays are NULL for single-end input. */
int sa_map_records(sa_index* index, const sa_map_options* opt, uint64_t n, const char* const* names1,
                   const char* const* seqs1, const char* const* quals1, const char* const* names2,
                   const char* const* seqs2, const char* const* quals2, char** out, size_t* out_len);

/* A read set parsed once (FASTQ or FASTA; reads2 may be NULL) and mapped by
 * sa_map_reads; it is the input of one request of the serving loop. */
typedef struct sa_reads sa_reads;
int sa_reads_load(const char* reads1, const char* reads2, sa_reads** out);
uint64_t sa_reads_count(const sa_reads* reads);
/* Fields of record i of a read set (mate 0: reads1, 1: reads2), valid while the
 * set lives: name (without "/1" or "/2", whose number goes to *end, else 0),
 * comment (NULL when the header has none), uppercase sequence, qualities (NULL
 * for FASTA input). The SequenceRecord of src/io/record.rs. */
int sa_reads_record(const sa_reads* reads, uint64_t i, int mate, const char** name, const char** comment,
                    const char** seq, const char** qual, int* end);
/* Same read set parsed on the index's device (FASTQ; FASTA read files are
 * parsed on the host): the files' bytes are copied in once, newlines found,
 * records located and checked, and the uppercase sequences staged in HBM as
 * sa_reads_upload leaves them; the host only gets back each record's
 * location (name, comment, sequence and quality offsets) to keep the fields
 * for output. Files ending in ".gz" are decompressed on the host first
 * (src/io/xopen.rs:9-20). Errors are those of sa_reads_load. *ingest_ms (may be
 * NULL) receives the device time from the copy of the files in to the last
 * kernel. Replaces src/io/fastq.rs:22-95 for the device path. */
int sa_reads_load_device(sa_index* index, const char* reads1, const char* reads2, sa_reads** out,
                         double* ingest_ms);
/* Stages the read set in the index's device memory, so that sa_map_reads starts
 * from reads already resident in HBM (no host-to-device copy per call). */
int sa_reads_upload(sa_index* index, sa_reads* reads);
void sa_reads_destroy(sa_reads* reads);

/* Wall-clock breakdown of one sa_map_reads call (milliseconds). */
typedef struct sa_run_stats {
    double total_ms, upload_ms, nam_ms, extend_ms, host_ms;
    uint64_t records, requests, rounds, nams;
    /* fused pass (sa_map_options.fused): batches it ran, its speculative
     * extensions, how many of them the decisions used, the passes that replayed
     * a recorded graph, and the batches whose anchors did not fit its buffers
     * (run by the plain stages instead) */
    uint64_t fused_batches, spec_requests, spec_used, graph_replays, fused_overflows;
    /* host replay of the read-level decisions: records decided and committed,
     * records decided as probes (each pass counts), and the thread time spent
     * on each and on writing the committed records' output text */
    uint64_t replay_committed, replay_probed;
    double replay_commit_ms, replay_probe_ms, replay_write_ms;
    /* probes the next pass took over unchanged (same generator and insert-size
     * state at their start, so the same decisions) */
    uint64_t replay_accepted;
} sa_run_stats;

/* Maps a read set. With out == NULL the output is produced and discarded
 * (measurement); otherwise as sa_map_files. stats may be NULL. */
int sa_map_reads(sa_index* index, const sa_map_options* opt, const sa_reads* reads, char** out, size_t* out_len,
                 sa_run_stats* stats);

/* Schedule autotuner: maps the first sample_pairs records of `reads` (0: all;
 * resident if uploaded, output dropped). Unless opt->seed_lanes is set, it
 * first times seeding on one engine with 1 and 64 lanes per read (a thread or
 * a wave per read) and keeps the faster; unless opt->lookup_lanes is set, it
 * times the lookup stage with 1, 8, 16 and 64 lanes per read orientation (a
 * thread, lane groups, or a wave per orientation) and keeps the fastest; then
 * it maps the sample with each batch pass (plain and fused;
 * the fused pass only in alignment mode), engine count in {1, 2, 3, 4, 6, 8}
 * and batch size in {16, 32, 64} Mbases, and writes the fastest choices into
 * opt->seed_lanes, opt->lookup_lanes, opt->fused, opt->slots and
 * opt->batch_bases. The outcome never depends on the schedule (the output is
 * identical for every choice); only the speed does. report (may be NULL)
 * receives the trials as JSON [{"seed_lanes", "seed_ms"}, ...,
 * {"lookup_lanes", "lookup_ms"}, ..., {"fused", "slots", "batch_bases", "ms"},
 * ...]. No counterpart in the reference, which has no device schedule. */
int sa_autotune(sa_index* index, const sa_reads* reads, uint64_t sample_pairs, sa_map_options* opt, char* report,
                size_t report_cap);

/* Serving loop: requests (read sets) are submitted as they arrive and mapped
 * with continuous batching. A request is admitted, in arrival order, once its
 * reads fit in a device arena of arena_bytes (0: min(4 GiB, a quarter of the
 * free device memory)); its reads stay in HBM until it completes. Each of the
 * opt->slots engines (sharing the index, which must outlive the server) forms its
 * next device batch from the oldest admitted chunks of any requests, up to
 * opt->batch_bases bases (0: 8M), so requests that arrive while others run share
 * batches and overlap on the device. Each request's output (SAM records without
 * header, PAF, or the abundance table) equals sa_map_records on that request
 * alone. Submit and wait may be called from any threads. This is the request loop
 * the reference lacks: its unit of work is Mapper::map_chunk (src/main.rs:873-988),
 * which each request runs as its own read set. */
typedef struct sa_server sa_server;
typedef struct sa_server_stats {
    uint64_t submitted, completed, batches, batch_requests, max_requests_per_batch;
    uint64_t arena_bytes, arena_peak;
    double busy_ms; /* engine time in batches, summed over engines */
} sa_server_stats;
int sa_server_create(sa_index* index, const sa_map_options* opt, uint64_t arena_bytes, sa_server** out);
/* Queues a request (arrays as for sa_map_records) and returns its ticket. */
int sa_server_submit(sa_server* server, uint64_t n, const char* const* names1, const char* const* seqs1,
                     const char* const* quals1, const char* const* names2, const char* const* seqs2,
                     const char* const* quals2, uint64_t* ticket);
/* Queues records [first, first + count) of a loaded read set as one request. */
int sa_server_submit_reads(sa_server* server, const sa_reads* reads, uint64_t first, uint64_t count,
                           uint64_t* ticket);
/* Waits for a request; *out (sa_free) receives its output unless out is NULL,
 * *latency_ms the time from the start of its submit call to its completion (all
 * of its output formatted). A ticket is waited once: a second wait on it, from
 * any thread, returns SA_ERR_ARG. */
int sa_server_wait(sa_server* server, uint64_t ticket, char** out, size_t* out_len, double* latency_ms);
int sa_server_get_stats(const sa_server* server, sa_server_stats* st);
/* As sa_index_kernel_times, over the server's engines. */
size_t sa_server_kernel_times(const sa_server* server, char* buf, size_t cap);
/* Stops the workers and frees the arena. Requests not yet completed fail: a
 * thread blocked in sa_server_wait returns an error before this returns, and
 * requests nobody waits for are dropped. */
void sa_server_destroy(sa_server* server);

/* Per-stage probes (parity tests, and a finer-grained drop-in for a caller that
 * keeps its own decision logic). Reads are concatenated ASCII with offsets[n+1]. */

/* QueryRandstrobe (src/seeding/mod.rs:10-16). */
typedef struct sa_qrand {
    uint64_t hash, hash_revcomp;
    uint32_t start, end;
} sa_qrand;

/* randstrobes_query (src/seeding/mod.rs:19-75) for n reads. out_fwd/out_rc are
 * sized offsets[n] each; read i's entries start at offsets[i]; counts[2i],
 * counts[2i+1] receive the forward / reverse-complement counts. */
int sa_randstrobes_query_batch(sa_index* index, const uint8_t* seqs, const uint64_t* offsets, uint32_t n,
                               sa_qrand* out_fwd, sa_qrand* out_rc, uint32_t* counts);

/* NAM (src/nam.rs:18-31) as produced by get_nams_by_chaining (src/nam.rs:138-166)
 * and the stable score sort of sort_nams (src/nam.rs:170), before the per-chunk
 * shuffle of the tied best (src/shuffle.rs:7-25, which needs the caller's RNG). */
typedef struct sa_nam {
    uint32_t nam_id, ref_id;
    uint32_t ref_start, ref_end;
    uint32_t query_start, query_end;
    uint32_t matching_bases, n_anchors;
    float score;
    uint32_t flags; /* bit0 is_revcomp, bit1 is_consistent, bits2-3 reverse_nam_if_needed outcome */
    uint64_t anchor_offset;
} sa_nam;

/* NAMs of n reads; *nams (released with sa_free) holds nam_offsets[n] records,
 * read i's at [nam_offsets[i], nam_offsets[i+1]). mcs as in sa_map_options. */
int sa_find_nams_batch(sa_index* index, const uint8_t* seqs, const uint64_t* offsets, uint32_t n, uint32_t mcs,
                       uint32_t rescue_distance, sa_nam** nams, uint64_t* nam_offsets);

/* Size of the per-device scratch array behind sa_find_nams_batch_noalloc. */
#define SA_NOALLOC_SCRATCH_BYTES (128ull << 20)

/* sa_find_nams_batch for callers that cannot allocate: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc. It has no
 * counterpart in the reference. Inputs and outputs are device memory the caller
 * owns: d_seqs (ASCII bases) with n + 1 offsets d_offsets; NAMs go to d_nams
 * (room for nam_cap records), read i's at [d_nam_offsets[i], d_nam_offsets[i+1])
 * (n + 1 entries); *n_nams receives the total. Work runs on `stream` (a
 * hipStream_t; NULL: the null stream), and the call returns once the outputs are
 * written (it waits twice on the stream to size the stages). Every intermediate
 * buffer comes from one static device array of SA_NOALLOC_SCRATCH_BYTES per
 * device; no hipMalloc, hipFree, pinned-memory or other allocation is made. A
 * batch that needs more scratch or more NAM room than given fails with SA_ERR_ARG
 * (split it). Calls on one device share that array: calls that overlap in time on
 * one device race and must be serialised by the caller. */
int sa_find_nams_batch_noalloc(sa_index* index, const uint8_t* d_seqs, const uint64_t* d_offsets, uint32_t n,
                               uint32_t mcs, uint32_t rescue_distance, sa_nam* d_nams, uint64_t nam_cap,
                               uint64_t* d_nam_offsets, uint64_t* n_nams, void* stream);

/* One NAM's extension in sa_map_batch: what extend_seed (src/mapper.rs:532-620)
 * returns for it (Alignment, src/mapper.rs:58-77: reference start and span,
 * query_start / query_end give the soft clips). The CIGAR (BAM-encoded ops with
 * =/X) is at [cigar_offset, cigar_offset + n_cigar) of the batch's CIGAR array. */
typedef struct sa_nam_extension {
    uint32_t found, gapped, score, ref_start, length, query_start, query_end, edit_distance;
    uint64_t cigar_offset;
    uint32_t n_cigar, extended; /* extended: 0 when this NAM was not extended (found then 0) */
} sa_nam_extension;

/* Batch drop-in for a caller that keeps its own decision loop (the reference's
 * align_single_end_read, src/mapper.rs:349-528, with its RNG and dropoff rules):
 * for n single-end reads, the NAMs of sa_find_nams_batch (stable score order,
 * before the per-chunk shuffle) and, for the first `speculative` NAMs of each
 * read that are consistent (Nam::is_consistent, src/nam.rs:49-66), the result of
 * extend_seed with opt's scores and --ssw choice; ext[x] belongs to nams[x].
 * A NAM outside that set is not extended (its entry has extended = 0) and the
 * caller aligns it itself. Work runs on `stream` (a hipStream_t; NULL: the
 * library's own): it is ordered after the work already queued there, and the
 * call returns once its results are copied out. *nams, *ext and *cigars are
 * released with sa_free. */
int sa_map_batch(sa_index* index, const sa_map_options* opt, const uint8_t* seqs, const uint64_t* offsets, uint32_t n,
                 uint32_t speculative, void* stream, sa_nam** nams, uint64_t* nam_offsets, sa_nam_extension** ext,
                 uint32_t** cigars);

/* Hit (src/hit.rs:19-28) and HitsDetails (src/hit.rs:32-52). */
typedef struct sa_hit {
    uint64_t query_start, query_end;
    uint64_t position; /* index entry: forward position for full hits, undirected for partial */
    uint64_t hash, hash_revcomp;
    uint32_t is_partial, is_filtered;
} sa_hit;
typedef struct sa_hits_details {
    uint64_t full_not_found, full_filtered, full_found;
    uint64_t partial_not_found, partial_filtered, partial_found;
    uint64_t rescued;
} sa_hits_details;

/* find_hits (src/hit.rs:270-333, with global and local rescue) for both
 * orientations of n reads (randstrobes_query first, src/seeding/mod.rs:19-75).
 * Orientation o (0 forward, 1 reverse complement) of read i has its hits at
 * [hit_offsets[2i+o], hit_offsets[2i+o+1]) of *hits (released with sa_free;
 * hit_offsets has 2n + 1 entries) and its details in details[2i+o]. */
int sa_find_hits_batch(sa_index* index, const uint8_t* seqs, const uint64_t* offsets, uint32_t n, uint32_t mcs,
                       uint32_t rescue_distance, sa_hit** hits, uint64_t* hit_offsets, sa_hits_details* details);

/* The chaining half of Chainer::get_chains (src/chainer.rs:182-267): orientation
 * choice, hits_to_anchors, sort + dedup, collinear_chaining, extract_chains, then
 * the stable score sort and consistency flags as sa_find_nams_batch returns them,
 * from hits and details laid out as sa_find_hits_batch returns them (an
 * orientation has at most as many hits as its read has bases). */
int sa_chain_batch(sa_index* index, const uint8_t* seqs, const uint64_t* offsets, uint32_t n, const sa_hit* hits,
                   const uint64_t* hit_offsets, const sa_hits_details* details, sa_nam** nams, uint64_t* nam_offsets);

/* Aligner::align (src/aligner.rs:80-151: SSW of ext/ssw/ssw.c plus the end-bonus
 * extension) of n (query, reference) pairs. Queries/references are ASCII,
 * concatenated with offsets. Each result: found, score, query_start, query_end,
 * ref_start, ref_end, edit_distance; CIGARs (BAM-encoded len<<4|op with =/X)
 * are concatenated in *cigars with cigar_offsets[n+1]. */
typedef struct sa_alignment {
    uint32_t found, score, query_start, query_end, ref_start, ref_end, edit_distance, n_cigar;
} sa_alignment;
int sa_align_ssw_batch(sa_ctx* ctx, const uint8_t* queries, const uint64_t* query_offsets, const uint8_t* refs,
                       const uint64_t* ref_offsets, uint32_t n, sa_alignment* out, uint32_t** cigars,
                       uint64_t* cigar_offsets);

/* Alignment scores (src/aligner.rs:11-30) plus the piecewise x-drop (src/main.rs:271). */
typedef struct sa_scores {
    int32_t match_score, mismatch_penalty, gap_open_penalty, gap_extend_penalty, end_bonus, xdrop;
} sa_scores;

/* Kinds of sa_extend_batch jobs: the aligner's entry points. */
#define SA_EXT_SSW 0       /* Aligner::align (src/aligner.rs:80-151): SSW + end-bonus extension */
#define SA_EXT_HAMMING 1   /* hamming_align (src/aligner.rs:223-259); not found when lengths differ */
#define SA_EXT_PIECEWISE 2 /* Aligner::align_piecewise (src/aligner.rs:153-163) = extend_piecewise
                              (src/piecewisealigner.rs:396-510) with the job's anchors, no pruning */
#define SA_EXT_GLOBAL 3    /* PiecewiseAligner::global_alignment (src/piecewisealigner.rs:84-120) */
#define SA_EXT_XDROP 4     /* xdrop_alignment, GlobalStartLocalEnd (src/piecewisealigner.rs:161-229) */
#define SA_EXT_XDROP_REV 5 /* xdrop_alignment, LocalStartGlobalEnd */

typedef struct sa_extend_job {
    uint32_t kind;          /* SA_EXT_* */
    uint32_t padding;       /* SA_EXT_PIECEWISE: end padding (src/mapper.rs:581 passes read length / 10) */
    uint64_t anchor_offset; /* SA_EXT_PIECEWISE: first of its n_anchors pairs in `anchors` */
    uint32_t n_anchors;
    uint32_t reserved;
} sa_extend_job;

/* Job i aligns query i (queries/query_offsets, ASCII) against reference window i
 * (refs/ref_offsets, ACGT, as refseq.decode returns it) with its kind's entry
 * point, on the device (the kernels of the mapping path). Piecewise anchors are
 * (ref_start, query_start) pairs in window coordinates, last anchor first, each
 * spanning k bases (the index's k). Results as sa_align_ssw_batch; coordinates
 * are window-relative; for SA_EXT_GLOBAL / SA_EXT_XDROP* `found` is set and
 * `score` holds the block aligner's i32 score (two's complement). */
int sa_extend_batch(sa_ctx* ctx, const sa_scores* scores, uint32_t k, const sa_extend_job* jobs, uint32_t n,
                    const uint8_t* queries, const uint64_t* query_offsets, const uint8_t* refs,
                    const uint64_t* ref_offsets, const uint32_t* anchors, uint64_t n_anchor_pairs, sa_alignment* out,
                    uint32_t** cigars, uint64_t* cigar_offsets);

/* Kernel time (ms) spent by the index's engine since creation, and the number
 * of device launches; for measurement. */
int sa_index_stats(const sa_index* index, double* kernel_ms);

/* Per-kernel totals since the index was created, as JSON text
 * {"k_seed": [milliseconds, launches, work_items, algorithmic_bytes, dp_cells], ...},
 * timed with HIP events on the engines' streams (algorithmic_bytes is 0 where the
 * kernel does not count them; dp_cells, query x reference cells of the forward SSW
 * pass, is counted for k_ssw_byte_fwd only). Returns the JSON length (truncated to
 * cap-1 in buf). */
size_t sa_index_kernel_times(const sa_index* index, char* buf, size_t cap);

/* Device copy bandwidth: copies `bytes` between two device buffers `iters` times
 * with a plain vector-copy kernel (k_hbm_copy) and reports read+write GB/s. The
 * HBM reference point of the roofline figures; no counterpart in the reference. */
int sa_hbm_copy_probe(sa_ctx* ctx, uint64_t bytes, int iters, double* gbps);

void sa_free(void* p);

#ifdef __cplusplus
}
#endif

#endif /* SA_HIP_H */
