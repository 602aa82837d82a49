// Host side of the drop-in: reference and read I/O, index construction, and
// the mapper that turns device NAMs and device extension results into SAM,
// PAF or abundance output with the reference's decision logic.
#pragma once
#include <cstdint>
#include <memory>
#include <optional>
#include <string>
#include <utility>
#include <vector>

#include "sa_engine.hpp"
#include "sa_extend.hpp"

namespace sa {

// src/packed_seq.rs:12-126
struct PackedRef {
    std::vector<std::string> names;
    std::vector<uint64_t> len;
    std::vector<uint64_t> word_off;  // n + 1
    std::vector<uint64_t> words;
    uint8_t base(uint32_t id, uint64_t i) const {
        return (uint8_t)"ACGT"[(words[word_off[id] + (i >> 5)] >> ((i & 31) * 2)) & 3];
    }
};

// src/io/fasta.rs:116-165
PackedRef read_fasta(const std::string& path);

// src/seeding/parameters.rs:181-221 (SeedingParameters::new(read_length))
SeedParams seed_params_for_read_length(uint64_t read_length);
// src/indexer.rs:294-300
int pick_bits(const PackedRef& ref, const SeedParams& p);
// Profile number of the read-length profile (src/seeding/parameters.rs:10-19,
// Noisy = 0, ReadLength50 = 1, ..., ReadLength400 = 7), as .sti files store it.
uint32_t profile_for_read_length(uint64_t read_length);

// Allocator whose resize() leaves new elements uninitialised: the index
// entries are written right after, in parallel, which also spreads the first
// touch of the pages over the threads.
template <typename T>
struct NoInitAlloc : std::allocator<T> {
    template <typename U>
    struct rebind { using other = NoInitAlloc<U>; };
    NoInitAlloc() = default;
    template <typename U>
    NoInitAlloc(const NoInitAlloc<U>&) {}
    template <typename U>
    void construct(U* p) noexcept { ::new ((void*)p) U; }
    template <typename U, typename... A>
    void construct(U* p, A&&... a) { ::new ((void*)p) U(std::forward<A>(a)...); }
};

struct HostIndex {
    SeedParams p;
    std::vector<RefEntry, NoInitAlloc<RefEntry>> entries;
    std::vector<uint64_t> bucket_starts;
};
// src/indexer.rs:18-170 (host construction with multiple threads)
void build_index(const PackedRef& ref, SeedParams p, int bits, double filter_fraction, int threads, HostIndex& out);

// src/io/record.rs + src/io/fastq.rs:22-95
struct Record {
    std::string name, comment, seq, qual;
    bool has_comment = false, has_qual = true;
    int end = 0;
};
bool read_records(const std::string& path, std::vector<Record>& out, std::string& err);
// The whole file (gunzipped when the name ends in ".gz", src/io/xopen.rs:9-20).
bool read_file_bytes(const std::string& path, std::string& out, std::string& err);
// FASTQ or FASTA read records of a file held in memory.
bool parse_records(const std::string& raw, std::vector<Record>& out, std::string& err, int threads);
struct FqLoc;
// Records from their locations in the file (as the device parser finds them);
// fails with the first record's error.
bool records_from_locs(const std::string& raw, const std::vector<FqLoc>& locs, std::vector<Record>& out,
                       std::string& err, int threads);
uint64_t estimate_read_length(const std::vector<Record>& recs);

struct Options {
    int mode = 0;  // 0 SAM, 1 PAF (-x), 2 abundances (--aemb)
    bool eqx = false, details = false, fastq_comments = false, only_mapped = false, use_ssw = false;
    bool has_rg = false;
    std::string rg_id;
    std::vector<std::string> rg_fields;
    uint64_t max_secondary = 0;
    float dropoff = 0.5f;
    uint32_t rescue_distance = 100;
    uint64_t max_tries = 20;
    uint32_t mcs = 0;
    uint64_t chunk_size = 1000000;
    int threads = 1;
    uint64_t batch_bases = 64ull << 20;  // bases per device batch (about 12 GB of device buffers per engine)
    int slots = 4;                        // engines working on batches concurrently
    uint32_t lookup_lanes = 16;           // lanes per read orientation in the lookup stage (1, 8, 16, 64)
    uint32_t seed_lanes = 64;             // lanes per read in seeding (1, 64)
    bool fused = false;                   // fused batch pass (Engine::map_fused) instead of find_nams
    uint32_t spec_per_read = 8;           // most speculative extensions per read in the fused pass
    // -A/-B/-O/-E/-L and --xdrop (src/main.rs:246-272; defaults src/aligner.rs:20-30)
    Scores scores{2, 8, 12, 1, 10, 500};
};

// src/io/sam.rs:177-201 (no @PG line)
std::string sam_header(const PackedRef& ref, const Options& o);

struct RunStats {
    double nam_ms = 0, extend_ms = 0, host_ms = 0, upload_ms = 0, total_ms = 0;
    uint64_t records = 0, requests = 0, rounds = 0, nams = 0;
    // fused pass: batches run by it, its speculative extensions, and how many of
    // them the decisions used
    uint64_t fused_batches = 0, spec_requests = 0, spec_used = 0;
    // fused passes that replayed a recorded graph; passes whose anchors did not fit
    uint64_t graph_replays = 0, fused_overflows = 0;
    // host replay: records decided and committed / decided as probes (counted
    // per pass) / probes taken over by the next pass unchanged, and thread time
    // in deciding and in writing output text
    uint64_t replay_committed = 0, replay_probed = 0, replay_accepted = 0;
    double replay_commit_ms = 0, replay_probe_ms = 0, replay_write_ms = 0;
};

// Maps all records (paired when r2 is non-empty) and appends the output
// (without header) to `out`, in input order. Abundances (mode 2) are written
// as the final table.
void map_records(Engine& eng, const PackedRef& ref, const std::vector<Record>& r1, const std::vector<Record>& r2,
                 const Options& o, std::string& out, RunStats* stats);

// A read set already in device memory: the sequences of all records
// concatenated (paired: r1 then r2 of each pair) at d_seq, and host offsets
// (records + 1) into it.
struct DeviceReads {
    const uint8_t* d_seq = nullptr;
    std::vector<uint64_t> off;
};

// Same over several engines that share one resident index: device batches of
// whole chunks are dealt to them and run concurrently. `out` may be null (the
// output is produced per chunk and dropped); `dreads` may be null (each batch
// uploads its reads).
void map_records(const std::vector<Engine*>& engines, const PackedRef& ref, const std::vector<Record>& r1,
                 const std::vector<Record>& r2, const Options& o, std::string* out, RunStats* stats,
                 const DeviceReads* dreads);

// Request loop with continuous batching over engines that share one resident
// index. Each request is a read set mapped as map_records maps it alone (its own
// chunks, RNG and insert-size state; output in record order). Requests are
// admitted in arrival order while their reads fit in a device arena (their reads
// stay there until the request completes); each engine forms its next device
// batch from the oldest admitted chunks of any requests, so requests that arrive
// while others run share batches and overlap on the device.
// Milliseconds on the steady clock the server's timings use.
double steady_ms();

class Server {
public:
    Server(std::vector<Engine*> engines, const PackedRef& ref, const Options& o, int device, uint64_t arena_bytes);
    ~Server();
    Server(const Server&) = delete;
    Server& operator=(const Server&) = delete;
    // Queues a request (paired when r2 is non-empty); returns its ticket. Its
    // latency counts from t_arrival (steady_ms() clock; < 0: now).
    uint64_t submit(std::vector<Record> r1, std::vector<Record> r2, double t_arrival = -1);
    // Times since submission, milliseconds.
    struct RequestTimes { double admit_ms = 0, done_ms = 0; };
    // Blocks until the request is done. Returns WAIT_OK, WAIT_BAD_TICKET (unknown,
    // or already held by another wait) or WAIT_FAILED (the request failed or the
    // server stopped first), with the reason in `err`.
    enum WaitStatus { WAIT_OK = 0, WAIT_BAD_TICKET = 1, WAIT_FAILED = 2 };
    WaitStatus wait(uint64_t ticket, std::string& out, RequestTimes& t, std::string& err);
    struct Stats {
        uint64_t submitted = 0, completed = 0, batches = 0, batch_requests = 0, max_requests_per_batch = 0;
        uint64_t arena_bytes = 0, arena_peak = 0;
        double busy_ms = 0;  // summed over engines
    };
    Stats stats() const;

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

}  // namespace sa
