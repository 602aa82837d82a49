// Device engine: owns the device-resident reference and index, the per-batch
// buffers, and launches the kernels of sa_kernels.hip. Host code above this
// layer (sa_mapper.cpp, the C ABI) never touches device pointers directly.
#pragma once
#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "sa_types.hpp"

struct ihipStream_t;
struct ihipEvent_t;

namespace sa {

struct Scores;
struct NamPass;

// Host copy of what the chainer/NAM stage returns for one batch of reads.
// A host array owned elsewhere (the engine's pinned staging memory).
template <typename T>
struct HostView {
    T* p = nullptr;
    size_t n = 0;
    T& operator[](size_t i) const { return p[i]; }
    size_t size() const { return n; }
    bool empty() const { return n == 0; }
    T* data() const { return p; }
    T* begin() const { return p; }
    T* end() const { return p + n; }
};

// NAMs of a batch. nams and nam_off point into the engine's pinned staging
// memory and stay valid until that engine's next find_nams.
struct NamBatch {
    HostView<NamRec> nams;          // per read, stable-sorted by score (pre-shuffle)
    HostView<uint64_t> nam_off;     // n_reads + 1
    std::vector<ReadNamInfo> info;  // n_reads
};

// Extensions computed by the fused batch pass ahead of the read-level
// decisions (SpecParams): views into the engine's pinned staging memory, valid
// until its next fused pass. Requests are in read order; result y's CIGAR is
// cigars[cigar_off[y], cigar_off[y + 1]). A result with a nonzero status has
// to be computed again by extend().
struct SpecBatch {
    HostView<ExtRequest> req;
    HostView<ExtResult> res;
    HostView<uint64_t> cigar_off;  // req.size() + 1
    HostView<uint32_t> cigars;
    uint64_t wanted = 0;  // requests the selection made (more than req.size() when the buffers were full)
    bool replayed = false;  // the pass ran as a replay of its recorded graph
    bool overflow = false;  // map_fused returned false: the anchors did not fit
};

// Per-stage timings of the last call, in milliseconds (HIP events on the engine stream).
struct StageTimes {
    float seed = 0, hits = 0, chain = 0, finalize = 0, extend = 0, scans = 0, fused = 0;
};

// Plain device buffers outside an engine (a read set staged once for many calls).
void* device_upload(int device, const void* src, uint64_t bytes);
void* device_alloc(int device, uint64_t bytes);
// Host -> device copy on a stream of the calling thread; returns when it is done.
void device_copy_in(int device, void* dst, const void* src, uint64_t bytes);
uint64_t device_free_bytes(int device);

struct SswRaw;
// ssw_align of one pair on the device (the SSW link shim): codes, an n x n
// matrix and ssw_align's arguments in; its result and M/I/D CIGAR out.
void ssw_raw_device(int device, const int8_t* read, int32_t readLen, const int8_t* ref, int32_t refLen,
                    const int8_t* mat, int32_t n, int32_t score_size, int32_t gapO, int32_t gapE, uint32_t flag,
                    uint32_t filters, int32_t filterd, int32_t maskLen, SswRaw& out, std::vector<uint32_t>& cigar);
// Copies `bytes` between two device buffers `iters` times with a plain vector-copy
// kernel; returns the achieved read+write GB/s (the HBM reference of the roofline).
double hbm_copy_probe(int device, uint64_t bytes, int iters);
void device_free(int device, void* p);

// Size of the static device array that backs the entry point that allocates
// nothing (Engine::find_nams_noalloc), and that array's address on `device`
// (looked up with hipGetSymbolAddress on the first call per device).
constexpr uint64_t NOALLOC_SCRATCH_BYTES = 128ull << 20;
uint8_t* noalloc_scratch(int device);

class Engine {
public:
    explicit Engine(int device);
    // A second engine (own stream and batch buffers) on the owner's device that
    // shares its resident reference and index; the owner must outlive it.
    struct ShareIndex {};
    Engine(ShareIndex, const Engine& owner);
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    // Packed reference: words of each contig start on a word boundary.
    void upload_reference(const uint64_t* words, uint64_t n_words, const uint64_t* word_off, const uint64_t* len,
                          uint32_t n_contigs);
    void upload_index(const RefEntry* entries, uint64_t n, const uint64_t* bucket_starts, const SeedParams& p);
    // The same from a byte stream in the .sti layout after the parameters
    // (src/index.rs:401-402): n 16-byte entries, then the bucket-starts length
    // (checked against p.bits) and the bucket starts. read(dst, bytes) fills host
    // memory; the bytes go through pinned staging memory straight to the device.
    void upload_index_streamed(uint64_t n, const SeedParams& p, const std::function<void(void*, uint64_t)>& read);
    // make_index (src/indexer.rs:18-170) on the device from the uploaded reference:
    // the same entries, bucket directory and filter cutoff as the host build
    // (contig lengths on the host; p.bits set to `bits`). Returns the parameters.
    SeedParams build_index(const uint64_t* len, uint32_t n_contigs, SeedParams p, int bits, double filter_fraction);
    // copies the device index back (n_entries entries, 2^bits + 1 bucket starts)
    void download_index(RefEntry* entries, uint64_t* bucket_starts) const;
    void setup_chain_tables();
    // pinned host memory of staging slot `slot` (kept, grown on demand)
    void* host_stage(int slot, uint64_t bytes);

    // Seeding -> hits -> chaining -> NAMs for reads held in host memory (or already
    // resident when `seq == nullptr`, see upload_reads). Keeps the reads and the chain
    // anchors on the device for the extension calls that follow.
    void upload_reads(const uint8_t* seq, const uint64_t* off, uint32_t n_reads);
    // Same for reads already in device memory: `d_seq` points at the first base of
    // the batch, `off` (host) holds n_reads + 1 offsets relative to it.
    void use_reads(const uint8_t* d_seq, const uint64_t* off, uint32_t n_reads);
    // Same for reads in device memory as segments: segment i holds the reads of
    // the batch's i-th part at d_base + seg[i] (seg_len[i] bytes); the parts are
    // copied together (a device-to-device copy) unless they are already adjacent.
    void gather_reads(const uint8_t* d_base, const uint64_t* seg, const uint64_t* seg_len, uint32_t n_segs,
                      const uint64_t* off, uint32_t n_reads);
    void find_nams(const MapParams& mp, const ChainParams& cp, NamBatch& out);
    // The two halves of find_nams on their own (per-stage drop-ins): the hits of
    // every read orientation (orientation t of read i at 2 * off[i] + t % 2 * L,
    // hc[2i + t % 2].n_hits of them; find_hits, src/hit.rs:270-333), and NAMs from
    // hits in that layout (src/chainer.rs:182-267 after find_hits).
    void find_hits(const MapParams& mp, std::vector<Hit>& hits, std::vector<HitCounts>& hc);

    // The fused batch pass: seeding -> lookups -> chaining -> NAMs -> the
    // speculative extensions of `sp` -> their traceback and CIGARs, with every
    // size that depends on the data kept on the device, so the whole pass is one
    // recorded graph: run directly the first time for a batch shape (reads,
    // bases, longest read, parameters), recorded as a hipGraph the second time
    // and replayed from then on, with one wait at its end. Its buffers have fixed
    // capacities sized from what this engine has seen; a batch whose anchors do
    // not fit returns false (nothing written; the next pass is sized larger) and
    // is run by find_nams instead, as is an engine's first batch. Speculative
    // requests beyond the buffers are left out (spec.wanted counts them). Reads
    // as for find_nams; the chain anchors stay for extend().
    bool map_fused(const MapParams& mp, const ChainParams& cp, const SpecParams& sp, const Scores& sc, NamBatch& out,
                   SpecBatch& spec);
    // Passes run by map_fused as a graph replay (not counting recordings).
    uint64_t fused_replays() const { return fused_replays_; }
    void chain_hits(const Hit* hits, const HitCounts* hc, const ChainParams& cp, NamBatch& out);

    // find_nams for reads and outputs in device memory, on the caller's stream,
    // with every intermediate buffer carved from `scratch` (scratch_bytes): no
    // device or pinned allocation, no timing events. NAMs go to d_nams (capacity
    // nam_cap records), read i's at [d_nam_off[i], d_nam_off[i + 1]). Throws
    // std::invalid_argument when the batch needs more scratch or NAM capacity.
    // Returns once the outputs are written.
    void find_nams_noalloc(const uint8_t* d_seq, const uint64_t* d_off, uint32_t n, const MapParams& mp,
                           const ChainParams& cp, NamRec* d_nams, uint64_t nam_cap, uint64_t* d_nam_off,
                           uint64_t* n_nams, ihipStream_t* st, uint8_t* scratch, uint64_t scratch_bytes);

    // Replaces the chain-anchor buffer (what find_nams leaves for piecewise
    // extension) with `n_pairs` (ref_start, query_start) pairs, for EXT_ALIGN_PW
    // requests that bring their own anchors.
    void set_chain_anchors(const uint32_t* pairs, uint64_t n_pairs);

    // Fine-grained stages for parity probes (valid after upload_reads).
    void randstrobes(std::vector<QRand>& qr, std::vector<uint32_t>& counts);

    // Extension requests over the resident reads. `cigars` receives each result's ops
    // at `cigar_off` positions computed by this call (compact).
    void extend(const ExtRequest* req, uint32_t n, const Scores& sc, std::vector<ExtResult>& res,
                std::vector<uint32_t>& cigars, std::vector<uint64_t>& cigar_off);

    const StageTimes& times() const { return times_; }
    // Per-kernel totals since creation: name -> (milliseconds, launches, work items).
    // bytes: algorithmic bytes (the compulsory reads and writes of the work items), 0 if not counted;
    // cells: dynamic-programming cells (query x reference) of the forward SSW pass, 0 elsewhere
    struct KernelStat { double ms = 0; uint64_t launches = 0, items = 0, bytes = 0, cells = 0; };
    const std::vector<std::pair<std::string, KernelStat>>& kernel_stats() const { return kstats_; }
    void add_kernel(const char* name, double ms, uint64_t items, uint64_t bytes = 0, uint64_t cells = 0);
    const SeedParams& params() const { return p_; }
    uint64_t n_entries() const { return n_entries_; }
    ihipStream_t* stream() const { return stream_; }
    // Waits until nothing of this engine runs on the device (errors ignored);
    // after a call failed part-way, before the engine's buffers are reused or freed.
    void drain() noexcept;
    // Runs the engine's work on `st` (null: its own stream) from now on; returns
    // the stream it used before. The caller keeps `st` alive while it is set.
    ihipStream_t* set_stream(ihipStream_t* st);
    // Kernel time accumulated over all find_nams/extend calls, milliseconds.
    double total_kernel_ms() const { return total_ms_; }
    uint64_t total_lookup_bytes() const { return lookup_bytes_; }

private:
    struct Fused;
    void fused_sequence(Fused& f);
    void chain_tables(const ChainParams& cp);
    void finish_nams(uint32_t n, const NamPass& o, ihipEvent_t** ev, NamBatch& out);
    void* grow(int slot, uint64_t bytes);
    void sync();  // waits for the engine's stream (blocking wait)
    ihipEvent_t* event();  // a timing event from the pool
    void release_event(ihipEvent_t* e);
    void scan_u64(const uint64_t* in, uint64_t* out, uint64_t n, uint64_t* total);

    int device_;
    ihipStream_t* stream_ = nullptr;
    ihipStream_t* own_stream_ = nullptr;
    ihipEvent_t* sync_ev_ = nullptr;
    std::vector<ihipEvent_t*> event_pool_;
    SeedParams p_{};
    uint64_t n_entries_ = 0;
    RefEntry* d_entries_ = nullptr;
    uint64_t* d_buckets_ = nullptr;
    uint64_t* d_words_ = nullptr;
    uint64_t* d_word_off_ = nullptr;
    uint64_t* d_len_ = nullptr;
    uint32_t n_contigs_ = 0;
    float* d_log2_ = nullptr;
    float* d_same_diag_ = nullptr;
    ChainParams tab_cp_{};
    uint32_t tab_k_ = 0;
    uint32_t n_reads_ = 0;
    uint64_t n_bases_ = 0;
    uint32_t max_len_ = 0;
    uint32_t min_len_ = 0;  // shortest non-empty read of the batch (0: none)
    std::vector<uint64_t> h_off_;
    const uint8_t* d_seq_ = nullptr;
    // offsets and chain anchors of the reads the extension calls work on (the
    // plain buffers, or the fused pass's after map_fused)
    const uint64_t* cur_off_ = nullptr;
    const uint32_t* cur_chain_ = nullptr;
    // largest anchors-per-base ratio of a plain NAM pass (sizes the fused pass)
    double anchors_per_base_ = 0;
    Fused* fused_ = nullptr;
    uint64_t fused_replays_ = 0;
    uint64_t alloc_gen_ = 0;   // bumped whenever a device buffer is reallocated
    bool recording_ = false;   // the fused pass is being recorded: no buffer may grow
    bool owns_index_ = true;
    // growable device buffers indexed by slot
    std::vector<void*> buf_;
    std::vector<uint64_t> cap_;
    void* h_stage_[3] = {nullptr, nullptr, nullptr};
    uint64_t h_stage_cap_[3] = {0, 0, 0};
    StageTimes times_;
    std::vector<std::pair<std::string, KernelStat>> kstats_;
    double total_ms_ = 0;
    uint64_t lookup_bytes_ = 0;
};

}  // namespace sa
