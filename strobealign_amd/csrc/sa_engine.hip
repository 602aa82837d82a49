// Device engine: buffer management, kernel launches and the device-side scans
// that size each stage's output.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <functional>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>

#include "sa_engine.hpp"
#include "sa_kernels.hip"

#define SA_CHECK(x)                                                                                  \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess)                                                                        \
            throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " +   \
                                     __FILE__ + ":" + std::to_string(__LINE__) + " (" #x ")");       \
    } while (0)

// Scratch of the entry point that allocates nothing (sa_find_nams_batch_noalloc):
// each device that loads this code object holds one copy until the process ends.
// Host code never passes `scratch` itself to a kernel (in host code the name is a
// stand-in, not a device address); it uses the address hipGetSymbolAddress returns
// for the current device.
static __device__ __align__(256) uint8_t scratch[sa::NOALLOC_SCRATCH_BYTES];

namespace sa {

namespace {
// device address of `scratch` per device, looked up once per device
std::mutex g_scratch_mu;
constexpr int MAX_DEVICES = 64;
uint8_t* g_scratch_ptr[MAX_DEVICES] = {};
}  // namespace

uint8_t* noalloc_scratch(int device) {
    if (device < 0 || device >= MAX_DEVICES) throw std::invalid_argument("device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        SA_CHECK(hipSetDevice(device));
        void* ptr = nullptr;
        SA_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) throw std::runtime_error("HIP error: no device address for the no-allocation scratch");
        g_scratch_ptr[device] = (uint8_t*)ptr;
    }
    return g_scratch_ptr[device];
}

namespace {
__global__ void k_pair_sum(const uint32_t* __restrict__ a, uint64_t* __restrict__ out, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (uint64_t)a[2 * i] + a[2 * i + 1];
}
__global__ void k_ncigar(const ExtResult* __restrict__ r, uint64_t* __restrict__ out, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = r[i].found ? r[i].n_cigar : 0;
}
__global__ void k_compact_cigars(const ExtRequest* __restrict__ req, const ExtResult* __restrict__ res,
                                 const uint32_t* __restrict__ slots, const uint64_t* __restrict__ coff,
                                 uint32_t* __restrict__ out, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !res[i].found) return;
    const uint32_t* s = slots + req[i].cigar_off;
    uint32_t* d = out + coff[i];
    for (uint32_t x = 0; x < res[i].n_cigar; ++x) d[x] = s[x];
}

inline uint32_t blocks(uint64_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

// Seeding of n reads: a wave per read up to SEED_WAVE_MAX_LEN bases (lanes 64)
// and a thread per read beyond, or a thread per read throughout (lanes 1).
void launch_seed(const uint8_t* d_seq, const uint64_t* d_off, uint32_t n, const SeedParams& p, uint32_t lanes,
                 hipStream_t st, Syncmer* d_sy, QRand* d_qr, uint32_t* d_qrn) {
    if (lanes != 1 && lanes != 64) throw std::invalid_argument("seed lanes must be 1 or 64");
    if (lanes == 64) {
        k_seed_wave<<<blocks(n, 4), 256, 0, st>>>(d_seq, d_off, n, p, d_qr, d_qrn);
        SA_CHECK(hipGetLastError());
    }
    k_seed<<<blocks(n, 256), 256, 0, st>>>(d_seq, d_off, n, p, d_sy, d_qr, d_qrn,
                                           lanes == 64 ? SEED_WAVE_MAX_LEN : 0);
    SA_CHECK(hipGetLastError());
}

enum Slot {
    S_SEQ, S_OFF, S_SY, S_QR, S_QRN, S_HITS, S_CNT, S_HC, S_ACAP, S_AOFF, S_ANCH, S_ANCH2, S_DP, S_PRED, S_CAND, S_USED,
    S_NAMS_IN, S_CHAIN, S_NAMN, S_ARAW, S_TOT, S_OUTOFF, S_NAMS, S_SCANTMP, S_SCANTOT, S_REQ, S_RES, S_CIGSLOT,
    S_NCIG, S_CIGOFF, S_CIGOUT, S_SCRATCH, S_JOBS, S_LJOBS, S_PJOBS, S_ENDS, S_FWD, S_LISTS, S_BCNT, S_XCNT, S_CBYTES, S_HBYTES, S_FINLISTS, S_PWROWS, S_PWDIRS, S_TIERS,
    // used by the fused pass only (it shares the others with the plain stages)
    F_SEQ, F_OFF, F_CTL, F_SUM0, F_SUM1, F_SUM2, F_SUM3, F_SUM4, F_SPECN, F_SPECS, F_ROFF, F_SOFF, F_LSCR, F_PSCR,
    S_COUNT
};
}  // namespace

void* device_upload(int device, const void* src, uint64_t bytes) {
    SA_CHECK(hipSetDevice(device));
    void* p = nullptr;
    SA_CHECK(hipMalloc(&p, bytes ? bytes : 16));
    if (bytes) SA_CHECK(hipMemcpy(p, src, bytes, hipMemcpyHostToDevice));
    return p;
}

void* device_alloc(int device, uint64_t bytes) {
    SA_CHECK(hipSetDevice(device));
    void* p = nullptr;
    SA_CHECK(hipMalloc(&p, bytes ? bytes : 16));
    return p;
}

void device_copy_in(int device, void* dst, const void* src, uint64_t bytes) {
    if (!bytes) return;
    SA_CHECK(hipSetDevice(device));
    // a stream of the calling thread's own, so that copies overlap the engines' work
    struct ThreadStream {
        hipStream_t st = nullptr;
        int dev = -1;
        ~ThreadStream() {
            if (st) (void)hipStreamDestroy(st);
        }
    };
    thread_local ThreadStream ts;
    if (ts.dev != device) {
        if (ts.st) (void)hipStreamDestroy(ts.st);
        ts.st = nullptr;
        SA_CHECK(hipStreamCreateWithFlags(&ts.st, hipStreamNonBlocking));
        ts.dev = device;
    }
    SA_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ts.st));
    SA_CHECK(hipStreamSynchronize(ts.st));
}

uint64_t device_free_bytes(int device) {
    SA_CHECK(hipSetDevice(device));
    size_t fr = 0, tot = 0;
    SA_CHECK(hipMemGetInfo(&fr, &tot));
    return fr;
}

void ssw_raw_device(int device, const int8_t* read, int32_t readLen, const int8_t* ref, int32_t refLen,
                    const int8_t* mat, int32_t n, int32_t score_size, int32_t gapO, int32_t gapE, uint32_t flag,
                    uint32_t filters, int32_t filterd, int32_t maskLen, SswRaw& out, std::vector<uint32_t>& cigar) {
    // one device context for all shim calls of the process, used one call at a time
    struct Ctx {
        int dev = -1;
        hipStream_t st = nullptr;
        uint8_t* buf = nullptr;
        uint64_t cap = 0;
    };
    static std::mutex mu;
    static Ctx c;
    std::lock_guard<std::mutex> lk(mu);
    SA_CHECK(hipSetDevice(device));
    if (c.dev != device) {
        if (c.st) (void)hipStreamDestroy(c.st);
        if (c.buf) (void)hipFree(c.buf);
        c = Ctx{};
        SA_CHECK(hipStreamCreateWithFlags(&c.st, hipStreamNonBlocking));
        c.dev = device;
    }
    const uint64_t cig_cap = (uint64_t)readLen + (uint64_t)refLen + 4;
    auto al = [](uint64_t x) { return (x + 255) & ~(uint64_t)255; };
    const uint64_t o_read = 0, o_ref = al(o_read + (uint64_t)readLen + 1), o_mat = al(o_ref + (uint64_t)refLen + 1);
    const uint64_t o_out = al(o_mat + (uint64_t)(n * n) + 1), o_cig = al(o_out + sizeof(SswRaw));
    const uint64_t o_scr = al(o_cig + 4 * cig_cap);
    const uint64_t scr = ssw_raw_scratch(readLen, refLen);
    const uint64_t need = o_scr + scr;
    if (need > c.cap) {
        if (c.buf) SA_CHECK(hipFree(c.buf));
        c.buf = nullptr;
        c.cap = 0;
        SA_CHECK(hipMalloc((void**)&c.buf, need));
        c.cap = need;
    }
    SA_CHECK(hipMemcpyAsync(c.buf + o_read, read, (uint64_t)readLen, hipMemcpyHostToDevice, c.st));
    SA_CHECK(hipMemcpyAsync(c.buf + o_ref, ref, (uint64_t)refLen, hipMemcpyHostToDevice, c.st));
    SA_CHECK(hipMemcpyAsync(c.buf + o_mat, mat, (uint64_t)(n * n), hipMemcpyHostToDevice, c.st));
    k_ssw_raw<<<1, 64, 0, c.st>>>((const int8_t*)(c.buf + o_read), readLen, (const int8_t*)(c.buf + o_ref), refLen,
                                  (const int8_t*)(c.buf + o_mat), n, score_size, gapO, gapE, flag, filters, filterd,
                                  maskLen, c.buf + o_scr, scr, (uint32_t*)(c.buf + o_cig), (uint32_t)cig_cap,
                                  (SswRaw*)(c.buf + o_out));
    SA_CHECK(hipGetLastError());
    SA_CHECK(hipMemcpyAsync(&out, c.buf + o_out, sizeof(SswRaw), hipMemcpyDeviceToHost, c.st));
    SA_CHECK(hipStreamSynchronize(c.st));
    cigar.assign((size_t)(out.status == 0 ? out.n_cigar : 0), 0u);
    if (!cigar.empty()) {
        SA_CHECK(hipMemcpyAsync(cigar.data(), c.buf + o_cig, cigar.size() * 4, hipMemcpyDeviceToHost, c.st));
        SA_CHECK(hipStreamSynchronize(c.st));
    }
}

__global__ void __launch_bounds__(256) k_hbm_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

double hbm_copy_probe(int device, uint64_t bytes, int iters) {
    SA_CHECK(hipSetDevice(device));
    const uint64_t n = bytes / 16;
    void *a = nullptr, *b = nullptr;
    SA_CHECK(hipMalloc(&a, n * 16 + 16));
    SA_CHECK(hipMalloc(&b, n * 16 + 16));
    SA_CHECK(hipMemset(a, 1, n * 16));
    hipEvent_t e0, e1;
    SA_CHECK(hipEventCreate(&e0));
    SA_CHECK(hipEventCreate(&e1));
    const uint32_t grid = 256 * 8 * 4;  // 8 XCDs x 32 CUs x 4 blocks of 256 per CU
    k_hbm_copy<<<grid, 256>>>((const uint4*)a, (uint4*)b, n);  // warm-up
    SA_CHECK(hipEventRecord(e0, nullptr));
    for (int i = 0; i < iters; ++i) k_hbm_copy<<<grid, 256>>>((const uint4*)a, (uint4*)b, n);
    SA_CHECK(hipEventRecord(e1, nullptr));
    SA_CHECK(hipEventSynchronize(e1));
    float ms = 0;
    SA_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return ms > 0 ? 2.0 * (double)n * 16 * iters / (ms * 1e-3) / 1e9 : 0.0;
}

void device_free(int device, void* p) {
    if (!p) return;
    (void)hipSetDevice(device);
    (void)hipFree(p);
}

// ------------------------------------------------------------ fused pass ---
namespace {
// Control words of the fused pass (u64 each, zeroed at its start, copied to the
// host at its end): anchor overflow flag (u32), speculative requests written
// (u32), NAMs, anchor capacity, requests and CIGAR slots the selection wanted,
// CIGAR ops of the results.
enum FusedCtl { C_FLAG, C_NREQ, C_NNAMS, C_ANCH, C_WANT, C_WANTS, C_NCIG, C_COUNT };
// Fixed grids of the launches whose counts exist only on the device.
constexpr uint32_t F_PREP_BLOCKS = 4096;   // x 4 waves, a request per wave
constexpr uint32_t F_SSW_BLOCKS = 1024;    // x 256 lanes
constexpr uint32_t F_FIN_SLOTS = 32768;    // traceback threads (scratch slots), at most
constexpr uint32_t F_LONG_SLOTS = 2048;    // one-thread SSW of long queries, at most
constexpr uint64_t F_FIN_BUDGET = 2ull << 30;   // bytes of traceback scratch, at most
constexpr uint64_t F_LONG_BUDGET = 2ull << 30;  // bytes of one-thread SSW scratch, at most
constexpr uint64_t F_PW_BUDGET = 4ull << 30;    // bytes of piecewise scratch
// extension counters (32-bit words): [0..3] prep (ssw, long, pw jobs; one unused),
// the SSW buckets, four 64-bit DP-cell counts (byte forward, byte reverse, word
// forward, word reverse), then the traceback classes
constexpr uint32_t EXT_CELLS = 4 + SSW_BUCKETS;
constexpr uint32_t EXT_FIN = EXT_CELLS + 8;
constexpr uint32_t EXT_NCNT = EXT_FIN + FIN_CLASSES;
static_assert(EXT_CELLS % 2 == 0, "cell counter alignment");
uint64_t round_up(uint64_t x, uint64_t m) { return (x + m - 1) / m * m; }
}  // namespace

struct Engine::Fused {
    // what the pass was sized and recorded for (len_min: no read of a batch is
    // shorter, so the forward SSW passes of fewer stripes are not recorded)
    uint32_t n_cap = 0, len_cap = 0, len_min = 0;
    uint64_t bases_cap = 0, a_cap = 0, req_cap = 0, slot_cap = 0;
    MapParams mp{};
    ChainParams cp{};
    SpecParams sp{};
    Scores sc{};
    const RefEntry* entries = nullptr;
    bool stale = true;  // resize before the next pass
    int runs = 0;       // passes with this shape (the first runs directly)
    // sizing from what the passes saw
    double req_per_read = 0, slots_per_req = 0;
    // The pass works in the engine's buffers (the slots it shares with the
    // plain stages are sized for whichever needs more). The recorded graph
    // holds their addresses, so it is valid while the engine's allocation
    // generation is the one it was recorded at.
    uint64_t gen = ~0ull;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t* h_ctl = nullptr;  // pinned: C_COUNT control words, then the EXT_NCNT extension counters
    int device = 0;

    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~Fused() {
        (void)hipSetDevice(device);
        drop_graph();
        if (h_ctl) (void)hipHostFree(h_ctl);
    }
};

Engine::Engine(int device) : device_(device) {
    int n = 0;
    SA_CHECK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) throw std::runtime_error("no HIP device " + std::to_string(device));
    SA_CHECK(hipSetDevice(device));
    // non-blocking: no implicit ordering with the null stream or other engines' streams
    SA_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    own_stream_ = stream_;
    // waits on this event sleep instead of spinning, leaving the host cores to the replay
    SA_CHECK(hipEventCreateWithFlags(&sync_ev_, hipEventBlockingSync | hipEventDisableTiming));
    buf_.assign(S_COUNT, nullptr);
    cap_.assign(S_COUNT, 0);
}

Engine::Engine(ShareIndex, const Engine& owner) : Engine(owner.device_) {
    // shares the owner's read-only reference, index and log2 table
    p_ = owner.p_;
    n_entries_ = owner.n_entries_;
    d_entries_ = owner.d_entries_;
    d_buckets_ = owner.d_buckets_;
    d_words_ = owner.d_words_;
    d_word_off_ = owner.d_word_off_;
    d_len_ = owner.d_len_;
    n_contigs_ = owner.n_contigs_;
    d_log2_ = owner.d_log2_;
    owns_index_ = false;
    SA_CHECK(hipMalloc(&d_same_diag_, 1024 * sizeof(float)));
}

Engine::~Engine() {
    (void)hipSetDevice(device_);  // destructor: errors are not reported
    drain();  // no kernel of this engine may outlive its buffers
    delete fused_;
    for (void* p : buf_)
        if (p) (void)hipFree(p);
    if (owns_index_)
        for (void* p : {(void*)d_entries_, (void*)d_buckets_, (void*)d_words_, (void*)d_word_off_, (void*)d_len_,
                        (void*)d_log2_})
            if (p) (void)hipFree(p);
    if (d_same_diag_) (void)hipFree(d_same_diag_);
    for (void* h : h_stage_)
        if (h) (void)hipHostFree(h);
    if (sync_ev_) (void)hipEventDestroy(sync_ev_);
    for (hipEvent_t e : event_pool_) (void)hipEventDestroy(e);
    if (own_stream_) (void)hipStreamDestroy(own_stream_);
}

hipStream_t Engine::set_stream(hipStream_t st) {
    hipStream_t prev = stream_;
    stream_ = st ? st : own_stream_;
    return prev;
}

hipEvent_t Engine::event() {
    if (!event_pool_.empty()) {
        hipEvent_t e = event_pool_.back();
        event_pool_.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    SA_CHECK(hipEventCreate(&e));
    return e;
}

void Engine::release_event(hipEvent_t e) { event_pool_.push_back(e); }

void Engine::sync() {
    SA_CHECK(hipEventRecord(sync_ev_, stream_));
    SA_CHECK(hipEventSynchronize(sync_ev_));
}

void Engine::add_kernel(const char* name, double ms, uint64_t items, uint64_t bytes, uint64_t cells) {
    for (auto& k : kstats_)
        if (k.first == name) {
            k.second.ms += ms;
            k.second.launches += 1;
            k.second.items += items;
            k.second.bytes += bytes;
            k.second.cells += cells;
            return;
        }
    kstats_.push_back({name, KernelStat{ms, 1, items, bytes, cells}});
}

void* Engine::grow(int slot, uint64_t bytes) {
    if (bytes == 0) bytes = 16;
    if (cap_[slot] < bytes) {
        if (recording_) throw std::logic_error("a device buffer would grow while the fused pass is recorded");
        ++alloc_gen_;
        void* old = buf_[slot];
        buf_[slot] = nullptr;
        cap_[slot] = 0;
        if (old) SA_CHECK(hipFree(old));
        uint64_t want = bytes + bytes / 4;
        void* p = nullptr;
        SA_CHECK(hipMalloc(&p, want));
        buf_[slot] = p;
        cap_[slot] = want;
    }
    return buf_[slot];
}

void Engine::drain() noexcept {
    (void)hipSetDevice(device_);
    if (stream_) (void)hipStreamSynchronize(stream_);
    (void)hipGetLastError();
}

void* Engine::host_stage(int slot, uint64_t bytes) {
    if (h_stage_cap_[slot] < bytes) {
        if (h_stage_[slot]) SA_CHECK(hipHostFree(h_stage_[slot]));
        const uint64_t want = bytes + bytes / 4 + 4096;
        SA_CHECK(hipHostMalloc(&h_stage_[slot], want, hipHostMallocDefault));
        h_stage_cap_[slot] = want;
    }
    return h_stage_[slot];
}

void Engine::scan_u64(const uint64_t* in, uint64_t* out, uint64_t n, uint64_t* total) {
    if (n == 0) { *total = 0; return; }
    size_t tmp = 0;
    SA_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, in, out, (int)n, stream_));
    void* t = grow(S_SCANTMP, tmp);
    SA_CHECK(hipcub::DeviceScan::ExclusiveSum(t, tmp, in, out, (int)n, stream_));
    uint64_t last[2];
    SA_CHECK(hipMemcpyAsync(&last[0], in + n - 1, 8, hipMemcpyDeviceToHost, stream_));
    SA_CHECK(hipMemcpyAsync(&last[1], out + n - 1, 8, hipMemcpyDeviceToHost, stream_));
    sync();
    *total = last[0] + last[1];
}

void Engine::upload_reference(const uint64_t* words, uint64_t n_words, const uint64_t* word_off, const uint64_t* len,
                              uint32_t n_contigs) {
    SA_CHECK(hipSetDevice(device_));
    SA_CHECK(hipMalloc(&d_words_, std::max<uint64_t>(n_words, 1) * 8));
    SA_CHECK(hipMalloc(&d_word_off_, (uint64_t)(n_contigs + 1) * 8));
    SA_CHECK(hipMalloc(&d_len_, std::max<uint64_t>(n_contigs, 1) * 8));
    SA_CHECK(hipMemcpy(d_words_, words, n_words * 8, hipMemcpyHostToDevice));
    SA_CHECK(hipMemcpy(d_word_off_, word_off, (uint64_t)(n_contigs + 1) * 8, hipMemcpyHostToDevice));
    SA_CHECK(hipMemcpy(d_len_, len, (uint64_t)n_contigs * 8, hipMemcpyHostToDevice));
    n_contigs_ = n_contigs;
}

void Engine::upload_index(const RefEntry* entries, uint64_t n, const uint64_t* bucket_starts, const SeedParams& p) {
    SA_CHECK(hipSetDevice(device_));
    p_ = p;
    p_.n_entries = n;
    n_entries_ = n;
    uint64_t nb = (1ull << p.bits) + 1;
    SA_CHECK(hipMalloc(&d_entries_, std::max<uint64_t>(n, 1) * sizeof(RefEntry)));
    SA_CHECK(hipMalloc(&d_buckets_, nb * 8));
    if (n) SA_CHECK(hipMemcpy(d_entries_, entries, n * sizeof(RefEntry), hipMemcpyHostToDevice));
    SA_CHECK(hipMemcpy(d_buckets_, bucket_starts, nb * 8, hipMemcpyHostToDevice));
    setup_chain_tables();
}

void Engine::upload_index_streamed(uint64_t n, const SeedParams& p, const std::function<void(void*, uint64_t)>& read) {
    SA_CHECK(hipSetDevice(device_));
    const uint64_t nb = (1ull << p.bits) + 1;
    SA_CHECK(hipMalloc(&d_entries_, std::max<uint64_t>(n, 1) * sizeof(RefEntry)));
    SA_CHECK(hipMalloc(&d_buckets_, nb * 8));
    // two pinned staging buffers: one fills from the reader while the other's
    // copy to the device runs
    const uint64_t chunk = 64ull << 20;
    uint8_t* stage[2] = {(uint8_t*)host_stage(0, chunk), (uint8_t*)host_stage(1, chunk)};
    hipEvent_t done[2] = {event(), event()};
    SA_CHECK(hipEventRecord(done[0], stream_));
    SA_CHECK(hipEventRecord(done[1], stream_));
    uint64_t turn = 0;
    auto stream_in = [&](uint8_t* dst, uint64_t total) {
        for (uint64_t at = 0; at < total; at += chunk, ++turn) {
            const uint64_t b = std::min(chunk, total - at);
            const int x = (int)(turn & 1);
            SA_CHECK(hipEventSynchronize(done[x]));
            read(stage[x], b);
            SA_CHECK(hipMemcpyAsync(dst + at, stage[x], b, hipMemcpyHostToDevice, stream_));
            SA_CHECK(hipEventRecord(done[x], stream_));
        }
    };
    try {
        stream_in((uint8_t*)d_entries_, n * sizeof(RefEntry));
        uint64_t nb_file = 0;
        read(&nb_file, 8);
        if (nb_file != nb)
            throw std::invalid_argument("The randstrobe starts vector has an unexpected size in the .sti file");
        stream_in((uint8_t*)d_buckets_, nb * 8);
    } catch (...) {
        sync();  // no copy out of the staging buffers may outlive this call
        release_event(done[0]);
        release_event(done[1]);
        throw;
    }
    sync();
    release_event(done[0]);
    release_event(done[1]);
    p_ = p;
    p_.n_entries = n;
    n_entries_ = n;
    setup_chain_tables();
}

void Engine::setup_chain_tables() {
    if (d_log2_) return;
    // log2(d + 1) for the chaining gap penalty, from the host libm (src/chainer.rs:294-298)
    std::vector<float> lg(65536);
    for (size_t d = 0; d < lg.size(); ++d) lg[d] = std::log2((float)d + 1.0f);
    SA_CHECK(hipMalloc(&d_log2_, lg.size() * sizeof(float)));
    SA_CHECK(hipMemcpy(d_log2_, lg.data(), lg.size() * sizeof(float), hipMemcpyHostToDevice));
    SA_CHECK(hipMalloc(&d_same_diag_, 1024 * sizeof(float)));
}

namespace {
// filter cutoff of src/indexer.rs:109-129 from the histogram of repeated-hash run
// lengths (slot 0: distinct hashes, slot c: runs of c entries, last slot: longer):
// the index_cutoff-th largest run length clamped to [30, 100] only needs the
// lengths up to 101.
uint64_t cutoff_from_hist(const unsigned long long* h, double filter_fraction) {
    const uint64_t unique = h[0];
    uint64_t total = 0;
    for (uint32_t c = 2; c < IX_HIST; ++c) total += h[c];
    uint64_t idx = (uint64_t)((double)unique * filter_fraction), fc = 30;
    if (idx < total) {
        for (uint32_t c = IX_HIST - 1; c >= 2; --c) {
            if (idx < h[c]) { fc = c; break; }
            idx -= h[c];
        }
    } else if (total > 0) {
        for (uint32_t c = 2; c < IX_HIST; ++c)
            if (h[c]) { fc = c; break; }
    }
    return std::min<uint64_t>(std::max<uint64_t>(fc, 30), 100);
}
}  // namespace

SeedParams Engine::build_index(const uint64_t* len, uint32_t n_contigs, SeedParams p, int bits,
                               double filter_fraction) {
    SA_CHECK(hipSetDevice(device_));
    p.bits = bits;
    // chunks of syncmer start positions per contig
    constexpr uint32_t C = 4096;
    std::vector<IxChunk> ch;
    std::vector<uint64_t> first_chunk(n_contigs + 1);
    for (uint32_t c = 0; c < n_contigs; ++c) {
        first_chunk[c] = ch.size();
        const uint64_t L = len[c];
        if (L < (uint64_t)p.w_max) continue;
        if (L >= (1ull << 31)) throw std::runtime_error("contigs longer than 2^31 bp are not supported");
        for (uint64_t s = 0; s < L; s += C) ch.push_back(IxChunk{c, (uint32_t)s, (uint32_t)std::min<uint64_t>(L, s + C), 0});
    }
    first_chunk[n_contigs] = ch.size();
    const uint32_t nch = (uint32_t)ch.size();
    IxChunk* d_ch = nullptr;
    uint64_t *d_cnt = nullptr, *d_off = nullptr, *d_syoff = nullptr;
    SA_CHECK(hipMalloc(&d_ch, std::max<uint64_t>(nch, 1) * sizeof(IxChunk)));
    SA_CHECK(hipMalloc(&d_cnt, (uint64_t)(nch + 1) * 8));
    SA_CHECK(hipMalloc(&d_off, (uint64_t)(nch + 1) * 8));
    SA_CHECK(hipMalloc(&d_syoff, (uint64_t)(n_contigs + 1) * 8));
    if (nch) SA_CHECK(hipMemcpyAsync(d_ch, ch.data(), nch * sizeof(IxChunk), hipMemcpyHostToDevice, stream_));
    k_ix_syncmers<true><<<blocks(nch, 256), 256, 0, stream_>>>(d_words_, d_word_off_, d_len_, d_ch, nch, p, d_cnt,
                                                                nullptr, nullptr);
    SA_CHECK(hipGetLastError());
    uint64_t n = 0;
    scan_u64(d_cnt, d_off, nch, &n);
    if (n >= (1ull << 31)) throw std::runtime_error("more than 2^31 index entries: build the index on the host");
    std::vector<uint64_t> off(nch);
    if (nch) SA_CHECK(hipMemcpy(off.data(), d_off, nch * 8, hipMemcpyDeviceToHost));
    std::vector<uint64_t> syoff(n_contigs + 1);
    for (uint32_t c = 0; c <= n_contigs; ++c) syoff[c] = first_chunk[c] < nch ? off[first_chunk[c]] : n;
    SA_CHECK(hipMemcpy(d_syoff, syoff.data(), (uint64_t)(n_contigs + 1) * 8, hipMemcpyHostToDevice));
    Syncmer* d_sy = nullptr;
    RefEntry* d_e = nullptr;
    SA_CHECK(hipMalloc(&d_sy, std::max<uint64_t>(n, 1) * sizeof(Syncmer)));
    k_ix_syncmers<false><<<blocks(nch, 256), 256, 0, stream_>>>(d_words_, d_word_off_, d_len_, d_ch, nch, p, nullptr,
                                                                 d_off, d_sy);
    SA_CHECK(hipGetLastError());
    SA_CHECK(hipMalloc(&d_e, std::max<uint64_t>(n, 1) * sizeof(RefEntry)));
    k_ix_randstrobes<<<blocks(n, 256), 256, 0, stream_>>>(d_sy, n, d_syoff, n_contigs, p, d_e);
    SA_CHECK(hipGetLastError());
    sync();
    SA_CHECK(hipFree(d_sy));
    SA_CHECK(hipFree(d_ch));
    SA_CHECK(hipFree(d_cnt));
    SA_CHECK(hipFree(d_off));
    SA_CHECK(hipFree(d_syoff));
    // order (hash_offset, position, contig): stable by (position, contig), then
    // stable by hash_offset
    uint64_t *k0 = nullptr, *k1 = nullptr;
    uint32_t *i0 = nullptr, *i1 = nullptr;
    const uint64_t nn = std::max<uint64_t>(n, 1);
    SA_CHECK(hipMalloc(&k0, nn * 8));
    SA_CHECK(hipMalloc(&k1, nn * 8));
    SA_CHECK(hipMalloc(&i0, nn * 4));
    SA_CHECK(hipMalloc(&i1, nn * 4));
    void* tmp = nullptr;
    size_t tmp_bytes = 0;
    if (n) {
        k_ix_poskey<<<blocks(n, 256), 256, 0, stream_>>>(d_e, n, k0, i0);
        SA_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, k0, k1, i0, i1, (int)n, 0, 64, stream_));
        SA_CHECK(hipMalloc(&tmp, tmp_bytes));
        SA_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, k0, k1, i0, i1, (int)n, 0, 64, stream_));
        k_ix_hashkey<<<blocks(n, 256), 256, 0, stream_>>>(d_e, i1, n, k0);
        SA_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, k0, k1, i1, i0, (int)n, 0, 64, stream_));
    }
    p.n_entries = n;
    p_ = p;
    n_entries_ = n;
    const uint64_t nb = (1ull << bits) + 1;
    if (d_entries_) SA_CHECK(hipFree(d_entries_));
    if (d_buckets_) SA_CHECK(hipFree(d_buckets_));
    SA_CHECK(hipMalloc(&d_entries_, nn * sizeof(RefEntry)));
    SA_CHECK(hipMalloc(&d_buckets_, nb * 8));
    if (n) k_ix_gather<<<blocks(n, 256), 256, 0, stream_>>>(d_e, i0, n, d_entries_);
    sync();
    SA_CHECK(hipFree(d_e));
    SA_CHECK(hipFree(k0));
    SA_CHECK(hipFree(k1));
    SA_CHECK(hipFree(i0));
    SA_CHECK(hipFree(i1));
    if (tmp) SA_CHECK(hipFree(tmp));
    // bucket directory
    uint64_t* d_first = nullptr;
    unsigned long long* d_hist = nullptr;
    SA_CHECK(hipMalloc(&d_first, 8));
    SA_CHECK(hipMalloc(&d_hist, IX_HIST * 8));
    k_ix_fill<<<blocks(nb, 256), 256, 0, stream_>>>(d_buckets_, nb, n);
    k_ix_fill<<<1, 1, 0, stream_>>>(d_first, 1, n);
    SA_CHECK(hipMemsetAsync(d_hist, 0, IX_HIST * 8, stream_));
    if (n) {
        k_ix_buckets<<<blocks(n, 256), 256, 0, stream_>>>(d_entries_, n, bits, d_buckets_, d_first);
        RefEntry e0;
        SA_CHECK(hipMemcpyAsync(&e0, d_entries_, sizeof e0, hipMemcpyDeviceToHost, stream_));
        sync();
        const uint64_t top0 = (e0.hash_offset & REF_HASH_MASK) >> (64 - bits);
        k_ix_buckets_head<<<blocks(top0 + 1, 256), 256, 0, stream_>>>(d_entries_, n, bits, d_buckets_, d_first);
        k_ix_runs<<<blocks(n, 256), 256, 0, stream_>>>(d_entries_, n, d_hist);
    }
    SA_CHECK(hipGetLastError());
    unsigned long long hist[IX_HIST];
    SA_CHECK(hipMemcpyAsync(hist, d_hist, sizeof hist, hipMemcpyDeviceToHost, stream_));
    sync();
    SA_CHECK(hipFree(d_first));
    SA_CHECK(hipFree(d_hist));
    p_.filter_cutoff = cutoff_from_hist(hist, filter_fraction);
    setup_chain_tables();
    return p_;
}

void Engine::download_index(RefEntry* entries, uint64_t* bucket_starts) const {
    SA_CHECK(hipSetDevice(device_));
    if (n_entries_) SA_CHECK(hipMemcpy(entries, d_entries_, n_entries_ * sizeof(RefEntry), hipMemcpyDeviceToHost));
    SA_CHECK(hipMemcpy(bucket_starts, d_buckets_, ((1ull << p_.bits) + 1) * 8, hipMemcpyDeviceToHost));
}

void Engine::upload_reads(const uint8_t* seq, const uint64_t* off, uint32_t n_reads) {
    SA_CHECK(hipSetDevice(device_));
    const uint64_t nb = off[n_reads];
    void* s = grow(S_SEQ, nb + 16);
    if (nb) SA_CHECK(hipMemcpyAsync(s, seq, nb, hipMemcpyHostToDevice, stream_));
    use_reads((const uint8_t*)s, off, n_reads);
}

void Engine::gather_reads(const uint8_t* d_base, const uint64_t* seg, const uint64_t* seg_len, uint32_t n_segs,
                          const uint64_t* off, uint32_t n_reads) {
    SA_CHECK(hipSetDevice(device_));
    // adjacent segments are copied (or used) as one
    std::vector<std::pair<uint64_t, uint64_t>> runs;
    for (uint32_t i = 0; i < n_segs; ++i) {
        if (!runs.empty() && runs.back().first + runs.back().second == seg[i])
            runs.back().second += seg_len[i];
        else if (seg_len[i])
            runs.push_back({seg[i], seg_len[i]});
    }
    if (runs.size() <= 1) {
        use_reads(d_base + (runs.empty() ? 0 : runs[0].first), off, n_reads);
        return;
    }
    uint8_t* s = (uint8_t*)grow(S_SEQ, off[n_reads] + 16);
    uint64_t at = 0;
    for (const auto& r : runs) {
        SA_CHECK(hipMemcpyAsync(s + at, d_base + r.first, r.second, hipMemcpyDeviceToDevice, stream_));
        at += r.second;
    }
    if (at != off[n_reads]) throw std::invalid_argument("gather_reads: segments do not cover the offsets");
    use_reads(s, off, n_reads);
}

void Engine::use_reads(const uint8_t* d_seq, const uint64_t* off, uint32_t n_reads) {
    SA_CHECK(hipSetDevice(device_));
    n_reads_ = n_reads;
    h_off_.assign(off, off + n_reads + 1);
    n_bases_ = h_off_[n_reads];
    max_len_ = 0;
    min_len_ = 0;
    for (uint32_t i = 0; i < n_reads; ++i) {
        const uint32_t L = (uint32_t)(off[i + 1] - off[i]);
        max_len_ = std::max<uint32_t>(max_len_, L);
        if (L && (min_len_ == 0 || L < min_len_)) min_len_ = L;
    }
    if (max_len_ >= 65536) throw std::runtime_error("reads longer than 65535 bp are not supported");
    d_seq_ = d_seq;
    void* o = grow(S_OFF, (uint64_t)(n_reads + 1) * 8);
    SA_CHECK(hipMemcpyAsync(o, off, (uint64_t)(n_reads + 1) * 8, hipMemcpyHostToDevice, stream_));
    cur_off_ = (const uint64_t*)o;
    sync();
}

void Engine::set_chain_anchors(const uint32_t* pairs, uint64_t n_pairs) {
    SA_CHECK(hipSetDevice(device_));
    void* d = grow(S_CHAIN, n_pairs * 8 + 8);
    if (n_pairs) SA_CHECK(hipMemcpyAsync(d, pairs, n_pairs * 8, hipMemcpyHostToDevice, stream_));
    cur_chain_ = (const uint32_t*)d;
    sync();
}

void Engine::randstrobes(std::vector<QRand>& qr, std::vector<uint32_t>& counts) {
    const uint32_t n = n_reads_;
    auto* d_seq = d_seq_;
    auto* d_off = (const uint64_t*)buf_[S_OFF];
    auto* d_sy = (Syncmer*)grow(S_SY, (p_.w_max < 24 ? 1 : n_bases_ + 1) * sizeof(Syncmer));
    auto* d_qr = (QRand*)grow(S_QR, (2 * n_bases_ + 1) * sizeof(QRand));
    auto* d_qrn = (uint32_t*)grow(S_QRN, (2ull * n + 1) * 4);
    if (n) launch_seed(d_seq, d_off, n, p_, 64, stream_, d_sy, d_qr, d_qrn);
    qr.resize(2 * n_bases_);
    counts.resize(2ull * n);
    if (n_bases_) SA_CHECK(hipMemcpyAsync(qr.data(), d_qr, 2 * n_bases_ * sizeof(QRand), hipMemcpyDeviceToHost, stream_));
    if (n) SA_CHECK(hipMemcpyAsync(counts.data(), d_qrn, 2ull * n * 4, hipMemcpyDeviceToHost, stream_));
    sync();
}

namespace {
// lanes per piecewise-extension job
constexpr int PW_LANES = 16;
// anchor capacities of the LDS chaining tiers
constexpr uint32_t CHAIN_CAP1 = 64, CHAIN_CAP2 = 256, CHAIN_CAP3 = 1024;

}  // namespace

// Device buffers and counts of one seeding -> NAM pass.
struct NamPass {
    uint32_t* qrn = nullptr;     // 2n: randstrobes per read orientation (null when hits were handed in)
    HitCounts* hc = nullptr;     // 2n
    uint32_t* namn = nullptr;    // 2n: NAMs per read orientation
    uint32_t* araw = nullptr;    // 2n: anchors before dedup
    uint32_t* cbytes = nullptr;  // 2n: chaining algorithmic bytes
    uint32_t* hbytes = nullptr;  // 2n: lookup algorithmic bytes (null when hits were handed in)
    uint64_t* outoff = nullptr;  // n: first NAM of each read
    NamRec* nams = nullptr;      // n_nams
    uint64_t n_nams = 0;
};

namespace {

// Device buffers of the lookup stage: hits of read orientation t at
// 2 * off[t / 2] + (t % 2) * L (L the read's length), hc[t].n_hits of them.
struct HitPass {
    Hit* hits = nullptr;
    HitCounts* hc = nullptr;
    uint64_t* acap = nullptr;  // 2n: anchor capacity per read orientation
};

// Seeding and lookups (k_seed, k_hits) for the n reads at d_seq with device
// offsets d_off (n_bases bases), on stream st; take(slot, bytes) supplies every
// device buffer. ev (or null) gets [0, 1] seed, [1, 2] hits.
template <typename Take>
void hit_pass(const IndexView& ix, const uint8_t* d_seq, const uint64_t* d_off, uint32_t n, uint64_t n_bases,
              const MapParams& mp, hipStream_t st, Take take, hipEvent_t* ev, HitPass& h, NamPass& o) {
    const SeedParams& p = ix.p;
    // the syncmer list is only materialised by the fallback for w_max >= 24
    auto* d_sy = (Syncmer*)take(S_SY, (p.w_max < 24 ? 1 : n_bases + 1) * sizeof(Syncmer));
    auto* d_qr = (QRand*)take(S_QR, (2 * n_bases + 1) * sizeof(QRand));
    o.qrn = (uint32_t*)take(S_QRN, (2ull * n + 1) * 4);
    h.hits = (Hit*)take(S_HITS, (2 * n_bases + 1) * sizeof(Hit));
    auto* d_cnt = (uint64_t*)take(S_CNT, (2 * n_bases + 1) * 8);
    o.hc = h.hc = (HitCounts*)take(S_HC, (2ull * n + 1) * sizeof(HitCounts));
    h.acap = (uint64_t*)take(S_ACAP, (2ull * n + 1) * 8);
    o.hbytes = (uint32_t*)take(S_HBYTES, (2ull * n + 1) * 4);
    if (ev) SA_CHECK(hipEventRecord(ev[0], st));
    launch_seed(d_seq, d_off, n, p, mp.seed_lanes, st, d_sy, d_qr, o.qrn);
    if (ev) SA_CHECK(hipEventRecord(ev[1], st));
    switch (mp.lookup_lanes) {
    case 1:
        k_hits<<<blocks(2ull * n, 256), 256, 0, st>>>(ix, mp, d_off, n, d_qr, o.qrn, h.hits, d_cnt, h.hc, h.acap,
                                                      o.hbytes);
        break;
    case 8:
        k_hits_grp<8><<<blocks(2ull * n * 8, 256), 256, 0, st>>>(ix, mp, d_off, n, d_qr, o.qrn, h.hits, d_cnt, h.hc,
                                                                 h.acap, o.hbytes);
        break;
    case 16:
        k_hits_grp<16><<<blocks(2ull * n * 16, 256), 256, 0, st>>>(ix, mp, d_off, n, d_qr, o.qrn, h.hits, d_cnt, h.hc,
                                                                   h.acap, o.hbytes);
        break;
    case 64:
        k_hits_grp<64><<<blocks(2ull * n * 64, 256), 256, 0, st>>>(ix, mp, d_off, n, d_qr, o.qrn, h.hits, d_cnt, h.hc,
                                                                   h.acap, o.hbytes);
        break;
    default:
        throw std::invalid_argument("lookup lanes must be 1, 8, 16 or 64");
    }
    SA_CHECK(hipGetLastError());
    if (ev) SA_CHECK(hipEventRecord(ev[2], st));
}

// Chaining -> NAMs from the hits of a lookup stage (or hits handed in): anchor
// capacities scanned into per-orientation offsets (scan(in, out, n) writes an
// exclusive prefix sum and returns its total: a host wait), the chaining tiers,
// then NAM counts scanned (the second host wait) and nams_out(N) supplies the
// NAM output for k_finalize. ev (or null) gets [3, 4] chain, [5, 6] finalize.
// abort (or null): a device flag that empties every orientation's results (the
// fused pass sets it when the anchors exceed their fixed buffer).
template <typename Take, typename Scan, typename NamsOut>
void chain_pass(const IndexView& ix, const RefView& ref, const float* log2tab, const float* same_diag,
                const uint8_t* d_seq, const uint64_t* d_off, uint32_t n, const ChainParams& cp, hipStream_t st,
                Take take, Scan scan, NamsOut nams_out, hipEvent_t* ev, const HitPass& h, NamPass& o,
                const uint32_t* abort = nullptr) {
    const uint32_t k = (uint32_t)ix.p.k;
    auto* d_aoff = (uint64_t*)take(S_AOFF, (2ull * n + 1) * 8);
    const uint64_t A = scan(h.acap, d_aoff, 2ull * n);
    ChainBufs b;
    b.anchors = (Anchor*)take(S_ANCH, (A + 1) * sizeof(Anchor));
    b.anchors_tmp = (Anchor*)take(S_ANCH2, (A + 1) * sizeof(Anchor));
    b.dp = (float*)take(S_DP, (A + 1) * 4);
    b.pred = (uint32_t*)take(S_PRED, (A + 1) * 4);
    b.cand = (uint32_t*)take(S_CAND, (A + 1) * 4);
    b.used = (uint8_t*)take(S_USED, A + 1);
    b.nams = (NamRec*)take(S_NAMS_IN, (A + 1) * sizeof(NamRec));
    b.chain_anchors = (uint32_t*)take(S_CHAIN, (4 * A + 4) * 4);
    b.alg_bytes = o.cbytes = (uint32_t*)take(S_CBYTES, (2ull * n + 1) * 4);
    o.namn = (uint32_t*)take(S_NAMN, (2ull * n + 1) * 4);
    o.araw = (uint32_t*)take(S_ARAW, (2ull * n + 1) * 4);
    // chaining tiers by anchor capacity, all in LDS: a 16-lane group per
    // orientation up to 64 anchors, a wave up to 256 and up to 1024; a thread
    // per orientation over global memory beyond
    ChainTiers tiers;
    auto* d_tiers = (uint32_t*)take(S_TIERS, (6ull * n + 4) * 4);
    tiers.count = d_tiers;
    for (int x = 0; x < CHAIN_TIERS - 1; ++x) tiers.list[x] = d_tiers + 4 + (uint64_t)x * 2 * n;
    tiers.cap[0] = CHAIN_CAP1;
    tiers.cap[1] = CHAIN_CAP2;
    tiers.cap[2] = CHAIN_CAP3;
    if (ev) SA_CHECK(hipEventRecord(ev[3], st));
    k_zero_u32<<<1, 64, 0, st>>>(tiers.count, 4);
    k_chain_t1<16, CHAIN_CAP1><<<blocks(2ull * n, 4), 64, 0, st>>>(ix, cp, d_off, n, h.hits, h.hc, h.acap, d_aoff, b,
                                                                   o.namn, o.araw, log2tab, tiers, abort);
    SA_CHECK(hipGetLastError());
    k_chain_tn<64, CHAIN_CAP2><<<8192, 64, 0, st>>>(ix, cp, d_off, h.hits, h.hc, d_aoff, b, o.namn, o.araw, log2tab,
                                                    tiers, 0);
    SA_CHECK(hipGetLastError());
    k_chain_tn<64, CHAIN_CAP3><<<1024, 64, 0, st>>>(ix, cp, d_off, h.hits, h.hc, d_aoff, b, o.namn, o.araw, log2tab,
                                                    tiers, 1);
    SA_CHECK(hipGetLastError());
    k_chain<<<256, 256, 0, st>>>(ix, cp, d_off, n, h.hits, h.hc, h.acap, d_aoff, b, o.namn, o.araw, log2tab,
                                 same_diag, tiers);
    SA_CHECK(hipGetLastError());
    if (ev) SA_CHECK(hipEventRecord(ev[4], st));
    auto* d_tot = (uint64_t*)take(S_TOT, (uint64_t)(n + 1) * 8);
    o.outoff = (uint64_t*)take(S_OUTOFF, (uint64_t)(n + 1) * 8);
    k_pair_sum<<<blocks(n, 256), 256, 0, st>>>(o.namn, d_tot, n);
    SA_CHECK(hipGetLastError());
    o.n_nams = scan(d_tot, o.outoff, n);
    o.nams = nams_out(o.n_nams);
    if (ev) SA_CHECK(hipEventRecord(ev[5], st));
    k_finalize<<<blocks(n, 256), 256, 0, st>>>(d_seq, d_off, n, ref, k, o.namn, d_aoff, b.nams, o.outoff, o.nams);
    SA_CHECK(hipGetLastError());
    if (ev) SA_CHECK(hipEventRecord(ev[6], st));
}

// Seeding -> lookups -> chaining -> NAMs: hit_pass then chain_pass.
template <typename Take, typename Scan, typename NamsOut>
void nam_pass(const IndexView& ix, const RefView& ref, const float* log2tab, const float* same_diag,
              const uint8_t* d_seq, const uint64_t* d_off, uint32_t n, uint64_t n_bases, const MapParams& mp,
              const ChainParams& cp, hipStream_t st, Take take, Scan scan, NamsOut nams_out, hipEvent_t* ev,
              NamPass& o) {
    HitPass h;
    hit_pass(ix, d_seq, d_off, n, n_bases, mp, st, take, ev, h, o);
    chain_pass(ix, ref, log2tab, same_diag, d_seq, d_off, n, cp, st, take, scan, nams_out, ev, h, o);
}

// same-diagonal chaining scores (src/chainer.rs:63-74) in host f32 arithmetic
void same_diag_table(uint32_t k, const ChainParams& cp, float* sd) {
    float lg[1025];
    for (uint32_t d = 0; d < 1025; ++d) lg[d] = std::log2((float)d + 1.0f);
    for (uint32_t i = 0; i < 1024; ++i) sd[i] = chain_score(i, i, k, cp, lg);
}
}  // namespace

void Engine::chain_tables(const ChainParams& cp) {
    const uint32_t k = (uint32_t)p_.k;
    if (tab_k_ != k || std::memcmp(&tab_cp_, &cp, sizeof cp) != 0) {
        float sd[1024];
        same_diag_table(k, cp, sd);
        SA_CHECK(hipMemcpyAsync(d_same_diag_, sd, sizeof sd, hipMemcpyHostToDevice, stream_));
        sync();
        tab_k_ = k;
        tab_cp_ = cp;
    }
}

// Longest read the mapping passes take. Single-end mapping of 520-2000 bp reads
// did not finish on the device and no cause is known yet, so batches with longer
// reads are refused with an error instead of being run. Seeding, hit lookup and
// the extension kernels themselves take longer queries.
constexpr uint32_t MAP_MAX_READ_LEN = 512;

static void require_map_len(uint32_t max_len) {
    if (max_len > MAP_MAX_READ_LEN)
        throw std::runtime_error("reads longer than 512 bp are not supported by the mapping path (longest: " +
                                 std::to_string(max_len) + " bp)");
}

void Engine::find_nams(const MapParams& mp, const ChainParams& cp, NamBatch& out) {
    SA_CHECK(hipSetDevice(device_));
    require_map_len(max_len_);
    const uint32_t n = n_reads_;
    out.nams = HostView<NamRec>{};
    out.nam_off = HostView<uint64_t>{(uint64_t*)host_stage(0, 8), 1};
    out.nam_off[0] = 0;
    out.info.clear();
    if (n == 0) return;
    chain_tables(cp);
    hipEvent_t ev[7];
    for (auto& e : ev) e = event();
    IndexView ix{d_entries_, d_buckets_, p_};
    RefView ref{d_words_, d_word_off_, d_len_, n_contigs_};
    NamPass o;
    nam_pass(
        ix, ref, d_log2_, d_same_diag_, d_seq_, (const uint64_t*)buf_[S_OFF], n, n_bases_, mp, cp, stream_,
        [&](int slot, uint64_t bytes) { return grow(slot, bytes); },
        [&](const uint64_t* in, uint64_t* outp, uint64_t m) {
            uint64_t t = 0;
            scan_u64(in, outp, m, &t);
            // the first scan totals the anchor capacities
            if (in == (const uint64_t*)buf_[S_ACAP] && n_bases_)
                anchors_per_base_ = std::max(anchors_per_base_, (double)t / (double)n_bases_);
            return t;
        },
        [&](uint64_t N) { return (NamRec*)grow(S_NAMS, (N + 1) * sizeof(NamRec)); }, ev, o);
    cur_chain_ = (const uint32_t*)buf_[S_CHAIN];
    finish_nams(n, o, ev, out);
}

void Engine::find_hits(const MapParams& mp, std::vector<Hit>& hits, std::vector<HitCounts>& hc) {
    SA_CHECK(hipSetDevice(device_));
    const uint32_t n = n_reads_;
    hits.clear();
    hc.assign(2ull * n, HitCounts{});
    if (n == 0) return;
    IndexView ix{d_entries_, d_buckets_, p_};
    HitPass h;
    NamPass o;
    hit_pass(ix, d_seq_, (const uint64_t*)buf_[S_OFF], n, n_bases_, mp, stream_,
             [&](int slot, uint64_t bytes) { return grow(slot, bytes); }, nullptr, h, o);
    hits.resize(2 * n_bases_);
    if (n_bases_) SA_CHECK(hipMemcpyAsync(hits.data(), h.hits, 2 * n_bases_ * sizeof(Hit), hipMemcpyDeviceToHost, stream_));
    SA_CHECK(hipMemcpyAsync(hc.data(), h.hc, 2ull * n * sizeof(HitCounts), hipMemcpyDeviceToHost, stream_));
    sync();
}

void Engine::chain_hits(const Hit* hits, const HitCounts* hc, const ChainParams& cp, NamBatch& out) {
    SA_CHECK(hipSetDevice(device_));
    const uint32_t n = n_reads_;
    out.nams = HostView<NamRec>{};
    out.nam_off = HostView<uint64_t>{(uint64_t*)host_stage(0, 8), 1};
    out.nam_off[0] = 0;
    out.info.clear();
    if (n == 0) return;
    chain_tables(cp);
    IndexView ix{d_entries_, d_buckets_, p_};
    RefView ref{d_words_, d_word_off_, d_len_, n_contigs_};
    const auto* d_off = (const uint64_t*)buf_[S_OFF];
    HitPass h;
    h.hits = (Hit*)grow(S_HITS, (2 * n_bases_ + 1) * sizeof(Hit));
    h.hc = (HitCounts*)grow(S_HC, (2ull * n + 1) * sizeof(HitCounts));
    h.acap = (uint64_t*)grow(S_ACAP, (2ull * n + 1) * 8);
    if (n_bases_) SA_CHECK(hipMemcpyAsync(h.hits, hits, 2 * n_bases_ * sizeof(Hit), hipMemcpyHostToDevice, stream_));
    SA_CHECK(hipMemcpyAsync(h.hc, hc, 2ull * n * sizeof(HitCounts), hipMemcpyHostToDevice, stream_));
    k_hit_runs<<<blocks(2ull * n, 256), 256, 0, stream_>>>(ix, d_off, n, h.hits, h.hc, h.acap);
    SA_CHECK(hipGetLastError());
    hipEvent_t ev[7];
    for (auto& e : ev) e = event();
    SA_CHECK(hipEventRecord(ev[0], stream_));
    SA_CHECK(hipEventRecord(ev[1], stream_));
    SA_CHECK(hipEventRecord(ev[2], stream_));
    NamPass o;
    o.hc = h.hc;
    chain_pass(
        ix, ref, d_log2_, d_same_diag_, d_seq_, d_off, n, cp, stream_,
        [&](int slot, uint64_t bytes) { return grow(slot, bytes); },
        [&](const uint64_t* in, uint64_t* outp, uint64_t m) {
            uint64_t t = 0;
            scan_u64(in, outp, m, &t);
            return t;
        },
        [&](uint64_t N) { return (NamRec*)grow(S_NAMS, (N + 1) * sizeof(NamRec)); }, ev, h, o);
    cur_chain_ = (const uint32_t*)buf_[S_CHAIN];
    finish_nams(n, o, ev, out);
}

void Engine::finish_nams(uint32_t n, const NamPass& o, hipEvent_t* ev, NamBatch& out) {
    const uint64_t N = o.n_nams;
    // everything the host reads back goes to pinned staging memory in one block
    const uint64_t n2 = 2ull * n;
    const uint64_t o_off = ((uint64_t)N * sizeof(NamRec) + 15) & ~15ull;
    const uint64_t o_hc = o_off + (((uint64_t)(n + 1) * 8 + 15) & ~15ull);
    const uint64_t o_u32 = o_hc + ((n2 * sizeof(HitCounts) + 15) & ~15ull);
    const uint64_t u32_bytes = (n2 * 4 + 15) & ~15ull;
    auto* hs = (uint8_t*)host_stage(0, o_u32 + 5 * u32_bytes);
    out.nams = HostView<NamRec>{(NamRec*)hs, N};
    out.nam_off = HostView<uint64_t>{(uint64_t*)(hs + o_off), (size_t)n + 1};
    const HitCounts* hc = (const HitCounts*)(hs + o_hc);
    uint32_t* qrn = (uint32_t*)(hs + o_u32);
    uint32_t* namn = (uint32_t*)(hs + o_u32 + u32_bytes);
    uint32_t* araw = (uint32_t*)(hs + o_u32 + 2 * u32_bytes);
    uint32_t* cbytes = (uint32_t*)(hs + o_u32 + 3 * u32_bytes);
    uint32_t* hbytes = (uint32_t*)(hs + o_u32 + 4 * u32_bytes);
    if (N) SA_CHECK(hipMemcpyAsync(out.nams.data(), o.nams, N * sizeof(NamRec), hipMemcpyDeviceToHost, stream_));
    SA_CHECK(hipMemcpyAsync(out.nam_off.data(), o.outoff, (uint64_t)n * 8, hipMemcpyDeviceToHost, stream_));
    SA_CHECK(hipMemcpyAsync(cbytes, o.cbytes, n2 * 4, hipMemcpyDeviceToHost, stream_));
    SA_CHECK(hipMemcpyAsync((void*)hc, o.hc, n2 * sizeof(HitCounts), hipMemcpyDeviceToHost, stream_));
    if (o.qrn) SA_CHECK(hipMemcpyAsync(qrn, o.qrn, n2 * 4, hipMemcpyDeviceToHost, stream_));
    else std::memset(qrn, 0, n2 * 4);
    SA_CHECK(hipMemcpyAsync(namn, o.namn, n2 * 4, hipMemcpyDeviceToHost, stream_));
    SA_CHECK(hipMemcpyAsync(araw, o.araw, n2 * 4, hipMemcpyDeviceToHost, stream_));
    if (o.hbytes) SA_CHECK(hipMemcpyAsync(hbytes, o.hbytes, n2 * 4, hipMemcpyDeviceToHost, stream_));
    else std::memset(hbytes, 0, n2 * 4);
    sync();
    out.nam_off[n] = N;
    out.info.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        ReadNamInfo& r = out.info[i];
        r.hits[0] = hc[2 * i];
        r.hits[1] = hc[2 * i + 1];
        r.n_randstrobes = qrn[2 * i] + qrn[2 * i + 1];
        r.n_anchors = araw[2 * i] + araw[2 * i + 1];
        r.n_nams = namn[2 * i] + namn[2 * i + 1];
        r.both_orientations = !better_orientation(hc[2 * i], hc[2 * i + 1]) && !better_orientation(hc[2 * i + 1], hc[2 * i]);
    }
    if (!ev) return;  // the fused pass times itself
    float t;
    // algorithmic bytes: read bases in, query randstrobes out (seed); randstrobes,
    // bucket_starts words and index entries in, hits out (lookups); hits/entries in,
    // NAMs out (chaining)
    uint64_t seed_bytes = n_bases_, chain_bytes = 0, hit_bytes = 0;
    for (uint64_t x = 0; x < 2ull * n; ++x) {
        seed_bytes += sizeof(QRand) * (uint64_t)qrn[x];
        chain_bytes += cbytes[x];
        hit_bytes += hbytes[x];
    }
    if (o.qrn) {
        SA_CHECK(hipEventElapsedTime(&t, ev[0], ev[1])); times_.seed = t; add_kernel("k_seed", t, n, seed_bytes);
        SA_CHECK(hipEventElapsedTime(&t, ev[1], ev[2])); times_.hits = t; add_kernel("k_hits", t, 2ull * n, hit_bytes);
    } else {
        times_.seed = times_.hits = 0;
    }
    SA_CHECK(hipEventElapsedTime(&t, ev[3], ev[4])); times_.chain = t; add_kernel("k_chain", t, 2ull * n, chain_bytes);
    SA_CHECK(hipEventElapsedTime(&t, ev[5], ev[6])); times_.finalize = t; add_kernel("k_finalize", t, n);
    total_ms_ += times_.seed + times_.hits + times_.chain + times_.finalize;
    for (int x = 0; x < 7; ++x) release_event(ev[x]);
}

namespace {
// n_bases (= off[n], the extent the per-base buffers cover) and the longest read
// of a batch whose offsets are in device memory: one block, written unconditionally.
__global__ void __launch_bounds__(256) k_batch_shape(const uint64_t* __restrict__ off, uint32_t n,
                                                     uint64_t* __restrict__ out) {
    __shared__ uint64_t m[256];
    uint64_t mx = 0;
    for (uint32_t i = threadIdx.x; i < n; i += 256) {
        const uint64_t L = off[i + 1] - off[i];
        mx = L > mx ? L : mx;
    }
    m[threadIdx.x] = mx;
    __syncthreads();
    for (uint32_t w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w && m[threadIdx.x + w] > m[threadIdx.x]) m[threadIdx.x] = m[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = off[n];
        out[1] = m[0];
    }
}
__global__ void k_set_u64(uint64_t* __restrict__ p, uint64_t v) { *p = v; }

// Raised when a batch does not fit the scratch array or the caller's NAM buffer.
struct TooLarge : std::invalid_argument {
    using std::invalid_argument::invalid_argument;
};
}  // namespace

void Engine::find_nams_noalloc(const uint8_t* d_seq, const uint64_t* d_off, uint32_t n, const MapParams& mp,
                               const ChainParams& cp, NamRec* d_nams, uint64_t nam_cap, uint64_t* d_nam_off,
                               uint64_t* n_nams, hipStream_t st, uint8_t* scratch, uint64_t scratch_bytes) {
    SA_CHECK(hipSetDevice(device_));
    *n_nams = 0;
    if (n == 0) {
        k_set_u64<<<1, 1, 0, st>>>(d_nam_off, 0);
        SA_CHECK(hipGetLastError());
        SA_CHECK(hipStreamSynchronize(st));
        return;
    }
    uint64_t used = 0;
    auto take = [&](uint64_t bytes) {
        const uint64_t at = used;
        used = (used + (bytes ? bytes : 16) + 255) & ~(uint64_t)255;
        if (used > scratch_bytes)
            throw TooLarge("batch needs more than the " + std::to_string(scratch_bytes >> 20) +
                           " MiB scratch of the no-allocation entry point (at least " + std::to_string(used) +
                           " bytes); split it into smaller batches");
        return (void*)(scratch + at);
    };
    // batch shape from the device offsets
    auto* shape = (uint64_t*)take(16);
    k_batch_shape<<<1, 256, 0, st>>>(d_off, n, shape);
    SA_CHECK(hipGetLastError());
    uint64_t h_shape[2] = {0, 0};
    SA_CHECK(hipMemcpyAsync(h_shape, shape, 16, hipMemcpyDeviceToHost, st));
    SA_CHECK(hipStreamSynchronize(st));
    const uint64_t n_bases = h_shape[0];
    if (h_shape[1] >= 65536) throw std::invalid_argument("reads longer than 65535 bp are not supported");
    const uint32_t k = (uint32_t)p_.k;
    if (tab_k_ != k || std::memcmp(&tab_cp_, &cp, sizeof cp) != 0) {
        float sd[1024];
        same_diag_table(k, cp, sd);
        SA_CHECK(hipMemcpyAsync(d_same_diag_, sd, sizeof sd, hipMemcpyHostToDevice, st));
        SA_CHECK(hipStreamSynchronize(st));
        tab_k_ = k;
        tab_cp_ = cp;
    }
    IndexView ix{d_entries_, d_buckets_, p_};
    RefView ref{d_words_, d_word_off_, d_len_, n_contigs_};
    NamPass o;
    nam_pass(
        ix, ref, d_log2_, d_same_diag_, d_seq, d_off, n, n_bases, mp, cp, st,
        [&](int slot, uint64_t bytes) { return slot == S_OUTOFF ? (void*)d_nam_off : take(bytes); },
        [&](const uint64_t* in, uint64_t* outp, uint64_t m) {
            size_t tmp = 0;
            SA_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, in, outp, (int)m, st));
            void* t = take(tmp);
            SA_CHECK(hipcub::DeviceScan::ExclusiveSum(t, tmp, in, outp, (int)m, st));
            uint64_t last[2];
            SA_CHECK(hipMemcpyAsync(&last[0], in + m - 1, 8, hipMemcpyDeviceToHost, st));
            SA_CHECK(hipMemcpyAsync(&last[1], outp + m - 1, 8, hipMemcpyDeviceToHost, st));
            SA_CHECK(hipStreamSynchronize(st));
            return last[0] + last[1];
        },
        [&](uint64_t N) {
            if (N > nam_cap)
                throw TooLarge("batch has " + std::to_string(N) + " NAMs, the output holds " + std::to_string(nam_cap));
            return d_nams;
        },
        nullptr, o);
    k_set_u64<<<1, 1, 0, st>>>(d_nam_off + n, o.n_nams);
    SA_CHECK(hipGetLastError());
    SA_CHECK(hipStreamSynchronize(st));
    *n_nams = o.n_nams;
}

namespace {
// Launches k_ssw_pass<G, BYTE, seg> for a run-time stripe count 1 <= seg <= MAX.
// dev_grid > 0: the slot counts are read on the device (k_ssw_pass with
// dev_counts) by a grid of dev_grid blocks.
template <int G, bool BYTE, int S, int MAX>
void launch_ssw_pass(uint32_t seg, const SswLaunchLists& ll, hipStream_t st, const uint8_t* seq, const uint64_t* off,
                     RefView ref, const Scores& sc, const ExtRequest* req, const ExtJob* jobs, uint32_t rev,
                     SswEnd* fwd, SswEnds* ends, uint32_t* lists, uint32_t* counts, uint32_t cap,
                     unsigned long long* cells, uint32_t dev_grid = 0) {
    if constexpr (S > MAX) {
        throw std::logic_error("SSW stripe count out of range");
    } else if (seg != S) {
        launch_ssw_pass<G, BYTE, S + 1, MAX>(seg, ll, st, seq, off, ref, sc, req, jobs, rev, fwd, ends, lists, counts,
                                             cap, cells, dev_grid);
    } else {
        const uint64_t count = ll.start[SSW_WCLASSES];
        const uint32_t grid = dev_grid ? dev_grid : blocks(count * G, 256);
        k_ssw_pass<G, BYTE, S><<<grid, 256, 0, st>>>(seq, off, ref, sc, req, jobs, ll, rev, fwd, ends, lists, counts,
                                                     cap, dev_grid ? 1u : 0u, cells);
    }
}
}  // namespace

void Engine::extend(const ExtRequest* req_in, uint32_t n, const Scores& sc, std::vector<ExtResult>& res,
                    std::vector<uint32_t>& cigars, std::vector<uint64_t>& cigar_off) {
    SA_CHECK(hipSetDevice(device_));
    res.assign(n, ExtResult{});
    cigars.clear();
    cigar_off.assign(n + 1, 0);
    if (n == 0) return;
    // CIGAR slots: at most 2*L+1 ops for a read of length L
    std::vector<ExtRequest> req(req_in, req_in + n);
    uint64_t slots = 0, prep_bytes = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const uint64_t Lq = h_off_[req[i].read_id + 1] - h_off_[req[i].read_id];
        // the aligner's entry points on explicit windows may need one op per query
        // and window position
        req[i].cigar_cap = req[i].kind >= EXT_HAMMING
                               ? (uint32_t)Lq + (req[i].win_end - req[i].win_start) + 8
                               : 2 * (uint32_t)Lq + 8;
        req[i].cigar_off = slots;
        slots += req[i].cigar_cap;
        // algorithmic bytes of the prep kernel: request in, result out, the query and
        // the 2-bit reference it is compared with (read window, or the rescue window)
        const uint64_t wl = req[i].kind == EXT_SEED ? Lq : (uint64_t)(req[i].win_end - req[i].win_start);
        prep_bytes += sizeof(ExtRequest) + sizeof(ExtResult) + Lq + (wl + 3) / 4;
    }
    auto* d_req = (ExtRequest*)grow(S_REQ, (uint64_t)n * sizeof(ExtRequest));
    auto* d_res = (ExtResult*)grow(S_RES, (uint64_t)n * sizeof(ExtResult));
    auto* d_slots = (uint32_t*)grow(S_CIGSLOT, slots * 4);
    auto* d_jobs = (ExtJob*)grow(S_JOBS, (uint64_t)n * sizeof(ExtJob));
    auto* d_ljobs = (ExtJob*)grow(S_LJOBS, (uint64_t)n * sizeof(ExtJob));
    auto* d_pjobs = (ExtJob*)grow(S_PJOBS, (uint64_t)n * sizeof(ExtJob));
    constexpr uint32_t NCNT = EXT_NCNT;
    auto* d_cnt = (uint32_t*)grow(S_XCNT, NCNT * 4);
    uint32_t* d_bcnt = d_cnt + 4;
    // SSW bucket lists sized for every request (the job count is known only on the device)
    const uint32_t cap = n;
    auto* d_lists = (uint32_t*)grow(S_LISTS, (uint64_t)SSW_BUCKETS * cap * 4);
    auto* d_ends = (SswEnds*)grow(S_ENDS, (uint64_t)n * sizeof(SswEnds));
    auto* d_fwd = (SswEnd*)grow(S_FWD, (uint64_t)n * sizeof(SswEnd));
    SA_CHECK(hipMemcpyAsync(d_req, req.data(), (uint64_t)n * sizeof(ExtRequest), hipMemcpyHostToDevice, stream_));
    SA_CHECK(hipMemsetAsync(d_cnt, 0, NCNT * 4, stream_));
    RefView ref{d_words_, d_word_off_, d_len_, n_contigs_};
    const auto* d_seq = d_seq_;
    const auto* d_off = cur_off_;
    const uint32_t* d_chain = cur_chain_;
    const uint32_t k = (uint32_t)p_.k;
    const uint64_t L = max_len_;
    // kernel times are read once the work is done (no wait per kernel)
    struct Timed { const char* name; uint64_t items, bytes, cells; hipEvent_t a, b; };
    std::vector<Timed> tl;
    auto timed = [&](const char* name, uint64_t items, auto&& launch, uint64_t bytes = 0, uint64_t cells = 0) {
        Timed x{name, items, bytes, cells, event(), event()};
        SA_CHECK(hipEventRecord(x.a, stream_));
        launch();
        SA_CHECK(hipGetLastError());
        SA_CHECK(hipEventRecord(x.b, stream_));
        tl.push_back(x);
    };
    uint32_t cnt[NCNT];
    auto read_counts = [&] {
        SA_CHECK(hipMemcpyAsync(cnt, d_cnt, sizeof cnt, hipMemcpyDeviceToHost, stream_));
        sync();
    };
    // 1. Hamming fast path, rescue prefilter, job lists; SSW jobs into forward-byte buckets
    timed("k_ext_prep", n, [&] {
        k_ext_prep<<<blocks((uint64_t)n * 64, 256), 256, 0, stream_>>>(d_seq, d_off, ref, sc, k, d_req, n, nullptr,
                                                                      d_res, d_slots, d_jobs, d_ljobs, d_pjobs, d_cnt);
    }, prep_bytes);
    auto* d_cells = (unsigned long long*)(d_cnt + EXT_CELLS);
    k_ssw_bucket<<<blocks(n, 256), 256, 0, stream_>>>(d_off, d_req, d_jobs, d_cnt, d_lists, d_bcnt, cap);
    SA_CHECK(hipGetLastError());
    read_counts();
    const uint32_t nssw = cnt[0], nlong = cnt[1], npw = cnt[2];
    const uint32_t* bc = cnt + 4;
    // thread-per-job kernels run in waves sized to the scratch budget
    const uint64_t budget = 16ull << 30;
    auto waves = [&](uint32_t njobs, uint64_t per, auto&& launch) {
        uint64_t wave = std::max<uint64_t>(1, std::min<uint64_t>(njobs, budget / per));
        // whole 64-job blocks: a kernel may interleave the scratch of a wave's jobs
        auto* d_scr = (uint8_t*)grow(S_SCRATCH, (wave + 63) / 64 * 64 * per);
        for (uint64_t first = 0; first < njobs; first += wave)
            launch((uint32_t)std::min<uint64_t>(wave, njobs - first), (uint32_t)first, d_scr);
    };
    // 2. SSW: score passes per lane group, bucketed by stripe count, then
    //    traceback per thread
    if (nssw) {
        // launches every non-empty bucket of one (mode, pass) family
        auto run = [&](bool word, bool rev) {
            const uint32_t nseg = word ? SSW_WORD_SEGS : SSW_BYTE_SEGS;
            uint32_t below = nseg;  // reverse passes: stripe counts below this one are still to run
            for (uint32_t step = 1; step <= nseg; ++step) {
                SswLaunchLists ll{};
                uint32_t at = 0, seg = step;
                if (!rev) {
                    // one launch per stripe count: its window classes, longest first
                    for (uint32_t x = 0; x < SSW_WCLASSES; ++x) {
                        const uint32_t b = ssw_bucket(word, false, seg, SSW_WCLASSES - 1 - x);
                        ll.list[x] = d_lists + (uint64_t)b * cap;
                        ll.start[x] = at;
                        at += bc[b];
                    }
                } else {
                    // reverse jobs have no window classes: the jobs of the largest
                    // stripe count left and of up to SSW_WCLASSES - 1 smaller ones
                    // share a launch of that stripe count, so that the few long
                    // rescue jobs of small counts do not run one launch each. A
                    // pass with more stripes than a query needs finds the same
                    // ends only while gap_open > gap_extend (with O <= E the lazy-F
                    // loop's early exit depends on the striping), so other scores
                    // run one launch per stripe count.
                    while (below >= 1 && bc[ssw_bucket(word, true, below, 0)] == 0) --below;
                    if (below == 0) break;
                    seg = below;
                    const uint32_t merge = sc.gap_open > sc.gap_extend ? SSW_WCLASSES : 1;
                    for (uint32_t x = 0; x < SSW_WCLASSES; ++x) {
                        const bool used = below >= 1 && x < merge;
                        const uint32_t b = ssw_bucket(word, true, used ? below : seg, 0);
                        ll.list[x] = d_lists + (uint64_t)b * cap;
                        ll.start[x] = at;
                        at += used ? bc[b] : 0;
                        if (used) --below;
                    }
                }
                ll.start[SSW_WCLASSES] = at;
                if (!at) continue;
                if (word)
                    launch_ssw_pass<8, false, 1, SSW_WORD_SEGS>(seg, ll, stream_, d_seq, d_off, ref, sc, d_req, d_jobs, rev,
                                                     d_fwd, d_ends, d_lists, d_bcnt, cap, d_cells);
                else
                    launch_ssw_pass<16, true, 1, SSW_BYTE_SEGS>(seg, ll, stream_, d_seq, d_off, ref, sc, d_req, d_jobs, rev,
                                                     d_fwd, d_ends, d_lists, d_bcnt, cap, d_cells);
                SA_CHECK(hipGetLastError());
            }
        };
        timed("k_ssw_byte_fwd", nssw, [&] { run(false, false); });
        read_counts();
        uint64_t nrb = 0, nfw = 0;
        for (uint32_t x = 0; x < SSW_BYTE_SEGS; ++x) nrb += bc[ssw_bucket(false, true, x + 1, 0)];
        for (uint32_t x = 0; x < SSW_WORD_SEGS; ++x)
            for (uint32_t wc = 0; wc < SSW_WCLASSES; ++wc) nfw += bc[ssw_bucket(true, false, x + 1, wc)];
        if (nrb) timed("k_ssw_byte_rev", nrb, [&] { run(false, true); });
        if (nfw) {
            timed("k_ssw_word_fwd", nfw, [&] { run(true, false); });
            read_counts();
            timed("k_ssw_word_rev", nfw, [&] { run(true, true); });
        }
        // banded traceback: band up to 64 without a retry (words: three band rows,
        // 17 words of packed direction cells per query row, the traceback ops)
        uint64_t per = 4ull * (3 * (2 * 64 + 4) + 17ull * L + 2 * (2ull * L + 64) + 64);
        per = (per + 255) & ~255ull;
        // jobs grouped by the band width their traceback starts with
        auto* d_fin = (uint32_t*)grow(S_FINLISTS, (uint64_t)FIN_CLASSES * nssw * 4);
        uint32_t* d_fcnt = d_cnt + EXT_FIN;
        k_fin_bucket<<<blocks(nssw, 256), 256, 0, stream_>>>(d_ends, nssw, nullptr, d_fin, d_fcnt, nssw);
        SA_CHECK(hipGetLastError());
        timed("k_ssw_finish", nssw, [&] {
            waves(nssw, per, [&](uint32_t cnt_, uint32_t first, uint8_t* scr) {
                k_ssw_finish<<<blocks(cnt_, 128), 128, 0, stream_>>>(d_seq, d_off, ref, sc, d_req, d_jobs, d_ends, cnt_,
                                                                    nullptr, d_res, d_slots, scr, per, first, d_fin,
                                                                    d_fcnt, nssw);
            });
        });
    }
    // 3. SSW of queries too long for a lane group, one thread each
    if (nlong) {
        uint64_t per = 16ull * ((L + 15) / 16 * 16) + 256 + 65536 + 8 * (2 * L + 64) + 129ull * L * 3;
        per = (per + 255) & ~255ull;
        timed("k_ext_long", nlong, [&] {
            waves(nlong, per, [&](uint32_t cnt_, uint32_t first, uint8_t* scr) {
                k_ext_long<<<blocks(cnt_, 128), 128, 0, stream_>>>(d_seq, d_off, ref, sc, d_req, d_ljobs, cnt_,
                                                                  nullptr, d_res, d_slots, scr, per, first);
            });
        });
    }
    // 4. piecewise extension, a group of PW_LANES lanes per job
    if (npw) {
        // per job: chain anchors, traceback ops (the DP itself uses the blocks below)
        uint64_t per = 16 * L + 256 + 16 * (2 * L + L / 10 + 16) + 16 * L + 4096;
        per = (per + 255) & ~255ull;
        // DP row buffer and direction bytes of the end extensions and gaps, one
        // block of each per job (a DP with more rows or columns falls back to one
        // thread with a large scratch)
        const uint32_t rcap = (uint32_t)(L + L / 10 + 3), drows = (uint32_t)(L + 2);
        const uint64_t pw_per = per + (uint64_t)rcap * sizeof(DpCol) + (uint64_t)rcap * drows;
        const uint64_t slice = std::max<uint64_t>(1, std::min<uint64_t>(npw, budget / pw_per));
        auto* scr = (uint8_t*)grow(S_SCRATCH, slice * per);
        auto* rows = (DpCol*)grow(S_PWROWS, slice * rcap * sizeof(DpCol));
        auto* dirs = (uint8_t*)grow(S_PWDIRS, slice * rcap * drows);
        timed("k_ext_pw", npw, [&] {
            for (uint64_t first = 0; first < npw; first += slice) {
                const uint32_t cnt_ = (uint32_t)std::min<uint64_t>(slice, npw - first);
                k_ext_pw_grp<PW_LANES><<<blocks((uint64_t)cnt_ * PW_LANES, 128), 128, 0, stream_>>>(
                    d_seq, d_off, ref, sc, k, d_req, d_pjobs, cnt_, nullptr, d_res, d_slots, d_chain, scr, per,
                    (uint32_t)first, rows, dirs, rcap, drows);
            }
        });
    }
    // 5. CIGAR offsets (found results only) and the results, in one wait
    auto* d_ncig = (uint64_t*)grow(S_NCIG, (uint64_t)n * 8);
    auto* d_coff = (uint64_t*)grow(S_CIGOFF, (uint64_t)(n + 1) * 8);
    k_ncigar<<<blocks(n, 256), 256, 0, stream_>>>(d_res, d_ncig, n);
    {
        size_t tmp = 0;
        SA_CHECK(hipcub::DeviceScan::InclusiveSum(nullptr, tmp, d_ncig, d_coff + 1, (int)n, stream_));
        void* t = grow(S_SCANTMP, tmp);
        SA_CHECK(hipcub::DeviceScan::InclusiveSum(t, tmp, d_ncig, d_coff + 1, (int)n, stream_));
        SA_CHECK(hipMemsetAsync(d_coff, 0, 8, stream_));
    }
    SA_CHECK(hipMemcpyAsync(res.data(), d_res, (uint64_t)n * sizeof(ExtResult), hipMemcpyDeviceToHost, stream_));
    SA_CHECK(hipMemcpyAsync(cigar_off.data(), d_coff, (uint64_t)(n + 1) * 8, hipMemcpyDeviceToHost, stream_));
    SA_CHECK(hipMemcpyAsync(cnt, d_cnt, sizeof cnt, hipMemcpyDeviceToHost, stream_));
    sync();
    // the DP cells each kind of score pass computed
    {
        uint64_t cells[4];
        std::memcpy(cells, cnt + EXT_CELLS, sizeof cells);
        static const char* const kind[4] = {"k_ssw_byte_fwd", "k_ssw_byte_rev", "k_ssw_word_fwd", "k_ssw_word_rev"};
        for (Timed& x : tl)
            for (int c = 0; c < 4; ++c)
                if (std::strcmp(x.name, kind[c]) == 0) x.cells = cells[c];
    }
    float ext_ms = 0;
    for (Timed& x : tl) {
        float t = 0;
        SA_CHECK(hipEventElapsedTime(&t, x.a, x.b));
        add_kernel(x.name, t, x.items, x.bytes, x.cells);
        total_ms_ += t;
        ext_ms += t;
        release_event(x.a);
        release_event(x.b);
    }
    times_.extend = ext_ms;
    // requests that ran out of scratch are re-run one at a time with a large slot
    std::vector<uint32_t> big;
    for (uint32_t i = 0; i < n; ++i)
        if (res[i].status == 2) big.push_back(i);
    if (!big.empty()) {
        const uint64_t bigs = 256ull << 20;
        auto* d_big = (uint8_t*)grow(S_SCRATCH, bigs);
        for (uint32_t i : big) {
            k_extend<<<1, 1, 0, stream_>>>(d_seq, d_off, ref, sc, k, d_req, 1, d_res, d_slots, d_chain, d_big, bigs, i);
            SA_CHECK(hipGetLastError());
        }
        k_ncigar<<<blocks(n, 256), 256, 0, stream_>>>(d_res, d_ncig, n);
        size_t tmp = 0;
        SA_CHECK(hipcub::DeviceScan::InclusiveSum(nullptr, tmp, d_ncig, d_coff + 1, (int)n, stream_));
        void* t = grow(S_SCANTMP, tmp);
        SA_CHECK(hipcub::DeviceScan::InclusiveSum(t, tmp, d_ncig, d_coff + 1, (int)n, stream_));
        SA_CHECK(hipMemcpyAsync(res.data(), d_res, (uint64_t)n * sizeof(ExtResult), hipMemcpyDeviceToHost, stream_));
        SA_CHECK(hipMemcpyAsync(cigar_off.data(), d_coff, (uint64_t)(n + 1) * 8, hipMemcpyDeviceToHost, stream_));
        sync();
    }
    for (uint32_t i = 0; i < n; ++i)
        if (res[i].status != 0) throw std::runtime_error("extension request failed (status)");
    // 6. the CIGARs, compacted on the device
    const uint64_t total = cigar_off[n];
    auto* d_cout = (uint32_t*)grow(S_CIGOUT, (total + 1) * 4);
    k_compact_cigars<<<blocks(n, 256), 256, 0, stream_>>>(d_req, d_res, d_slots, d_coff, d_cout, n);
    SA_CHECK(hipGetLastError());
    cigars.resize(total);
    if (total) SA_CHECK(hipMemcpyAsync(cigars.data(), d_cout, total * 4, hipMemcpyDeviceToHost, stream_));
    sync();
}


// Every launch of one fused pass on the engine stream, from the reads in
// F_SEQ / F_OFF (n_cap reads, the ones past the batch empty); either run
// directly or recorded. Nothing in it waits for the device or reads a count
// on the host.
void Engine::fused_sequence(Fused& f) {
    auto take = [&](int slot, uint64_t bytes) { return grow(slot, bytes); };
    hipStream_t st = stream_;
    const uint32_t n = f.n_cap;
    const auto* d_seq = (const uint8_t*)buf_[F_SEQ];
    const auto* d_off = (const uint64_t*)buf_[F_OFF];
    auto* ctl = (uint64_t*)take(F_CTL, C_COUNT * 8);
    auto* flag = (uint32_t*)(ctl + C_FLAG);
    auto* nreq = (uint32_t*)(ctl + C_NREQ);
    k_zero_u32<<<1, 64, 0, st>>>((uint32_t*)ctl, 2 * C_COUNT);
    SA_CHECK(hipGetLastError());
    // exclusive scan of m values, total to *total (and *fl set when it exceeds cap)
    auto dscan = [&](const uint64_t* in, uint64_t* out, uint64_t m, uint64_t* total, uint64_t capv, uint32_t* fl,
                     int sum_slot) {
        const uint32_t nb = (uint32_t)std::max<uint64_t>(1, (m + SCAN_PER - 1) / SCAN_PER);
        auto* sums = (uint64_t*)take(sum_slot, (uint64_t)nb * 8);
        k_scan_sums<<<nb, 256, 0, st>>>(in, m, sums);
        k_scan_top<<<1, 256, 0, st>>>(sums, nb, total, capv, fl);
        k_scan_apply<<<nb, 256, 0, st>>>(in, m, sums, out);
        SA_CHECK(hipGetLastError());
    };
    IndexView ix{d_entries_, d_buckets_, p_};
    RefView ref{d_words_, d_word_off_, d_len_, n_contigs_};
    NamPass o;
    HitPass h;
    hit_pass(ix, d_seq, d_off, n, f.bases_cap, f.mp, st, take, nullptr, h, o);
    int scans = 0;
    chain_pass(
        ix, ref, d_log2_, d_same_diag_, d_seq, d_off, n, f.cp, st, take,
        [&](const uint64_t* in, uint64_t* outp, uint64_t m) -> uint64_t {
            // anchors: checked against their fixed capacity; NAMs fit the same
            if (scans++ == 0) dscan(in, outp, m, ctl + C_ANCH, f.a_cap, flag, F_SUM0);
            else dscan(in, outp, m, ctl + C_NNAMS, 0, nullptr, F_SUM1);
            return f.a_cap;
        },
        [&](uint64_t N) { return (NamRec*)take(S_NAMS, (N + 1) * sizeof(NamRec)); }, nullptr, h, o, flag);
    const auto* nam_cnt = (const uint64_t*)buf_[S_TOT];
    // speculative requests
    auto* rq_n = (uint64_t*)take(F_SPECN, (uint64_t)(n + 1) * 8);
    auto* rq_s = (uint64_t*)take(F_SPECS, (uint64_t)(n + 1) * 8);
    auto* roff = (uint64_t*)take(F_ROFF, (uint64_t)(n + 1) * 8);
    auto* soff = (uint64_t*)take(F_SOFF, (uint64_t)(n + 1) * 8);
    k_spec_count<<<blocks(n, 256), 256, 0, st>>>(o.nams, o.outoff, nam_cnt, d_off, n, f.sp, rq_n, rq_s);
    SA_CHECK(hipGetLastError());
    dscan(rq_n, roff, n, ctl + C_WANT, 0, nullptr, F_SUM2);
    dscan(rq_s, soff, n, ctl + C_WANTS, 0, nullptr, F_SUM3);
    const uint32_t cap = (uint32_t)f.req_cap;
    auto* d_req = (ExtRequest*)take(S_REQ, (uint64_t)cap * sizeof(ExtRequest));
    k_spec_write<<<blocks(n, 256), 256, 0, st>>>(o.nams, o.outoff, nam_cnt, d_off, n, f.sp, roff, soff, f.req_cap,
                                                  f.slot_cap, d_req, nreq);
    SA_CHECK(hipGetLastError());
    // extension of the requests [0, *nreq): Engine::extend with the job counts
    // read by the kernels
    const Scores& sc = f.sc;
    const uint32_t k = (uint32_t)p_.k;
    auto* d_res = (ExtResult*)take(S_RES, (uint64_t)cap * sizeof(ExtResult));
    auto* d_slots = (uint32_t*)take(S_CIGSLOT, f.slot_cap * 4 + 4);
    auto* d_jobs = (ExtJob*)take(S_JOBS, (uint64_t)cap * sizeof(ExtJob));
    auto* d_ljobs = (ExtJob*)take(S_LJOBS, (uint64_t)cap * sizeof(ExtJob));
    auto* d_pjobs = (ExtJob*)take(S_PJOBS, (uint64_t)cap * sizeof(ExtJob));
    auto* d_cnt = (uint32_t*)take(S_XCNT, EXT_NCNT * 4);
    uint32_t* d_bcnt = d_cnt + 4;
    // SSW buffers only when the requests extend with SSW
    const bool ssw_jobs = f.sp.paired || f.sp.use_ssw;
    auto* d_lists = (uint32_t*)take(S_LISTS, ssw_jobs ? (uint64_t)SSW_BUCKETS * cap * 4 : 16);
    auto* d_ends = (SswEnds*)take(S_ENDS, ssw_jobs ? (uint64_t)cap * sizeof(SswEnds) : 16);
    auto* d_fwd = (SswEnd*)take(S_FWD, ssw_jobs ? (uint64_t)cap * sizeof(SswEnd) : 16);
    const auto* d_chain = (const uint32_t*)buf_[S_CHAIN];
    k_zero_u32<<<1, 64, 0, st>>>(d_cnt, EXT_NCNT);
    k_ext_prep<<<F_PREP_BLOCKS, 256, 0, st>>>(d_seq, d_off, ref, sc, k, d_req, 0, nreq, d_res, d_slots, d_jobs,
                                               d_ljobs, d_pjobs, d_cnt);
    SA_CHECK(hipGetLastError());
    // the speculative requests are extend_seed only: SSW jobs exist when they
    // extend with SSW (paired-end, or single-end with --ssw), piecewise jobs
    // otherwise
    const bool ssw = ssw_jobs;
    if (ssw) {
        k_ssw_bucket<<<1024, 256, 0, st>>>(d_off, d_req, d_jobs, d_cnt, d_lists, d_bcnt, cap);
        SA_CHECK(hipGetLastError());
        // every stripe count a job of this shape can have, byte then word,
        // forward then reverse (a pass only adds jobs to the buckets of later
        // launches): forward passes from the shortest read's count to the
        // longest's, reverse passes from 1 (an alignment may end anywhere)
        const uint32_t Lq = std::min<uint32_t>(f.len_cap, SSW_GROUP_MAX_QUERY);
        const uint32_t Lm = std::min<uint32_t>(std::max<uint32_t>(f.len_min, 1), Lq);
        for (int word = 0; word < 2; ++word)
            for (int rev = 0; rev < 2; ++rev) {
                const uint32_t g = word ? 8 : 16;
                const uint32_t s0 = rev ? 1 : (Lm + g - 1) / g, s1 = (Lq + g - 1) / g;
                for (uint32_t seg = s0; seg <= s1; ++seg) {
                    SswLaunchLists ll{};
                    for (uint32_t x = 0; x < SSW_WCLASSES; ++x) {
                        const uint32_t b = rev ? ssw_bucket(word, true, seg, 0)
                                               : ssw_bucket(word, false, seg, SSW_WCLASSES - 1 - x);
                        ll.list[x] = d_lists + (uint64_t)b * cap;
                    }
                    if (word)
                        launch_ssw_pass<8, false, 1, SSW_WORD_SEGS>(seg, ll, st, d_seq, d_off, ref, sc, d_req,
                                                                    d_jobs, rev, d_fwd, d_ends, d_lists, d_bcnt,
                                                                    cap, (unsigned long long*)(d_cnt + EXT_CELLS),
                                                                    F_SSW_BLOCKS);
                    else
                        launch_ssw_pass<16, true, 1, SSW_BYTE_SEGS>(seg, ll, st, d_seq, d_off, ref, sc, d_req,
                                                                    d_jobs, rev, d_fwd, d_ends, d_lists, d_bcnt,
                                                                    cap, (unsigned long long*)(d_cnt + EXT_CELLS),
                                                                    F_SSW_BLOCKS);
                    SA_CHECK(hipGetLastError());
                }
            }
    }
    // banded traceback of the lane-group jobs (queries of at most
    // SSW_GROUP_MAX_QUERY; scratch per slot as in Engine::extend)
    const uint64_t L = f.len_cap;
    const uint32_t Lq = std::min<uint32_t>(f.len_cap, SSW_GROUP_MAX_QUERY);
    uint64_t per = 4ull * (3 * (2 * 64 + 4) + 17ull * Lq + 2 * (2ull * Lq + 64) + 64);
    per = (per + 255) & ~255ull;
    if (ssw) {
        const uint32_t fin_slots =
            (uint32_t)std::max<uint64_t>(128, std::min<uint64_t>(F_FIN_SLOTS, F_FIN_BUDGET / per) / 128 * 128);
        auto* d_fin = (uint32_t*)take(S_FINLISTS, (uint64_t)FIN_CLASSES * cap * 4);
        uint32_t* d_fcnt = d_cnt + EXT_FIN;
        k_fin_bucket<<<256, 256, 0, st>>>(d_ends, 0, d_cnt, d_fin, d_fcnt, cap);
        auto* fscr = (uint8_t*)take(S_SCRATCH, (uint64_t)fin_slots * per);
        k_ssw_finish<<<fin_slots / 128, 128, 0, st>>>(d_seq, d_off, ref, sc, d_req, d_jobs, d_ends, 0, d_cnt, d_res,
                                                      d_slots, fscr, per, 0, d_fin, d_fcnt, cap);
        SA_CHECK(hipGetLastError());
    }
    // one-thread SSW: only queries too long (or scored too high) for a lane group
    if (ssw && (f.len_cap > SSW_GROUP_MAX_QUERY || (uint64_t)sc.match * f.len_cap > 30000 || sc.match > 127 ||
                sc.mismatch > 127)) {
        uint64_t lper = 16ull * ((L + 15) / 16 * 16) + 256 + 65536 + 8 * (2 * L + 64) + 129ull * L * 3;
        lper = (lper + 255) & ~255ull;
        const uint32_t long_slots =
            (uint32_t)std::max<uint64_t>(128, std::min<uint64_t>(F_LONG_SLOTS, F_LONG_BUDGET / lper) / 128 * 128);
        auto* lscr = (uint8_t*)take(F_LSCR, (uint64_t)long_slots * lper);
        k_ext_long<<<long_slots / 128, 128, 0, st>>>(d_seq, d_off, ref, sc, d_req, d_ljobs, 0, d_cnt + 1, d_res,
                                                     d_slots, lscr, lper, 0);
        SA_CHECK(hipGetLastError());
    }
    // piecewise extension: single-end seeds without --ssw
    if (!f.sp.paired && !f.sp.use_ssw) {
        uint64_t pper = 16 * L + 256 + 16 * (2 * L + L / 10 + 16) + 16 * L + 4096;
        pper = (pper + 255) & ~255ull;
        const uint32_t rcap = (uint32_t)(L + L / 10 + 3), drows = (uint32_t)(L + 2);
        const uint64_t pw_per = pper + (uint64_t)rcap * sizeof(DpCol) + (uint64_t)rcap * drows;
        // lane-group slots: whole blocks of 128 lanes
        const uint64_t groups = std::max<uint64_t>(128 / PW_LANES, F_PW_BUDGET / pw_per / (128 / PW_LANES) * (128 / PW_LANES));
        auto* pscr = (uint8_t*)take(F_PSCR, groups * pper);
        auto* rows = (DpCol*)take(S_PWROWS, groups * rcap * sizeof(DpCol));
        auto* dirs = (uint8_t*)take(S_PWDIRS, groups * rcap * drows);
        k_ext_pw_grp<PW_LANES><<<(uint32_t)(groups * PW_LANES / 128), 128, 0, st>>>(
            d_seq, d_off, ref, sc, k, d_req, d_pjobs, 0, d_cnt + 2, d_res, d_slots, d_chain, pscr, pper, 0, rows,
            dirs, rcap, drows);
        SA_CHECK(hipGetLastError());
    }
    // CIGARs packed in request order
    auto* d_ncig = (uint64_t*)take(S_NCIG, (uint64_t)cap * 8);
    auto* d_coff = (uint64_t*)take(S_CIGOFF, (uint64_t)(cap + 1) * 8);
    k_ncigar_dev<<<1024, 256, 0, st>>>(d_res, nreq, d_ncig, cap);
    dscan(d_ncig, d_coff, cap, ctl + C_NCIG, 0, nullptr, F_SUM4);
    auto* d_cout = (uint32_t*)take(S_CIGOUT, f.slot_cap * 4 + 4);
    k_compact_cigars_dev<<<1024, 256, 0, st>>>(d_req, d_res, d_slots, d_coff, d_cout, nreq);
    SA_CHECK(hipGetLastError());
    SA_CHECK(hipMemcpyAsync(f.h_ctl, ctl, C_COUNT * 8, hipMemcpyDeviceToHost, st));
    SA_CHECK(hipMemcpyAsync(f.h_ctl + C_COUNT, d_cnt, EXT_NCNT * 4, hipMemcpyDeviceToHost, st));
}

bool Engine::map_fused(const MapParams& mp, const ChainParams& cp, const SpecParams& sp, const Scores& sc,
                       NamBatch& out, SpecBatch& spec) {
    SA_CHECK(hipSetDevice(device_));
    require_map_len(max_len_);
    const uint32_t n = n_reads_;
    spec = SpecBatch{};
    // the first batch of an engine (and any empty one) runs the plain pass, which
    // measures the anchors per base that size this one
    if (n == 0 || anchors_per_base_ <= 0) return false;
    if (!fused_) {
        fused_ = new Fused;
        fused_->device = device_;
        SA_CHECK(hipHostMalloc((void**)&fused_->h_ctl, C_COUNT * 8 + EXT_NCNT * 4 + 64, hipHostMallocDefault));
    }
    Fused& f = *fused_;
    chain_tables(cp);
    auto same = [](const auto& a, const auto& b) { return std::memcmp(&a, &b, sizeof a) == 0; };
    const bool fits = !f.stale && n <= f.n_cap && n_bases_ <= f.bases_cap && max_len_ <= f.len_cap &&
                      (min_len_ == 0 || min_len_ >= f.len_min) &&
                      same(mp, f.mp) && same(cp, f.cp) && same(sp, f.sp) && same(sc, f.sc) &&
                      d_entries_ == f.entries;
    if (!fits) {
        f.drop_graph();
        f.runs = 0;
        f.stale = false;
        f.mp = mp;
        f.cp = cp;
        f.sp = sp;
        f.sc = sc;
        f.entries = d_entries_;
        // room for batches a little larger than this one, so that one recorded
        // shape serves the batches of a run
        f.n_cap = (uint32_t)round_up(n + n / 8 + 64, 1024);
        f.bases_cap = round_up(n_bases_ + n_bases_ / 8 + 4096, 1 << 16);
        f.len_cap = std::max<uint32_t>(f.len_cap, (uint32_t)round_up(max_len_, 16));
        // a little room below this batch's shortest read (a multiple of 16)
        const uint32_t lmin = min_len_ > 32 ? (min_len_ - 16) / 16 * 16 : 1;
        f.len_min = f.len_min ? std::min(f.len_min, lmin) : lmin;
        f.a_cap = (uint64_t)(anchors_per_base_ * 1.5 * (double)f.bases_cap) + 4096;
        const double rpr = std::max(f.req_per_read, sp.paired ? 1.25 : 2.0);
        f.req_cap = round_up((uint64_t)(rpr * 1.25 * f.n_cap) + 1024, 1024);
        const double avg_len = (double)n_bases_ / (double)n;
        const double spr = std::max(f.slots_per_req, 2.0 * avg_len + 8.0);
        f.slot_cap = (uint64_t)(spr * 1.25 * (double)f.req_cap) + 4096;
        if (f.req_cap > 0xFFFFFFFFull / 2) throw std::runtime_error("fused pass: too many speculative requests");
    } else if (f.gen != alloc_gen_) {
        // a buffer moved (the plain stages grew it): record the pass again; its
        // buffers are at least as large as when it ran directly
        f.drop_graph();
    }
    // the batch's reads into the pass's fixed buffers; the reads past the batch are empty
    auto* fseq = (uint8_t*)grow(F_SEQ, f.bases_cap + 16);
    auto* foff = (uint64_t*)grow(F_OFF, (uint64_t)(f.n_cap + 1) * 8);
    if (n_bases_) SA_CHECK(hipMemcpyAsync(fseq, d_seq_, n_bases_, hipMemcpyDeviceToDevice, stream_));
    SA_CHECK(hipMemcpyAsync(foff, cur_off_, (uint64_t)(n + 1) * 8, hipMemcpyDeviceToDevice, stream_));
    if (f.n_cap > n) k_pad_off<<<blocks(f.n_cap - n, 256), 256, 0, stream_>>>(foff, n, f.n_cap);
    SA_CHECK(hipGetLastError());
    hipEvent_t e0 = event(), e1 = event();
    try {
        SA_CHECK(hipEventRecord(e0, stream_));
        if (f.runs == 0) {
            // run directly: grows the buffers to this shape
            fused_sequence(f);
            f.gen = alloc_gen_;
        } else {
            if (!f.exec) {
                SA_CHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
                recording_ = true;
                try {
                    fused_sequence(f);
                } catch (...) {
                    recording_ = false;
                    hipGraph_t g = nullptr;
                    (void)hipStreamEndCapture(stream_, &g);
                    if (g) (void)hipGraphDestroy(g);
                    throw;
                }
                recording_ = false;
                SA_CHECK(hipStreamEndCapture(stream_, &f.graph));
                SA_CHECK(hipGraphInstantiate(&f.exec, f.graph, nullptr, nullptr, 0));
                f.gen = alloc_gen_;
            } else {
                ++fused_replays_;
                spec.replayed = true;
            }
            SA_CHECK(hipGraphLaunch(f.exec, stream_));
        }
    } catch (...) {
        // nothing of a pass that failed part-way may still run once this returns
        drain();
        f.drop_graph();
        f.stale = true;
        release_event(e0);
        release_event(e1);
        throw;
    }
    ++f.runs;
    SA_CHECK(hipEventRecord(e1, stream_));
    sync();
    float ms = 0;
    SA_CHECK(hipEventElapsedTime(&ms, e0, e1));
    release_event(e0);
    release_event(e1);
    const uint64_t* ctl = f.h_ctl;
    const uint32_t overflow = (uint32_t)ctl[C_FLAG];
    const uint64_t n_req = (uint32_t)ctl[C_NREQ], n_nams = ctl[C_NNAMS], anchors = ctl[C_ANCH];
    const uint64_t wanted = ctl[C_WANT], wanted_slots = ctl[C_WANTS], n_cig = ctl[C_NCIG];
    // what the next shape is sized from
    anchors_per_base_ = std::max(anchors_per_base_, (double)anchors / (double)std::max<uint64_t>(1, n_bases_));
    f.req_per_read = std::max(f.req_per_read, (double)wanted / (double)n);
    if (wanted) f.slots_per_req = std::max(f.slots_per_req, (double)wanted_slots / (double)wanted);
    if (wanted > f.req_cap || wanted_slots > f.slot_cap) f.stale = true;
    times_ = StageTimes{};
    times_.fused = ms;
    total_ms_ += ms;
    add_kernel("fused_pass", ms, n);
    if (overflow) {
        f.stale = true;
        spec.overflow = true;
        return false;
    }
    // speculative results: requests, results, CIGAR offsets and CIGARs
    const uint64_t b_req = round_up(n_req * sizeof(ExtRequest), 64), b_res = round_up(n_req * sizeof(ExtResult), 64);
    const uint64_t b_off = round_up((n_req + 1) * 8, 64);
    auto* hs = (uint8_t*)host_stage(2, b_req + b_res + b_off + n_cig * 4 + 64);
    spec.req = HostView<ExtRequest>{(ExtRequest*)hs, n_req};
    spec.res = HostView<ExtResult>{(ExtResult*)(hs + b_req), n_req};
    spec.cigar_off = HostView<uint64_t>{(uint64_t*)(hs + b_req + b_res), n_req + 1};
    spec.cigars = HostView<uint32_t>{(uint32_t*)(hs + b_req + b_res + b_off), n_cig};
    spec.wanted = wanted;
    if (n_req) {
        SA_CHECK(hipMemcpyAsync(spec.req.data(), buf_[S_REQ], n_req * sizeof(ExtRequest), hipMemcpyDeviceToHost,
                                stream_));
        SA_CHECK(hipMemcpyAsync(spec.res.data(), buf_[S_RES], n_req * sizeof(ExtResult), hipMemcpyDeviceToHost,
                                stream_));
        SA_CHECK(hipMemcpyAsync(spec.cigar_off.data(), buf_[S_CIGOFF], n_req * 8, hipMemcpyDeviceToHost, stream_));
    }
    if (n_cig) SA_CHECK(hipMemcpyAsync(spec.cigars.data(), buf_[S_CIGOUT], n_cig * 4, hipMemcpyDeviceToHost, stream_));
    // the NAMs (finish_nams waits for all of the copies)
    NamPass o;
    o.qrn = (uint32_t*)buf_[S_QRN];
    o.hc = (HitCounts*)buf_[S_HC];
    o.namn = (uint32_t*)buf_[S_NAMN];
    o.araw = (uint32_t*)buf_[S_ARAW];
    o.cbytes = (uint32_t*)buf_[S_CBYTES];
    o.hbytes = (uint32_t*)buf_[S_HBYTES];
    o.outoff = (uint64_t*)buf_[S_OUTOFF];
    o.nams = (NamRec*)buf_[S_NAMS];
    o.n_nams = n_nams;
    finish_nams(n, o, nullptr, out);
    spec.cigar_off[n_req] = n_cig;
    // the extension calls that follow work on this pass's reads and anchors
    d_seq_ = fseq;
    cur_off_ = foff;
    cur_chain_ = (const uint32_t*)buf_[S_CHAIN];
    return true;
}

}  // namespace sa
