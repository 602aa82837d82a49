// FASTQ parsing on the device (src/io/fastq.rs:22-95): the file's bytes are
// copied in once, its newlines found and numbered in two passes over 16-byte
// vectors (counts per block, an exclusive scan, positions written by each
// thread at its offset), a thread per record applies fq_record to its four
// lines, and the uppercase sequences are gathered into the staged read layout
// by one 16-lane group per read. The host gets back only the record locations.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

#include "sa_engine.hpp"
#include "sa_ingest.hpp"

#define SA_ICHECK(x)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (x);                                                                           \
        if (e_ != hipSuccess)                                                                          \
            throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " +     \
                                     __FILE__ + ":" + std::to_string(__LINE__) + " (" #x ")");         \
    } while (0)

namespace sa {

namespace {

constexpr uint32_t NL_THREADS = 256, NL_BYTES = 16;  // a thread reads one 16-byte vector

__device__ __forceinline__ uint32_t nl_mask(uint4 v, uint64_t base, uint64_t n) {
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
    uint32_t m = 0;
#pragma unroll
    for (int b = 0; b < 16; ++b)
        if (((w[b / 4] >> (8 * (b % 4))) & 0xFFu) == '\n' && base + b < n) m |= 1u << b;
    return m;
}

// newlines per block of NL_THREADS x 16 bytes
__global__ void __launch_bounds__(NL_THREADS) k_nl_count(const uint4* __restrict__ raw, uint64_t n,
                                                          uint32_t* __restrict__ cnt) {
    const uint64_t t = (uint64_t)blockIdx.x * NL_THREADS + threadIdx.x;
    uint32_t c = 0;
    if (t * NL_BYTES < n) c = __builtin_popcount(nl_mask(raw[t], t * NL_BYTES, n));
    using Red = hipcub::BlockReduce<uint32_t, NL_THREADS>;
    __shared__ typename Red::TempStorage tmp;
    const uint32_t s = Red(tmp).Sum(c);
    if (threadIdx.x == 0) cnt[blockIdx.x] = s;
}

// newline positions, numbered in file order from the blocks' offsets
__global__ void __launch_bounds__(NL_THREADS) k_nl_write(const uint4* __restrict__ raw, uint64_t n,
                                                          const uint64_t* __restrict__ blk_off,
                                                          uint64_t* __restrict__ nl) {
    const uint64_t t = (uint64_t)blockIdx.x * NL_THREADS + threadIdx.x;
    uint32_t m = 0;
    if (t * NL_BYTES < n) m = nl_mask(raw[t], t * NL_BYTES, n);
    using Scan = hipcub::BlockScan<uint32_t, NL_THREADS>;
    __shared__ typename Scan::TempStorage tmp;
    uint32_t before;
    Scan(tmp).ExclusiveSum((uint32_t)__builtin_popcount(m), before);
    uint64_t at = blk_off[blockIdx.x] + before;
    while (m) {
        const uint32_t b = (uint32_t)__builtin_ctz(m);
        m &= m - 1;
        nl[at++] = t * NL_BYTES + b;
    }
}

// one thread per record
__global__ void __launch_bounds__(256) k_fq_records(const uint8_t* __restrict__ raw, uint64_t n,
                                                    const uint64_t* __restrict__ nl, uint64_t n_nl, uint64_t nrec,
                                                    FqLoc* __restrict__ loc) {
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrec) return;
    FqLine l[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) l[j] = fq_line(nl, n_nl, n, 4 * r + j);
    loc[r] = fq_record(raw, l);
}

// sequence lengths of read x in the staged layout (reads of both files interleaved)
__global__ void k_fq_lengths(const FqLoc* __restrict__ a, const FqLoc* __restrict__ b, uint64_t nrec,
                             uint64_t* __restrict__ len) {
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t nx = b ? 2 * nrec : nrec;
    if (x >= nx) return;
    len[x] = b ? (x & 1 ? b[x >> 1].seq_len : a[x >> 1].seq_len) : a[x].seq_len;
}

// uppercase sequences into the staged layout, 16 lanes per read
__global__ void __launch_bounds__(256) k_fq_copy(const uint8_t* __restrict__ ra, const uint8_t* __restrict__ rb,
                                                 const FqLoc* __restrict__ a, const FqLoc* __restrict__ b,
                                                 uint64_t nrec, const uint64_t* __restrict__ off,
                                                 uint8_t* __restrict__ seq) {
    const uint64_t g = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / 16;
    const uint32_t lane = threadIdx.x % 16;
    const uint64_t nx = b ? 2 * nrec : nrec;
    if (g >= nx) return;
    const bool second = b && (g & 1);
    const FqLoc& r = second ? b[g >> 1] : (b ? a[g >> 1] : a[g]);
    const uint8_t* src = (second ? rb : ra) + r.seq_off;
    uint8_t* dst = seq + off[g];
    for (uint32_t i = lane; i < r.seq_len; i += 16) dst[i] = ascii_upper(src[i]);
}

inline uint32_t nblocks(uint64_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

struct DevBuf {
    void* p = nullptr;
    explicit DevBuf(uint64_t bytes) { SA_ICHECK(hipMalloc(&p, bytes ? bytes : 16)); }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    template <typename T>
    T* as() const { return (T*)p; }
};

void exclusive_sum(const uint64_t* in, uint64_t* out, uint64_t n, hipStream_t st) {
    size_t tmp = 0;
    SA_ICHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, in, out, (int)n, st));
    DevBuf t(tmp);
    SA_ICHECK(hipcub::DeviceScan::ExclusiveSum(t.p, tmp, in, out, (int)n, st));
}

// the newline table and the records of one file resident at d_raw
struct FileParse {
    std::unique_ptr<DevBuf> nl, loc;
    uint64_t nrec = 0;
};

FileParse parse_file(const uint8_t* d_raw, uint64_t n, bool last_is_newline, hipStream_t st) {
    FileParse f;
    const uint64_t vecs = (n + NL_BYTES - 1) / NL_BYTES;
    const uint32_t nb = nblocks(vecs, NL_THREADS);
    DevBuf cnt((uint64_t)nb * 4 + 4), blk((uint64_t)nb * 8 + 8);
    if (nb) {
        k_nl_count<<<nb, NL_THREADS, 0, st>>>((const uint4*)d_raw, n, cnt.as<uint32_t>());
        SA_ICHECK(hipGetLastError());
    }
    // the blocks' offsets: a scan of the counts on the host (one more entry for the total)
    std::vector<uint32_t> hc(nb);
    if (nb) SA_ICHECK(hipMemcpyAsync(hc.data(), cnt.p, (uint64_t)nb * 4, hipMemcpyDeviceToHost, st));
    SA_ICHECK(hipStreamSynchronize(st));
    std::vector<uint64_t> ho(nb + 1, 0);
    for (uint32_t b = 0; b < nb; ++b) ho[b + 1] = ho[b] + hc[b];
    const uint64_t n_nl = ho[nb];
    SA_ICHECK(hipMemcpyAsync(blk.p, ho.data(), (uint64_t)(nb + 1) * 8, hipMemcpyHostToDevice, st));
    f.nl = std::make_unique<DevBuf>(n_nl * 8 + 8);
    if (nb) {
        k_nl_write<<<nb, NL_THREADS, 0, st>>>((const uint4*)d_raw, n, blk.as<uint64_t>(), f.nl->as<uint64_t>());
        SA_ICHECK(hipGetLastError());
    }
    f.nrec = fq_records(n_nl, n, last_is_newline);
    f.loc = std::make_unique<DevBuf>(f.nrec * sizeof(FqLoc) + sizeof(FqLoc));
    if (f.nrec) {
        k_fq_records<<<nblocks(f.nrec, 256), 256, 0, st>>>(d_raw, n, f.nl->as<uint64_t>(), n_nl, f.nrec,
                                                           f.loc->as<FqLoc>());
        SA_ICHECK(hipGetLastError());
    }
    SA_ICHECK(hipStreamSynchronize(st));  // the scratch buffers above go out of scope
    return f;
}

}  // namespace

double ingest_fastq(int device, const std::string& r1, const std::string* r2, std::vector<FqLoc>& loc1,
                    std::vector<FqLoc>& loc2, void** d_seq, std::vector<uint64_t>& off) {
    SA_ICHECK(hipSetDevice(device));
    hipStream_t st = nullptr;
    SA_ICHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{st};
    hipEvent_t e0, e1;
    SA_ICHECK(hipEventCreate(&e0));
    SA_ICHECK(hipEventCreate(&e1));
    struct EventGuard {
        hipEvent_t a, b;
        ~EventGuard() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); }
    } eg{e0, e1};
    SA_ICHECK(hipEventRecord(e0, st));
    // the files' bytes, padded to whole 16-byte vectors
    // the files' bytes through two pinned staging buffers: the copy of one
    // chunk into pinned memory overlaps the transfer of the previous one
    constexpr uint64_t CHUNK = 32ull << 20;
    struct Pinned {
        void* p[2] = {nullptr, nullptr};
        hipEvent_t done[2] = {nullptr, nullptr};
        ~Pinned() {
            for (int i = 0; i < 2; ++i) {
                if (done[i]) (void)hipEventDestroy(done[i]);
                if (p[i]) (void)hipHostFree(p[i]);
            }
        }
    } pin;
    for (int i = 0; i < 2; ++i) {
        SA_ICHECK(hipHostMalloc(&pin.p[i], CHUNK, hipHostMallocDefault));
        SA_ICHECK(hipEventCreateWithFlags(&pin.done[i], hipEventDisableTiming));
    }
    uint64_t nchunk = 0;
    auto upload = [&](const std::string& s) {
        auto b = std::make_unique<DevBuf>((s.size() + 31) / 16 * 16);
        SA_ICHECK(hipMemsetAsync(b->as<uint8_t>() + s.size() / 16 * 16, 0, (s.size() + 31) / 16 * 16 - s.size() / 16 * 16,
                                 st));
        for (uint64_t at = 0; at < s.size(); at += CHUNK, ++nchunk) {
            const int x = (int)(nchunk & 1);
            const uint64_t len = std::min<uint64_t>(CHUNK, s.size() - at);
            SA_ICHECK(hipEventSynchronize(pin.done[x]));  // the buffer's previous transfer is done
            std::memcpy(pin.p[x], s.data() + at, len);
            SA_ICHECK(hipMemcpyAsync(b->as<uint8_t>() + at, pin.p[x], len, hipMemcpyHostToDevice, st));
            SA_ICHECK(hipEventRecord(pin.done[x], st));
        }
        return b;
    };
    auto raw1 = upload(r1);
    std::unique_ptr<DevBuf> raw2 = r2 ? upload(*r2) : nullptr;
    FileParse f1 = parse_file(raw1->as<uint8_t>(), r1.size(), !r1.empty() && r1.back() == '\n', st);
    FileParse f2;
    if (r2) f2 = parse_file(raw2->as<uint8_t>(), r2->size(), !r2->empty() && r2->back() == '\n', st);
    loc1.resize(f1.nrec);
    loc2.resize(r2 ? f2.nrec : 0);
    if (f1.nrec) SA_ICHECK(hipMemcpyAsync(loc1.data(), f1.loc->p, f1.nrec * sizeof(FqLoc), hipMemcpyDeviceToHost, st));
    if (r2 && f2.nrec)
        SA_ICHECK(hipMemcpyAsync(loc2.data(), f2.loc->p, f2.nrec * sizeof(FqLoc), hipMemcpyDeviceToHost, st));
    SA_ICHECK(hipStreamSynchronize(st));
    *d_seq = nullptr;
    off.assign(1, 0);
    // the sequences are staged only when both files parse and pair up
    for (const auto* v : {&loc1, &loc2})
        for (const FqLoc& r : *v)
            if (r.err) return 0.0;
    if (r2 && f1.nrec != f2.nrec) return 0.0;
    const uint64_t nx = r2 ? 2 * f1.nrec : f1.nrec;
    DevBuf len(nx * 8 + 8), doff(nx * 8 + 8);
    if (nx) {
        k_fq_lengths<<<nblocks(nx, 256), 256, 0, st>>>(f1.loc->as<FqLoc>(), r2 ? f2.loc->as<FqLoc>() : nullptr,
                                                       f1.nrec, len.as<uint64_t>());
        SA_ICHECK(hipGetLastError());
        exclusive_sum(len.as<uint64_t>(), doff.as<uint64_t>(), nx, st);
    }
    off.resize(nx + 1);
    if (nx) SA_ICHECK(hipMemcpyAsync(off.data(), doff.p, nx * 8, hipMemcpyDeviceToHost, st));
    SA_ICHECK(hipStreamSynchronize(st));
    uint64_t total = 0;
    for (uint64_t x = 0; x < nx; ++x) total += r2 ? (x & 1 ? loc2[x >> 1].seq_len : loc1[x >> 1].seq_len) : loc1[x].seq_len;
    off[nx] = total;
    void* seq = device_alloc(device, total + 16);
    try {
        SA_ICHECK(hipMemsetAsync((uint8_t*)seq + total, 0, 16, st));
        if (nx) {
            k_fq_copy<<<nblocks(nx * 16, 256), 256, 0, st>>>(raw1->as<uint8_t>(), r2 ? raw2->as<uint8_t>() : nullptr,
                                                             f1.loc->as<FqLoc>(), r2 ? f2.loc->as<FqLoc>() : nullptr,
                                                             f1.nrec, doff.as<uint64_t>(), (uint8_t*)seq);
            SA_ICHECK(hipGetLastError());
        }
        SA_ICHECK(hipEventRecord(e1, st));
        SA_ICHECK(hipStreamSynchronize(st));
    } catch (...) {
        device_free(device, seq);
        throw;
    }
    *d_seq = seq;
    float ms = 0;
    SA_ICHECK(hipEventElapsedTime(&ms, e0, e1));
    return ms;
}

}  // namespace sa
