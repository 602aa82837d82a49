// Query randstrobes of one read on one wave (64 lanes): the same records as
// read_randstrobes (src/seeding/mod.rs:19-75), computed a lane per read
// position and then a lane per syncmer.
//
// Syncmers (src/seeding/syncmers.rs:171-237) in chunks of 64 positions, lane j
// at position base + j:
//  - the last 32 bases as a 2-bit packed word (newest lowest), built in five
//    doubling steps from the lanes 1, 2, 4, 8, 16 positions back (from the
//    previous chunk's words where the lane index would go below 0);
//  - the last N at or before the position, an inclusive max-scan carried across
//    chunks: the scanner restarts after an N, so an s-mer is hashed once the s
//    bases up to it are clean and the window of W = k - s + 1 s-mer hashes is
//    full once the k bases up to it are;
//  - the canonical s-mer hash, the minimum over the W hashes that end at the
//    W positions up to this one (read from the lanes before, or the previous
//    chunk's), and the test that the (t-1)-th oldest of them is that minimum;
//  - the canonical k-mer hash of each syncmer, appended to a list in LDS in
//    position order (ballot and prefix count).
// Randstrobes (src/seeding/strobes.rs:145-176): a lane per syncmer picks the
// second strobe from the list with pick_strobe2, forward and in the reverse-
// complement order (src/seeding/mod.rs:53-57), and writes its record.
#pragma once
#include "sa_seed.hpp"

namespace sa {

// Longest read the wave form takes (its syncmer list lives in LDS); longer
// reads are seeded one thread each.
constexpr uint32_t SEED_WAVE_MAX_LEN = 512;

struct SeedWaveLds {
    uint64_t hash[SEED_WAVE_MAX_LEN];
    uint32_t pos_fwd[SEED_WAVE_MAX_LEN];  // position << 1 | is_forward
};

// Reverse complement of the 2-bit packed n-mer in the low 2n bits of x (first
// base highest), in the same packing.
__device__ inline uint64_t rc_packed(uint64_t x, int n) {
    x = ~x;
    x = ((x >> 2) & 0x3333333333333333ull) | ((x & 0x3333333333333333ull) << 2);
    x = ((x >> 4) & 0x0F0F0F0F0F0F0F0Full) | ((x & 0x0F0F0F0F0F0F0F0Full) << 4);
    x = __builtin_bswap64(x);
    return x >> (64 - 2 * n);
}

// Value of lane (lane - d) of `cur`, or of the previous chunk's `prev` where
// that lane index would be negative.
__device__ inline uint64_t back64(uint64_t cur, uint64_t prev, uint32_t lane, uint32_t d) {
    const int src = (int)((lane - d) & 63u);
    const uint64_t a = __shfl(cur, src), b = __shfl(prev, src);
    return lane >= d ? a : b;
}

// The forward and reverse-complement randstrobes of read seq[0, len) (len at
// most SEED_WAVE_MAX_LEN) into out0 / out1, their count into *n01 (both
// orientations have the same count). Every lane of the wave calls it.
__device__ inline void read_randstrobes_wave(const uint8_t* __restrict__ seq, uint32_t len, const SeedParams& p,
                                             QRand* __restrict__ out0, QRand* __restrict__ out1, uint32_t* n0,
                                             uint32_t* n1, SeedWaveLds& lds, uint32_t lane) {
    if (len < (uint32_t)p.w_max) {
        if (lane == 0) *n0 = *n1 = 0;
        return;
    }
    const int k = p.k, s = p.s;
    const uint32_t W = (uint32_t)(k - s + 1), tback = W - (uint32_t)p.t;  // the tested hash: tback positions back
    const uint64_t kmask = k == 32 ? ~0ull : ((1ull << (2 * k)) - 1);
    const uint64_t smask = (1ull << (2 * s)) - 1;
    uint64_t pv[5] = {0, 0, 0, 0, 0};  // previous chunk's packed words of 1, 2, 4, 8, 16 bases
    uint64_t ph = 0;                    // previous chunk's s-mer hash
    int32_t last_n = -1;                // last N before this chunk
    uint32_t n = 0;                     // syncmers so far
    for (uint32_t base = 0; base < len; base += 64) {
        const uint32_t i = base + lane;
        const bool in = i < len;
        const uint32_t c = in ? nt2(seq[i]) : 4u;
        int32_t ln = c >= 4 ? (int32_t)i : -1;
SA_UNROLL
        for (uint32_t d = 1; d < 64; d <<= 1) {
            const int32_t u = __shfl_up(ln, d);
            if (lane >= d) ln = ln > u ? ln : u;
        }
        ln = ln > last_n ? ln : last_n;
        last_n = __shfl(ln, 63);
        // the last 32 bases, newest lowest
        uint64_t v = c & 3u, cur[5];
SA_UNROLL
        for (uint32_t m = 0; m < 5; ++m) {
            cur[m] = v;
            const uint32_t d = 1u << m;
            v = (back64(v, pv[m], lane, d) << (2 * d)) | v;
        }
SA_UNROLL
        for (uint32_t m = 0; m < 5; ++m) pv[m] = cur[m];
        const uint64_t xs0 = v & smask, xs1 = rc_packed(xs0, s);
        const uint64_t hs = xxh64(xs0 < xs1 ? xs0 : xs1);
        // minimum of the W s-mer hashes ending here, and the one tback back
        uint64_t mn = hs, at = hs;
        for (uint32_t d = 1; d < W; ++d) {
            const uint64_t h = back64(hs, ph, lane, d);
            mn = h < mn ? h : mn;
            if (d == tback) at = h;
        }
        ph = hs;
        const bool sync = in && (int32_t)i - ln >= k && at == mn;
        uint64_t hk = 0;
        bool fwd = false;
        if (sync) {
            const uint64_t xk0 = v & kmask, xk1 = rc_packed(xk0, k);
            hk = xxh64(xk0 < xk1 ? xk0 : xk1);
            fwd = xk0 <= xk1;
        }
        const uint64_t bal = __ballot(sync);
        if (sync) {
            const uint32_t at_n = n + (uint32_t)__popcll(bal & ((1ull << lane) - 1));
            lds.hash[at_n] = hk;
            lds.pos_fwd[at_n] = ((i + 1 - (uint32_t)k) << 1) | (fwd ? 1u : 0u);
        }
        n += (uint32_t)__popcll(bal);
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const uint32_t k32 = (uint32_t)k;
    auto fwd_at = [&](uint32_t x) { return Syncmer{lds.hash[x], lds.pos_fwd[x], 0}; };
    // reverse complement: the syncmers last first, position len - pos - k, orientation toggled
    auto rc_at = [&](uint32_t x) {
        const uint32_t pf = lds.pos_fwd[n - 1 - x];
        return Syncmer{lds.hash[n - 1 - x], ((len - (pf >> 1) - k32) << 1) | ((pf & 1u) ^ 1u), 0};
    };
    for (uint32_t a = lane; a < n; a += 64) {
        {
            const uint32_t b = pick_strobe2(a, n, p, fwd_at);
            const Syncmer s1 = fwd_at(a), s2 = fwd_at(b);
            const uint64_t f1 = s1.pos_fwd & 1, f2 = s2.pos_fwd & 1;
            out0[a] = QRand{randstrobe_hash(s1.hash, s2.hash, f1, f2, p),
                            randstrobe_hash(s2.hash, s1.hash, f1 ^ 1, f2 ^ 1, p), s1.pos_fwd >> 1,
                            (s2.pos_fwd >> 1) + k32};
        }
        {
            const uint32_t b = pick_strobe2(a, n, p, rc_at);
            const Syncmer s1 = rc_at(a), s2 = rc_at(b);
            const uint64_t f1 = s1.pos_fwd & 1, f2 = s2.pos_fwd & 1;
            out1[a] = QRand{randstrobe_hash(s1.hash, s2.hash, f1, f2, p),
                            randstrobe_hash(s2.hash, s1.hash, f1 ^ 1, f2 ^ 1, p), s1.pos_fwd >> 1,
                            (s2.pos_fwd >> 1) + k32};
        }
    }
    if (lane == 0) *n0 = *n1 = n;
}

}  // namespace sa
