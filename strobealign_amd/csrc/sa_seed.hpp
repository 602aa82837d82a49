// Per-item algorithms for seeding, index lookups, hit finding, anchor
// chaining and NAM extraction. Each function handles one read (or one read
// orientation) and is called by one GPU thread from the kernels in
// sa_kernels.hip; the functions are __host__ __device__ only so that the
// index builder can share the seeding code on the host.
#pragma once
#include "sa_types.hpp"

namespace sa {

// ----------------------------------------------------------------- hash ---
// XXH64 of the 8 little-endian bytes of x with seed 0 (src/seeding/hash.rs:11-27).
SA_HD uint64_t rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }

SA_HD uint64_t xxh64(uint64_t input) {
    const uint64_t P1 = 0x9E3779B185EBCA87ULL, P2 = 0xC2B2AE3D27D4EB4FULL, P3 = 0x165667B19E3779F9ULL,
                   P4 = 0x85EBCA77C2B2AE63ULL, P5 = 0x27D4EB2F165667C5ULL;
    uint64_t h = P5 + 8;
    h ^= rotl64(input * P2, 31) * P1;
    h = rotl64(h, 27) * P1 + P4;
    h ^= h >> 33;
    h *= P2;
    h ^= h >> 29;
    h *= P3;
    h ^= h >> 32;
    return h;
}

// ASCII -> 2-bit code, 4 for anything that is not ACGTU (src/seeding/syncmers.rs:14-40).
SA_HD uint32_t nt2(uint8_t c) {
    switch (c) {
        case 'A': case 'a': return 0;
        case 'C': case 'c': return 1;
        case 'G': case 'g': return 2;
        case 'T': case 't': case 'U': case 'u': return 3;
        default: return 4;
    }
}

// Complement used for read reverse complements (src/revcomp.rs:1-29).
// Nucleotide letters without a switch (a switch over a per-lane value becomes a
// branch tree in device code). d = lowercase letter - 0x60: a=1, c=3, g=7, t=20,
// u=21; for a/c/g/t, ((x >> 1) ^ (x >> 2)) & 3 gives 0, 1, 2, 3.
SA_HD bool nt_is_acgt(uint8_t c) {
    const uint32_t d = (uint32_t)(c | 0x20u) - 0x60u;
    return d < 32u && ((0x10008Au >> d) & 1u);
}
SA_HD bool nt_is_u(uint8_t c) { return (c | 0x20u) == 'u'; }
SA_HD uint32_t nt_acgt_code(uint8_t c) {
    const uint32_t x = c | 0x20u;
    return ((x >> 1) ^ (x >> 2)) & 3u;
}

SA_HD uint8_t comp_ascii(uint8_t c) {
    // A<->T, C<->G, U->A, anything else N
    const uint8_t tgca = (uint8_t)((0x41434754u >> (8 * nt_acgt_code(c))) & 0xFFu);
    return nt_is_acgt(c) ? tgca : (nt_is_u(c) ? (uint8_t)'A' : (uint8_t)'N');
}

// ------------------------------------------------------------- syncmers ---
// Canonical open syncmers (src/seeding/syncmers.rs:171-237). A k-mer is emitted
// when the canonical s-mer at offset t-1 holds the minimum xxh64 of the window
// of k-s+1 s-mers (the reference's rolling minimum equals the exact window minimum).
// `code(i)` returns 0..3 or 4 (restart). Emits via `emit(hash, is_forward, position)`.
template <typename Code, typename Emit>
SA_HD void syncmer_scan(uint32_t len, const SeedParams& p, Code code, Emit emit) {
    const int k = p.k, s = p.s, t = p.t;
    const uint64_t kmask = (k == 32) ? ~0ULL : ((1ULL << (2 * k)) - 1);
    const uint64_t smask = (1ULL << (2 * s)) - 1;
    const int kshift = (k - 1) * 2, sshift = (s - 1) * 2;
    const uint32_t w = (uint32_t)(k - s + 1);
    uint64_t ring[32];
    uint64_t xk0 = 0, xk1 = 0, xs0 = 0, xs1 = 0;
    uint32_t l = 0, nq = 0, head = 0;  // nq: hashes held in ring; head: slot of the oldest
    for (uint32_t i = 0; i < len; ++i) {
        uint32_t c = code(i);
        if (c < 4) {
            xk0 = ((xk0 << 2) | c) & kmask;
            xk1 = (xk1 >> 2) | ((uint64_t)(3 - c) << kshift);
            xs0 = ((xs0 << 2) | c) & smask;
            xs1 = (xs1 >> 2) | ((uint64_t)(3 - c) << sshift);
            ++l;
            if (l < (uint32_t)s) continue;
            uint64_t hs = xxh64(xs0 < xs1 ? xs0 : xs1);
            if (nq < w) {
                ring[(head + nq) % w] = hs;
                ++nq;
                if (nq < w) continue;
            } else {
                ring[head] = hs;
                head = (head + 1) % w;
            }
            uint64_t mn = ~0ULL;
            for (uint32_t j = 0; j < w; ++j) mn = ring[j] < mn ? ring[j] : mn;
            if (ring[(head + t - 1) % w] == mn) {
                uint64_t yk = xk0 < xk1 ? xk0 : xk1;
                emit(xxh64(yk), xk0 <= xk1, i + 1 - (uint32_t)k);
            }
        } else {
            l = 0;
            nq = 0;
            head = 0;
            xk0 = xk1 = xs0 = xs1 = 0;
        }
    }
}

// syncmer_scan with the window length W = k-s+1 fixed at compile time, so the
// window lives in registers (a shift register) instead of an indexed ring.
template <int W, typename Code, typename Emit>
SA_HD void syncmer_scan_w(uint32_t len, const SeedParams& p, Code code, Emit emit) {
    const int k = p.k, s = p.s;
    const uint32_t tm1 = (uint32_t)p.t - 1;
    const uint64_t kmask = (k == 32) ? ~0ULL : ((1ULL << (2 * k)) - 1);
    const uint64_t smask = (1ULL << (2 * s)) - 1;
    const int kshift = (k - 1) * 2, sshift = (s - 1) * 2;
    uint64_t hv[W];  // s-mer hashes of the window, oldest first
SA_UNROLL
    for (int j = 0; j < W; ++j) hv[j] = 0;
    uint64_t xk0 = 0, xk1 = 0, xs0 = 0, xs1 = 0;
    uint32_t l = 0, nq = 0;
    for (uint32_t i = 0; i < len; ++i) {
        const uint32_t c = code(i);
        if (c < 4) {
            xk0 = ((xk0 << 2) | c) & kmask;
            xk1 = (xk1 >> 2) | ((uint64_t)(3 - c) << kshift);
            xs0 = ((xs0 << 2) | c) & smask;
            xs1 = (xs1 >> 2) | ((uint64_t)(3 - c) << sshift);
            ++l;
            if (l < (uint32_t)s) continue;
            const uint64_t hs = xxh64(xs0 < xs1 ? xs0 : xs1);
SA_UNROLL
            for (int j = 0; j + 1 < W; ++j) hv[j] = hv[j + 1];
            hv[W - 1] = hs;
            if (nq < (uint32_t)W) ++nq;
            if (nq < (uint32_t)W) continue;
            uint64_t mn = hv[0], at = hv[0];
SA_UNROLL
            for (int j = 1; j < W; ++j) {
                mn = hv[j] < mn ? hv[j] : mn;
                at = (uint32_t)j == tm1 ? hv[j] : at;
            }
            if (at == mn) {
                const uint64_t yk = xk0 < xk1 ? xk0 : xk1;
                emit(xxh64(yk), xk0 <= xk1, i + 1 - (uint32_t)k);
            }
        } else {
            l = 0;
            nq = 0;
            xk0 = xk1 = xs0 = xs1 = 0;
        }
    }
}

// src/seeding/strobes.rs:112-124 (bit 8, the strobe-2 orientation, is masked off)
SA_HD uint64_t randstrobe_hash(uint64_t h1, uint64_t h2, uint64_t f1, uint64_t f2, const SeedParams& p) {
    return ((h1 & p.main_mask) | (h2 & ~p.fwd_main_mask) | (f1 << p.partial_pos) | (f2 << 8)) & (REF_HASH_MASK << 1);
}

// Syncmer stored in scratch: hash and (position << 1 | is_forward).
struct Syncmer {
    uint64_t hash;
    uint32_t pos_fwd;
    uint32_t pad;
};

// Second strobe for syncmer a of n (src/seeding/strobes.rs:145-176); `at(i)` is syncmer i.
template <typename At>
SA_HD uint32_t pick_strobe2(uint32_t a, uint32_t n, const SeedParams& p, At at) {
    uint32_t avail = n - a;
    if (avail > (uint32_t)p.w_max + 1) avail = (uint32_t)p.w_max + 1;
    Syncmer s1 = at(a);
    uint32_t max_pos = (s1.pos_fwd >> 1) + (uint32_t)p.max_dist;
    uint32_t best = a, min_val = 0xFFFFFFFFu;
    for (uint32_t i = (uint32_t)p.w_min; i < avail; ++i) {
        Syncmer c = at(a + i);
        if ((c.pos_fwd >> 1) > max_pos) break;
        uint32_t ones = (uint32_t)__builtin_popcountll((s1.hash ^ c.hash) & p.q);
        if (ones < min_val) { min_val = ones; best = a + i; }
    }
    return best;
}

// Query randstrobes of one read without a syncmer list (same output as
// read_randstrobes): the last RING syncmers are kept in a register shift
// register (slot j = the j-th most recent). When syncmer m arrives, the forward
// randstrobe of syncmer m - w_max is complete (its candidates are slots
// w_max - i), and the reverse-complement randstrobe whose first strobe is m can
// be formed at once (its candidates are the earlier syncmers m - i, slot i;
// src/seeding/mod.rs:53-57 reverses the list). After the last syncmer, w_max
// sentinel steps (position "infinitely far", which ends a candidate scan like the
// end of the list) flush the forward randstrobes of the tail. Reverse-complement
// records are produced last first and reversed in place at the end.
template <int RING>
SA_HD void read_randstrobes_stream(const uint8_t* seq, uint32_t len, const SeedParams& p, QRand* out0, QRand* out1,
                                   uint32_t* n0, uint32_t* n1) {
    *n0 = 0;
    *n1 = 0;
    if (len < (uint32_t)p.w_max) return;
    const uint32_t k = (uint32_t)p.k;
    const uint32_t w_min = (uint32_t)p.w_min, w_max = (uint32_t)p.w_max, max_dist = (uint32_t)p.max_dist;
    uint64_t rh[RING];
    uint32_t rp[RING];  // position << 1 | is_forward; 0xFFFFFFFF = sentinel
SA_UNROLL
    for (int j = 0; j < RING; ++j) {
        rh[j] = 0;
        rp[j] = 0xFFFFFFFFu;
    }
    uint32_t m = 0, nf = 0;
    auto fwd_strobe = [&]() {
        // randstrobe of syncmer a = m - w_max (slot w_max): candidates a + i = slot w_max - i
        uint64_t h1 = 0;
        uint32_t p1 = 0;
SA_UNROLL
        for (int j = 0; j < RING; ++j)
            if ((uint32_t)j == w_max) { h1 = rh[j]; p1 = rp[j]; }
        const uint32_t max_pos = (p1 >> 1) + max_dist;
        uint32_t best_slot = w_max, min_val = 0xFFFFFFFFu;
        bool stop = false;
SA_UNROLL
        for (int j = RING - 1; j >= 0; --j) {
            // slot j holds candidate i = w_max - j
            if ((uint32_t)j < w_max && w_max - (uint32_t)j >= w_min && !stop) {
                if (rp[j] == 0xFFFFFFFFu || (rp[j] >> 1) > max_pos) {
                    stop = true;
                } else {
                    const uint32_t ones = (uint32_t)__builtin_popcountll((h1 ^ rh[j]) & p.q);
                    if (ones < min_val) { min_val = ones; best_slot = (uint32_t)j; }
                }
            }
        }
        uint64_t h2 = 0;
        uint32_t p2 = 0;
SA_UNROLL
        for (int j = 0; j < RING; ++j)
            if ((uint32_t)j == best_slot) { h2 = rh[j]; p2 = rp[j]; }
        const uint64_t f1 = p1 & 1, f2 = p2 & 1;
        out0[nf++] = QRand{randstrobe_hash(h1, h2, f1, f2, p), randstrobe_hash(h2, h1, f1 ^ 1, f2 ^ 1, p), p1 >> 1,
                           (p2 >> 1) + k};
    };
    auto push = [&](uint64_t h, uint32_t pf) {
SA_UNROLL
        for (int j = RING - 1; j > 0; --j) {
            rh[j] = rh[j - 1];
            rp[j] = rp[j - 1];
        }
        rh[0] = h;
        rp[0] = pf;
    };
    auto on_syncmer = [&](uint64_t h, bool fwd, uint32_t pos) {
        push(h, (pos << 1) | (fwd ? 1u : 0u));
        // reverse-complement randstrobe with first strobe m: candidates m - i = slot i
        const uint32_t max_gap = max_dist;  // pos(m) - pos(m - i) > max_dist ends the scan
        uint32_t best_slot = 0, min_val = 0xFFFFFFFFu;
        bool stop = false;
SA_UNROLL
        for (int j = 1; j < RING; ++j) {
            if ((uint32_t)j >= w_min && (uint32_t)j <= w_max && (uint32_t)j <= m && !stop) {
                if (pos - (rp[j] >> 1) > max_gap) {
                    stop = true;
                } else {
                    const uint32_t ones = (uint32_t)__builtin_popcountll((h ^ rh[j]) & p.q);
                    if (ones < min_val) { min_val = ones; best_slot = (uint32_t)j; }
                }
            }
        }
        uint64_t h2 = h;
        uint32_t p2 = rp[0];
SA_UNROLL
        for (int j = 1; j < RING; ++j)
            if ((uint32_t)j == best_slot) { h2 = rh[j]; p2 = rp[j]; }
        const uint32_t pos1 = len - pos - k, pos2 = len - (p2 >> 1) - k;
        const uint64_t f1 = (fwd ? 1u : 0u) ^ 1u, f2 = (p2 & 1) ^ 1u;
        out1[m] = QRand{randstrobe_hash(h, h2, f1, f2, p), randstrobe_hash(h2, h, f1 ^ 1, f2 ^ 1, p), pos1, pos2 + k};
        if (m >= w_max) fwd_strobe();
        ++m;
    };
    if (p.k - p.s + 1 == 5) syncmer_scan_w<5>(len, p, [&](uint32_t i) { return nt2(seq[i]); }, on_syncmer);
    else syncmer_scan(len, p, [&](uint32_t i) { return nt2(seq[i]); }, on_syncmer);
    const uint32_t n = m;
    // flush: the forward randstrobes of the last min(n, w_max) syncmers
    for (uint32_t e = 0; e < w_max; ++e) {
        push(0, 0xFFFFFFFFu);
        if (m >= w_max && m - w_max < n) fwd_strobe();
        ++m;
    }
    for (uint32_t x = 0; x < n / 2; ++x) {
        const QRand t = out1[x];
        out1[x] = out1[n - 1 - x];
        out1[n - 1 - x] = t;
    }
    *n0 = n;
    *n1 = n;
}

// Query randstrobes of one read (src/seeding/mod.rs:19-75). `sy` is scratch
// for up to len syncmers; out[0]/out[1] receive up to len entries each.
SA_HD void read_randstrobes(const uint8_t* seq, uint32_t len, const SeedParams& p, Syncmer* sy, QRand* out0,
                            QRand* out1, uint32_t* n0, uint32_t* n1) {
    *n0 = 0;
    *n1 = 0;
    if (len < (uint32_t)p.w_max) return;
    uint32_t n = 0;
    syncmer_scan(len, p, [&](uint32_t i) { return nt2(seq[i]); },
                 [&](uint64_t h, bool fwd, uint32_t pos) { sy[n++] = Syncmer{h, (pos << 1) | (fwd ? 1u : 0u), 0}; });
    const uint32_t k = (uint32_t)p.k;
    for (uint32_t a = 0; a < n; ++a) {
        uint32_t b = pick_strobe2(a, n, p, [&](uint32_t i) { return sy[i]; });
        const Syncmer &s1 = sy[a], &s2 = sy[b];
        uint64_t f1 = s1.pos_fwd & 1, f2 = s2.pos_fwd & 1;
        out0[a] = QRand{randstrobe_hash(s1.hash, s2.hash, f1, f2, p), randstrobe_hash(s2.hash, s1.hash, f1 ^ 1, f2 ^ 1, p),
                        s1.pos_fwd >> 1, (s2.pos_fwd >> 1) + k};
    }
    // reverse complement: syncmers reversed, position L-pos-k, orientation toggled
    auto rc_at = [&](uint32_t i) {
        Syncmer s = sy[n - 1 - i];
        uint32_t pos = len - (s.pos_fwd >> 1) - k;
        return Syncmer{s.hash, (pos << 1) | ((s.pos_fwd & 1) ^ 1), 0};
    };
    for (uint32_t a = 0; a < n; ++a) {
        uint32_t b = pick_strobe2(a, n, p, rc_at);
        Syncmer s1 = rc_at(a), s2 = rc_at(b);
        uint64_t f1 = s1.pos_fwd & 1, f2 = s2.pos_fwd & 1;
        out1[a] = QRand{randstrobe_hash(s1.hash, s2.hash, f1, f2, p), randstrobe_hash(s2.hash, s1.hash, f1 ^ 1, f2 ^ 1, p),
                        s1.pos_fwd >> 1, (s2.pos_fwd >> 1) + k};
    }
    *n0 = n;
    *n1 = n;
}

// --------------------------------------------------------------- index ---
struct IndexView {
    const RefEntry* e;
    const uint64_t* bs;  // bucket_starts, 2^bits + 1 entries
    SeedParams p;
    // reads made through this copy (the lookup stage's algorithmic bytes)
    mutable uint32_t bucket_reads = 0, entry_reads = 0;
};

SA_HD uint64_t ehash(const IndexView& ix, uint64_t i) {
    ++ix.entry_reads;
    return ix.e[i].hash_offset & REF_HASH_MASK;
}

// Lower bound of `masked` among entries [lo, hi) under `mask` (src/partition.rs:27-86).
SA_HD uint64_t lower_bound_masked(const IndexView& ix, uint64_t lo, uint64_t hi, uint64_t mask, uint64_t masked) {
    while (lo < hi) {
        uint64_t mid = lo + (hi - lo) / 2;
        if ((ehash(ix, mid) & mask) < masked) lo = mid + 1; else hi = mid;
    }
    return lo;
}

SA_HD uint64_t upper_bound_masked(const IndexView& ix, uint64_t lo, uint64_t hi, uint64_t mask, uint64_t masked) {
    while (lo < hi) {
        uint64_t mid = lo + (hi - lo) / 2;
        if ((ehash(ix, mid) & mask) <= masked) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// First entry whose masked hash equals hash&mask, searching the bucket of `hash`
// from `start` (or the bucket start); NONE64 when absent (src/index.rs:159-197).
// *bucket_end (if given) receives the end of that bucket.
SA_HD uint64_t get_masked_from(const IndexView& ix, uint64_t hash, uint64_t mask, uint64_t start,
                               uint64_t* bucket_end = nullptr) {
    uint64_t masked = hash & mask;
    uint64_t top = hash >> (64 - ix.p.bits);
    uint64_t ps = start == NONE64 ? ix.bs[top] : start;
    uint64_t pe = ix.bs[top + 1];
    if (bucket_end) *bucket_end = pe;
    ix.bucket_reads += start == NONE64 ? 2 : 1;
    if (ps >= pe) return NONE64;
    if (pe - ps < 4) {
        for (uint64_t i = ps; i < pe; ++i) {
            uint64_t h = ehash(ix, i) & mask;
            if (h == masked) return i;
            if (h > masked) return NONE64;
        }
        return NONE64;
    }
    uint64_t pos = lower_bound_masked(ix, ps, pe, mask, masked);
    if (pos < pe && (ehash(ix, pos) & mask) == masked) return pos;
    return NONE64;
}

// src/index.rs:272-295
SA_HD uint64_t get_count(const IndexView& ix, uint64_t position, uint64_t mask) {
    uint64_t key = ehash(ix, position);
    uint64_t mk = key & mask;
    uint64_t top = key >> (64 - ix.p.bits);
    uint64_t pe = ix.bs[top + 1];
    ++ix.bucket_reads;
    if (pe - position < 8) {
        uint64_t count = 1;
        for (uint64_t i = position + 1; i < pe; ++i) {
            if ((ehash(ix, i) & mask) == mk) ++count; else break;
        }
        return count;
    }
    return upper_bound_masked(ix, position, pe, mask, mk) - position;
}

// src/index.rs:254-266
SA_HD uint64_t get_count_full(const IndexView& ix, uint64_t position, uint64_t hash_rc) {
    uint64_t rc = get_masked_from(ix, hash_rc, REF_HASH_MASK, NONE64);
    uint64_t rev = rc == NONE64 ? 0 : get_count(ix, rc, REF_HASH_MASK);
    return rev + get_count(ix, position, REF_HASH_MASK);
}

// src/index.rs:298-307 (the guard uses filter_cutoff, the probe uses cutoff).
// bucket_end: the end of position's bucket. An entry at or past it lies in a
// later bucket, whose entries differ from position's in the hash's top bits,
// so the probe is not read there (it cannot compare equal).
SA_HD bool too_frequent_forward(const IndexView& ix, uint64_t position, uint64_t cutoff, uint64_t bucket_end) {
    if (position + ix.p.filter_cutoff < ix.p.n_entries) {
        if (position + cutoff >= bucket_end) return false;
        return ehash(ix, position) == ehash(ix, position + cutoff);
    }
    return false;
}

// src/index.rs:309-323 (bucket_end as for too_frequent_forward)
SA_HD bool too_frequent(const IndexView& ix, uint64_t position, uint64_t cutoff, uint64_t hash_rc,
                        uint64_t bucket_end) {
    if (too_frequent_forward(ix, position, cutoff, bucket_end)) return true;
    uint64_t rc_end = 0;
    uint64_t rc = get_masked_from(ix, hash_rc, REF_HASH_MASK, NONE64, &rc_end);
    if (rc != NONE64) {
        if (too_frequent_forward(ix, rc, cutoff, rc_end)) return true;
        return get_count(ix, position, REF_HASH_MASK) + get_count(ix, rc, REF_HASH_MASK) > cutoff;
    }
    return false;
}

// src/index.rs:325-338. When the bucket bits lie within main_mask's leading
// ones, entries of different buckets differ under main_mask too, so a probe
// past position's bucket (bucket_end) compares unequal without being read.
SA_HD bool too_frequent_partial(const IndexView& ix, uint64_t position, uint64_t cutoff, uint64_t bucket_end) {
    if (position + cutoff < ix.p.n_entries) {
        const uint64_t m = ix.p.main_mask;
        const bool bucket_in_mask = ~m == 0 || (uint64_t)ix.p.bits <= (uint64_t)__builtin_clzll(~m);
        if (bucket_in_mask && position + cutoff >= bucket_end) return false;
        return (ehash(ix, position) & m) == (ehash(ix, position + cutoff) & m);
    }
    return false;
}

// Number of consecutive entries from `pos` whose masked hash equals entry pos's
// (galloping search; the runs the chainer walks in src/chainer.rs:304-360).
SA_HD uint64_t run_length(const IndexView& ix, uint64_t pos, uint64_t mask) {
    uint64_t key = ehash(ix, pos) & mask, n = ix.p.n_entries;
    uint64_t step = 1, lo = pos + 1, hi = pos + 1;
    while (hi < n && (ehash(ix, hi) & mask) == key) {
        lo = hi + 1;
        hi = pos + (step <<= 1);
        if (hi > n) hi = n;
    }
    if (hi >= n) hi = n;
    return upper_bound_masked(ix, lo, hi, mask, key) - pos;
}

// ----------------------------------------------------------------- hits ---
constexpr uint32_t HIT_PARTIAL = 1, HIT_FILTERED = 2;

// Count used by rescue (src/hit.rs:102-107), cached in the caller's array.
SA_HD uint64_t hit_count(const IndexView& ix, const Hit& h) {
    return (h.flags & HIT_PARTIAL) ? get_count(ix, h.position, ix.p.main_mask) : get_count_full(ix, h.position, h.hash_rc);
}

// Un-filter the `to_rescue` hits in [start, end) with the smallest counts <= threshold,
// ties by position in the list (stable sort, src/hit.rs:90-125). `cnt` caches counts.
SA_HD uint32_t rescue_least_frequent(const IndexView& ix, Hit* hits, uint64_t* cnt, uint32_t start, uint32_t end,
                                     uint32_t to_rescue, uint64_t threshold) {
    for (uint32_t i = start; i < end; ++i)
        if (cnt[i] == NONE64) cnt[i] = hit_count(ix, hits[i]);
    uint32_t rescued = 0;
    uint64_t last_c = 0;
    int64_t last_i = -1;
    for (uint32_t r = 0; r < to_rescue; ++r) {
        int64_t bi = -1;
        uint64_t bc = 0;
        for (uint32_t i = start; i < end; ++i) {
            uint64_t c = cnt[i];
            if (c > threshold) continue;
            bool after = last_i < 0 || c > last_c || (c == last_c && (int64_t)i > last_i);
            if (!after) continue;
            if (bi < 0 || c < bc) { bi = i; bc = c; }
        }
        if (bi < 0) break;
        rescued += (hits[bi].flags & HIT_FILTERED) ? 1 : 0;
        hits[bi].flags &= ~HIT_FILTERED;
        last_c = bc;
        last_i = bi;
    }
    return rescued;
}

// find_hits for one orientation (src/hit.rs:129-333). Returns number of hits.
SA_HD uint32_t find_hits(const IndexView& ix, const QRand* qr, uint32_t n, const MapParams& mp, Hit* hits,
                         uint64_t* cnt, HitCounts& d) {
    d = HitCounts{0, 0, 0, 0, 0, 0, 0, 0};
    const uint64_t cutoff = ix.p.filter_cutoff;
    const uint32_t k = (uint32_t)ix.p.k;
    uint32_t nh = 0;
    auto partial = [&](const QRand& r) {
        uint64_t pend = 0;
        uint64_t pos = get_masked_from(ix, r.hash, ix.p.main_mask, NONE64, &pend);
        if (pos != NONE64) {
            bool f = too_frequent_partial(ix, pos, cutoff, pend);
            if (f) d.partial_filtered++; else d.partial_found++;
            hits[nh++] = Hit{r.hash, r.hash_rc, pos, NONE64, 0, r.start, r.start + k, HIT_PARTIAL | (f ? HIT_FILTERED : 0u)};
        } else {
            d.partial_not_found++;
        }
    };
    if (mp.mcs != 3) {
        for (uint32_t i = 0; i < n; ++i) {
            const QRand& r = qr[i];
            uint64_t pend = 0;
            uint64_t pos = get_masked_from(ix, r.hash, REF_HASH_MASK, NONE64, &pend);
            if (pos != NONE64) {
                bool f = too_frequent(ix, pos, cutoff, r.hash_rc, pend);
                if (f) d.full_filtered++; else d.full_found++;
                hits[nh++] = Hit{r.hash, r.hash_rc, pos, NONE64, 0, r.start, r.end, f ? HIT_FILTERED : 0u};
            } else {
                d.full_not_found++;
                if (mp.mcs == 0) partial(r);
            }
        }
    }
    if (mp.mcs == 3 || (mp.mcs == 1 && d.full_filtered + d.full_found == 0))
        for (uint32_t i = 0; i < n; ++i) partial(qr[i]);
    for (uint32_t i = 0; i < nh; ++i) cnt[i] = NONE64;
    uint32_t total = d.partial_filtered + d.partial_found + d.full_filtered + d.full_found;
    uint32_t nonrep = d.full_found + d.partial_found;
    float frac = total > 0 ? (float)nonrep / (float)total : 1.0f;
    uint32_t rd = mp.rescue_distance;
    if (rd > 0) {
        if (frac < 0.7f && nonrep < 5) d.rescued += rescue_least_frequent(ix, hits, cnt, 0, nh, 5, 1000);
        // local rescue (src/hit.rs:232-267)
        uint32_t last_start = 0, first_filtered = 0;
        // The interval list depends only on the filtered flags before local rescue; rescuing
        // inside an interval never changes the flags that delimit later intervals (an interval
        // only spans filtered hits), so processing each interval as soon as it is found is
        // equivalent to collecting them first.
        for (uint32_t i = 0; i < nh; ++i) {
            if (hits[i].flags & HIT_FILTERED) continue;
            if (hits[i].query_start > last_start + rd) {
                uint32_t to = (hits[i].query_start - last_start) / rd;
                d.rescued += rescue_least_frequent(ix, hits, cnt, first_filtered, i, to, 10000);
            }
            last_start = hits[i].query_start;
            first_filtered = i + 1;
        }
        if (nh > 0 && hits[nh - 1].query_start - last_start > rd) {
            uint32_t to = (hits[nh - 1].query_start - last_start) / rd;
            d.rescued += rescue_least_frequent(ix, hits, cnt, first_filtered, nh, to, 10000);
        }
    }
    d.n_hits = nh;
    return nh;
}

// Anchor run of each unfiltered hit and the anchor capacity they need.
SA_HD uint64_t hit_runs(const IndexView& ix, Hit* hits, uint32_t nh) {
    uint64_t cap = 0;
    for (uint32_t i = 0; i < nh; ++i) {
        Hit& h = hits[i];
        h.run_start = NONE64;
        h.run_len = 0;
        if (h.flags & HIT_FILTERED) continue;
        if (h.flags & HIT_PARTIAL) {
            uint64_t fe = get_masked_from(ix, h.hash, ix.p.fwd_main_mask, h.position);
            if (fe == NONE64) continue;
            h.run_start = fe;
            h.run_len = (uint32_t)run_length(ix, fe, ix.p.fwd_main_mask);
            cap += h.run_len;
        } else {
            h.run_start = h.position;
            h.run_len = (uint32_t)run_length(ix, h.position, REF_HASH_MASK);
            cap += 2ull * h.run_len;
        }
    }
    return cap;
}

// src/hit.rs:76-81
SA_HD bool better_orientation(const HitCounts& a, const HitCounts& b) {
    uint32_t t = a.full_found + a.full_filtered, ot = b.full_found + b.full_filtered;
    return t >= ot * 2 && t > ot + 5;
}

// ------------------------------------------------------------ chaining ---
struct Anchor {
    uint64_t rk;  // ref_id << 32 | ref_start
    uint32_t q;   // query_start
    uint32_t pad;
};

SA_HD bool anchor_less(const Anchor& a, const Anchor& b) { return a.rk < b.rk || (a.rk == b.rk && a.q < b.q); }

template <typename T, typename Less>
SA_HD void heap_sort(T* a, uint32_t n, Less less) {
    auto sift = [&](uint32_t root, uint32_t end) {
        while (true) {
            uint32_t child = 2 * root + 1;
            if (child >= end) break;
            if (child + 1 < end && less(a[child], a[child + 1])) ++child;
            if (!less(a[root], a[child])) break;
            T t = a[root]; a[root] = a[child]; a[child] = t;
            root = child;
        }
    };
    if (n < 2) return;
    for (uint32_t i = n / 2; i-- > 0;) sift(i, n);
    for (uint32_t end = n - 1; end > 0; --end) {
        T t = a[0]; a[0] = a[end]; a[end] = t;
        sift(0, end);
    }
}

// Anchors of one hit (src/chainer.rs:304-384): a full hit walks its run of
// equal hashes and keeps the entries whose span length differs from the query
// span by no more than the smallest difference so far (in table order), two
// anchors each; a partial hit gives one anchor per entry. Start anchors go to
// s[*ns], end anchors to e[*ne] (hits arrive in query order, so each sequence is
// nearly sorted by reference position on its own); adds the entries read to
// *entries.
SA_HD void hit_anchors(const IndexView& ix, const Hit& h, Anchor* s, uint32_t* ns, Anchor* e, uint32_t* ne,
                       uint32_t* entries) {
    if (h.run_start == NONE64) return;
    const uint32_t k = (uint32_t)ix.p.k;
    *entries += h.run_len;
    if (h.flags & HIT_PARTIAL) {
        for (uint64_t j = h.run_start; j < h.run_start + h.run_len; ++j) {
            const RefEntry& x = ix.e[j];
            s[(*ns)++] = Anchor{((uint64_t)x.ref_index << 32) | x.position, h.query_start, 0};
        }
    } else {
        uint32_t min_diff = 0xFFFFFFFFu;
        const uint32_t ql = h.query_end - h.query_start;
        for (uint64_t j = h.run_start; j < h.run_start + h.run_len; ++j) {
            const RefEntry& x = ix.e[j];
            const uint32_t rs = x.position, re = rs + (uint32_t)(x.hash_offset & 0xFF) + k;
            const uint32_t rl = re - rs;
            const uint32_t diff = ql > rl ? ql - rl : rl - ql;
            if (diff <= min_diff) {
                s[(*ns)++] = Anchor{((uint64_t)x.ref_index << 32) | rs, h.query_start, 0};
                e[(*ne)++] = Anchor{((uint64_t)x.ref_index << 32) | (re - k), h.query_end - k, 0};
                min_diff = diff;
            }
        }
    }
}

// Hits -> anchors for one read orientation: start anchors in s, end anchors in e.
SA_HD void hits_to_anchors(const IndexView& ix, const Hit* hits, uint32_t nh, Anchor* s, uint32_t* ns, Anchor* e,
                           uint32_t* ne, uint32_t* entries) {
    *ns = *ne = 0;
    *entries = 0;
    for (uint32_t i = 0; i < nh; ++i) hit_anchors(ix, hits[i], s, ns, e, ne, entries);
}

// ratio = (float)mx / (float)mn > R, the diagonal-ratio test of
// src/chainer.rs:126-130, without the division: with R = S * 2^e (S its 24-bit
// significand) the correctly rounded quotient exceeds R exactly when mx / mn
// exceeds the midpoint (2S + 1) * 2^(e-1) between R and the next float, or
// equals it and the tie rounds up (S odd); that is mx * 2^(1-e) > or >= mn *
// (2S + 1). Exact while mx < 2^24 (its float conversion exact) and the products
// fit in 64 bits; the division is used otherwise.
struct RatioTest {
    uint64_t mul_mn;   // 2S + 1
    uint32_t shift;    // 1 - e
    uint32_t tie_up;   // S odd
    uint32_t exact;    // this form applies to R at all (R finite, positive, 1 - e in 0..39)
    float R;
};
SA_HD RatioTest make_ratio_test(float R) {
    RatioTest t{0, 0, 0, 0, R};
    if (!(R > 0.0f) || R > 3.0e38f) return t;
    uint32_t bits;
    __builtin_memcpy(&bits, &R, 4);
    const int32_t ex = (int32_t)((bits >> 23) & 0xFF);
    if (ex == 0) return t;  // subnormal
    const uint64_t S = (bits & 0x7FFFFF) | 0x800000;
    const int32_t e = ex - 127 - 23;
    if (1 - e < 0 || 1 - e > 39) return t;
    t.mul_mn = 2 * S + 1;
    t.shift = (uint32_t)(1 - e);
    t.tie_up = (uint32_t)(S & 1);
    t.exact = 1;
    return t;
}
SA_HD bool ratio_exceeds(uint32_t mx, uint32_t mn, const RatioTest& t) {
    if (t.exact && mx < (1u << 24)) {
        const uint64_t lhs = (uint64_t)mx << t.shift, rhs = (uint64_t)mn * t.mul_mn;
        return t.tie_up ? lhs >= rhs && mx > 0 : lhs > rhs;
    }
    return (float)mx / (float)mn > t.R;
}

// src/chainer.rs:286-302; log2(dd+1) comes from a host-filled table (host libm).
SA_HD float chain_score(uint32_t dq, uint32_t dr, uint32_t k, const ChainParams& cp, const float* log2tab) {
    uint32_t dd = dr > dq ? dr - dq : dq - dr;
    uint32_t dg = dq < dr ? dq : dr;
    float score = (float)(k < dg ? k : dg);
    uint32_t dg2 = dg > k ? dg - k : 0;
    float lin = cp.diag_diff_penalty * (float)dd + cp.gap_length_penalty * (float)dg2;
    float lg = dd >= 1 ? log2tab[dd] : 0.0f;
    score -= lin + 0.5f * lg;
    return score;
}

struct ChainScratch {
    Anchor* anchors;
    float* dp;
    uint32_t* pred;
    uint32_t* cand;
    uint8_t* used;
    uint32_t cap;
};

// Collinear chaining + chain extraction for one orientation
// (src/chainer.rs:85-180, 387-448). Writes NAMs (nam_id relative to this
// orientation) and their anchors (last first). Returns the number of NAMs.
SA_HD uint32_t chain_orientation(const ChainScratch& s, uint32_t n, uint32_t k, uint32_t read_len, bool is_revcomp,
                                 const ChainParams& cp, const float* log2tab, const float* same_diag, NamRec* nams,
                                 uint32_t* chain_anchors, uint64_t chain_anchor_base, uint32_t* n_chain_anchors = nullptr) {
    if (n_chain_anchors) *n_chain_anchors = 0;
    if (n == 0) return 0;
    Anchor* A = s.anchors;
    uint64_t max_ref_gap = cp.max_ref_gap >= 0 ? (uint64_t)cp.max_ref_gap : read_len;
    auto score_of = [&](uint32_t dq, uint32_t dr) {
        if (dq == dr && dq < 1024) return same_diag[dq];
        return chain_score(dq, dr, k, cp, log2tab);
    };
    float best = 0.0f;
    uint32_t best_index = NONE32;
    for (uint32_t i = 0; i < n; ++i) {
        float dpi = (float)k;
        uint32_t pi = NONE32;
        uint32_t lookup_end = i > cp.max_lookback ? i - cp.max_lookback : 0;
        const Anchor ai = A[i];
        const uint32_t ai_ref = (uint32_t)(ai.rk >> 32), ai_rs = (uint32_t)ai.rk;
        for (uint32_t j = i; j-- > lookup_end;) {
            const Anchor aj = A[j];
            if ((uint32_t)(aj.rk >> 32) != ai_ref) break;
            if (ai.q < aj.q) continue;
            uint32_t dq = ai.q - aj.q;
            uint32_t dr = ai_rs - (uint32_t)aj.rk;
            if (dr >= max_ref_gap) break;
            uint32_t mx = dq > dr ? dq : dr, mn = dq < dr ? dq : dr;
            float ratio = (float)mx / (float)mn;
            if (ratio > cp.max_diagonal_ratio) continue;
            float ns = s.dp[j] + score_of(dq, dr);
            if (ns >= dpi) {
                dpi = ns;
                pi = j;
                if (dq == dr) break;
            }
        }
        if (best_index != NONE32 && (uint32_t)(A[best_index].rk >> 32) == ai_ref) {
            const Anchor aj = A[best_index];
            if (ai.q > aj.q) {
                uint32_t dq = ai.q - aj.q;
                uint32_t dr = ai_rs - (uint32_t)aj.rk;
                uint32_t mx = dq > dr ? dq : dr, mn = dq < dr ? dq : dr;
                float ratio = (float)mx / (float)mn;
                if (dr < max_ref_gap && dr > 0 && ratio <= cp.max_diagonal_ratio) {
                    float ns = s.dp[best_index] + score_of(dq, dr);
                    if (ns > dpi) { dpi = ns; pi = best_index; }
                }
            }
        }
        s.dp[i] = dpi;
        s.pred[i] = pi;
        if (dpi > best) { best = dpi; best_index = i; }
    }
    // extract chains: candidates with dp >= valid, stable by dp descending
    float valid = best * cp.valid_score_threshold;
    uint32_t nc = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (s.dp[i] >= valid) s.cand[nc++] = i;
        s.used[i] = 0;
    }
    const float* dp = s.dp;
    heap_sort(s.cand, nc, [&](uint32_t a, uint32_t b) { return dp[a] > dp[b] || (dp[a] == dp[b] && a < b); });
    uint32_t nn = 0;
    uint64_t ca = 0;  // chain anchors written
    for (uint32_t c = 0; c < nc; ++c) {
        uint32_t i = s.cand[c];
        if (s.used[i]) continue;
        uint32_t j = i;
        bool overlaps = false;
        uint64_t start = ca;
        chain_anchors[2 * ca] = (uint32_t)A[i].rk;
        chain_anchors[2 * ca + 1] = A[i].q;
        ++ca;
        uint32_t matching = k, cov = (uint32_t)A[i].rk;
        while (s.pred[j] != NONE32) {
            j = s.pred[j];
            if (s.used[j]) { overlaps = true; break; }
            chain_anchors[2 * ca] = (uint32_t)A[j].rk;
            chain_anchors[2 * ca + 1] = A[j].q;
            ++ca;
            s.used[j] = 1;
            uint32_t rj = (uint32_t)A[j].rk;
            uint32_t gap = cov > rj ? cov - rj : 0;
            matching += gap < k ? gap : k;
            cov = rj;
        }
        if (overlaps) { ca = start; continue; }
        NamRec r;
        r.nam_id = nn;
        r.query_start = A[j].q;
        r.query_end = A[i].q + k;
        r.ref_start = (uint32_t)A[j].rk;
        r.ref_end = (uint32_t)A[i].rk + k;
        r.matching_bases = matching;
        r.ref_id = (uint32_t)(A[i].rk >> 32);
        r.n_anchors = (uint32_t)(ca - start);
        r.score = dp[i] + (float)r.n_anchors * cp.matches_weight;
        r.flags = is_revcomp ? 1u : 0u;
        r.anchor_off = chain_anchor_base + start;
        nams[nn++] = r;
    }
    if (n_chain_anchors) *n_chain_anchors = (uint32_t)ca;
    return nn;
}

SA_HD bool anchors_ascending(const Anchor* a, uint32_t n) {
    for (uint32_t i = 1; i < n; ++i)
        if (anchor_less(a[i], a[i - 1])) return false;
    return true;
}

// Sort + dedup anchors (src/chainer.rs:228-229): the order is total up to equal
// anchors, which dedup removes, so any sort gives the reference's result. Input:
// start anchors a[0, ns) and end anchors tmp[0, ne). When both are ascending (a
// read with one placement per hit), they are merged back to front into a[0, ns +
// ne), which never overwrites an unread start anchor. Otherwise a natural merge
// sort: ascending runs are merged pairwise, ping-ponging between a and tmp, so
// every pass reads and writes each array front to back. The deduplicated anchors
// end in *out (a or tmp); returns their count.
SA_HD uint32_t sort_dedup_anchors(Anchor* a, uint32_t ns, Anchor* tmp, uint32_t ne, Anchor** out) {
    uint32_t n = ns + ne;
    Anchor* src = a;
    Anchor* dst = tmp;
    if (anchors_ascending(a, ns) && anchors_ascending(tmp, ne)) {
        uint32_t x = ns, y = ne, w = n;
        while (y > 0) a[--w] = (x > 0 && anchor_less(tmp[y - 1], a[x - 1])) ? a[--x] : tmp[--y];
    } else {
        for (uint32_t i = 0; i < ne; ++i) a[ns + i] = tmp[i];
        while (n > 1) {
            uint32_t runs = 0;
            for (uint32_t i = 0; i < n;) {
                uint32_t m = i + 1;
                while (m < n && !anchor_less(src[m], src[m - 1])) ++m;
                if (m == n && runs == 0) break;  // a single run: sorted
                uint32_t e = m;
                if (m < n) {
                    ++e;
                    while (e < n && !anchor_less(src[e], src[e - 1])) ++e;
                }
                uint32_t x = i, y = m, w = i;
                while (x < m && y < e) dst[w++] = anchor_less(src[y], src[x]) ? src[y++] : src[x++];
                while (x < m) dst[w++] = src[x++];
                while (y < e) dst[w++] = src[y++];
                ++runs;
                i = e;
            }
            if (runs == 0) break;
            Anchor* t = src;
            src = dst;
            dst = t;
            if (runs == 1) break;
        }
    }
    *out = src;
    if (n == 0) return 0;
    uint32_t m = 1;
    for (uint32_t i = 1; i < n; ++i)
        if (src[i].rk != src[m - 1].rk || src[i].q != src[m - 1].q) src[m++] = src[i];
    return m;
}

// ------------------------------------------------------------- reference ---
struct RefView {
    const uint64_t* words;     // all contigs, each starting on a word boundary
    const uint64_t* word_off;  // first word of each contig
    const uint64_t* len;       // contig lengths
    uint32_t n;
};

SA_HD uint32_t ref_bits(const RefView& r, uint32_t id, uint64_t i) {
    return (uint32_t)(r.words[r.word_off[id] + (i >> 5)] >> ((i & 31) * 2)) & 3;
}
SA_HD uint8_t ref_char(const RefView& r, uint32_t id, uint64_t i) {
    return (uint8_t)((0x54474341u >> (8 * ref_bits(r, id, i))) & 0xFFu);  // "ACGT"
}

// Read character i of the read or of its reverse complement.
SA_HD uint8_t read_char(const uint8_t* seq, uint32_t len, bool rc, uint32_t i) {
    return rc ? comp_ascii(seq[len - 1 - i]) : seq[i];
}

// First and last k-mer of the NAM equal the reference (src/nam.rs:49-66).
SA_HD bool nam_consistent(const RefView& ref, const uint8_t* seq, uint32_t len, bool rc, uint32_t ref_id,
                          uint32_t rs, uint32_t re, uint32_t qs, uint32_t qe, uint32_t k) {
    for (uint32_t t = 0; t < k; ++t)
        if (read_char(seq, len, rc, qs + t) != ref_char(ref, ref_id, rs + t)) return false;
    for (uint32_t t = 0; t < k; ++t)
        if (read_char(seq, len, rc, qe - k + t) != ref_char(ref, ref_id, re - k + t)) return false;
    return true;
}

// reverse_nam_if_needed outcome (src/nam.rs:94-135): 0 consistent as is,
// 1 consistent after flipping orientation and query coordinates, 2 neither.
SA_HD uint32_t nam_reverse_state(const RefView& ref, const uint8_t* seq, uint32_t len, const NamRec& n, uint32_t k) {
    bool rc = n.flags & 1;
    if (nam_consistent(ref, seq, len, rc, n.ref_id, n.ref_start, n.ref_end, n.query_start, n.query_end, k)) return 0;
    uint32_t qs = len - n.query_end, qe = len - n.query_start;
    if (nam_consistent(ref, seq, len, !rc, n.ref_id, n.ref_start, n.ref_end, qs, qe, k)) return 1;
    return 2;
}

}  // namespace sa
