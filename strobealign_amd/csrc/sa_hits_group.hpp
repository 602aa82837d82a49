// find_hits of one read orientation per lane group (device only).
//
// src/hit.rs:129-333 looks up a read orientation's randstrobes one after another
// and then rescues filtered hits. Every lookup of a randstrobe (bucket words,
// the search in the bucket, the repetitiveness probes of too_frequent) depends
// only on that randstrobe, so G consecutive lanes of a wave run G randstrobes'
// lookup chains side by side instead of one lane running them all in turn:
//   - lookups: a lane per randstrobe; the hits are placed in query order by a
//     group prefix sum over the lanes that found one;
//   - rescue (src/hit.rs:90-125, 232-267): the counts of a range are looked up a
//     lane per hit; each pick of the least frequent hit after the previous pick
//     is a group minimum of (count, index); the local-rescue intervals come from
//     a ballot over the unfiltered hits, walked in order;
//   - anchor runs (hit_runs): a lane per hit.
// Each lane makes exactly the lookups the sequential loop makes for its items,
// so the hits, the details counts and the counted bytes equal item_hits'.
#pragma once
#include <hip/hip_runtime.h>

#include "sa_chain_group.hpp"

namespace sa {

// group minimum of a 64-bit key
template <int G>
__device__ __forceinline__ uint64_t grp_min64(uint64_t v) {
    for (int d = G / 2; d > 0; d >>= 1) {
        const uint64_t o = __shfl_xor(v, d, G);
        v = o < v ? o : v;
    }
    return v;
}

template <int G>
__device__ __forceinline__ uint64_t grp_sum64(uint64_t v) {
    for (int d = G / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, G);
    return v;
}

// One randstrobe's lookups as find_hits makes them: the full lookup (when
// `full`), then the partial one when the full one found nothing and
// `partial_fallback` is set, or directly when `full` is not set. kind: 0 none,
// 1 full found, 2 full filtered, 3 partial found, 4 partial filtered; nf_full /
// nf_partial count the misses.
SA_HD uint32_t lookup_one(const IndexView& ix, const QRand& r, bool full, bool partial_fallback, Hit& h,
                          uint32_t& nf_full, uint32_t& nf_partial) {
    const uint64_t cutoff = ix.p.filter_cutoff;
    bool try_partial = !full;
    if (full) {
        uint64_t pend = 0;
        const uint64_t pos = get_masked_from(ix, r.hash, REF_HASH_MASK, NONE64, &pend);
        if (pos != NONE64) {
            const bool f = too_frequent(ix, pos, cutoff, r.hash_rc, pend);
            h = Hit{r.hash, r.hash_rc, pos, NONE64, 0, r.start, r.end, f ? HIT_FILTERED : 0u};
            return f ? 2u : 1u;
        }
        ++nf_full;
        try_partial = partial_fallback;
    }
    if (try_partial) {
        uint64_t pend = 0;
        const uint64_t pos = get_masked_from(ix, r.hash, ix.p.main_mask, NONE64, &pend);
        if (pos != NONE64) {
            const bool f = too_frequent_partial(ix, pos, cutoff, pend);
            h = Hit{r.hash, r.hash_rc, pos, NONE64, 0, r.start, r.start + (uint32_t)ix.p.k,
                    HIT_PARTIAL | (f ? HIT_FILTERED : 0u)};
            return f ? 4u : 3u;
        }
        ++nf_partial;
    }
    return 0;
}

// rescue_least_frequent over hits [start, end) on the group: returns the number
// of hits un-filtered (every lane gets it).
template <int G>
__device__ uint32_t rescue_grp(const Grp<G>& gp, const IndexView& ix, Hit* hits, uint64_t* cnt, uint32_t start,
                               uint32_t end, uint32_t to_rescue, uint64_t threshold) {
    for (uint32_t i = start + gp.r; i < end; i += G)
        if (cnt[i] == NONE64) cnt[i] = hit_count(ix, hits[i]);
    lds_fence();
    uint32_t rescued = 0;
    // (count, index) of the last pick; the next pick is the smallest key above it
    uint64_t last = 0;
    bool first = true;
    for (uint32_t x = 0; x < to_rescue; ++x) {
        uint64_t best = NONE64;
        for (uint32_t i = start + gp.r; i < end; i += G) {
            const uint64_t c = cnt[i];
            if (c > threshold) continue;
            const uint64_t key = (c << 32) | i;  // c <= threshold < 2^32
            if ((first || key > last) && key < best) best = key;
        }
        best = grp_min64<G>(best);
        if (best == NONE64) break;
        const uint32_t bi = (uint32_t)best;
        // every lane reads the flags before the owner lane clears them
        const uint32_t fl = hits[bi].flags;
        rescued += (fl & HIT_FILTERED) ? 1u : 0u;
        lds_fence();
        if (gp.r == (bi - start) % G) hits[bi].flags = fl & ~HIT_FILTERED;
        last = best;
        first = false;
    }
    lds_fence();
    return rescued;
}

// find_hits + hit_runs for one orientation on the group (all lanes get nh, d
// and the anchor capacity).
template <int G>
__device__ uint32_t find_hits_grp(const Grp<G>& gp, const IndexView& ix, const QRand* qr, uint32_t n,
                                  const MapParams& mp, Hit* hits, uint64_t* cnt, HitCounts& d, uint64_t& acap) {
    uint32_t nh = 0;
    uint32_t c_found[5] = {0, 0, 0, 0, 0}, nf_full = 0, nf_partial = 0;
    auto pass = [&](bool full, bool partial_fallback) {
        for (uint32_t c0 = 0; c0 < n; c0 += G) {
            const uint32_t i = c0 + gp.r;
            Hit h;
            uint32_t kind = 0;
            if (i < n) kind = lookup_one(ix, qr[i], full, partial_fallback, h, nf_full, nf_partial);
            ++c_found[kind];
            const uint32_t at = nh + gp.excl_sum(kind != 0 ? 1u : 0u);
            if (kind != 0) hits[at] = h;
            nh += gp.sum(kind != 0 ? 1u : 0u);
        }
    };
    if (mp.mcs != 3) pass(true, mp.mcs == 0);
    const uint32_t full_any = gp.sum(c_found[1] + c_found[2]);
    if (mp.mcs == 3 || (mp.mcs == 1 && full_any == 0)) pass(false, false);
    d.full_found = gp.sum(c_found[1]);
    d.full_filtered = gp.sum(c_found[2]);
    d.partial_found = gp.sum(c_found[3]);
    d.partial_filtered = gp.sum(c_found[4]);
    d.full_not_found = gp.sum(nf_full);
    d.partial_not_found = gp.sum(nf_partial);
    d.rescued = 0;
    for (uint32_t i = gp.r; i < nh; i += G) cnt[i] = NONE64;
    lds_fence();
    const uint32_t total = d.partial_filtered + d.partial_found + d.full_filtered + d.full_found;
    const uint32_t nonrep = d.full_found + d.partial_found;
    const float frac = total > 0 ? (float)nonrep / (float)total : 1.0f;
    const uint32_t rd = mp.rescue_distance;
    if (rd > 0) {
        if (frac < 0.7f && nonrep < 5) d.rescued += rescue_grp<G>(gp, ix, hits, cnt, 0, nh, 5, 1000);
        // local rescue: the unfiltered hits in order; an interval of filtered hits
        // between two of them is rescued when they lie more than rd apart
        uint32_t last_start = 0, first_filtered = 0;
        for (uint32_t c0 = 0; c0 < nh; c0 += G) {
            const uint32_t i = c0 + gp.r;
            bool unf = false;
            uint32_t qs = 0;
            if (i < nh) {
                const Hit& h = hits[i];
                unf = !(h.flags & HIT_FILTERED);
                qs = h.query_start;
            }
            uint64_t m = gp.ballot(unf);
            while (m) {
                const uint32_t j = (uint32_t)__builtin_ctzll(m);
                m &= m - 1;
                const uint32_t qj = gp.bcast(qs, j);
                if (qj > last_start + rd) {
                    const uint32_t to = (qj - last_start) / rd;
                    d.rescued += rescue_grp<G>(gp, ix, hits, cnt, first_filtered, c0 + j, to, 10000);
                }
                last_start = qj;
                first_filtered = c0 + j + 1;
            }
        }
        if (nh > 0) {
            const uint32_t ql = hits[nh - 1].query_start;
            if (ql - last_start > rd)
                d.rescued += rescue_grp<G>(gp, ix, hits, cnt, first_filtered, nh, (ql - last_start) / rd, 10000);
        }
    }
    d.n_hits = nh;
    // anchor runs of the unfiltered hits
    uint64_t cap = 0;
    for (uint32_t i = gp.r; i < nh; i += G) cap += hit_runs(ix, hits + i, 1);
    acap = grp_sum64<G>(cap);
    return nh;
}

}  // namespace sa
