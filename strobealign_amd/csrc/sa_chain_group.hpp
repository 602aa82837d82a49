// Chaining of one read orientation per lane group, in LDS (device only).
//
// src/chainer.rs:182-267 turns the hits of a read orientation into anchors
// (304-384), sorts and deduplicates them (228-229), runs the collinear-chaining
// DP (85-180) and extracts the chains as NAMs (387-448). Here G consecutive
// lanes of a wave (G = 16: four orientations per wave; G = 64: one) do all of
// that for one orientation with every intermediate array in LDS:
//   - hits -> anchors: a lane per hit walks the hit's run of index entries (the
//     length-difference filter is sequential within a run), a group prefix sum
//     places its anchors;
//   - sort: a bitonic sort of the (ref_id, ref_start, query_start) keys, padded
//     to a power of two; dedup: a compaction by group prefix sums;
//   - DP: anchors in order, the lookback window j = i-1, i-2, ... spread over the
//     lanes (chunks of G): break conditions by ballot, the sequential "take if
//     >= the running best, stop after a taken same-diagonal predecessor" rule by a
//     group prefix maximum and two ballots, so the result equals the scalar loop;
//   - extraction: candidates by compaction, their stable order by dp by a rank
//     count, and the greedy backtrack (a pointer chase in LDS) on the group's
//     first lane.
// An orientation is handled here when its anchor capacity (2 per entry of a
// full hit's run, 1 per entry of a partial hit's, computed by the lookups) is
// at most CAP; the kernel of the first tier lists the larger ones for the next
// tiers (a wave per orientation with a larger CAP, then the thread-per-
// orientation form over global memory).
#pragma once
#include <hip/hip_runtime.h>

#include "sa_items.hpp"

namespace sa {

// Orientations deferred by the first chaining tier, per later tier.
struct ChainTiers {
    uint32_t* list[2];   // tier 2 (wave, LDS) and tier 3 (thread, global memory)
    uint32_t* count;     // count[0], count[1]
    uint64_t cap1, cap2; // anchor capacities handled by tiers 1 and 2
};

template <int CAP>
struct ChainLds {
    uint32_t ref[CAP], rs[CAP], q[CAP];
    float dp[CAP];
    uint16_t pred[CAP], cand[CAP], order[CAP];
    uint8_t used[CAP];
};
constexpr uint16_t NONE16 = 0xFFFF;

__device__ __forceinline__ void lds_fence() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Collective operations of one group of G consecutive lanes (every lane of the
// group takes part; control flow around them is uniform within the group).
template <int G>
struct Grp {
    uint32_t r;     // lane within the group
    uint32_t base;  // first lane of the group within the wave
    __device__ __forceinline__ uint64_t ballot(bool p) const {
        const uint64_t m = __ballot(p);
        return G == 64 ? m : (m >> base) & ((1ull << G) - 1);
    }
    __device__ __forceinline__ uint32_t excl_sum(uint32_t v) const {
        uint32_t x = v;
        for (int d = 1; d < G; d <<= 1) {
            const uint32_t y = __shfl_up(x, d, G);
            if (r >= (uint32_t)d) x += y;
        }
        return x - v;
    }
    __device__ __forceinline__ uint32_t sum(uint32_t v) const {
        for (int d = G / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, G);
        return v;
    }
    // maximum over the lanes before this one (-inf for the first)
    __device__ __forceinline__ float excl_max(float v) const {
        float x = v;
        for (int d = 1; d < G; d <<= 1) {
            const float y = __shfl_up(x, d, G);
            if (r >= (uint32_t)d) x = fmaxf(x, y);
        }
        const float e = __shfl_up(x, 1, G);
        return r == 0 ? -__builtin_huge_valf() : e;
    }
    template <typename T>
    __device__ __forceinline__ T bcast(T v, uint32_t src) const {
        return __shfl(v, (int)src, G);
    }
};

SA_HD bool key_less(uint32_t ar, uint32_t as, uint32_t aq, uint32_t br, uint32_t bs, uint32_t bq) {
    return ar < br || (ar == br && (as < bs || (as == bs && aq < bq)));
}

// Anchors of a hit (src/chainer.rs:304-384) into the group's arrays at `pos`, or
// only counted when `write` is false.
template <int CAP>
__device__ uint32_t grp_hit_anchors(const IndexView& ix, const Hit& h, bool write, uint32_t pos, ChainLds<CAP>& L) {
    const uint32_t k = (uint32_t)ix.p.k;
    uint32_t c = 0;
    if (h.flags & HIT_PARTIAL) {
        for (uint64_t j = h.run_start; j < h.run_start + h.run_len; ++j) {
            if (write) {
                const RefEntry x = ix.e[j];
                L.ref[pos + c] = x.ref_index;
                L.rs[pos + c] = x.position;
                L.q[pos + c] = h.query_start;
            }
            ++c;
        }
        return c;
    }
    uint32_t min_diff = 0xFFFFFFFFu;
    const uint32_t ql = h.query_end - h.query_start;
    for (uint64_t j = h.run_start; j < h.run_start + h.run_len; ++j) {
        const RefEntry x = ix.e[j];
        const uint32_t rl = (uint32_t)(x.hash_offset & 0xFF) + k;
        const uint32_t diff = ql > rl ? ql - rl : rl - ql;
        if (diff <= min_diff) {
            if (write) {
                L.ref[pos + c] = x.ref_index;
                L.rs[pos + c] = x.position;
                L.q[pos + c] = h.query_start;
                L.ref[pos + c + 1] = x.ref_index;
                L.rs[pos + c + 1] = x.position + rl - k;
                L.q[pos + c + 1] = h.query_end - k;
            }
            c += 2;
            min_diff = diff;
        }
    }
    return c;
}

// One orientation t on one group: same outputs as item_chain.
template <int G, int CAP>
__device__ void chain_grp(const Grp<G>& gp, ChainLds<CAP>& L, uint32_t t, const IndexView& ix, const ChainParams& cp,
                          const uint64_t* __restrict__ off, const Hit* __restrict__ hits,
                          const HitCounts* __restrict__ hc, const uint64_t* __restrict__ aoff, const ChainBufs& b,
                          uint32_t* __restrict__ nam_n, uint32_t* __restrict__ anchors_raw,
                          const float* __restrict__ log2tab, const float* __restrict__ same_diag) {
    const uint32_t r = gp.r;
    const uint32_t i_read = t >> 1, o = t & 1;
    const uint64_t ro = off[i_read];
    const uint32_t L_read = (uint32_t)(off[i_read + 1] - ro);
    const Hit* H = hits + 2 * ro + (uint64_t)o * L_read;
    const uint32_t nh = hc[t].n_hits;
    const uint64_t ab = aoff[t];
    const uint32_t k = (uint32_t)ix.p.k;

    // 1. hits -> anchors (unsorted; any order sorts to the same list)
    uint32_t na = 0, entries = 0;
    for (uint32_t h0 = 0; h0 < nh; h0 += G) {
        const uint32_t h = h0 + r;
        Hit hh;
        uint32_t c = 0;
        const bool has = h < nh && H[h].run_start != NONE64;
        if (has) {
            hh = H[h];
            entries += hh.run_len;
            c = grp_hit_anchors<CAP>(ix, hh, false, 0, L);
        }
        const uint32_t pos = na + gp.excl_sum(c);
        if (has && c && pos + c <= CAP) grp_hit_anchors<CAP>(ix, hh, true, pos, L);
        na += gp.sum(c);
    }
    entries = gp.sum(entries);
    if (na > CAP) na = CAP;  // capacity bound: cannot be exceeded (2 per entry at most)
    lds_fence();

    // 2. bitonic sort, padded to a power of two with keys above every anchor
    uint32_t P = 1;
    while (P < na) P <<= 1;
    for (uint32_t x = na + r; x < P; x += G) {
        L.ref[x] = 0xFFFFFFFFu;
        L.rs[x] = 0xFFFFFFFFu;
        L.q[x] = 0xFFFFFFFFu;
    }
    lds_fence();
    for (uint32_t k2 = 2; k2 <= P; k2 <<= 1) {
        for (uint32_t j2 = k2 >> 1; j2 > 0; j2 >>= 1) {
            for (uint32_t x = r; x < P / 2; x += G) {
                const uint32_t i1 = ((x & ~(j2 - 1)) << 1) | (x & (j2 - 1));
                const uint32_t i2 = i1 + j2;
                const bool up = (i1 & k2) == 0;
                const uint32_t ar = L.ref[i1], as = L.rs[i1], aq = L.q[i1];
                const uint32_t br = L.ref[i2], bs = L.rs[i2], bq = L.q[i2];
                if (key_less(br, bs, bq, ar, as, aq) == up) {
                    L.ref[i1] = br; L.rs[i1] = bs; L.q[i1] = bq;
                    L.ref[i2] = ar; L.rs[i2] = as; L.q[i2] = aq;
                }
            }
            lds_fence();
        }
    }

    // 3. dedup (equal anchors are adjacent): compaction in place
    uint32_t m = 0;
    uint32_t pr = 0xFFFFFFFFu, ps = 0xFFFFFFFFu, pq = 0xFFFFFFFFu;  // last anchor of the previous chunk
    for (uint32_t x0 = 0; x0 < na; x0 += G) {
        const uint32_t x = x0 + r;
        uint32_t er = 0, es = 0, eq = 0;
        if (x < na) { er = L.ref[x]; es = L.rs[x]; eq = L.q[x]; }
        uint32_t qr_ = __shfl_up(er, 1, G), qs_ = __shfl_up(es, 1, G), qq_ = __shfl_up(eq, 1, G);
        if (r == 0) { qr_ = pr; qs_ = ps; qq_ = pq; }
        const bool keep = x < na && (x == 0 || er != qr_ || es != qs_ || eq != qq_);
        pr = gp.bcast(er, G - 1);
        ps = gp.bcast(es, G - 1);
        pq = gp.bcast(eq, G - 1);
        const uint32_t pos = m + gp.excl_sum(keep ? 1u : 0u);
        lds_fence();
        if (keep) { L.ref[pos] = er; L.rs[pos] = es; L.q[pos] = eq; }
        m += gp.sum(keep ? 1u : 0u);
        lds_fence();
    }

    // 4. collinear chaining DP (src/chainer.rs:85-180)
    const uint64_t max_ref_gap = cp.max_ref_gap >= 0 ? (uint64_t)cp.max_ref_gap : L_read;
    auto score_of = [&](uint32_t dq, uint32_t dr) {
        if (dq == dr && dq < 1024) return same_diag[dq];
        return chain_score(dq, dr, k, cp, log2tab);
    };
    const float NEG = -__builtin_huge_valf();
    float best = 0.0f;
    uint32_t best_i = NONE32, b_ref = 0, b_rs = 0, b_q = 0;
    float b_dp = 0.0f;
    for (uint32_t i = 0; i < m; ++i) {
        const uint32_t ai_ref = L.ref[i], ai_rs = L.rs[i], ai_q = L.q[i];
        float dpi = (float)k;
        uint32_t pi = NONE32;
        const uint32_t lookup_end = i > cp.max_lookback ? i - cp.max_lookback : 0;
        const uint32_t span = i - lookup_end;
        for (uint32_t c0 = 0; c0 < span; c0 += G) {
            const uint32_t l = c0 + r;
            bool brk = false, cand = false, diag = false;
            float ns = NEG;
            if (l < span) {
                const uint32_t j = i - 1 - l;
                if (L.ref[j] != ai_ref) {
                    brk = true;
                } else {
                    const uint32_t aj_q = L.q[j];
                    if (ai_q >= aj_q) {
                        const uint32_t dq = ai_q - aj_q, dr = ai_rs - L.rs[j];
                        if (dr >= max_ref_gap) {
                            brk = true;
                        } else {
                            const uint32_t mx = dq > dr ? dq : dr, mn = dq < dr ? dq : dr;
                            const float ratio = (float)mx / (float)mn;
                            if (!(ratio > cp.max_diagonal_ratio)) {
                                cand = true;
                                ns = L.dp[j] + score_of(dq, dr);
                                diag = dq == dr;
                            }
                        }
                    }
                }
            }
            // the scalar loop visits l = 0, 1, ... and stops at the first break
            const uint64_t bm = gp.ballot(brk);
            const uint32_t fb = bm ? (uint32_t)__builtin_ctzll(bm) : (uint32_t)G;
            cand = cand && r < fb;
            // taken where ns >= the running best (k or an earlier candidate)
            const float before = fmaxf(dpi, gp.excl_max(cand ? ns : NEG));
            const bool taken = cand && ns >= before;
            const uint64_t sm = gp.ballot(taken && diag);  // taken on the same diagonal: the loop stops
            const uint32_t fs = sm ? (uint32_t)__builtin_ctzll(sm) : (uint32_t)G;
            const uint64_t tm = gp.ballot(taken && r <= fs);
            if (tm) {
                const uint32_t last = 63u - (uint32_t)__builtin_clzll(tm);
                dpi = gp.bcast(ns, last);
                pi = i - 1 - (c0 + last);
            }
            if (sm || bm) break;
        }
        if (best_i != NONE32 && b_ref == ai_ref && ai_q > b_q) {
            const uint32_t dq = ai_q - b_q, dr = ai_rs - b_rs;
            const uint32_t mx = dq > dr ? dq : dr, mn = dq < dr ? dq : dr;
            const float ratio = (float)mx / (float)mn;
            if (dr < max_ref_gap && dr > 0 && ratio <= cp.max_diagonal_ratio) {
                const float ns = b_dp + score_of(dq, dr);
                if (ns > dpi) {
                    dpi = ns;
                    pi = best_i;
                }
            }
        }
        if (r == 0) {
            L.dp[i] = dpi;
            L.pred[i] = pi == NONE32 ? NONE16 : (uint16_t)pi;
        }
        if (dpi > best) {
            best = dpi;
            best_i = i;
            b_ref = ai_ref;
            b_rs = ai_rs;
            b_q = ai_q;
            b_dp = dpi;
        }
        lds_fence();
    }

    // 5. chain extraction (src/chainer.rs:387-448): candidates with dp >= the
    // threshold, stable by dp descending, then the greedy backtrack
    const float valid = best * cp.valid_score_threshold;
    uint32_t nc = 0;
    for (uint32_t x0 = 0; x0 < m; x0 += G) {
        const uint32_t x = x0 + r;
        const bool c = x < m && L.dp[x] >= valid;
        const uint32_t pos = nc + gp.excl_sum(c ? 1u : 0u);
        if (c) L.cand[pos] = (uint16_t)x;
        if (x < m) L.used[x] = 0;
        nc += gp.sum(c ? 1u : 0u);
    }
    lds_fence();
    for (uint32_t c = r; c < nc; c += G) {
        const uint32_t x = L.cand[c];
        const float dx = L.dp[x];
        uint32_t rank = 0;
        for (uint32_t d = 0; d < nc; ++d) {
            const float dy = L.dp[L.cand[d]];
            rank += (dy > dx || (dy == dx && d < c)) ? 1u : 0u;
        }
        L.order[rank] = (uint16_t)x;
    }
    lds_fence();
    if (r == 0) {
        uint32_t nn = 0;
        uint64_t ca = 0;
        uint32_t* cha = b.chain_anchors + 4 * ab;
        NamRec* nams = b.nams + ab;
        for (uint32_t c = 0; c < nc; ++c) {
            const uint32_t i = L.order[c];
            if (L.used[i]) continue;
            uint32_t j = i;
            bool overlaps = false;
            const uint64_t start = ca;
            cha[2 * ca] = L.rs[i];
            cha[2 * ca + 1] = L.q[i];
            ++ca;
            uint32_t matching = k, cov = L.rs[i];
            while (L.pred[j] != NONE16) {
                j = L.pred[j];
                if (L.used[j]) {
                    overlaps = true;
                    break;
                }
                const uint32_t rj = L.rs[j];
                cha[2 * ca] = rj;
                cha[2 * ca + 1] = L.q[j];
                ++ca;
                L.used[j] = 1;
                const uint32_t gap = cov > rj ? cov - rj : 0;
                matching += gap < k ? gap : k;
                cov = rj;
            }
            if (overlaps) {
                ca = start;
                continue;
            }
            NamRec rec;
            rec.nam_id = nn;
            rec.query_start = L.q[j];
            rec.query_end = L.q[i] + k;
            rec.ref_start = L.rs[j];
            rec.ref_end = L.rs[i] + k;
            rec.matching_bases = matching;
            rec.ref_id = L.ref[i];
            rec.n_anchors = (uint32_t)(ca - start);
            rec.score = L.dp[i] + (float)rec.n_anchors * cp.matches_weight;
            rec.flags = o ? 1u : 0u;
            rec.anchor_off = 2 * ab + start;
            nams[nn++] = rec;
        }
        nam_n[t] = nn;
        anchors_raw[t] = na;
        b.alg_bytes[t] = (uint32_t)sizeof(Hit) * nh + (uint32_t)sizeof(RefEntry) * entries +
                         (uint32_t)sizeof(NamRec) * nn + 8u * (uint32_t)ca;
    }
    lds_fence();
}

// Orientation selection of src/chainer.rs:206-216: false if the other
// orientation of the read is chained alone.
SA_HD bool orientation_chained(const HitCounts* hc, uint32_t t) {
    const uint32_t i = t >> 1, o = t & 1;
    const HitCounts& h0 = hc[2 * i];
    const HitCounts& h1 = hc[2 * i + 1];
    if (better_orientation(h0, h1)) return o == 0;
    if (better_orientation(h1, h0)) return o == 1;
    return true;
}

}  // namespace sa
