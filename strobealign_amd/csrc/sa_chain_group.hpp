// Chaining of one read orientation per lane group, in LDS (device only).
//
// src/chainer.rs:182-267 turns the hits of a read orientation into anchors
// (304-384), sorts and deduplicates them (228-229), runs the collinear-chaining
// DP (85-180) and extracts the chains as NAMs (387-448). Here G consecutive
// lanes of a wave (G = 16: four orientations per wave; G = 64: one) do all of
// that for one orientation with every intermediate array in LDS:
//   - hits -> anchors: a lane per hit walks the hit's run of index entries once
//     (the length-difference filter is sequential within a run) and writes its
//     anchors into the slots its capacity reserves, placed by a group prefix sum;
//   - sort: a bitonic sort of the (ref_id, ref_start, query_start) keys, padded
//     to a power of two; dedup: a compaction by group prefix sums;
//   - DP: anchors in order, the lookback window j = i-1, i-2, ... spread over the
//     lanes (chunks of G): break conditions by ballot, the sequential "take if
//     >= the running best, stop after a taken same-diagonal predecessor" rule by a
//     group prefix maximum and two ballots, so the result equals the scalar loop;
//   - extraction: candidates by compaction, their stable order by dp by a rank
//     count, and the greedy backtrack by the whole group, G links at a time
//     along runs of consecutive anchors.
// An orientation is handled here when its anchor capacity (2 per entry of a
// full hit's run, 1 per entry of a partial hit's, computed by the lookups) is
// at most CAP; the kernel of the first tier lists the larger ones for the next
// tiers (a wave per orientation with a larger CAP, then the thread-per-
// orientation form over global memory).
#pragma once
#include <hip/hip_runtime.h>

#include "sa_items.hpp"

namespace sa {

// Orientations deferred by the first chaining tier, per later tier: list[x]
// (count[x] entries) holds those for tier x + 2, whose anchor capacities are
// (cap[x], cap[x + 1]]; the last list goes to the thread-per-orientation form.
constexpr int CHAIN_TIERS = 4;
struct ChainTiers {
    uint32_t* list[CHAIN_TIERS - 1];
    uint32_t* count;
    uint64_t cap[CHAIN_TIERS - 1];  // anchor capacities handled by the LDS tiers 1..3
};



// An anchor in LDS: (ref_id, ref_start, query_start) and its chaining score,
// read and written as one 16-byte access.
struct alignas(16) AnchorL {
    uint32_t ref, rs, q;
    float dp;
};
template <int CAP>
struct ChainLds {
    AnchorL a[CAP];
    uint16_t pred[CAP], cand[CAP], order[CAP];
    uint8_t used[CAP];
};
constexpr uint16_t NONE16 = 0xFFFF;

__device__ __forceinline__ void lds_fence() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// DPP move of a 32-bit value (lanes whose source lies outside the 16-lane row
// keep `old`)
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_u(uint32_t old, uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ float dpp_f(float old, float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old),
                                                                 __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}

// Collective operations of one group of G consecutive lanes (every lane of the
// group takes part; control flow around them is uniform within the group). A
// 16-lane group is a DPP row: its scans and reductions are row shifts and
// butterflies in the VALU; a 64-lane group uses cross-lane permutes.
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
        if constexpr (G == 8) {
            // row shifts cross into the other half-row: lanes below d take 0
            uint32_t y = dpp_u<0x111>(0, x);
            x += r >= 1 ? y : 0u;
            y = dpp_u<0x112>(0, x);
            x += r >= 2 ? y : 0u;
            y = dpp_u<0x114>(0, x);
            x += r >= 4 ? y : 0u;
        } else if constexpr (G == 16) {
            x += dpp_u<0x111>(0, x);  // row_shr:1
            x += dpp_u<0x112>(0, x);  // row_shr:2
            x += dpp_u<0x114>(0, x);  // row_shr:4
            x += dpp_u<0x118>(0, x);  // row_shr:8
        } else {
            for (int d = 1; d < G; d <<= 1) {
                const uint32_t y = __shfl_up(x, d, G);
                if (r >= (uint32_t)d) x += y;
            }
        }
        return x - v;
    }
    __device__ __forceinline__ uint32_t sum(uint32_t v) const {
        if constexpr (G == 8) {
            v += dpp_u<0xB1>(0, v);   // quad_perm [1,0,3,2]
            v += dpp_u<0x4E>(0, v);   // quad_perm [2,3,0,1]
            v += dpp_u<0x141>(0, v);  // row_half_mirror
        } else if constexpr (G == 16) {
            v += dpp_u<0xB1>(0, v);   // quad_perm [1,0,3,2]
            v += dpp_u<0x4E>(0, v);   // quad_perm [2,3,0,1]
            v += dpp_u<0x141>(0, v);  // row_half_mirror
            v += dpp_u<0x140>(0, v);  // row_mirror
        } else {
            for (int d = G / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, G);
        }
        return v;
    }
    // maximum over the lanes before this one (-inf for the first)
    __device__ __forceinline__ float excl_max(float v) const {
        const float NEG = -__builtin_huge_valf();
        float x = v;
        if constexpr (G == 8) {
            float y = dpp_f<0x111>(NEG, x);
            x = fmaxf(x, r >= 1 ? y : NEG);
            y = dpp_f<0x112>(NEG, x);
            x = fmaxf(x, r >= 2 ? y : NEG);
            y = dpp_f<0x114>(NEG, x);
            x = fmaxf(x, r >= 4 ? y : NEG);
            y = dpp_f<0x111>(NEG, x);
            return r >= 1 ? y : NEG;
        } else if constexpr (G == 16) {
            x = fmaxf(x, dpp_f<0x111>(NEG, x));
            x = fmaxf(x, dpp_f<0x112>(NEG, x));
            x = fmaxf(x, dpp_f<0x114>(NEG, x));
            x = fmaxf(x, dpp_f<0x118>(NEG, x));
            return dpp_f<0x111>(NEG, x);
        } else {
            for (int d = 1; d < G; d <<= 1) {
                const float y = __shfl_up(x, d, G);
                if (r >= (uint32_t)d) x = fmaxf(x, y);
            }
            const float e = __shfl_up(x, 1, G);
            return r == 0 ? NEG : e;
        }
    }
    __device__ __forceinline__ float max(float v) const {
        if constexpr (G == 8) {
            v = fmaxf(v, dpp_f<0xB1>(v, v));
            v = fmaxf(v, dpp_f<0x4E>(v, v));
            v = fmaxf(v, dpp_f<0x141>(v, v));
        } else if constexpr (G == 16) {
            v = fmaxf(v, dpp_f<0xB1>(v, v));
            v = fmaxf(v, dpp_f<0x4E>(v, v));
            v = fmaxf(v, dpp_f<0x141>(v, v));
            v = fmaxf(v, dpp_f<0x140>(v, v));
        } else {
            for (int d = G / 2; d > 0; d >>= 1) v = fmaxf(v, __shfl_xor(v, d, G));
        }
        return v;
    }
    template <typename T>
    __device__ __forceinline__ T bcast(T v, uint32_t src) const {
        return __shfl(v, (int)src, G);
    }
    // a value equal on every lane of the group; on a whole wave it is moved to
    // a scalar register, so the code that depends on it runs once per wave
    __device__ __forceinline__ uint32_t uni(uint32_t v) const {
        if constexpr (G == 64) return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
        else return v;
    }
    __device__ __forceinline__ float uni(float v) const {
        if constexpr (G == 64)
            return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
        else return v;
    }
    // lane src's value to every lane of the group (src the same on all of them)
    __device__ __forceinline__ float lane(float v, uint32_t src) const {
        if constexpr (G == 64)
            return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), (int)src));
        else return __shfl(v, (int)(base + src));
    }
};

SA_HD bool key_less(uint32_t ar, uint32_t as, uint32_t aq, uint32_t br, uint32_t bs, uint32_t bq) {
    return ar < br || (ar == br && (as < bs || (as == bs && aq < bq)));
}

// The key above every anchor: the sort's padding, and the slots of the entries
// the length filter drops.
constexpr AnchorL ANCHOR_PAD{0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0.0f};

// Anchors of a hit (src/chainer.rs:304-384) into its slots [pos, pos + its
// capacity) of the group's array: one per entry of a partial hit's run; two per
// entry of a full hit's that passes the length filter, ANCHOR_PAD in the two
// slots of an entry it drops. Returns the anchors written.
template <int CAP>
__device__ uint32_t grp_hit_anchors(const IndexView& ix, const Hit& h, uint32_t pos, ChainLds<CAP>& L) {
    const uint32_t k = (uint32_t)ix.p.k;
    if (h.flags & HIT_PARTIAL) {
        for (uint32_t e = 0; e < h.run_len; ++e) {
            const RefEntry x = ix.e[h.run_start + e];
            L.a[pos + e] = AnchorL{x.ref_index, x.position, h.query_start, 0.0f};
        }
        return h.run_len;
    }
    uint32_t min_diff = 0xFFFFFFFFu, c = 0;
    const uint32_t ql = h.query_end - h.query_start;
    for (uint32_t e = 0; e < h.run_len; ++e) {
        const RefEntry x = ix.e[h.run_start + e];
        const uint32_t rl = (uint32_t)(x.hash_offset & 0xFF) + k;
        const uint32_t diff = ql > rl ? ql - rl : rl - ql;
        AnchorL a0 = ANCHOR_PAD, a1 = ANCHOR_PAD;
        if (diff <= min_diff) {
            a0 = AnchorL{x.ref_index, x.position, h.query_start, 0.0f};
            a1 = AnchorL{x.ref_index, x.position + rl - k, h.query_end - k, 0.0f};
            c += 2;
            min_diff = diff;
        }
        L.a[pos + 2 * e] = a0;
        L.a[pos + 2 * e + 1] = a1;
    }
    return c;
}

__device__ __forceinline__ bool anchor_lt(const AnchorL& a, const AnchorL& b) {
    return a.ref < b.ref || (a.ref == b.ref && (a.rs < b.rs || (a.rs == b.rs && a.q < b.q)));
}

// One orientation t on one group: same outputs as item_chain.
template <int G, int CAP>
__device__ void chain_grp(const Grp<G>& gp, ChainLds<CAP>& L, uint32_t t, const IndexView& ix, const ChainParams& cp,
                          const uint64_t* __restrict__ off, const Hit* __restrict__ hits,
                          const HitCounts* __restrict__ hc, const uint64_t* __restrict__ aoff, const ChainBufs& b,
                          uint32_t* __restrict__ nam_n, uint32_t* __restrict__ anchors_raw,
                          const float* __restrict__ log2tab) {
    const uint32_t r = gp.r;
    t = gp.uni(t);
    const uint32_t i_read = t >> 1, o = t & 1;
    const uint64_t ro = off[i_read];
    const uint32_t L_read = (uint32_t)(off[i_read + 1] - ro);
    const Hit* H = hits + 2 * ro + (uint64_t)o * L_read;
    const uint32_t nh = hc[t].n_hits;
    const uint64_t ab = aoff[t];
    const uint32_t k = (uint32_t)ix.p.k;

    // 1. hits -> anchors (unsorted; any order sorts to the same list), each
    // hit's in the slots its capacity reserves (2 per entry of a full hit's run,
    // 1 per entry of a partial hit's, the capacities the lookups summed), placed
    // by a group prefix sum of the capacities: each entry is read once
    uint32_t na = 0, entries = 0, ncap = 0;
    for (uint32_t h0 = 0; h0 < nh; h0 += G) {
        const uint32_t h = h0 + r;
        Hit hh;
        bool has = false;
        if (h < nh) {
            hh = H[h];
            has = hh.run_start != NONE64;
        }
        const uint32_t c = has ? ((hh.flags & HIT_PARTIAL) ? hh.run_len : 2 * hh.run_len) : 0u;
        const uint32_t pos = ncap + gp.excl_sum(c);
        if (has) {
            entries += hh.run_len;
            if (pos + c <= CAP) {
                na += grp_hit_anchors<CAP>(ix, hh, pos, L);
            } else {
                // beyond the tier's capacity bound, which the lookups' sum of
                // the same capacities keeps the orientation within
                for (uint32_t x = pos; x < CAP; ++x) L.a[x] = ANCHOR_PAD;
            }
        }
        ncap += gp.sum(c);
    }
    entries = gp.sum(entries);
    na = gp.uni(gp.sum(na));
    ncap = gp.uni(ncap < (uint32_t)CAP ? ncap : (uint32_t)CAP);

    // 2. bitonic sort of the capacity slots, padded to a power of two: the
    // anchors come first, the padding of dropped entries after them
    uint32_t P = 1;
    while (P < ncap) P <<= 1;
    for (uint32_t x = ncap + r; x < P; x += G) L.a[x] = ANCHOR_PAD;
    lds_fence();
    for (uint32_t k2 = 2; k2 <= P; k2 <<= 1) {
        for (uint32_t j2 = k2 >> 1; j2 > 0; j2 >>= 1) {
            for (uint32_t x = r; x < P / 2; x += G) {
                const uint32_t i1 = ((x & ~(j2 - 1)) << 1) | (x & (j2 - 1));
                const uint32_t i2 = i1 + j2;
                const bool up = (i1 & k2) == 0;
                const AnchorL A = L.a[i1], B = L.a[i2];
                if (anchor_lt(B, A) == up) {
                    L.a[i1] = B;
                    L.a[i2] = A;
                }
            }
            lds_fence();
        }
    }

    // 3. dedup (equal anchors are adjacent): compaction in place
    uint32_t m = 0;
    AnchorL prev{0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0.0f};  // last anchor of the previous chunk
    for (uint32_t x0 = 0; x0 < na; x0 += G) {
        const uint32_t x = x0 + r;
        AnchorL e{0, 0, 0, 0.0f};
        if (x < na) e = L.a[x];
        uint32_t pr = __shfl_up(e.ref, 1, G), ps = __shfl_up(e.rs, 1, G), pq = __shfl_up(e.q, 1, G);
        if (r == 0) { pr = prev.ref; ps = prev.rs; pq = prev.q; }
        const bool keep = x < na && (x == 0 || e.ref != pr || e.rs != ps || e.q != pq);
        prev.ref = gp.bcast(e.ref, G - 1);
        prev.rs = gp.bcast(e.rs, G - 1);
        prev.q = gp.bcast(e.q, G - 1);
        const uint32_t pos = m + gp.excl_sum(keep ? 1u : 0u);
        lds_fence();
        if (keep) L.a[pos] = e;
        m += gp.sum(keep ? 1u : 0u);
        lds_fence();
    }
    m = gp.uni(m);

    // 4. collinear chaining DP (src/chainer.rs:85-180). What the scalar loop
    // carries from one anchor to the next (the previous anchor and its score,
    // the best anchor) is equal on every lane of the group: on a whole wave it
    // is held in scalar registers and the common steps run as scalar code.
    // A same-diagonal predecessor (dq == dr >= 1: the quotient is exactly 1)
    // passes the ratio test exactly when 1 does, and its score is chain_score's
    // with dd = 0, the value of the precomputed same-diagonal table
    // (src/chainer.rs:63-74).
    const uint64_t max_ref_gap = cp.max_ref_gap >= 0 ? (uint64_t)cp.max_ref_gap : L_read;
    // dr >= max_ref_gap for a 32-bit dr, in 32 bits (no dr reaches a larger gap)
    const bool gap_fits = max_ref_gap <= 0xFFFFFFFFull;
    const uint32_t gap32 = gap_fits ? (uint32_t)max_ref_gap : 0u;
    auto gap_breaks = [&](uint32_t dr) { return gap_fits && dr >= gap32; };
    const RatioTest rt = make_ratio_test(cp.max_diagonal_ratio);
    const bool diag_ratio_ok = !ratio_exceeds(1u, 1u, rt);
    const float NEG = -__builtin_huge_valf();
    float best = 0.0f;
    uint32_t best_i = NONE32;
    uint32_t b_ref = 0, b_rs = 0, b_q = 0;  // the best anchor
    float b_dp = 0.0f;
    uint32_t p_ref = 0, p_rs = 0, p_q = 0;  // anchor i - 1
    float p_dp = 0.0f;
    AnchorL nx = L.a[0];  // anchor i, read one step ahead
    for (uint32_t i = 0; i < m; ++i) {
        const uint32_t ai_ref = gp.uni(nx.ref), ai_rs = gp.uni(nx.rs), ai_q = gp.uni(nx.q);
        if (i + 1 < m) nx = L.a[i + 1];
        float dpi = (float)k;
        uint32_t pi = NONE32;
        const uint32_t lookup_end = i > cp.max_lookback ? i - cp.max_lookback : 0;
        const uint32_t span = i - lookup_end;
        // The scalar loop's first predecessor j = i - 1 settles most steps: a
        // break there ends the loop with nothing taken; taken on the same
        // diagonal, the loop stops.
        bool settled = span == 0;
        if (!settled) {
            const uint32_t dq = ai_q - p_q, dr = ai_rs - p_rs;
            if (p_ref != ai_ref || (ai_q >= p_q && gap_breaks(dr))) {
                settled = true;
            } else if (ai_q >= p_q && dq == dr && diag_ratio_ok) {
                const float ns = p_dp + chain_score(dq, dq, k, cp, log2tab);
                if (ns >= dpi) {
                    dpi = ns;
                    pi = i - 1;
                    settled = true;
                }
            }
        }
        // otherwise the lookback window j = i-1, i-2, ... spread over the lanes
        for (uint32_t c0 = 0; c0 < span && !settled; c0 += G) {
            const uint32_t l = c0 + r;
            bool brk = false, cand = false, diag = false;
            float ns = NEG;
            if (l < span) {
                const AnchorL aj = L.a[i - 1 - l];
                const uint32_t dq = ai_q - aj.q, dr = ai_rs - aj.rs;
                const bool qok = ai_q >= aj.q;
                // the scalar loop: ref_id change breaks; a later query start is
                // skipped; otherwise dr >= max_ref_gap breaks
                brk = aj.ref != ai_ref || (qok && gap_breaks(dr));
                if (!brk && qok) {
                    const uint32_t mx = dq > dr ? dq : dr, mn = dq < dr ? dq : dr;
                    diag = dq == dr;
                    if (!ratio_exceeds(mx, mn, rt)) {
                        cand = true;
                        ns = aj.dp + chain_score(dq, dr, k, cp, log2tab);
                    }
                }
            }
            // the scalar loop visits l = 0, 1, ... and stops at the first break
            const uint64_t bm = gp.ballot(brk);
            const uint32_t fb = bm ? (uint32_t)__builtin_ctzll(bm) : (uint32_t)G;
            cand = cand && r < fb;
            // taken where ns >= the running best (k or an earlier candidate);
            // after a taken same-diagonal predecessor the loop stops
            const float before = fmaxf(dpi, gp.excl_max(cand ? ns : NEG));
            const bool taken = cand && ns >= before;
            const uint64_t sm = gp.ballot(taken && diag);
            const uint32_t fs = sm ? (uint32_t)__builtin_ctzll(sm) : (uint32_t)G;
            const uint64_t tm = gp.ballot(taken && r <= fs);
            if (tm) {
                // the last taken value is the running maximum up to the stop
                const uint32_t last = 63u - (uint32_t)__builtin_clzll(tm);
                dpi = gp.max(cand && r <= fs ? ns : NEG);
                pi = i - 1 - (c0 + last);
            }
            if (sm || bm) break;
        }
        dpi = gp.uni(dpi);
        pi = gp.uni(pi);
        // the best-index re-check; when the loop took the best anchor itself it
        // recomputes the same value, which cannot be strictly larger
        if (best_i != NONE32 && best_i != pi && b_ref == ai_ref && ai_q > b_q) {
            const uint32_t dq = ai_q - b_q, dr = ai_rs - b_rs;
            const uint32_t mx = dq > dr ? dq : dr, mn = dq < dr ? dq : dr;
            // ratio <= R: dq >= 1 and dr >= 1 here, so the quotient is a number
            if (!gap_breaks(dr) && dr > 0 && !ratio_exceeds(mx, mn, rt)) {
                const float ns = b_dp + chain_score(dq, dr, k, cp, log2tab);
                if (ns > dpi) {
                    dpi = ns;
                    pi = best_i;
                }
            }
        }
        if (r == 0) {
            L.a[i].dp = dpi;
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
        p_ref = ai_ref;
        p_rs = ai_rs;
        p_q = ai_q;
        p_dp = dpi;
        lds_fence();
    }

    // 5. chain extraction (src/chainer.rs:387-448): candidates with dp >= the
    // threshold, stable by dp descending, then the greedy backtrack
    const float valid = best * cp.valid_score_threshold;
    uint32_t nc = 0;
    for (uint32_t x0 = 0; x0 < m; x0 += G) {
        const uint32_t x = x0 + r;
        const bool c = x < m && L.a[x].dp >= valid;
        const uint32_t pos = nc + gp.excl_sum(c ? 1u : 0u);
        if (c) L.cand[pos] = (uint16_t)x;
        if (x < m) L.used[x] = 0;
        nc += gp.sum(c ? 1u : 0u);
    }
    lds_fence();
    nc = gp.uni(nc);
    // the stable rank of each candidate (the candidates with a larger score, or
    // an equal one earlier in the list): the group holds the scores of G
    // candidates at a time in registers and every lane counts against them
    for (uint32_t c0 = 0; c0 < nc; c0 += G) {
        const uint32_t c = c0 + r;
        const uint16_t xi = c < nc ? L.cand[c] : (uint16_t)0;
        const float dx = c < nc ? L.a[xi].dp : 0.0f;
        uint32_t rank = 0;
        for (uint32_t d0 = 0; d0 < nc; d0 += G) {
            const float dyl = d0 + r < nc ? L.a[L.cand[d0 + r]].dp : 0.0f;
            const uint32_t dn = nc - d0 < (uint32_t)G ? nc - d0 : (uint32_t)G;
            for (uint32_t s = 0; s < dn; ++s) {
                const float dy = gp.lane(dyl, s);
                rank += (dy > dx || (dy == dx && d0 + s < c)) ? 1u : 0u;
            }
        }
        if (c < nc) L.order[rank] = xi;
    }
    lds_fence();
    // The greedy backtrack, by the group together. From a chain's head the
    // links are followed G nodes at a time where they run through consecutive
    // anchors (pred[x] == x - 1, the common case along a diagonal): lane l
    // takes node j - l. A run ends at the first node whose link goes
    // elsewhere, or before the first node already used: the chain then
    // overlaps an earlier one and is dropped (the nodes it visited stay used).
    uint32_t nn = 0, ca = 0;
    uint32_t* cha = b.chain_anchors + 4 * ab;
    NamRec* nams = b.nams + ab;
    for (uint32_t c = 0; c < nc; ++c) {
        const uint32_t i = gp.uni((uint32_t)L.order[c]);
        if (gp.uni((uint32_t)L.used[i])) continue;
        const AnchorL hd = L.a[i];
        const uint32_t ai_ref = gp.uni(hd.ref), ai_rs = gp.uni(hd.rs), ai_q = gp.uni(hd.q);
        const float ai_dp = gp.uni(hd.dp);
        const uint32_t start = ca;
        if (r == 0) {
            cha[2 * ca] = ai_rs;
            cha[2 * ca + 1] = ai_q;
        }
        ++ca;
        uint32_t matching = k, cov = ai_rs, last_q = ai_q;
        bool overlaps = false;
        for (uint32_t j = gp.uni((uint32_t)L.pred[i]); j != NONE16;) {
            const bool valid = r <= j;
            const uint32_t x = j - r;
            uint32_t px = NONE16, used = 0;
            AnchorL ax{0, 0, 0, 0.0f};
            if (valid) {
                px = L.pred[x];
                used = L.used[x];
                ax = L.a[x];
            }
            const bool link = valid && px != NONE16 && px + 1 == x;
            const uint64_t um = gp.ballot(used != 0), lm = gp.ballot(!link);
            const uint32_t fu = um ? (uint32_t)__builtin_ctzll(um) : (uint32_t)G;
            const uint32_t fl = lm ? (uint32_t)__builtin_ctzll(lm) : (uint32_t)G;
            const bool stop = fu <= fl && fu < (uint32_t)G;  // a used node before the run's end
            const uint32_t R = stop ? fu : (fl + 1 < (uint32_t)G ? fl + 1 : (uint32_t)G);
            // every lane takes part in the shuffle (a lane whose source lane is
            // inactive would read nothing defined); lane 0 then takes cov
            const uint32_t up = __shfl_up(ax.rs, 1, G);
            const uint32_t prev = r == 0 ? cov : up;
            uint32_t g = 0;
            if (r < R) {
                cha[2 * (ca + r)] = ax.rs;
                cha[2 * (ca + r) + 1] = ax.q;
                L.used[x] = 1;
                const uint32_t gap = prev > ax.rs ? prev - ax.rs : 0;
                g = gap < k ? gap : k;
            }
            matching += gp.sum(g);
            ca += R;
            lds_fence();
            if (R) {
                cov = gp.uni(gp.bcast(ax.rs, R - 1));
                last_q = gp.uni(gp.bcast(ax.q, R - 1));
            }
            if (stop) {
                overlaps = true;
                break;
            }
            j = gp.uni(gp.bcast(px, R - 1));
        }
        if (overlaps) {
            ca = start;
            continue;
        }
        if (r == 0) {
            NamRec rec;
            rec.nam_id = nn;
            rec.query_start = last_q;
            rec.query_end = ai_q + k;
            rec.ref_start = cov;
            rec.ref_end = ai_rs + k;
            rec.matching_bases = matching;
            rec.ref_id = ai_ref;
            rec.n_anchors = ca - start;
            rec.score = ai_dp + (float)rec.n_anchors * cp.matches_weight;
            rec.flags = o ? 1u : 0u;
            rec.anchor_off = 2 * ab + start;
            nams[nn] = rec;
        }
        ++nn;
    }
    if (r == 0) {
        nam_n[t] = nn;
        anchors_raw[t] = na;
        b.alg_bytes[t] = (uint32_t)sizeof(Hit) * nh + (uint32_t)sizeof(RefEntry) * entries +
                         (uint32_t)sizeof(NamRec) * nn + 8u * ca;
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
