// Oracle: hit finding, anchors, collinear chaining, NAMs, fastrand RNG.
// TEST INFRASTRUCTURE ONLY (see oracle.hpp).
#include "oracle.hpp"

#include <atomic>
#include <cstdlib>

#include <algorithm>
#include <cmath>

namespace orc {

static const size_t GLOBAL_RESCUE_COUNT = 5;      // src/hit.rs:11
static const size_t LOCAL_RESCUE_THRESHOLD = 10000;  // src/hit.rs:14
static const size_t GLOBAL_RESCUE_THRESHOLD = 1000;  // src/hit.rs:15

// src/hit.rs:90-125
static size_t rescue_least_frequent(const Index& index, std::vector<Hit>& hits, size_t start, size_t end,
                                    size_t to_rescue, size_t threshold) {
    std::vector<std::pair<size_t, size_t>> hc;
    for (size_t i = start; i < end; ++i) {
        size_t cnt = hits[i].is_partial ? index.get_count(hits[i].position, index.params.randstrobe.main_hash_mask)
                                        : index.get_count_full(hits[i].position, hits[i].hash_revcomp);
        if (cnt <= threshold) hc.push_back({i, cnt});
    }
    std::stable_sort(hc.begin(), hc.end(), [](const auto& a, const auto& b) { return a.second < b.second; });
    size_t rescued = 0;
    for (size_t j = 0; j < hc.size() && j < to_rescue; ++j) {
        rescued += hits[hc[j].first].is_filtered ? 1 : 0;
        hits[hc[j].first].is_filtered = false;
    }
    return rescued;
}

// src/hit.rs:129-225
static std::vector<Hit> find_all_hits(const std::vector<QueryRandstrobe>& qr, const Index& index,
                                      size_t filter_cutoff, Mcs mcs, HitsDetails& d) {
    std::vector<Hit> hits;
    const size_t k = (size_t)index.k();
    auto partial_lookup = [&](const QueryRandstrobe& r) {
        size_t pos = index.get_partial(r.hash);
        if (pos != SIZE_MAX) {
            bool f = index.is_too_frequent_partial(pos, filter_cutoff);
            if (f) d.partial_filtered++; else d.partial_found++;
            hits.push_back({r.start, r.start + k, pos, r.hash, r.hash_revcomp, true, f});
        } else {
            d.partial_not_found++;
        }
    };
    if (mcs != Mcs::FirstStrobe) {
        for (const QueryRandstrobe& r : qr) {
            size_t pos = index.get_full_forward(r.hash);
            if (pos != SIZE_MAX) {
                bool f = index.is_too_frequent(pos, filter_cutoff, r.hash_revcomp);
                if (f) d.full_filtered++; else d.full_found++;
                hits.push_back({r.start, r.end, pos, r.hash, r.hash_revcomp, false, f});
            } else {
                d.full_not_found++;
                if (mcs == Mcs::Always) partial_lookup(r);
            }
        }
    }
    if (mcs == Mcs::FirstStrobe || (mcs == Mcs::Rescue && d.full_filtered + d.full_found == 0)) {
        for (const QueryRandstrobe& r : qr) partial_lookup(r);
    }
    return hits;
}

// src/hit.rs:232-267
static size_t rescue_all_least_frequent(const Index& index, std::vector<Hit>& hits, size_t rescue_distance) {
    size_t last_unfiltered_start = 0, first_filtered = 0;
    struct Iv { size_t s, e, n; };
    std::vector<Iv> iv;
    for (size_t i = 0; i < hits.size(); ++i) {
        const Hit& h = hits[i];
        if (h.is_filtered) continue;
        if (h.query_start > last_unfiltered_start + rescue_distance)
            iv.push_back({first_filtered, i, (h.query_start - last_unfiltered_start) / rescue_distance});
        last_unfiltered_start = h.query_start;
        first_filtered = i + 1;
    }
    if (!hits.empty() && hits.back().query_start - last_unfiltered_start > rescue_distance)
        iv.push_back({first_filtered, hits.size(), (hits.back().query_start - last_unfiltered_start) / rescue_distance});
    size_t rescued = 0;
    for (const Iv& x : iv) rescued += rescue_least_frequent(index, hits, x.s, x.e, x.n, LOCAL_RESCUE_THRESHOLD);
    return rescued;
}

// src/hit.rs:270-333
std::vector<Hit> find_hits(const std::vector<QueryRandstrobe>& qr, const Index& index, Mcs mcs,
                           size_t filter_cutoff, size_t rescue_distance, HitsDetails& d) {
    std::vector<Hit> hits = find_all_hits(qr, index, filter_cutoff, mcs, d);
    size_t total = d.partial_filtered + d.partial_found + d.full_filtered + d.full_found;
    size_t nonrep = d.full_found + d.partial_found;
    float frac = total > 0 ? (float)nonrep / (float)total : 1.0f;
    if (rescue_distance > 0) {
        if (frac < 0.7f && nonrep < GLOBAL_RESCUE_COUNT)
            d.rescued += rescue_least_frequent(index, hits, 0, hits.size(), GLOBAL_RESCUE_COUNT, GLOBAL_RESCUE_THRESHOLD);
        d.rescued += rescue_all_least_frequent(index, hits, rescue_distance);
    }
    return hits;
}

// src/chainer.rs:286-302 (f32 throughout; log2 from the host libm)
float chaining_score(size_t dq, size_t dr, size_t k, const ChainingParams& p) {
    size_t dd = dr > dq ? dr - dq : dq - dr;
    size_t dg = std::min(dq, dr);
    float score = (float)std::min(k, dg);
    size_t dg2 = dg > k ? dg - k : 0;
    volatile float a = p.diag_diff_penalty * (float)dd;
    volatile float b = p.gap_length_penalty * (float)dg2;
    float lin = a + b;
    float logp = dd >= 1 ? std::log2f((float)dd + 1.0f) : 0.0f;
    volatile float half = 0.5f * logp;
    score -= lin + half;
    return score;
}

struct ChainingResult {
    float best_score = 0;
    std::vector<float> dp;
    std::vector<size_t> pred;
    std::vector<Anchor> anchors;
};

// src/chainer.rs:85-180
static ChainingResult collinear_chaining(std::vector<Anchor> anchors, size_t k, size_t read_len,
                                         const ChainingParams& p, const std::vector<float>& pre) {
    ChainingResult res;
    size_t n = anchors.size();
    if (n == 0) return res;
    size_t max_ref_gap = p.max_ref_gap >= 0 ? (size_t)p.max_ref_gap : read_len;
    auto score_of = [&](size_t dq, size_t dr) {
        if (dq == dr && dq < pre.size()) return pre[dq];
        return chaining_score(dq, dr, k, p);
    };
    std::vector<float> dp(n, (float)k);
    std::vector<size_t> pred(n, SIZE_MAX);
    float best = 0.0f;
    size_t best_index = SIZE_MAX;
    for (size_t i = 0; i < n; ++i) {
        size_t lookup_end = i > p.max_lookback ? i - p.max_lookback : 0;
        const Anchor& ai = anchors[i];
        for (size_t j = i; j-- > lookup_end;) {
            const Anchor& aj = anchors[j];
            if (ai.ref_id != aj.ref_id) break;
            if (ai.query_start < aj.query_start) continue;
            size_t dq = ai.query_start - aj.query_start;
            size_t dr = ai.ref_start - aj.ref_start;
            if (dr >= max_ref_gap) break;
            float ratio = (float)std::max(dq, dr) / (float)std::min(dq, dr);
            if (ratio > p.max_diagonal_ratio) continue;
            float ns = dp[j] + score_of(dq, dr);
            if (ns >= dp[i]) {
                dp[i] = ns;
                pred[i] = j;
                if (dq == dr) break;
            }
        }
        if (best_index != SIZE_MAX && ai.ref_id == anchors[best_index].ref_id) {
            const Anchor& aj = anchors[best_index];
            if (ai.query_start > aj.query_start) {
                size_t dq = ai.query_start - aj.query_start;
                size_t dr = ai.ref_start - aj.ref_start;
                float ratio = (float)std::max(dq, dr) / (float)std::min(dq, dr);
                if (dr < max_ref_gap && dr > 0 && ratio <= p.max_diagonal_ratio) {
                    float ns = dp[best_index] + score_of(dq, dr);
                    if (ns > dp[i]) { dp[i] = ns; pred[i] = best_index; }
                }
            }
        }
        if (dp[i] > best) { best = dp[i]; best_index = i; }
    }
    res.best_score = best;
    res.dp = std::move(dp);
    res.pred = std::move(pred);
    res.anchors = std::move(anchors);
    return res;
}

static std::vector<float> precompute_scores(size_t k, const ChainingParams& p) {
    std::vector<float> pre(1024);
    for (size_t i = 0; i < 1024; ++i) pre[i] = chaining_score(i, i, k, p);
    return pre;
}

float collinear_chaining_best(const std::vector<Anchor>& anchors, size_t k, size_t read_len, const ChainingParams& p) {
    return collinear_chaining(anchors, k, read_len, p, precompute_scores(k, p)).best_score;
}

// src/chainer.rs:387-448
static void extract_chains(const ChainingResult& r, size_t k, bool is_revcomp, float valid_thr, float mw,
                           std::vector<Nam>& chains) {
    size_t n = r.anchors.size();
    float valid = r.best_score * valid_thr;
    std::vector<std::pair<size_t, float>> cand;
    for (size_t i = 0; i < n; ++i)
        if (r.dp[i] >= valid) cand.push_back({i, r.dp[i]});
    std::stable_sort(cand.begin(), cand.end(), [](const auto& a, const auto& b) { return a.second > b.second; });
    std::vector<bool> used(n, false);
    for (const auto& c : cand) {
        size_t i = c.first;
        if (used[i]) continue;
        size_t j = i;
        bool overlaps = false;
        std::vector<Anchor> ca{r.anchors[i]};
        size_t matching = k, cov = r.anchors[i].ref_start;
        while (r.pred[j] != SIZE_MAX) {
            j = r.pred[j];
            if (used[j]) { overlaps = true; break; }
            ca.push_back(r.anchors[j]);
            used[j] = true;
            size_t gap = cov > r.anchors[j].ref_start ? cov - r.anchors[j].ref_start : 0;
            matching += std::min(gap, k);
            cov = r.anchors[j].ref_start;
        }
        if (overlaps) continue;
        const Anchor& first = r.anchors[j];
        const Anchor& last = r.anchors[i];
        Nam nam;
        nam.nam_id = chains.size();
        nam.query_start = first.query_start;
        nam.query_end = last.query_start + k;
        nam.ref_start = first.ref_start;
        nam.ref_end = last.ref_start + k;
        nam.matching_bases = matching;
        nam.ref_id = last.ref_id;
        nam.score = c.second + (float)ca.size() * mw;
        nam.is_revcomp = is_revcomp;
        nam.anchors = std::move(ca);
        chains.push_back(std::move(nam));
    }
}

// src/chainer.rs:304-384
static std::vector<Anchor> hits_to_anchors(const std::vector<Hit>& hits, const Index& index) {
    std::vector<Anchor> a;
    const size_t k = (size_t)index.k();
    const uint64_t fmask = index.params.randstrobe.forward_main_hash_mask;
    for (const Hit& h : hits) {
        if (h.is_filtered) continue;
        if (h.is_partial) {
            size_t fe = index.get_masked_from(h.hash, fmask, h.position);
            if (fe == SIZE_MAX) continue;
            uint64_t fh = index.randstrobes[fe].hash_offset & fmask;
            for (size_t i = fe; i < index.size(); ++i) {
                const RefRandstrobe& e = index.randstrobes[i];
                if ((e.hash_offset & fmask) != fh) break;
                a.push_back({e.ref_index, e.position, h.query_start});
            }
        } else {
            size_t min_diff = SIZE_MAX;
            uint64_t fh = index.randstrobes[h.position].hash();
            for (size_t i = h.position; i < index.size(); ++i) {
                const RefRandstrobe& e = index.randstrobes[i];
                if (e.hash() != fh) break;
                size_t rs = e.position, re = rs + e.strobe2_offset() + k;
                size_t ql = h.query_end - h.query_start, rl = re - rs;
                size_t diff = ql > rl ? ql - rl : rl - ql;
                if (diff <= min_diff) {
                    a.push_back({e.ref_index, rs, h.query_start});
                    a.push_back({e.ref_index, re - k, h.query_end - k});
                    min_diff = diff;
                }
            }
        }
    }
    return a;
}

// src/nam.rs:138-166 + src/chainer.rs:182-267
std::vector<Nam> get_nams_by_chaining(const uint8_t* seq, size_t len, const Index& index,
                                      const ChainingParams& cp, size_t rescue_distance, Mcs mcs,
                                      NamDetails& details) {
    std::vector<QueryRandstrobe> qr[2];
    {
        StageTimer st(stage_times().randstrobes);
        randstrobes_query(seq, len, index.params, qr);
    }
    std::vector<Hit> hits[2];
    HitsDetails hd[2];
    {
        StageTimer st(stage_times().find_hits);
        for (int o = 0; o < 2; ++o) hits[o] = find_hits(qr[o], index, mcs, index.filter_cutoff, rescue_distance, hd[o]);
    }
    auto better = [](const HitsDetails& a, const HitsDetails& b) {
        size_t t = a.full_found + a.full_filtered, ot = b.full_found + b.full_filtered;
        return t >= ot * 2 && t > ot + 5;
    };
    std::vector<int> orient;
    if (better(hd[0], hd[1])) orient = {0};
    else if (better(hd[1], hd[0])) orient = {1};
    else orient = {0, 1};
    static thread_local std::vector<float> pre;
    static thread_local size_t pre_k = 0;
    static thread_local ChainingParams pre_p;
    const size_t k = (size_t)index.k();
    if (pre.empty() || pre_k != k || pre_p.diag_diff_penalty != cp.diag_diff_penalty ||
        pre_p.gap_length_penalty != cp.gap_length_penalty) {
        pre = precompute_scores(k, cp);
        pre_k = k;
        pre_p = cp;
    }
    std::vector<Nam> chains;
    size_t n_anchors = 0;
    for (int o : orient) {
        std::vector<Anchor> a;
        {
            StageTimer st(stage_times().find_hits);
            a = hits_to_anchors(hits[o], index);
        }
        StageTimer st(stage_times().chaining);
        n_anchors += a.size();
        std::sort(a.begin(), a.end());
        a.erase(std::unique(a.begin(), a.end()), a.end());
        ChainingResult r = collinear_chaining(std::move(a), k, len, cp, pre);
        extract_chains(r, k, o == 1, cp.valid_score_threshold, cp.matches_weight, chains);
    }
    HitsDetails both = hd[0];
    both.full_not_found += hd[1].full_not_found;
    both.full_filtered += hd[1].full_filtered;
    both.full_found += hd[1].full_found;
    both.partial_not_found += hd[1].partial_not_found;
    both.partial_filtered += hd[1].partial_filtered;
    both.partial_found += hd[1].partial_found;
    both.rescued += hd[1].rescued;
    details.hits = both;
    details.n_reads = 1;
    details.n_randstrobes = qr[0].size() + qr[1].size();
    details.n_anchors = n_anchors;
    details.n_nams = chains.size();
    details.both_orientations = orient.size() > 1;
    return chains;
}

// ORC_RNG_VARIANT (test infrastructure: plants a mistake in the generator to
// measure what the golden files pin): 0 fastrand as restated; 1 another wyrand
// increment; 2 modulo in place of Lemire's bounded sampling; 3 the shuffle
// drawing from [0, i) instead of [0, i]. Bounded draws with n >= 2 are counted.
int rng_variant() {
    static const int v = [] {
        const char* e = std::getenv("ORC_RNG_VARIANT");
        return e ? std::atoi(e) : 0;
    }();
    return v;
}
std::atomic<uint64_t>& rng_draws() {
    static std::atomic<uint64_t> n{0};
    return n;
}

// fastrand 2.3.0 (wyrand v4.2 constants) -- parity unpinned, see oracle.hpp.
uint64_t Rng::gen_u64() {
    s += rng_variant() == 1 ? 0x9e3779b97f4a7c15ULL : 0x2d358dccaa6c78a5ULL;
    unsigned __int128 t = (unsigned __int128)s * (unsigned __int128)(s ^ 0x8bb84b93962eacc9ULL);
    return (uint64_t)t ^ (uint64_t)(t >> 64);
}

// Lemire's nearly divisionless bounded sampling, 32-bit variant.
uint32_t Rng::gen_mod_u32(uint32_t n) {
    if (n >= 2) rng_draws()++;
    if (rng_variant() == 2) return (uint32_t)gen_u64() % n;
    uint32_t r = (uint32_t)gen_u64();
    uint64_t m = (uint64_t)r * n;
    uint32_t hi = (uint32_t)(m >> 32), lo = (uint32_t)m;
    if (lo < n) {
        uint32_t t = (uint32_t)(-n) % n;
        while (lo < t) {
            r = (uint32_t)gen_u64();
            m = (uint64_t)r * n;
            hi = (uint32_t)(m >> 32);
            lo = (uint32_t)m;
        }
    }
    return hi;
}

uint64_t Rng::gen_mod_u64(uint64_t n) {
    if (n >= 2) rng_draws()++;
    if (rng_variant() == 2) return gen_u64() % n;
    uint64_t r = gen_u64();
    unsigned __int128 m = (unsigned __int128)r * n;
    uint64_t hi = (uint64_t)(m >> 64), lo = (uint64_t)m;
    if (lo < n) {
        uint64_t t = (uint64_t)(-n) % n;
        while (lo < t) {
            r = gen_u64();
            m = (unsigned __int128)r * n;
            hi = (uint64_t)(m >> 64);
            lo = (uint64_t)m;
        }
    }
    return hi;
}

// src/nam.rs:168-188, src/shuffle.rs:7-25 (fastrand shuffle: swap(i, usize(..=i)) for i in 1..n)
void sort_nams(std::vector<Nam>& nams, Rng& rng) {
    StageTimer st(stage_times().sort_nams);
    std::stable_sort(nams.begin(), nams.end(), [](const Nam& a, const Nam& b) { return a.score > b.score; });
    if (nams.empty()) return;
    float best = nams[0].score;
    size_t n_best = 0;
    while (n_best < nams.size() && nams[n_best].score == best) ++n_best;
    if (n_best > 1)
        for (size_t i = 1; i < n_best; ++i)
            std::swap(nams[i], nams[rng.gen_mod_u64(rng_variant() == 3 ? i : i + 1)]);
}

StageTimes& stage_times() {
    thread_local StageTimes t;
    return t;
}

}  // namespace orc
