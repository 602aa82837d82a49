// Checks two_diagonal_score (strobealign_amd/csrc/sa_extend.hpp), the lower
// bound that sends an SSW job straight to the 16-bit pass, on the CPU:
//  - against a direct enumeration of the alignments it stands for (a segment on
//    d1, a segment on d2, or a segment on d1, one gap, a segment on d2), on
//    short queries;
//  - as a lower bound of the local affine-gap score (Gotoh, a gap of g
//    positions costing gap_open + (g - 1) * gap_extend, as in the SSW passes),
//    on queries up to 200 bases;
//  - on constructed reads (a window copy with one insertion or deletion) that
//    it must score in full.
// Prints "ok <cases>" or the first failing case, and exits non-zero on failure.
#include <algorithm>
#include <cstdio>
#include <random>
#include <string>
#include <vector>

#include "sa_extend.hpp"

using namespace sa;

struct Query {
    const std::string* s;
    uint8_t operator[](uint32_t i) const { return (uint8_t)(*s)[i]; }
};

static int32_t sub(char q, char r, const Scores& sc) {
    const int32_t a = ssw_code((uint8_t)q), b = ssw_code((uint8_t)r);
    return (a == b && a < 4) ? sc.match : -sc.mismatch;
}

// Every alignment of the family, summed cell by cell.
static int32_t enumerate(const std::string& q, const std::string& w, int64_t d1, int64_t d2, const Scores& sc) {
    const int64_t L = (int64_t)q.size(), W = (int64_t)w.size();
    auto on = [&](int64_t x, int64_t d) { return x + d >= 0 && x + d < W; };
    auto seg = [&](int64_t s, int64_t e, int64_t d, int32_t& v) {  // [s, e] on d
        v = 0;
        for (int64_t x = s; x <= e; ++x) {
            if (!on(x, d)) return false;
            v += sub(q[x], w[x + d], sc);
        }
        return true;
    };
    int32_t best = 0, v, v2;
    for (int64_t s = 0; s < L; ++s)
        for (int64_t e = s; e < L; ++e) {
            if (seg(s, e, d1, v)) best = std::max(best, v);
            if (seg(s, e, d2, v)) best = std::max(best, v);
        }
    const int64_t g = d1 > d2 ? d1 - d2 : d2 - d1;
    if (g == 0 || g > TWO_DIAGONAL_MAX_GAP) return best;
    const int64_t off = d2 > d1 ? 1 : 1 + g;
    const int32_t cost = sc.gap_open + (int32_t)(g - 1) * sc.gap_extend;
    for (int64_t s = 0; s < L; ++s)
        for (int64_t e = s; e < L; ++e) {
            if (!seg(s, e, d1, v)) continue;
            for (int64_t e2 = e + off; e2 < L; ++e2)
                if (seg(e + off, e2, d2, v2)) best = std::max(best, v + v2 - cost);
        }
    return best;
}

// Local alignment score with affine gaps (Gotoh).
static int32_t local_affine(const std::string& q, const std::string& w, const Scores& sc) {
    const size_t L = q.size(), W = w.size();
    const int32_t NEG = -(1 << 28);
    std::vector<int32_t> H(W + 1, 0), E(W + 1, NEG), Hn(W + 1), En(W + 1);
    int32_t best = 0;
    for (size_t i = 1; i <= L; ++i) {
        Hn[0] = 0;
        En[0] = NEG;
        int32_t F = NEG;
        for (size_t j = 1; j <= W; ++j) {
            En[j] = std::max(E[j] - sc.gap_extend, H[j] - sc.gap_open);          // gap in the reference (down the query)
            F = std::max(F - sc.gap_extend, Hn[j - 1] - sc.gap_open);            // gap in the query (along the reference)
            Hn[j] = std::max({0, H[j - 1] + sub(q[i - 1], w[j - 1], sc), En[j], F});
            best = std::max(best, Hn[j]);
        }
        std::swap(H, Hn);
        std::swap(E, En);
    }
    return best;
}

int main() {
    std::mt19937 rng(2024);
    const char* acgt = "ACGT";
    auto rnd = [&](size_t n) {
        std::string s(n, 'A');
        for (auto& c : s) c = acgt[rng() % 4];
        return s;
    };
    const Scores scs[2] = {{2, 8, 12, 1, 10, 0}, {1, 4, 6, 1, 10, 0}};
    // A packed window: one contig of the window's bases.
    std::vector<uint64_t> words;
    uint64_t word_off = 0, len = 0;
    auto pack = [&](const std::string& w) {
        words.assign(w.size() / 32 + 1, 0);
        for (size_t i = 0; i < w.size(); ++i)
            words[i / 32] |= (uint64_t)(std::string("ACGT").find(w[i])) << ((i % 32) * 2);
        len = w.size();
        return RefView{words.data(), &word_off, &len, 1};
    };
    uint64_t cases = 0;
    for (int it = 0; it < 6000; ++it) {
        const Scores& sc = scs[it & 1];
        const bool small = it < 3000;
        const size_t L = small ? 1 + rng() % 24 : 1 + rng() % 200;
        const std::string w = rnd(L + rng() % (small ? 12 : 60));
        // a read drawn from the window around d1, then d2, with substitutions and Ns
        const int64_t d1 = (int64_t)(rng() % (w.size() + 4)) - 4, d2 = d1 + (int64_t)(rng() % 13) - 6;
        std::string q(L, 'A');
        const size_t cut = rng() % (L + 1);
        for (size_t x = 0; x < L; ++x) {
            const int64_t p = (int64_t)x + (x < cut ? d1 : d2);
            q[x] = (p >= 0 && p < (int64_t)w.size() && rng() % 100 < 92) ? w[p] : "ACGTN"[rng() % 5];
        }
        const RefView rv = pack(w);
        const RefWin win{&rv, 0, 0, (uint32_t)w.size()};
        const int32_t b = two_diagonal_score(Query{&q}, win, (uint32_t)L, d1, d2, sc);
        const int32_t sw = local_affine(q, w, sc);
        if (b > sw) {
            std::printf("bound %d above the local score %d: L=%zu W=%zu d1=%lld d2=%lld\n", b, sw, L, w.size(),
                        (long long)d1, (long long)d2);
            return 1;
        }
        if (small) {
            const int32_t e = enumerate(q, w, d1, d2, sc);
            if (e != b) {
                std::printf("bound %d, enumeration %d: L=%zu W=%zu d1=%lld d2=%lld\n", b, e, L, w.size(),
                            (long long)d1, (long long)d2);
                return 1;
            }
        }
        ++cases;
    }
    // one indel of g bases in an exact copy: the bound is the whole alignment
    for (int it = 0; it < 400; ++it) {
        const Scores& sc = scs[it & 1];
        const size_t L = 60 + rng() % 140, cut = 10 + rng() % (L - 20), g = 1 + rng() % 8;
        const std::string w = rnd(L + 2 * g + 20);
        const int64_t d1 = 10;
        const bool del = it & 2;  // reference bases skipped (d2 > d1) or query bases inserted (d2 < d1)
        std::string q;
        int32_t want;
        if (del) {
            q = w.substr(d1, cut) + w.substr(d1 + cut + g, L - cut);
            want = (int32_t)L * sc.match - sc.gap_open - (int32_t)(g - 1) * sc.gap_extend;
        } else {
            q = w.substr(d1, cut) + std::string(g, 'N') + w.substr(d1 + cut, L - cut - g);
            want = (int32_t)(L - g) * sc.match - sc.gap_open - (int32_t)(g - 1) * sc.gap_extend;
        }
        const int64_t d2 = del ? d1 + (int64_t)g : d1 - (int64_t)g;
        const RefView rv = pack(w);
        const RefWin win{&rv, 0, 0, (uint32_t)w.size()};
        const int32_t b = two_diagonal_score(Query{&q}, win, (uint32_t)q.size(), d1, d2, sc);
        const int32_t sw = local_affine(q, w, sc);
        if (b < want || b > sw) {
            std::printf("indel case: bound %d, alignment %d, local score %d (g=%zu, %s)\n", b, want, sw, g,
                        del ? "deletion" : "insertion");
            return 1;
        }
        ++cases;
    }
    std::printf("ok %llu\n", (unsigned long long)cases);
    return 0;
}
