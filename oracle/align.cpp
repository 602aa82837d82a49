// Oracle: CIGAR, ungapped (Hamming) alignment, SSW restatement, piecewise extension.
// TEST INFRASTRUCTURE ONLY (see oracle.hpp).
#include "oracle.hpp"

#include <emmintrin.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>

namespace orc {

// ---------------------------------------------------------------- cigar ---
// src/cigar.rs:96-105 (drop empty ops, merge equal neighbours)
void Cigar::push(uint8_t op, size_t len) {
    if (len == 0) return;
    if (ops.empty() || ops.back().first != op) ops.push_back({op, (uint32_t)len});
    else ops.back().second += (uint32_t)len;
}

// src/cigar.rs:112-125
Cigar Cigar::with_m() const {
    Cigar c;
    for (auto& p : ops) c.push((p.first == 7 || p.first == 8) ? 0 : p.first, p.second);
    return c;
}

// src/cigar.rs:150-188 (ops of length 0 from SSW are skipped by push)
Cigar Cigar::with_eqx(const uint8_t* q, const uint8_t* r) const {
    Cigar c;
    size_t qi = 0, ri = 0;
    for (auto& p : ops) {
        uint8_t op = p.first;
        size_t len = p.second;
        switch (op) {
            case 0:
                for (size_t t = 0; t < len; ++t) { c.push(q[qi] == r[ri] ? 7 : 8, 1); ++qi; ++ri; }
                break;
            case 1: case 4: c.push(op, len); qi += len; break;
            case 2: case 3: c.push(op, len); ri += len; break;
            case 5: case 6: c.push(op, len); break;
            default: c.push(op, len); qi += len; ri += len; break;
        }
    }
    return c;
}

// src/cigar.rs:136-147
size_t Cigar::edit_distance() const {
    size_t d = 0;
    for (auto& p : ops)
        if (p.first == 8 || p.first == 1 || p.first == 2) d += p.second;
    return d;
}

std::string Cigar::str() const {
    static const char L[] = "MIDNSHP=X";
    std::string s;
    for (auto& p : ops) { s += std::to_string(p.second); s += L[p.first]; }
    return s;
}

// ------------------------------------------------------------- hamming ---
// src/aligner.rs:181-219
void highest_scoring_segment(const uint8_t* q, const uint8_t* r, size_t n, int match, int mismatch,
                             int end_bonus, size_t& bs, size_t& be, int& bscore) {
    size_t start = 0;
    int score = end_bonus;
    bs = 0; be = 0; bscore = 0;
    for (size_t i = 0; i < n; ++i) {
        if (q[i] == r[i]) score += match; else score -= mismatch;
        if (score < 0) { start = i + 1; score = 0; }
        if (score > bscore) { bs = start; bscore = score; be = i + 1; }
    }
    if (score + end_bonus > bscore) { bscore = score + end_bonus; be = n; bs = start; }
}

// src/aligner.rs:223-259
std::optional<AlignmentInfo> hamming_align(const uint8_t* q, const uint8_t* r, size_t len, const Scores& sc) {
    size_t s, e;
    int score;
    highest_scoring_segment(q, r, len, sc.match, sc.mismatch, sc.end_bonus, s, e, score);
    AlignmentInfo a;
    size_t mm = 0;
    for (size_t i = s; i < e; ++i) {
        if (q[i] != r[i]) { ++mm; a.cigar.push(8, 1); } else a.cigar.push(7, 1);
    }
    a.score = (uint32_t)score;
    a.edit_distance = mm;
    a.ref_start = s; a.ref_end = e; a.query_start = s; a.query_end = e;
    return a;
}

// src/aligner.rs:261-291
AlignmentInfo hamming_align_global(const uint8_t* q, const uint8_t* r, size_t n, int match, int mismatch) {
    AlignmentInfo a;
    int score = 0;
    for (size_t i = 0; i < n; ++i) {
        if (q[i] == r[i]) { score += match; a.cigar.push(7, 1); }
        else { score -= mismatch; a.edit_distance++; a.cigar.push(8, 1); }
    }
    a.ref_start = 0; a.ref_end = n; a.query_start = 0; a.query_end = n;
    a.score = (uint32_t)std::max(0, score);
    return a;
}

// ------------------------------------------------------------------ SSW ---
// Striped Smith-Waterman as in ext/ssw/ssw.c:197-588.
namespace {
struct EndPos { int score; int32_t ref; int32_t read; };

// The striped passes on SSE2 vectors: 16 unsigned bytes with bias (BYTE) or 8
// signed 16-bit words, one vector per stripe segment, so the CPU baseline runs
// at the speed of the reference's own SIMD kernels. Saturating arithmetic, the
// lazy-F loop that does not revisit E and stops once F cannot raise any H of the
// next stripe, the column maximum, the best column's H kept for the end position
// on the read: the results of ext/ssw/ssw.c, whose scores depend on the striping.
template <bool BYTE>
struct StripedSse {
    static __m128i adds(__m128i a, __m128i b) { return BYTE ? _mm_adds_epu8(a, b) : _mm_adds_epi16(a, b); }
    static __m128i subs(__m128i a, __m128i b) { return BYTE ? _mm_subs_epu8(a, b) : _mm_subs_epu16(a, b); }
    static __m128i vmax(__m128i a, __m128i b) { return BYTE ? _mm_max_epu8(a, b) : _mm_max_epi16(a, b); }
    static __m128i set1(int v) { return BYTE ? _mm_set1_epi8((char)v) : _mm_set1_epi16((short)v); }
    static __m128i shift_up(__m128i v) { return BYTE ? _mm_slli_si128(v, 1) : _mm_slli_si128(v, 2); }
    // all lanes of a <= b (unsigned bytes or non-negative words): a - b saturates to 0
    static bool all_le(__m128i a, __m128i b) {
        const __m128i d = subs(a, b);
        return _mm_movemask_epi8(BYTE ? _mm_cmpeq_epi8(d, _mm_setzero_si128())
                                      : _mm_cmpeq_epi16(d, _mm_setzero_si128())) == 0xFFFF;
    }
    static int hmax(__m128i v) {
        alignas(16) uint16_t w[8];
        alignas(16) uint8_t b[16];
        int m = 0;
        if (BYTE) {
            _mm_store_si128((__m128i*)b, v);
            for (int l = 0; l < 16; ++l) m = std::max<int>(m, b[l]);
        } else {
            _mm_store_si128((__m128i*)w, v);
            for (int l = 0; l < 8; ++l) m = std::max<int>(m, (int16_t)w[l]);
        }
        return m;
    }
    static constexpr int LANES = BYTE ? 16 : 8;

    static std::vector<__m128i> profile(const int8_t* read, int len, const int8_t* mat, int n, int bias) {
        const int segLen = (len + LANES - 1) / LANES;
        std::vector<__m128i> p((size_t)n * segLen);
        for (int nt = 0; nt < n; ++nt)
            for (int i = 0; i < segLen; ++i) {
                alignas(16) int16_t w[8];
                alignas(16) uint8_t b[16];
                for (int l = 0; l < LANES; ++l) {
                    const int j = i + l * segLen;
                    const int v = j >= len ? 0 : mat[nt * n + read[j]];
                    if (BYTE) b[l] = (uint8_t)(v + bias);
                    else w[l] = (int16_t)v;
                }
                p[(size_t)nt * segLen + i] = BYTE ? _mm_load_si128((const __m128i*)b) : _mm_load_si128((const __m128i*)w);
            }
        return p;
    }

    static EndPos run(const int8_t* ref, bool reverse, int refLen, int readLen, int gapO, int gapE,
                      const std::vector<__m128i>& prof, int terminate, int bias) {
        const int segLen = (readLen + LANES - 1) / LANES;
        int maxv = 0;
        int32_t end_read = readLen - 1;
        int32_t end_ref = BYTE ? -1 : 0;
        const __m128i zero = _mm_setzero_si128(), vGapO = set1(gapO), vGapE = set1(gapE), vBias = set1(bias);
        std::vector<__m128i> HS(segLen, zero), HL(segLen, zero), E(segLen, zero), Hmax(segLen, zero);
        __m128i vMaxScore = zero, vMaxMark = zero;
        const int begin = reverse ? refLen - 1 : 0, end = reverse ? -1 : refLen, step = reverse ? -1 : 1;
        for (int i = begin; i != end; i += step) {
            __m128i vF = zero, vMaxCol = zero;
            __m128i vH = shift_up(HS[segLen - 1]);
            const __m128i* vP = &prof[(size_t)ref[i] * segLen];
            std::swap(HS, HL);
            for (int j = 0; j < segLen; ++j) {
                vH = BYTE ? subs(adds(vH, vP[j]), vBias) : adds(vH, vP[j]);
                __m128i e = E[j];
                vH = vmax(vH, e);
                vH = vmax(vH, vF);
                vMaxCol = vmax(vMaxCol, vH);
                HS[j] = vH;
                vH = subs(vH, vGapO);
                e = vmax(subs(e, vGapE), vH);
                E[j] = e;
                vF = vmax(subs(vF, vGapE), vH);
                vH = HL[j];
            }
            // lazy F: at most LANES passes, F carried one lane up per pass, E left
            // as it is; stops at the first segment where F, after the extension,
            // cannot exceed that segment's updated H minus the opening
            bool done = false;
            for (int k = 0; k < LANES && !done; ++k) {
                vF = shift_up(vF);
                for (int j = 0; j < segLen; ++j) {
                    vH = vmax(HS[j], vF);
                    vMaxCol = vmax(vMaxCol, vH);
                    HS[j] = vH;
                    vH = subs(vH, vGapO);
                    vF = subs(vF, vGapE);
                    if (all_le(vF, vH)) { done = true; break; }
                }
            }
            vMaxScore = vmax(vMaxScore, vMaxCol);
            if (!all_le(vMaxScore, vMaxMark) || !all_le(vMaxMark, vMaxScore)) {
                vMaxMark = vMaxScore;
                const int temp = hmax(vMaxScore);
                if (temp > maxv) {
                    maxv = temp;
                    if (BYTE && maxv + bias >= 255) break;
                    end_ref = i;
                    Hmax = HS;
                }
            }
            if (hmax(vMaxCol) == terminate) break;
        }
        for (int s = 0; s < segLen; ++s) {
            alignas(16) int16_t w[8];
            alignas(16) uint8_t b[16];
            if (BYTE) _mm_store_si128((__m128i*)b, Hmax[s]);
            else _mm_store_si128((__m128i*)w, Hmax[s]);
            for (int l = 0; l < LANES; ++l) {
                const int v = BYTE ? b[l] : w[l];
                if (v == maxv) {
                    const int32_t t = s + l * segLen;
                    if (t < end_read) end_read = t;
                }
            }
        }
        EndPos r;
        r.score = BYTE ? (maxv + bias >= 255 ? 255 : maxv) : maxv;
        r.ref = end_ref;
        r.read = end_read;
        return r;
    }
};

// ext/ssw/ssw.c:590-774: banded global alignment with band doubling and traceback.
// Returns false when the traceback meets an unset cell.
bool banded_sw(const int8_t* ref, const int8_t* read, int refLen, int readLen, int score, int gapO, int gapE,
               int bw, const int8_t* mat, int n, std::vector<uint32_t>& out) {
    int len = std::max(refLen, readLen);
    int maxv = 0;
    std::vector<int> hb, eb, hc;
    std::vector<int8_t> dir;
    int width = 0, width_d = 0;
    auto U = [&](int i, int j) { int x = i - bw; x = x > 0 ? x : 0; return j - x + 1; };
    auto D = [&](int i, int j, int p) { int x = i - bw; x = x > 0 ? x : 0; return (j - x) * 3 + p; };
    do {
        width = bw * 2 + 3;
        width_d = bw * 2 + 1;
        hb.assign(width + 1, 0);
        eb.assign(width + 1, 0);
        hc.assign(width + 1, 0);
        dir.assign((size_t)width_d * readLen * 3 + 3, 0);
        for (int i = 0; i < readLen; ++i) {
            int beg = std::max(0, i - bw), end = std::min(refLen - 1, i + bw);
            int edge = std::min(end + 1, width - 1);
            int f = 0;
            hb[0] = eb[0] = hb[edge] = eb[edge] = hc[0] = 0;
            int8_t* dl = &dir[(size_t)width_d * i * 3];
            int u = 0;
            for (int j = beg; j <= end; ++j) {
                u = U(i, j);
                int e = U(i - 1, j), b = U(i, j - 1), d = U(i - 1, j - 1);
                int de = D(i, j, 0), df = D(i, j, 1), dh = D(i, j, 2);
                int t1 = i == 0 ? -gapO : hb[e] - gapO;
                int t2 = i == 0 ? -gapE : eb[e] - gapE;
                eb[u] = t1 > t2 ? t1 : t2;
                dl[de] = t1 > t2 ? 3 : 2;
                t1 = hc[b] - gapO;
                t2 = f - gapE;
                f = t1 > t2 ? t1 : t2;
                dl[df] = t1 > t2 ? 5 : 4;
                int e1 = eb[u] > 0 ? eb[u] : 0, f1 = f > 0 ? f : 0;
                t1 = e1 > f1 ? e1 : f1;
                t2 = hb[d] + mat[ref[j] * n + read[i]];
                hc[u] = t1 > t2 ? t1 : t2;
                if (hc[u] > maxv) maxv = hc[u];
                if (t1 <= t2) dl[dh] = 1;
                else dl[dh] = e1 > f1 ? dl[de] : dl[df];
            }
            for (int j = 1; j <= u; ++j) hb[j] = hc[j];
        }
        bw *= 2;
    } while (maxv < score && bw <= len);
    bw /= 2;
    static const uint32_t code_M = 0, code_I = 1, code_D = 2;
    std::vector<uint32_t> c;
    int i = readLen - 1, j = refLen - 1, e = 0;
    char op = 'M', prev = 'M';
    int t2 = 2;
    int row = readLen - 1;
    auto enc = [&](int l, char o) { return ((uint32_t)l << 4) | (o == 'M' ? code_M : o == 'I' ? code_I : code_D); };
    while (i >= 0 && j > 0) {
        int idx = D(i, j, t2);
        int8_t v = dir[(size_t)width_d * row * 3 + idx];
        switch (v) {
            case 1: --i; --j; t2 = 2; --row; op = 'M'; break;
            case 2: --i; t2 = 0; --row; op = 'I'; break;
            case 3: --i; t2 = 2; --row; op = 'I'; break;
            case 4: --j; t2 = 1; op = 'D'; break;
            case 5: --j; t2 = 2; op = 'D'; break;
            default: return false;
        }
        if (op == prev) ++e;
        else { c.push_back(enc(e, prev)); prev = op; e = 1; }
    }
    if (op == 'M') c.push_back(enc(e + 1, op));
    else { c.push_back(enc(e, op)); c.push_back(enc(1, 'M')); }
    out.assign(c.rbegin(), c.rend());
    return true;
}
}  // namespace

// ext/ssw/ssw.c:818-922 with score_size = 2 (u8 first, i16 on overflow)
SswResult ssw_align(const int8_t* read, int readLen, const int8_t* ref, int refLen, const int8_t* mat, int n,
                    int gapO, int gapE, int maskLen) {
    (void)maskLen;
    SswResult r;
    int bias = 0;
    for (int i = 0; i < n * n; ++i) bias = std::min<int>(bias, mat[i]);
    bias = -bias;
    using B = StripedSse<true>;
    using W = StripedSse<false>;
    EndPos best = B::run(ref, false, refLen, readLen, gapO, gapE, B::profile(read, readLen, mat, n, bias), 255, bias);
    bool word = false;
    if (best.score == 255) {
        best = W::run(ref, false, refLen, readLen, gapO, gapE, W::profile(read, readLen, mat, n, 0), 65535, 0);
        word = true;
    }
    r.score1 = (uint16_t)best.score;
    r.ref_end1 = best.ref;
    r.read_end1 = best.read;
    std::vector<int8_t> rev(read, read + r.read_end1 + 1);
    std::reverse(rev.begin(), rev.end());
    int rlen = r.read_end1 + 1;
    EndPos br;
    if (!word) br = B::run(ref, true, r.ref_end1 + 1, rlen, gapO, gapE, B::profile(rev.data(), rlen, mat, n, bias), (uint8_t)r.score1, bias);
    else br = W::run(ref, true, r.ref_end1 + 1, rlen, gapO, gapE, W::profile(rev.data(), rlen, mat, n, 0), r.score1, 0);
    r.ref_begin1 = br.ref;
    r.read_begin1 = r.read_end1 - br.read;
    if (r.score1 > br.score) r.flag = 2;
    int refL = r.ref_end1 - r.ref_begin1 + 1, readL = r.read_end1 - r.read_begin1 + 1;
    if (r.ref_begin1 < 0 || r.read_begin1 < 0) return r;  // no alignment (the caller rejects it)
    int bw = std::abs(refL - readL) + 1;
    if (!banded_sw(ref + r.ref_begin1, read + r.read_begin1, refL, readL, r.score1, gapO, gapE, bw, mat, n, r.cigar))
        r.flag = 1;
    return r;
}

// src/ssw.rs:8-22 translate: ACGT -> 0..3, U -> 0 (sic), other -> 4
static int8_t ssw_code(uint8_t c) {
    switch (c) {
        case 'A': case 'a': case 'U': case 'u': return 0;
        case 'C': case 'c': return 1;
        case 'G': case 'g': return 2;
        case 'T': case 't': return 3;
        default: return 4;
    }
}

// src/ssw.rs:156-191 + src/aligner.rs:80-151
std::optional<AlignmentInfo> aligner_align(const uint8_t* query, size_t qlen, const uint8_t* refseq, size_t rlen,
                                           const Scores& sc) {
    if (rlen > 2000) return std::nullopt;
    if (qlen == 0) return std::nullopt;
    std::vector<int8_t> q(qlen), r(rlen);
    for (size_t i = 0; i < qlen; ++i) q[i] = ssw_code(query[i]);
    for (size_t i = 0; i < rlen; ++i) r[i] = ssw_code(refseq[i]);
    int8_t mat[25];
    for (int a = 0; a < 5; ++a)
        for (int b = 0; b < 5; ++b) mat[a * 5 + b] = (a == b && a < 4) ? (int8_t)sc.match : (int8_t)-sc.mismatch;
    int mask_len = std::max<int>((int)qlen / 2, 15);
    SswResult s = ssw_align(q.data(), (int)qlen, r.data(), (int)rlen, mat, 5, sc.gap_open, sc.gap_extend, mask_len);
    if (s.ref_begin1 == -1 || s.read_begin1 == -1) return std::nullopt;
    if (s.flag != 0) return std::nullopt;
    AlignmentInfo a;
    a.score = s.score1;
    a.ref_start = (size_t)s.ref_begin1;
    a.ref_end = (size_t)s.ref_end1 + 1;
    a.query_start = (size_t)s.read_begin1;
    a.query_end = (size_t)s.read_end1 + 1;
    Cigar raw;
    for (uint32_t v : s.cigar) raw.ops.push_back({(uint8_t)(v & 0xf), v >> 4});
    a.cigar = raw.with_eqx(query + a.query_start, refseq + a.ref_start);
    a.edit_distance = a.cigar.edit_distance();

    // end-bonus extension towards the query start
    size_t qs = a.query_start, rs = a.ref_start;
    int score = (int)a.score;
    size_t edits = a.edit_distance;
    Cigar front;
    while (qs > 0 && rs > 0) {
        --qs; --rs;
        if (query[qs] == refseq[rs]) { score += sc.match; front.push(7, 1); }
        else { score -= sc.mismatch; front.push(8, 1); ++edits; }
    }
    int swb = score + sc.end_bonus;
    if (qs == 0 && swb > (int)a.score) {
        if (a.query_start > 0) {
            Cigar c = front.reversed();
            c.extend(a.cigar);
            a.cigar = c;
        }
        a.query_start = 0;
        a.ref_start = rs;
        a.score = (uint32_t)swb;
        a.edit_distance = edits;
    }
    // ... and towards the query end
    size_t qe = a.query_end, re = a.ref_end;
    score = (int)a.score;
    edits = a.edit_distance;
    Cigar back;
    while (qe < qlen && re < rlen) {
        if (query[qe] == refseq[re]) { score += sc.match; back.push(7, 1); }
        else { score -= sc.mismatch; back.push(8, 1); ++edits; }
        ++qe; ++re;
    }
    swb = score + sc.end_bonus;
    if (qe == qlen && swb > (int)a.score) {
        if (a.query_end < qlen) a.cigar.extend(back);
        a.query_end = qlen;
        a.ref_end = re;
        a.score = (uint32_t)swb;
        a.edit_distance = edits;
    }
    return a;
}

// ------------------------------------------------------------ piecewise ---
// Affine-gap DP standing in for block-aligner (parity unpinned): gap of
// length L costs gap_open + (L-1)*gap_extend, as block-aligner's Gaps and the
// profile gap settings in src/piecewisealigner.rs:589-624 both give.
// Traceback order at ties (walking back from the end): deletion, then insertion, then diagonal;
// a gap run is closed as soon as opening it explains the score.
namespace {
const int NEG = INT_MIN / 4;

struct Dp {
    int m, n;  // query length (rows), ref length (cols)
    std::vector<int> H, E, F;  // E: horizontal (consumes ref -> D), F: vertical (consumes query -> I)
    int at(const std::vector<int>& v, int i, int j) const { return v[(size_t)i * (n + 1) + j]; }
};

// Fills the DP for a global start at (0,0). sub(i,j) scores query[i-1] vs ref[j-1].
template <typename Sub>
Dp fill(int m, int n, int go, int ge, Sub sub) {
    Dp d;
    d.m = m; d.n = n;
    size_t N = (size_t)(m + 1) * (n + 1);
    d.H.assign(N, NEG); d.E.assign(N, NEG); d.F.assign(N, NEG);
    auto ix = [&](int i, int j) { return (size_t)i * (n + 1) + j; };
    d.H[0] = 0;
    for (int j = 1; j <= n; ++j) { d.E[ix(0, j)] = -(go + (j - 1) * ge); d.H[ix(0, j)] = d.E[ix(0, j)]; }
    for (int i = 1; i <= m; ++i) { d.F[ix(i, 0)] = -(go + (i - 1) * ge); d.H[ix(i, 0)] = d.F[ix(i, 0)]; }
    for (int i = 1; i <= m; ++i)
        for (int j = 1; j <= n; ++j) {
            int e = std::max(d.E[ix(i, j - 1)] - ge, d.H[ix(i, j - 1)] - go);
            int f = std::max(d.F[ix(i - 1, j)] - ge, d.H[ix(i - 1, j)] - go);
            int h = d.H[ix(i - 1, j - 1)] + sub(i, j);
            d.E[ix(i, j)] = e;
            d.F[ix(i, j)] = f;
            d.H[ix(i, j)] = std::max(h, std::max(e, f));
        }
    return d;
}

// Traceback from (i,j) to (0,0); returns ops in forward order (M as =/X by equality).
template <typename Sub>
Cigar trace(const Dp& d, int i, int j, int go, int ge, Sub sub, const uint8_t* q, const uint8_t* r) {
    auto ix = [&](int a, int b) { return (size_t)a * (d.n + 1) + b; };
    std::vector<uint8_t> rev;
    int state = 0;  // 0 H, 1 E (D run), 2 F (I run)
    while (i > 0 || j > 0) {
        if (state == 0) {
            int h = d.H[ix(i, j)];
            if (j > 0 && h == d.E[ix(i, j)]) {
                state = 1;
            } else if (i > 0 && h == d.F[ix(i, j)]) {
                state = 2;
            } else {
                rev.push_back(q[i - 1] == r[j - 1] ? 7 : 8);
                --i; --j;
            }
        } else if (state == 2) {
            rev.push_back(1);
            bool open = (i == 1 && j == 0) ? true : (d.F[ix(i, j)] == d.H[ix(i - 1, j)] - go);
            if (i - 1 == 0 && j == 0) open = true;
            --i;
            if (open) state = 0;
        } else {
            rev.push_back(2);
            bool open = (d.E[ix(i, j)] == d.H[ix(i, j - 1)] - go);
            if (i == 0 && j - 1 == 0) open = true;
            --j;
            if (open) state = 0;
        }
    }
    Cigar c;
    for (size_t t = rev.size(); t-- > 0;) c.push(rev[t], 1);
    return c;
}
}  // namespace

// src/piecewisealigner.rs:84-130
PwResult pw_global_alignment(const uint8_t* q, size_t ql, const uint8_t* r, size_t rl, const Scores& sc) {
    PwResult res;
    if (ql == 0 || rl == 0) return res;
    auto sub = [&](int i, int j) { return q[i - 1] == r[j - 1] ? sc.match : -sc.mismatch; };
    Dp d = fill((int)ql, (int)rl, sc.gap_open, sc.gap_extend, sub);
    res.score = d.H[(size_t)ql * (rl + 1) + rl];
    res.query_end = ql;
    res.ref_end = rl;
    res.cigar = trace(d, (int)ql, (int)rl, sc.gap_open, sc.gap_extend, sub, q, r);
    return res;
}

// src/piecewisealigner.rs:161-229. Global start, free end (max cell), end bonus on
// the last query position. LocalStartGlobalEnd runs on reversed sequences.
PwResult pw_xdrop_alignment(const uint8_t* q0, size_t ql, const uint8_t* r0, size_t rl, const Scores& sc,
                            bool gsle, int xdrop) {
    (void)xdrop;
    PwResult res;
    if (ql == 0 || rl == 0) return res;
    std::vector<uint8_t> q(q0, q0 + ql), r(r0, r0 + rl);
    if (!gsle) { std::reverse(q.begin(), q.end()); std::reverse(r.begin(), r.end()); }
    int m = (int)ql, n = (int)rl;
    auto sub = [&](int i, int j) {
        int s = q[i - 1] == r[j - 1] ? sc.match : -sc.mismatch;
        if (i == m) s += sc.end_bonus;
        return s;
    };
    Dp d = fill(m, n, sc.gap_open, sc.gap_extend, sub);
    int bi = 0, bj = 0, best = 0;
    for (int i = 0; i <= m; ++i)
        for (int j = 0; j <= n; ++j) {
            int h = d.H[(size_t)i * (n + 1) + j];
            if (h > best) { best = h; bi = i; bj = j; }
        }
    Cigar c = trace(d, bi, bj, sc.gap_open, sc.gap_extend, sub, q.data(), r.data());
    res.score = best;
    if (gsle) {
        res.query_start = 0; res.query_end = (size_t)bi; res.ref_start = 0; res.ref_end = (size_t)bj;
        res.cigar = c;
    } else {
        res.query_start = ql - (size_t)bi; res.query_end = ql;
        res.ref_start = rl - (size_t)bj; res.ref_end = rl;
        res.cigar = c.reversed();
    }
    return res;
}

// src/piecewisealigner.rs:645-719
void remove_spurious_anchors(std::vector<Anchor>& a) {
    if (a.size() < 2) return;
    const long tol = 5;
    long tracked = 0;
    long dev_start = -1;
    size_t i = 1;
    while (i < a.size()) {
        long qd = (long)a[i - 1].query_start - (long)a[i].query_start;
        long rd = (long)a[i - 1].ref_start - (long)a[i].ref_start;
        long indel = qd - rd;
        if (std::labs(indel) > tol) {
            if (dev_start >= 0 && std::labs(tracked + indel) <= tol) {
                a.erase(a.begin() + dev_start, a.begin() + i);
                i = (size_t)dev_start;
                dev_start = -1;
                tracked = 0;
                continue;
            }
            dev_start = (long)i;
            tracked = indel;
        }
        ++i;
    }
    size_t maxp = (size_t)std::ceil((float)a.size() * 0.20f);
    for (size_t t = 1; t < std::min(a.size(), maxp); ++t) {
        long qd = (long)a[t].query_start - (long)a[t - 1].query_start;
        long rd = (long)a[t].ref_start - (long)a[t - 1].ref_start;
        if (qd - rd != 0) { a.erase(a.begin(), a.begin() + t); break; }
    }
    size_t lo = (a.size() > maxp ? a.size() - maxp : 0) + 1;
    for (size_t t = a.size(); t-- > lo;) {
        long qd = (long)a[t].query_start - (long)a[t - 1].query_start;
        long rd = (long)a[t].ref_start - (long)a[t - 1].ref_start;
        if (qd - rd != 0) { a.erase(a.begin() + t, a.end()); break; }
    }
}

// src/piecewisealigner.rs:396-510 (anchors: last anchor first)
std::optional<AlignmentInfo> extend_piecewise(const uint8_t* q, size_t ql, const uint8_t* r, size_t rl,
                                              const std::vector<Anchor>& anchors, size_t padding,
                                              const Scores& sc, size_t k, int xdrop) {
    if (anchors.empty()) return std::nullopt;
    Cigar cigar;
    int score = 0;
    size_t ref_start, query_start;
    {
        const Anchor& fa = anchors.back();
        if (fa.query_start > 0 && fa.ref_start > 0) {
            size_t qlen = fa.query_start;
            size_t rs = fa.ref_start > qlen + padding ? fa.ref_start - (qlen + padding) : 0;
            PwResult pre = pw_xdrop_alignment(q, qlen, r + rs, fa.ref_start - rs, sc, false, xdrop);
            if (pre.score == 0) { query_start = fa.query_start; ref_start = fa.ref_start; }
            else { cigar = pre.cigar; score = pre.score; query_start = pre.query_start; ref_start = pre.ref_start + rs; }
        } else {
            query_start = fa.query_start;
            ref_start = fa.ref_start;
            score = fa.query_start == 0 ? sc.end_bonus : 0;
        }
    }
    score += (int)k * sc.match;
    cigar.push(7, k);
    for (size_t i = anchors.size() - 1; i >= 1; --i) {
        size_t cqs = anchors[i - 1].query_start, crs = anchors[i - 1].ref_start;
        size_t pqe = anchors[i].query_start + k, pre = anchors[i].ref_start + k;
        long qd = (long)cqs - (long)pqe, rd = (long)crs - (long)pre;
        if (rd > 0 && qd > 0) {
            const uint8_t* qp = q + pqe;
            const uint8_t* rp = r + pre;
            bool done = false;
            if (rd == qd) {
                AlignmentInfo h = hamming_align_global(qp, rp, (size_t)qd, sc.match, sc.mismatch);
                uint32_t thr = (uint32_t)sc.match * (uint32_t)std::ceil((float)qd * 0.85f);
                if (h.score > 0 && h.score > thr) {
                    score += (int)h.score;
                    cigar.extend(h.cigar);
                    score += (int)k * sc.match;
                    cigar.push(7, k);
                    done = true;
                }
            }
            if (!done) {
                PwResult g = pw_global_alignment(qp, (size_t)qd, rp, (size_t)rd, sc);
                score += g.score;
                cigar.extend(g.cigar);
                score += (int)k * sc.match;
                cigar.push(7, k);
            }
        } else {
            if (rd < qd) {
                size_t ins = (size_t)(qd - rd);
                score += -sc.gap_open + ((int)ins - 1) * -sc.gap_extend;
                cigar.push(1, ins);
                size_t mp = (size_t)((long)k + rd);
                score += (int)mp * sc.match;
                cigar.push(7, mp);
            } else if (rd > qd) {
                size_t del = (size_t)(rd - qd);
                score += -sc.gap_open + ((int)del - 1) * -sc.gap_extend;
                cigar.push(2, del);
                size_t mp = (size_t)((long)k + qd);
                score += (int)mp * sc.match;
                cigar.push(7, mp);
            } else {
                size_t mp = (size_t)((long)k + qd);
                score += (int)mp * sc.match;
                cigar.push(7, mp);
            }
        }
    }
    size_t ref_end, query_end;
    {
        const Anchor& la = anchors.front();
        size_t lqe = la.query_start + k, lre = la.ref_start + k;
        if (lqe < ql && lre < rl) {
            size_t qlen = ql - lqe;
            size_t re = std::min(rl, lre + qlen + padding);
            PwResult post = pw_xdrop_alignment(q + lqe, qlen, r + lre, re - lre, sc, true, xdrop);
            if (post.score == 0) { query_end = lqe; ref_end = lre; }
            else { cigar.extend(post.cigar); score += post.score; query_end = post.query_end + lqe; ref_end = post.ref_end + lre; }
        } else {
            query_end = lqe;
            ref_end = lre;
            score += lqe == ql ? sc.end_bonus : 0;
        }
    }
    if (score <= 0) return std::nullopt;
    AlignmentInfo a;
    a.cigar = cigar;
    a.edit_distance = cigar.edit_distance();
    a.ref_start = ref_start; a.ref_end = ref_end; a.query_start = query_start; a.query_end = query_end;
    a.score = (uint32_t)score;
    return a;
}

}  // namespace orc
