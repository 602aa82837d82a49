// Per-request base-level extension: ungapped (Hamming) alignment, SSW with
// the reference's striped lane semantics, banded global traceback, and
// piecewise extension between chain anchors. One request per GPU thread;
// all working memory lives in a per-request scratch slot.
#pragma once
#include "sa_seed.hpp"

namespace sa {

struct Scores {
    int32_t match, mismatch, gap_open, gap_extend, end_bonus, xdrop;
};

// CIGAR writer with the merge-equal / drop-empty rule of src/cigar.rs:96-105.
// Ops are BAM-encoded (len << 4 | op); op codes 0 M, 1 I, 2 D, 4 S, 7 =, 8 X.
struct CigarOut {
    uint32_t* ops;
    uint32_t cap, n;
    bool overflow;
    uint32_t last = 15;  // op of the last entry (15: none); entries past cap are only counted
    SA_HD void reset() {
        n = 0;
        overflow = false;
        last = 15;
    }
    SA_HD void push(uint32_t op, uint32_t len) {
        if (len == 0) return;
        if (n > 0 && last == op) {
            if (n <= cap) ops[n - 1] += len << 4;
            return;
        }
        last = op;
        if (n >= cap) { overflow = true; ++n; return; }
        ops[n++] = (len << 4) | op;
    }
    SA_HD uint32_t edit_distance() const {
        uint32_t d = 0;
        for (uint32_t i = 0; i < n && i < cap; ++i) {
            uint32_t op = ops[i] & 0xF;
            if (op == 8 || op == 1 || op == 2) d += ops[i] >> 4;
        }
        return d;
    }
};

// Views of the query (read or its reverse complement) and of a reference window.
// Both keep the last aligned 8-byte block (query) or packed 64-bit word
// (reference) they read: consecutive accesses, the common pattern of the
// aligners, then cost no memory access. The read buffers they point into are
// padded by at least 8 bytes, so the aligned block read never leaves them.
struct QueryView {
    const uint8_t* seq;
    uint32_t len;
    bool rc;
    mutable uintptr_t cbi = ~(uintptr_t)0;  // address >> 3 of the cached block
    mutable uint64_t cb = 0;
    SA_HD uint8_t byte_at(uint32_t pos) const {
        const uintptr_t addr = (uintptr_t)(seq + pos);
        if ((addr >> 3) != cbi) {
            cbi = addr >> 3;
            cb = *(const uint64_t*)(addr & ~(uintptr_t)7);
        }
        return (uint8_t)(cb >> ((addr & 7) * 8));
    }
    SA_HD uint8_t operator[](uint32_t i) const { return rc ? comp_ascii(byte_at(len - 1 - i)) : byte_at(i); }
};
struct RefWin {
    const RefView* ref;
    uint32_t id;
    uint64_t start;  // absolute position of window index 0
    uint32_t len;
    mutable const uint64_t* base = nullptr;  // packed words of contig id
    mutable uint64_t cwi = ~0ull;            // index of the cached word
    mutable uint64_t cw = 0;
    SA_HD uint32_t code(uint32_t i) const {
        const uint64_t a = start + i;
        if ((a >> 5) != cwi) {
            if (!base) base = ref->words + ref->word_off[id];
            cwi = a >> 5;
            cw = base[cwi];
        }
        return (uint32_t)(cw >> ((a & 31) * 2)) & 3;
    }
    SA_HD uint8_t operator[](uint32_t i) const { return (uint8_t)((0x54474341u >> (8 * code(i))) & 0xFFu); }
};

// src/ssw.rs:8-22 translate: A/U -> 0, C 1, G 2, T 3, other 4
SA_HD int32_t ssw_code(uint8_t c) {
    return nt_is_acgt(c) ? (int32_t)nt_acgt_code(c) : (nt_is_u(c) ? 0 : 4);
}
// ssw_code of the complement of c (comp_ascii), without the ASCII round trip
SA_HD int32_t ssw_code_comp(uint8_t c) {
    return nt_is_acgt(c) ? 3 - (int32_t)nt_acgt_code(c) : (nt_is_u(c) ? 0 : 4);
}
SA_HD int32_t ssw_mat(int32_t a, int32_t b, const Scores& sc) { return (a == b && a < 4) ? sc.match : -sc.mismatch; }

// Maximum-sum segment summary: total, best prefix, best suffix and best
// segment (each of the last three at least 0, the empty segment).
struct SegSum {
    int32_t tot = 0, pre = 0, suf = 0, best = 0;
    SA_HD void push(int32_t v) {
        tot += v;
        pre = pre > tot ? pre : tot;
        suf = suf + v > 0 ? suf + v : 0;
        best = best > suf ? best : suf;
    }
    // this segment followed by b
    SA_HD void join(const SegSum& b) {
        const int32_t p = tot + b.pre > pre ? tot + b.pre : pre;
        const int32_t s = b.tot + suf > b.suf ? b.tot + suf : b.suf;
        int32_t m = best > b.best ? best : b.best;
        m = m > suf + b.pre ? m : suf + b.pre;
        tot += b.tot;
        pre = p;
        suf = s;
        best = m;
    }
};

// Gap lengths up to which two_diagonal_score joins its two diagonals.
constexpr int64_t TWO_DIAGONAL_MAX_GAP = 4096;

// Best score of an alignment of the query against the window [0, w.len) made of
// a segment on diagonal d1 (query position x against window position x + d1),
// one gap and a segment on diagonal d2 further down the query, or of a segment
// on either diagonal alone: a lower bound of the local score the SSW passes find
// (a gap of g positions costs gap_open + (g - 1) * gap_extend there; a cell off
// the window ends every segment through it). One pass: the d1 scan trails the
// d2 scan by the query offset at which the segment after the gap starts
// (1 when d2 > d1, a reference gap; 1 + g when d2 < d1, a query gap).
template <typename Q>
SA_HD int32_t two_diagonal_score(const Q& q, const RefWin& w, uint32_t L, int64_t d1, int64_t d2,
                                 const Scores& sc) {
    constexpr int32_t NEG = -(1 << 28);
    const int64_t g = d1 > d2 ? d1 - d2 : d2 - d1;
    const bool join = g > 0 && g <= TWO_DIAGONAL_MAX_GAP;
    const uint32_t off = !join ? 0u : d2 > d1 ? 1u : 1u + (uint32_t)g;
    const int32_t cost = join ? sc.gap_open + (int32_t)(g - 1) * sc.gap_extend : 0;
    const RefWin w1 = w, w2 = w;  // one cached reference word per diagonal
    auto cell = [&](const RefWin& wd, uint32_t x, int64_t d, bool& ok) -> int32_t {
        const int64_t p = (int64_t)x + d;
        ok = p >= 0 && p < (int64_t)wd.len;
        return !ok ? 0 : ssw_code(q[x]) == (int32_t)wd.code((uint32_t)p) ? sc.match : -sc.mismatch;
    };
    int32_t best = 0, a = NEG, s2 = NEG;
    const uint32_t zend = L + off;
    for (uint32_t z = 0; z < zend; ++z) {
        int32_t ax = NEG;  // best segment on d1 ending at z - off
        bool ok;
        if (z >= off && z - off < L) {
            const int32_t c = cell(w1, z - off, d1, ok);
            a = ok ? (a > 0 ? a : 0) + c : NEG;
            best = best > a ? best : a;
            ax = a;
        }
        if (z < L) {
            const int32_t c = cell(w2, z, d2, ok);
            int32_t from = s2 > 0 ? s2 : 0;
            if (join && ax > NEG / 2 && ax - cost > from) from = ax - cost;
            s2 = ok ? from + c : NEG;
            best = best > s2 ? best : s2;
        }
    }
    return best;
}

// Substitution scores of the SSW passes, called as score(reference code, query
// code): the reference's 5x5 match/mismatch matrix (src/ssw.rs:139-149), or an
// n x n matrix as ssw_init takes it (row = reference code, ext/ssw/ssw.c:179, 656).
struct MatchMismatch {
    int32_t match, mismatch;
    SA_HD int32_t operator()(int32_t r, int32_t q) const { return (r == q && r < 4) ? match : -mismatch; }
};
struct MatrixScore {
    const int8_t* m;
    int32_t n;
    SA_HD int32_t operator()(int32_t r, int32_t q) const { return m[r * n + q]; }
};

// ------------------------------------------------------------------ SSW ---
struct SswEnd {
    int32_t score, ref, read;
};

// One striped pass of ext/ssw/ssw.c:197-588 with LANES lanes per vector
// (16 unsigned bytes with bias 8, or 8 signed words without bias). The query
// is `qcode(j)` for j < qlen; the reference index runs forward over [0, rlen)
// or, for the reverse pass, backward from rlen-1 to 0. H/E/Hmax live in
// `buf` (4 * segLen * LANES ints).
// `bias` is the byte form's offset (minus the smallest substitution score); when
// `maxcol` is set it receives each processed column's maximum (0 for columns the
// pass did not reach), from which ssw_align picks its second-best end.
template <int LANES, bool BYTE, typename QC, typename RC, typename SM>
SA_HD SswEnd ssw_pass_sm(QC qcode, int32_t qlen, RC rcode, int32_t rlen, bool reverse, int32_t gapO, int32_t gapE,
                         int32_t byte_bias, const SM& score, int32_t terminate, int32_t* buf, int32_t* maxcol) {
    const int32_t bias = BYTE ? byte_bias : 0;
    if (maxcol)
        for (int32_t i = 0; i < rlen; ++i) maxcol[i] = 0;
    const int32_t segLen = (qlen + LANES - 1) / LANES;
    const int32_t W = segLen * LANES;
    int32_t* HS = buf;
    int32_t* HL = buf + W;
    int32_t* E = buf + 2 * W;
    int32_t* Hmax = buf + 3 * W;
    for (int32_t i = 0; i < 4 * W; ++i) buf[i] = 0;
    auto sat = [](int32_t v) { return BYTE ? (v < 0 ? 0 : (v > 255 ? 255 : v)) : (v < -32768 ? -32768 : (v > 32767 ? 32767 : v)); };
    auto subs = [](int32_t a, int32_t b) { int32_t v = a - b; return v < 0 ? 0 : v; };
    int32_t maxv = 0, end_read = qlen - 1, end_ref = BYTE ? -1 : 0;
    int32_t vMaxScore[LANES], vMaxMark[LANES];
    for (int l = 0; l < LANES; ++l) vMaxScore[l] = vMaxMark[l] = 0;
    int32_t begin = reverse ? rlen - 1 : 0, end = reverse ? -1 : rlen, step = reverse ? -1 : 1;
    for (int32_t i = begin; i != end; i += step) {
        int32_t vF[LANES], vMaxCol[LANES], vH[LANES];
        for (int l = LANES - 1; l > 0; --l) vH[l] = HS[(segLen - 1) * LANES + l - 1];
        vH[0] = 0;
        for (int l = 0; l < LANES; ++l) { vF[l] = 0; vMaxCol[l] = 0; }
        int32_t* t = HS; HS = HL; HL = t;
        const int32_t rc = rcode(i);
        for (int32_t j = 0; j < segLen; ++j) {
            for (int l = 0; l < LANES; ++l) {
                int32_t qi = j + l * segLen;
                int32_t prof = qi >= qlen ? 0 : score(rc, qcode(qi));
                int32_t h = BYTE ? subs(sat(vH[l] + prof + bias), bias) : sat(vH[l] + prof);
                int32_t e = E[j * LANES + l];
                h = h > e ? h : e;
                h = h > vF[l] ? h : vF[l];
                vMaxCol[l] = vMaxCol[l] > h ? vMaxCol[l] : h;
                HS[j * LANES + l] = h;
                int32_t hg = subs(h, gapO);
                int32_t en = subs(e, gapE);
                E[j * LANES + l] = en > hg ? en : hg;
                int32_t fn = subs(vF[l], gapE);
                vF[l] = fn > hg ? fn : hg;
                vH[l] = HL[j * LANES + l];
            }
        }
        bool done = false;
        for (int kk = 0; kk < LANES && !done; ++kk) {
            for (int l = LANES - 1; l > 0; --l) vF[l] = vF[l - 1];
            vF[0] = 0;
            for (int32_t j = 0; j < segLen; ++j) {
                bool all_le = true;
                for (int l = 0; l < LANES; ++l) {
                    int32_t h = HS[j * LANES + l];
                    h = h > vF[l] ? h : vF[l];
                    vMaxCol[l] = vMaxCol[l] > h ? vMaxCol[l] : h;
                    HS[j * LANES + l] = h;
                    int32_t hg = subs(h, gapO);
                    vF[l] = subs(vF[l], gapE);
                    if (vF[l] > hg) all_le = false;
                }
                if (all_le) { done = true; break; }
            }
        }
        bool changed = false;
        for (int l = 0; l < LANES; ++l) {
            vMaxScore[l] = vMaxScore[l] > vMaxCol[l] ? vMaxScore[l] : vMaxCol[l];
            if (vMaxMark[l] != vMaxScore[l]) changed = true;
        }
        if (changed) {
            int32_t temp = 0;
            for (int l = 0; l < LANES; ++l) { vMaxMark[l] = vMaxScore[l]; temp = temp > vMaxScore[l] ? temp : vMaxScore[l]; }
            if (temp > maxv) {
                maxv = temp;
                if (BYTE && maxv + bias >= 255) break;
                end_ref = i;
                for (int32_t x = 0; x < W; ++x) Hmax[x] = HS[x];
            }
        }
        int32_t colmax = 0;
        for (int l = 0; l < LANES; ++l) colmax = colmax > vMaxCol[l] ? colmax : vMaxCol[l];
        if (maxcol) maxcol[i] = colmax;
        if (colmax == terminate) break;
    }
    for (int32_t idx = 0; idx < W; ++idx) {
        if (Hmax[idx] == maxv) {
            int32_t t = idx / LANES + (idx % LANES) * segLen;
            if (t < end_read) end_read = t;
        }
    }
    SswEnd r;
    r.score = BYTE ? (maxv + bias >= 255 ? 255 : maxv) : maxv;
    r.ref = end_ref;
    r.read = end_read;
    return r;
}

template <int LANES, bool BYTE, typename QC, typename RC>
SA_HD SswEnd ssw_pass(QC qcode, int32_t qlen, RC rcode, int32_t rlen, bool reverse, const Scores& sc, int32_t terminate,
                      int32_t* buf) {
    return ssw_pass_sm<LANES, BYTE>(qcode, qlen, rcode, rlen, reverse, sc.gap_open, sc.gap_extend, sc.mismatch,
                                    MatchMismatch{sc.match, sc.mismatch}, terminate, buf, nullptr);
}

// A lane's working memory in 32-bit words: word k at p[k * stride]. With
// stride 64 the 64 lanes of a wave interleave their words, so that lanes at the
// same k touch one 256-byte block (one job per lane, coalesced); stride 1 is
// plain contiguous memory.
struct LaneMem {
    uint32_t* p;
    uint32_t stride;
    uint64_t words;
    SA_HD uint32_t& operator[](uint64_t k) const { return p[k * stride]; }
};

SA_HD LaneMem contiguous_mem(uint8_t* scratch, uint64_t bytes) {
    return LaneMem{(uint32_t*)scratch, 1, bytes / 4};
}

// Direction codes of the banded DP (ext/ssw/ssw.c:605-660), three per cell:
// E (2 or 3), F (4 or 5) and H (1, or the cell's E or F code). ssw.c keeps them
// in a byte array, cell j of row i at byte 3 * (width_d * i + j - x_i) + p.
// Here a cell is 4 bits (0: never written, else 1 + [E == 3] + 2 [F == 5] +
// 4 kind, kind 0 H = 1, 1 H = E, 2 H = F), 8 cells per word, a band row in
// ceil(width_d / 8) words; a traceback read outside a row's cells finds the
// cell and code that ssw.c's byte index names, so reads that leave the band
// resolve as they do there.
SA_HD uint32_t dir_cell(int32_t de, int32_t df, bool diag, bool e_over_f) {
    const uint32_t kind = diag ? 0u : (e_over_f ? 1u : 2u);
    return 1u + (de == 3 ? 1u : 0u) + (df == 5 ? 2u : 0u) + 4u * kind;
}

struct DirView {
    LaneMem m;
    uint64_t base;    // first word of the rows
    int32_t width_d;  // cells per band row
    int32_t wd;       // words per band row
    int32_t rows;
    // code `part` of cell `cell` of row `row`: ssw.c's byte
    // 3 * (width_d * row + cell) + part, which lies in another row when cell is
    // outside [0, width_d)
    SA_HD int32_t code(int32_t row, int32_t cell, int32_t part) const {
        if (cell < 0 || cell >= width_d) {
            const int64_t byte = 3ll * ((int64_t)width_d * row + cell) + part;
            if (byte < 0) return 0;
            const int64_t rb = 3ll * width_d;
            row = (int32_t)(byte / rb);
            const int32_t rem = (int32_t)(byte % rb);
            cell = rem / 3;
            part = rem % 3;
        }
        if (row >= rows) return 0;
        const uint32_t n = (m[base + (uint64_t)row * wd + (uint32_t)cell / 8] >> (4 * (cell % 8))) & 15u;
        if (n == 0) return 0;
        const uint32_t v = n - 1;
        const int32_t de = 2 + (int32_t)(v & 1), df = 4 + (int32_t)((v >> 1) & 1), kind = (int32_t)(v >> 2);
        return part == 0 ? de : part == 1 ? df : (kind == 0 ? 1 : kind == 1 ? de : df);
    }
};

// One pass of the banded DP of ext/ssw/ssw.c:605-660 for band width bw <= BWMAX
// with the three band rows (previous H, E carried in place, current H) in
// registers: entry u of row i holds column j = x_i + u - 1, x_i = max(i - bw, 0),
// so the previous row's entries for column j and j - 1 sit at u + (x_i - x_{i-1})
// and one below, and every index is static once the cell loop is unrolled over u.
// The reference codes of the row's columns x_i .. x_i + 2*BWMAX are kept in
// registers too and slide by one when x_i advances, so each row reads one new
// reference base. Each row's direction cells are packed in registers and stored
// as whole words (cells past the row's end as 0, as ssw.c's cleared array reads
// there). Returns the same maximum as the scratch form (band_pass_mem),
// including its resets of the row entries at 0 and `edge`.
template <int BWMAX, typename QC, typename RC, typename SM>
SA_HD int32_t band_pass_reg(QC qc, RC rc, int32_t refLen, int32_t readLen, const Scores& sc, const SM& score,
                            int32_t bw, const DirView& dv) {
    constexpr int WM = 2 * BWMAX + 3;  // largest row width
    constexpr int UM = 2 * BWMAX + 1;  // largest number of cells in a row
    constexpr int WDM = (UM + 7) / 8;  // largest number of words in a row
    const int32_t gapO = sc.gap_open, gapE = sc.gap_extend;
    const int32_t width = bw * 2 + 3;
    int32_t hb[WM + 1], eb[WM + 1], hc[WM + 1];
    SA_UNROLL
    for (int x = 0; x <= WM; ++x) hb[x] = eb[x] = hc[x] = 0;
    int32_t rw[UM];  // rw[u - 1]: reference code of column x_i + u - 1 (0 past the end)
    SA_UNROLL
    for (int u = 0; u < UM; ++u) rw[u] = u < refLen ? rc(u) : 0;
    int32_t maxv = 0, xprev = 0;
    for (int32_t i = 0; i < readLen; ++i) {
        const int32_t xi = i - bw > 0 ? i - bw : 0;
        const bool shift = xi != xprev;
        xprev = xi;
        if (shift) {
            SA_UNROLL
            for (int u = 0; u < UM - 1; ++u) rw[u] = rw[u + 1];
            rw[UM - 1] = xi + UM - 1 < refLen ? rc(xi + UM - 1) : 0;
        }
        const int32_t end = refLen - 1 < i + bw ? refLen - 1 : i + bw;
        const int32_t edge = end + 1 < width - 1 ? end + 1 : width - 1;
        hb[0] = eb[0] = hc[0] = 0;
        SA_UNROLL
        for (int x = 1; x < WM; ++x)
            if (x == edge) hb[x] = eb[x] = 0;
        const int32_t uend = end - xi + 1;
        const int32_t qi = qc(i);
        int32_t f = 0;
        uint32_t wv[WDM];
        SA_UNROLL
        for (int w = 0; w < WDM; ++w) wv[w] = 0;
        SA_UNROLL
        for (int u = 1; u <= UM; ++u) {
            if (u <= uend) {
                const int32_t hbe = shift ? hb[u + 1] : hb[u];
                const int32_t ebe = shift ? eb[u + 1] : eb[u];
                const int32_t hbd = shift ? hb[u] : hb[u - 1];
                int32_t t1 = i == 0 ? -gapO : hbe - gapO;
                int32_t t2 = i == 0 ? -gapE : ebe - gapE;
                const int32_t ev = t1 > t2 ? t1 : t2;
                eb[u] = ev;
                const int32_t de = t1 > t2 ? 3 : 2;
                t1 = hc[u - 1] - gapO;
                t2 = f - gapE;
                f = t1 > t2 ? t1 : t2;
                const int32_t df = t1 > t2 ? 5 : 4;
                const int32_t e1 = ev > 0 ? ev : 0, f1 = f > 0 ? f : 0;
                t1 = e1 > f1 ? e1 : f1;
                t2 = hbd + score(rw[u - 1], qi);
                const int32_t h = t1 > t2 ? t1 : t2;
                hc[u] = h;
                maxv = h > maxv ? h : maxv;
                wv[(u - 1) / 8] |= dir_cell(de, df, t1 <= t2, e1 > f1) << (4 * ((u - 1) % 8));
            }
        }
        const uint64_t rowp = dv.base + (uint64_t)i * dv.wd;
        SA_UNROLL
        for (int w = 0; w < WDM; ++w)
            if (w < dv.wd) dv.m[rowp + w] = wv[w];
        SA_UNROLL
        for (int u = 1; u <= UM; ++u)
            if (u <= uend) hb[u] = hc[u];
    }
    return maxv;
}

// The same pass with the band rows in the lane's memory (any band width).
template <typename QC, typename RC, typename SM>
SA_HD int32_t band_pass_mem(QC qc, RC rc, int32_t refLen, int32_t readLen, const Scores& sc, const SM& score,
                            int32_t bw, LaneMem m, uint64_t hb, uint64_t eb, uint64_t hc, const DirView& dv) {
    const int32_t gapO = sc.gap_open, gapE = sc.gap_extend;
    const int32_t width = bw * 2 + 3;
    auto H = [&](uint64_t a, int32_t x) -> uint32_t& { return m[a + (uint32_t)x]; };
    auto U = [&](int32_t i, int32_t j) { int32_t x = i - bw; x = x > 0 ? x : 0; return j - x + 1; };
    int32_t maxv = 0;
    for (int32_t x = 0; x <= width; ++x) H(hb, x) = H(eb, x) = H(hc, x) = 0;
    for (int32_t i = 0; i < readLen; ++i) {
        int32_t beg = i - bw > 0 ? i - bw : 0, end = refLen - 1 < i + bw ? refLen - 1 : i + bw;
        int32_t edge = end + 1 < width - 1 ? end + 1 : width - 1;
        int32_t f = 0;
        H(hb, 0) = H(eb, 0) = H(hb, edge) = H(eb, edge) = H(hc, 0) = 0;
        int32_t u = 0;
        const int32_t qi = qc(i);
        const uint64_t rowp = dv.base + (uint64_t)i * dv.wd;
        uint32_t word = 0;
        int32_t wi = 0;  // words of the row stored
        for (int32_t j = beg; j <= end; ++j) {
            u = U(i, j);
            int32_t e = U(i - 1, j), b = U(i, j - 1), d = U(i - 1, j - 1);
            int32_t t1 = i == 0 ? -gapO : (int32_t)H(hb, e) - gapO;
            int32_t t2 = i == 0 ? -gapE : (int32_t)H(eb, e) - gapE;
            const int32_t ev = t1 > t2 ? t1 : t2;
            H(eb, u) = (uint32_t)ev;
            const int32_t de = t1 > t2 ? 3 : 2;
            t1 = (int32_t)H(hc, b) - gapO;
            t2 = f - gapE;
            f = t1 > t2 ? t1 : t2;
            const int32_t df = t1 > t2 ? 5 : 4;
            int32_t e1 = ev > 0 ? ev : 0, f1 = f > 0 ? f : 0;
            t1 = e1 > f1 ? e1 : f1;
            t2 = (int32_t)H(hb, d) + score(rc(j), qi);
            const int32_t h = t1 > t2 ? t1 : t2;
            H(hc, u) = (uint32_t)h;
            if (h > maxv) maxv = h;
            word |= dir_cell(de, df, t1 <= t2, e1 > f1) << (4 * ((u - 1) % 8));
            if ((u - 1) % 8 == 7) {
                dv.m[rowp + (uint32_t)wi++] = word;
                word = 0;
            }
        }
        // the partial word and the rest of the row: cells never written
        for (; wi < dv.wd; ++wi) {
            dv.m[rowp + (uint32_t)wi] = word;
            word = 0;
        }
        for (int32_t j = 1; j <= u; ++j) H(hb, j) = H(hc, j);
    }
    return maxv;
}

// Banded global alignment with band doubling and traceback (ext/ssw/ssw.c:590-774)
// in the lane's memory m. Emits M/I/D ops in forward order through
// `emit(op, len)`. Returns 0 ok, 1 traceback failure, 2 memory too small.
template <typename QC, typename RC, typename Emit, typename SM>
SA_HD int banded_sw(QC qc, RC rc, int32_t refLen, int32_t readLen, int32_t score, const Scores& sc, const SM& sm,
                    int32_t bw, LaneMem m, Emit emit) {
    int32_t len = refLen > readLen ? refLen : readLen;
    int32_t maxv = 0;
    int32_t width = 0, width_d = 0;
    DirView dv{m, 0, 0, 0, readLen};
    uint64_t used = 0;
    do {
        width = bw * 2 + 3;
        width_d = bw * 2 + 1;
        // words: the three band rows of the memory form, then the direction rows
        const uint64_t rows_at = 3ull * (uint64_t)(width + 1);
        dv = DirView{m, rows_at, width_d, (width_d + 7) / 8, readLen};
        used = rows_at + (uint64_t)dv.wd * (uint64_t)readLen;
        if (used > m.words) return 2;
        if (bw <= 1) maxv = band_pass_reg<1>(qc, rc, refLen, readLen, sc, sm, bw, dv);
        else if (bw <= 2) maxv = band_pass_reg<2>(qc, rc, refLen, readLen, sc, sm, bw, dv);
        else if (bw <= 4) maxv = band_pass_reg<4>(qc, rc, refLen, readLen, sc, sm, bw, dv);
        else if (bw <= 8) maxv = band_pass_reg<8>(qc, rc, refLen, readLen, sc, sm, bw, dv);
        else if (bw <= 16) maxv = band_pass_reg<16>(qc, rc, refLen, readLen, sc, sm, bw, dv);
        else if (bw <= 32) maxv = band_pass_reg<32>(qc, rc, refLen, readLen, sc, sm, bw, dv);
        else maxv = band_pass_mem(qc, rc, refLen, readLen, sc, sm, bw, m, 0, (uint64_t)(width + 1),
                                  2ull * (uint64_t)(width + 1), dv);
        bw *= 2;
    } while (maxv < score && bw <= len);
    bw /= 2;
    // traceback (ops collected backwards after the direction rows, emitted forwards)
    if (m.words - used < 4) return 2;
    const uint64_t room = m.words - used - 1;
    const uint64_t st = used;
    uint64_t nst = 0;
    int32_t i = readLen - 1, j = refLen - 1, e = 0, t2 = 2, row = readLen - 1;
    char op = 'M', prev = 'M';
    auto code = [](char o) { return o == 'M' ? 0u : (o == 'I' ? 1u : 2u); };
    while (i >= 0 && j > 0) {
        const int32_t x = i - bw > 0 ? i - bw : 0;
        const int32_t v = dv.code(row, j - x, t2);
        switch (v) {
            case 1: --i; --j; t2 = 2; --row; op = 'M'; break;
            case 2: --i; t2 = 0; --row; op = 'I'; break;
            case 3: --i; t2 = 2; --row; op = 'I'; break;
            case 4: --j; t2 = 1; op = 'D'; break;
            case 5: --j; t2 = 2; op = 'D'; break;
            default: return 1;
        }
        if (op == prev) ++e;
        else {
            if (nst >= room) return 2;
            m[st + nst++] = ((uint32_t)e << 4) | code(prev);
            prev = op;
            e = 1;
        }
    }
    if (nst + 2 >= room) return 2;
    if (op == 'M') m[st + nst++] = ((uint32_t)(e + 1) << 4) | code(op);
    else { m[st + nst++] = ((uint32_t)e << 4) | code(op); m[st + nst++] = (1u << 4) | 0u; }
    for (uint64_t x = nst; x-- > 0;) {
        const uint32_t v = m[st + x];
        emit(v & 0xF, v >> 4);
    }
    return 0;
}

struct AlignInfo {
    bool found;
    uint32_t score, query_start, query_end, ref_start, ref_end, edit_distance;
    int status;  // 0 ok, 1 cigar overflow, 2 scratch overflow
};

// Outcome of the SSW score passes (ext/ssw/ssw.c:818-888): best score, its end
// (first column reaching it, then smallest query row) and the start found by the
// reverse pass.
struct SswEnds {
    int32_t score1, ref_end1, read_end1, ref_begin1, read_begin1;
    bool rev_short;  // score1 > reverse-pass score (the reference returns flag 2)
};

// Banded traceback, with_eqx CIGAR and the end-bonus extensions (ext/ssw/ssw.c:
// 895-915 then src/ssw.rs:181-188 and src/aligner.rs:90-148), given the passes' result.
SA_HD AlignInfo ssw_finish(const QueryView& q, const RefWin& r, const SswEnds& e, const Scores& sc, LaneMem m,
                           CigarOut& cig) {
    AlignInfo a{false, 0, 0, 0, 0, 0, 0, 0};
    cig.reset();
    const int32_t score1 = e.score1, ref_end1 = e.ref_end1, read_end1 = e.read_end1;
    const int32_t ref_begin1 = e.ref_begin1, read_begin1 = e.read_begin1;
    if (ref_begin1 < 0 || read_begin1 < 0) return a;
    if (e.rev_short) return a;
    const int32_t refL = ref_end1 - ref_begin1 + 1, readL = read_end1 - read_begin1 + 1;
    int32_t bw = refL > readL ? refL - readL + 1 : readL - refL + 1;
    // decide the end-bonus extensions first so the CIGAR can be written front to back
    uint32_t qs = (uint32_t)read_begin1, rs = (uint32_t)ref_begin1;
    uint32_t qe = (uint32_t)read_end1 + 1, re = (uint32_t)ref_end1 + 1;
    // The SSW part is written after room reserved for a possible front extension
    // (at most qs ops), then moved down once the extensions are decided.
    uint32_t qi = qs, ri = rs, edits = 0;
    CigarOut mainc = cig;
    uint32_t front_room = qs <= rs ? qs : 0;
    if (front_room >= cig.cap) { a.status = 1; return a; }
    mainc.ops = cig.ops + front_room;
    mainc.cap = cig.cap - front_room;
    mainc.reset();
    // emit main alignment into mainc
    {
        auto emit_main = [&](uint32_t op, uint32_t len) {
            if (op == 0) {
                for (uint32_t t = 0; t < len; ++t) {
                    bool eq = q[qi] == r[ri];
                    mainc.push(eq ? 7u : 8u, 1);
                    edits += eq ? 0 : 1;
                    ++qi; ++ri;
                }
            } else if (op == 1) {
                mainc.push(1, len); qi += len; edits += len;
            } else {
                mainc.push(2, len); ri += len; edits += len;
            }
        };
        // reference bases are 2-bit codes, which are their SSW codes
        int st = banded_sw([&](int32_t i) { return ssw_code(q[qs + (uint32_t)i]); },
                           [&](int32_t j) { return (int32_t)r.code(rs + (uint32_t)j); }, refL, readL, score1, sc,
                           MatchMismatch{sc.match, sc.mismatch}, bw, m, emit_main);
        if (st == 1) return a;  // flag 1: no alignment
        if (st == 2) { a.status = 2; return a; }
    }
    if (mainc.overflow) { a.status = 1; return a; }
    int32_t score = score1;
    // front extension towards query start
    uint32_t fq = qs, fr = rs;
    int32_t fscore = score;
    uint32_t fedits = edits;
    while (fq > 0 && fr > 0) {
        --fq; --fr;
        if (q[fq] == r[fr]) fscore += sc.match; else { fscore -= sc.mismatch; ++fedits; }
    }
    bool front = fq == 0 && fscore + sc.end_bonus > score;
    // final CIGAR: [front ops][main ops][back ops]
    cig.reset();
    if (front && qs > 0) {
        for (uint32_t t = 0; t < qs; ++t) cig.push(q[t] == r[rs - qs + t] ? 7u : 8u, 1);
    }
    // append main ops (they were written at an offset; copy down with merging)
    for (uint32_t x = 0; x < mainc.n; ++x) {
        uint32_t v = mainc.ops[x];
        cig.push(v & 0xF, v >> 4);
    }
    if (front) {
        qs = 0;
        rs = fr;
        score = fscore + sc.end_bonus;
        edits = fedits;
    }
    // back extension towards query end
    uint32_t bq = qe, br2 = re;
    int32_t bscore = score;
    uint32_t bedits = edits;
    while (bq < q.len && br2 < r.len) {
        if (q[bq] == r[br2]) bscore += sc.match; else { bscore -= sc.mismatch; ++bedits; }
        ++bq; ++br2;
    }
    if (bq == q.len && bscore + sc.end_bonus > score) {
        if (qe < q.len)
            for (uint32_t t = 0; t < bq - qe; ++t) cig.push(q[qe + t] == r[re + t] ? 7u : 8u, 1);
        qe = q.len;
        re = br2;
        score = bscore + sc.end_bonus;
        edits = bedits;
    }
    if (cig.overflow) { a.status = 1; return a; }
    a.found = true;
    a.score = (uint32_t)score;
    a.query_start = qs;
    a.query_end = qe;
    a.ref_start = rs;
    a.ref_end = re;
    a.edit_distance = edits;
    return a;
}

// Scalar SSW: both passes run in one thread (a loop over the vector lanes), then
// ssw_finish. Used for queries longer than the lane-parallel kernel supports.
SA_HD AlignInfo ssw_aligner_align(const QueryView& q, const RefWin& r, const Scores& sc, uint8_t* scratch,
                                  uint64_t scratch_bytes, CigarOut& cig) {
    AlignInfo a{false, 0, 0, 0, 0, 0, 0, 0};
    cig.reset();
    if (r.len > 2000 || q.len == 0) return a;
    const int32_t qlen = (int32_t)q.len, rlen = (int32_t)r.len;
    uint64_t ssw_bytes = 4ull * 4 * (uint64_t)(((qlen + 15) / 16) * 16);
    ssw_bytes = (ssw_bytes + 255) & ~255ull;
    if (ssw_bytes + 64 > scratch_bytes) { a.status = 2; return a; }
    int32_t* buf = (int32_t*)scratch;
    auto qf = [&](int32_t j) { return ssw_code(q[(uint32_t)j]); };
    auto rf = [&](int32_t i) { return ssw_code(r[(uint32_t)i]); };
    SswEnd best = ssw_pass<16, true>(qf, qlen, rf, rlen, false, sc, 255, buf);
    bool word = false;
    if (best.score == 255) {
        best = ssw_pass<8, false>(qf, qlen, rf, rlen, false, sc, 65535, buf);
        word = true;
    }
    SswEnds e;
    e.score1 = best.score;
    e.ref_end1 = best.ref;
    e.read_end1 = best.read;
    auto qrev = [&](int32_t j) { return ssw_code(q[(uint32_t)(e.read_end1 - j)]); };
    SswEnd br = word ? ssw_pass<8, false>(qrev, e.read_end1 + 1, rf, e.ref_end1 + 1, true, sc, e.score1, buf)
                     : ssw_pass<16, true>(qrev, e.read_end1 + 1, rf, e.ref_end1 + 1, true, sc, (uint8_t)e.score1, buf);
    e.ref_begin1 = br.ref;
    e.read_begin1 = e.read_end1 - br.read;
    e.rev_short = e.score1 > br.score;
    return ssw_finish(q, r, e, sc, contiguous_mem(scratch + ssw_bytes, scratch_bytes - ssw_bytes), cig);
}

// Outcome of ssw_align (the s_align fields of ext/ssw/ssw.h; the CIGAR goes to
// the caller's buffer).
struct SswRaw {
    int32_t score1, score2, ref_begin1, ref_end1, read_begin1, read_end1, ref_end2;
    int32_t n_cigar;  // ops written
    int32_t flag;     // s_align.flag: 1 traceback failed, 2 the reverse pass fell short
    int32_t status;   // 0 result; 1 ssw_align returns NULL (no profile for the scores it needs);
                      // 2 scratch too small; 3 CIGAR buffer too small
};

// Bytes of scratch ssw_align_raw needs at most (passes, column maxima, and the
// banded DP at the widest band the doubling can reach).
SA_HD uint64_t ssw_raw_scratch(int32_t readLen, int32_t refLen) {
    const uint64_t W = (uint64_t)((readLen + 15) / 16) * 16;
    const uint64_t len = (uint64_t)(readLen > refLen ? readLen : refLen);
    const uint64_t bw = len + 1;  // the band doubling stops once the width exceeds len
    const uint64_t band = 3 * (2 * bw + 4) * 4 + (2 * bw + 1) * (uint64_t)readLen * 3 + 3 + 8 * (uint64_t)(readLen + refLen + 8);
    return 16 * W + 4 * (uint64_t)refLen + band + 64;
}

// ssw_align (ext/ssw/ssw.c:818-922) on translated codes with an n x n matrix:
// the forward pass (byte lanes, or word lanes when score_size is 1, or when the
// byte scores saturate and score_size is 2), the second-best end outside maskLen
// columns of the best (ssw.c:357-380, 560-583), the reverse pass for the start, and the
// banded traceback as M/I/D ops (len << 4 | op), each step gated by `flag`,
// `filters` and `filterd` as there.
SA_HD void ssw_align_raw(const int8_t* read, int32_t readLen, const int8_t* ref, int32_t refLen, const int8_t* mat,
                         int32_t n, int32_t score_size, int32_t gapO, int32_t gapE, uint32_t flag, uint32_t filters,
                         int32_t filterd, int32_t maskLen, uint8_t* scratch, uint64_t scratch_bytes, uint32_t* cigar,
                         uint32_t cigar_cap, SswRaw& o) {
    o = SswRaw{0, 0, -1, 0, -1, 0, 0, 0, 0, 0};
    const bool has_byte = score_size == 0 || score_size == 2, has_word = score_size == 1 || score_size == 2;
    if (!has_byte && !has_word) { o.status = 1; return; }
    const uint64_t W = (uint64_t)((readLen + 15) / 16) * 16;
    if (16 * W + 4 * (uint64_t)refLen + 64 > scratch_bytes) { o.status = 2; return; }
    int32_t* buf = (int32_t*)scratch;
    int32_t* maxcol = buf + 4 * W;
    uint8_t* rest = (uint8_t*)(maxcol + refLen);
    rest = (uint8_t*)(((uintptr_t)rest + 15) & ~(uintptr_t)15);
    const uint64_t rest_bytes = scratch_bytes - (uint64_t)(rest - scratch);
    int32_t bias = 0;
    for (int32_t i = 0; i < n * n; ++i) bias = mat[i] < bias ? mat[i] : bias;
    bias = -bias;
    const MatrixScore sm{mat, n};
    auto qf = [&](int32_t j) { return (int32_t)read[j]; };
    auto rf = [&](int32_t i) { return (int32_t)ref[i]; };
    SswEnd best;
    bool word = false;
    if (has_byte) {
        best = ssw_pass_sm<16, true>(qf, readLen, rf, refLen, false, gapO, gapE, bias, sm, 255, buf, maxcol);
        if (best.score == 255) {
            if (!has_word) { o.status = 1; return; }
            best = ssw_pass_sm<8, false>(qf, readLen, rf, refLen, false, gapO, gapE, 0, sm, 65535, buf, maxcol);
            word = true;
        }
    } else {
        best = ssw_pass_sm<8, false>(qf, readLen, rf, refLen, false, gapO, gapE, 0, sm, 65535, buf, maxcol);
        word = true;
    }
    o.score1 = best.score;
    o.ref_end1 = best.ref;
    o.read_end1 = best.read;
    if (maskLen >= 15) {
        int32_t s2 = 0, r2 = 0;
        int32_t edge = best.ref - maskLen > 0 ? best.ref - maskLen : 0;
        for (int32_t i = 0; i < edge; ++i)
            if (maxcol[i] > s2) { s2 = maxcol[i]; r2 = i; }
        edge = best.ref + maskLen > refLen ? refLen : best.ref + maskLen;
        // the byte pass scans from edge + 1, the word pass from edge (ssw.c:374 and 578)
        for (int32_t i = word ? edge : edge + 1; i < refLen; ++i)
            if (maxcol[i] > s2) { s2 = maxcol[i]; r2 = i; }
        o.score2 = s2;
        o.ref_end2 = r2;
    } else {
        o.score2 = 0;
        o.ref_end2 = -1;
    }
    if (flag == 0 || (flag == 2 && (uint32_t)o.score1 < filters)) return;
    const int32_t re1 = o.read_end1;
    auto qrev = [&](int32_t j) { return (int32_t)read[re1 - j]; };
    const SswEnd br = word ? ssw_pass_sm<8, false>(qrev, re1 + 1, rf, o.ref_end1 + 1, true, gapO, gapE, 0, sm,
                                                   (uint16_t)o.score1, buf, nullptr)
                           : ssw_pass_sm<16, true>(qrev, re1 + 1, rf, o.ref_end1 + 1, true, gapO, gapE, bias, sm,
                                                   (uint8_t)o.score1, buf, nullptr);
    o.ref_begin1 = br.ref;
    o.read_begin1 = re1 - br.read;
    if (o.score1 > br.score) o.flag = 2;
    if ((7 & flag) == 0 || ((2 & flag) != 0 && (uint32_t)o.score1 < filters) ||
        ((4 & flag) != 0 && (o.ref_end1 - o.ref_begin1 > filterd || o.read_end1 - o.read_begin1 > filterd)))
        return;
    const int32_t rL = o.ref_end1 - o.ref_begin1 + 1, qL = o.read_end1 - o.read_begin1 + 1;
    const int32_t bw = (rL > qL ? rL - qL : qL - rL) + 1;
    const int32_t qb = o.read_begin1, rb = o.ref_begin1;
    uint32_t nc = 0;
    bool over = false;
    auto emit = [&](uint32_t op, uint32_t len) {
        if (nc < cigar_cap) cigar[nc] = (len << 4) | op;
        else over = true;
        ++nc;
    };
    const Scores gaps{0, 0, gapO, gapE, 0, 0};
    const int st = banded_sw([&](int32_t i) { return (int32_t)read[qb + i]; }, [&](int32_t j) { return (int32_t)ref[rb + j]; },
                             rL, qL, o.score1, gaps, sm, bw, contiguous_mem(rest, rest_bytes), emit);
    if (st == 2) { o.status = 2; return; }
    if (st == 1) { o.flag = 1; return; }
    if (over) { o.status = 3; return; }
    o.n_cigar = (int32_t)nc;
}

// src/aligner.rs:181-259 (segment scores with the end bonus at both ends)
template <typename Q, typename R>
SA_HD AlignInfo hamming_align(const Q& q, const R& r, uint32_t n, const Scores& sc, CigarOut& cig) {
    uint32_t start = 0, bs = 0, be = 0;
    int32_t score = sc.end_bonus, best = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (q[i] == r[i]) score += sc.match; else score -= sc.mismatch;
        if (score < 0) { start = i + 1; score = 0; }
        if (score > best) { bs = start; best = score; be = i + 1; }
    }
    if (score + sc.end_bonus > best) { best = score + sc.end_bonus; be = n; bs = start; }
    cig.reset();
    uint32_t mm = 0;
    for (uint32_t i = bs; i < be; ++i) {
        bool eq = q[i] == r[i];
        mm += eq ? 0 : 1;
        cig.push(eq ? 7u : 8u, 1);
    }
    AlignInfo a{true, (uint32_t)best, bs, be, bs, be, mm, cig.overflow ? 1 : 0};
    return a;
}

// hamming_align over a mismatch bit mask (bit i set: position i mismatches),
// visiting mismatches only: between two mismatches the score rises by `match`
// per base, so the running maximum of a run of matches is reached at its last
// base, and a mismatch can only lower the score (and reset it below zero).
// Same result as hamming_align (src/aligner.rs:181-259) when match > 0; callers
// use hamming_align otherwise.
SA_HD AlignInfo hamming_align_bits(const uint64_t* m, uint32_t n, const Scores& sc, CigarOut& cig) {
    uint32_t start = 0, bs = 0, be = 0;
    int32_t score = sc.end_bonus, best = 0;
    uint32_t prev = 0;  // first position after the previous mismatch
    auto run_to = [&](uint32_t end) {  // matches at [prev, end)
        if (end > prev) {
            score += (int32_t)(end - prev) * sc.match;
            if (score > best) { bs = start; best = score; be = end; }
        }
    };
    for (uint32_t w = 0; w * 64 < n; ++w) {
        uint64_t bits = m[w];
        while (bits) {
            const uint32_t x = w * 64 + (uint32_t)__builtin_ctzll(bits);
            bits &= bits - 1;
            if (x >= n) break;
            run_to(x);
            score -= sc.mismatch;
            if (score < 0) { start = x + 1; score = 0; }
            if (score > best) { bs = start; best = score; be = x + 1; }  // only while the end bonus lasts
            prev = x + 1;
        }
    }
    run_to(n);
    if (score + sc.end_bonus > best) { best = score + sc.end_bonus; be = n; bs = start; }
    cig.reset();
    uint32_t mm = 0, p = bs;
    for (uint32_t w = bs / 64; w * 64 < be; ++w) {
        uint64_t bits = m[w];
        while (bits) {
            const uint32_t x = w * 64 + (uint32_t)__builtin_ctzll(bits);
            bits &= bits - 1;
            if (x < bs) continue;
            if (x >= be) break;
            if (x > p) cig.push(7u, x - p);
            cig.push(8u, 1);
            ++mm;
            p = x + 1;
        }
    }
    if (be > p) cig.push(7u, be - p);
    AlignInfo a{true, (uint32_t)best, bs, be, bs, be, mm, cig.overflow ? 1 : 0};
    return a;
}

// --------------------------------------------------------- piecewise DP ---
// Affine-gap DP (gap of length L costs open + (L-1)*extend) with a global start
// at (0,0); `gsle`: local end (max cell, end bonus on the last query row);
// otherwise global (cell (m,n)). Traceback preference at ties, walking back:
// deletion, then insertion, then diagonal; a gap run closes as soon as opening
// it explains the score. Ops are written to `tb` in traceback order.
struct DpOut {
    int32_t score;
    uint32_t bi, bj;  // end cell
    uint32_t ntb;     // ops in tb
    int status;
};

// One DP column as the row being computed sees it: H and F (updated in place)
// and the column's reference base, read once per DP.
struct alignas(16) DpCol {
    int32_t h, f, r, pad;
};
// Where the DP keeps its row and its direction bytes: row element j at
// p[j * stride] for j < cap, direction byte of cell (i, j) at
// dir[(i * cap + j) * stride] for i < dir_rows. The piecewise kernel gives each
// job one block of each (stride 1), filled by the job's lane group; without
// blocks both are placed in the scratch (the direction bytes first, row pitch
// n + 1).
struct DpRows {
    DpCol* p = nullptr;
    uint32_t stride = 1, cap = 0;
    uint8_t* dir = nullptr;
    uint32_t dir_rows = 0;
};

template <typename Q, typename R>
SA_HD DpOut affine_dp(const Q& q, uint32_t m, const R& r, uint32_t n, const Scores& sc, bool local_end,
                      uint8_t* scratch, uint64_t scratch_bytes, uint8_t* tb, uint32_t tb_cap,
                      DpRows rows = DpRows{}) {
    DpOut o{0, 0, 0, 0, 0};
    const int32_t go = sc.gap_open, ge = sc.gap_extend;
    const int32_t NEG = -(1 << 28);
    uint64_t W = (uint64_t)n + 1;
    const uint64_t dir_bytes = ((uint64_t)(m + 1) * W + 15) & ~15ull;
    if (!rows.p) {
        if (dir_bytes + W * sizeof(DpCol) + 16 > scratch_bytes) { o.status = 2; return o; }
        rows = DpRows{(DpCol*)(scratch + dir_bytes), 1, (uint32_t)W, scratch, m + 1};
    } else if (W > rows.cap || m + 1 > rows.dir_rows) {
        o.status = 2;
        return o;
    }
    // direction bytes: bit0 H==E, bit1 H==F, bit2 E opened, bit3 F opened
    const uint64_t st = rows.stride, pitch = rows.cap;
    uint8_t* const dirs = rows.dir;
    auto dir_at = [&](uint64_t i, uint64_t j) -> uint8_t& { return dirs[(i * pitch + j) * st]; };
    DpCol* row = rows.p;
    int32_t best = 0;
    uint32_t bi = 0, bj = 0;
    // row 0
    row[0] = DpCol{0, NEG, 0, 0};
    dir_at(0, 0) = 0;
    for (uint32_t j = 1; j <= n; ++j) {
        const int32_t h0 = -(go + (int32_t)(j - 1) * ge);
        row[j * st] = DpCol{h0, NEG, (int32_t)r[j - 1], 0};
        dir_at(0, j) = (uint8_t)(1 | (j == 1 ? 4 : 0));
        if (local_end && h0 > best) { best = h0; bi = 0; bj = j; }
    }
    // row i overwrites row i - 1 in place: `diag` keeps the previous row's H of
    // column j - 1, `left` the current row's
    for (uint32_t i = 1; i <= m; ++i) {
        uint8_t* dr = &dir_at(i, 0);
        int32_t Fi0 = -(go + (int32_t)(i - 1) * ge);
        int32_t diag = row[0].h;
        row[0].h = Fi0;
        row[0].f = Fi0;
        dr[0] = (uint8_t)(2 | (i == 1 ? 8 : 0));
        if (local_end && Fi0 > best) { best = Fi0; bi = i; bj = 0; }
        int32_t E = NEG, left = Fi0;
        const uint8_t qc = q[i - 1];
        const int32_t bonus = (local_end && i == m) ? sc.end_bonus : 0;
        for (uint32_t j = 1; j <= n; ++j) {
            const DpCol prev = row[j * st];
            int32_t eo = left - go, ee = E - ge;
            E = eo > ee ? eo : ee;
            int32_t fo = prev.h - go, fe = prev.f - ge;
            int32_t f = fo > fe ? fo : fe;
            int32_t h = diag + ((int32_t)qc == prev.r ? sc.match : -sc.mismatch) + bonus;
            int32_t hh = h > E ? h : E;
            hh = hh > f ? hh : f;
            row[j * st].h = hh;
            row[j * st].f = f;
            dr[j * st] = (uint8_t)((hh == E ? 1 : 0) | (hh == f ? 2 : 0) | (E == eo ? 4 : 0) | (f == fo ? 8 : 0));
            if (local_end && hh > best) { best = hh; bi = i; bj = j; }
            diag = prev.h;
            left = hh;
        }
    }
    if (!local_end) { best = row[(uint64_t)n * st].h; bi = m; bj = n; }
    o.score = best;
    o.bi = bi;
    o.bj = bj;
    // traceback
    uint32_t i = bi, j = bj, nt = 0;
    int state = 0;
    while (i > 0 || j > 0) {
        uint8_t d = dir_at(i, j);
        if (state == 0) {
            if (j > 0 && (d & 1)) state = 1;
            else if (i > 0 && (d & 2)) state = 2;
            else {
                if (nt >= tb_cap) { o.status = 2; return o; }
                tb[nt++] = q[i - 1] == r[j - 1] ? 7 : 8;
                --i; --j;
            }
        } else if (state == 2) {
            if (nt >= tb_cap) { o.status = 2; return o; }
            tb[nt++] = 1;
            bool open = (d & 8) != 0;
            --i;
            if (open) state = 0;
        } else {
            if (nt >= tb_cap) { o.status = 2; return o; }
            tb[nt++] = 2;
            bool open = (d & 4) != 0;
            --j;
            if (open) state = 0;
        }
    }
    o.ntb = nt;
    return o;
}

// Reversed views used by the LocalStartGlobalEnd mode.
template <typename V>
struct RevView {
    const V* v;
    uint32_t len;
    SA_HD uint8_t operator[](uint32_t i) const { return (*v)[len - 1 - i]; }
};

// Offset view (sub-range of another view).
template <typename V>
struct SubView {
    const V* v;
    uint32_t off;
    SA_HD uint8_t operator[](uint32_t i) const { return (*v)[off + i]; }
};

// src/piecewisealigner.rs:645-719 on (ref_start, query_start) pairs, last anchor first.
SA_HD uint32_t remove_spurious_anchors(uint32_t* a, uint32_t n) {
    if (n < 2) return n;
    auto erase = [&](uint32_t from, uint32_t to) {  // [from, to)
        uint32_t d = to - from;
        for (uint32_t x = to; x < n; ++x) { a[2 * (x - d)] = a[2 * x]; a[2 * (x - d) + 1] = a[2 * x + 1]; }
        n -= d;
    };
    const int64_t tol = 5;
    int64_t tracked = 0, dev_start = -1;
    uint32_t i = 1;
    while (i < n) {
        int64_t qd = (int64_t)a[2 * (i - 1) + 1] - (int64_t)a[2 * i + 1];
        int64_t rd = (int64_t)a[2 * (i - 1)] - (int64_t)a[2 * i];
        int64_t indel = qd - rd;
        int64_t ai = indel < 0 ? -indel : indel;
        if (ai > tol) {
            int64_t s = tracked + indel;
            if (dev_start >= 0 && (s < 0 ? -s : s) <= tol) {
                erase((uint32_t)dev_start, i);
                i = (uint32_t)dev_start;
                dev_start = -1;
                tracked = 0;
                continue;
            }
            dev_start = i;
            tracked = indel;
        }
        ++i;
    }
    // max_prune_count = ceil(len * 0.20) in f32
    float fl = (float)n * 0.20f;
    uint32_t maxp = (uint32_t)fl;
    if ((float)maxp < fl) ++maxp;
    for (uint32_t t = 1; t < (n < maxp ? n : maxp); ++t) {
        int64_t qd = (int64_t)a[2 * t + 1] - (int64_t)a[2 * (t - 1) + 1];
        int64_t rd = (int64_t)a[2 * t] - (int64_t)a[2 * (t - 1)];
        if (qd - rd != 0) { erase(0, t); break; }
    }
    uint32_t lo = (n > maxp ? n - maxp : 0) + 1;
    for (uint32_t t = n; t-- > lo;) {
        int64_t qd = (int64_t)a[2 * t + 1] - (int64_t)a[2 * (t - 1) + 1];
        int64_t rd = (int64_t)a[2 * t] - (int64_t)a[2 * (t - 1)];
        if (qd - rd != 0) { erase(t, n); break; }
    }
    return n;
}

// Emit DP traceback ops in forward order into `cig`; `reversed_space` means the DP
// ran on reversed sequences (its traceback order already is forward order).
SA_HD void emit_tb(const uint8_t* tb, uint32_t ntb, bool reversed_space, CigarOut& cig) {
    if (reversed_space) for (uint32_t x = 0; x < ntb; ++x) cig.push(tb[x], 1);
    else for (uint32_t x = ntb; x-- > 0;) cig.push(tb[x], 1);
}

// The DP of extend_piecewise in one thread.
struct DpScalar {
    template <typename Q, typename R>
    SA_HD DpOut operator()(const Q& q, uint32_t m, const R& r, uint32_t n, const Scores& sc, bool local_end,
                           uint8_t* scratch, uint64_t scratch_bytes, uint8_t* tb, uint32_t tb_cap, DpRows rows) const {
        return affine_dp(q, m, r, n, sc, local_end, scratch, scratch_bytes, tb, tb_cap, rows);
    }
};

// src/piecewisealigner.rs:396-510. `anch` holds (ref_start, query_start) pairs,
// last anchor first, with absolute reference coordinates; `win` is the decoded
// reference window (its start is the decode start). `dp` runs the affine DPs.
template <typename Dp = DpScalar>
SA_HD AlignInfo extend_piecewise(const QueryView& q, const RefWin& win, uint32_t* anch, uint32_t na, uint32_t padding,
                                 const Scores& sc, uint32_t k, uint8_t* scratch, uint64_t scratch_bytes,
                                 CigarOut& cig, DpRows rows = DpRows{}, const Dp& dp = Dp{}) {
    AlignInfo res{false, 0, 0, 0, 0, 0, 0, 0};
    cig.reset();
    if (na == 0) return res;
    const uint32_t ql = q.len;
    const uint32_t rl = win.len;
    // anchors relative to the window
    auto A_rs = [&](uint32_t x) { return anch[2 * x] - (uint32_t)win.start; };
    auto A_qs = [&](uint32_t x) { return anch[2 * x + 1]; };
    uint8_t* tb = scratch;
    uint32_t tb_cap = 2 * (ql + rl) + 16;
    uint64_t tb_bytes = ((uint64_t)tb_cap + 15) & ~15ull;
    if (tb_bytes >= scratch_bytes) { res.status = 2; return res; }
    uint8_t* dps = scratch + tb_bytes;
    uint64_t dps_bytes = scratch_bytes - tb_bytes;
    int32_t score = 0;
    uint32_t ref_start, query_start;
    {
        uint32_t fq = A_qs(na - 1), fr = A_rs(na - 1);
        if (fq > 0 && fr > 0) {
            uint32_t qlen = fq;
            uint32_t rs = fr > qlen + padding ? fr - (qlen + padding) : 0;
            uint32_t rlen = fr - rs;
            SubView<QueryView> qv{&q, 0};
            SubView<RefWin> rv{&win, rs};
            RevView<SubView<QueryView>> qr{&qv, qlen};
            RevView<SubView<RefWin>> rr{&rv, rlen};
            DpOut d = dp(qr, qlen, rr, rlen, sc, true, dps, dps_bytes, tb, tb_cap, rows);
            if (d.status) { res.status = d.status; return res; }
            if (d.score == 0) { query_start = fq; ref_start = fr; }
            else {
                emit_tb(tb, d.ntb, true, cig);
                score = d.score;
                query_start = qlen - d.bi;
                ref_start = rlen - d.bj + rs;
            }
        } else {
            query_start = fq;
            ref_start = fr;
            score = fq == 0 ? sc.end_bonus : 0;
        }
    }
    score += (int32_t)k * sc.match;
    cig.push(7, k);
    for (uint32_t i = na - 1; i >= 1; --i) {
        uint32_t cqs = A_qs(i - 1), crs = A_rs(i - 1);
        uint32_t pqe = A_qs(i) + k, pre = A_rs(i) + k;
        int64_t qd = (int64_t)cqs - (int64_t)pqe, rd = (int64_t)crs - (int64_t)pre;
        if (rd > 0 && qd > 0) {
            bool done = false;
            if (rd == qd) {
                int32_t hs = 0;
                for (uint32_t t = 0; t < (uint32_t)qd; ++t) hs += q[pqe + t] == win[pre + t] ? sc.match : -sc.mismatch;
                uint32_t hsc = hs > 0 ? (uint32_t)hs : 0;
                float fl = (float)qd * 0.85f;
                uint32_t c = (uint32_t)fl;
                if ((float)c < fl) ++c;
                if (hsc > 0 && hsc > (uint32_t)sc.match * c) {
                    score += (int32_t)hsc;
                    for (uint32_t t = 0; t < (uint32_t)qd; ++t) cig.push(q[pqe + t] == win[pre + t] ? 7u : 8u, 1);
                    score += (int32_t)k * sc.match;
                    cig.push(7, k);
                    done = true;
                }
            }
            if (!done) {
                SubView<QueryView> qv{&q, pqe};
                SubView<RefWin> rv{&win, pre};
                DpOut d = dp(qv, (uint32_t)qd, rv, (uint32_t)rd, sc, false, dps, dps_bytes, tb, tb_cap, rows);
                if (d.status) { res.status = d.status; return res; }
                score += d.score;
                emit_tb(tb, d.ntb, false, cig);
                score += (int32_t)k * sc.match;
                cig.push(7, k);
            }
        } else if (rd < qd) {
            uint32_t ins = (uint32_t)(qd - rd);
            score += -sc.gap_open + ((int32_t)ins - 1) * -sc.gap_extend;
            cig.push(1, ins);
            uint32_t mp = (uint32_t)((int64_t)k + rd);
            score += (int32_t)mp * sc.match;
            cig.push(7, mp);
        } else if (rd > qd) {
            uint32_t del = (uint32_t)(rd - qd);
            score += -sc.gap_open + ((int32_t)del - 1) * -sc.gap_extend;
            cig.push(2, del);
            uint32_t mp = (uint32_t)((int64_t)k + qd);
            score += (int32_t)mp * sc.match;
            cig.push(7, mp);
        } else {
            uint32_t mp = (uint32_t)((int64_t)k + qd);
            score += (int32_t)mp * sc.match;
            cig.push(7, mp);
        }
    }
    uint32_t ref_end, query_end;
    {
        uint32_t lqe = A_qs(0) + k, lre = A_rs(0) + k;
        if (lqe < ql && lre < rl) {
            uint32_t qlen = ql - lqe;
            uint32_t re = rl < lre + qlen + padding ? rl : lre + qlen + padding;
            SubView<QueryView> qv{&q, lqe};
            SubView<RefWin> rv{&win, lre};
            DpOut d = dp(qv, qlen, rv, re - lre, sc, true, dps, dps_bytes, tb, tb_cap, rows);
            if (d.status) { res.status = d.status; return res; }
            if (d.score == 0) { query_end = lqe; ref_end = lre; }
            else {
                emit_tb(tb, d.ntb, false, cig);
                score += d.score;
                query_end = d.bi + lqe;
                ref_end = d.bj + lre;
            }
        } else {
            query_end = lqe;
            ref_end = lre;
            score += lqe == ql ? sc.end_bonus : 0;
        }
    }
    if (cig.overflow) { res.status = 1; return res; }
    if (score <= 0) return res;
    res.found = true;
    res.score = (uint32_t)score;
    res.query_start = query_start;
    res.query_end = query_end;
    res.ref_start = ref_start;
    res.ref_end = ref_end;
    res.edit_distance = cig.edit_distance();
    return res;
}

}  // namespace sa
