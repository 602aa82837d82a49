// Lane-group Smith-Waterman score passes for CDNA (device only).
//
// ext/ssw/ssw.c:197-588 runs the striped (Farrar) recurrence on one SSE register:
// 16 unsigned-byte lanes, or 8 signed-word lanes once the byte scores saturate.
// Its lazy-F loop does not feed corrected H values back into E, so the scores
// it reports depend on the striping. Here one SSE lane is one GPU lane: a group
// of G = 16 (byte) or 8 (word) consecutive lanes of a wave runs one alignment,
// lane l keeping the l-th lane of every stripe segment in registers. The
// SSE byte shift that carries H and F into the next stripe is a DPP row shift
// (row_shr:1, groups never straddle a 16-lane DPP row), the per-column maximum
// is a DPP butterfly inside the group and the lazy-F exit test is a ballot, so
// a wave runs 64 / G alignments with no LDS and no scratch memory. The stripe
// count SEG is a compile-time constant of each kernel instance (jobs are
// bucketed by it), so H, E and the profile live in registers.
#pragma once
#include <hip/hip_runtime.h>

#include "sa_extend.hpp"

namespace sa {

// lane l <- lane l-1 inside each group of G lanes, lane 0 of the group <- 0
template <int G>
__device__ __forceinline__ int32_t grp_shift_up(int32_t v, uint32_t l) {
    int32_t s = __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, true);  // row_shr:1
    if (G == 8) s = l == 0 ? 0 : s;
    return s;
}

template <int G>
__device__ __forceinline__ int32_t grp_max(int32_t v) {
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false));   // quad_perm [1,0,3,2]
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false));   // quad_perm [2,3,0,1]
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, false));  // row_half_mirror
    if (G == 16) v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x140, 0xF, 0xF, false));  // row_mirror
    return v;
}

template <int G>
__device__ __forceinline__ int32_t grp_min(int32_t v) {
    v = min(v, __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false));
    v = min(v, __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false));
    v = min(v, __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, false));
    if (G == 16) v = min(v, __builtin_amdgcn_update_dpp(0, v, 0x140, 0xF, 0xF, false));
    return v;
}

// Reference window reader: one 64-bit load per 32 columns (2-bit codes, the
// PackedSeq layout), only words inside the window are touched.
struct RefCursor {
    const uint64_t* w;  // words of the contig
    uint64_t start;     // window start within the contig
    int64_t ci;         // index of the cached word
    uint64_t cur;
    __device__ __forceinline__ int32_t at(int32_t i) {
        const uint64_t a = start + (uint64_t)i;
        const int64_t wi = (int64_t)(a >> 5);
        if (wi != ci) {
            ci = wi;
            cur = w[wi];
        }
        return (int32_t)((cur >> ((a & 31) * 2)) & 3);
    }
};

// Column loop of one striped pass (the scalar form is ssw_pass in
// sa_extend.hpp) with SEG stripe segments per lane, known at compile time so
// that H and E stay in registers with static indices. A group whose lazy-F loop
// has ended holds vF = 0, which makes further lazy-F steps no-ops for it, so the
// wave leaves that loop when all its groups have.
//
// The query profile of the lane (ext/ssw/ssw.c:140-160, qP_byte/qP_word) is held
// as signed bytes: byte j of prof[c][j / 4] is the score of segment j, query
// position j + l*SEG, against reference base c (0 past the query end). The
// best-column copy of H (pvHmax) is replaced by the smallest query position that
// reaches the new maximum, found when the maximum is set: the value the reference
// derives from that copy once the pass ends.
template <int G, bool BYTE, int SEG>
__device__ SswEnd ssw_columns(const uint32_t (&prof)[4][(SEG + 3) / 4], int32_t qlen, RefCursor rc_cur, int32_t rlen,
                              bool reverse, const Scores& sc, int32_t terminate, uint32_t l, uint32_t gbase,
                              int32_t& cols) {
    constexpr int W = (SEG + 3) / 4;
    const int32_t bias = BYTE ? sc.mismatch : 0;
    const int32_t gapO = sc.gap_open, gapE = sc.gap_extend;
    const int32_t hcap = 255 - bias;  // byte lanes saturate at 255 with the bias added
    int32_t H[SEG], E[SEG];
#pragma unroll
    for (int j = 0; j < SEG; ++j) H[j] = E[j] = 0;
    int32_t maxv = 0, end_ref = BYTE ? -1 : 0;
    int32_t end_read = 0;  // smallest query position holding maxv in the best column
    const int32_t begin = reverse ? rlen - 1 : 0, end = reverse ? -1 : rlen, step = reverse ? -1 : 1;
    cols = 0;  // columns computed (the pass may stop before the window's end)
    for (int32_t i = begin; i != end; i += step) {
        ++cols;
        int32_t vH = grp_shift_up<G>(H[SEG - 1], l);
        int32_t vF = 0, maxcol = 0;
        const int32_t rc = rc_cur.at(i);
        uint32_t p[W];
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const uint32_t x0 = (rc & 1) ? prof[1][w] : prof[0][w];
            const uint32_t x1 = (rc & 1) ? prof[3][w] : prof[2][w];
            p[w] = (rc & 2) ? x1 : x0;
        }
#pragma unroll
        for (int j = 0; j < SEG; ++j) {
            int32_t h = vH + (int32_t)(int8_t)(p[j >> 2] >> (8 * (j & 3)));
            if (BYTE) h = min(max(h, 0), hcap);
            const int32_t e = E[j];
            h = max(max(h, e), vF);
            maxcol = max(maxcol, h);
            vH = H[j];
            H[j] = h;
            const int32_t hg = h - gapO;
            E[j] = max(max(e - gapE, hg), 0);
            vF = max(max(vF - gapE, hg), 0);
        }
        // lazy-F loop (E is not updated)
        bool all_done = false;
        for (int kk = 0; kk < G && !all_done; ++kk) {
            vF = grp_shift_up<G>(vF, l);
#pragma unroll
            for (int j = 0; j < SEG; ++j) {
                const int32_t h = max(H[j], vF);
                maxcol = max(maxcol, h);
                H[j] = h;
                const int32_t hg = max(h - gapO, 0);
                vF = max(vF - gapE, 0);
                const uint64_t live = __ballot(vF > hg);
                vF = ((live >> gbase) & ((1ull << G) - 1)) == 0 ? 0 : vF;
                if (live == 0) {
                    all_done = true;
                    break;
                }
            }
        }
        const int32_t colmax = grp_max<G>(maxcol);
        if (colmax > maxv) {
            maxv = colmax;
            if (BYTE && maxv + bias >= 255) break;
            end_ref = i;
            int32_t cand = 0x7fffffff;
            int32_t lseg = (int32_t)l * SEG;
            asm volatile("" : "+v"(lseg));  // one add per segment, not one register each
#pragma unroll
            for (int j = SEG - 1; j >= 0; --j) cand = H[j] == maxv ? j + lseg : cand;
            end_read = grp_min<G>(cand);
        }
        if (colmax == terminate) break;
    }
    SswEnd o;
    o.score = BYTE ? (maxv + bias >= 255 ? 255 : maxv) : maxv;
    o.ref = end_ref;
    o.read = end_read < qlen - 1 ? end_read : qlen - 1;
    return o;
}

// Builds the lane's profile for SEG segments and runs the column loop.
template <int G, bool BYTE, int SEG>
__device__ __forceinline__ SswEnd ssw_pass_seg(const QueryView& q, int32_t ql, bool rev, int32_t qend, RefCursor rc_cur, int32_t rl,
                               const Scores& sc, int32_t term, uint32_t l, uint32_t gbase, int32_t& cols) {
    constexpr int W = (SEG + 3) / 4;
    uint32_t prof[4][W];
#pragma unroll
    for (int w = 0; w < W; ++w) prof[0][w] = prof[1][w] = prof[2][w] = prof[3][w] = 0;
    const uint32_t mt = (uint32_t)(sc.match & 0xFF), mm = (uint32_t)((-sc.mismatch) & 0xFF);
    // the lane's query bytes, all loads issued before the first is used
    // (positions past the query end read a byte inside it and are not used);
    // query position of the pass -> read position, read as given or reverse
    // complemented
    uint8_t qb[SEG];
#pragma unroll
    for (int j = 0; j < SEG; ++j) {
        const int32_t qi = j + (int32_t)l * SEG;
        const int32_t qc = qi < ql ? qi : (ql > 0 ? ql - 1 : 0);
        const uint32_t t = (uint32_t)(rev ? qend - qc : qc);
        qb[j] = q.seq[q.rc ? q.len - 1 - t : t];
    }
#pragma unroll
    for (int j = 0; j < SEG; ++j) {
        const int32_t qi = j + (int32_t)l * SEG;
        if (qi < ql) {
            const int32_t c = q.rc ? ssw_code_comp(qb[j]) : ssw_code(qb[j]);
#pragma unroll
            for (int b = 0; b < 4; ++b) prof[b][j >> 2] |= (c == b ? mt : mm) << (8 * (j & 3));
        }
    }
    return ssw_columns<G, BYTE, SEG>(prof, ql, rc_cur, rl, rev, sc, term, l, gbase, cols);
}

}  // namespace sa
