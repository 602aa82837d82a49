// Piecewise extension of one job per lane group (device only).
//
// extend_piecewise (src/piecewisealigner.rs:396-510) runs its affine-gap DPs
// (the ends and the gaps between anchors, affine_dp) one after another; the end
// DPs span the query's unanchored tail against a padded window, up to hundreds
// of rows and columns. Here G lanes fill a DP together along anti-diagonals, in
// stripes of G rows: lane r holds row i0 + r + 1 of the stripe and, at step s,
// computes its column s - r + 1. The cell above and the one above-left come
// from lane r - 1 by a DPP row shift (its last two cells of the previous step),
// the cell to the left is the lane's own previous cell, and the stripe's first
// lane reads the row above from the row buffer, which the stripe's last row
// overwrites (a column is read there before it is written). The cells, their
// direction bits and the best end (the largest score, the first in row-major
// order among equal ones, as the row-major scalar loop keeps it) are those of
// affine_dp; the traceback and everything outside the DPs run on every lane of
// the group alike, so the group leaves each DP with the same state in each lane.
#pragma once
#include <hip/hip_runtime.h>

#include "sa_chain_group.hpp"

namespace sa {

// affine_dp's contract on a group of G lanes (G <= 16: the group is within one
// DPP row). rows: the row buffer and direction bytes with stride 1 (pitch
// rows.cap). Small DPs run the scalar form on every lane.
template <int G, typename Q, typename R>
__device__ DpOut affine_dp_grp(const Grp<G>& gp, const Q& q, uint32_t m, const R& r, uint32_t n, const Scores& sc,
                               bool local_end, uint8_t* scratch, uint64_t scratch_bytes, uint8_t* tb, uint32_t tb_cap,
                               const DpRows& rows) {
    static_assert(G == 8 || G == 16, "lane groups within one DPP row");
    if (m < 4 || n == 0 || m >= 65535 || n >= 65535)
        return affine_dp(q, m, r, n, sc, local_end, scratch, scratch_bytes, tb, tb_cap, rows);
    DpOut o{0, 0, 0, 0, 0};
    const int32_t go = sc.gap_open, ge = sc.gap_extend;
    const int32_t NEG = -(1 << 28);
    if ((uint64_t)n + 1 > rows.cap || m + 1 > rows.dir_rows) {
        o.status = 2;
        return o;
    }
    DpCol* row = rows.p;
    uint8_t* dir = rows.dir;
    const uint64_t pitch = rows.cap;
    // row 0 (its reference bases too) and column 0, lanes over j and over i
    for (uint32_t j = gp.r; j <= n; j += G) {
        const int32_t h0 = j == 0 ? 0 : -(go + (int32_t)(j - 1) * ge);
        row[j] = DpCol{h0, NEG, j == 0 ? 0 : (int32_t)r[j - 1], 0};
        dir[j] = j == 0 ? (uint8_t)0 : (uint8_t)(1 | (j == 1 ? 4 : 0));
    }
    for (uint32_t i = gp.r + 1; i <= m; i += G) dir[(uint64_t)i * pitch] = (uint8_t)(2 | (i == 1 ? 8 : 0));
    lds_fence();
    // best end as a key: score (biased), then the smaller row, then the smaller
    // column; the start (0, 0) with 0 wins ties at 0, as nothing above 0 replaces it
    auto key = [](int32_t s, uint32_t i, uint32_t j) {
        return ((uint64_t)(uint32_t)(s + 0x40000000) << 32) | ((uint64_t)(0xFFFFu - i) << 16) | (0xFFFFu - j);
    };
    uint64_t best = key(0, 0, 0);
    for (uint32_t i0 = 0; i0 < m; i0 += G) {
        const uint32_t i = i0 + gp.r + 1;
        const bool row_ok = i <= m;
        const uint32_t last_r = (m - i0 < (uint32_t)G ? m - i0 : (uint32_t)G) - 1;
        const int32_t Fi0 = -(go + (int32_t)(i - 1) * ge);  // H(i, 0)
        const int32_t qc = row_ok ? (int32_t)q[i - 1] : -1;
        const int32_t bonus = (local_end && i == m) ? sc.end_bonus : 0;
        // the lane's last cell H(i, j), the one before it, F(i, j), and the
        // reference base of its column (what the lane below takes next step)
        int32_t hl = Fi0, hp = Fi0, fl = NEG, rl = 0, E = NEG;
        int32_t up0 = i0 == 0 ? 0 : -(go + (int32_t)(i0 - 1) * ge);  // lane 0: H(i0, j - 1)
        const uint32_t steps = n + last_r;
        for (uint32_t s = 0; s < steps; ++s) {
            // the lane above's state after the previous step
            int32_t up_h = (int32_t)dpp_u<0x111>(0, (uint32_t)hl);
            int32_t up_f = (int32_t)dpp_u<0x111>(0, (uint32_t)fl);
            int32_t diag = (int32_t)dpp_u<0x111>(0, (uint32_t)hp);
            int32_t rb = (int32_t)dpp_u<0x111>(0, (uint32_t)rl);
            const int32_t j = (int32_t)s - (int32_t)gp.r + 1;
            if (gp.r == 0 && j >= 1 && j <= (int32_t)n) {
                const DpCol u = row[j];
                up_h = u.h;
                up_f = u.f;
                rb = u.r;
                diag = up0;
                up0 = u.h;
            }
            if (row_ok && j >= 1 && j <= (int32_t)n) {
                const int32_t eo = hl - go, ee = E - ge;
                E = eo > ee ? eo : ee;
                const int32_t fo = up_h - go, fe = up_f - ge;
                const int32_t f = fo > fe ? fo : fe;
                const int32_t h = diag + (qc == rb ? sc.match : -sc.mismatch) + bonus;
                int32_t hh = h > E ? h : E;
                hh = hh > f ? hh : f;
                dir[(uint64_t)i * pitch + (uint32_t)j] =
                    (uint8_t)((hh == E ? 1 : 0) | (hh == f ? 2 : 0) | (E == eo ? 4 : 0) | (f == fo ? 8 : 0));
                if (local_end) {
                    const uint64_t kk = key(hh, i, (uint32_t)j);
                    best = kk > best ? kk : best;
                }
                if (gp.r == last_r) {
                    row[j].h = hh;
                    row[j].f = f;
                }
                hp = hl;
                hl = hh;
                fl = f;
                rl = rb;
            }
        }
        lds_fence();  // the stripe's last row is in the row buffer
    }
    if (local_end) {
        for (int d = G / 2; d > 0; d >>= 1) {
            const uint64_t x = __shfl_xor(best, d, G);
            best = x > best ? x : best;
        }
        o.score = (int32_t)(uint32_t)(best >> 32) - 0x40000000;
        o.bi = 0xFFFFu - (uint32_t)((best >> 16) & 0xFFFFu);
        o.bj = 0xFFFFu - (uint32_t)(best & 0xFFFFu);
    } else {
        o.score = row[n].h;
        o.bi = m;
        o.bj = n;
    }
    // traceback, the same on every lane
    uint32_t ti = o.bi, tj = o.bj, nt = 0;
    int state = 0;
    while (ti > 0 || tj > 0) {
        const uint8_t d = dir[(uint64_t)ti * pitch + tj];
        if (state == 0) {
            if (tj > 0 && (d & 1)) state = 1;
            else if (ti > 0 && (d & 2)) state = 2;
            else {
                if (nt >= tb_cap) { o.status = 2; return o; }
                tb[nt++] = q[ti - 1] == r[tj - 1] ? 7 : 8;
                --ti; --tj;
            }
        } else if (state == 2) {
            if (nt >= tb_cap) { o.status = 2; return o; }
            tb[nt++] = 1;
            const bool open = (d & 8) != 0;
            --ti;
            if (open) state = 0;
        } else {
            if (nt >= tb_cap) { o.status = 2; return o; }
            tb[nt++] = 2;
            const bool open = (d & 4) != 0;
            --tj;
            if (open) state = 0;
        }
    }
    o.ntb = nt;
    return o;
}

// The DP of extend_piecewise on a lane group.
template <int G>
struct DpGroup {
    Grp<G> gp;
    template <typename Q, typename R>
    __device__ DpOut operator()(const Q& q, uint32_t m, const R& r, uint32_t n, const Scores& sc, bool local_end,
                                uint8_t* scratch, uint64_t scratch_bytes, uint8_t* tb, uint32_t tb_cap,
                                DpRows rows) const {
        return affine_dp_grp<G>(gp, q, m, r, n, sc, local_end, scratch, scratch_bytes, tb, tb_cap, rows);
    }
};

}  // namespace sa
