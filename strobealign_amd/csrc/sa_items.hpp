// Work items of the mapping kernels: one call handles one read, one read
// orientation or one extension request. The __global__ kernels in
// sa_kernels.hip map threads onto these items.
#pragma once
#include "sa_extend.hpp"

namespace sa {

// --------------------------------------------------------------- seeding ---
// With skip_len > 0, reads of at most skip_len bases are left to the wave
// form (k_seed_wave); with 0, every read is seeded here, empty ones included
// (their counts must be written: the next stage reads them).
SA_HD void item_seed(uint32_t i, const uint8_t* seq, const uint64_t* off,
                                             uint32_t n, SeedParams p, Syncmer* sy,
                                             QRand* qr, uint32_t* qr_n, uint32_t skip_len) {
    if (i >= n) return;
    uint64_t o = off[i];
    uint32_t L = (uint32_t)(off[i + 1] - o);
    if (skip_len > 0 && L <= skip_len) return;
    if (p.w_max < 24)
        read_randstrobes_stream<24>(seq + o, L, p, qr + 2 * o, qr + 2 * o + L, &qr_n[2 * i], &qr_n[2 * i + 1]);
    else
        read_randstrobes(seq + o, L, p, sy + o, qr + 2 * o, qr + 2 * o + L, &qr_n[2 * i], &qr_n[2 * i + 1]);
}

// ------------------------------------------------------------------ hits ---
SA_HD void item_hits(uint32_t t, IndexView ix, MapParams mp, const uint64_t* off,
                                             uint32_t n, const QRand* qr,
                                             const uint32_t* qr_n, Hit* hits,
                                             uint64_t* cnt, HitCounts* hc,
                                             uint64_t* acap, uint32_t* lookup_bytes) {
    if (t >= 2 * n) return;
    uint32_t i = t >> 1, o = t & 1;
    uint64_t ro = off[i];
    uint32_t L = (uint32_t)(off[i + 1] - ro);
    uint64_t base = 2 * ro + (uint64_t)o * L;
    HitCounts d;
    uint32_t nh = find_hits(ix, qr + base, qr_n[t], mp, hits + base, cnt + base, d);
    hc[t] = d;
    acap[t] = hit_runs(ix, hits + base, nh);
    // query randstrobes in, bucket_starts words and index entries read, hits out
    lookup_bytes[t] = (uint32_t)sizeof(QRand) * qr_n[t] + 8u * ix.bucket_reads + (uint32_t)sizeof(RefEntry) * ix.entry_reads +
                      (uint32_t)sizeof(Hit) * nh;
}

// -------------------------------------------------------------- chaining ---
struct ChainBufs {
    Anchor* anchors;
    Anchor* anchors_tmp;  // second buffer of the anchor sort
    float* dp;
    uint32_t* pred;
    uint32_t* cand;
    uint8_t* used;
    NamRec* nams;
    uint32_t* chain_anchors;
    uint32_t* alg_bytes;  // per read orientation: hits + index entries read, NAMs + chain anchors written
};

SA_HD void item_chain(uint32_t t, IndexView ix, ChainParams cp, const uint64_t* off,
                                              uint32_t n, const Hit* hits,
                                              const HitCounts* hc, const uint64_t* acap,
                                              const uint64_t* aoff, ChainBufs b,
                                              uint32_t* nam_n, uint32_t* anchors_raw,
                                              const float* log2tab,
                                              const float* same_diag) {
    if (t >= 2 * n) return;
    uint32_t i = t >> 1, o = t & 1;
    const HitCounts& h0 = hc[2 * i];
    const HitCounts& h1 = hc[2 * i + 1];
    bool sel;
    if (better_orientation(h0, h1)) sel = (o == 0);
    else if (better_orientation(h1, h0)) sel = (o == 1);
    else sel = true;
    nam_n[t] = 0;
    anchors_raw[t] = 0;
    b.alg_bytes[t] = 0;
    if (!sel) return;
    uint64_t ro = off[i];
    uint32_t L = (uint32_t)(off[i + 1] - ro);
    uint64_t base = 2 * ro + (uint64_t)o * L;
    uint64_t ab = aoff[t];
    ChainScratch s{b.anchors + ab, b.dp + ab, b.pred + ab, b.cand + ab, b.used + ab, (uint32_t)acap[t]};
    uint32_t entries = 0, n_chain = 0, ns = 0, ne = 0;
    Anchor* tmp = b.anchors_tmp + ab;
    hits_to_anchors(ix, hits + base, hc[t].n_hits, s.anchors, &ns, tmp, &ne, &entries);
    anchors_raw[t] = ns + ne;
    const uint32_t na = sort_dedup_anchors(s.anchors, ns, tmp, ne, &s.anchors);
    const uint32_t nn = chain_orientation(s, na, (uint32_t)ix.p.k, L, o == 1, cp, log2tab, same_diag, b.nams + ab,
                                          b.chain_anchors + 2 * (2 * ab), 2 * ab, &n_chain);
    nam_n[t] = nn;
    b.alg_bytes[t] = (uint32_t)sizeof(Hit) * hc[t].n_hits + (uint32_t)sizeof(RefEntry) * entries +
                     (uint32_t)sizeof(NamRec) * nn + 8u * n_chain;
}

// Merge both orientations, number NAMs like src/chainer.rs:435-436, stable-sort by
// score (src/nam.rs:170) and record the consistency checks of src/nam.rs:49-135.
SA_HD void item_finalize(uint32_t i, const uint8_t* seq, const uint64_t* off,
                                                 uint32_t n, RefView ref, uint32_t k,
                                                 const uint32_t* nam_n,
                                                 const uint64_t* aoff, const NamRec* nams_in,
                                                 const uint64_t* out_off, NamRec* nams_out) {
    if (i >= n) return;
    uint32_t n0 = nam_n[2 * i], n1 = nam_n[2 * i + 1];
    NamRec* out = nams_out + out_off[i];
    const NamRec* a0 = nams_in + aoff[2 * i];
    const NamRec* a1 = nams_in + aoff[2 * i + 1];
    for (uint32_t x = 0; x < n0; ++x) out[x] = a0[x];
    for (uint32_t x = 0; x < n1; ++x) {
        out[n0 + x] = a1[x];
        out[n0 + x].nam_id += n0;
    }
    uint32_t m = n0 + n1;
    // stable: order by (score desc, nam_id asc); nam_id equals the concatenation index
    heap_sort(out, m, [](const NamRec& a, const NamRec& b) {
        return a.score > b.score || (a.score == b.score && a.nam_id < b.nam_id);
    });
    uint64_t ro = off[i];
    uint32_t L = (uint32_t)(off[i + 1] - ro);
    for (uint32_t x = 0; x < m; ++x) {
        NamRec& r = out[x];
        bool cons = nam_consistent(ref, seq + ro, L, r.flags & 1, r.ref_id, r.ref_start, r.ref_end, r.query_start,
                                   r.query_end, k);
        uint32_t rev = cons ? 0u : nam_reverse_state(ref, seq + ro, L, r, k);
        r.flags = (r.flags & 1u) | (cons ? 2u : 0u) | (rev << 2);
    }
}

// ------------------------------------------------------------- extension ---
#if defined(__HIP_DEVICE_COMPILE__)
#define SA_ATOMIC_ADD(p, v) atomicAdd((p), (v))
#define SA_ATOMIC_MIN(p, v) atomicMin((p), (v))
#else
#define SA_ATOMIC_ADD(p, v) ((*(p) += (v)) - (v))
#define SA_ATOMIC_MIN(p, v) (*(p) = *(p) < (v) ? *(p) : (v))
#endif

// Longest query the lane-parallel SSW kernel handles; longer ones use the scalar form.
constexpr uint32_t SSW_GROUP_MAX_QUERY = 512;

SA_HD void write_found(ExtResult& out, const AlignInfo& a, uint64_t ref_base, bool gapped, uint32_t n_cigar) {
    out.found = 1;
    out.gapped = gapped ? 1 : 0;
    out.score = a.score;
    out.ref_start = (uint32_t)(ref_base + a.ref_start);
    out.length = a.ref_end - a.ref_start;
    out.query_start = a.query_start;
    out.query_end = a.query_end;
    out.edit_distance = a.edit_distance;
    out.n_cigar = n_cigar;
}

// The piecewise aligner's entry points on an explicit window (EXT_ALIGN_PW,
// EXT_PW_GLOBAL, EXT_PW_XDROP; see ExtKind). Coordinates are window-relative.
SA_HD AlignInfo pw_probe(const ExtRequest& r, const QueryView& q, const RefWin& w, const Scores& sc, uint32_t k,
                         const uint32_t* chain_anchors, uint8_t* scr, uint64_t scratch_per, CigarOut& cig,
                         DpRows rows = DpRows{}) {
    AlignInfo a{false, 0, 0, 0, 0, 0, 0, 0};
    cig.reset();
    if (r.kind == EXT_ALIGN_PW) {
        uint32_t* an = (uint32_t*)scr;
        const uint64_t an_bytes = ((uint64_t)r.n_anchors * 8 + 15) & ~15ull;
        if (an_bytes >= scratch_per) {
            a.status = 2;
            return a;
        }
        for (uint32_t x = 0; x < 2 * r.n_anchors; ++x) an[x] = chain_anchors[2 * r.anchor_off + x];
        return extend_piecewise(q, w, an, r.n_anchors, r.ref_start, sc, k, scr + an_bytes, scratch_per - an_bytes,
                                cig, rows);
    }
    const uint32_t ql = q.len, rl = w.len;
    const bool rev = r.kind == EXT_PW_XDROP && (r.flags & 1) != 0;
    const bool local_end = r.kind == EXT_PW_XDROP;
    if (local_end && (ql == 0 || rl == 0)) {
        a.found = true;
        return a;
    }
    const uint32_t tb_cap = ql + rl + 16;
    const uint64_t tb_bytes = ((uint64_t)tb_cap + 15) & ~15ull;
    if (tb_bytes >= scratch_per) {
        a.status = 2;
        return a;
    }
    uint8_t* tb = scr;
    DpOut d;
    if (rev) {
        RevView<QueryView> qr{&q, ql};
        RevView<RefWin> rr{&w, rl};
        d = affine_dp(qr, ql, rr, rl, sc, true, scr + tb_bytes, scratch_per - tb_bytes, tb, tb_cap, rows);
    } else {
        d = affine_dp(q, ql, w, rl, sc, local_end, scr + tb_bytes, scratch_per - tb_bytes, tb, tb_cap, rows);
    }
    if (d.status) {
        a.status = d.status;
        return a;
    }
    emit_tb(tb, d.ntb, rev, cig);
    a.found = true;
    a.score = (uint32_t)d.score;
    if (rev) {
        a.query_start = ql - d.bi;
        a.query_end = ql;
        a.ref_start = rl - d.bj;
        a.ref_end = rl;
    } else {
        a.query_start = 0;
        a.query_end = d.bi;
        a.ref_start = 0;
        a.ref_end = d.bj;
    }
    a.edit_distance = cig.edit_distance();
    if (cig.overflow) a.status = 1;
    return a;
}

// First stage of extend_seed (src/mapper.rs:532-620) and rescue_align
// (src/mapper.rs:1094-1153): the ungapped (Hamming) case and the rescue
// substring prefilter are finished here; gapped alignments become jobs for the
// SSW kernel (ssw_jobs, or long_jobs for queries above SSW_GROUP_MAX_QUERY) or
// the piecewise kernel (pw_jobs).
SA_HD void item_ext_prep(uint32_t t, const uint8_t* seq, const uint64_t* off, RefView ref, Scores sc, uint32_t k,
                         const ExtRequest* req, uint32_t nreq, ExtResult* res, uint32_t* cigars, ExtJob* ssw_jobs,
                         ExtJob* long_jobs, ExtJob* pw_jobs, uint32_t* counters) {
    if (t >= nreq) return;
    const ExtRequest r = req[t];
    ExtResult out{0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    const uint64_t ro = off[r.read_id];
    const uint32_t L = (uint32_t)(off[r.read_id + 1] - ro);
    QueryView q{seq + ro, L, (r.flags & 4) != 0};
    const uint64_t rlen = ref.len[r.ref_id];
    auto push_ssw = [&](uint64_t ws, uint64_t we, uint32_t jflags) {
        if (we - ws > 2000 || L == 0) return;  // Aligner::align returns None (src/aligner.rs:83-86)
        ExtJob j{t, r.ref_id, ws, (uint32_t)(we - ws), jflags};
        if (L > SSW_GROUP_MAX_QUERY || (uint64_t)sc.match * L > 30000 || sc.match > 127 || sc.mismatch > 127)
            long_jobs[SA_ATOMIC_ADD(&counters[1], 1u)] = j;
        else ssw_jobs[SA_ATOMIC_ADD(&counters[0], 1u)] = j;
    };
    if (r.kind == EXT_SEED) {
        const bool use_ssw = r.flags & 1, consistent = (r.flags & 2) != 0;
        uint64_t prs = r.ref_start > r.query_start ? r.ref_start - r.query_start : 0;
        uint64_t pre = (uint64_t)r.ref_end + L - r.query_end;
        if (pre > rlen) pre = rlen;
        if (prs + L == pre && consistent) {
            RefWin w{&ref, r.ref_id, prs, L};
            uint32_t hd = 0;
            for (uint32_t x = 0; x < L; ++x) hd += q[x] != w[x];
            if ((float)hd / (float)L < 0.05f) {
                CigarOut cig{cigars + r.cigar_off, r.cigar_cap, 0, false};
                AlignInfo a = hamming_align(q, w, L, sc, cig);
                out.status = (uint32_t)a.status;
                if (a.found && a.status == 0) write_found(out, a, prs, false, cig.n);
                res[t] = out;
                return;
            }
        }
        uint32_t padding = L / 10;
        if (use_ssw) {
            uint64_t ws = prs > padding ? prs - padding : 0;
            uint64_t we = pre + padding < rlen ? pre + padding : rlen;
            // the NAM's start and end diagonals, in window coordinates
            uint32_t jflags = 0;
            if (L <= SSW_GROUP_MAX_QUERY && (uint64_t)sc.match * L <= 30000 && sc.match <= 127 && sc.mismatch <= 127 &&
                sc.gap_open > sc.gap_extend) {
                const int64_t d1 = (int64_t)prs - (int64_t)ws;
                const int64_t d2 = (int64_t)r.ref_end - (int64_t)r.query_end - (int64_t)ws;
                const RefWin w{&ref, r.ref_id, ws, (uint32_t)(we - ws)};
                jflags = two_diagonal_score(q, w, L, d1, d2, sc) >= 255 - sc.mismatch ? SSW_WORD_FIRST : 0u;
            }
            push_ssw(ws, we, jflags);
        } else {
            uint64_t ds = r.ref_start > (uint64_t)L + padding ? r.ref_start - ((uint64_t)L + padding) : 0;
            uint64_t de = (uint64_t)r.ref_end + L + padding;
            if (de > rlen) de = rlen;
            pw_jobs[SA_ATOMIC_ADD(&counters[2], 1u)] = ExtJob{t, r.ref_id, ds, (uint32_t)(de - ds), 0};
        }
    } else if (r.kind == EXT_HAMMING) {
        const RefWin w{&ref, r.ref_id, r.win_start, r.win_end - r.win_start};
        if (w.len == L) {
            CigarOut cig{cigars + r.cigar_off, r.cigar_cap, 0, false};
            AlignInfo a = hamming_align(q, w, L, sc, cig);
            out.status = (uint32_t)a.status;
            if (a.found && a.status == 0) write_found(out, a, r.win_start, false, cig.n);
        }
    } else if (r.kind >= EXT_ALIGN_PW) {
        pw_jobs[SA_ATOMIC_ADD(&counters[2], 1u)] = ExtJob{t, r.ref_id, r.win_start, r.win_end - r.win_start, 0};
    } else {
        // has_shared_substring (src/mapper.rs:1156-1170) unless bit3 asks for a plain alignment
        uint32_t ws = r.win_start, we = r.win_end;
        RefWin w{&ref, r.ref_id, ws, we - ws};
        uint32_t sub = 2 * k / 3, step = k / 3;
        bool shared = (r.flags & 8) != 0, found = false;
        uint32_t hit_i = 0, hit_p = 0;
        uint32_t lim = L > sub ? L - sub : 0;
        for (uint32_t i0 = 0; i0 < lim && !shared; i0 += step) {
            for (uint32_t p = 0; p + sub <= w.len && !shared; ++p) {
                uint32_t x = 0;
                while (x < sub && w[p + x] == q[i0 + x]) ++x;
                if (x == sub) {
                    shared = found = true;
                    hit_i = i0;
                    hit_p = p;
                }
            }
        }
        // best ungapped segment on the diagonal of the first shared substring
        uint32_t jflags = 0;
        if (found) {
            const int64_t d = (int64_t)hit_p - (int64_t)hit_i;
            const int64_t x0 = d < 0 ? -d : 0;
            const int64_t x1 = (int64_t)L < (int64_t)(we - ws) - d ? (int64_t)L : (int64_t)(we - ws) - d;
            const uint32_t n = x1 > x0 ? (uint32_t)(x1 - x0) : 0;
            const RefWin wd{&ref, r.ref_id, (uint64_t)((int64_t)ws + d + x0), n};
            SegSum s;
            for (uint32_t j = 0; j < n; ++j)
                s.push(ssw_code(q[(uint32_t)x0 + j]) == (int32_t)wd.code(j) ? sc.match : -sc.mismatch);
            jflags = s.best >= 255 - sc.mismatch ? SSW_WORD_FIRST : 0u;
        }
        if (shared) push_ssw(ws, we, jflags);
    }
    res[t] = out;
}

// SSW job in one thread (scalar lanes): queries longer than SSW_GROUP_MAX_QUERY.
// Job `job` of the list with scratch slot `slot` (scratch_per bytes each).
SA_HD void item_ext_long(uint32_t job, uint32_t slot, const uint8_t* seq, const uint64_t* off, RefView ref, Scores sc,
                         const ExtRequest* req, const ExtJob* jobs, ExtResult* res, uint32_t* cigars,
                         uint8_t* scratch, uint64_t scratch_per) {
    const ExtJob j = jobs[job];
    const ExtRequest r = req[j.req];
    const uint64_t ro = off[r.read_id];
    const uint32_t L = (uint32_t)(off[r.read_id + 1] - ro);
    QueryView q{seq + ro, L, (r.flags & 4) != 0};
    CigarOut cig{cigars + r.cigar_off, r.cigar_cap, 0, false};
    ExtResult out{0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    RefWin w{&ref, r.ref_id, j.ws, j.wlen};
    AlignInfo a = ssw_aligner_align(q, w, sc, scratch + (uint64_t)slot * scratch_per, scratch_per, cig);
    out.status = (uint32_t)a.status;
    if (a.found && a.status == 0) write_found(out, a, j.ws, true, cig.n);
    res[j.req] = out;
}

// Traceback, CIGAR and end bonus of an SSW job whose score passes are done.
// Traceback order: SSW jobs grouped by the band width their banded traceback
// starts with (|aligned reference length - aligned query length| + 1: 1, 2,
// 3-4, 5-8, more; last the jobs that stop before it), so that the lanes of a
// wave run the same band-pass variant.
// Classes by the band width the traceback starts with, widest first (their
// jobs take longest, so their waves start first), jobs without a traceback last.
constexpr uint32_t FIN_CLASSES = 8;
SA_HD uint32_t fin_class(const SswEnds& e) {
    if (e.ref_begin1 < 0 || e.read_begin1 < 0 || e.rev_short) return FIN_CLASSES - 1;
    const int32_t refL = e.ref_end1 - e.ref_begin1 + 1, readL = e.read_end1 - e.read_begin1 + 1;
    const int32_t bw = (refL > readL ? refL - readL : readL - refL) + 1;
    return bw > 32 ? 0u : bw > 16 ? 1u : bw > 8 ? 2u : bw > 4 ? 3u : bw > 2 ? 4u : bw > 1 ? 5u : 6u;
}
// lists: FIN_CLASSES lists of `cap` entries; counts: their lengths (zeroed before)
SA_HD void item_fin_bucket(uint32_t t, const SswEnds* ends, uint32_t* lists, uint32_t* counts, uint32_t cap) {
    const uint32_t c = fin_class(ends[t]);
    lists[(uint64_t)c * cap + SA_ATOMIC_ADD(&counts[c], 1u)] = t;
}

// The job at position pos of the class lists taken in order, with scratch slot
// `slot` (the slots of a wave's 64 threads interleave their scratch words, so
// the scratch covers whole waves of scratch_per bytes per slot).
SA_HD void item_ssw_finish(uint32_t pos, uint32_t slot, const uint8_t* seq, const uint64_t* off, RefView ref,
                           Scores sc, const ExtRequest* req, const ExtJob* jobs, const SswEnds* ends, ExtResult* res,
                           uint32_t* cigars, uint8_t* scratch, uint64_t scratch_per, const uint32_t* lists,
                           const uint32_t* counts, uint32_t cap) {
    uint32_t c = 0;
    while (c + 1 < FIN_CLASSES && pos >= counts[c]) pos -= counts[c++];
    const uint32_t job = lists[(uint64_t)c * cap + pos];
    const ExtJob j = jobs[job];
    const ExtRequest r = req[j.req];
    const uint64_t ro = off[r.read_id];
    const uint32_t L = (uint32_t)(off[r.read_id + 1] - ro);
    QueryView q{seq + ro, L, (r.flags & 4) != 0};
    CigarOut cig{cigars + r.cigar_off, r.cigar_cap, 0, false};
    ExtResult out{0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    RefWin w{&ref, r.ref_id, j.ws, j.wlen};
    const LaneMem m{(uint32_t*)(scratch + (uint64_t)(slot / 64) * 64 * scratch_per) + slot % 64, 64, scratch_per / 4};
    AlignInfo a = ssw_finish(q, w, ends[job], sc, m, cig);
    out.status = (uint32_t)a.status;
    if (a.found && a.status == 0) write_found(out, a, j.ws, true, cig.n);
    res[j.req] = out;
}

// extend_seed (src/mapper.rs:532-620) and rescue_align (src/mapper.rs:1094-1153) for one
// request in one thread: the fallback for requests whose scratch slot was too small.
SA_HD void item_extend(uint32_t t, const uint8_t* seq, const uint64_t* off,
                                               RefView ref, Scores sc, uint32_t k,
                                               const ExtRequest* req, uint32_t nreq,
                                               ExtResult* res, uint32_t* cigars,
                                               const uint32_t* chain_anchors, uint8_t* scratch,
                                               uint64_t scratch_per, uint32_t first) {
    if (t >= nreq) return;
    const uint32_t ri = first + t;
    const ExtRequest r = req[ri];
    ExtResult out{0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    uint8_t* scr = scratch + (uint64_t)t * scratch_per;
    const uint64_t ro = off[r.read_id];
    const uint32_t L = (uint32_t)(off[r.read_id + 1] - ro);
    QueryView q{seq + ro, L, (r.flags & 4) != 0};
    CigarOut cig{cigars + r.cigar_off, r.cigar_cap, 0, false};
    const uint64_t rlen = ref.len[r.ref_id];
    if (r.kind == EXT_SEED) {
        const bool use_ssw = r.flags & 1, consistent = (r.flags & 2) != 0;
        uint64_t prs = r.ref_start > r.query_start ? r.ref_start - r.query_start : 0;
        uint64_t pre = (uint64_t)r.ref_end + L - r.query_end;
        if (pre > rlen) pre = rlen;
        bool gapped = true;
        AlignInfo a{false, 0, 0, 0, 0, 0, 0, 0};
        uint64_t result_ref_start = 0;
        if (prs + L == pre && consistent) {
            RefWin w{&ref, r.ref_id, prs, L};
            uint32_t hd = 0;
            for (uint32_t x = 0; x < L; ++x) hd += q[x] != w[x];
            if ((float)hd / (float)L < 0.05f) {
                a = hamming_align(q, w, L, sc, cig);
                result_ref_start = prs + a.ref_start;
                gapped = false;
            }
        }
        if (gapped) {
            uint32_t padding = L / 10;
            if (use_ssw) {
                uint64_t ws = prs > padding ? prs - padding : 0;
                uint64_t we = pre + padding < rlen ? pre + padding : rlen;
                RefWin w{&ref, r.ref_id, ws, (uint32_t)(we - ws)};
                a = ssw_aligner_align(q, w, sc, scr, scratch_per, cig);
                result_ref_start = ws + a.ref_start;
            } else {
                uint64_t ds = r.ref_start > (uint64_t)L + padding ? r.ref_start - ((uint64_t)L + padding) : 0;
                uint64_t de = (uint64_t)r.ref_end + L + padding;
                if (de > rlen) de = rlen;
                RefWin w{&ref, r.ref_id, ds, (uint32_t)(de - ds)};
                // copy the chain anchors into scratch, prune them, then extend
                uint32_t* an = (uint32_t*)scr;
                uint64_t an_bytes = ((uint64_t)r.n_anchors * 8 + 15) & ~15ull;
                if (an_bytes >= scratch_per) {
                    a.status = 2;
                } else {
                    for (uint32_t x = 0; x < 2 * r.n_anchors; ++x) an[x] = chain_anchors[2 * r.anchor_off + x];
                    uint32_t na = remove_spurious_anchors(an, r.n_anchors);
                    a = extend_piecewise(q, w, an, na, padding, sc, k, scr + an_bytes, scratch_per - an_bytes, cig);
                    result_ref_start = a.ref_start + ds;
                }
            }
        }
        out.status = (uint32_t)a.status;
        if (a.found && a.status == 0) {
            out.found = 1;
            out.gapped = gapped ? 1 : 0;
            out.score = a.score;
            out.ref_start = (uint32_t)result_ref_start;
            out.length = a.ref_end - a.ref_start;
            out.query_start = a.query_start;
            out.query_end = a.query_end;
            out.edit_distance = a.edit_distance;
            out.n_cigar = cig.n;
        }
    } else if (r.kind == EXT_HAMMING) {
        const RefWin w{&ref, r.ref_id, r.win_start, r.win_end - r.win_start};
        if (w.len == L) {
            AlignInfo a = hamming_align(q, w, L, sc, cig);
            out.status = (uint32_t)a.status;
            if (a.found && a.status == 0) write_found(out, a, r.win_start, false, cig.n);
        }
    } else if (r.kind >= EXT_ALIGN_PW) {
        const RefWin w{&ref, r.ref_id, r.win_start, r.win_end - r.win_start};
        AlignInfo a = pw_probe(r, q, w, sc, k, chain_anchors, scr, scratch_per, cig);
        out.status = (uint32_t)a.status;
        if (a.found && a.status == 0) write_found(out, a, r.win_start, true, cig.n);
    } else {
        // rescue: shared-substring test, then SSW over the window
        uint32_t ws = r.win_start, we = r.win_end;
        RefWin w{&ref, r.ref_id, ws, we - ws};
        uint32_t sub = 2 * k / 3, step = k / 3;
        bool shared = (r.flags & 8) != 0;  // bit3: plain alignment without the substring precheck
        uint32_t lim = L > sub ? L - sub : 0;
        for (uint32_t i0 = 0; i0 < lim && !shared; i0 += step) {
            for (uint32_t p = 0; p + sub <= w.len && !shared; ++p) {
                uint32_t x = 0;
                while (x < sub && w[p + x] == q[i0 + x]) ++x;
                if (x == sub) shared = true;
            }
        }
        if (shared) {
            AlignInfo a = ssw_aligner_align(q, w, sc, scr, scratch_per, cig);
            out.status = (uint32_t)a.status;
            if (a.found && a.status == 0) {
                out.found = 1;
                out.gapped = 1;
                out.score = a.score;
                out.ref_start = ws + a.ref_start;
                out.length = a.ref_end - a.ref_start;
                out.query_start = a.query_start;
                out.query_end = a.query_end;
                out.edit_distance = a.edit_distance;
                out.n_cigar = cig.n;
            }
        }
    }
    res[ri] = out;
}


// ---------------------------------------------------------- index build ---
// src/indexer.rs:18-170 on the device. The reference is cut into chunks of
// positions; a chunk's thread scans its bases plus the k - 1 that follow, which
// gives the syncmers starting in the chunk exactly as a scan of the whole contig
// does (a syncmer decision depends only on the k bases of its k-mer, and a
// packed reference has no wildcard restarts).
struct IxChunk {
    uint32_t contig, start, end;  // syncmer start positions [start, end) of the contig
    uint32_t pad;
};
template <bool COUNT>
SA_HD void item_ix_syncmers(uint32_t t, const uint64_t* words, const uint64_t* word_off, const uint64_t* len,
                            const IxChunk* chunks, uint32_t nchunks, SeedParams p, uint64_t* counts,
                            const uint64_t* offsets, Syncmer* out) {
    if (t >= nchunks) return;
    const IxChunk ch = chunks[t];
    const uint64_t* w = words + word_off[ch.contig];
    const uint32_t L = (uint32_t)len[ch.contig];
    const uint32_t stop = ch.end + (uint32_t)p.k - 1 < L ? ch.end + (uint32_t)p.k - 1 : L;
    const uint32_t base = ch.start;
    uint64_t n = 0;
    Syncmer* o = COUNT ? nullptr : out + offsets[t];
    auto code = [&](uint32_t i) { const uint32_t a = base + i; return (uint32_t)(w[a >> 5] >> ((a & 31) * 2)) & 3u; };
    auto emit = [&](uint64_t h, bool fwd, uint32_t pos) {
        if (!COUNT) o[n] = Syncmer{h, ((base + pos) << 1) | (fwd ? 1u : 0u), 0};
        ++n;
    };
    if (p.k - p.s + 1 == 5) syncmer_scan_w<5>(stop - base, p, code, emit);
    else syncmer_scan(stop - base, p, code, emit);
    if (COUNT) counts[t] = n;
}

// Randstrobe of syncmer a (src/seeding/strobes.rs:145-176) as the index entry
// (src/indexer.rs:143-160); sy_off: first syncmer of each contig (n_contigs + 1).
SA_HD void item_ix_randstrobes(uint32_t a, const Syncmer* sy, uint64_t n, const uint64_t* sy_off, uint32_t n_contigs,
                               SeedParams p, RefEntry* out) {
    if (a >= n) return;
    uint32_t lo = 0, hi = n_contigs;  // the contig c with sy_off[c] <= a < sy_off[c + 1]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (sy_off[mid] <= a) lo = mid; else hi = mid;
    }
    const uint32_t c = lo;
    const uint64_t cs = sy_off[c];
    const uint32_t m = (uint32_t)(sy_off[c + 1] - cs);
    const uint32_t la = (uint32_t)(a - cs);
    const uint32_t b = pick_strobe2(la, m, p, [&](uint32_t i) { return sy[cs + i]; });
    const Syncmer s1 = sy[a], s2 = sy[cs + b];
    const uint64_t h = randstrobe_hash(s1.hash, s2.hash, s1.pos_fwd & 1, s2.pos_fwd & 1, p);
    const uint32_t p1 = s1.pos_fwd >> 1, p2 = s2.pos_fwd >> 1;
    out[a] = RefEntry{(h & REF_HASH_MASK) | (uint64_t)((p2 - p1) & 0xFF), p1, c};
}

// Sort keys: the secondary key (position, then contig) and the hash key of the
// entry a permutation points at.
SA_HD void item_ix_poskey(uint32_t i, const RefEntry* e, uint64_t n, uint64_t* key, uint32_t* idx) {
    if (i >= n) return;
    key[i] = ((uint64_t)e[i].position << 32) | e[i].ref_index;
    idx[i] = i;
}
SA_HD void item_ix_hashkey(uint32_t i, const RefEntry* e, const uint32_t* perm, uint64_t n, uint64_t* key) {
    if (i >= n) return;
    key[i] = e[perm[i]].hash_offset;
}
SA_HD void item_ix_gather(uint32_t i, const RefEntry* e, const uint32_t* perm, uint64_t n, RefEntry* out) {
    if (i >= n) return;
    out[i] = e[perm[i]];
}

// Bucket directory (src/indexer.rs:77-129 as the host build writes it): slot t
// holds the first entry, after entry 0, that starts a new hash with top bits
// >= t; slot 0 is 0; slots past the last hash hold n. `bs` is filled with n
// before. Entry pos starting a new hash writes slots (top(pos - 1), top(pos)].
SA_HD void item_ix_buckets(uint32_t i, const RefEntry* e, uint64_t n, int bits, uint64_t* bs, uint64_t* first_new) {
    const uint64_t pos = i;
    if (pos == 0 || pos >= n) return;
    const uint64_t prev = e[pos - 1].hash_offset & REF_HASH_MASK, cur = e[pos].hash_offset & REF_HASH_MASK;
    if (prev == cur) return;
    const uint64_t t0 = prev >> (64 - bits), t1 = cur >> (64 - bits);
    for (uint64_t t = t0 + 1; t <= t1; ++t) bs[t] = pos;
    if (t0 == (e[0].hash_offset & REF_HASH_MASK) >> (64 - bits)) SA_ATOMIC_MIN(first_new, pos);
}
// Slots 1 .. top(entry 0) take the first new hash (or n), as the host build's
// first push does.
SA_HD void item_ix_buckets_head(uint32_t t, const RefEntry* e, uint64_t n, int bits, uint64_t* bs,
                                const uint64_t* first_new) {
    if (n == 0) return;
    const uint64_t top0 = (e[0].hash_offset & REF_HASH_MASK) >> (64 - bits);
    if (t == 0) bs[0] = 0;
    else if (t <= top0) bs[t] = *first_new < n ? *first_new : n;
}

SA_HD void item_ix_fill(uint32_t i, uint64_t* a, uint64_t n, uint64_t v) {
    if (i < n) a[i] = v;
}

// Per hash run starting at entry i: one distinct hash, and its length (at most
// IX_HIST - 1, longer runs in the last slot) in the histogram of repeated hashes.
constexpr uint32_t IX_HIST = 102;
SA_HD void item_ix_runs(uint32_t i, const RefEntry* e, uint64_t n, unsigned long long* hist) {
    if (i >= n) return;
    const uint64_t h = e[i].hash_offset & REF_HASH_MASK;
    if (i > 0 && (e[i - 1].hash_offset & REF_HASH_MASK) == h) return;
    uint64_t len = 1;
    while (len < IX_HIST - 1 && i + len < n && (e[i + len].hash_offset & REF_HASH_MASK) == h) ++len;
    SA_ATOMIC_ADD(&hist[0], 1ull);  // slot 0: distinct hashes
    if (len >= 2) SA_ATOMIC_ADD(&hist[len], 1ull);
}

}  // namespace sa
