// HIP kernels of the mapping path: seeding -> index lookups/hits -> anchors ->
// collinear chaining -> NAMs, then extension. Most threads run one work item of
// sa_items.hpp (a read, a read orientation, an extension request or job); the
// SSW score passes run one alignment per lane group (sa_ssw_group.hpp).
#include <hip/hip_runtime.h>

#include "sa_chain_group.hpp"
#include "sa_hits_group.hpp"
#include "sa_pw_group.hpp"
#include "sa_items.hpp"
#include "sa_seed_wave.hpp"
#include "sa_ssw_group.hpp"

namespace sa {

// Seeding a thread per read (reads of at most skip_len bases are skipped: the
// wave form seeds them).
__global__ void __launch_bounds__(256) k_seed(const uint8_t* __restrict__ seq, const uint64_t* __restrict__ off,
                                             uint32_t n, SeedParams p, Syncmer* __restrict__ sy,
                                             QRand* __restrict__ qr, uint32_t* __restrict__ qr_n,
                                             uint32_t skip_len) {
    item_seed(blockIdx.x * blockDim.x + threadIdx.x, seq, off, n, p, sy, qr, qr_n, skip_len);
}

// Seeding a wave per read (sa_seed_wave.hpp) for reads of at most
// SEED_WAVE_MAX_LEN bases; longer ones are left to k_seed.
__global__ void __launch_bounds__(256) k_seed_wave(const uint8_t* __restrict__ seq, const uint64_t* __restrict__ off,
                                                  uint32_t n, SeedParams p, QRand* __restrict__ qr,
                                                  uint32_t* __restrict__ qr_n) {
    __shared__ SeedWaveLds lds[4];
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const uint32_t i = blockIdx.x * 4 + w;
    if (i >= n) return;
    const uint64_t o = off[i];
    const uint32_t L = (uint32_t)(off[i + 1] - o);
    if (L > SEED_WAVE_MAX_LEN) return;
    read_randstrobes_wave(seq + o, L, p, qr + 2 * o, qr + 2 * o + L, &qr_n[2 * i], &qr_n[2 * i + 1], lds[w], lane);
}

__global__ void __launch_bounds__(256) k_hits(IndexView ix, MapParams mp, const uint64_t* __restrict__ off,
                                             uint32_t n, const QRand* __restrict__ qr,
                                             const uint32_t* __restrict__ qr_n, Hit* __restrict__ hits,
                                             uint64_t* __restrict__ cnt, HitCounts* __restrict__ hc,
                                             uint64_t* __restrict__ acap, uint32_t* __restrict__ lookup_bytes) {
    item_hits(blockIdx.x * blockDim.x + threadIdx.x, ix, mp, off, n, qr, qr_n, hits, cnt, hc, acap, lookup_bytes);
}

// find_hits on lane groups (sa_hits_group.hpp): one read orientation per G
// lanes, same outputs as k_hits.
template <int G>
__global__ void __launch_bounds__(256) k_hits_grp(IndexView ix, MapParams mp, const uint64_t* __restrict__ off,
                                                  uint32_t n, const QRand* __restrict__ qr,
                                                  const uint32_t* __restrict__ qr_n, Hit* __restrict__ hits,
                                                  uint64_t* __restrict__ cnt, HitCounts* __restrict__ hc,
                                                  uint64_t* __restrict__ acap, uint32_t* __restrict__ lookup_bytes) {
    const uint32_t lane = threadIdx.x % 64;
    const Grp<G> gp{lane % G, lane / G * G};
    const uint32_t t = (blockIdx.x * blockDim.x + threadIdx.x) / G;
    if (t >= 2 * n) return;
    const uint32_t i = t >> 1, o = t & 1;
    const uint64_t ro = off[i];
    const uint32_t L = (uint32_t)(off[i + 1] - ro);
    const uint64_t base = 2 * ro + (uint64_t)o * L;
    IndexView lx = ix;
    HitCounts d;
    uint64_t cap = 0;
    const uint32_t nq = qr_n[t];
    const uint32_t nh = find_hits_grp<G>(gp, lx, qr + base, nq, mp, hits + base, cnt + base, d, cap);
    const uint32_t br = gp.sum(lx.bucket_reads), er = gp.sum(lx.entry_reads);
    if (gp.r == 0) {
        hc[t] = d;
        acap[t] = cap;
        // query randstrobes in, bucket_starts words and index entries read, hits out
        lookup_bytes[t] = (uint32_t)sizeof(QRand) * nq + 8u * br + (uint32_t)sizeof(RefEntry) * er +
                          (uint32_t)sizeof(Hit) * nh;
    }
}
template __global__ void k_hits_grp<8>(IndexView, MapParams, const uint64_t*, uint32_t, const QRand*, const uint32_t*,
                                       Hit*, uint64_t*, HitCounts*, uint64_t*, uint32_t*);
template __global__ void k_hits_grp<16>(IndexView, MapParams, const uint64_t*, uint32_t, const QRand*,
                                        const uint32_t*, Hit*, uint64_t*, HitCounts*, uint64_t*, uint32_t*);
template __global__ void k_hits_grp<64>(IndexView, MapParams, const uint64_t*, uint32_t, const QRand*,
                                        const uint32_t*, Hit*, uint64_t*, HitCounts*, uint64_t*, uint32_t*);

// Chaining, first tier: one read orientation per group of G lanes with its
// anchors in LDS (sa_chain_group.hpp), for anchor capacities up to tiers.cap[0]
// (CAP); orientations not chained (src/chainer.rs:206-216) get empty results,
// larger ones are listed for the later tiers (which run only what is listed).
template <int G, int CAP>
__global__ void __launch_bounds__(64) k_chain_t1(IndexView ix, ChainParams cp, const uint64_t* __restrict__ off,
                                                 uint32_t n, const Hit* __restrict__ hits,
                                                 const HitCounts* __restrict__ hc, const uint64_t* __restrict__ acap,
                                                 const uint64_t* __restrict__ aoff, ChainBufs b,
                                                 uint32_t* __restrict__ nam_n, uint32_t* __restrict__ anchors_raw,
                                                 const float* __restrict__ log2tab, ChainTiers tiers,
                                                 const uint32_t* __restrict__ abort) {
    __shared__ ChainLds<CAP> lds[64 / G];
    const uint32_t g = threadIdx.x / G;
    const Grp<G> gp{threadIdx.x % G, g * G};
    const uint32_t t = blockIdx.x * (64 / G) + g;
    if (t >= 2 * n) return;
    // abort (when given and set): the anchors do not fit their buffer; every
    // orientation gets empty results and the batch is run again another way
    if (!orientation_chained(hc, t) || (abort && *abort)) {
        if (gp.r == 0) {
            nam_n[t] = 0;
            anchors_raw[t] = 0;
            b.alg_bytes[t] = 0;
        }
        return;
    }
    const uint64_t cap = acap[t];
    if (cap > tiers.cap[0]) {
        if (gp.r == 0) {
            const uint32_t tier = cap > tiers.cap[2] ? 2u : cap > tiers.cap[1] ? 1u : 0u;
            tiers.list[tier][atomicAdd(&tiers.count[tier], 1u)] = t;
        }
        return;
    }
    chain_grp<G, CAP>(gp, lds[g], t, ix, cp, off, hits, hc, aoff, b, nam_n, anchors_raw, log2tab);
}

// Chaining, later LDS tiers: the orientations of list `which`, a group of G
// lanes each, grid-stride over the list.
template <int G, int CAP>
__global__ void __launch_bounds__(64) k_chain_tn(IndexView ix, ChainParams cp, const uint64_t* __restrict__ off,
                                                 const Hit* __restrict__ hits, const HitCounts* __restrict__ hc,
                                                 const uint64_t* __restrict__ aoff, ChainBufs b,
                                                 uint32_t* __restrict__ nam_n, uint32_t* __restrict__ anchors_raw,
                                                 const float* __restrict__ log2tab, ChainTiers tiers,
                                                 uint32_t which) {
    __shared__ ChainLds<CAP> lds[64 / G];
    const uint32_t g = threadIdx.x / G;
    const Grp<G> gp{threadIdx.x % G, g * G};
    const uint32_t cnt = tiers.count[which];
    for (uint32_t x = blockIdx.x * (64 / G) + g; x < cnt; x += gridDim.x * (64 / G))
        chain_grp<G, CAP>(gp, lds[g], tiers.list[which][x], ix, cp, off, hits, hc, aoff, b, nam_n, anchors_raw,
                          log2tab);
}

// Chaining, last tier: the listed orientations above the LDS tiers, a thread each
// with its work arrays in global memory (item_chain), grid-stride over the list.
__global__ void __launch_bounds__(256) k_chain(IndexView ix, ChainParams cp, const uint64_t* __restrict__ off,
                                              uint32_t n, const Hit* __restrict__ hits,
                                              const HitCounts* __restrict__ hc, const uint64_t* __restrict__ acap,
                                              const uint64_t* __restrict__ aoff, ChainBufs b,
                                              uint32_t* __restrict__ nam_n, uint32_t* __restrict__ anchors_raw,
                                              const float* __restrict__ log2tab,
                                              const float* __restrict__ same_diag, ChainTiers tiers) {
    const uint32_t cnt = tiers.count[CHAIN_TIERS - 2];
    for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < cnt; x += gridDim.x * blockDim.x)
        item_chain(tiers.list[CHAIN_TIERS - 2][x], ix, cp, off, n, hits, hc, acap, aoff, b, nam_n, anchors_raw, log2tab, same_diag);
}

// Anchor runs and capacities of hits handed in from outside (the chaining
// stage on its own): hit_runs per read orientation, as k_hits ends.
__global__ void __launch_bounds__(256) k_hit_runs(IndexView ix, const uint64_t* __restrict__ off, uint32_t n,
                                                  Hit* __restrict__ hits, const HitCounts* __restrict__ hc,
                                                  uint64_t* __restrict__ acap) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * n) return;
    const uint32_t i = t >> 1, o = t & 1;
    const uint64_t ro = off[i];
    const uint32_t L = (uint32_t)(off[i + 1] - ro);
    acap[t] = hit_runs(ix, hits + 2 * ro + (uint64_t)o * L, hc[t].n_hits);
}

__global__ void __launch_bounds__(64) k_zero_u32(uint32_t* __restrict__ p, uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n; i += 64) p[i] = 0;
}

__global__ void __launch_bounds__(256) k_finalize(const uint8_t* __restrict__ seq, const uint64_t* __restrict__ off,
                                                 uint32_t n, RefView ref, uint32_t k,
                                                 const uint32_t* __restrict__ nam_n,
                                                 const uint64_t* __restrict__ aoff, const NamRec* __restrict__ nams_in,
                                                 const uint64_t* __restrict__ out_off, NamRec* __restrict__ nams_out) {
    item_finalize(blockIdx.x * blockDim.x + threadIdx.x, seq, off, n, ref, k, nam_n, aoff, nams_in, out_off, nams_out);
}

__global__ void __launch_bounds__(128) k_extend(const uint8_t* __restrict__ seq, const uint64_t* __restrict__ off,
                                               RefView ref, Scores sc, uint32_t k,
                                               const ExtRequest* __restrict__ req, uint32_t nreq,
                                               ExtResult* __restrict__ res, uint32_t* __restrict__ cigars,
                                               const uint32_t* __restrict__ chain_anchors, uint8_t* __restrict__ scratch,
                                               uint64_t scratch_per, uint32_t first) {
    item_extend(blockIdx.x * blockDim.x + threadIdx.x, seq, off, ref, sc, k, req, nreq, res, cigars, chain_anchors,
                scratch, scratch_per, first);
}

// Mismatch bits as a sequence view for hamming_align: q[i] == r[i] exactly when
// bit i of the mask is clear.
struct MismatchBits {
    uint64_t w[8];
    __device__ uint8_t operator[](uint32_t i) const { return (uint8_t)((w[i >> 6] >> (i & 63)) & 1u); }
};
struct ZeroBits {
    __device__ uint8_t operator[](uint32_t) const { return 0; }
};

__device__ inline SegSum shfl_seg(const SegSum& s, int d, bool up) {
    SegSum b;
    b.tot = up ? __shfl_up(s.tot, d) : __shfl_down(s.tot, d);
    b.pre = up ? __shfl_up(s.pre, d) : __shfl_down(s.pre, d);
    b.suf = up ? __shfl_up(s.suf, d) : __shfl_down(s.suf, d);
    b.best = up ? __shfl_up(s.best, d) : __shfl_down(s.best, d);
    return b;
}

// two_diagonal_score for one wave (every lane returns the score), queries of at
// most 512 bases whose cell scores sum to at most 30000 in magnitude. Query
// positions are dealt to lanes in contiguous slices of at most 8; max-sum-segment
// summaries scanned across the wave give each lane the best segment ending
// before its slice on d1 and the best prefix after it on d2, and the segment
// after a gap is read from the lane that holds its first position.
template <typename Q>
__device__ int32_t two_diagonal_score_wave(const Q& q, const RefWin& w, uint32_t L, int64_t d1, int64_t d2,
                                           const Scores& sc, uint32_t lane) {
    constexpr int32_t OUT = -(1 << 20);  // a cell off the window: below any sum of the other cells
    const uint32_t per = (L + 63) / 64;
    const uint32_t x0 = min(L, lane * per), n = min(L, x0 + per) - x0;
    const RefWin w1 = w, w2 = w;
    int32_t c1[8], c2[8];
    SegSum s1, s2;
    SA_UNROLL
    for (uint32_t k = 0; k < 8; ++k) {
        c1[k] = c2[k] = OUT;
        if (k < n) {
            const int32_t qc = ssw_code(q[x0 + k]);
            const int64_t p1 = (int64_t)(x0 + k) + d1, p2 = (int64_t)(x0 + k) + d2;
            if (p1 >= 0 && p1 < (int64_t)w.len) c1[k] = qc == (int32_t)w1.code((uint32_t)p1) ? sc.match : -sc.mismatch;
            if (p2 >= 0 && p2 < (int64_t)w.len) c2[k] = qc == (int32_t)w2.code((uint32_t)p2) ? sc.match : -sc.mismatch;
            s1.push(c1[k]);
            s2.push(c2[k]);
        }
    }
    // inclusive scans: d1 from lane 0 up, d2 from lane 63 down
    SA_UNROLL
    for (int d = 1; d < 64; d <<= 1) {
        SegSum l1 = shfl_seg(s1, d, true);
        const SegSum r2 = shfl_seg(s2, d, false);
        if (lane >= (uint32_t)d) {
            l1.join(s1);
            s1 = l1;
        }
        if (lane + d < 64) s2.join(r2);
    }
    const SegSum e1 = shfl_seg(s1, 1, true), e2 = shfl_seg(s2, 1, false);
    int32_t suf = lane >= 1 ? e1.suf : 0, pre = lane + 1 < 64 ? e2.pre : 0;
    int32_t best = 0, a[8], b[8];
    SA_UNROLL
    for (uint32_t k = 0; k < 8; ++k) {  // a[k]: best segment on d1 ending at x0 + k
        a[k] = OUT;
        if (k < n) {
            a[k] = suf + c1[k];
            suf = max(a[k], 0);
            best = max(best, a[k]);
        }
    }
    SA_UNROLL
    for (int k = 7; k >= 0; --k) {  // b[k]: best segment on d2 starting at x0 + k
        b[k] = OUT;
        if ((uint32_t)k < n) {
            b[k] = c2[k] + pre;
            pre = max(b[k], 0);
            best = max(best, b[k]);
        }
    }
    const int64_t g = d1 > d2 ? d1 - d2 : d2 - d1;
    if (g > 0 && g <= TWO_DIAGONAL_MAX_GAP) {
        const uint32_t off = d2 > d1 ? 1u : 1u + (uint32_t)g;
        const int32_t cost = sc.gap_open + (int32_t)(g - 1) * sc.gap_extend;
        const uint32_t q0 = off / per, r0 = off % per;
        SA_UNROLL
        for (uint32_t k = 0; k < 8; ++k) {
            if (k >= per) break;
            // position x0 + k + off: index (r0 + k) mod per of lane lane + q0 + (r0 + k) / per
            const uint32_t i = r0 + k, carry = i >= per ? 1u : 0u, idx = i - carry * per;
            int32_t v = b[0];
            SA_UNROLL
            for (uint32_t j = 1; j < 8; ++j) v = j == idx ? b[j] : v;
            v = __shfl(v, (int)((lane + q0 + carry) & 63u));
            if (k < n && x0 + k + off < L) best = max(best, a[k] + v - cost);
        }
    }
    SA_UNROLL
    for (int d = 32; d >= 1; d >>= 1) best = max(best, __shfl_xor(best, d));
    return best;
}

// First stage of extension, one wave per request (item_ext_prep is the same
// logic in one thread): the ungapped case of extend_seed (src/mapper.rs:567-588)
// compares the read with the reference in 64-base ballots, and the rescue
// prefilter has_shared_substring (src/mapper.rs:1156-1170) tests 64 window
// positions per step. Gapped alignments are queued as jobs.
__device__ inline void ext_prep_one(uint32_t t, uint32_t lane, const uint8_t* __restrict__ seq,
                                    const uint64_t* __restrict__ off, const RefView& ref, const Scores& sc, uint32_t k,
                                    const ExtRequest* __restrict__ req, ExtResult* __restrict__ res,
                                    uint32_t* __restrict__ cigars, ExtJob* __restrict__ ssw_jobs,
                                    ExtJob* __restrict__ long_jobs, ExtJob* __restrict__ pw_jobs,
                                    uint32_t* __restrict__ counters) {
    const ExtRequest r = req[t];
    const uint64_t ro = off[r.read_id];
    const uint32_t L = (uint32_t)(off[r.read_id + 1] - ro);
    QueryView q{seq + ro, L, (r.flags & 4) != 0};
    const uint64_t rlen = ref.len[r.ref_id];
    ExtResult out{0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    auto push_ssw = [&](uint64_t ws, uint64_t we, uint32_t jflags) {
        if (we - ws > 2000 || L == 0) return;  // Aligner::align returns None (src/aligner.rs:83-86)
        ExtJob j{t, r.ref_id, ws, (uint32_t)(we - ws), jflags};
        if (L > SSW_GROUP_MAX_QUERY || (uint64_t)sc.match * L > 30000 || sc.match > 127 || sc.mismatch > 127)
            long_jobs[atomicAdd(&counters[1], 1u)] = j;
        else
            ssw_jobs[atomicAdd(&counters[0], 1u)] = j;
    };
    if (r.kind == EXT_SEED) {
        const bool use_ssw = r.flags & 1, consistent = (r.flags & 2) != 0;
        const uint64_t prs = r.ref_start > r.query_start ? r.ref_start - r.query_start : 0;
        uint64_t pre = (uint64_t)r.ref_end + L - r.query_end;
        if (pre > rlen) pre = rlen;
        if (prs + L == pre && consistent) {
            RefWin w{&ref, r.ref_id, prs, L};
            if (L <= 512) {
                MismatchBits m;
                uint32_t hd = 0;
#pragma unroll
                for (uint32_t c = 0; c < 8; ++c) {
                    const uint32_t pos = c * 64 + lane;
                    m.w[c] = __ballot(pos < L && q[pos] != w[pos]);
                    hd += (uint32_t)__popcll(m.w[c]);
                }
                if ((float)hd / (float)L < 0.05f) {
                    if (lane == 0) {
                        CigarOut cig{cigars + r.cigar_off, r.cigar_cap, 0, false};
                        AlignInfo a = sc.match > 0 ? hamming_align_bits(m.w, L, sc, cig)
                                                   : hamming_align(ZeroBits{}, m, L, sc, cig);
                        out.status = (uint32_t)a.status;
                        if (a.found && a.status == 0) write_found(out, a, prs, false, cig.n);
                        res[t] = out;
                    }
                    return;
                }
            } else {
                bool done = false;
                if (lane == 0) {
                    uint32_t hd = 0;
                    for (uint32_t x = 0; x < L; ++x) hd += q[x] != w[x];
                    if ((float)hd / (float)L < 0.05f) {
                        CigarOut cig{cigars + r.cigar_off, r.cigar_cap, 0, false};
                        AlignInfo a = hamming_align(q, w, L, sc, cig);
                        out.status = (uint32_t)a.status;
                        if (a.found && a.status == 0) write_found(out, a, prs, false, cig.n);
                        res[t] = out;
                        done = true;
                    }
                }
                if (__ballot(done)) return;
            }
        }
        const uint32_t padding = L / 10;
        const uint64_t ws = prs > padding ? prs - padding : 0;
        const uint64_t we = pre + padding < rlen ? pre + padding : rlen;
        uint32_t jflags = 0;
        if (use_ssw && L <= SSW_GROUP_MAX_QUERY && (uint64_t)sc.match * L <= 30000 && sc.match <= 127 &&
            sc.mismatch <= 127 &&
            sc.gap_open > sc.gap_extend) {
            // lower bound of the SSW score from the NAM's start and end
            // diagonals, in window coordinates (only lane-group jobs use it)
            const int64_t d1 = (int64_t)prs - (int64_t)ws;
            const int64_t d2 = (int64_t)r.ref_end - (int64_t)r.query_end - (int64_t)ws;
            const RefWin w{&ref, r.ref_id, ws, (uint32_t)(we - ws)};
            jflags = two_diagonal_score_wave(q, w, L, d1, d2, sc, lane) >= 255 - sc.mismatch ? SSW_WORD_FIRST : 0u;
        }
        if (lane == 0) {
            if (use_ssw) {
                push_ssw(ws, we, jflags);
            } else {
                const uint64_t ds = r.ref_start > (uint64_t)L + padding ? r.ref_start - ((uint64_t)L + padding) : 0;
                uint64_t de = (uint64_t)r.ref_end + L + padding;
                if (de > rlen) de = rlen;
                pw_jobs[atomicAdd(&counters[2], 1u)] = ExtJob{t, r.ref_id, ds, (uint32_t)(de - ds), 0};
            }
            res[t] = out;
        }
    } else if (r.kind == EXT_HAMMING) {
        // hamming_align (src/aligner.rs:223-259) of the whole query against the window
        if (lane == 0) {
            const RefWin w{&ref, r.ref_id, r.win_start, r.win_end - r.win_start};
            if (w.len == L) {
                CigarOut cig{cigars + r.cigar_off, r.cigar_cap, 0, false};
                AlignInfo a = hamming_align(q, w, L, sc, cig);
                out.status = (uint32_t)a.status;
                if (a.found && a.status == 0) write_found(out, a, r.win_start, false, cig.n);
            }
            res[t] = out;
        }
    } else if (r.kind >= EXT_ALIGN_PW) {
        // the piecewise aligner's entry points: a job for the piecewise kernel
        if (lane == 0) {
            pw_jobs[atomicAdd(&counters[2], 1u)] = ExtJob{t, r.ref_id, r.win_start, r.win_end - r.win_start, 0};
            res[t] = out;
        }
    } else {
        // has_shared_substring unless bit3 asks for a plain alignment
        const uint32_t ws = r.win_start, we = r.win_end;
        RefWin w{&ref, r.ref_id, ws, we - ws};
        const uint32_t sub = 2 * k / 3, step = k / 3;
        bool shared = (r.flags & 8) != 0;
        bool found = false;
        uint32_t hit_i = 0, hit_p = 0;  // the first shared substring: query and window offsets
        const uint32_t lim = L > sub ? L - sub : 0;
        for (uint32_t i0 = 0; i0 < lim && !shared; i0 += step) {
            for (uint32_t p0 = 0; p0 + sub <= w.len && !shared; p0 += 64) {
                const uint32_t p = p0 + lane;
                bool hit = false;
                if (p + sub <= w.len) {
                    uint32_t x = 0;
                    while (x < sub && w[p + x] == q[i0 + x]) ++x;
                    hit = x == sub;
                }
                const uint64_t hm = __ballot(hit);
                if (hm) {
                    shared = found = true;
                    hit_i = i0;
                    hit_p = p0 + (uint32_t)__builtin_ctzll(hm);
                }
            }
        }
        // lower bound of the SSW score: best ungapped segment on the diagonal of
        // that substring, as for seed jobs
        uint32_t jflags = 0;
        if (found) {
            const int64_t d = (int64_t)hit_p - (int64_t)hit_i;  // window offset of query position 0
            const int64_t x0 = d < 0 ? -d : 0;
            const int64_t x1 = (int64_t)L < (int64_t)(we - ws) - d ? (int64_t)L : (int64_t)(we - ws) - d;
            const uint32_t n = x1 > x0 ? (uint32_t)(x1 - x0) : 0;
            const RefWin wd{&ref, r.ref_id, (uint64_t)((int64_t)ws + d + x0), n};
            const uint32_t per = (n + 63) / 64;
            const uint32_t j0 = min(n, lane * per), j1 = min(n, j0 + per);
            SegSum s;
            for (uint32_t j = j0; j < j1; ++j)
                s.push(ssw_code(q[(uint32_t)x0 + j]) == (int32_t)wd.code(j) ? sc.match : -sc.mismatch);
            SA_UNROLL
            for (int dd = 1; dd < 64; dd <<= 1) {
                SegSum b;
                b.tot = __shfl_down(s.tot, dd);
                b.pre = __shfl_down(s.pre, dd);
                b.suf = __shfl_down(s.suf, dd);
                b.best = __shfl_down(s.best, dd);
                s.join(b);
            }
            jflags = s.best >= 255 - sc.mismatch ? SSW_WORD_FIRST : 0u;  // lane 0 holds the whole diagonal
        }
        if (lane == 0) {
            if (shared) push_ssw(ws, we, jflags);
            res[t] = out;
        }
    }
}

// The requests [0, nreq), or [0, *dnreq) when dnreq is given, a wave each, grid-stride.
__global__ void __launch_bounds__(256) k_ext_prep(const uint8_t* __restrict__ seq, const uint64_t* __restrict__ off,
                                                 RefView ref, Scores sc, uint32_t k, const ExtRequest* __restrict__ req,
                                                 uint32_t nreq, const uint32_t* __restrict__ dnreq,
                                                 ExtResult* __restrict__ res, uint32_t* __restrict__ cigars,
                                                 ExtJob* __restrict__ ssw_jobs, ExtJob* __restrict__ long_jobs,
                                                 ExtJob* __restrict__ pw_jobs, uint32_t* __restrict__ counters) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t n = dnreq ? *dnreq : nreq;
    for (uint32_t t = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; t < n; t += gridDim.x * (blockDim.x >> 6))
        ext_prep_one(t, lane, seq, off, ref, sc, k, req, res, cigars, ssw_jobs, long_jobs, pw_jobs, counters);
}

// Piecewise job `job` of the list with scratch slot t on the lane group gp.
template <int G>
__device__ inline void ext_pw_job(const Grp<G>& gp, uint32_t job, uint32_t t, const uint8_t* __restrict__ seq,
                                  const uint64_t* __restrict__ off, const RefView& ref, const Scores& sc, uint32_t k,
                                  const ExtRequest* __restrict__ req, const ExtJob* __restrict__ jobs,
                                  ExtResult* __restrict__ res, uint32_t* __restrict__ cigars,
                                  const uint32_t* __restrict__ chain_anchors, uint8_t* __restrict__ scratch,
                                  uint64_t scratch_per, DpCol* __restrict__ rows, uint8_t* __restrict__ dirs,
                                  uint32_t row_cap, uint32_t dir_rows) {
    const ExtJob j = jobs[job];
    const ExtRequest r = req[j.req];
    uint8_t* scr = scratch + (uint64_t)t * scratch_per;
    const uint64_t ro = off[r.read_id];
    const uint32_t L = (uint32_t)(off[r.read_id + 1] - ro);
    QueryView q{seq + ro, L, (r.flags & 4) != 0};
    CigarOut cig{cigars + r.cigar_off, r.cigar_cap, 0, false};
    ExtResult out{0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    RefWin w{&ref, r.ref_id, j.ws, j.wlen};
    const DpRows dr{rows + (uint64_t)t * row_cap, 1, row_cap, dirs + (uint64_t)t * row_cap * dir_rows, dir_rows};
    uint32_t* an = (uint32_t*)scr;
    const uint64_t an_bytes = ((uint64_t)r.n_anchors * 8 + 15) & ~15ull;
    if (r.kind != EXT_SEED) {
        const AlignInfo a = pw_probe(r, q, w, sc, k, chain_anchors, scr, scratch_per, cig, dr);
        out.status = (uint32_t)a.status;
        if (a.found && a.status == 0) write_found(out, a, j.ws, true, cig.n);
    } else if (an_bytes >= scratch_per) {
        out.status = 2;
    } else {
        for (uint32_t x = 0; x < 2 * r.n_anchors; ++x) an[x] = chain_anchors[2 * r.anchor_off + x];
        const uint32_t na = remove_spurious_anchors(an, r.n_anchors);
        const AlignInfo a = extend_piecewise(q, w, an, na, L / 10, sc, k, scr + an_bytes, scratch_per - an_bytes, cig,
                                             dr, DpGroup<G>{gp});
        out.status = (uint32_t)a.status;
        if (a.found && a.status == 0) write_found(out, a, j.ws, true, cig.n);
    }
    if (gp.r == 0) res[j.req] = out;
}

// Piecewise extension with G lanes per job (sa_pw_group.hpp); the lane group
// with scratch slot x keeps its DP row buffer at rows + x * row_cap and its
// direction bytes at dirs + x * row_cap * dir_rows. The lanes run the job's
// scalar steps alike and fill its DPs together; lane 0 writes the result.
// Job counts as for k_ext_long below.
template <int G>
__global__ void __launch_bounds__(128) k_ext_pw_grp(const uint8_t* __restrict__ seq, const uint64_t* __restrict__ off,
                                                    RefView ref, Scores sc, uint32_t k,
                                                    const ExtRequest* __restrict__ req,
                                                    const ExtJob* __restrict__ jobs, uint32_t njobs,
                                                    const uint32_t* __restrict__ dn,
                                                    ExtResult* __restrict__ res, uint32_t* __restrict__ cigars,
                                                    const uint32_t* __restrict__ chain_anchors,
                                                    uint8_t* __restrict__ scratch, uint64_t scratch_per, uint32_t first,
                                                    DpCol* __restrict__ rows, uint8_t* __restrict__ dirs,
                                                    uint32_t row_cap, uint32_t dir_rows) {
    const uint32_t lane = threadIdx.x % 64;
    const Grp<G> gp{lane % G, lane / G * G};
    const uint32_t t = (blockIdx.x * blockDim.x + threadIdx.x) / G;
    if (dn) {
        const uint32_t n = *dn;
        for (uint32_t x = t; x < n; x += gridDim.x * blockDim.x / G)
            ext_pw_job<G>(gp, x, t, seq, off, ref, sc, k, req, jobs, res, cigars, chain_anchors, scratch, scratch_per,
                          rows, dirs, row_cap, dir_rows);
    } else if (t < njobs) {
        ext_pw_job<G>(gp, first + t, t, seq, off, ref, sc, k, req, jobs, res, cigars, chain_anchors, scratch,
                      scratch_per, rows, dirs, row_cap, dir_rows);
    }
}

// Job counts: a kernel either runs the host-given range [first, first + njobs),
// a thread (lane group) per job with scratch slot = its index in the launch, or,
// when `dn` is given, the device-side count *dn from job 0, grid-stride, each
// thread (lane group) keeping its own scratch slot across its jobs. The second
// form lets a launch be recorded once with a fixed grid and replayed for
// batches whose job counts exist only on the device.
__global__ void __launch_bounds__(128) k_ext_long(const uint8_t* __restrict__ seq, const uint64_t* __restrict__ off,
                                                 RefView ref, Scores sc, const ExtRequest* __restrict__ req,
                                                 const ExtJob* __restrict__ jobs, uint32_t njobs,
                                                 const uint32_t* __restrict__ dn, ExtResult* __restrict__ res,
                                                 uint32_t* __restrict__ cigars, uint8_t* __restrict__ scratch,
                                                 uint64_t scratch_per, uint32_t first) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (dn) {
        const uint32_t n = *dn;
        for (uint32_t x = tid; x < n; x += gridDim.x * blockDim.x)
            item_ext_long(x, tid, seq, off, ref, sc, req, jobs, res, cigars, scratch, scratch_per);
    } else if (tid < njobs) {
        item_ext_long(first + tid, tid, seq, off, ref, sc, req, jobs, res, cigars, scratch, scratch_per);
    }
}

__global__ void __launch_bounds__(128) k_ssw_finish(const uint8_t* __restrict__ seq, const uint64_t* __restrict__ off,
                                                   RefView ref, Scores sc, const ExtRequest* __restrict__ req,
                                                   const ExtJob* __restrict__ jobs, const SswEnds* __restrict__ ends,
                                                   uint32_t njobs, const uint32_t* __restrict__ dn,
                                                   ExtResult* __restrict__ res, uint32_t* __restrict__ cigars,
                                                   uint8_t* __restrict__ scratch, uint64_t scratch_per,
                                                   uint32_t first, const uint32_t* __restrict__ lists,
                                                   const uint32_t* __restrict__ counts, uint32_t cap) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (dn) {
        const uint32_t n = *dn;
        for (uint32_t x = tid; x < n; x += gridDim.x * blockDim.x)
            item_ssw_finish(x, tid, seq, off, ref, sc, req, jobs, ends, res, cigars, scratch, scratch_per, lists,
                            counts, cap);
    } else if (tid < njobs) {
        item_ssw_finish(first + tid, tid, seq, off, ref, sc, req, jobs, ends, res, cigars, scratch, scratch_per,
                        lists, counts, cap);
    }
}

__global__ void __launch_bounds__(256) k_fin_bucket(const SswEnds* __restrict__ ends, uint32_t njobs,
                                                   const uint32_t* __restrict__ dn, uint32_t* __restrict__ lists,
                                                   uint32_t* __restrict__ counts, uint32_t cap) {
    const uint32_t n = dn ? *dn : njobs;
    for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < n; x += gridDim.x * blockDim.x)
        item_fin_bucket(x, ends, lists, counts, cap);
}

// SSW score-pass buckets: jobs are listed by (pass, stripe count) so that each
// launch of k_ssw_pass runs one compile-time stripe count, and the forward passes
// also by reference-window class, since a lane group runs as many columns as its
// window is long and a wave as long as its longest group (seed windows are
// ~1.2 L, rescue windows up to 2000).
constexpr uint32_t SSW_WCLASSES = 4;
SA_HD uint32_t ssw_wclass(uint32_t wlen) { return wlen < 256 ? 0u : wlen < 512 ? 1u : wlen < 1024 ? 2u : 3u; }
// Stripe counts of queries up to SSW_GROUP_MAX_QUERY (512): 32 of 16 byte lanes,
// 64 of 8 word lanes. Buckets: byte forward 0..127 (32 stripe counts x 4
// classes), byte reverse 128..159, word forward 160..415 (64 x 4), word reverse
// 416..479.
constexpr uint32_t SSW_BYTE_SEGS = 32, SSW_WORD_SEGS = 64;
constexpr uint32_t SSW_BYTE_REV0 = SSW_BYTE_SEGS * SSW_WCLASSES;
constexpr uint32_t SSW_WORD_FWD0 = SSW_BYTE_REV0 + SSW_BYTE_SEGS;
constexpr uint32_t SSW_WORD_REV0 = SSW_WORD_FWD0 + SSW_WORD_SEGS * SSW_WCLASSES;
constexpr uint32_t SSW_BUCKETS = SSW_WORD_REV0 + SSW_WORD_SEGS;
SA_HD uint32_t ssw_bucket(bool word, bool rev, uint32_t seg, uint32_t wc) {
    if (rev) return (word ? SSW_WORD_REV0 : SSW_BYTE_REV0) + seg - 1;
    return (word ? SSW_WORD_FWD0 : 0u) + (seg - 1) * SSW_WCLASSES + wc;
}

// Forward byte-pass buckets of the SSW jobs (queries of at most SSW_GROUP_MAX_QUERY);
// the job count is counters[0] on the device (grid-stride). Also
// sums the jobs' forward DP cells (query length x window length) into *cells.
__global__ void __launch_bounds__(256) k_ssw_bucket(const uint64_t* __restrict__ off,
                                                   const ExtRequest* __restrict__ req,
                                                   const ExtJob* __restrict__ jobs, const uint32_t* __restrict__ counters,
                                                   uint32_t* __restrict__ lists, uint32_t* __restrict__ counts,
                                                   uint32_t cap) {
    const uint32_t n = counters[0];
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const uint32_t rid = req[jobs[t].req].read_id;
        const uint32_t L = (uint32_t)(off[rid + 1] - off[rid]);
        const bool word = (jobs[t].flags & SSW_WORD_FIRST) != 0;
        const uint32_t b = ssw_bucket(word, false, word ? (L + 7) / 8 : (L + 15) / 16, ssw_wclass(jobs[t].wlen));
        lists[(uint64_t)b * cap + atomicAdd(&counts[b], 1u)] = t;
    }
}

// One striped score pass (ext/ssw/ssw.c:197-588) per group of G lanes: the
// forward pass over the whole window, or the reverse pass that finds the start
// (ext/ssw/ssw.c:818-888). A saturated forward byte pass sends the job to the
// word buckets; a finished forward pass sends it to the reverse bucket of its
// stripe count; the reverse pass writes the job's SswEnds.
// The job lists of one launch: the window classes of one stripe count, longest
// windows first, so that each wave's groups have similar column counts and the
// long ones start early.
struct SswLaunchLists {
    const uint32_t* list[SSW_WCLASSES];
    uint32_t start[SSW_WCLASSES + 1];  // slot ranges of the lists
};

//
// With dev_counts set, the launch was recorded before the counts existed: the
// slot ranges come from the bucket counts on the device (the window classes of
// stripe count SEG, longest first, for a forward pass; the one reverse bucket
// of SEG for a reverse pass) and the lane groups stride over the slots.
template <int G, bool BYTE, int SEG>
__device__ inline void ssw_pass_job(uint32_t job, uint32_t l, const uint8_t* __restrict__ seq,
                                    const uint64_t* __restrict__ off, const RefView& ref, const Scores& sc,
                                    const ExtRequest* __restrict__ req, const ExtJob* __restrict__ jobs, uint32_t rev,
                                    SswEnd* __restrict__ fwd, SswEnds* __restrict__ ends, uint32_t* __restrict__ lists,
                                    uint32_t* __restrict__ counts, uint32_t cap, uint64_t& cells);

// DP cells a pass computed (query rows x columns run), one counter per pass kind:
// [byte forward, byte reverse, word forward, word reverse]
__device__ __forceinline__ uint32_t ssw_cell_kind(bool byte, uint32_t rev) { return (byte ? 0u : 2u) + (rev ? 1u : 0u); }

template <int G, bool BYTE, int SEG>
__global__ void __launch_bounds__(256) k_ssw_pass(const uint8_t* __restrict__ seq, const uint64_t* __restrict__ off,
                                                 RefView ref, Scores sc, const ExtRequest* __restrict__ req,
                                                 const ExtJob* __restrict__ jobs, SswLaunchLists ll,
                                                 uint32_t rev, SswEnd* __restrict__ fwd,
                                                 SswEnds* __restrict__ ends, uint32_t* __restrict__ lists,
                                                 uint32_t* __restrict__ counts, uint32_t cap, uint32_t dev_counts,
                                                 unsigned long long* __restrict__ cells) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t l = tid % G;
    uint32_t stride = 0xFFFFFFFFu;
    if (dev_counts) {
        uint32_t at = 0;
        SA_UNROLL
        for (uint32_t x = 0; x < SSW_WCLASSES; ++x) {
            ll.start[x] = at;
            if (!rev) at += counts[ssw_bucket(!BYTE, false, SEG, SSW_WCLASSES - 1 - x)];
            else if (x == 0) at += counts[ssw_bucket(!BYTE, true, SEG, 0)];
        }
        ll.start[SSW_WCLASSES] = at;
        stride = gridDim.x * blockDim.x / G;
    }
    const uint32_t total = ll.start[SSW_WCLASSES];
    uint64_t my_cells = 0;
    for (uint32_t slot = tid / G; slot < total; slot += stride) {
        uint32_t c = 0;
        SA_UNROLL
        for (uint32_t x = 1; x < SSW_WCLASSES; ++x) c += slot >= ll.start[x] ? 1u : 0u;
        ssw_pass_job<G, BYTE, SEG>(ll.list[c][slot - ll.start[c]], l, seq, off, ref, sc, req, jobs, rev, fwd, ends,
                                   lists, counts, cap, my_cells);
        if (stride == 0xFFFFFFFFu) break;
    }
    // every lane is here: one atomic per wave
    SA_UNROLL
    for (int s = 32; s > 0; s >>= 1) my_cells += __shfl_xor(my_cells, s);
    if ((threadIdx.x & 63u) == 0 && my_cells) atomicAdd(cells + ssw_cell_kind(BYTE, rev), (unsigned long long)my_cells);
}

template <int G, bool BYTE, int SEG>
__device__ inline void ssw_pass_job(uint32_t job, uint32_t l, const uint8_t* __restrict__ seq,
                                    const uint64_t* __restrict__ off, const RefView& ref, const Scores& sc,
                                    const ExtRequest* __restrict__ req, const ExtJob* __restrict__ jobs, uint32_t rev,
                                    SswEnd* __restrict__ fwd, SswEnds* __restrict__ ends, uint32_t* __restrict__ lists,
                                    uint32_t* __restrict__ counts, uint32_t cap, uint64_t& cells) {
    const ExtJob j = jobs[job];
    const ExtRequest r = req[j.req];
    const uint64_t ro = off[r.read_id];
    const int32_t L = (int32_t)(off[r.read_id + 1] - ro);
    QueryView q{seq + ro, (uint32_t)L, (r.flags & 4) != 0};
    RefCursor w{ref.words + ref.word_off[j.ref_id], j.ws, -1, 0};
    const uint32_t gbase = (__lane_id() & 63u) & ~(uint32_t)(G - 1);
    int32_t cols = 0;
    if (!rev) {
        const SswEnd f = ssw_pass_seg<G, BYTE, SEG>(q, L, false, 0, w, (int32_t)j.wlen, sc, BYTE ? 255 : 65535, l,
                                                    gbase, cols);
        if (l == 0) {
            cells += (uint64_t)L * (uint32_t)cols;
            uint32_t b;
            if (BYTE && f.score == 255) {
                b = ssw_bucket(true, false, (uint32_t)(L + 7) / 8, ssw_wclass(j.wlen));
            } else {
                fwd[job] = f;
                b = ssw_bucket(!BYTE, true, (uint32_t)(f.read + 1 + G - 1) / G, 0);
            }
            lists[(uint64_t)b * cap + atomicAdd(&counts[b], 1u)] = job;
        }
    } else {
        const SswEnd f = fwd[job];
        const SswEnd b = ssw_pass_seg<G, BYTE, SEG>(q, f.read + 1, true, f.read, w, f.ref + 1, sc, f.score, l, gbase,
                                                    cols);
        if (l == 0) {
            cells += (uint64_t)(f.read + 1) * (uint32_t)cols;
            ends[job] = SswEnds{f.score, f.ref, f.read, b.ref, f.read - b.read, f.score > b.score};
        }
    }
}

// ssw_align of one (query, reference) pair in one thread: the device side of the
// SSW link shim (ssw_init / ssw_align), which serves single calls for conformance.
__global__ void __launch_bounds__(64) k_ssw_raw(const int8_t* __restrict__ read, int32_t readLen,
                                               const int8_t* __restrict__ ref, int32_t refLen,
                                               const int8_t* __restrict__ mat, int32_t n, int32_t score_size,
                                               int32_t gapO, int32_t gapE, uint32_t flag, uint32_t filters,
                                               int32_t filterd, int32_t maskLen, uint8_t* __restrict__ scratch,
                                               uint64_t scratch_bytes, uint32_t* __restrict__ cigar,
                                               uint32_t cigar_cap, SswRaw* __restrict__ out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    SswRaw o;
    ssw_align_raw(read, readLen, ref, refLen, mat, n, score_size, gapO, gapE, flag, filters, filterd, maskLen, scratch,
                  scratch_bytes, cigar, cigar_cap, o);
    *out = o;
}


// ------------------------------------------------- index construction ---
template <bool COUNT>
__global__ void __launch_bounds__(256) k_ix_syncmers(const uint64_t* __restrict__ words,
                                                    const uint64_t* __restrict__ word_off,
                                                    const uint64_t* __restrict__ len,
                                                    const IxChunk* __restrict__ chunks, uint32_t nchunks, SeedParams p,
                                                    uint64_t* __restrict__ counts, const uint64_t* __restrict__ offsets,
                                                    Syncmer* __restrict__ out) {
    item_ix_syncmers<COUNT>(blockIdx.x * blockDim.x + threadIdx.x, words, word_off, len, chunks, nchunks, p, counts,
                            offsets, out);
}
__global__ void __launch_bounds__(256) k_ix_randstrobes(const Syncmer* __restrict__ sy, uint64_t n,
                                                       const uint64_t* __restrict__ sy_off, uint32_t n_contigs,
                                                       SeedParams p, RefEntry* __restrict__ out) {
    item_ix_randstrobes(blockIdx.x * blockDim.x + threadIdx.x, sy, n, sy_off, n_contigs, p, out);
}
__global__ void __launch_bounds__(256) k_ix_poskey(const RefEntry* __restrict__ e, uint64_t n, uint64_t* __restrict__ key,
                                                  uint32_t* __restrict__ idx) {
    item_ix_poskey(blockIdx.x * blockDim.x + threadIdx.x, e, n, key, idx);
}
__global__ void __launch_bounds__(256) k_ix_hashkey(const RefEntry* __restrict__ e, const uint32_t* __restrict__ perm,
                                                   uint64_t n, uint64_t* __restrict__ key) {
    item_ix_hashkey(blockIdx.x * blockDim.x + threadIdx.x, e, perm, n, key);
}
__global__ void __launch_bounds__(256) k_ix_gather(const RefEntry* __restrict__ e, const uint32_t* __restrict__ perm,
                                                  uint64_t n, RefEntry* __restrict__ out) {
    item_ix_gather(blockIdx.x * blockDim.x + threadIdx.x, e, perm, n, out);
}
__global__ void __launch_bounds__(256) k_ix_fill(uint64_t* __restrict__ a, uint64_t n, uint64_t v) {
    item_ix_fill(blockIdx.x * blockDim.x + threadIdx.x, a, n, v);
}
__global__ void __launch_bounds__(256) k_ix_buckets(const RefEntry* __restrict__ e, uint64_t n, int bits,
                                                   uint64_t* __restrict__ bs, uint64_t* __restrict__ first_new) {
    item_ix_buckets(blockIdx.x * blockDim.x + threadIdx.x, e, n, bits, bs, first_new);
}
__global__ void __launch_bounds__(256) k_ix_buckets_head(const RefEntry* __restrict__ e, uint64_t n, int bits,
                                                        uint64_t* __restrict__ bs,
                                                        const uint64_t* __restrict__ first_new) {
    item_ix_buckets_head(blockIdx.x * blockDim.x + threadIdx.x, e, n, bits, bs, first_new);
}
__global__ void __launch_bounds__(256) k_ix_runs(const RefEntry* __restrict__ e, uint64_t n,
                                                unsigned long long* __restrict__ hist) {
    item_ix_runs(blockIdx.x * blockDim.x + threadIdx.x, e, n, hist);
}


// ------------------------------------------------------------ fused pass ---
// Kernels of the batch pass that is recorded once as a graph and replayed:
// every size that depends on the data stays on the device.

// Exclusive prefix sum of m u64 values (m fixed when the launches are
// recorded), in three launches: per-block sums of SCAN_PER values, one block
// scanning those sums (total to *total; *flag set when it exceeds cap), and each
// block scanning its values from its sum's offset.
constexpr uint32_t SCAN_PER = 2048;  // 256 threads x 8 values

__device__ inline uint64_t block_sum_256(uint64_t v, uint64_t* sh) {
    SA_UNROLL
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d);
    if ((threadIdx.x & 63u) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    const uint64_t t = sh[0] + sh[1] + sh[2] + sh[3];
    __syncthreads();
    return t;
}

// Exclusive scan of one value per thread over a block of 256; returns the
// thread's exclusive prefix and the block total in *tot.
__device__ inline uint64_t block_excl_256(uint64_t v, uint64_t* sh, uint64_t* tot) {
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    uint64_t inc = v;
    SA_UNROLL
    for (int d = 1; d < 64; d <<= 1) {
        const uint64_t u = __shfl_up(inc, d);
        if (lane >= (uint32_t)d) inc += u;
    }
    if (lane == 63) sh[w] = inc;
    __syncthreads();
    uint64_t before = 0;
    for (uint32_t x = 0; x < w; ++x) before += sh[x];
    *tot = sh[0] + sh[1] + sh[2] + sh[3];
    __syncthreads();
    return before + inc - v;
}

__global__ void __launch_bounds__(256) k_scan_sums(const uint64_t* __restrict__ in, uint64_t m,
                                                  uint64_t* __restrict__ sums) {
    __shared__ uint64_t sh[4];
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_PER;
    uint64_t v = 0;
    SA_UNROLL
    for (uint32_t x = 0; x < SCAN_PER / 256; ++x) {
        const uint64_t i = base + x * 256 + threadIdx.x;
        if (i < m) v += in[i];
    }
    const uint64_t t = block_sum_256(v, sh);
    if (threadIdx.x == 0) sums[blockIdx.x] = t;
}

__global__ void __launch_bounds__(256) k_scan_top(uint64_t* __restrict__ sums, uint32_t nb, uint64_t* __restrict__ total,
                                                 uint64_t cap, uint32_t* __restrict__ flag) {
    __shared__ uint64_t sh[4];
    uint64_t carry = 0;
    for (uint32_t b0 = 0; b0 < nb; b0 += 256) {
        const uint32_t i = b0 + threadIdx.x;
        const uint64_t v = i < nb ? sums[i] : 0;
        uint64_t tot = 0;
        const uint64_t e = block_excl_256(v, sh, &tot);
        if (i < nb) sums[i] = carry + e;
        carry += tot;
    }
    if (threadIdx.x == 0) {
        *total = carry;
        if (flag && carry > cap) *flag = 1;
    }
}

__global__ void __launch_bounds__(256) k_scan_apply(const uint64_t* __restrict__ in, uint64_t m,
                                                   const uint64_t* __restrict__ sums, uint64_t* __restrict__ out) {
    __shared__ uint64_t sh[4];
    constexpr uint32_t PER = SCAN_PER / 256;
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_PER + (uint64_t)threadIdx.x * PER;
    uint64_t v[PER], s = 0;
    SA_UNROLL
    for (uint32_t x = 0; x < PER; ++x) {
        v[x] = base + x < m ? in[base + x] : 0;
        s += v[x];
    }
    uint64_t tot = 0;
    uint64_t e = sums[blockIdx.x] + block_excl_256(s, sh, &tot);
    SA_UNROLL
    for (uint32_t x = 0; x < PER; ++x) {
        if (base + x < m) out[base + x] = e;
        e += v[x];
    }
}

// Offsets of the padding reads: reads n..cap-1 are empty (off[i] = off[n]).
__global__ void __launch_bounds__(256) k_pad_off(uint64_t* __restrict__ off, uint32_t n, uint32_t cap) {
    const uint32_t i = n + 1 + blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= cap) off[i] = off[n];
}

// The speculative extend_seed requests of read i (SpecParams), handed to emit(nam,
// flipped, consistent) in order; returns how many.
template <typename Emit>
__device__ inline uint32_t spec_select(const SpecParams& sp, const NamRec* __restrict__ nams, uint64_t cnt, Emit emit) {
    if (cnt == 0) return 0;
    const float top = nams[0].score;
    uint64_t ntie = 1;
    while (ntie < cnt && nams[ntie].score == top) ++ntie;
    uint32_t made = 0;
    if (!sp.paired) {
        const uint64_t lim = cnt < (ntie > sp.max_tries ? ntie : sp.max_tries) ? cnt
                                                                               : (ntie > sp.max_tries ? ntie : sp.max_tries);
        for (uint64_t x = 0; x < lim && made < sp.max_per_read; ++x) {
            if (nams[x].score / top < sp.dropoff) break;
            if (nams[x].flags & 2u) {
                emit(nams[x], false, true);
                ++made;
            }
        }
    } else {
        // reverse_nam_if_needed outcome 1 (consistent once flipped): the
        // proper-pair path extends the NAM as it is, the other path flipped
        for (uint64_t x = 0; x < ntie; ++x) {
            const uint32_t rev = (nams[x].flags >> 2) & 3u;
            if (made + (rev == 1 ? 2u : 1u) > sp.max_per_read) break;
            emit(nams[x], false, rev == 0);
            ++made;
            if (rev == 1) {
                emit(nams[x], true, true);
                ++made;
            }
        }
    }
    return made;
}

// Per read: speculative request count and their CIGAR slots (2 L + 8 each).
__global__ void __launch_bounds__(256) k_spec_count(const NamRec* __restrict__ nams, const uint64_t* __restrict__ nam_off,
                                                   const uint64_t* __restrict__ nam_cnt, const uint64_t* __restrict__ off,
                                                   uint32_t n, SpecParams sp, uint64_t* __restrict__ nreq,
                                                   uint64_t* __restrict__ nslots) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t made = spec_select(sp, nams + nam_off[i], nam_cnt[i], [](const NamRec&, bool, bool) {});
    const uint64_t L = off[i + 1] - off[i];
    nreq[i] = made;
    nslots[i] = made * (2 * L + 8);
}

// Writes the requests of every read whose requests and CIGAR slots fit the
// buffers (a prefix of the reads, since the offsets grow with i) and raises
// *n_req to their end.
__global__ void __launch_bounds__(256) k_spec_write(const NamRec* __restrict__ nams, const uint64_t* __restrict__ nam_off,
                                                   const uint64_t* __restrict__ nam_cnt, const uint64_t* __restrict__ off,
                                                   uint32_t n, SpecParams sp, const uint64_t* __restrict__ roff,
                                                   const uint64_t* __restrict__ soff, uint64_t req_cap,
                                                   uint64_t slot_cap, ExtRequest* __restrict__ req,
                                                   uint32_t* __restrict__ n_req) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t L = (uint32_t)(off[i + 1] - off[i]);
    const uint32_t made = spec_select(sp, nams + nam_off[i], nam_cnt[i], [](const NamRec&, bool, bool) {});
    if (!made || roff[i] + made > req_cap || soff[i] + made * (2ull * L + 8) > slot_cap) return;
    uint64_t at = roff[i], slot = soff[i];
    spec_select(sp, nams + nam_off[i], nam_cnt[i], [&](const NamRec& m, bool flip, bool cons) {
        ExtRequest q;
        q.kind = EXT_SEED;
        q.read_id = i;
        const bool rc = ((m.flags & 1u) != 0) != flip;
        q.flags = (sp.paired || sp.use_ssw ? 1u : 0u) | (cons ? 2u : 0u) | (rc ? 4u : 0u);
        q.ref_id = m.ref_id;
        q.ref_start = m.ref_start;
        q.ref_end = m.ref_end;
        q.query_start = flip ? L - m.query_end : m.query_start;
        q.query_end = flip ? L - m.query_start : m.query_end;
        q.anchor_off = m.anchor_off;
        q.n_anchors = m.n_anchors;
        q.win_start = q.win_end = 0;
        q.cigar_cap = 2 * L + 8;
        q.cigar_off = slot;
        req[at++] = q;
        slot += q.cigar_cap;
    });
    atomicMax(n_req, (uint32_t)at);
}

// CIGAR op counts of the results [0, *n) and zeros up to cap (the scan's input).
__global__ void __launch_bounds__(256) k_ncigar_dev(const ExtResult* __restrict__ r, const uint32_t* __restrict__ n,
                                                   uint64_t* __restrict__ out, uint32_t cap) {
    const uint32_t m = *n;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cap; i += gridDim.x * blockDim.x)
        out[i] = i < m && r[i].found ? r[i].n_cigar : 0;
}

// The CIGARs of results [0, *n) packed at their scanned offsets.
__global__ void __launch_bounds__(256) k_compact_cigars_dev(const ExtRequest* __restrict__ req,
                                                           const ExtResult* __restrict__ res,
                                                           const uint32_t* __restrict__ slots,
                                                           const uint64_t* __restrict__ coff,
                                                           uint32_t* __restrict__ out, const uint32_t* __restrict__ n) {
    const uint32_t m = *n;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
        if (!res[i].found) continue;
        const uint32_t* s = slots + req[i].cigar_off;
        uint32_t* d = out + coff[i];
        for (uint32_t x = 0; x < res[i].n_cigar; ++x) d[x] = s[x];
    }
}

}  // namespace sa
