// Oracle: read-level decisions (src/mapper.rs, src/maponly.rs), SAM/PAF text.
// TEST INFRASTRUCTURE ONLY (see oracle.hpp).
#include "oracle.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <set>

namespace orc {

namespace {

const uint16_t PAIRED = 1, PROPER_PAIR = 2, UNMAP = 4, MUNMAP = 8, REVERSE = 0x10, MREVERSE = 0x20, READ1 = 0x40,
               READ2 = 0x80, SECONDARY = 0x100;  // src/io/sam.rs:7-18

// src/revcomp.rs:1-29
std::string revcomp(const std::string& s) {
    std::string r(s.size(), 'N');
    for (size_t i = 0; i < s.size(); ++i) {
        char c = s[s.size() - 1 - i], o;
        switch (c) {
            case 'A': case 'a': o = 'T'; break;
            case 'C': case 'c': o = 'G'; break;
            case 'G': case 'g': o = 'C'; break;
            case 'T': case 't': case 'U': case 'u': o = 'A'; break;
            default: o = 'N';
        }
        r[i] = o;
    }
    return r;
}

struct Read {
    const std::string* seq;
    std::string rc;
    explicit Read(const std::string& s) : seq(&s), rc(revcomp(s)) {}
    size_t len() const { return seq->size(); }
};

// src/mapper.rs:58-77
struct Alignment {
    size_t ref_id = 0, ref_start = 0;
    Cigar cigar;
    size_t edit_distance = 0, soft_clip_left = 0, soft_clip_right = 0;
    uint32_t score = 0;
    size_t length = 0;
    bool is_revcomp = false, gapped = false;
    size_t global_ed() const { return edit_distance + soft_clip_left + soft_clip_right; }
};

// src/details.rs:48-93
struct Details {
    NamDetails nam;
    size_t inconsistent_nams = 0, mate_rescue = 0, tried_alignment = 0, gapped = 0, best_alignments = 0;
};

struct SamRecord {
    std::string qname;
    uint16_t flags = 0;
    const std::string* rname = nullptr;
    long pos = -1;
    uint8_t mapq = 0;
    bool has_cigar = false;
    Cigar cigar;
    std::string rnext;  // empty = none
    long mate_pos = -1;
    bool has_tlen = false;
    int32_t tlen = 0;
    bool has_seq = false, has_qual = false;
    std::string seq, qual;
    bool has_nm = false;
    uint32_t nm = 0, as = 0;
    bool has_details = false;
    Details det;
    const std::string* rg = nullptr;
    const std::string* extra = nullptr;
    bool is_mapped() const { return (flags & UNMAP) == 0; }
};

// src/io/sam.rs:53-138
void write_sam(const SamRecord& r, std::string& out) {
    out += r.qname; out += '\t';
    out += std::to_string(r.flags); out += '\t';
    out += r.rname ? *r.rname : std::string("*"); out += '\t';
    out += std::to_string(r.pos >= 0 ? r.pos + 1 : 0); out += '\t';
    out += std::to_string((unsigned)r.mapq); out += '\t';
    out += r.has_cigar ? r.cigar.str() : std::string("*"); out += '\t';
    out += r.rnext.empty() ? std::string("*") : r.rnext; out += '\t';
    out += std::to_string(r.mate_pos >= 0 ? r.mate_pos + 1 : 0); out += '\t';
    out += std::to_string(r.has_tlen ? r.tlen : 0); out += '\t';
    bool sec = (r.flags & SECONDARY) != 0;
    out += (r.has_seq && !sec && !r.seq.empty()) ? r.seq : std::string("*"); out += '\t';
    out += (r.has_qual && !sec && !r.qual.empty()) ? r.qual : std::string("*");
    if (r.has_nm) { out += "\tNM:i:"; out += std::to_string(r.nm); out += "\tAS:i:"; out += std::to_string(r.as); }
    if (r.has_details) {
        out += "\tna:i:" + std::to_string(r.det.nam.n_nams) + "\tal:i:" + std::to_string(r.det.tried_alignment) +
               "\tga:i:" + std::to_string(r.det.gapped) + "\tX0:i:" + std::to_string(r.det.best_alignments);
        if (r.flags & PAIRED) out += "\tmr:i:" + std::to_string(r.det.mate_rescue);
    }
    if (r.rg) { out += "\tRG:Z:"; out += *r.rg; }
    if (r.extra) { out += '\t'; out += *r.extra; }
    out += '\n';
}

// src/math.rs:1-7 (f32, host libm)
float normal_pdf(float x, float mu, float sigma) {
    float a = (x - mu) / sigma;
    return (float)0.3989422804014327 / sigma * std::exp(-0.5f * a * a);
}

// src/insertsize.rs:3-50
struct InsertSize {
    size_t sample_size = 0;
    float mu = 300.0f, sigma = 100.0f, variance = 10000.0f, sse = 10000.0f;
    void update(size_t insert_size) {
        if (insert_size >= 2000) return;
        if (sample_size > 0) {
            float x = (float)insert_size;
            float e = x - mu;
            mu += e / (float)sample_size;
            sse += e * (x - mu);
            if (sample_size > 1) variance = sse / ((float)sample_size - 1.0f);
            else variance = sse;
            sigma = std::sqrt(variance);
        }
        sample_size += 1;
    }
};

struct Ctx {
    const Mapper& m;
    Rng& rng;
    InsertSize& isize;
    size_t k() const { return (size_t)m.index->k(); }
    const PackedSeq& ref(size_t id) const { return (*m.refs)[id].seq; }
};

// src/nam.rs:49-66
bool is_consistent(const Nam& nam, const Read& read, const Ctx& c) {
    size_t k = c.k();
    const PackedSeq& r = c.ref(nam.ref_id);
    std::string rs = r.decode(nam.ref_start, nam.ref_start + k), re = r.decode(nam.ref_end - k, nam.ref_end);
    const std::string& seq = nam.is_revcomp ? read.rc : *read.seq;
    return seq.compare(nam.query_start, k, rs) == 0 && seq.compare(nam.query_end - k, k, re) == 0;
}

// src/nam.rs:94-135
bool reverse_nam_if_needed(Nam& nam, const Read& read, const Ctx& c) {
    size_t k = c.k();
    const PackedSeq& r = c.ref(nam.ref_id);
    std::string rs = r.decode(nam.ref_start, nam.ref_start + k), re = r.decode(nam.ref_end - k, nam.ref_end);
    const std::string& seq = nam.is_revcomp ? read.rc : *read.seq;
    const std::string& seq_rc = nam.is_revcomp ? *read.seq : read.rc;
    if (seq.compare(nam.query_start, k, rs) == 0 && seq.compare(nam.query_end - k, k, re) == 0) return true;
    size_t L = read.len();
    size_t qs = L - nam.query_end, qe = L - nam.query_start;
    if (seq_rc.compare(qs, k, rs) == 0 && seq_rc.compare(qe - k, k, re) == 0) {
        nam.is_revcomp = !nam.is_revcomp;
        nam.query_start = qs;
        nam.query_end = qe;
        return true;
    }
    return false;
}

// src/mapper.rs:532-620
std::optional<Alignment> extend_seed(Nam& nam, const Read& read, bool consistent, bool use_ssw, const Ctx& c) {
    StageTimer st(stage_times().extend);
    const std::string& query = nam.is_revcomp ? read.rc : *read.seq;
    const PackedSeq& refseq = c.ref(nam.ref_id);
    const Scores& sc = c.m.sc;
    size_t qlen = query.size();
    size_t prs = nam.ref_start > nam.query_start ? nam.ref_start - nam.query_start : 0;
    size_t pre = std::min(nam.ref_end + qlen - nam.query_end, refseq.len);
    AlignmentInfo info;
    size_t result_ref_start = 0;
    bool gapped = true;
    if (prs + qlen == pre && consistent) {
        std::string seg = refseq.decode(prs, pre);
        size_t hd = 0;
        for (size_t i = 0; i < qlen; ++i) hd += query[i] != seg[i];
        if ((float)hd / (float)qlen < 0.05f) {
            info = *hamming_align((const uint8_t*)query.data(), (const uint8_t*)seg.data(), qlen, sc);
            result_ref_start = prs + info.ref_start;
            gapped = false;
        }
    }
    if (gapped) {
        size_t padding = read.len() / 10;
        if (use_ssw) {
            size_t rs = prs > padding ? prs - padding : 0;
            size_t re = std::min(pre + padding, refseq.len);
            std::string seg = refseq.decode(rs, re);
            auto a = aligner_align((const uint8_t*)query.data(), qlen, (const uint8_t*)seg.data(), seg.size(), sc);
            if (!a) return std::nullopt;
            info = *a;
            result_ref_start = rs + info.ref_start;
        } else {
            remove_spurious_anchors(nam.anchors);
            size_t ds = nam.ref_start > qlen + padding ? nam.ref_start - (qlen + padding) : 0;
            size_t de = std::min(nam.ref_end + qlen + padding, refseq.len);
            std::string dec = refseq.decode(ds, de);
            std::vector<Anchor> adj = nam.anchors;
            for (Anchor& a : adj) a.ref_start -= ds;
            auto a = extend_piecewise((const uint8_t*)query.data(), qlen, (const uint8_t*)dec.data(), dec.size(), adj,
                                      padding, sc, c.k(), c.m.xdrop);
            if (!a) return std::nullopt;
            info = *a;
            result_ref_start = info.ref_start + ds;
        }
    }
    Alignment al;
    al.cigar = info.cigar;
    al.edit_distance = info.edit_distance;
    al.soft_clip_left = info.query_start;
    al.soft_clip_right = qlen - info.query_end;
    al.score = info.score;
    al.ref_start = result_ref_start;
    al.length = info.ref_end - info.ref_start;
    al.is_revcomp = nam.is_revcomp;
    al.ref_id = nam.ref_id;
    al.gapped = gapped;
    return al;
}

// --- SAM record construction (src/mapper.rs:115-345)
SamRecord make_unmapped(const Record& rec, const Details& d, const Ctx& c) {
    SamRecord r;
    r.qname = rec.name;
    r.flags = UNMAP;
    r.has_seq = true;
    r.seq = rec.seq;
    r.has_qual = rec.has_qual;
    r.qual = rec.qual;
    if (c.m.so.details) { r.has_details = true; r.det = d; }
    if (c.m.so.has_rg) r.rg = &c.m.so.rg_id;
    if (c.m.so.fastq_comments && rec.has_comment) r.extra = &rec.comment;
    return r;
}

SamRecord make_mapped(const Alignment& a, const Record& rec, uint8_t mapq, bool secondary, const Details& d,
                      const Ctx& c) {
    SamRecord r;
    r.qname = rec.name;
    if (a.is_revcomp) r.flags |= REVERSE;
    if (secondary) { mapq = 0; r.flags |= SECONDARY; }
    r.has_seq = true;
    r.has_qual = rec.has_qual;
    if (a.is_revcomp) {
        r.seq = revcomp(rec.seq);
        if (rec.has_qual) r.qual.assign(rec.qual.rbegin(), rec.qual.rend());
    } else {
        r.seq = rec.seq;
        r.qual = rec.qual;
    }
    Cigar cg;
    cg.push(4, a.soft_clip_left);
    cg.extend(a.cigar);
    cg.push(4, a.soft_clip_right);
    r.has_cigar = true;
    r.cigar = c.m.so.eqx ? cg : cg.with_m();
    r.rname = &(*c.m.refs)[a.ref_id].name;
    r.pos = (long)a.ref_start;
    r.mapq = mapq;
    r.has_nm = true;
    r.nm = (uint32_t)a.edit_distance;
    r.as = a.score;
    if (c.m.so.details) { r.has_details = true; r.det = d; }
    if (c.m.so.has_rg) r.rg = &c.m.so.rg_id;
    if (c.m.so.fastq_comments && rec.has_comment) r.extra = &rec.comment;
    return r;
}

SamRecord make_record(const Alignment* a, const Record& rec, uint8_t mapq, bool secondary, const Details& d,
                      const Ctx& c) {
    return a ? make_mapped(*a, rec, mapq, secondary, d, c) : make_unmapped(rec, d, c);
}

// src/mapper.rs:245-345
std::vector<SamRecord> make_paired_records(const Alignment* al[2], const Record* recs[2], const uint8_t mapq[2],
                                           const Details det[2], bool secondary, bool proper, const Ctx& c) {
    std::vector<SamRecord> v;
    v.push_back(make_record(al[0], *recs[0], mapq[0], secondary, det[0], c));
    v.push_back(make_record(al[1], *recs[1], mapq[1], secondary, det[1], c));
    v[0].flags |= READ1;
    v[1].flags |= READ2;
    for (int i = 0; i < 2; ++i) {
        v[i].flags |= PAIRED;
        if (proper) v[i].flags |= PROPER_PAIR;
        const Alignment* mate = al[1 - i];
        if (mate) {
            if (mate->is_revcomp) v[i].flags |= MREVERSE;
            v[i].mate_pos = (long)mate->ref_start;
            v[i].rnext = (*c.m.refs)[mate->ref_id].name;
            if (al[i]) {
                const Alignment* th = al[i];
                if (th->ref_id == mate->ref_id) {
                    v[i].rnext = "=";
                    long tl;
                    if (mate->ref_start > th->ref_start) tl = (long)(mate->ref_start - th->ref_start + mate->length);
                    else if (th->ref_start > mate->ref_start) tl = -(long)(th->ref_start - mate->ref_start + th->length);
                    else if (i == 0) tl = -(long)mate->length;
                    else tl = (long)th->length;
                    v[i].has_tlen = true;
                    v[i].tlen = (int32_t)tl;
                }
            } else {
                v[i].rnext = "=";
                v[i].rname = &(*c.m.refs)[mate->ref_id].name;
                v[i].pos = (long)mate->ref_start;
            }
        } else {
            v[i].flags |= MUNMAP;
            if (al[i]) { v[i].rnext = "="; v[i].mate_pos = (long)al[i]->ref_start; }
        }
    }
    if (secondary) {
        if (!v[1].is_mapped()) v.pop_back();
        else if (!v[0].is_mapped()) { v[0] = std::move(v.back()); v.pop_back(); }
    }
    return v;
}

// src/mapper.rs:1302-1313
uint8_t mapping_quality(const std::vector<Nam>& nams) {
    if (nams.size() <= 1) return 60;
    float s1 = nams[0].score, s2 = nams[1].score;
    float mm = (float)std::min<size_t>(nams[0].anchors.size(), 10) / 10.0f;
    float u = 40.0f * (1.0f - s2 / s1) * mm * std::log(s1);
    u = std::min(u, 60.0f);
    if (!(u > 0.0f)) return 0;  // Rust `as u8` saturates (NaN -> 0)
    return (uint8_t)u;
}

// ----------------------------------------------------------------- SE ---
// src/mapper.rs:349-528
void align_single_end(const Record& rec, const Ctx& c, std::vector<SamRecord>& out) {
    NamDetails nd;
    std::vector<Nam> nams = get_nams_by_chaining((const uint8_t*)rec.seq.data(), rec.seq.size(), *c.m.index, c.m.cp,
                                                 c.m.mp.rescue_distance, c.m.mp.mcs, nd);
    sort_nams(nams, c.rng);
    Details det;
    det.nam = nd;
    if (nams.empty()) { out.push_back(make_unmapped(rec, det, c)); return; }
    std::vector<Alignment> alignments;
    float nam_max_score = nams[0].score;
    size_t best_ed = SIZE_MAX;
    uint32_t best_score = 0, second = 0;
    size_t best_index = 0;
    uint32_t n_best = 0;
    std::optional<Alignment> best;
    Read read(rec.seq);
    for (size_t tries = 0; tries < nams.size() && tries < c.m.mp.max_tries; ++tries) {
        Nam& nam = nams[tries];
        float dropoff = nam.score / nam_max_score;
        if ((tries > 1 && best_ed == 0) || dropoff < c.m.mp.dropoff_threshold) break;
        bool cons = is_consistent(nam, read, c);
        if (!cons) { det.inconsistent_nams++; continue; }
        auto al = extend_seed(nam, read, cons, c.m.mp.use_ssw, c);
        if (!al) continue;
        det.tried_alignment++;
        det.gapped += al->gapped ? 1 : 0;
        if (al->score >= best_score) {
            second = best_score;
            bool upd = false;
            if (al->score > best_score) { n_best = 1; upd = true; }
            else {
                n_best++;
                if (c.rng.gen_mod_u32(n_best) == 0) upd = true;
            }
            if (upd) {
                best_score = al->score;
                best = *al;
                best_index = alignments.size();
                if (c.m.mp.max_secondary == 0) best_ed = al->global_ed();
            }
        } else if (al->score > second) {
            second = al->score;
        }
        if (c.m.mp.max_secondary > 0) alignments.push_back(*al);
    }
    if (!best) { out.push_back(make_unmapped(rec, det, c)); return; }
    uint32_t num = 60 * (best_score - second);
    uint8_t mapq = (uint8_t)((num + best_score - 1) / best_score);
    out.push_back(make_mapped(*best, rec, mapq, false, det, c));
    if (c.m.mp.max_secondary > 0) {
        alignments[best_index] = alignments.back();
        alignments.pop_back();
        std::stable_sort(alignments.begin(), alignments.end(),
                         [](const Alignment& a, const Alignment& b) { return a.score > b.score; });
        for (size_t i = 0; i < alignments.size() && i < c.m.mp.max_secondary; ++i) {
            const Alignment& a = alignments[i];
            uint32_t diff = a.score > best_score ? a.score - best_score : 0;
            if (diff > (uint32_t)(2 * c.m.sc.mismatch + c.m.sc.gap_open)) break;
            out.push_back(make_mapped(a, rec, mapq, true, det, c));
        }
    }
}

// ----------------------------------------------------------------- PE ---
struct ScoredPair {
    double score;
    std::optional<Alignment> a1, a2;
};

// src/mapper.rs:1156-1170
bool has_shared_substring(const std::string& rs, const std::string& ref, size_t k) {
    size_t sub = 2 * k / 3, step = k / 3;
    size_t lim = rs.size() > sub ? rs.size() - sub : 0;
    for (size_t i = 0; i < lim; i += step)
        if (ref.find(rs.data() + i, 0, sub) != std::string::npos) return true;
    return false;
}

// src/mapper.rs:1094-1153
std::optional<Alignment> rescue_align(const Nam& mate, const Read& read, float mu, float sigma, const Ctx& c) {
    StageTimer st(stage_times().extend);
    size_t L = read.len();
    size_t k = c.k();
    const std::string* rtmp;
    size_t rs, re;
    float w = mu + 5.0f * sigma;
    size_t wi = w > 0 ? (size_t)w : 0;
    if (mate.is_revcomp) {
        rtmp = read.seq;
        size_t p = mate.projected_ref_start();
        rs = p > wi ? p - wi : 0;
        re = p + L / 2;
    } else {
        rtmp = &read.rc;
        size_t b = mate.ref_end + L - mate.query_end;
        rs = b > L / 2 ? b - L / 2 : 0;
        re = b + wi;
    }
    const PackedSeq& ref = c.ref(mate.ref_id);
    rs = std::min(rs, ref.len);
    re = std::min(re, ref.len);
    if (re < rs + k) return std::nullopt;
    std::string seg = ref.decode(rs, re);
    if (!has_shared_substring(*rtmp, seg, k)) return std::nullopt;
    auto info = aligner_align((const uint8_t*)rtmp->data(), rtmp->size(), (const uint8_t*)seg.data(), seg.size(), c.m.sc);
    if (!info) return std::nullopt;
    Alignment a;
    a.ref_id = mate.ref_id;
    a.ref_start = rs + info->ref_start;
    a.edit_distance = info->edit_distance;
    a.soft_clip_left = info->query_start;
    a.soft_clip_right = L - info->query_end;
    a.score = info->score;
    a.length = info->ref_end - info->ref_start;
    a.cigar = info->cigar;
    a.is_revcomp = !mate.is_revcomp;
    a.gapped = true;
    return a;
}

// src/mapper.rs:1018-1091
std::vector<ScoredPair> rescue_read(const Read& read2, const Read& read1, std::vector<Nam>& nams1, Details det[2],
                                    float mu, float sigma, const Ctx& c) {
    float max1 = nams1[0].score;
    std::vector<Alignment> al1;
    std::vector<std::optional<Alignment>> al2;
    for (size_t t = 0; t < nams1.size() && t < c.m.mp.max_tries; ++t) {
        Nam& nam = nams1[t];
        if (nam.score / max1 < c.m.mp.dropoff_threshold) break;
        bool cons = reverse_nam_if_needed(nam, read1, c);
        det[0].inconsistent_nams += !cons;
        auto a = extend_seed(nam, read1, cons, true, c);
        if (a) {
            det[0].gapped += a->gapped;
            al1.push_back(*a);
            det[0].tried_alignment++;
            auto b = rescue_align(nam, read2, mu, sigma, c);
            if (b) det[1].mate_rescue++;
            al2.push_back(b);
        }
    }
    std::vector<ScoredPair> pairs;
    for (const Alignment& a1 : al1) {
        for (const auto& a2 : al2) {
            if (a2) {
                size_t dist = a1.ref_start > a2->ref_start ? a1.ref_start - a2->ref_start : a2->ref_start - a1.ref_start;
                float score = (float)(a1.score + a2->score);
                if ((a1.is_revcomp ^ a2->is_revcomp) && (float)dist < mu + 4.0f * sigma)
                    score += std::log(normal_pdf((float)dist, mu, sigma));
                else
                    score -= 10.0f;
                pairs.push_back({(double)score, a1, *a2});
            } else {
                pairs.push_back({(double)a1.score - 10.0, a1, std::nullopt});
            }
        }
    }
    return pairs;
}

// src/mapper.rs:1184-1198
bool is_proper_pair(const Alignment* a1, const Alignment* a2, float mu, float sigma) {
    if (!a1 || !a2) return false;
    long dist = (long)a2->ref_start - (long)a1->ref_start;
    bool same = a1->ref_id == a2->ref_id;
    bool r1r2 = !a1->is_revcomp && a2->is_revcomp && dist >= 0;
    bool r2r1 = !a2->is_revcomp && a1->is_revcomp && dist <= 0;
    float lim = mu + sigma * 6.0f;
    size_t limi = lim > 0 ? (size_t)lim : 0;
    return same && (size_t)std::labs(dist) <= limi && (r1r2 || r2r1);
}

// src/mapper.rs:1200-1220
bool is_proper_nam_pair(const Nam& n1, const Nam& n2, float mu, float sigma) {
    if (n1.ref_id != n2.ref_id || n1.is_revcomp == n2.is_revcomp) return false;
    size_t a = n1.projected_ref_start(), b = n2.projected_ref_start();
    bool r1r2 = n2.is_revcomp && a <= b && (float)(b - a) < mu + 10.0f * sigma;
    bool r2r1 = n1.is_revcomp && b <= a && (float)(a - b) < mu + 10.0f * sigma;
    return r1r2 || r2r1;
}

struct NamPair { float score; const Nam* n1; const Nam* n2; };

// src/mapper.rs:1224-1299
std::vector<NamPair> best_scoring_nam_pairs(const std::vector<Nam>& n1, const std::vector<Nam>& n2, float mu, float sigma) {
    std::vector<NamPair> pairs;
    if (n1.empty() && n2.empty()) return pairs;
    std::set<size_t> added1, added2;
    size_t best_joint = 0;
    for (size_t a = 0; a < n1.size() && a < 1000; ++a) {
        for (size_t b = 0; b < n2.size() && b < 1000; ++b) {
            size_t joint = n1[a].anchors.size() + n2[b].anchors.size();
            if (joint < best_joint / 2) break;
            if (is_proper_nam_pair(n1[a], n2[b], mu, sigma)) {
                pairs.push_back({n1[a].score + n2[b].score, &n1[a], &n2[b]});
                added1.insert(n1[a].nam_id);
                added2.insert(n2[b].nam_id);
                best_joint = std::max(joint, best_joint);
            }
        }
    }
    if (!n1.empty()) {
        size_t bh = best_joint > 0 ? best_joint : n1[0].anchors.size();
        for (const Nam& n : n1) {
            if (n.anchors.size() < bh / 2) break;
            if (added1.count(n.nam_id)) continue;
            pairs.push_back({n.score, &n, nullptr});
        }
    }
    if (!n2.empty()) {
        size_t bh = best_joint > 0 ? best_joint : n2[0].anchors.size();
        for (const Nam& n : n2) {
            if (n.anchors.size() < bh / 2) break;
            if (added2.count(n.nam_id)) continue;
            pairs.push_back({n.score, nullptr, &n});
        }
    }
    std::stable_sort(pairs.begin(), pairs.end(), [](const NamPair& x, const NamPair& y) { return x.score > y.score; });
    return pairs;
}

// src/mapper.rs:1456-1470
float top_dropoff(const std::vector<Nam>& nams) {
    const Nam& n = nams[0];
    if (n.anchors.size() <= 2) return 1.0f;
    if (nams.size() > 1) return (float)nams[1].anchors.size() / (float)n.anchors.size();
    return 0.0f;
}

struct PairedResult {
    bool proper = false;
    Alignment p1, p2;
    std::vector<ScoredPair> pairs;
};

// src/mapper.rs:778-1014
PairedResult extend_paired_seeds(std::vector<Nam> nams[2], const Read& r1, const Read& r2, Details det[2],
                                 const InsertSize& isz, const Ctx& c) {
    PairedResult res;
    float mu = isz.mu, sigma = isz.sigma;
    float dropoff = c.m.mp.dropoff_threshold;
    if (nams[0].empty() && nams[1].empty()) return res;
    if (!nams[0].empty() && nams[1].empty()) {
        res.pairs = rescue_read(r2, r1, nams[0], det, mu, sigma, c);
        return res;
    }
    if (nams[0].empty() && !nams[1].empty()) {
        Details sw[2] = {det[1], det[0]};
        res.pairs = rescue_read(r1, r2, nams[1], sw, mu, sigma, c);
        det[0] = sw[1];
        det[1] = sw[0];
        for (auto& p : res.pairs) std::swap(p.a1, p.a2);
        return res;
    }
    if (top_dropoff(nams[0]) < dropoff && top_dropoff(nams[1]) < dropoff &&
        is_proper_nam_pair(nams[0][0], nams[1][0], mu, sigma)) {
        Nam m1 = nams[0][0], m2 = nams[1][0];
        bool c1 = is_consistent(m1, r1, c);
        det[0].inconsistent_nams += !c1;
        bool c2 = is_consistent(m2, r2, c);
        det[1].inconsistent_nams += !c2;
        auto a1 = extend_seed(m1, r1, c1, true, c);
        auto a2 = extend_seed(m2, r2, c2, true, c);
        if (a1 && a2) {
            det[0].tried_alignment++;
            det[0].gapped += a1->gapped;
            det[1].tried_alignment++;
            det[1].gapped += a2->gapped;
            res.proper = true;
            res.p1 = *a1;
            res.p2 = *a2;
            return res;
        }
    }
    const Read* reads[2] = {&r1, &r2};
    std::map<size_t, std::optional<Alignment>> cache[2];
    std::optional<Alignment> indv[2];
    for (int i = 0; i < 2; ++i) {
        bool cons = reverse_nam_if_needed(nams[i][0], *reads[i], c);
        det[i].inconsistent_nams += !cons;
        indv[i] = extend_seed(nams[i][0], *reads[i], cons, true, c);
        det[i].tried_alignment++;
        det[i].gapped += indv[i] ? indv[i]->gapped : 0;
        cache[i][nams[i][0].nam_id] = indv[i];
    }
    std::vector<NamPair> np = best_scoring_nam_pairs(nams[0], nams[1], mu, sigma);
    std::vector<ScoredPair>& out = res.pairs;
    float max_score = np[0].score;
    for (const NamPair& p : np) {
        float sd = p.score / max_score;
        if (out.size() >= c.m.mp.max_tries || sd < dropoff) break;
        std::optional<Alignment> als[2];
        const Nam* ns[2] = {p.n1, p.n2};
        for (int i = 0; i < 2; ++i) {
            std::optional<Alignment> al;
            if (ns[i]) {
                Nam th = *ns[i];
                auto it = cache[i].find(th.nam_id);
                if (it == cache[i].end()) {
                    bool cons = reverse_nam_if_needed(th, *reads[i], c);
                    det[i].inconsistent_nams += !cons;
                    al = extend_seed(th, *reads[i], cons, true, c);
                    det[i].tried_alignment++;
                    det[i].gapped += al ? al->gapped : 0;
                    cache[i][th.nam_id] = al;
                } else {
                    al = it->second;
                }
            } else {
                Nam other = *ns[1 - i];
                det[1 - i].inconsistent_nams += !reverse_nam_if_needed(other, *reads[1 - i], c);
                al = rescue_align(other, *reads[i], mu, sigma, c);
                if (al) { det[i].mate_rescue++; det[i].tried_alignment++; }
            }
            if ((al ? al->score : 0) > (indv[i] ? indv[i]->score : 0)) indv[i] = al;
            als[i] = al;
        }
        if (!als[0] || !als[1]) continue;
        const Alignment& a1 = *als[0];
        const Alignment& a2 = *als[1];
        float lim = mu + 10.0f * sigma;
        size_t limi = lim > 0 ? (size_t)lim : 0;
        bool r1r2 = a2.is_revcomp && !a1.is_revcomp && a1.ref_start <= a2.ref_start && (a2.ref_start - a1.ref_start) < limi;
        bool r2r1 = a1.is_revcomp && !a2.is_revcomp && a2.ref_start <= a1.ref_start && (a1.ref_start - a2.ref_start) < limi;
        double cs;
        if (r1r2 || r2r1) {
            size_t x = a1.ref_start > a2.ref_start ? a1.ref_start - a2.ref_start : a2.ref_start - a1.ref_start;
            cs = (double)a1.score + (double)a2.score +
                 std::max(-20.0 + 0.001, (double)std::log(normal_pdf((float)x, mu, sigma)));
        } else {
            cs = (double)a1.score + (double)a2.score - 20.0;
        }
        out.push_back({cs, a1, a2});
    }
    if (indv[0] && indv[1]) out.push_back({(double)indv[0]->score + (double)indv[1]->score - 20.0, indv[0], indv[1]});
    return res;
}

// src/mapper.rs:1343-1373
void dedup_pairs(std::vector<ScoredPair>& v) {
    if (v.size() < 2) return;
    auto same = [](const std::optional<Alignment>& a, const std::optional<Alignment>& b) {
        if (!a && !b) return true;
        if (!a || !b) return false;
        return a->ref_start == b->ref_start && a->ref_id == b->ref_id;
    };
    size_t k = 0, j = 1;
    for (size_t i = 1; i < v.size(); ++i) {
        if (!same(v[k].a1, v[i].a1) || !same(v[k].a2, v[i].a2)) { v[j] = v[i]; ++j; k = i; }
    }
    v.resize(j);
}

// src/mapper.rs:1436-1453
uint8_t alignment_quality(const std::vector<ScoredPair>& v) {
    if (v.size() <= 1) return 60;
    float s1 = (float)v[0].score, s2 = (float)v[1].score;
    if (s1 > 0 && s2 > 0) {
        float d = std::min(s1 - s2, 60.0f);
        if (!(d > 0)) return 0;
        return (uint8_t)d;
    }
    if (s1 > 0 && s2 <= 0) return 60;
    return 1;
}

// src/mapper.rs:624-768
void align_paired_end(const Record& rec1, const Record& rec2, const Ctx& c, std::vector<SamRecord>& out) {
    Details det[2];
    std::vector<Nam> nams[2];
    const Record* recs[2] = {&rec1, &rec2};
    for (int i = 0; i < 2; ++i) {
        NamDetails nd;
        nams[i] = get_nams_by_chaining((const uint8_t*)recs[i]->seq.data(), recs[i]->seq.size(), *c.m.index, c.m.cp,
                                       c.m.mp.rescue_distance, c.m.mp.mcs, nd);
        sort_nams(nams[i], c.rng);
        det[i].nam = nd;
    }
    Read r1(rec1.seq), r2(rec2.seq);
    PairedResult pr = extend_paired_seeds(nams, r1, r2, det, c.isize, c);
    float mu = c.isize.mu, sigma = c.isize.sigma;
    if (pr.proper) {
        bool ps = is_proper_pair(&pr.p1, &pr.p2, mu, sigma);
        if (ps && c.isize.sample_size < 400 && pr.p1.edit_distance + pr.p2.edit_distance < 3) {
            size_t d = pr.p1.ref_start > pr.p2.ref_start ? pr.p1.ref_start - pr.p2.ref_start : pr.p2.ref_start - pr.p1.ref_start;
            c.isize.update(d);
        }
        uint8_t mq[2] = {mapping_quality(nams[0]), mapping_quality(nams[1])};
        det[0].best_alignments = 1;
        det[1].best_alignments = 1;
        const Alignment* al[2] = {&pr.p1, &pr.p2};
        for (auto& r : make_paired_records(al, recs, mq, det, false, ps, c)) out.push_back(std::move(r));
        return;
    }
    std::vector<ScoredPair>& v = pr.pairs;
    std::stable_sort(v.begin(), v.end(), [](const ScoredPair& a, const ScoredPair& b) { return a.score > b.score; });
    dedup_pairs(v);
    size_t nb = 0;
    while (nb < v.size() && v[nb].score == v[0].score) ++nb;
    det[0].best_alignments = nb;
    det[1].best_alignments = nb;
    if (nb > 1) {
        size_t ri = (size_t)c.rng.gen_mod_u64(nb);
        std::swap(v[0], v[ri]);
    }
    uint8_t mq[2];
    if (v.size() == 1 && !is_proper_pair(v[0].a1 ? &*v[0].a1 : nullptr, v[0].a2 ? &*v[0].a2 : nullptr, mu, sigma)) {
        mq[0] = mapping_quality(nams[0]);
        mq[1] = mapping_quality(nams[1]);
    } else {
        uint8_t j = alignment_quality(v);
        mq[0] = mq[1] = j;
    }
    // src/mapper.rs:1376-1434 aligned_pairs_to_sam
    if (v.empty()) {
        SamRecord a = make_unmapped(rec1, det[0], c), b = make_unmapped(rec2, det[1], c);
        a.flags |= PAIRED | READ1 | MUNMAP;
        b.flags |= PAIRED | READ2 | MUNMAP;
        out.push_back(std::move(a));
        out.push_back(std::move(b));
        return;
    }
    double dropoff = (double)(uint8_t)(2 * c.m.sc.mismatch + c.m.sc.gap_open);
    if (c.m.mp.max_secondary == 0) {
        const Alignment* al[2] = {v[0].a1 ? &*v[0].a1 : nullptr, v[0].a2 ? &*v[0].a2 : nullptr};
        bool ps = is_proper_pair(al[0], al[1], mu, sigma);
        for (auto& r : make_paired_records(al, recs, mq, det, false, ps, c)) out.push_back(std::move(r));
    } else {
        bool secondary = false;
        double smax = v[0].score;
        for (size_t i = 0; i < v.size() && i < c.m.mp.max_secondary; ++i) {
            if (smax - v[i].score < dropoff) {
                const Alignment* al[2] = {v[i].a1 ? &*v[i].a1 : nullptr, v[i].a2 ? &*v[i].a2 : nullptr};
                bool ps = is_proper_pair(al[0], al[1], mu, sigma);
                uint8_t emq[2] = {secondary ? (uint8_t)0 : mq[0], secondary ? (uint8_t)0 : mq[1]};
                for (auto& r : make_paired_records(al, recs, emq, det, secondary, ps, c)) out.push_back(std::move(r));
            } else {
                break;
            }
            secondary = true;
        }
    }
}

// -------------------------------------------------------------- maponly ---
// src/io/paf.rs:37-60 + src/maponly.rs:94-110
void write_paf(const Nam& nam, const Record& rec, int end, const std::vector<RefSeq>& refs, int mapq, std::string& out) {
    out += rec.name;
    if (end == 1) out += "/1";
    else if (end == 2) out += "/2";
    char buf[512];
    snprintf(buf, sizeof buf, "\t%zu\t%zu\t%zu\t%c\t", rec.seq.size(), nam.query_start, nam.query_end,
             nam.is_revcomp ? '-' : '+');
    out += buf;
    out += refs[nam.ref_id].name;
    snprintf(buf, sizeof buf, "\t%zu\t%zu\t%zu\t%zu\t%zu\t%d\n", refs[nam.ref_id].seq.len, nam.ref_start, nam.ref_end,
             nam.matching_bases, nam.ref_end - nam.ref_start, mapq);
    out += buf;
}

struct PafPair { const Nam* n1; const Nam* n2; double score; };

// src/maponly.rs:380-458
std::vector<std::pair<Nam, Nam>> find_pairs(const std::vector<Nam*>& fwd, const std::vector<Nam*>& rev, float mu,
                                            float sigma, bool swap, std::vector<double>& scores) {
    std::vector<std::pair<Nam, Nam>> out;
    float md = std::ceil(mu + 10.0f * sigma);
    size_t max_dist = md > 0 ? (size_t)md : 0;
    size_t rp = 0;
    bool have_last = false;
    size_t last_id = 0;
    double last_score = 0;
    for (const Nam* f : fwd) {
        while (rp < rev.size() && (rev[rp]->ref_id < f->ref_id ||
                                   (rev[rp]->ref_id == f->ref_id && rev[rp]->projected_ref_start() < f->projected_ref_start())))
            ++rp;
        if (rp == rev.size()) break;
        if (rev[rp]->ref_id > f->ref_id) continue;
        bool have = false;
        size_t best_id = 0;
        double best_score = 0;
        for (size_t i = rp; i < rev.size() && rev[i]->ref_id == f->ref_id &&
                            rev[i]->projected_ref_start() - f->projected_ref_start() <= max_dist;
             ++i) {
            const Nam* r = rev[i];
            size_t x = f->ref_start > r->ref_start ? f->ref_start - r->ref_start : r->ref_start - f->ref_start;
            double s = (double)f->score + (double)r->score +
                       std::max(0.001, (double)std::log(normal_pdf((float)x, mu, sigma) + 1.0f));
            if (!have || s > best_score) { have = true; best_id = i; best_score = s; }
        }
        if (!have) continue;
        if (have_last && last_id == best_id) {
            if (best_score <= last_score) continue;
            out.pop_back();
            scores.pop_back();
        }
        if (swap) out.push_back({*rev[best_id], *f});
        else out.push_back({*f, *rev[best_id]});
        scores.push_back(best_score);
        have_last = true;
        last_id = best_id;
        last_score = best_score;
        rp = best_id;
    }
    return out;
}

struct NamPairOwned { Nam n1, n2; double score; };

// src/maponly.rs:314-375. The in-place partition mirrors split_nams_by_orientation; tie order of
// Rust's sort_unstable is not reproducible here (parity unpinned for tied keys).
std::vector<NamPairOwned> get_nam_pairs(std::vector<Nam>& n1, std::vector<Nam>& n2, float mu, float sigma, bool both1,
                                        bool both2) {
    std::vector<NamPairOwned> res;
    if (n1.empty() || n2.empty()) return res;
    auto split = [](std::vector<Nam>& v, bool both, size_t& lo_end) {
        if (both) {
            size_t left = 0, right = v.size();
            while (left < right) {
                if (v[left].is_revcomp) { --right; std::swap(v[left], v[right]); }
                else ++left;
            }
            lo_end = left;
        } else if (v[0].is_revcomp) {
            lo_end = 0;
        } else {
            lo_end = v.size();
        }
    };
    size_t e1, e2;
    split(n1, both1, e1);
    split(n2, both2, e2);
    auto key_sort = [](std::vector<Nam>& v, size_t a, size_t b) {
        std::sort(v.begin() + a, v.begin() + b, [](const Nam& x, const Nam& y) {
            if (x.ref_id != y.ref_id) return x.ref_id < y.ref_id;
            return x.projected_ref_start() < y.projected_ref_start();
        });
    };
    auto run = [&](size_t fa, size_t fb, std::vector<Nam>& fv, size_t ra, size_t rb, std::vector<Nam>& rv, bool swap) {
        if (fa == fb || ra == rb) return;
        key_sort(fv, fa, fb);
        key_sort(rv, ra, rb);
        std::vector<Nam*> f, r;
        for (size_t i = fa; i < fb; ++i) f.push_back(&fv[i]);
        for (size_t i = ra; i < rb; ++i) r.push_back(&rv[i]);
        std::vector<double> sc;
        auto pr = find_pairs(f, r, mu, sigma, swap, sc);
        for (size_t i = 0; i < pr.size(); ++i) res.push_back({pr[i].first, pr[i].second, sc[i]});
    };
    run(0, e1, n1, e2, n2.size(), n2, false);
    run(0, e2, n2, e1, n1.size(), n1, true);
    std::stable_sort(res.begin(), res.end(), [](const NamPairOwned& a, const NamPairOwned& b) { return a.score > b.score; });
    return res;
}

}  // namespace

std::vector<SpecExtension> speculative_extensions(const Mapper& m, const std::string& seq, size_t spec) {
    NamDetails nd;
    std::vector<Nam> nams = get_nams_by_chaining((const uint8_t*)seq.data(), seq.size(), *m.index, m.cp,
                                                 m.mp.rescue_distance, m.mp.mcs, nd);
    std::stable_sort(nams.begin(), nams.end(), [](const Nam& a, const Nam& b) { return a.score > b.score; });
    Rng rng(0);
    InsertSize isize;
    Ctx c{m, rng, isize};
    Read read(seq);
    std::vector<SpecExtension> out;
    for (size_t i = 0, done = 0; i < nams.size() && done < spec; ++i) {
        Nam nam = nams[i];
        if (!is_consistent(nam, read, c)) continue;
        ++done;
        SpecExtension e;
        e.nam_index = i;
        auto al = extend_seed(nam, read, true, m.mp.use_ssw, c);
        if (al) {
            e.found = true;
            e.score = al->score;
            e.ref_start = al->ref_start;
            e.length = al->length;
            e.query_start = al->soft_clip_left;
            e.query_end = seq.size() - al->soft_clip_right;
            e.edit_distance = al->edit_distance;
            e.gapped = al->gapped;
            e.cigar = al->cigar.str();
        }
        out.push_back(e);
    }
    return out;
}

// src/main.rs:873-988
void Mapper::map_chunk(const std::vector<std::pair<Record, std::optional<Record>>>& chunk, std::string& out) {
    Rng rng(0);
    InsertSize isize;
    Ctx c{*this, rng, isize};
    for (const auto& pr : chunk) {
        const Record& r1 = pr.first;
        if (mode == 0) {
            std::vector<SamRecord> recs;
            if (pr.second) {
                align_paired_end(r1, *pr.second, c, recs);
                if (!mp.output_unmapped && !recs[0].is_mapped() && !recs[1].is_mapped()) recs.clear();
            } else {
                align_single_end(r1, c, recs);
                if (!mp.output_unmapped && !recs[0].is_mapped()) recs.clear();
            }
            for (const SamRecord& r : recs) write_sam(r, out);
        } else {
            const Index& idx = *index;
            if (!pr.second) {
                NamDetails nd;
                std::vector<Nam> nams = get_nams_by_chaining((const uint8_t*)r1.seq.data(), r1.seq.size(), idx, cp,
                                                             mp.rescue_distance, mp.mcs, nd);
                sort_nams(nams, rng);
                if (mode == 1) {
                    if (!nams.empty()) write_paf(nams[0], r1, 0, *refs, mapping_quality(nams), out);
                } else {
                    size_t nb = 0;
                    while (nb < nams.size() && nams[nb].score == nams[0].score) ++nb;
                    double w = (double)r1.seq.size() / (double)nb;
                    for (size_t i = 0; i < nb; ++i) abundances[nams[i].ref_id] += w;
                }
            } else {
                const Record& r2 = *pr.second;
                NamDetails d1, d2;
                std::vector<Nam> n1 = get_nams_by_chaining((const uint8_t*)r1.seq.data(), r1.seq.size(), idx, cp,
                                                           mp.rescue_distance, mp.mcs, d1);
                std::vector<Nam> n2 = get_nams_by_chaining((const uint8_t*)r2.seq.data(), r2.seq.size(), idx, cp,
                                                           mp.rescue_distance, mp.mcs, d2);
                if (n1.empty() && n2.empty()) continue;
                std::vector<NamPairOwned> np = get_nam_pairs(n1, n2, isize.mu, isize.sigma, d1.both_orientations,
                                                             d2.both_orientations);
                sort_nams(n1, rng);
                sort_nams(n2, rng);
                // src/maponly.rs:277-304
                double indiv = (n1.empty() ? 0.0 : (double)n1[0].score) + (n2.empty() ? 0.0 : (double)n2[0].score);
                bool pair = !np.empty() && np[0].score >= indiv / 2.0;
                if (pair && isize.sample_size < 400) {
                    const Nam& a = np[0].n1;
                    const Nam& b = np[0].n2;
                    isize.update(a.ref_start > b.ref_start ? a.ref_start - b.ref_start : b.ref_start - a.ref_start);
                }
                if (mode == 1) {
                    if (pair) {
                        write_paf(np[0].n1, r1, 1, *refs, 255, out);
                        write_paf(np[0].n2, r2, 2, *refs, 255, out);
                    } else {
                        if (!n1.empty()) write_paf(n1[0], r1, 1, *refs, 255, out);
                        if (!n2.empty()) write_paf(n2[0], r2, 2, *refs, 255, out);
                    }
                } else {
                    if (pair) {
                        size_t nb = 0;
                        while (nb < np.size() && np[nb].score == np[0].score) ++nb;
                        double w1 = (double)r1.seq.size() / (double)nb, w2 = (double)r2.seq.size() / (double)nb;
                        for (size_t i = 0; i < nb; ++i) {
                            abundances[np[i].n1.ref_id] += w1;
                            abundances[np[i].n2.ref_id] += w2;
                        }
                    } else {
                        for (int e = 0; e < 2; ++e) {
                            const std::vector<Nam>& v = e == 0 ? n1 : n2;
                            size_t L = e == 0 ? r1.seq.size() : r2.seq.size();
                            size_t nb = 0;
                            while (nb < v.size() && v[nb].score == v[0].score) ++nb;
                            double w = (double)L / (double)nb;
                            for (size_t i = 0; i < nb; ++i) abundances[v[i].ref_id] += w;
                        }
                    }
                }
            }
        }
    }
}

// src/io/sam.rs:177-201 (no @PG line: outputs are compared with --no-PG)
std::string sam_header(const std::vector<RefSeq>& refs, const SamOpts& so, const std::vector<std::string>& rg_fields) {
    std::string h = "@HD\tVN:1.6\tSO:unsorted\n";
    for (const RefSeq& r : refs) h += "@SQ\tSN:" + r.name + "\tLN:" + std::to_string(r.seq.len) + "\n";
    if (so.has_rg) {
        h += "@RG\tID:" + so.rg_id;
        for (const std::string& f : rg_fields) h += "\t" + f;
        h += "\n";
    }
    return h;
}

// src/main.rs:1001-1012
size_t estimate_read_length(const std::vector<Record>& recs) {
    size_t s = 0, n = 0;
    for (size_t i = 0; i < recs.size() && i < 500; ++i) { s += recs[i].seq.size(); ++n; }
    return n ? s / n : 0;
}

}  // namespace orc
