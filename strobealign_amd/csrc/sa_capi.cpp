// extern "C" entry points declared in include/sa_hip.h.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

#include "sa_hip.h"
#include "sa_host.hpp"
#include "sa_ingest.hpp"

struct sa_ctx {
    int device;
};

struct sa_reads {
    std::vector<sa::Record> r1, r2;
    // device-resident copy (sa_reads_upload)
    int device = -1;
    void* d_seq = nullptr;
    sa::DeviceReads dr;
    ~sa_reads() {
        if (d_seq) sa::device_free(device, d_seq);
    }
};

struct sa_index {
    int device;
    std::unique_ptr<sa::Engine> eng;
    std::vector<std::unique_ptr<sa::Engine>> slots;  // more engines sharing eng's index
    sa::PackedRef ref;
    sa::SeedParams p;
    ~sa_index() { slots.clear(); }  // before eng, whose index they share
};

struct sa_server {
    sa_index* index;
    std::vector<std::unique_ptr<sa::Engine>> engines;  // share the index's resident index
    std::unique_ptr<sa::Server> srv;
    ~sa_server() { srv.reset(); }  // workers stop before their engines go
};

namespace {
thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

template <typename F>
int guard(F f) {
    try {
        g_err.clear();
        return f();
    } catch (const std::bad_alloc&) {
        return fail(SA_ERR_INTERNAL, "out of memory");
    } catch (const std::invalid_argument& e) {
        return fail(SA_ERR_ARG, e.what());
    } catch (const std::exception& e) {
        std::string m = e.what();
        return fail(m.rfind("HIP error", 0) == 0 ? SA_ERR_DEVICE : SA_ERR_INTERNAL, m);
    }
}

sa::SeedParams from_c(const sa_seeding_params& c) {
    sa::SeedParams p{};
    p.k = c.k; p.s = c.s; p.t = c.t; p.w_min = c.w_min; p.w_max = c.w_max; p.max_dist = c.max_dist; p.q = c.q;
    p.main_mask = c.main_hash_mask; p.fwd_main_mask = c.forward_main_hash_mask;
    p.partial_pos = c.partial_orientation_pos; p.bits = c.bits; p.filter_cutoff = c.filter_cutoff;
    return p;
}

void to_c(const sa::SeedParams& p, sa_seeding_params* c) {
    c->k = p.k; c->s = p.s; c->t = p.t; c->w_min = p.w_min; c->w_max = p.w_max; c->max_dist = p.max_dist; c->q = p.q;
    c->main_hash_mask = p.main_mask; c->forward_main_hash_mask = p.fwd_main_mask;
    c->partial_orientation_pos = p.partial_pos; c->bits = p.bits; c->filter_cutoff = p.filter_cutoff;
}

sa::Options options_of(const sa_map_options* o) {
    sa::Options r;
    if (!o) return r;
    r.mode = o->mode;
    r.eqx = o->eqx; r.details = o->details; r.fastq_comments = o->fastq_comments; r.only_mapped = o->only_mapped;
    r.use_ssw = o->use_ssw;
    if (o->rg_id) { r.has_rg = true; r.rg_id = o->rg_id; }
    for (int i = 0; i < o->n_rg; ++i) r.rg_fields.push_back(o->rg[i]);
    if (!r.rg_fields.empty() && !r.has_rg) { r.has_rg = true; r.rg_id = "1"; }
    r.max_secondary = o->max_secondary; r.dropoff = o->dropoff; r.rescue_distance = o->rescue_distance;
    r.max_tries = o->max_tries; r.mcs = o->mcs; r.chunk_size = o->chunk_size;
    r.threads = o->threads > 0 ? o->threads : 1;
    if (o->batch_bases) r.batch_bases = o->batch_bases;
    if (o->slots > 0) r.slots = o->slots;
    r.scores = sa::Scores{o->match_score, o->mismatch_penalty, o->gap_open_penalty, o->gap_extend_penalty,
                          o->end_bonus, o->xdrop};
    // the reference takes u8 scores and penalties (src/aligner.rs:11-19) and
    // negates them into i8 matrix entries (src/ssw.rs:143-144)
    for (int32_t v : {o->match_score, o->mismatch_penalty, o->gap_open_penalty, o->gap_extend_penalty})
        if (v < 0 || v > 127) throw std::invalid_argument("scores and penalties must be in 0..127");
    if (o->end_bonus < 0) throw std::invalid_argument("end_bonus must be >= 0");
    if (o->lookup_lanes) {
        if (o->lookup_lanes != 1 && o->lookup_lanes != 8 && o->lookup_lanes != 16 && o->lookup_lanes != 64)
            throw std::invalid_argument("lookup_lanes must be 0, 1, 8, 16 or 64");
        r.lookup_lanes = (uint32_t)o->lookup_lanes;
    }
    if (o->seed_lanes) {
        if (o->seed_lanes != 1 && o->seed_lanes != 64) throw std::invalid_argument("seed_lanes must be 0, 1 or 64");
        r.seed_lanes = (uint32_t)o->seed_lanes;
    }
    if (o->fused != 0 && o->fused != 1) throw std::invalid_argument("fused must be 0 or 1");
    r.fused = o->fused == 1;
    return r;
}

// The index's engine plus enough engines sharing its index for `n` slots.
std::vector<sa::Engine*> engines(sa_index* ix, int n) {
    std::vector<sa::Engine*> v{ix->eng.get()};
    while ((int)ix->slots.size() + 1 < n)
        ix->slots.push_back(std::make_unique<sa::Engine>(sa::Engine::ShareIndex{}, *ix->eng));
    for (int i = 0; i + 1 < n; ++i) v.push_back(ix->slots[(size_t)i].get());
    return v;
}

char* dup_out(const std::string& s, size_t* len) {
    char* p = (char*)std::malloc(s.size() + 1);
    if (!p) throw std::bad_alloc();
    std::memcpy(p, s.data(), s.size());
    p[s.size()] = 0;
    *len = s.size();
    return p;
}
// Per-kernel totals summed over engines, as JSON text (see sa_index_kernel_times).
size_t kernel_times_json(const std::vector<const sa::Engine*>& engs, char* buf, size_t cap) {
    std::vector<std::pair<std::string, sa::Engine::KernelStat>> all;
    for (const sa::Engine* e : engs)
        for (const auto& k : e->kernel_stats()) {
            bool found = false;
            for (auto& a : all)
                if (a.first == k.first) {
                    a.second.ms += k.second.ms;
                    a.second.launches += k.second.launches;
                    a.second.items += k.second.items;
                    a.second.bytes += k.second.bytes;
                    a.second.cells += k.second.cells;
                    found = true;
                }
            if (!found) all.push_back(k);
        }
    std::string s = "{";
    bool first = true;
    for (const auto& k : all) {
        char tmp[256];
        std::snprintf(tmp, sizeof tmp, "%s\"%s\": [%.6f, %llu, %llu, %llu, %llu]", first ? "" : ", ",
                      k.first.c_str(), k.second.ms, (unsigned long long)k.second.launches,
                      (unsigned long long)k.second.items, (unsigned long long)k.second.bytes,
                      (unsigned long long)k.second.cells);
        s += tmp;
        first = false;
    }
    s += "}";
    if (buf && cap) {
        size_t n = std::min(cap - 1, s.size());
        std::memcpy(buf, s.data(), n);
        buf[n] = 0;
    }
    return s.size();
}

sa::Record make_record(const char* name, const char* seq, const char* qual) {
    sa::Record r;
    std::string h = name ? name : "";
    size_t p = h.find_first_of(" \t");
    if (p == std::string::npos) r.name = h;
    else { r.name = h.substr(0, p); r.comment = h.substr(p + 1); r.has_comment = true; }
    r.seq = seq;
    for (char& c : r.seq) c = (char)toupper((unsigned char)c);
    if (qual) r.qual = qual; else r.has_qual = false;
    return r;
}
}  // namespace

namespace sa {
// error text for sa_last_error from entry points outside this file (the SSW shim)
void set_last_error(const std::string& msg) { g_err = msg; }
}  // namespace sa

extern "C" {

size_t sa_last_error(char* buf, size_t cap) {
    if (buf && cap) {
        size_t n = std::min(cap - 1, g_err.size());
        std::memcpy(buf, g_err.data(), n);
        buf[n] = 0;
    }
    return g_err.size();
}

const char* sa_version(void) { return "0.1.0 (gfx950)"; }

int sa_ctx_create(int device, sa_ctx** out) {
    if (!out) return fail(SA_ERR_ARG, "out is NULL");
    return guard([&] {
        // validates the device by creating (and dropping) an engine on it
        sa::Engine probe(device);
        *out = new sa_ctx{device};
        return SA_OK;
    });
}

void sa_ctx_destroy(sa_ctx* ctx) { delete ctx; }

int sa_seeding_params_for_read_length(uint64_t read_length, sa_seeding_params* out) {
    if (!out) return fail(SA_ERR_ARG, "out is NULL");
    to_c(sa::seed_params_for_read_length(read_length), out);
    return SA_OK;
}

int sa_index_build(sa_ctx* ctx, const char* fasta_path, uint64_t read_length, int bits, int threads, sa_index** out) {
    if (!ctx || !fasta_path || !out) return fail(SA_ERR_ARG, "NULL argument");
    return guard([&] {
        auto ix = std::make_unique<sa_index>();
        ix->device = ctx->device;
        try {
            ix->ref = sa::read_fasta(fasta_path);
        } catch (const std::exception& e) {
            return fail(SA_ERR_IO, e.what());
        }
        sa::SeedParams sp = sa::seed_params_for_read_length(read_length);
        if (bits <= 0) bits = sa::pick_bits(ix->ref, sp);
        sa::HostIndex hi;
        sa::build_index(ix->ref, sp, bits, 0.0002, threads > 0 ? threads : 1, hi);
        ix->p = hi.p;
        ix->eng = std::make_unique<sa::Engine>(ctx->device);
        ix->eng->upload_reference(ix->ref.words.data(), ix->ref.words.size(), ix->ref.word_off.data(),
                                  ix->ref.len.data(), (uint32_t)ix->ref.names.size());
        ix->eng->upload_index(hi.entries.data(), hi.entries.size(), hi.bucket_starts.data(), hi.p);
        *out = ix.release();
        return SA_OK;
    });
}

int sa_index_build_device(sa_ctx* ctx, const char* fasta_path, uint64_t read_length, int bits, sa_index** out) {
    if (!ctx || !fasta_path || !out) return fail(SA_ERR_ARG, "NULL argument");
    return guard([&] {
        auto ix = std::make_unique<sa_index>();
        ix->device = ctx->device;
        try {
            ix->ref = sa::read_fasta(fasta_path);
        } catch (const std::exception& e) {
            return fail(SA_ERR_IO, e.what());
        }
        sa::SeedParams sp = sa::seed_params_for_read_length(read_length);
        if (bits <= 0) bits = sa::pick_bits(ix->ref, sp);
        if (bits > 32) return fail(SA_ERR_ARG, "bits must be in 1..32");
        ix->eng = std::make_unique<sa::Engine>(ctx->device);
        ix->eng->upload_reference(ix->ref.words.data(), ix->ref.words.size(), ix->ref.word_off.data(),
                                  ix->ref.len.data(), (uint32_t)ix->ref.names.size());
        ix->p = ix->eng->build_index(ix->ref.len.data(), (uint32_t)ix->ref.names.size(), sp, bits, 0.0002);
        *out = ix.release();
        return SA_OK;
    });
}

int sa_index_download(const sa_index* index, void* entries, uint64_t* bucket_starts) {
    if (!index || (!entries && index->p.n_entries) || !bucket_starts) return fail(SA_ERR_ARG, "NULL argument");
    return guard([&] {
        index->eng->download_index((sa::RefEntry*)entries, bucket_starts);
        return SA_OK;
    });
}

int sa_index_upload(sa_ctx* ctx, const sa_seeding_params* params, const void* entries, uint64_t n,
                    const uint64_t* bucket_starts, const uint64_t* packed_ref_words, uint64_t n_words,
                    const uint64_t* contig_word_offsets, const uint64_t* contig_lengths,
                    const char* const* contig_names, uint32_t n_contigs, sa_index** out) {
    if (!ctx || !params || (!entries && n) || !bucket_starts || !packed_ref_words || !contig_word_offsets ||
        !contig_lengths || !out)
        return fail(SA_ERR_ARG, "NULL argument");
    if (params->bits < 1 || params->bits > 32) return fail(SA_ERR_ARG, "bits must be in 1..32");
    return guard([&] {
        auto ix = std::make_unique<sa_index>();
        ix->device = ctx->device;
        ix->p = from_c(*params);
        ix->p.n_entries = n;
        ix->ref.words.assign(packed_ref_words, packed_ref_words + n_words);
        ix->ref.word_off.assign(contig_word_offsets, contig_word_offsets + n_contigs + 1);
        ix->ref.len.assign(contig_lengths, contig_lengths + n_contigs);
        for (uint32_t i = 0; i < n_contigs; ++i)
            ix->ref.names.push_back(contig_names ? contig_names[i] : ("contig" + std::to_string(i)));
        ix->eng = std::make_unique<sa::Engine>(ctx->device);
        ix->eng->upload_reference(packed_ref_words, n_words, contig_word_offsets, contig_lengths, n_contigs);
        ix->eng->upload_index((const sa::RefEntry*)entries, n, bucket_starts, ix->p);
        *out = ix.release();
        return SA_OK;
    });
}

int sa_index_save_sti(const sa_index* index, uint64_t read_length, const char* path) {
    if (!index || !path) return fail(SA_ERR_ARG, "NULL argument");
    return guard([&] {
        const sa::SeedParams& p = index->p;
        const uint64_t n = p.n_entries, nb = (1ull << p.bits) + 1;
        std::vector<sa::RefEntry> e(std::max<uint64_t>(n, 1));
        std::vector<uint64_t> b(nb);
        index->eng->download_index(e.data(), b.data());
        FILE* f = std::fopen(path, "wb");
        if (!f) return fail(SA_ERR_IO, std::string("cannot create ") + path);
        bool ok = true;
        auto put = [&](const void* x, size_t bytes) { ok = ok && std::fwrite(x, 1, bytes, f) == bytes; };
        auto u32 = [&](uint32_t v) { put(&v, 4); };
        auto u64 = [&](uint64_t v) { put(&v, 8); };
        put("STI\x01", 4);
        u32(7);  // STI_FILE_FORMAT_VERSION
        u64(8);  // reserved chunk: its length, then its contents
        u64(0);
        u32((uint32_t)p.filter_cutoff);
        u32((uint32_t)p.bits);
        u32(sa::profile_for_read_length(read_length));
        u32((uint32_t)p.k);
        u32((uint32_t)p.s);
        u32((uint32_t)p.w_min);
        u32((uint32_t)p.w_max);
        u32((uint32_t)p.q);
        u32((uint32_t)p.max_dist);
        u64(p.main_mask);
        u64(n);
        put(e.data(), n * sizeof(sa::RefEntry));
        u64(nb);
        put(b.data(), nb * 8);
        ok = (std::fclose(f) == 0) && ok;
        if (!ok) return fail(SA_ERR_IO, std::string("cannot write ") + path);
        return SA_OK;
    });
}

int sa_index_load_sti(sa_ctx* ctx, const char* fasta_path, const char* sti_path, uint64_t read_length, int bits,
                      sa_index** out) {
    if (!ctx || !fasta_path || !sti_path || !out) return fail(SA_ERR_ARG, "NULL argument");
    return guard([&] {
        auto ix = std::make_unique<sa_index>();
        ix->device = ctx->device;
        try {
            ix->ref = sa::read_fasta(fasta_path);
        } catch (const std::exception& e) {
            return fail(SA_ERR_IO, e.what());
        }
        sa::SeedParams sp = sa::seed_params_for_read_length(read_length);
        if (bits <= 0) bits = sa::pick_bits(ix->ref, sp);
        std::unique_ptr<FILE, int (*)(FILE*)> f(std::fopen(sti_path, "rb"), &std::fclose);
        if (!f) return fail(SA_ERR_IO, std::string("cannot open ") + sti_path);
        auto read = [&](void* dst, uint64_t bytes) {
            if (std::fread(dst, 1, bytes, f.get()) != bytes)
                throw std::runtime_error(std::string("When reading the .sti (index) file: unexpected end of ") +
                                         sti_path);
        };
        auto u32 = [&] { uint32_t v; read(&v, 4); return v; };
        auto u64 = [&] { uint64_t v; read(&v, 8); return v; };
        // src/index.rs:412-495
        char magic[4];
        read(magic, 4);
        if (std::memcmp(magic, "STI\x01", 4) != 0)
            return fail(SA_ERR_IO, "Signature at the beginning of the .sti file is incorrect.");
        const uint32_t ver = u32();
        if (ver != 7)
            return fail(SA_ERR_IO, "The .sti file format version is " + std::to_string(ver) +
                                       ", but this version of strobealign can only read files that have version 7");
        std::vector<uint8_t> reserved(u64());
        if (!reserved.empty()) read(reserved.data(), reserved.size());
        const uint32_t filter_cutoff = u32();
        const uint32_t sti_bits = u32();
        const uint32_t profile = u32();
        const uint32_t k = u32(), s = u32(), w_min = u32(), w_max = u32(), q = u32(), max_dist = u32();
        const uint64_t main_mask = u64();
        if (profile > 7) return fail(SA_ERR_IO, "The .sti (index) file uses an unknown profile");
        if ((int)sti_bits != bits || profile != sa::profile_for_read_length(read_length) || (int)k != sp.k ||
            (int)s != sp.s || (int)w_min != sp.w_min || (int)w_max != sp.w_max || q != sp.q ||
            (int)max_dist != sp.max_dist || main_mask != sp.main_mask)
            return fail(SA_ERR_ARG, "Index parameters in .sti file and those specified on command line differ");
        sp.bits = bits;
        sp.filter_cutoff = filter_cutoff;
        const uint64_t n = u64();
        ix->eng = std::make_unique<sa::Engine>(ctx->device);
        ix->eng->upload_reference(ix->ref.words.data(), ix->ref.words.size(), ix->ref.word_off.data(),
                                  ix->ref.len.data(), (uint32_t)ix->ref.names.size());
        ix->eng->upload_index_streamed(n, sp, read);
        ix->p = ix->eng->params();
        *out = ix.release();
        return SA_OK;
    });
}

void sa_index_destroy(sa_index* index) { delete index; }

int sa_index_info(const sa_index* index, sa_seeding_params* params, uint64_t* n_entries, uint32_t* n_contigs) {
    if (!index) return fail(SA_ERR_ARG, "index is NULL");
    if (params) to_c(index->p, params);
    if (n_entries) *n_entries = index->eng->n_entries();
    if (n_contigs) *n_contigs = (uint32_t)index->ref.names.size();
    return SA_OK;
}

void sa_map_options_default(sa_map_options* o) {
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->dropoff = 0.5f;
    o->rescue_distance = 100;
    o->max_tries = 20;
    o->chunk_size = 1000000;
    o->threads = 1;
    o->match_score = 2;
    o->mismatch_penalty = 8;
    o->gap_open_penalty = 12;
    o->gap_extend_penalty = 1;
    o->end_bonus = 10;
    o->xdrop = 500;
}

int sa_map_files(sa_index* index, const sa_map_options* opt, const char* reads1, const char* reads2, char** out,
                 size_t* out_len) {
    if (!index || !reads1 || !out || !out_len) return fail(SA_ERR_ARG, "NULL argument");
    return guard([&] {
        std::string err;
        std::vector<sa::Record> r1, r2;
        if (!sa::read_records(reads1, r1, err)) return fail(SA_ERR_IO, err);
        if (reads2 && !sa::read_records(reads2, r2, err)) return fail(SA_ERR_IO, err);
        sa::Options o = options_of(opt);
        std::string s;
        if (o.mode == 0) s = sa::sam_header(index->ref, o);
        sa::map_records(engines(index, o.slots), index->ref, r1, r2, o, &s, nullptr, nullptr);
        *out = dup_out(s, out_len);
        return SA_OK;
    });
}

int sa_map_records(sa_index* index, const sa_map_options* opt, uint64_t n, const char* const* names1,
                   const char* const* seqs1, const char* const* quals1, const char* const* names2,
                   const char* const* seqs2, const char* const* quals2, char** out, size_t* out_len) {
    if (!index || !names1 || !seqs1 || !out || !out_len) return fail(SA_ERR_ARG, "NULL argument");
    if ((names2 == nullptr) != (seqs2 == nullptr)) return fail(SA_ERR_ARG, "names2 and seqs2 must both be set");
    return guard([&] {
        std::vector<sa::Record> r1, r2;
        for (uint64_t i = 0; i < n; ++i) {
            r1.push_back(make_record(names1[i], seqs1[i], quals1 ? quals1[i] : nullptr));
            if (seqs2) r2.push_back(make_record(names2[i], seqs2[i], quals2 ? quals2[i] : nullptr));
        }
        sa::Options o = options_of(opt);
        std::string s;
        sa::map_records(engines(index, o.slots), index->ref, r1, r2, o, &s, nullptr, nullptr);
        *out = dup_out(s, out_len);
        return SA_OK;
    });
}

int sa_reads_load(const char* reads1, const char* reads2, sa_reads** out) {
    if (!reads1 || !out) return fail(SA_ERR_ARG, "NULL argument");
    return guard([&] {
        auto r = std::make_unique<sa_reads>();
        std::string err;
        if (!sa::read_records(reads1, r->r1, err)) return fail(SA_ERR_IO, err);
        if (reads2 && !sa::read_records(reads2, r->r2, err)) return fail(SA_ERR_IO, err);
        if (reads2 && r->r1.size() != r->r2.size()) return fail(SA_ERR_IO, "reads1 and reads2 differ in length");
        *out = r.release();
        return SA_OK;
    });
}

uint64_t sa_reads_count(const sa_reads* reads) { return reads ? reads->r1.size() : 0; }

int sa_reads_record(const sa_reads* reads, uint64_t i, int mate, const char** name, const char** comment,
                    const char** seq, const char** qual, int* end) {
    if (!reads || !name || !comment || !seq || !qual || !end) return fail(SA_ERR_ARG, "NULL argument");
    const std::vector<sa::Record>& v = mate ? reads->r2 : reads->r1;
    if ((mate != 0 && mate != 1) || i >= v.size()) return fail(SA_ERR_ARG, "no such record");
    const sa::Record& r = v[i];
    *name = r.name.c_str();
    *comment = r.has_comment ? r.comment.c_str() : nullptr;
    *seq = r.seq.c_str();
    *qual = r.has_qual ? r.qual.c_str() : nullptr;
    *end = r.end;
    return SA_OK;
}

int sa_reads_load_device(sa_index* index, const char* reads1, const char* reads2, sa_reads** out, double* ingest_ms) {
    if (!index || !reads1 || !out) return fail(SA_ERR_ARG, "NULL argument");
    return guard([&] {
        auto r = std::make_unique<sa_reads>();
        std::string err, raw1, raw2;
        if (!sa::read_file_bytes(reads1, raw1, err)) return fail(SA_ERR_IO, err);
        if (reads2 && !sa::read_file_bytes(reads2, raw2, err)) return fail(SA_ERR_IO, err);
        const int threads = 16;
        if (ingest_ms) *ingest_ms = 0;
        const bool fasta = (!raw1.empty() && raw1[0] == '>') || (reads2 && !raw2.empty() && raw2[0] == '>');
        if (fasta) {
            // FASTA read files: parsed on the host, then staged
            if (!sa::parse_records(raw1, r->r1, err, threads)) return fail(SA_ERR_IO, err);
            if (reads2 && !sa::parse_records(raw2, r->r2, err, threads)) return fail(SA_ERR_IO, err);
            if (reads2 && r->r1.size() != r->r2.size()) return fail(SA_ERR_IO, "reads1 and reads2 differ in length");
            sa_reads* p = r.get();
            const int rc = sa_reads_upload(index, p);
            if (rc != SA_OK) return rc;
            *out = r.release();
            return SA_OK;
        }
        std::vector<sa::FqLoc> l1, l2;
        void* d_seq = nullptr;
        std::vector<uint64_t> off;
        const double ms = sa::ingest_fastq(index->device, raw1, reads2 ? &raw2 : nullptr, l1, l2, &d_seq, off);
        struct Own {
            int dev;
            void* p;
            ~Own() {
                if (p) sa::device_free(dev, p);
            }
        } own{index->device, d_seq};
        if (!sa::records_from_locs(raw1, l1, r->r1, err, threads)) return fail(SA_ERR_IO, err);
        if (reads2 && !sa::records_from_locs(raw2, l2, r->r2, err, threads)) return fail(SA_ERR_IO, err);
        if (reads2 && r->r1.size() != r->r2.size()) return fail(SA_ERR_IO, "reads1 and reads2 differ in length");
        if (!d_seq) return fail(SA_ERR_INTERNAL, "device parse staged no sequences");
        r->device = index->device;
        r->d_seq = d_seq;
        own.p = nullptr;
        r->dr.d_seq = (const uint8_t*)d_seq;
        r->dr.off = std::move(off);
        if (ingest_ms) *ingest_ms = ms;
        *out = r.release();
        return SA_OK;
    });
}

int sa_reads_upload(sa_index* index, sa_reads* reads) {
    if (!index || !reads) return fail(SA_ERR_ARG, "NULL argument");
    return guard([&] {
        const bool paired = !reads->r2.empty();
        const uint64_t n = reads->r1.size();
        std::vector<uint64_t> off(1, 0);
        off.reserve((paired ? 2 : 1) * n + 1);
        for (uint64_t i = 0; i < n; ++i) {
            off.push_back(off.back() + reads->r1[i].seq.size());
            if (paired) off.push_back(off.back() + reads->r2[i].seq.size());
        }
        std::string seq;
        seq.resize(off.back() + 16);  // zero padding after the last base
        for (uint64_t i = 0; i < n; ++i) {
            const uint64_t x = paired ? 2 * i : i;
            std::memcpy(&seq[off[x]], reads->r1[i].seq.data(), reads->r1[i].seq.size());
            if (paired) std::memcpy(&seq[off[x + 1]], reads->r2[i].seq.data(), reads->r2[i].seq.size());
        }
        if (reads->d_seq) sa::device_free(reads->device, reads->d_seq);
        reads->d_seq = nullptr;
        reads->d_seq = sa::device_upload(index->device, seq.data(), seq.size());
        reads->device = index->device;
        reads->dr.d_seq = (const uint8_t*)reads->d_seq;
        reads->dr.off = std::move(off);
        return SA_OK;
    });
}

void sa_reads_destroy(sa_reads* reads) { delete reads; }

int sa_map_reads(sa_index* index, const sa_map_options* opt, const sa_reads* reads, char** out, size_t* out_len,
                 sa_run_stats* stats) {
    if (!index || !reads) return fail(SA_ERR_ARG, "NULL argument");
    return guard([&] {
        sa::Options o = options_of(opt);
        std::string s;
        if (o.mode == 0 && out) s = sa::sam_header(index->ref, o);
        sa::RunStats st;
        const sa::DeviceReads* dr = reads->d_seq && reads->device == index->device ? &reads->dr : nullptr;
        sa::map_records(engines(index, o.slots), index->ref, reads->r1, reads->r2, o, out ? &s : nullptr, &st, dr);
        if (stats) {
            stats->total_ms = st.total_ms; stats->upload_ms = st.upload_ms; stats->nam_ms = st.nam_ms;
            stats->extend_ms = st.extend_ms; stats->host_ms = st.host_ms; stats->records = st.records;
            stats->requests = st.requests; stats->rounds = st.rounds; stats->nams = st.nams;
            stats->fused_batches = st.fused_batches; stats->spec_requests = st.spec_requests;
            stats->spec_used = st.spec_used; stats->graph_replays = st.graph_replays;
            stats->fused_overflows = st.fused_overflows;
            stats->replay_committed = st.replay_committed; stats->replay_probed = st.replay_probed;
            stats->replay_commit_ms = st.replay_commit_ms; stats->replay_probe_ms = st.replay_probe_ms;
            stats->replay_write_ms = st.replay_write_ms; stats->replay_accepted = st.replay_accepted;
        }
        if (out) {
            if (!out_len) return fail(SA_ERR_ARG, "out_len is NULL");
            *out = dup_out(s, out_len);
        }
        return SA_OK;
    });
}

int sa_autotune(sa_index* index, const sa_reads* reads, uint64_t sample_pairs, sa_map_options* opt, char* report,
                size_t report_cap) {
    if (!index || !reads || !opt) return fail(SA_ERR_ARG, "NULL argument");
    return guard([&] {
        const bool paired = !reads->r2.empty();
        const uint64_t n = std::min<uint64_t>(reads->r1.size(), sample_pairs ? sample_pairs : reads->r1.size());
        // the sample: the leading records, resident when the read set is
        std::vector<sa::Record> s1, s2;
        const bool whole = n == reads->r1.size();
        if (!whole) {
            s1.assign(reads->r1.begin(), reads->r1.begin() + (ptrdiff_t)n);
            if (paired) s2.assign(reads->r2.begin(), reads->r2.begin() + (ptrdiff_t)n);
        }
        const std::vector<sa::Record>& r1 = whole ? reads->r1 : s1;
        const std::vector<sa::Record>& r2 = whole ? reads->r2 : s2;
        sa::DeviceReads dr;
        const sa::DeviceReads* pdr = nullptr;
        if (reads->d_seq && reads->device == index->device) {
            dr.d_seq = reads->dr.d_seq;
            dr.off.assign(reads->dr.off.begin(), reads->dr.off.begin() + (ptrdiff_t)((paired ? 2 : 1) * n + 1));
            pdr = &dr;
        }
        std::string rep = "[";
        // the stage forms first (unless the caller fixed them), each timed by
        // its kernel's events on one engine: seeding a thread or a wave per read
        auto kernel_ms = [&](sa::Engine& e, const char* name) {
            double ms = 0;
            for (const auto& k : e.kernel_stats())
                if (k.first == name) ms += k.second.ms;
            return ms;
        };
        uint32_t seed_lanes = options_of(opt).seed_lanes;
        if (opt->seed_lanes == 0) {
            double best_s = 1e300;
            for (uint32_t ln : {1u, 64u}) {
                sa::Options o = options_of(opt);
                o.slots = 1;
                o.seed_lanes = ln;
                const std::vector<sa::Engine*> eng = engines(index, 1);
                const double before = kernel_ms(*eng[0], "k_seed");
                sa::RunStats st;
                sa::map_records(eng, index->ref, r1, r2, o, nullptr, &st, pdr);
                const double ms = kernel_ms(*eng[0], "k_seed") - before;
                char tmp[128];
                std::snprintf(tmp, sizeof tmp, "%s{\"seed_lanes\": %u, \"seed_ms\": %.3f}",
                              rep.size() > 1 ? ", " : "", ln, ms);
                rep += tmp;
                if (ms < best_s) {
                    best_s = ms;
                    seed_lanes = ln;
                }
            }
        }
        // then the lookup stage: a thread, lane groups or a wave per read orientation
        uint32_t lanes = options_of(opt).lookup_lanes;
        if (opt->lookup_lanes == 0) {
            double best_k = 1e300;
            for (uint32_t ln : {1u, 8u, 16u, 64u}) {
                sa::Options o = options_of(opt);
                o.slots = 1;
                o.seed_lanes = seed_lanes;
                o.lookup_lanes = ln;
                const std::vector<sa::Engine*> eng = engines(index, 1);
                auto lookup_ms = [&] {
                    double ms = 0;
                    for (const auto& k : eng[0]->kernel_stats())
                        if (k.first == "k_hits") ms += k.second.ms;
                    return ms;
                };
                const double before = lookup_ms();
                sa::RunStats st;
                sa::map_records(eng, index->ref, r1, r2, o, nullptr, &st, pdr);
                const double ms = lookup_ms() - before;
                char tmp[128];
                std::snprintf(tmp, sizeof tmp, "%s{\"lookup_lanes\": %u, \"lookup_ms\": %.3f}",
                              rep.size() > 1 ? ", " : "", ln, ms);
                rep += tmp;
                if (ms < best_k) {
                    best_k = ms;
                    lanes = ln;
                }
            }
        }
        const int slot_set[] = {1, 2, 3, 4, 6, 8};
        const uint64_t batch_set[] = {16ull << 20, 32ull << 20, 64ull << 20};
        double best = 1e300;
        int best_slots = 0, best_fused = 0;
        uint64_t best_batch = 0;
        // the fused pass only changes alignment mode (the other modes extend nothing)
        const int n_pass = options_of(opt).mode == 0 ? 2 : 1;
        for (int fz = 0; fz < n_pass; ++fz)
            for (int sl : slot_set)
                for (uint64_t bb : batch_set) {
                    sa::Options o = options_of(opt);
                    o.slots = sl;
                    o.batch_bases = bb;
                    o.lookup_lanes = lanes;
                    o.seed_lanes = seed_lanes;
                    o.fused = fz == 1;
                    double ms = 0;
                    // the first run grows the buffers (and sizes and records the fused pass);
                    // a schedule the device memory cannot hold is skipped
                    try {
                        for (int rep_i = 0; rep_i < 2; ++rep_i) {
                            sa::RunStats st;
                            sa::map_records(engines(index, sl), index->ref, r1, r2, o, nullptr, &st, pdr);
                            ms = st.total_ms;
                        }
                    } catch (const std::runtime_error& e) {
                        if (std::string(e.what()).find("out of memory") == std::string::npos) throw;
                        ms = 1e300;
                    }
                    char tmp[160];
                    if (ms < 1e300)
                        std::snprintf(tmp, sizeof tmp,
                                      "%s{\"fused\": %d, \"slots\": %d, \"batch_bases\": %llu, \"ms\": %.2f}",
                                      rep.size() > 1 ? ", " : "", fz, sl, (unsigned long long)bb, ms);
                    else
                        std::snprintf(tmp, sizeof tmp,
                                      "%s{\"fused\": %d, \"slots\": %d, \"batch_bases\": %llu, \"ms\": null}",
                                      rep.size() > 1 ? ", " : "", fz, sl, (unsigned long long)bb);
                    rep += tmp;
                    if (ms < best) {
                        best = ms;
                        best_slots = sl;
                        best_batch = bb;
                        best_fused = fz;
                    }
                }
        rep += "]";
        opt->slots = best_slots;
        opt->batch_bases = best_batch;
        opt->lookup_lanes = (int32_t)lanes;
        opt->seed_lanes = (int32_t)seed_lanes;
        opt->fused = best_fused;
        if (report && report_cap) {
            const size_t m = std::min(report_cap - 1, rep.size());
            std::memcpy(report, rep.data(), m);
            report[m] = 0;
        }
        return SA_OK;
    });
}

int sa_server_create(sa_index* index, const sa_map_options* opt, uint64_t arena_bytes, sa_server** out) {
    if (!index || !out) return fail(SA_ERR_ARG, "NULL argument");
    return guard([&] {
        sa::Options o = options_of(opt);
        if (!opt || opt->batch_bases == 0) o.batch_bases = 8ull << 20;  // serving default: short batches
        if (arena_bytes == 0) arena_bytes = std::min<uint64_t>(4ull << 30, sa::device_free_bytes(index->device) / 4);
        auto s = std::make_unique<sa_server>();
        s->index = index;
        std::vector<sa::Engine*> v;
        for (int i = 0; i < o.slots; ++i) {
            s->engines.push_back(std::make_unique<sa::Engine>(sa::Engine::ShareIndex{}, *index->eng));
            v.push_back(s->engines.back().get());
        }
        s->srv = std::make_unique<sa::Server>(v, index->ref, o, index->device, arena_bytes);
        *out = s.release();
        return SA_OK;
    });
}

void sa_server_destroy(sa_server* server) { delete server; }

int sa_server_submit(sa_server* server, uint64_t n, const char* const* names1, const char* const* seqs1,
                     const char* const* quals1, const char* const* names2, const char* const* seqs2,
                     const char* const* quals2, uint64_t* ticket) {
    if (!server || !ticket || (n && (!names1 || !seqs1))) return fail(SA_ERR_ARG, "NULL argument");
    if ((names2 == nullptr) != (seqs2 == nullptr)) return fail(SA_ERR_ARG, "names2 and seqs2 must both be set");
    const double t_in = sa::steady_ms();
    return guard([&] {
        std::vector<sa::Record> r1, r2;
        r1.reserve(n);
        for (uint64_t i = 0; i < n; ++i) {
            r1.push_back(make_record(names1[i], seqs1[i], quals1 ? quals1[i] : nullptr));
            if (seqs2) r2.push_back(make_record(names2[i], seqs2[i], quals2 ? quals2[i] : nullptr));
        }
        *ticket = server->srv->submit(std::move(r1), std::move(r2), t_in);
        return SA_OK;
    });
}

int sa_server_submit_reads(sa_server* server, const sa_reads* reads, uint64_t first, uint64_t count,
                           uint64_t* ticket) {
    if (!server || !reads || !ticket) return fail(SA_ERR_ARG, "NULL argument");
    if (first > reads->r1.size() || count > reads->r1.size() - first) return fail(SA_ERR_ARG, "range out of bounds");
    const double t_in = sa::steady_ms();
    return guard([&] {
        const auto b = reads->r1.begin() + (ptrdiff_t)first;
        std::vector<sa::Record> r1(b, b + (ptrdiff_t)count), r2;
        if (!reads->r2.empty()) {
            const auto b2 = reads->r2.begin() + (ptrdiff_t)first;
            r2.assign(b2, b2 + (ptrdiff_t)count);
        }
        *ticket = server->srv->submit(std::move(r1), std::move(r2), t_in);
        return SA_OK;
    });
}

int sa_server_wait(sa_server* server, uint64_t ticket, char** out, size_t* out_len, double* latency_ms) {
    if (!server) return fail(SA_ERR_ARG, "NULL argument");
    if (out && !out_len) return fail(SA_ERR_ARG, "out_len is NULL");
    return guard([&] {
        std::string s, err;
        sa::Server::RequestTimes t;
        const sa::Server::WaitStatus w = server->srv->wait(ticket, s, t, err);
        if (w != sa::Server::WAIT_OK) return fail(w == sa::Server::WAIT_BAD_TICKET ? SA_ERR_ARG : SA_ERR_INTERNAL, err);
        if (latency_ms) *latency_ms = t.done_ms;
        if (out) *out = dup_out(s, out_len);
        return SA_OK;
    });
}

int sa_server_get_stats(const sa_server* server, sa_server_stats* st) {
    if (!server || !st) return fail(SA_ERR_ARG, "NULL argument");
    const sa::Server::Stats s = server->srv->stats();
    st->submitted = s.submitted;
    st->completed = s.completed;
    st->batches = s.batches;
    st->batch_requests = s.batch_requests;
    st->max_requests_per_batch = s.max_requests_per_batch;
    st->arena_bytes = s.arena_bytes;
    st->arena_peak = s.arena_peak;
    st->busy_ms = s.busy_ms;
    return SA_OK;
}

int sa_randstrobes_query_batch(sa_index* index, const uint8_t* seqs, const uint64_t* offsets, uint32_t n,
                               sa_qrand* out_fwd, sa_qrand* out_rc, uint32_t* counts) {
    if (!index || (!seqs && n) || !offsets || !out_fwd || !out_rc || !counts) return fail(SA_ERR_ARG, "NULL argument");
    return guard([&] {
        index->eng->upload_reads(seqs, offsets, n);
        std::vector<sa::QRand> qr;
        std::vector<uint32_t> cnt;
        index->eng->randstrobes(qr, cnt);
        for (uint32_t i = 0; i < n; ++i) {
            uint64_t o = offsets[i], L = offsets[i + 1] - o;
            counts[2 * i] = cnt[2 * i];
            counts[2 * i + 1] = cnt[2 * i + 1];
            for (uint32_t x = 0; x < cnt[2 * i]; ++x) std::memcpy(&out_fwd[o + x], &qr[2 * o + x], sizeof(sa_qrand));
            for (uint32_t x = 0; x < cnt[2 * i + 1]; ++x) std::memcpy(&out_rc[o + x], &qr[2 * o + L + x], sizeof(sa_qrand));
        }
        return SA_OK;
    });
}

int sa_find_nams_batch(sa_index* index, const uint8_t* seqs, const uint64_t* offsets, uint32_t n, uint32_t mcs,
                       uint32_t rescue_distance, sa_nam** nams, uint64_t* nam_offsets) {
    if (!index || (!seqs && n) || !offsets || !nams || !nam_offsets) return fail(SA_ERR_ARG, "NULL argument");
    if (mcs > 3) return fail(SA_ERR_ARG, "mcs must be 0..3");
    return guard([&] {
        index->eng->upload_reads(seqs, offsets, n);
        sa::NamBatch nb;
        sa::MapParams mp{mcs, rescue_distance};
        sa::ChainParams cp{50, 0.1f, 0.05f, 0.7f, -1, 0.01f, 10.0f};
        index->eng->find_nams(mp, cp, nb);
        static_assert(sizeof(sa_nam) == sizeof(sa::NamRec), "sa_nam layout");
        sa_nam* p = (sa_nam*)std::malloc(std::max<size_t>(1, nb.nams.size()) * sizeof(sa_nam));
        if (!p) throw std::bad_alloc();
        if (!nb.nams.empty()) std::memcpy(p, nb.nams.data(), nb.nams.size() * sizeof(sa_nam));
        for (uint32_t i = 0; i <= n; ++i) nam_offsets[i] = nb.nam_off[i];
        *nams = p;
        return SA_OK;
    });
}

int sa_map_batch(sa_index* index, const sa_map_options* opt, const uint8_t* seqs, const uint64_t* offsets, uint32_t n,
                 uint32_t speculative, void* stream, sa_nam** nams, uint64_t* nam_offsets, sa_nam_extension** ext,
                 uint32_t** cigars) {
    if (!index || (!seqs && n) || !offsets || !nams || !nam_offsets || !ext || !cigars)
        return fail(SA_ERR_ARG, "NULL argument");
    return guard([&] {
        const sa::Options o = options_of(opt);
        sa::Engine& eng = *index->eng;
        // the caller's stream for the duration of the call
        struct StreamScope {
            sa::Engine& e;
            ihipStream_t* prev;
            ~StreamScope() { e.set_stream(prev); }
        } scope{eng, eng.set_stream((ihipStream_t*)stream)};
        eng.upload_reads(seqs, offsets, n);
        sa::NamBatch nb;
        sa::ChainParams cp{50, 0.1f, 0.05f, 0.7f, -1, 0.01f, 10.0f};
        eng.find_nams(sa::MapParams{o.mcs, o.rescue_distance, o.lookup_lanes, o.seed_lanes}, cp, nb);
        const uint64_t N = nb.nams.size();
        sa_nam* pn = (sa_nam*)std::malloc(std::max<uint64_t>(1, N) * sizeof(sa_nam));
        sa_nam_extension* pe = (sa_nam_extension*)std::calloc(std::max<uint64_t>(1, N), sizeof(sa_nam_extension));
        if (!pn || !pe) {
            std::free(pn);
            std::free(pe);
            throw std::bad_alloc();
        }
        if (N) std::memcpy(pn, nb.nams.data(), N * sizeof(sa_nam));
        for (uint32_t i = 0; i <= n; ++i) nam_offsets[i] = nb.nam_off[i];
        // extend_seed requests (the mapping path's EXT_SEED) for the first
        // `speculative` consistent NAMs of each read
        std::vector<sa::ExtRequest> req;
        std::vector<uint64_t> which;
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t done = 0;
            for (uint64_t x = nb.nam_off[i]; x < nb.nam_off[i + 1] && done < speculative; ++x) {
                const sa::NamRec& r = nb.nams[x];
                if (!(r.flags & 2)) continue;  // not consistent as it is
                ++done;
                sa::ExtRequest q;
                std::memset(&q, 0, sizeof q);
                q.kind = sa::EXT_SEED;
                q.read_id = i;
                q.flags = (o.use_ssw ? 1u : 0u) | 2u | ((r.flags & 1) ? 4u : 0u);
                q.ref_id = r.ref_id;
                q.ref_start = r.ref_start;
                q.ref_end = r.ref_end;
                q.query_start = r.query_start;
                q.query_end = r.query_end;
                q.anchor_off = r.anchor_off;
                q.n_anchors = r.n_anchors;
                req.push_back(q);
                which.push_back(x);
            }
        }
        std::vector<sa::ExtResult> res;
        std::vector<uint32_t> cig;
        std::vector<uint64_t> coff;
        eng.extend(req.data(), (uint32_t)req.size(), o.scores, res, cig, coff);
        for (size_t y = 0; y < req.size(); ++y) {
            const sa::ExtResult& a = res[y];
            sa_nam_extension& e = pe[which[y]];
            e.extended = 1;
            e.found = a.found;
            if (!a.found) continue;
            e.gapped = a.gapped;
            e.score = a.score;
            e.ref_start = a.ref_start;
            e.length = a.length;
            e.query_start = a.query_start;
            e.query_end = a.query_end;
            e.edit_distance = a.edit_distance;
            e.cigar_offset = coff[y];
            e.n_cigar = a.n_cigar;
        }
        uint32_t* pc = (uint32_t*)std::malloc(std::max<size_t>(1, cig.size()) * 4);
        if (!pc) {
            std::free(pn);
            std::free(pe);
            throw std::bad_alloc();
        }
        if (!cig.empty()) std::memcpy(pc, cig.data(), cig.size() * 4);
        *nams = pn;
        *ext = pe;
        *cigars = pc;
        return SA_OK;
    });
}

int sa_find_hits_batch(sa_index* index, const uint8_t* seqs, const uint64_t* offsets, uint32_t n, uint32_t mcs,
                       uint32_t rescue_distance, sa_hit** hits, uint64_t* hit_offsets, sa_hits_details* details) {
    if (!index || (!seqs && n) || !offsets || !hits || !hit_offsets || !details) return fail(SA_ERR_ARG, "NULL argument");
    if (mcs > 3) return fail(SA_ERR_ARG, "mcs must be 0..3");
    return guard([&] {
        index->eng->upload_reads(seqs, offsets, n);
        std::vector<sa::Hit> h;
        std::vector<sa::HitCounts> hc;
        index->eng->find_hits(sa::MapParams{mcs, rescue_distance}, h, hc);
        uint64_t total = 0;
        for (uint32_t t = 0; t < 2 * n; ++t) total += hc[t].n_hits;
        sa_hit* p = (sa_hit*)std::malloc(std::max<uint64_t>(1, total) * sizeof(sa_hit));
        if (!p) throw std::bad_alloc();
        uint64_t at = 0;
        for (uint32_t t = 0; t < 2 * n; ++t) {
            const uint64_t ro = offsets[t / 2] - offsets[0], L = offsets[t / 2 + 1] - offsets[t / 2];
            const sa::Hit* src = h.data() + 2 * ro + (t & 1) * L;
            hit_offsets[t] = at;
            for (uint32_t x = 0; x < hc[t].n_hits; ++x) {
                const sa::Hit& a = src[x];
                p[at++] = sa_hit{a.query_start, a.query_end, a.position, a.hash, a.hash_rc,
                                 (a.flags & sa::HIT_PARTIAL) ? 1u : 0u, (a.flags & sa::HIT_FILTERED) ? 1u : 0u};
            }
            const sa::HitCounts& c = hc[t];
            details[t] = sa_hits_details{c.full_not_found, c.full_filtered, c.full_found, c.partial_not_found,
                                         c.partial_filtered, c.partial_found, c.rescued};
        }
        hit_offsets[2 * n] = at;
        *hits = p;
        return SA_OK;
    });
}

int sa_chain_batch(sa_index* index, const uint8_t* seqs, const uint64_t* offsets, uint32_t n, const sa_hit* hits,
                   const uint64_t* hit_offsets, const sa_hits_details* details, sa_nam** nams, uint64_t* nam_offsets) {
    if (!index || (!seqs && n) || !offsets || (!hits && n) || !hit_offsets || !details || !nams || !nam_offsets)
        return fail(SA_ERR_ARG, "NULL argument");
    return guard([&] {
        // device layout: orientation t of read i at 2 * off[i] + t % 2 * L
        const uint64_t nb = offsets[n] - offsets[0];
        std::vector<sa::Hit> h(2 * nb + 1);
        std::vector<sa::HitCounts> hc(2ull * n);
        for (uint32_t t = 0; t < 2 * n; ++t) {
            const uint64_t ro = offsets[t / 2] - offsets[0], L = offsets[t / 2 + 1] - offsets[t / 2];
            const uint64_t cnt = hit_offsets[t + 1] - hit_offsets[t];
            if (hit_offsets[t + 1] < hit_offsets[t] || cnt > L)
                throw std::invalid_argument("read orientation " + std::to_string(t) +
                                            " has more hits than its read has bases");
            for (uint64_t x = 0; x < cnt; ++x) {
                const sa_hit& a = hits[hit_offsets[t] + x];
                sa::Hit& d = h[2 * ro + (t & 1) * L + x];
                d = sa::Hit{a.hash, a.hash_revcomp, a.position, sa::NONE64, 0, (uint32_t)a.query_start,
                            (uint32_t)a.query_end,
                            (a.is_partial ? sa::HIT_PARTIAL : 0u) | (a.is_filtered ? sa::HIT_FILTERED : 0u)};
            }
            const sa_hits_details& c = details[t];
            hc[t] = sa::HitCounts{(uint32_t)c.full_not_found, (uint32_t)c.full_filtered, (uint32_t)c.full_found,
                                  (uint32_t)c.partial_not_found, (uint32_t)c.partial_filtered,
                                  (uint32_t)c.partial_found, (uint32_t)c.rescued, (uint32_t)cnt};
        }
        std::vector<uint64_t> off(offsets, offsets + n + 1);
        for (auto& x : off) x -= offsets[0];
        index->eng->upload_reads(seqs + offsets[0], off.data(), n);
        sa::NamBatch nb_;
        sa::ChainParams cp{50, 0.1f, 0.05f, 0.7f, -1, 0.01f, 10.0f};
        index->eng->chain_hits(h.data(), hc.data(), cp, nb_);
        sa_nam* p = (sa_nam*)std::malloc(std::max<size_t>(1, nb_.nams.size()) * sizeof(sa_nam));
        if (!p) throw std::bad_alloc();
        if (!nb_.nams.empty()) std::memcpy(p, nb_.nams.data(), nb_.nams.size() * sizeof(sa_nam));
        for (uint32_t i = 0; i <= n; ++i) nam_offsets[i] = nb_.nam_off[i];
        *nams = p;
        return SA_OK;
    });
}

int sa_find_nams_batch_noalloc(sa_index* index, const uint8_t* d_seqs, const uint64_t* d_offsets, uint32_t n,
                               uint32_t mcs, uint32_t rescue_distance, sa_nam* d_nams, uint64_t nam_cap,
                               uint64_t* d_nam_offsets, uint64_t* n_nams, void* stream) {
    if (!index || (!d_seqs && n) || !d_offsets || (!d_nams && nam_cap) || !d_nam_offsets || !n_nams)
        return fail(SA_ERR_ARG, "NULL argument");
    if (mcs > 3) return fail(SA_ERR_ARG, "mcs must be 0..3");
    static_assert(sa::NOALLOC_SCRATCH_BYTES == SA_NOALLOC_SCRATCH_BYTES, "scratch size");
    return guard([&] {
        uint8_t* scr = sa::noalloc_scratch(index->device);
        sa::MapParams mp{mcs, rescue_distance};
        sa::ChainParams cp{50, 0.1f, 0.05f, 0.7f, -1, 0.01f, 10.0f};
        index->eng->find_nams_noalloc(d_seqs, d_offsets, n, mp, cp, (sa::NamRec*)d_nams, nam_cap, d_nam_offsets,
                                      n_nams, (ihipStream_t*)stream, scr, sa::NOALLOC_SCRATCH_BYTES);
        return SA_OK;
    });
}

namespace {
// Runs n extension jobs of the aligner's entry points on the device: query i
// against reference window i (ACGT; each window becomes its own contig of a
// reference made for the call), through the same kernels as the mapping path.
int extend_jobs(sa_ctx* ctx, const sa::Scores& sc, uint32_t k, const sa_extend_job* jobs, uint32_t n,
                const uint8_t* queries, const uint64_t* query_offsets, const uint8_t* refs,
                const uint64_t* ref_offsets, const uint32_t* anchors, uint64_t n_anchor_pairs, sa_alignment* out,
                uint32_t** cigars, uint64_t* cigar_offsets) {
    std::vector<uint64_t> words, woff{0}, len;
    for (uint32_t i = 0; i < n; ++i) {
        const uint64_t L = ref_offsets[i + 1] - ref_offsets[i];
        for (uint64_t x = 0; x < L; ++x) {
            const uint32_t c = sa::nt2(refs[ref_offsets[i] + x]);
            if (c > 3) throw std::invalid_argument("reference windows must be ACGT");
            if ((x & 31) == 0) words.push_back(0);
            words.back() |= (uint64_t)c << ((x & 31) * 2);
        }
        woff.push_back(words.size());
        len.push_back(L);
    }
    sa::Engine eng(ctx->device);
    eng.upload_reference(words.data(), words.size(), woff.data(), len.data(), n);
    sa::SeedParams sp = sa::seed_params_for_read_length(150);
    sp.k = (int32_t)k;
    uint64_t bs[3] = {0, 0, 0};
    sp.bits = 1;
    eng.upload_index(nullptr, 0, bs, sp);
    eng.upload_reads(queries, query_offsets, n);
    eng.set_chain_anchors(anchors, n_anchor_pairs);
    std::vector<sa::ExtRequest> req(n);
    for (uint32_t i = 0; i < n; ++i) {
        const sa_extend_job& j = jobs[i];
        sa::ExtRequest& r = req[i];
        std::memset(&r, 0, sizeof r);
        r.read_id = i;
        r.ref_id = i;
        r.win_start = 0;
        r.win_end = (uint32_t)len[i];
        switch (j.kind) {
            case SA_EXT_SSW:
                r.kind = sa::EXT_RESCUE;
                r.flags = 8;  // plain alignment: no shared-substring precheck, query as given
                break;
            case SA_EXT_HAMMING: r.kind = sa::EXT_HAMMING; break;
            case SA_EXT_PIECEWISE:
                if (j.anchor_offset + j.n_anchors > n_anchor_pairs || j.n_anchors == 0)
                    throw std::invalid_argument("piecewise job anchors outside the anchor array");
                r.kind = sa::EXT_ALIGN_PW;
                r.anchor_off = j.anchor_offset;
                r.n_anchors = j.n_anchors;
                r.ref_start = j.padding;
                break;
            case SA_EXT_GLOBAL: r.kind = sa::EXT_PW_GLOBAL; break;
            case SA_EXT_XDROP: r.kind = sa::EXT_PW_XDROP; break;
            case SA_EXT_XDROP_REV:
                r.kind = sa::EXT_PW_XDROP;
                r.flags = 1;
                break;
            default: throw std::invalid_argument("unknown extension job kind " + std::to_string(j.kind));
        }
    }
    std::vector<sa::ExtResult> res;
    std::vector<uint32_t> cig;
    std::vector<uint64_t> coff;
    eng.extend(req.data(), n, sc, res, cig, coff);
    for (uint32_t i = 0; i < n; ++i) {
        out[i].found = res[i].found;
        out[i].score = res[i].score;
        out[i].query_start = res[i].query_start;
        out[i].query_end = res[i].query_end;
        out[i].ref_start = res[i].ref_start;
        out[i].ref_end = res[i].ref_start + res[i].length;
        out[i].edit_distance = res[i].edit_distance;
        out[i].n_cigar = res[i].found ? res[i].n_cigar : 0;
        cigar_offsets[i] = coff[i];
    }
    cigar_offsets[n] = coff[n];
    uint32_t* p = (uint32_t*)std::malloc(std::max<size_t>(1, cig.size()) * 4);
    if (!p) throw std::bad_alloc();
    if (!cig.empty()) std::memcpy(p, cig.data(), cig.size() * 4);
    *cigars = p;
    return SA_OK;
}
}  // namespace

int sa_extend_batch(sa_ctx* ctx, const sa_scores* scores, uint32_t k, const sa_extend_job* jobs, uint32_t n,
                    const uint8_t* queries, const uint64_t* query_offsets, const uint8_t* refs,
                    const uint64_t* ref_offsets, const uint32_t* anchors, uint64_t n_anchor_pairs, sa_alignment* out,
                    uint32_t** cigars, uint64_t* cigar_offsets) {
    if (!ctx || !scores || (!jobs && n) || !query_offsets || !ref_offsets || !out || !cigars || !cigar_offsets ||
        (!anchors && n_anchor_pairs))
        return fail(SA_ERR_ARG, "NULL argument");
    return guard([&] {
        const sa::Scores sc{scores->match_score, scores->mismatch_penalty, scores->gap_open_penalty,
                            scores->gap_extend_penalty, scores->end_bonus, scores->xdrop};
        for (int32_t v : {sc.match, sc.mismatch, sc.gap_open, sc.gap_extend})
            if (v < 0 || v > 127) throw std::invalid_argument("scores and penalties must be in 0..127");
        if (k == 0 || k > 32) throw std::invalid_argument("k must be in 1..32");
        return extend_jobs(ctx, sc, k, jobs, n, queries, query_offsets, refs, ref_offsets, anchors, n_anchor_pairs,
                           out, cigars, cigar_offsets);
    });
}

int sa_align_ssw_batch(sa_ctx* ctx, const uint8_t* queries, const uint64_t* query_offsets, const uint8_t* refs,
                       const uint64_t* ref_offsets, uint32_t n, sa_alignment* out, uint32_t** cigars,
                       uint64_t* cigar_offsets) {
    if (!ctx || !query_offsets || !ref_offsets || !out || !cigars || !cigar_offsets) return fail(SA_ERR_ARG, "NULL argument");
    return guard([&] {
        std::vector<sa_extend_job> jobs(n);
        for (auto& j : jobs) {
            std::memset(&j, 0, sizeof j);
            j.kind = SA_EXT_SSW;
        }
        return extend_jobs(ctx, sa::Scores{2, 8, 12, 1, 10, 500}, 20, jobs.data(), n, queries, query_offsets, refs,
                           ref_offsets, nullptr, 0, out, cigars, cigar_offsets);
    });
}

int sa_index_stats(const sa_index* index, double* kernel_ms) {
    if (!index) return fail(SA_ERR_ARG, "index is NULL");
    if (kernel_ms) {
        *kernel_ms = index->eng->total_kernel_ms();
        for (const auto& e : index->slots) *kernel_ms += e->total_kernel_ms();
    }
    return SA_OK;
}

size_t sa_index_kernel_times(const sa_index* index, char* buf, size_t cap) {
    if (!index) return 0;
    std::vector<const sa::Engine*> v{index->eng.get()};
    for (const auto& e : index->slots) v.push_back(e.get());
    return kernel_times_json(v, buf, cap);
}

size_t sa_server_kernel_times(const sa_server* server, char* buf, size_t cap) {
    if (!server) return 0;
    std::vector<const sa::Engine*> v;
    for (const auto& e : server->engines) v.push_back(e.get());
    return kernel_times_json(v, buf, cap);
}

int sa_hbm_copy_probe(sa_ctx* ctx, uint64_t bytes, int iters, double* gbps) {
    if (!ctx || !gbps || iters <= 0 || bytes < 16) return fail(SA_ERR_ARG, "bad argument");
    return guard([&] {
        *gbps = sa::hbm_copy_probe(ctx->device, bytes, iters);
        return SA_OK;
    });
}

void sa_free(void* p) { std::free(p); }

}  // extern "C"
