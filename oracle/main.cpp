// Oracle command-line driver (src/main.rs:322-849 subset): chunking, per-chunk
// RNG/insert-size state, ordered output, N worker threads.
// TEST INFRASTRUCTURE ONLY (see oracle.hpp).
#include "oracle.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <thread>

using namespace orc;

static void usage() {
    std::fprintf(stderr,
                 "usage: sa_oracle [-t N] [-x|--aemb] [-r N] [-N N] [-U] [--eqx] [--ssw] [-C] [--details]\n"
                 "                 [--rg-id ID] [--rg TAG] [--chunk-size N] [--mcs MODE] [-o OUT]\n"
                 "                 [-A N] [-B N] [-O N] [-E N] [-L N] [--xdrop N]\n"
                 "                 [--dump-index PATH] [--paf-out PATH] [--time] ref.fa reads1 [reads2]\n"
                 "  --paf-out PATH: also map in -x mode with the same index and write the PAF to PATH\n");
}

int main(int argc, char** argv) {
    size_t threads = 1, chunk_size = 1000000, read_length = 0;
    int mode = 0;
    bool time_it = false;
    bool rng_stats = false;  // print the number of bounded random draws to stderr
    size_t t1_chunks = 0;  // with --time: also map the first N chunks again on one thread
    MappingParams mp;
    SamOpts so;
    Scores sc;  // src/main.rs:246-272 (-A/-B/-O/-E/-L, --xdrop)
    int xdrop = 500;
    std::vector<std::string> rg_fields;
    std::string out_path, dump_index, dump_nams, paf_out;
    std::vector<std::string> pos;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto val = [&](const std::string& name) -> std::string {
            size_t eq = a.find('=');
            if (eq != std::string::npos && a.compare(0, eq, name) == 0) return a.substr(eq + 1);
            if (i + 1 >= argc) { usage(); std::exit(2); }
            return argv[++i];
        };
        auto is = [&](const char* name) { return a == name || a.rfind(std::string(name) + "=", 0) == 0; };
        if (a == "-t") threads = std::stoul(val("-t"));
        else if (a == "-x") mode = 1;
        else if (a == "--aemb") mode = 2;
        else if (a == "-r") read_length = std::stoul(val("-r"));
        else if (a == "-N") mp.max_secondary = std::stoul(val("-N"));
        else if (a == "-U") mp.output_unmapped = false;
        else if (a == "--eqx") so.eqx = true;
        else if (a == "--ssw") mp.use_ssw = true;
        else if (a == "-C") so.fastq_comments = true;
        else if (a == "--details") so.details = true;
        else if (a == "--no-PG") {}
        else if (a == "--time") time_it = true;
        else if (a == "--rng-stats") rng_stats = true;
        else if (is("--t1-chunks")) t1_chunks = std::stoul(val("--t1-chunks"));
        else if (is("--rg-id")) { so.has_rg = true; so.rg_id = val("--rg-id"); }
        else if (is("--rg")) rg_fields.push_back(val("--rg"));
        else if (is("--chunk-size")) chunk_size = std::stoul(val("--chunk-size"));
        else if (is("--mcs")) {
            std::string m = val("--mcs");
            mp.mcs = m == "rescue" ? Mcs::Rescue : m == "off" ? Mcs::Off : m == "first-strobe" ? Mcs::FirstStrobe : Mcs::Always;
        } else if (a == "-A") sc.match = std::stoi(val("-A"));
        else if (a == "-B") sc.mismatch = std::stoi(val("-B"));
        else if (a == "-O") sc.gap_open = std::stoi(val("-O"));
        else if (a == "-E") sc.gap_extend = std::stoi(val("-E"));
        else if (a == "-L") sc.end_bonus = std::stoi(val("-L"));
        else if (is("--xdrop")) xdrop = std::stoi(val("--xdrop"));
        else if (a == "-o") out_path = val("-o");
        else if (is("--dump-index")) dump_index = val("--dump-index");
        else if (is("--dump-nams")) dump_nams = val("--dump-nams");
        else if (is("--paf-out")) paf_out = val("--paf-out");
        else if (!a.empty() && a[0] == '-' && a.size() > 1) { usage(); return 2; }
        else pos.push_back(a);
    }
    if (!rg_fields.empty() && !so.has_rg) { so.has_rg = true; so.rg_id = "1"; }
    if (pos.size() < 2) { usage(); return 2; }
    std::string err;
    std::vector<Record> r1, r2;
    if (!read_reads(pos[1], r1, err)) { std::fprintf(stderr, "error: %s\n", err.c_str()); return 1; }
    if (pos.size() > 2 && !read_reads(pos[2], r2, err)) { std::fprintf(stderr, "error: %s\n", err.c_str()); return 1; }
    if (read_length == 0) read_length = estimate_read_length(r1);
    SeedingParams sp = seeding_params_for_read_length(read_length);
    auto t0 = std::chrono::steady_clock::now();
    std::vector<RefSeq> refs = read_fasta(pos[0]);
    if (refs.empty()) { std::fprintf(stderr, "error: No sequences found in the reference FASTA\n"); return 1; }
    int bits = pick_bits(refs, sp.syncmer);
    Index index = make_index(refs, sp, bits, 0.0002);
    auto t1 = std::chrono::steady_clock::now();
    if (!dump_index.empty()) {
        std::ofstream f(dump_index, std::ios::binary);
        uint64_t n = index.randstrobes.size(), nb = index.bucket_starts.size();
        uint64_t hdr[4] = {n, nb, (uint64_t)index.bits, (uint64_t)index.filter_cutoff};
        f.write((const char*)hdr, sizeof hdr);
        f.write((const char*)index.randstrobes.data(), (std::streamsize)(n * 16));
        f.write((const char*)index.bucket_starts.data(), (std::streamsize)(nb * 8));
    }

    if (!dump_nams.empty()) {
        // NAMs of every read after the stable score sort, before the shuffle of the tied best
        FILE* f = std::fopen(dump_nams.c_str(), "wb");
        ChainingParams cp;
        for (size_t i = 0; i < r1.size(); ++i) {
            NamDetails nd;
            std::vector<Nam> nams = get_nams_by_chaining((const uint8_t*)r1[i].seq.data(), r1[i].seq.size(), index, cp,
                                                         mp.rescue_distance, mp.mcs, nd);
            std::stable_sort(nams.begin(), nams.end(), [](const Nam& a, const Nam& b) { return a.score > b.score; });
            for (const Nam& n : nams) {
                uint32_t sb;
                std::memcpy(&sb, &n.score, 4);
                std::fprintf(f, "%zu\t%zu\t%zu\t%zu\t%zu\t%zu\t%zu\t%zu\t%zu\t%08x\t%u\n", i, n.nam_id, n.ref_id,
                             n.ref_start, n.ref_end, n.query_start, n.query_end, n.matching_bases, n.anchors.size(), sb,
                             (unsigned)n.is_revcomp);
            }
        }
        std::fclose(f);
        return 0;
    }
    Mapper m;
    m.index = &index;
    m.refs = &refs;
    m.mp = mp;
    m.so = so;
    m.sc = sc;
    m.xdrop = xdrop;
    m.mode = mode;
    m.abundances.assign(refs.size(), 0.0);

    // src/main.rs:591-604: a chunk ends with the record that makes its nucleotide count exceed chunk_size
    using Pair = std::pair<Record, std::optional<Record>>;
    std::vector<std::vector<Pair>> chunks;
    {
        std::vector<Pair> cur;
        size_t nt = 0;
        size_t n = r2.empty() ? r1.size() : std::min(r1.size(), r2.size());
        for (size_t i = 0; i < n; ++i) {
            if (r2.empty()) { nt += r1[i].seq.size(); cur.push_back({r1[i], std::nullopt}); }
            else { nt += r1[i].seq.size() + r2[i].seq.size(); cur.push_back({r1[i], r2[i]}); }
            if (nt > chunk_size) { chunks.push_back(std::move(cur)); cur.clear(); nt = 0; }
        }
        if (!cur.empty()) chunks.push_back(std::move(cur));
    }
    std::vector<std::string> outs(chunks.size());
    std::vector<std::vector<double>> abund(threads, std::vector<double>(refs.size(), 0.0));
    std::vector<orc::StageTimes> stages(threads);
    std::atomic<size_t> next{0};
    auto t2 = std::chrono::steady_clock::now();
    auto worker = [&](size_t tid) {
        Mapper mm = m;
        mm.abundances.assign(refs.size(), 0.0);
        for (;;) {
            size_t c = next.fetch_add(1);
            if (c >= chunks.size()) break;
            mm.map_chunk(chunks[c], outs[c]);
        }
        abund[tid] = mm.abundances;
        stages[tid] = orc::stage_times();
    };
    std::vector<std::thread> th;
    for (size_t t = 0; t < threads; ++t) th.emplace_back(worker, t);
    for (auto& t : th) t.join();
    auto t3 = std::chrono::steady_clock::now();
    // single-thread rate on the leading chunks (same index, output dropped)
    double t1_s = 0;
    size_t t1_records = 0;
    if (time_it && t1_chunks) {
        Mapper mm = m;
        mm.abundances.assign(refs.size(), 0.0);
        std::string sink;
        auto a = std::chrono::steady_clock::now();
        for (size_t c = 0; c < chunks.size() && c < t1_chunks; ++c) {
            sink.clear();
            mm.map_chunk(chunks[c], sink);
            t1_records += chunks[c].size();
        }
        t1_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count();
    }

    FILE* out = out_path.empty() ? stdout : std::fopen(out_path.c_str(), "wb");
    if (!out) { std::fprintf(stderr, "error: cannot open %s\n", out_path.c_str()); return 1; }
    if (mode == 0) {
        std::string h = sam_header(refs, so, rg_fields);
        std::fwrite(h.data(), 1, h.size(), out);
    }
    for (const std::string& s : outs) std::fwrite(s.data(), 1, s.size(), out);
    if (mode == 2) {
        std::vector<double> ab(refs.size(), 0.0);
        for (size_t t = 0; t < threads; ++t)
            for (size_t i = 0; i < refs.size(); ++i) ab[i] += abund[t][i];
        for (size_t i = 0; i < refs.size(); ++i) std::fprintf(out, "%s\t%.6f\n", refs[i].name.c_str(), ab[i] / (double)refs[i].seq.len);
    }
    if (out != stdout) std::fclose(out);
    if (!paf_out.empty()) {
        // the same chunks once more in mapping-only mode (fresh per-chunk state)
        Mapper mp_paf = m;
        mp_paf.mode = 1;
        std::vector<std::string> pafs(chunks.size());
        std::atomic<size_t> nxt{0};
        std::vector<std::thread> pth;
        for (size_t t = 0; t < threads; ++t)
            pth.emplace_back([&] {
                Mapper mm = mp_paf;
                for (;;) {
                    size_t c = nxt.fetch_add(1);
                    if (c >= chunks.size()) break;
                    mm.map_chunk(chunks[c], pafs[c]);
                }
            });
        for (auto& t : pth) t.join();
        FILE* pf = std::fopen(paf_out.c_str(), "wb");
        if (!pf) { std::fprintf(stderr, "error: cannot open %s\n", paf_out.c_str()); return 1; }
        for (const std::string& x : pafs) std::fwrite(x.data(), 1, x.size(), pf);
        std::fclose(pf);
    }
    if (time_it) {
        double ti = std::chrono::duration<double>(t1 - t0).count();
        double tm = std::chrono::duration<double>(t3 - t2).count();
        size_t n = r2.empty() ? r1.size() : std::min(r1.size(), r2.size());
        orc::StageTimes s;
        for (const auto& x : stages) {
            s.randstrobes += x.randstrobes; s.find_hits += x.find_hits; s.chaining += x.chaining;
            s.sort_nams += x.sort_nams; s.extend += x.extend;
        }
        std::fprintf(stderr,
                     "{\"index_s\": %.4f, \"map_s\": %.4f, \"records\": %zu, \"threads\": %zu, \"stages_s\": "
                     "{\"randstrobes\": %.4f, \"find_hits\": %.4f, \"chaining\": %.4f, \"sort_nams\": %.4f, "
                     "\"extend\": %.4f}, \"t1_s\": %.4f, \"t1_records\": %zu}\n",
                     ti, tm, n, threads, s.randstrobes, s.find_hits, s.chaining, s.sort_nams, s.extend, t1_s,
                     t1_records);
    }
    if (rng_stats) std::fprintf(stderr, "rng_draws %llu\n", (unsigned long long)orc::rng_draws().load());
    return 0;
}
