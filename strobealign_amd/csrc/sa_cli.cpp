// strobealign-hip: command-line drop-in for the mapping path (the subset of
// src/main.rs:61-283 flags that the path uses). Index construction runs on the
// host; seeding, lookups, chaining and extension run on the GPU.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include "sa_host.hpp"

using namespace sa;

static void usage() {
    std::fprintf(stderr,
                 "usage: strobealign-hip [-t N] [-x|--aemb] [-r N] [-N N] [-U] [--eqx] [--ssw] [-C] [--details]\n"
                 "                       [--rg-id ID] [--rg TAG] [--chunk-size N] [--mcs MODE] [-o OUT]\n"
                 "                       [-A N] [-B N] [-O N] [-E N] [-L N] [--xdrop N]\n"
                 "                       [--device D] [--batch-bases N] [--slots N] [--dump-nams PATH] [--time]\n"
                 "                       ref.fa reads1 [reads2]\n");
}

static double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int main(int argc, char** argv) {
    Options o;
    uint64_t read_length = 0;
    int device = 0;
    bool time_it = false;
    std::string out_path, dump_nams;
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
        if (a == "-t") o.threads = std::stoi(val("-t"));
        else if (a == "-x") o.mode = 1;
        else if (a == "--aemb") o.mode = 2;
        else if (a == "-r") read_length = std::stoull(val("-r"));
        else if (a == "-N") o.max_secondary = std::stoull(val("-N"));
        else if (a == "-U") o.only_mapped = true;
        else if (a == "--eqx") o.eqx = true;
        else if (a == "--ssw") o.use_ssw = true;
        else if (a == "-C") o.fastq_comments = true;
        else if (a == "--details") o.details = true;
        else if (a == "--no-PG") {}
        else if (a == "--time") time_it = true;
        else if (is("--rg-id")) { o.has_rg = true; o.rg_id = val("--rg-id"); }
        else if (is("--rg")) o.rg_fields.push_back(val("--rg"));
        else if (is("--chunk-size")) o.chunk_size = std::stoull(val("--chunk-size"));
        else if (is("--batch-bases")) o.batch_bases = std::stoull(val("--batch-bases"));
        else if (is("--slots")) o.slots = std::stoi(val("--slots"));
        else if (is("--device")) device = std::stoi(val("--device"));
        else if (is("--dump-nams")) dump_nams = val("--dump-nams");
        else if (is("--mcs")) {
            std::string m = val("--mcs");
            o.mcs = m == "rescue" ? 1 : m == "off" ? 2 : m == "first-strobe" ? 3 : 0;
        } else if (a == "-A") o.scores.match = std::stoi(val("-A"));
        else if (a == "-B") o.scores.mismatch = std::stoi(val("-B"));
        else if (a == "-O") o.scores.gap_open = std::stoi(val("-O"));
        else if (a == "-E") o.scores.gap_extend = std::stoi(val("-E"));
        else if (a == "-L") o.scores.end_bonus = std::stoi(val("-L"));
        else if (is("--xdrop")) o.scores.xdrop = std::stoi(val("--xdrop"));
        else if (a == "-o") out_path = val("-o");
        else if (!a.empty() && a[0] == '-' && a.size() > 1) { usage(); return 2; }
        else pos.push_back(a);
    }
    if (!o.rg_fields.empty() && !o.has_rg) { o.has_rg = true; o.rg_id = "1"; }
    if (pos.size() < 2) { usage(); return 2; }
    try {
        std::string err;
        std::vector<Record> r1, r2;
        if (!read_records(pos[1], r1, err)) throw std::runtime_error(err);
        if (pos.size() > 2 && !read_records(pos[2], r2, err)) throw std::runtime_error(err);
        if (read_length == 0) read_length = estimate_read_length(r1);
        SeedParams sp = seed_params_for_read_length(read_length);
        double t0 = now_s();
        PackedRef ref = read_fasta(pos[0]);
        int bits = pick_bits(ref, sp);
        HostIndex idx;
        build_index(ref, sp, bits, 0.0002, std::max(1, o.threads), idx);
        double t1 = now_s();
        Engine eng(device);
        eng.upload_reference(ref.words.data(), ref.words.size(), ref.word_off.data(), ref.len.data(),
                             (uint32_t)ref.names.size());
        eng.upload_index(idx.entries.data(), idx.entries.size(), idx.bucket_starts.data(), idx.p);
        double t2 = now_s();
        if (!dump_nams.empty()) {
            // NAMs of every read (stable-sorted, before the per-chunk shuffle), one line each
            std::vector<uint8_t> seq;
            std::vector<uint64_t> off{0};
            for (size_t i = 0; i < r1.size(); ++i) {
                seq.insert(seq.end(), r1[i].seq.begin(), r1[i].seq.end());
                off.push_back(seq.size());
            }
            eng.upload_reads(seq.data(), off.data(), (uint32_t)r1.size());
            NamBatch nb;
            MapParams mp{o.mcs, o.rescue_distance};
            ChainParams cp{50, 0.1f, 0.05f, 0.7f, -1, 0.01f, 10.0f};
            eng.find_nams(mp, cp, nb);
            FILE* f = std::fopen(dump_nams.c_str(), "wb");
            for (size_t i = 0; i < r1.size(); ++i)
                for (uint64_t x = nb.nam_off[i]; x < nb.nam_off[i + 1]; ++x) {
                    const NamRec& n = nb.nams[x];
                    uint32_t sb;
                    std::memcpy(&sb, &n.score, 4);
                    std::fprintf(f, "%zu\t%u\t%u\t%u\t%u\t%u\t%u\t%u\t%u\t%08x\t%u\n", i, n.nam_id, n.ref_id,
                                 n.ref_start, n.ref_end, n.query_start, n.query_end, n.matching_bases, n.n_anchors, sb,
                                 n.flags & 1);
                }
            std::fclose(f);
            return 0;
        }
        std::string out;
        RunStats st;
        std::vector<std::unique_ptr<Engine>> extra;
        std::vector<Engine*> engs{&eng};
        for (int i = 1; i < o.slots; ++i) {
            extra.push_back(std::make_unique<Engine>(Engine::ShareIndex{}, eng));
            engs.push_back(extra.back().get());
        }
        map_records(engs, ref, r1, r2, o, &out, &st, nullptr);
        double t3 = now_s();
        FILE* f = out_path.empty() ? stdout : std::fopen(out_path.c_str(), "wb");
        if (!f) throw std::runtime_error("cannot open " + out_path);
        if (o.mode == 0) {
            std::string h = sam_header(ref, o);
            std::fwrite(h.data(), 1, h.size(), f);
        }
        std::fwrite(out.data(), 1, out.size(), f);
        if (f != stdout) std::fclose(f);
        if (time_it)
            std::fprintf(stderr,
                         "{\"index_s\": %.4f, \"upload_s\": %.4f, \"map_s\": %.4f, \"records\": %llu, \"nams\": %llu, "
                         "\"requests\": %llu, \"rounds\": %llu, \"nam_ms\": %.2f, \"extend_ms\": %.2f, "
                         "\"host_ms\": %.2f, \"reads_upload_ms\": %.2f}\n",
                         t1 - t0, t2 - t1, t3 - t2, (unsigned long long)st.records, (unsigned long long)st.nams,
                         (unsigned long long)st.requests, (unsigned long long)st.rounds, st.nam_ms, st.extend_ms,
                         st.host_ms, st.upload_ms);
        if (time_it) {
            // per-kernel totals: milliseconds, launches, work items
            std::string k = "{";
            std::vector<std::pair<std::string, Engine::KernelStat>> all;
            for (Engine* e : engs)
                for (const auto& ks : e->kernel_stats()) {
                    bool found = false;
                    for (auto& a : all)
                        if (a.first == ks.first) {
                            a.second.ms += ks.second.ms;
                            a.second.launches += ks.second.launches;
                            a.second.items += ks.second.items;
                            found = true;
                        }
                    if (!found) all.push_back(ks);
                }
            for (const auto& ks : all) {
                char b[160];
                std::snprintf(b, sizeof b, "%s\"%s\": [%.3f, %llu, %llu]", k.size() > 1 ? ", " : "", ks.first.c_str(),
                              ks.second.ms, (unsigned long long)ks.second.launches,
                              (unsigned long long)ks.second.items);
                k += b;
            }
            std::fprintf(stderr, "%s}\n", k.c_str());
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
