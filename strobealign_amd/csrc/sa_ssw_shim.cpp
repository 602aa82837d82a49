// extern "C" SSW entry points declared in include/ssw.h: the reference's SSW
// binding (src/ssw/raw.rs) served by the device kernel k_ssw_raw.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "sa_engine.hpp"
#include "sa_extend.hpp"
#include "ssw.h"

struct s_profile {
    std::vector<int8_t> read, mat;
    int32_t n = 0;
    int8_t score_size = 0;
};

namespace sa {
void set_last_error(const std::string& msg);
}

namespace {
int shim_device() {
    const char* d = std::getenv("SA_DEVICE");
    return d ? std::atoi(d) : 0;
}
}  // namespace

extern "C" {

s_profile* ssw_init(const int8_t* read, int32_t read_length, const int8_t* mat, int32_t n, int8_t score_size) {
    if (read_length < 0 || n <= 0 || !mat || (read_length > 0 && !read)) return nullptr;
    try {
        s_profile* p = new s_profile;
        p->read.assign(read, read + read_length);
        p->mat.assign(mat, mat + (size_t)n * (size_t)n);
        p->n = n;
        p->score_size = score_size;
        return p;
    } catch (const std::exception&) {
        return nullptr;
    }
}

void init_destroy(s_profile* p) { delete p; }

s_align* ssw_align(const s_profile* prof, const int8_t* refseq, int32_t ref_length, uint8_t weight_gap_open,
                   uint8_t weight_gap_extend, uint8_t flag, uint16_t filters, int32_t filterd, int32_t mask_length) {
    if (!prof || ref_length < 0 || (ref_length > 0 && !refseq)) return nullptr;
    if (mask_length < 15)
        std::fprintf(stderr, "ssw_align: mask_length < 15, so no second-best alignment is reported.\n");
    try {
        sa::SswRaw r;
        std::vector<uint32_t> cig;
        sa::ssw_raw_device(shim_device(), prof->read.data(), (int32_t)prof->read.size(), refseq, ref_length,
                           prof->mat.data(), prof->n, prof->score_size, weight_gap_open, weight_gap_extend, flag,
                           filters, filterd, mask_length, r, cig);
        if (r.status == 1) return nullptr;
        if (r.status != 0) {
            sa::set_last_error(r.status == 2 ? "ssw_align: device scratch too small" : "ssw_align: CIGAR buffer too small");
            return nullptr;
        }
        s_align* a = (s_align*)std::calloc(1, sizeof(s_align));
        if (!a) return nullptr;
        a->score1 = (uint16_t)r.score1;
        a->score2 = (uint16_t)r.score2;
        a->ref_begin1 = r.ref_begin1;
        a->ref_end1 = r.ref_end1;
        a->read_begin1 = r.read_begin1;
        a->read_end1 = r.read_end1;
        a->ref_end2 = r.ref_end2;
        a->flag = (uint16_t)r.flag;
        if (!cig.empty()) {
            a->cigar = (uint32_t*)std::malloc(cig.size() * 4);
            if (!a->cigar) {
                std::free(a);
                return nullptr;
            }
            std::memcpy(a->cigar, cig.data(), cig.size() * 4);
            a->cigarLen = (int32_t)cig.size();
        }
        return a;
    } catch (const std::exception& e) {
        sa::set_last_error(e.what());
        return nullptr;
    }
}

void align_destroy(s_align* a) {
    if (!a) return;
    std::free(a->cigar);
    std::free(a);
}

}  // extern "C"
