// The library's SSW code (ssw_align_raw: striped passes, banded traceback with
// packed direction cells) compiled for the host, for a CPU test against
// vectors written by the reference's ssw.c.
#include <cstdlib>
#include <vector>

#include "sa_extend.hpp"

extern "C" int ssw_raw_host(const int8_t* read, int32_t readLen, const int8_t* ref, int32_t refLen, const int8_t* mat,
                            int32_t n, int32_t score_size, int32_t gapO, int32_t gapE, uint32_t flag, uint32_t filters,
                            int32_t filterd, int32_t maskLen, uint32_t* cigar, uint32_t cigar_cap, int32_t* out) {
    std::vector<uint64_t> scratch((sa::ssw_raw_scratch(readLen, refLen) + 7) / 8 + 8);
    sa::SswRaw o;
    sa::ssw_align_raw(read, readLen, ref, refLen, mat, n, score_size, gapO, gapE, flag, filters, filterd, maskLen,
                      (uint8_t*)scratch.data(), scratch.size() * 8, cigar, cigar_cap, o);
    const int32_t v[10] = {o.score1, o.score2, o.ref_begin1, o.ref_end1, o.read_begin1, o.read_end1, o.ref_end2,
                           o.n_cigar, o.flag, o.status};
    for (int i = 0; i < 10; ++i) out[i] = v[i];
    return 0;
}
