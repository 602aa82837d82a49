/* Link-compatible replacement for the reference's vendored SSW library.
 *
 * Same four entry points, argument meanings and s_align layout as the binding
 * the reference declares in src/ssw/raw.rs:1-51 (mirroring ext/ssw/ssw.h), so
 * src/ssw.rs can link libsa_hip instead of libssw.a unchanged. ssw_align runs
 * on the GPU (device 0, or the one named by the SA_DEVICE environment variable),
 * one alignment per call: this entry point is for conformance; batched callers
 * use sa_align_ssw_batch / sa_map_* of sa_hip.h. */
#ifndef SA_SSW_H
#define SA_SSW_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* src/ssw/raw.rs:3-17 */
typedef struct {
    uint16_t score1;     /* best score */
    uint16_t score2;     /* second-best score outside maskLen of the best end (0 if maskLen < 15) */
    int32_t ref_begin1;  /* 0-based start on the reference, -1 when not computed */
    int32_t ref_end1;    /* 0-based end on the reference */
    int32_t read_begin1; /* 0-based start on the query, -1 when not computed */
    int32_t read_end1;   /* 0-based end on the query */
    int32_t ref_end2;    /* end of the second best (-1 if maskLen < 15) */
    uint32_t* cigar;     /* len << 4 | op (0 M, 1 I, 2 D), or NULL */
    int32_t cigarLen;
    uint16_t flag;       /* 0 ok, 1 traceback failed, 2 the start may miss a small part */
} s_align;

/* src/ssw/raw.rs:19-24: opaque query profile */
typedef struct s_profile s_profile;

/* Query profile of `read` (codes 0..n-1, read_length of them) under the n x n
 * matrix `mat` (row: reference code, column: query code). score_size: 0 byte
 * scores only, 1 word scores only, 2 byte with word fallback. Both arrays are
 * copied. Replaces ssw_init (src/ssw/raw.rs:28-34, ext/ssw/ssw.c:789-806). */
s_profile* ssw_init(const int8_t* read, int32_t read_length, const int8_t* mat, int32_t n, int8_t score_size);

/* Releases a profile (src/ssw/raw.rs:36, ext/ssw/ssw.c:808-812). */
void init_destroy(s_profile* p);

/* Smith-Waterman of the profile's query against `refseq` (codes, ref_length
 * of them) with gap open/extend penalties; flag/filters/filterd/mask_length as
 * in ext/ssw/ssw.c:818-922 (bit 0-2 of flag: compute the start and the CIGAR;
 * flag 2 / bit 1: only when score1 >= filters; bit 2: only when both spans are
 * <= filterd). Returns NULL where the reference does (no profile for the scores
 * needed, e.g. byte scores saturate with score_size 0) or on a device error
 * (message through sa_last_error). Replaces ssw_align (src/ssw/raw.rs:38-48). */
s_align* ssw_align(const s_profile* prof, const int8_t* refseq, int32_t ref_length, uint8_t weight_gap_open,
                   uint8_t weight_gap_extend, uint8_t flag, uint16_t filters, int32_t filterd, int32_t mask_length);

/* Releases a result and its CIGAR (src/ssw/raw.rs:50, ext/ssw/ssw.c:924-927). */
void align_destroy(s_align* a);

#ifdef __cplusplus
}
#endif

#endif /* SA_SSW_H */
