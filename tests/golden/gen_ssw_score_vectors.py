"""Write SSW vectors under the alignment scores a user can pick (-A/-B/-O/-E),
from the reference's own ext/ssw/ssw.c (compiled by `make -C oracle ref` into
oracle/_ref/libssw_ref.so, in the development container only).

The matrix is the one src/ssw.rs:136-153 builds from Scores (match on the
diagonal of the 4 bases, -mismatch everywhere else, N included), the
call the one src/ssw.rs:156-191 makes (score_size 2, flag 0x0f, mask_len
max(len/2, 15)). Gap open/extend pairs include linear gaps (open == extend) and
small openings, where F runs across many stripes and the lazy-F loop's exit
test decides the score; long queries with high match scores drive the 16-bit
pass. Each vector: match, mismatch, gap_open, gap_extend, query and reference
codes, score1, ref_begin1, ref_end1, read_begin1, read_end1 and the BAM cigar.

usage: python tests/golden/gen_ssw_score_vectors.py [N] > tests/golden/ssw_score_vectors.json
"""
import ctypes
import json
import random
import sys

from gen_ssw_vectors import LIB, SAlign, mutate

GAPS = [(2, 2), (1, 1), (3, 1), (4, 2), (6, 1), (3, 3), (12, 1), (5, 5), (2, 1), (8, 4)]


def cases(n, seed=20261205):
    rng = random.Random(seed)
    for i in range(n):
        m, x = rng.randint(1, 4), rng.randint(1, 9)
        go, ge = GAPS[i % len(GAPS)]
        L = rng.choice([rng.randint(10, 60), rng.randint(60, 200), rng.randint(150, 300)])
        ref = [rng.randint(0, 3) for _ in range(L + rng.randint(0, 80))]
        st = rng.randint(0, max(0, len(ref) - L))
        q = mutate(ref[st:st + L], rng, rng.choice([0.01, 0.05, 0.15]), rng.choice([0.02, 0.05, 0.1]))
        if q:
            yield m, x, go, ge, q, ref


def main():
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 500
    lib = ctypes.CDLL(LIB)
    lib.ssw_init.restype = ctypes.c_void_p
    lib.ssw_init.argtypes = [ctypes.c_char_p, ctypes.c_int32, ctypes.c_char_p, ctypes.c_int32, ctypes.c_int8]
    lib.ssw_align.restype = ctypes.POINTER(SAlign)
    lib.ssw_align.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int32, ctypes.c_uint8, ctypes.c_uint8,
                              ctypes.c_uint8, ctypes.c_uint16, ctypes.c_int32, ctypes.c_int32]
    lib.init_destroy.argtypes = [ctypes.c_void_p]
    lib.align_destroy.argtypes = [ctypes.POINTER(SAlign)]
    out = []
    for m, x, go, ge, q, ref in cases(n):
        mat = bytes([m if (a == b and a < 4) else (-x) & 0xFF for a in range(5) for b in range(5)])
        qb, rb = bytes(q), bytes(ref)
        p = lib.ssw_init(qb, len(q), mat, 5, 2)
        a = lib.ssw_align(p, rb, len(rb), go, ge, 0x0F, 0, 2**31 - 1, max(len(q) // 2, 15))
        A = a.contents
        cig = [A.cigar[i] for i in range(A.cigarLen)] if A.cigar else []
        out.append({"sc": [m, x, go, ge], "q": "".join(map(str, q)), "r": "".join(map(str, ref)),
                    "res": [A.score1, A.ref_begin1, A.ref_end1, A.read_begin1, A.read_end1], "cigar": cig})
        lib.align_destroy(a)
        lib.init_destroy(p)
    json.dump(out, sys.stdout, separators=(",", ":"))


if __name__ == "__main__":
    main()
