"""Write ssw_align vectors over its whole argument space from the reference's
own ext/ssw/ssw.c (compiled by `make -C oracle ref` into
oracle/_ref/libssw_ref.so, in the development container only).

Each vector holds the ssw_init / ssw_align arguments (query and reference codes,
the n x n matrix, score_size, gap penalties, flag, filters, filterd, mask_len)
and every s_align field: score1, score2, ref_begin1, ref_end1, read_begin1,
read_end1, ref_end2, flag and the BAM cigar, or null where ssw_align returns
NULL. The cases vary the matrix (the reference's 2/-8 matrix, other
match/mismatch pairs, random 4x4 and 5x5 matrices), score_size 0/1/2, the gap
penalties, flag 0x0f/0x08/0x00/0x02/0x04/0x06/0x01 with filters and filterd,
and mask_len above and below 15.

usage: python tests/golden/gen_ssw_shim_vectors.py [N] > tests/golden/ssw_shim_vectors.json
"""
import ctypes
import json
import os
import random
import sys

from gen_ssw_vectors import LIB, SAlign, mutate


def matrices(rng):
    yield 5, [(2 if (a == b and a < 4) else -8) for a in range(5) for b in range(5)]
    m, x = rng.randint(1, 6), rng.randint(1, 9)
    yield 5, [(m if (a == b and a < 4) else -x) for a in range(5) for b in range(5)]
    yield 4, [(rng.randint(1, 5) if a == b else rng.randint(-6, -1)) for a in range(4) for b in range(4)]
    yield 5, [(rng.randint(1, 4) if a == b else rng.randint(-5, 0)) for a in range(5) for b in range(5)]


def cases(n, seed=20261204):
    rng = random.Random(seed)
    while n > 0:
        for nmat, mat in matrices(rng):
            L = rng.choice([rng.randint(1, 30), rng.randint(30, 160), rng.randint(150, 400)])
            ref = [rng.randint(0, 3) for _ in range(L + rng.randint(0, 120))]
            st = rng.randint(0, max(0, len(ref) - L))
            q = mutate(ref[st:st + L], rng, rng.choice([0, 0.02, 0.1, 0.4]), rng.choice([0, 0.02, 0.08]))
            if nmat == 5 and rng.random() < 0.1:
                q = [4 if rng.random() < 0.1 else c for c in q]
            if not q:
                continue
            flag, filters, filterd = rng.choice([(0x0F, 0, 2**31 - 1), (0x08, 0, 2**31 - 1), (0x00, 0, 0),
                                                 (0x02, rng.randint(0, 300), 0), (0x04, 0, rng.randint(0, 200)),
                                                 (0x06, rng.randint(0, 300), rng.randint(0, 200)),
                                                 (0x01, 0, 2**31 - 1)])
            yield dict(q=q, r=ref, n=nmat, mat=mat, score_size=rng.choice([2, 2, 2, 0, 1]),
                       gap_open=rng.choice([12, 5, 3, 8]), gap_extend=rng.choice([1, 2, 1]), flag=flag,
                       filters=filters, filterd=filterd,
                       mask_len=rng.choice([max(len(q) // 2, 15), 15, 40, 10]))
            n -= 1


def main():
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 600
    lib = ctypes.CDLL(LIB)
    lib.ssw_init.restype = ctypes.c_void_p
    lib.ssw_init.argtypes = [ctypes.c_char_p, ctypes.c_int32, ctypes.c_char_p, ctypes.c_int32, ctypes.c_int8]
    lib.ssw_align.restype = ctypes.POINTER(SAlign)
    lib.ssw_align.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int32, ctypes.c_uint8, ctypes.c_uint8,
                              ctypes.c_uint8, ctypes.c_uint16, ctypes.c_int32, ctypes.c_int32]
    lib.init_destroy.argtypes = [ctypes.c_void_p]
    lib.align_destroy.argtypes = [ctypes.POINTER(SAlign)]
    out = []
    for c in cases(n):
        qb, rb = bytes(c["q"]), bytes(c["r"])
        mb = bytes(x & 0xFF for x in c["mat"])
        p = lib.ssw_init(qb, len(qb), mb, c["n"], c["score_size"])
        a = lib.ssw_align(p, rb, len(rb), c["gap_open"], c["gap_extend"], c["flag"], c["filters"], c["filterd"],
                          c["mask_len"])
        res = None
        if a:
            A = a.contents
            res = [A.score1, A.score2, A.ref_begin1, A.ref_end1, A.read_begin1, A.read_end1, A.ref_end2, A.flag,
                   [A.cigar[i] for i in range(A.cigarLen)] if A.cigar else []]
            lib.align_destroy(a)
        lib.init_destroy(p)
        v = dict(c)
        v["q"] = "".join(map(str, c["q"]))
        v["r"] = "".join(map(str, c["r"]))
        v["res"] = res
        out.append(v)
    json.dump(out, sys.stdout, separators=(",", ":"))


if __name__ == "__main__":
    main()
