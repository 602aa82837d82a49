"""Write SSW input/output vectors from the reference's own ext/ssw/ssw.c.

The library is compiled by `make -C oracle ref` into oracle/_ref/libssw_ref.so
(in the development container only; the reference sources are not shipped).
Each vector holds a translated query/reference (codes 0..4), and ssw_align's
score1, ref_begin1, ref_end1, read_begin1, read_end1, flag and BAM cigar, called
with the arguments src/ssw.rs:156-191 uses (match 2, mismatch 8, gap 12/1,
flag 0x0f, mask_len max(len/2, 15)).

usage: python tests/golden/gen_ssw_vectors.py [N] > tests/golden/ssw_vectors.json
"""
import ctypes
import json
import os
import random
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "..", "..", "oracle", "_ref", "libssw_ref.so")


class SAlign(ctypes.Structure):
    _fields_ = [("score1", ctypes.c_uint16), ("score2", ctypes.c_uint16), ("ref_begin1", ctypes.c_int32),
                ("ref_end1", ctypes.c_int32), ("read_begin1", ctypes.c_int32), ("read_end1", ctypes.c_int32),
                ("ref_end2", ctypes.c_int32), ("cigar", ctypes.POINTER(ctypes.c_uint32)),
                ("cigarLen", ctypes.c_int32), ("flag", ctypes.c_uint16)]


MAT = bytes([(2 if (a == b and a < 4) else 256 - 8) for a in range(5) for b in range(5)])


def mutate(s, rng, sub, indel):
    out, i = [], 0
    while i < len(s):
        r = rng.random()
        if r < indel / 2:
            i += rng.randint(1, 3)
            continue
        if r < indel:
            out += [rng.randint(0, 3) for _ in range(rng.randint(1, 3))]
        c = s[i]
        if rng.random() < sub:
            c = (c + rng.randint(1, 3)) % 4
        out.append(c)
        i += 1
    return out


def cases(n, seed=20261203):
    rng = random.Random(seed)
    for _ in range(n):
        L = rng.choice([rng.randint(1, 40), rng.randint(40, 200), rng.randint(100, 320)])
        ref = [rng.randint(0, 3) for _ in range(L + rng.randint(0, 80))]
        st = rng.randint(0, max(0, len(ref) - L))
        q = mutate(ref[st:st + L], rng, rng.choice([0, 0.01, 0.05, 0.2, 0.6]), rng.choice([0, 0.01, 0.05, 0.1]))
        if rng.random() < 0.05:
            q = [4 if rng.random() < 0.1 else c for c in q]
        if q:
            yield q, ref


def main():
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 400
    lib = ctypes.CDLL(LIB)
    lib.ssw_init.restype = ctypes.c_void_p
    lib.ssw_init.argtypes = [ctypes.c_char_p, ctypes.c_int32, ctypes.c_char_p, ctypes.c_int32, ctypes.c_int8]
    lib.ssw_align.restype = ctypes.POINTER(SAlign)
    lib.ssw_align.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int32, ctypes.c_uint8, ctypes.c_uint8,
                              ctypes.c_uint8, ctypes.c_uint16, ctypes.c_int32, ctypes.c_int32]
    lib.init_destroy.argtypes = [ctypes.c_void_p]
    lib.align_destroy.argtypes = [ctypes.POINTER(SAlign)]
    out = []
    for q, ref in cases(n):
        qb, rb = bytes(q), bytes(ref)
        p = lib.ssw_init(qb, len(q), MAT, 5, 2)
        a = lib.ssw_align(p, rb, len(rb), 12, 1, 0x0F, 0, 2**31 - 1, max(len(q) // 2, 15))
        A = a.contents
        cig = [A.cigar[i] for i in range(A.cigarLen)] if A.cigar else []
        out.append({"q": "".join(map(str, q)), "r": "".join(map(str, ref)),
                    "res": [A.score1, A.ref_begin1, A.ref_end1, A.read_begin1, A.read_end1, A.flag], "cigar": cig})
        lib.align_destroy(a)
        lib.init_destroy(p)
    json.dump(out, sys.stdout, separators=(",", ":"))


if __name__ == "__main__":
    main()
