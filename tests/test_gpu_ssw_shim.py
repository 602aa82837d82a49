"""The link-compatible SSW entry points (include/ssw.h) against vectors written
by the reference's own ext/ssw/ssw.c over its argument space (matrices,
score_size, gaps, flag/filters/filterd, mask_len). Needs an MI355X."""
import ctypes
import json
import os

import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


class SAlign(ctypes.Structure):
    _fields_ = [("score1", ctypes.c_uint16), ("score2", ctypes.c_uint16), ("ref_begin1", ctypes.c_int32),
                ("ref_end1", ctypes.c_int32), ("read_begin1", ctypes.c_int32), ("read_end1", ctypes.c_int32),
                ("ref_end2", ctypes.c_int32), ("cigar", ctypes.POINTER(ctypes.c_uint32)),
                ("cigarLen", ctypes.c_int32), ("flag", ctypes.c_uint16)]


def shim():
    import strobealign_amd
    lib = strobealign_amd.load()
    lib.ssw_init.restype = ctypes.c_void_p
    lib.ssw_init.argtypes = [ctypes.c_char_p, ctypes.c_int32, ctypes.c_char_p, ctypes.c_int32, ctypes.c_int8]
    lib.ssw_align.restype = ctypes.POINTER(SAlign)
    lib.ssw_align.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int32, ctypes.c_uint8, ctypes.c_uint8,
                              ctypes.c_uint8, ctypes.c_uint16, ctypes.c_int32, ctypes.c_int32]
    lib.init_destroy.argtypes = [ctypes.c_void_p]
    lib.init_destroy.restype = None
    lib.align_destroy.argtypes = [ctypes.POINTER(SAlign)]
    lib.align_destroy.restype = None
    return lib


def test_ssw_shim_matches_reference_ssw_c():
    print("case: SSW link shim, 600 ssw.c vectors", flush=True)
    lib = shim()
    with open(os.path.join(GOLDEN, "ssw_shim_vectors.json")) as f:
        vecs = json.load(f)
    bad = []
    for i, v in enumerate(vecs):
        q = bytes(int(c) for c in v["q"])
        r = bytes(int(c) for c in v["r"])
        m = bytes(x & 0xFF for x in v["mat"])
        p = lib.ssw_init(q, len(q), m, v["n"], v["score_size"])
        assert p
        a = lib.ssw_align(p, r, len(r), v["gap_open"], v["gap_extend"], v["flag"], v["filters"], v["filterd"],
                          v["mask_len"])
        got = None
        if a:
            A = a.contents
            got = [A.score1, A.score2, A.ref_begin1, A.ref_end1, A.read_begin1, A.read_end1, A.ref_end2, A.flag,
                   [A.cigar[x] for x in range(A.cigarLen)] if A.cigar else []]
            lib.align_destroy(a)
        lib.init_destroy(p)
        if got != v["res"]:
            bad.append((i, got, v["res"]))
    assert not bad, bad[:3]
    assert sum(1 for v in vecs if v["res"] is None) > 10  # NULL returns covered
    assert sum(1 for v in vecs if v["res"] and v["res"][0] >= 255) > 50  # word fallback covered


def test_ssw_shim_under_user_scores():
    """The device SSW under the -A/-B/-O/-E matrices and gaps a user can pick
    (ssw.c vectors of tests/golden/gen_ssw_score_vectors.py): linear gaps and
    small openings, where the lazy-F loop decides scores, and the 16-bit pass."""
    print("case: SSW link shim, 500 ssw.c vectors under user scores", flush=True)
    lib = shim()
    with open(os.path.join(GOLDEN, "ssw_score_vectors.json")) as f:
        vecs = json.load(f)
    bad = []
    for i, v in enumerate(vecs):
        m, x, go, ge = v["sc"]
        q = bytes(int(c) for c in v["q"])
        r = bytes(int(c) for c in v["r"])
        mat = bytes(m if (a == b and a < 4) else (-x) & 0xFF for a in range(5) for b in range(5))
        p = lib.ssw_init(q, len(q), mat, 5, 2)
        assert p
        a = lib.ssw_align(p, r, len(r), go, ge, 0x0F, 0, 2**31 - 1, max(len(q) // 2, 15))
        assert a
        A = a.contents
        got = [A.score1, A.ref_begin1, A.ref_end1, A.read_begin1, A.read_end1]
        cig = [A.cigar[k] for k in range(A.cigarLen)] if A.cigar else []
        lib.align_destroy(a)
        lib.init_destroy(p)
        if (got, cig) != (v["res"], v["cigar"]):
            bad.append((i, v["sc"], got, v["res"]))
    assert not bad, bad[:3]
