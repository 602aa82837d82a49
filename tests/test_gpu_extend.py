"""sa_extend_batch: the aligner's entry points on the device (through the C ABI)
against the reference's known answers and the oracle. Needs an MI355X.

The known answers are the reference's own unit tests of the piecewise aligner
(src/piecewisealigner.rs:725-1711: global_alignment, xdrop_alignment in both
directions, extend_piecewise) and of hamming_align (src/aligner.rs:374-470);
the random cases compare every job kind with the oracle's functions."""
import random

import pytest

from test_oracle import EXTEND_PIECEWISE_KATS, GLOBAL_KATS, XDROP_KATS, ham, ssw

pytestmark = pytest.mark.gpu

L = "MIDNSHP=X"


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def ctx():
    import strobealign_amd
    strobealign_amd.load()
    c = strobealign_amd.Context(0)
    yield c
    c.close()


def cigar(ops):
    return "".join("%d%s" % (o >> 4, L[o & 15]) for o in ops)


def test_piecewise_kats_on_device(ctx):
    """global_alignment, xdrop_alignment (GlobalStartLocalEnd = mode 1,
    LocalStartGlobalEnd = mode 2) and extend_piecewise known answers."""
    say("case: block-aligner and extend_piecewise known answers on the device")
    G, X = ctx.EXT_GLOBAL, ctx.EXT_XDROP
    got = ctx.extend([(G, q, r, None, 0) for q, r, *_ in GLOBAL_KATS])
    for (q, r, score, qe, re_, cig), g in zip(GLOBAL_KATS, got):
        assert (g[0], g[1], g[2], g[3], g[4], g[5], cigar(g[7])) == (1, score, 0, qe, 0, re_, cig), (q, r)
    for q, r, mode, eb, exp, cig in XDROP_KATS:
        kind = X if mode == 1 else ctx.EXT_XDROP_REV
        g = ctx.extend([(kind, q, r, None, 0)], scores=(2, 8, 12, 1, eb))[0]
        assert [g[1], g[2], g[3], g[4], g[5]] == exp and cigar(g[7]) == cig, (q, r, mode)
    got = ctx.extend([(ctx.EXT_PIECEWISE, q, r, a, 5) for q, r, a, _ in EXTEND_PIECEWISE_KATS], k=5)
    for (q, r, a, exp), g in zip(EXTEND_PIECEWISE_KATS, got):
        if exp is None:
            assert g[0] == 0
        else:
            assert g[0] == 1 and ([g[1], g[2], g[3], g[4], g[5], g[6]], cigar(g[7])) == exp, (q, r)


def test_hamming_kats_on_device(ctx):
    """hamming_align known answers (src/aligner.rs:374-470). The device packs
    reference windows as 2-bit ACGT, so the cases with X/Y stand-ins compare an
    equivalent ACGT pair (a mismatch is a mismatch whatever the letters)."""
    say("case: hamming_align known answers on the device")
    cases = [
        (b"AACGGG", b"AATGGG", 1, 1, 0, ("2=1X3=", 1, 4, 0, 6)),
        (b"ACGGG", b"ATGGG", 1, 4, 0, ("3=", 0, 3, 2, 5)),
        (b"NAACCG", b"TAACCG", 3, 7, 0, ("5=", 0, 15, 1, 6)),
        (b"AACCGN", b"AACCGT", 3, 7, 0, ("5=", 0, 15, 0, 5)),
        (b"NAAAAAAAAAAAAAA", b"TAAAATTTTTTTTTT", 3, 7, 0, ("4=", 0, 12, 1, 5)),
        (b"", b"", 7, 5, 0, ("", 0, 0, 0, 0)),
    ]
    for q, r, m, mm, eb, exp in cases:
        g = ctx.extend([(ctx.EXT_HAMMING, q, r, None, 0)], scores=(m, mm, 12, 1, eb))[0]
        assert g[0] == 1 and (cigar(g[7]), g[6], g[1], g[2], g[3]) == exp, (q, r)
    # lengths differ: None
    assert ctx.extend([(ctx.EXT_HAMMING, b"ACGT", b"ACG", None, 0)])[0][0] == 0


def _mutate(rng, s, sub, indel):
    out = bytearray()
    for c in s:
        x = rng.random()
        if x < indel / 2:
            continue
        if x < indel:
            out.append(rng.choice(b"ACGT"))
        out.append(rng.choice(b"ACGT") if rng.random() < sub else c)
    return bytes(out)


def _anchors(q, r, k, rng):
    """Exact k-mer matches of q in r, increasing and non-overlapping in both,
    last first (a chain as the chainer hands it over)."""
    out, qi, rmin = [], 0, 0
    while qi + k <= len(q):
        p = r.find(q[qi:qi + k], rmin)
        if p >= 0:
            out.append((p, qi))
            qi += k + rng.randint(0, 8)
            rmin = p + k
        else:
            qi += 1
    return out[::-1]


@pytest.mark.parametrize("seed", [1, 2])
def test_extend_batch_random_vs_oracle(ctx, orc, seed):
    """Random (query, window) pairs of every job kind in one batch, against the
    oracle's global/x-drop alignment, extend_piecewise, hamming_align and
    Aligner::align; default and non-default scores."""
    say("case: sa_extend_batch random jobs vs oracle, seed %d" % seed)
    rng = random.Random(seed)
    scores = (2, 8, 12, 1, 10) if seed == 1 else (1, 4, 6, 1, 5)
    jobs, exp = [], []
    for i in range(600):
        kind = i % 6
        ql = rng.randint(1, 120)
        q = bytes(rng.choice(b"ACGT") for _ in range(ql))
        if kind == 4:  # hamming: same length, some mismatches
            r = _mutate(rng, q, 0.05, 0.0)[:ql].ljust(ql, b"A")
            jobs.append((ctx.EXT_HAMMING, q, r, None, 0))
            cig, ed, s, qs, qe = ham(orc, q, r, scores[0], scores[1], scores[4])
            exp.append((1, s, qs, qe, qs, qe, ed, cig))
            continue
        core = _mutate(rng, q, 0.04, 0.03)
        r = bytes(rng.choice(b"ACGT") for _ in range(rng.randint(0, 40))) + core + \
            bytes(rng.choice(b"ACGT") for _ in range(rng.randint(0, 40)))
        r = r or b"A"
        if kind in (0, 1, 2):  # global, x-drop both ways
            k = (ctx.EXT_GLOBAL, ctx.EXT_XDROP, ctx.EXT_XDROP_REV)[kind]
            jobs.append((k, q, r, None, 0))
            out, cig = _pw_scores(orc, q, r, kind, scores)
            exp.append((1, out[0], out[1], out[2], out[3], out[4], None, cig))
        elif kind == 3:  # piecewise over an exact-match chain
            an = _anchors(q, r, 5, rng)
            if not an:
                an = [(0, 0)]
            jobs.append((ctx.EXT_PIECEWISE, q, r, an, 5))
            e = _ext_pw_scores(orc, q, r, an, scores)
            exp.append(None if e is None else (1, e[0][0], e[0][1], e[0][2], e[0][3], e[0][4], e[0][5], e[1]))
        else:  # SSW (default scores only: the aligner's matrix)
            jobs.append((ctx.EXT_SSW, q, r[:2000], None, 0))
            e = ssw(orc, q, r[:2000]) if scores == (2, 8, 12, 1, 10) else None
            exp.append(("ssw", e))
    got = ctx.extend(jobs, scores=scores, k=5)
    n_checked = 0
    for idx, (j, e, g) in enumerate(zip(jobs, exp, got)):
        if isinstance(e, tuple) and e and e[0] == "ssw":
            if scores != (2, 8, 12, 1, 10):
                continue
            if e[1] is None:
                assert g[0] == 0, idx
            else:
                out, cig = e[1]
                assert [g[1], g[2], g[3], g[4], g[5], g[6]] == list(out[:6]) and cigar(g[7]) == cig, idx
            n_checked += 1
            continue
        if e is None:
            assert g[0] == 0, (idx, j[0])
            n_checked += 1
            continue
        assert g[0] == 1, (idx, j[0])
        assert (g[1], g[2], g[3], g[4], g[5]) == tuple(e[1:6]), (idx, j[0], g, e)
        if e[6] is not None:
            assert g[6] == e[6], (idx, j[0])
        assert cigar(g[7]) == e[7], (idx, j[0], cigar(g[7]), e[7])
        n_checked += 1
    assert n_checked >= 500


def _pw_scores(orc, q, r, mode, scores):
    import ctypes
    sc = (ctypes.c_int * 5)(*scores)
    out = (ctypes.c_int64 * 5)()
    buf = ctypes.create_string_buffer(4096)
    orc.orc_pw_align(q, ctypes.c_uint64(len(q)), r, ctypes.c_uint64(len(r)), sc, mode, 500, out, buf,
                     ctypes.c_uint64(4096))
    return list(out), buf.value.decode()


def _ext_pw_scores(orc, q, r, anchors, scores):
    import ctypes
    arr = (ctypes.c_uint64 * (2 * len(anchors)))(*[x for a in anchors for x in a])
    sc = (ctypes.c_int * 5)(*scores)
    out = (ctypes.c_int64 * 6)()
    buf = ctypes.create_string_buffer(4096)
    ok = orc.orc_extend_piecewise(q, ctypes.c_uint64(len(q)), r, ctypes.c_uint64(len(r)), arr,
                                  ctypes.c_int64(len(anchors)), ctypes.c_uint64(5), sc, ctypes.c_uint64(5), 500, out,
                                  buf, ctypes.c_uint64(4096))
    return (list(out), buf.value.decode()) if ok else None


def test_block_aligner_small_cases_vs_oracle(ctx, orc):
    """global and x-drop alignment on the device over the small cases whose
    oracle scores the brute-force checks in test_oracle.py establish (every
    alignment enumerated): same score, ends and CIGAR."""
    import itertools
    say("case: block-aligner small cases on the device vs oracle")
    rng = random.Random(5)
    cases = [(bytes(a), bytes(b)) for m in range(1, 6) for n in range(1, 6)
             for a in itertools.product(b"AC", repeat=m) for b in [rng.choice(list(itertools.product(b"AC", repeat=n)))]]
    cases += [(bytes(rng.choice(b"ACGT") for _ in range(rng.randint(1, 7))),
               bytes(rng.choice(b"ACGT") for _ in range(rng.randint(1, 7)))) for _ in range(300)]
    kinds = ((ctx.EXT_GLOBAL, 0), (ctx.EXT_XDROP, 1), (ctx.EXT_XDROP_REV, 2))
    for scores in ((2, 8, 12, 1, 10), (3, 2, 3, 1, 0), (1, 4, 6, 1, 5)):
        jobs = [(kind, q, r, None, 0) for q, r in cases for kind, _ in kinds]
        got = ctx.extend(jobs, scores=scores)
        at = 0
        for q, r in cases:
            for _, mode in kinds:
                g = got[at]
                at += 1
                out, cig = _pw_scores(orc, q, r, mode, scores)
                assert (g[0], g[1], g[2], g[3], g[4], g[5], cigar(g[7])) == (1, out[0], out[1], out[2], out[3], out[4],
                                                                             cig), (q, r, mode, scores)
