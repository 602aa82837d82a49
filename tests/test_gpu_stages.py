"""The per-stage drop-ins on the device (through the C ABI): find_hits
(sa_find_hits_batch) against the oracle's find_hits for every MCS strategy,
the chaining stage on its own (sa_chain_batch) against the whole NAM stage, and
NAM lists for every MCS strategy against the oracle. Needs an MI355X."""
import ctypes
import os
import sys

import pytest

from conftest import ORACLE_BIN, ROOT

pytestmark = pytest.mark.gpu

MCS = {0: "always", 1: "rescue", 2: "off", 3: "first-strobe"}


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def sa():
    import strobealign_amd
    strobealign_amd.load()
    return strobealign_amd


@pytest.fixture(scope="module")
def ctx(sa):
    c = sa.Context(0)
    yield c
    c.close()


def fastq_seqs(path, n):
    with open(path, "rb") as f:
        lines = f.read().split(b"\n")
    return [lines[i] for i in range(1, len(lines), 4)][:n]


@pytest.fixture(scope="module")
def workloads(tmp_path_factory):
    sys.path.insert(0, ROOT)
    from strobealign_amd import synth
    d = str(tmp_path_factory.mktemp("stages"))
    small = synth.make("small", os.path.join(d, "small"), 400)
    pe = synth.make("pe100m", os.path.join(d, "pe"), 400)
    rep = os.path.join(d, "rep.fa")
    rep_reads = repeat_reference(rep, 11)
    # the 100 Mbp repetitive reference (both mates), and 120 copies of one
    # segment, whose hits exceed the filter cutoff and are rescued
    return [(small["ref"], fastq_seqs(small["reads1"], 400)),
            (pe["ref"], fastq_seqs(pe["reads1"], 200) + fastq_seqs(pe["reads2"], 200)),
            (rep, rep_reads)]


def repeat_reference(path, seed):
    """Two contigs: 120 copies (0.5% divergence, random spacers) of a 5 kbp
    segment, and 200 kbp of random sequence; returns 300 reads of 150 bp, 200
    from the copies and 100 from the random contig, 1% substitutions."""
    import numpy as np
    sys.path.insert(0, ROOT)
    from strobealign_amd import synth
    rng = np.random.default_rng(seed)
    seg = rng.integers(0, 4, 5000, dtype=np.uint8)
    parts = []
    for _ in range(120):
        c = seg.copy()
        m = rng.random(len(c)) < 0.005
        c[m] = (c[m] + rng.integers(1, 4, int(m.sum()), dtype=np.uint8)) % 4
        parts += [c, rng.integers(0, 4, int(rng.integers(10, 200)), dtype=np.uint8)]
    rep, rnd = np.concatenate(parts), rng.integers(0, 4, 200_000, dtype=np.uint8)
    synth.write_fasta(path, ["rep", "rnd"], [rep, rnd])
    reads = []
    for i in range(300):
        src = rep if i < 200 else rnd
        p = int(rng.integers(0, len(src) - 150))
        r = src[p:p + 150].copy()
        m = rng.random(150) < 0.01
        r[m] = (r[m] + 1) % 4
        if rng.random() < 0.5:
            r = (3 - r)[::-1]
        reads.append(b"ACGT"[0:0].join(bytes([b"ACGT"[x]]) for x in r))
    return reads


def oracle_hits(orc, fasta, seq, mcs):
    cap = max(1, len(seq))
    out = (ctypes.c_uint64 * (2 * cap * 7))()
    cnt = (ctypes.c_uint64 * 2)()
    det = (ctypes.c_uint64 * 14)()
    rc = orc.orc_find_hits(fasta.encode(), ctypes.c_uint64(150), seq, ctypes.c_uint64(len(seq)), mcs,
                           ctypes.c_uint64(100), ctypes.c_uint64(cap), out, cnt, det)
    assert rc == 0
    hits = [[tuple(out[(o * cap + i) * 7 + f] for f in range(7)) for i in range(cnt[o])] for o in range(2)]
    return hits, [tuple(det[o * 7 + f] for f in range(7)) for o in range(2)]


@pytest.mark.parametrize("mcs", [0, 1, 2, 3])
def test_find_hits_vs_oracle(sa, ctx, orc, workloads, mcs):
    """Hits (query span, index position, hashes, partial and filtered flags, in
    order) and the details counters of both orientations."""
    n_rescued = n_partial = 0
    for fasta, seqs in workloads:
        say("case: find_hits mcs=%s on %s (%d reads)" % (MCS[mcs], fasta, len(seqs)))
        with sa.Index(ctx, fasta, read_length=150, threads=8) as ix:
            hits, dets = ix.find_hits(seqs, mcs=mcs)
        for i, s in enumerate(seqs):
            eh, ed = oracle_hits(orc, fasta, s, mcs)
            for o in range(2):
                assert hits[2 * i + o] == [tuple(h) for h in eh[o]], (i, o)
                assert dets[2 * i + o] == ed[o], (i, o)
                n_rescued += ed[o][6]
                n_partial += sum(h[5] for h in eh[o])
    if mcs in (0, 1):
        assert n_partial > 0  # partial lookups exercised
    assert n_rescued > 0  # rescue of filtered hits exercised


@pytest.mark.parametrize("mcs", [0, 1, 2, 3])
def test_lookup_lanes_give_the_same_nams(sa, ctx, workloads, mcs):
    """The lookup stage's forms (one thread per read orientation making its
    lookups in turn, or 8, 16 or 64 lanes with a lane per randstrobe, rescue on
    the group) give the same NAM lists, f32 score bits included; the hits of
    the 16-lane default are checked against the oracle above."""
    for fasta, seqs in workloads:
        say("case: lookup lanes 1/8/16/64, mcs=%s on %s" % (MCS[mcs], fasta))
        with sa.Index(ctx, fasta, read_length=150, threads=8) as ix:
            res = {ln: ix.map_batch(seqs, sa.options(mcs=mcs, lookup_lanes=ln), speculative=0)[0]
                   for ln in (1, 8, 16, 64)}
        for ln in (1, 8, 64):
            assert res[ln] == res[16], ln
        assert sum(len(x) for x in res[16]) > len(seqs) // 2


def test_seed_lanes_give_the_same_nams(sa, ctx, workloads):
    """Seeding a thread per read and a wave per read (a lane per position, then
    per syncmer) give the same NAM lists, f32 score bits included; empty reads
    and reads shorter than the strobe window among them (their counts must be
    written by either form, or the lookups read stale ones)."""
    for fasta, seqs in workloads:
        say("case: seed lanes 1/64 on %s" % fasta)
        seqs = [b""] + seqs[:50] + [b"", b"ACGTACGT", b""] + seqs[50:] + [b""]
        with sa.Index(ctx, fasta, read_length=150, threads=8) as ix:
            res = {}
            for ln in (1, 64, 1):  # the thread form again, over buffers the wave form has filled
                res[ln] = ix.map_batch(seqs, sa.options(seed_lanes=ln), speculative=0)[0]
                assert res[ln][0] == [] and res[ln][-1] == []
        assert res[1] == res[64]
        assert sum(len(x) for x in res[64]) > len(seqs) // 2


def test_lookup_lanes_out_of_range(sa, ctx, workloads):
    fasta, seqs = workloads[0]
    with sa.Index(ctx, fasta, read_length=150) as ix:
        with pytest.raises(sa.SaError, match="lookup_lanes"):
            ix.map_batch(seqs[:4], sa.options(lookup_lanes=32), speculative=0)


@pytest.mark.parametrize("mcs", [0, 1, 2, 3])
def test_chain_stage_equals_nam_stage(sa, ctx, workloads, mcs):
    """sa_chain_batch over sa_find_hits_batch's output gives the NAM records of
    sa_find_nams_batch, byte for byte."""
    for fasta, seqs in workloads:
        say("case: chain stage mcs=%s on %s" % (MCS[mcs], fasta))
        with sa.Index(ctx, fasta, read_length=150, threads=8) as ix:
            hits, dets = ix.find_hits(seqs, mcs=mcs)
            got = ix.chain_raw(seqs, hits, dets)
            exp = ix.nams_raw(seqs, mcs=mcs)
        assert got[1] == exp[1]
        assert got[0] == exp[0]
        assert len(exp[0]) > 48 * len(seqs) // 2


@pytest.mark.parametrize("mcs", [1, 2, 3])
def test_nams_vs_oracle_other_mcs(sa, ctx, workloads, tmp_path, mcs):
    """NAM lists (f32 score bits included) for the MCS strategies other than the
    default, against the oracle's --dump-nams."""
    import subprocess
    fasta, seqs = workloads[1]
    say("case: NAM lists mcs=%s vs oracle" % MCS[mcs])
    fq = str(tmp_path / "r.fq")
    with open(fq, "wb") as f:
        for i, s in enumerate(seqs):
            f.write(b"@r%d\n%s\n+\n%s\n" % (i, s, b"?" * len(s)))
    dump = str(tmp_path / "o.nams")
    subprocess.run([ORACLE_BIN, "--mcs", MCS[mcs], "--dump-nams", dump, "-r", "150", fasta, fq], check=True)
    exp = [l.split("\t") for l in open(dump).read().splitlines()]
    with sa.Index(ctx, fasta, read_length=150, threads=8) as ix:
        got = ix.nams(seqs, mcs=mcs)
    flat = []
    for i, nams in enumerate(got):
        for n in nams:
            sb = ctypes.c_uint32.from_buffer(ctypes.c_float(n[8])).value
            flat.append([str(i)] + [str(x) for x in n[:8]] + ["%08x" % sb, str(n[9])])
    assert flat == exp


def oracle_spec(orc, fasta, seq, use_ssw, spec, scores=None):
    out = (ctypes.c_uint64 * (10 * 64))()
    cig = ctypes.create_string_buffer(1 << 16)
    sc = (ctypes.c_int * 5)(*scores) if scores else None
    m = orc.orc_speculative_extensions(fasta.encode(), ctypes.c_uint64(150), seq, int(use_ssw), ctypes.c_uint64(spec),
                                       sc, ctypes.c_uint64(64), out, cig, ctypes.c_uint64(1 << 16))
    assert m >= 0
    res, at = {}, 0
    for i in range(m):
        w = [out[10 * i + f] for f in range(10)]
        text = cig.raw[at:at + w[9]].decode()
        at += w[9]
        res[w[0]] = (w[1], w[2], w[3], w[4], w[5], w[6], w[7], w[8], text) if w[1] else (0,)
    return res


@pytest.mark.parametrize("use_ssw", [False, True])
def test_map_batch_speculative_extensions_vs_oracle(sa, ctx, orc, workloads, use_ssw):
    """sa_map_batch: the NAMs of sa_find_nams_batch and, for the first three
    consistent NAMs of each read, extend_seed's result (Hamming, piecewise or SSW;
    the CIGAR with =/X) equal the oracle's; the same on a stream of the caller."""
    L = "MIDNSHP=X"
    n_gapped = 0
    for fasta, seqs in workloads[:2]:
        say("case: sa_map_batch speculative extensions (ssw=%s) on %s" % (use_ssw, fasta))
        with sa.Index(ctx, fasta, read_length=150, threads=8) as ix:
            opts = sa.options(use_ssw=use_ssw)
            nams, exts = ix.map_batch(seqs, opts, speculative=3)
            assert [x for x in nams] == ix.nams(seqs)
            hip = ctypes.CDLL("libamdhip64.so.7")
            st = ctypes.c_void_p()
            assert hip.hipStreamCreate(ctypes.byref(st)) == 0
            try:
                nams2, exts2 = ix.map_batch(seqs, opts, speculative=3, stream=st.value)
            finally:
                hip.hipStreamDestroy(st)
            assert (nams2, exts2) == (nams, exts)
        for i, s in enumerate(seqs):
            exp = oracle_spec(orc, fasta, s, use_ssw, 3)
            got = {x: e for x, e in enumerate(exts[i]) if e is not None}
            assert sorted(got) == sorted(exp), i
            for x, e in got.items():
                if not e[0]:
                    assert exp[x] == (0,), (i, x)
                    continue
                cig = "".join("%d%s" % (o >> 4, L[o & 15]) for o in e[8])
                assert (1,) + e[1:8] + (cig,) == exp[x], (i, x)
                n_gapped += e[1]
    assert n_gapped > 0
