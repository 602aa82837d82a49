"""The fused batch pass (sa_map_options.fused): seeding -> lookups -> chaining ->
NAMs -> the speculative extensions as one recorded graph, replayed per batch.
Its output must be the oracle's byte for byte on every workload shape (single-
and paired-end, piecewise and SSW extension, fixed and variable read length,
non-default scores, several engines), including batches whose anchors exceed
the pass's fixed buffers and run the plain stages instead. Small batches make
each run go through the engine's first (plain) batch, the pass run directly,
its recording and its replays. Needs an MI355X."""
import os
import random
import subprocess
import sys

import pytest

from conftest import ORACLE_BIN, ROOT

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def oracle(*args):
    return subprocess.run([ORACLE_BIN, *args], check=True, capture_output=True).stdout


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


def synth():
    sys.path.insert(0, ROOT)
    from strobealign_amd import synth as s
    return s


def run(sa, ix, r1, r2=None, **kw):
    reads = sa.Reads(r1, r2)
    try:
        return ix.map_reads(reads, sa.options(threads=8, **kw))
    finally:
        reads.close()


def check_fused(st, batches_at_least=3):
    """The run went through the pass run directly, its recording and at least
    one replay, and the decisions used its extensions."""
    assert st.fused_batches >= batches_at_least, st.fused_batches
    assert st.graph_replays >= 1
    assert 0 < st.spec_used <= st.spec_requests


def test_fused_paired_end_100m(sa, ctx, tmp_path_factory):
    """100 Mbp repetitive reference, 3000 pairs 2x150, 100 kbase batches on one
    and on three engines, seeding a wave per read (default) and a thread per
    read (which then also seeds the empty reads that pad each batch)."""
    w = synth().make("pe100m", str(tmp_path_factory.mktemp("fpe")), 3000)
    exp = oracle("-t", "8", "--no-PG", "--chunk-size=20000", w["ref"], w["reads1"], w["reads2"])
    with sa.Index(ctx, w["ref"], read_length=150, threads=8) as ix:
        for slots, seed_lanes in ((1, 0), (3, 0), (1, 1), (3, 1)):
            say("case: fused pass, 100 Mbp PE 3000 pairs, %d engine(s), seed lanes %d" % (slots, seed_lanes))
            out, st = run(sa, ix, w["reads1"], w["reads2"], fused=True, batch_bases=100_000, slots=slots,
                          chunk_size=20_000, seed_lanes=seed_lanes)
            assert out == exp, (slots, seed_lanes)
            check_fused(st)
            assert st.fused_overflows == 0


def test_fused_single_end_small(sa, ctx, tmp_path_factory):
    """Configuration 1 (1 Mbp, 10,000 SE 150 bp reads with indels): piecewise
    (default) and SSW (--ssw) extension from the fused pass."""
    w = synth().make("small", str(tmp_path_factory.mktemp("fsmall")), 10_000)
    with sa.Index(ctx, w["ref"], read_length=150) as ix:
        for flags, kw in (((), {}), (("--ssw",), {"use_ssw": True})):
            say("case: fused pass, 1 Mbp SE 10000 reads %s" % (flags or "piecewise",))
            exp = oracle("--no-PG", "--chunk-size=30000", *flags, w["ref"], w["reads1"])
            out, st = run(sa, ix, w["reads1"], fused=True, batch_bases=150_000, chunk_size=30_000, slots=1, **kw)
            assert out == exp, flags
            check_fused(st)


def test_fused_single_end_variable_length(sa, ctx, tmp_path_factory):
    """Configuration 3's read shape: 3000 SE reads of 50-500 bp (15% from
    repeats) on the 100 Mbp reference, r250 profile; piecewise and --ssw
    (queries up to 500 bp through every stripe count of the lane groups)."""
    w = synth().make("se_var", str(tmp_path_factory.mktemp("fsevar")), 3000)
    with open(w["reads1"], "rb") as f:
        lines = f.read().split(b"\n")
    lens = [len(lines[i]) for i in range(1, len(lines), 4)]
    est = sum(lens[:500]) // len(lens[:500])
    with sa.Index(ctx, w["ref"], read_length=est, threads=8) as ix:
        for flags, kw in (((), {}), (("--ssw",), {"use_ssw": True})):
            say("case: fused pass, 100 Mbp SE 3000 reads of 50-500 bp %s" % (flags or "piecewise",))
            exp = oracle("-t", "8", "--no-PG", "--chunk-size=30000", *flags, w["ref"], w["reads1"])
            out, st = run(sa, ix, w["reads1"], fused=True, batch_bases=60_000, chunk_size=30_000, slots=2, **kw)
            assert out == exp, flags
            check_fused(st)


def test_fused_non_default_scores(sa, ctx, tmp_path_factory):
    """-A 1 -B 4 -O 6 -E 1 -L 5 on paired-end and single-end reads."""
    s = synth()
    w = s.make("pe100m", str(tmp_path_factory.mktemp("fsc")), 2000)
    scores = (1, 4, 6, 1, 5)
    flags = ("-A", "1", "-B", "4", "-O", "6", "-E", "1", "-L", "5")
    with sa.Index(ctx, w["ref"], read_length=150, threads=8) as ix:
        say("case: fused pass, non-default scores, PE and SE")
        exp = oracle("-t", "8", "--no-PG", "--chunk-size=20000", *flags, w["ref"], w["reads1"], w["reads2"])
        out, st = run(sa, ix, w["reads1"], w["reads2"], fused=True, batch_bases=100_000, chunk_size=20_000,
                      scores=scores, slots=1)
        assert out == exp
        check_fused(st)
        exp = oracle("-t", "8", "--no-PG", "--chunk-size=15000", *flags, w["ref"], w["reads1"])
        out, st = run(sa, ix, w["reads1"], fused=True, batch_bases=60_000, chunk_size=15_000, scores=scores, slots=1)
        assert out == exp
        check_fused(st)


def test_fused_anchor_overflow_falls_back(sa, ctx, tmp_path_factory):
    """The engine's first batch is mostly reads from nowhere in the reference,
    so the fused pass is sized for far fewer anchors per base than the reads
    after it have: those batches overflow the anchor buffer, run the plain
    stages, and the pass is resized; the SAM is the oracle's throughout."""
    s = synth()
    d = str(tmp_path_factory.mktemp("fover"))
    w = s.make("small", d, 4000)
    rng = random.Random(5)
    mixed = os.path.join(d, "mixed.fq")
    with open(w["reads1"], "rb") as f:
        real = f.read().split(b"\n")
    with open(mixed, "wb") as f:
        for i in range(1200):
            if i % 100 == 0:  # a few real reads, so the first batch has some anchors
                f.write(b"\n".join(real[4 * i:4 * i + 4]) + b"\n")
            else:
                seq = bytes(rng.choice(b"ACGT") for _ in range(150))
                f.write(b"@nowhere%d\n%s\n+\n%s\n" % (i, seq, b"?" * 150))
        f.write(b"\n".join(real[4 * 1200:]).rstrip(b"\n") + b"\n")
    say("case: fused pass, anchor overflow after a batch of unmappable reads")
    exp = oracle("--no-PG", "--chunk-size=30000", w["ref"], mixed)
    with sa.Index(ctx, w["ref"], read_length=150) as ix:
        out, st = run(sa, ix, mixed, fused=True, batch_bases=150_000, chunk_size=30_000, slots=1)
    assert out == exp
    assert st.fused_overflows >= 1
    assert st.fused_batches >= 2


def test_fused_human_scale_sample(sa, ctx, tmp_path_factory):
    """2x250 pairs at 3% substitutions and 0.3% indels on a 50 Mbp reference
    from the human-scale generator (45% repeats of 300 bp-50 kbp at 0-15%
    divergence; the full 3.1 Gbp run is the human-scale test's): extension-heavy
    paired-end reads through the fused pass."""
    w = synth().make_human_small(str(tmp_path_factory.mktemp("fhum")), 2000)
    exp = oracle("-t", "8", "--no-PG", "--chunk-size=50000", w["ref"], w["reads1"], w["reads2"])
    with sa.Index(ctx, w["ref"], read_length=250, threads=8) as ix:
        say("case: fused pass, human-like 50 Mbp, 2000 pairs 2x250")
        out, st = run(sa, ix, w["reads1"], w["reads2"], fused=True, batch_bases=200_000, chunk_size=50_000, slots=1)
    assert out == exp
    check_fused(st)
