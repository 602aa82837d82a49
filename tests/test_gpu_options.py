"""Parity of the HIP path with the oracle on output options and inputs that
the workload configurations leave out: --details, -U, references with
lowercase and IUPAC/N bases, and reads longer than the lane-group SSW kernels
take (the one-thread extension). Needs an MI355X.

Single-end mapping of reads of 520-2000 bp has not finished on the device:
its first case produced no output for three minutes where the oracle takes
0.3 s, and no cause has been found in the code yet. The mapping passes refuse
reads above 512 bp with an error until one is, and the long-read parity cases
are skipped; reads of at most 512 bp are unaffected."""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import ORACLE_BIN, ROOT

pytestmark = pytest.mark.gpu


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


def oracle(*args):
    return subprocess.run([ORACLE_BIN, *args], check=True, capture_output=True).stdout


def wildcard_reference(src, dst):
    """The one-contig reference at src rewritten to dst with: a 2 kbp run of N,
    runs of 1-40 IUPAC codes every ~20 kbp, lowercase stretches, a second
    contig that is all N but a short ACGT island, and 61-column lines. The
    reference's reader keeps every byte it does not know as a wildcard, which
    the packed sequence turns into a position-seeded base (src/io/fasta.rs:
    116-165)."""
    with open(src, "rb") as f:
        seq = bytearray(b"".join(l.strip() for l in f if not l.startswith(b">")))
    rng = np.random.default_rng(5)
    seq[300_000:302_000] = b"N" * 2000
    codes = b"RYKMSWBDHVNn"
    for p in range(10_000, len(seq) - 100, 20_000):
        n = int(rng.integers(1, 41))
        seq[p:p + n] = bytes(codes[int(c)] for c in rng.integers(0, len(codes), n))
    for p in range(5_000, len(seq) - 3000, 50_000):
        seq[p:p + 3000] = seq[p:p + 3000].lower()
    second = b"N" * 5000 + bytes(seq[600_000:601_000]) + b"n" * 5000
    with open(dst, "wb") as f:
        for name, s in ((b"chr1 wildcards", bytes(seq)), (b"chrN", second)):
            f.write(b">" + name + b"\n")
            for i in range(0, len(s), 61):
                f.write(s[i:i + 61] + b"\n")


@pytest.fixture(scope="module")
def small(tmp_path_factory):
    sys.path.insert(0, ROOT)
    from strobealign_amd import synth
    d = str(tmp_path_factory.mktemp("options"))
    w = synth.make("small", os.path.join(d, "small"), 3000)
    wild = os.path.join(d, "wild.fa")
    wildcard_reference(w["ref"], wild)
    # single-end reads of 520-2000 bp from the unmodified reference
    longr = os.path.join(d, "long.fq")
    with open(w["ref"], "rb") as f:
        contig = np.frombuffer(b"".join(l.strip() for l in f if not l.startswith(b">")), dtype=np.uint8)
    code = {ord(c): i for i, c in enumerate("ACGT")}
    contigs = [np.array([code[c] for c in contig.tobytes()], dtype=np.uint8)]
    g = np.random.default_rng(77)
    synth.single_end(g, contigs, 400, g.integers(520, 2001, 400), 0.01, 0.002, longr, prefix="long")
    return {"ref": w["ref"], "wild": wild, "reads": w["reads1"], "long": longr}


# (oracle flags, map_files keywords)
OPTIONS = [((), {}),
           (("--details",), {"details": True}),
           (("-U",), {"only_mapped": True}),
           (("--details", "-U", "-N", "2"), {"details": True, "only_mapped": True, "max_secondary": 2}),
           (("--ssw", "--details"), {"use_ssw": True, "details": True})]
IDS = ["default", "details", "U", "details_U_N2", "ssw_details"]


@pytest.mark.parametrize("flags,kw", OPTIONS, ids=IDS)
def test_wildcard_reference_vs_oracle(sa, ctx, small, flags, kw):
    """3000 single-end 150 bp reads (1% substitutions, 0.1% indels) on the
    reference with wildcards, lowercase and an almost all-N contig; SAM byte
    for byte, the @SQ lengths counting every wildcard."""
    say("case: wildcard reference, 3000 SE reads, %s" % (" ".join(flags) or "default"))
    exp = oracle("--no-PG", *flags, small["wild"], small["reads"])
    with sa.Index(ctx, small["wild"], read_length=150) as ix:
        got = ix.map_files(small["reads"], **kw)
    assert b"@SQ\tSN:chrN\tLN:11000" in got
    assert got == exp
    if "only_mapped" in kw:
        assert all(int(l.split(b"\t")[1]) & 4 == 0 for l in got.splitlines() if not l.startswith(b"@"))
    if "details" in kw:
        assert b"\tna:i:" in got


LONG_READS_OPEN = ("single-end mapping of 520-2000 bp reads did not finish on the device; cause not yet found, "
                   "so the mapping passes refuse reads above 512 bp")


@pytest.mark.skip(reason=LONG_READS_OPEN)
@pytest.mark.parametrize("flags,kw", [((), {}), (("--ssw", "--details", "-U"),
                                                 {"use_ssw": True, "details": True, "only_mapped": True})],
                         ids=["default", "ssw_details_U"])
def test_long_reads_vs_oracle(sa, ctx, small, flags, kw):
    """400 single-end reads of 520-2000 bp (1% substitutions, 0.2% indels):
    queries past the lane-group SSW limit take the one-thread extension, on
    the plain reference and the one with wildcards."""
    for ref in (small["ref"], small["wild"]):
        say("case: long reads on %s, %s" % (os.path.basename(ref), " ".join(flags) or "default"))
        exp = oracle("--no-PG", *flags, ref, small["long"])
        with sa.Index(ctx, ref, read_length=150) as ix:
            got = ix.map_files(small["long"], **kw)
        recs = [l.split(b"\t") for l in got.splitlines() if not l.startswith(b"@")]
        assert sum(int(r[1]) & 4 == 0 for r in recs) > 300
        assert got == exp


def test_long_reads_are_refused(sa, ctx, small, tmp_path):
    """Reads above 512 bp are refused with an error before any mapping pass
    runs (single-end and paired-end, plain and fused), and the index then maps
    short reads as before."""
    say("case: reads above 512 bp refused by the mapping passes")
    with open(small["long"], "rb") as f:
        lines = f.read().splitlines(keepends=True)
    one = str(tmp_path / "one_long.fq")
    with open(one, "wb") as f:
        f.writelines(lines[:4])
    with sa.Index(ctx, small["ref"], read_length=150) as ix:
        for kw in ({}, {"fused": True}, {"use_ssw": True}, {"mode": 1}):
            with pytest.raises(sa.SaError, match="longer than 512 bp"):
                ix.map_files(one, **kw)
        with pytest.raises(sa.SaError, match="longer than 512 bp"):
            ix.map_files(one, one)
        exp = oracle("--no-PG", small["ref"], small["reads"])
        assert ix.map_files(small["reads"]) == exp


def test_paired_end_details_and_only_mapped_vs_oracle(sa, ctx, tmp_path_factory):
    """2x150 pairs on the 100 Mbp repetitive reference with --details and -U
    (the pair flags, rescue counters and mate fields of dropped records), and
    the same pairs with --details -N 2."""
    sys.path.insert(0, ROOT)
    from strobealign_amd import synth
    say("case: PE 2x150, 2000 pairs, --details -U and --details -N 2")
    w = synth.make("pe100m", str(tmp_path_factory.mktemp("pe_opts")), 2000)
    with sa.Index(ctx, w["ref"], read_length=150, threads=8) as ix:
        for flags, kw in ((("--details", "-U"), {"details": True, "only_mapped": True}),
                          (("--details", "-N", "2"), {"details": True, "max_secondary": 2})):
            exp = oracle("-t", "8", "--no-PG", *flags, w["ref"], w["reads1"], w["reads2"])
            got = ix.map_files(w["reads1"], w["reads2"], threads=8, **kw)
            assert b"\tna:i:" in got
            assert got == exp, flags
