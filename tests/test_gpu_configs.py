"""Parity of the HIP path with the oracle on the workload configurations not
covered elsewhere: non-default alignment scores (-A/-B/-O/-E/-L) and the
human-scale reference. Needs an MI355X."""
import os
import subprocess
import sys

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


def score_flags(sc):
    return ["-A", str(sc[0]), "-B", str(sc[1]), "-O", str(sc[2]), "-E", str(sc[3]), "-L", str(sc[4])]


# (match, mismatch, gap_open, gap_extend, end_bonus): a milder scheme with a
# smaller byte-pass bias, and gap_open == gap_extend (linear gaps), under which
# the SSW launches run one stripe count each and no job starts in the 16-bit pass
SCORES = [(1, 4, 6, 1, 5), (3, 5, 2, 2, 0)]


@pytest.mark.parametrize("sc", SCORES, ids=["A1B4O6E1L5", "A3B5O2E2L0"])
def test_nondefault_scores_vs_oracle(sa, ctx, tmp_path_factory, sc):
    """Scores flow into every extension kind: single-end piecewise and --ssw on
    the 1 Mbp reference with indels, paired-end SSW and rescue on the 100 Mbp
    repetitive reference, and the secondary-alignment dropoff (-N 3)
    2 * mismatch + gap_open (src/mapper.rs:509-511, 746)."""
    sys.path.insert(0, ROOT)
    from strobealign_amd import synth
    say("case: scores %s, 1 Mbp SE 3000 reads (piecewise, --ssw, -N 3)" % (sc,))
    d = str(tmp_path_factory.mktemp("scores"))
    small = synth.make("small", os.path.join(d, "small"), 3000)
    with sa.Index(ctx, small["ref"], read_length=150) as ix:
        for flags, kw in (((), {}), (("--ssw",), {"use_ssw": True}), (("-N", "3"), {"max_secondary": 3})):
            exp = oracle("--no-PG", *score_flags(sc), *flags, small["ref"], small["reads1"])
            got = ix.map_files(small["reads1"], scores=sc, **kw)
            assert got == exp, flags
    say("case: scores %s, 100 Mbp PE 3000 pairs (SSW, rescue, -N 2)" % (sc,))
    pe = synth.make("pe100m", os.path.join(d, "pe"), 3000)
    with sa.Index(ctx, pe["ref"], read_length=150, threads=8) as ix:
        for flags, kw in (((), {}), (("-N", "2"), {"max_secondary": 2})):
            exp = oracle("-t", "8", "--no-PG", *score_flags(sc), *flags, pe["ref"], pe["reads1"], pe["reads2"])
            got = ix.map_files(pe["reads1"], pe["reads2"], threads=8, scores=sc, **kw)
            assert got == exp, flags


def test_scores_out_of_range_are_rejected(sa, ctx, tmp_path_factory):
    """The reference takes u8 scores and negates them into i8 matrix entries
    (src/ssw.rs:143-144): a penalty above 127 is an argument error, not a
    silently wrapped matrix."""
    sys.path.insert(0, ROOT)
    from strobealign_amd import synth
    small = synth.make("small", str(tmp_path_factory.mktemp("scores_bad")), 10)
    with sa.Index(ctx, small["ref"], read_length=150) as ix:
        with pytest.raises(sa.SaError, match="0..127"):
            ix.map_files(small["reads1"], scores=(2, 200, 12, 1, 10))


@pytest.mark.timeout(1200)
def test_human_scale_vs_oracle(sa, ctx, tmp_path_factory):
    """Configuration 5: the seed-11 3.1 Gbp reference (24 contigs, 45% repeats
    of 300 bp-50 kbp segments at 0-15% divergence) with its full index in HBM
    (bits 28, about 620M 16-byte entries, built on the device), 2,000 pairs
    2x250 at 3% substitutions and 0.3% indels; SAM and -x PAF byte for byte
    against the oracle (src/mapper.rs:624-768, src/maponly.rs:114-193). The
    configuration's 10M pairs are a throughput figure (bench.py --workload
    human); 2,000 pairs keep the oracle's side of this test to about a minute."""
    sys.path.insert(0, ROOT)
    from strobealign_amd import synth
    d = str(tmp_path_factory.mktemp("human"))
    say("case: human scale, generating the 3.1 Gbp reference and 2000 pairs")
    w = synth.make_human(d, 2000)
    say("case: human scale, oracle SAM and -x (index on 16 host threads)")
    osam, opaf = os.path.join(d, "o.sam"), os.path.join(d, "o.paf")
    env = dict(os.environ, OMP_NUM_THREADS="16")
    subprocess.run([ORACLE_BIN, "-t", "16", "--no-PG", "-r", "250", "-o", osam, "--paf-out", opaf, w["ref"],
                    w["reads1"], w["reads2"]], check=True, env=env)
    say("case: human scale, device index and GPU mapping")
    with sa.Index(ctx, w["ref"], read_length=250, device_build=True) as ix:
        p, n, contigs = ix.params()
        assert (p.bits, contigs) == (28, 24) and n > 500_000_000
        got_sam = ix.map_files(w["reads1"], w["reads2"], threads=16)
        got_paf = ix.map_files(w["reads1"], w["reads2"], mode=1, threads=16)
    with open(osam, "rb") as f:
        exp_sam = f.read()
    with open(opaf, "rb") as f:
        exp_paf = f.read()
    recs = [l.split(b"\t") for l in got_sam.splitlines() if not l.startswith(b"@")]
    assert len(recs) == 4000 and any(b"I" in r[5] or b"D" in r[5] for r in recs)  # gapped alignments present
    assert got_sam == exp_sam
    assert got_paf == exp_paf
