"""FASTQ parsing on the device (sa_reads_load_device) against the host reader:
the same records (fields through the record accessor), the same errors, and
the staged sequences mapping to the same SAM as reads parsed on the host and
uploaded. Needs an MI355X."""
import gzip
import os
import sys

import pytest

from conftest import GOLDEN, ROOT
from test_reads_input import CASES

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


@pytest.fixture(scope="module")
def phix(sa, ctx):
    ix = sa.Index(ctx, os.path.join(GOLDEN, "phix.fasta"), read_length=150)
    yield ix
    ix.close()


def records(reads):
    return [reads.record(i, m) for i in range(len(reads)) for m in ((0, 1) if reads.paired else (0,))]


def write(tmp_path, name, data):
    p = str(tmp_path / name)
    with open(p, "wb") as f:
        f.write(data)
    return p


def test_edge_case_files_parse_alike(sa, phix, tmp_path):
    """Every host-reader case (CRLF, suffixes, comments, blanks, an empty last
    record, unterminated lines, FASTA, empty file) gives the same records."""
    say("case: device FASTQ parse, edge-case files")
    for i, (data, exp) in enumerate(CASES):
        p = write(tmp_path, "c%d.fq" % i, data)
        dev = sa.Reads(p, device=phix)
        assert records(dev) == records(sa.Reads(p)) == exp, i


@pytest.mark.parametrize("data,msg", [
    (b"@a\nACGT\n+\nIIII", "lengths differ"),
    (b"a\nA\n+\nI\n", "must start with '@'"),
    (b"@a\nA\n+\nI\n\n", "must start with '@'"),
    (b"@ x\nA\n+\nI\n", "identifier is empty"),
])
def test_errors_alike(sa, phix, tmp_path, data, msg):
    p = write(tmp_path, "e.fq", data)
    say("case: device FASTQ parse error %r" % msg)
    with pytest.raises(sa.SaError, match=msg):
        sa.Reads(p, device=phix)


def test_golden_reads_parse_and_map_alike(sa, phix):
    """phiX pairs and the tagged reads: same records, and the staged sequences
    map to the same SAM (with -C, so the comments reach the output)."""
    say("case: device FASTQ parse, phiX goldens")
    r1, r2 = os.path.join(GOLDEN, "phix.1.fastq"), os.path.join(GOLDEN, "phix.2.fastq")
    dev, host = sa.Reads(r1, r2, device=phix), sa.Reads(r1, r2)
    assert records(dev) == records(host) and len(dev) > 10
    assert dev.ingest_ms > 0
    host.upload(phix)
    o = sa.options(threads=4)
    assert phix.map_reads(dev, o)[0] == phix.map_reads(host, o)[0]
    t = os.path.join(GOLDEN, "phix.tags.fastq")
    dev, host = sa.Reads(t, device=phix), sa.Reads(t)
    assert records(dev) == records(host)
    host.upload(phix)
    o = sa.options(threads=4, fastq_comments=True)
    assert phix.map_reads(dev, o)[0] == phix.map_reads(host, o)[0]


def test_workload_parse_and_map_alike(sa, ctx, tmp_path):
    """20,000 pairs of the 100 Mbp workload, one file gzipped: same records and
    the same SAM as the host reader's staged reads."""
    sys.path.insert(0, ROOT)
    from strobealign_amd import synth
    say("case: device FASTQ parse, 100 Mbp PE 20000 pairs (reads2 gzipped)")
    w = synth.make("pe100m", str(tmp_path / "pe"), 20000)
    gz = str(tmp_path / "r2.fq.gz")
    with open(w["reads2"], "rb") as f, gzip.open(gz, "wb", compresslevel=1) as g:
        g.write(f.read())
    with sa.Index(ctx, w["ref"], read_length=150, threads=8) as ix:
        dev, host = sa.Reads(w["reads1"], gz, device=ix), sa.Reads(w["reads1"], w["reads2"])
        assert len(dev) == len(host) == 20000
        assert records(dev) == records(host)
        host.upload(ix)
        o = sa.options(threads=8)
        assert ix.map_reads(dev, o)[0] == ix.map_reads(host, o)[0]
