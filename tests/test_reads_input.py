"""Read-file input on the host (no GPU needed): the FASTQ semantics of
src/io/fastq.rs:22-95 (the lines as BufRead::read_line / read_until return
them), FASTA reads (src/io/fasta.rs:60-113), header splitting
(src/io/mod.rs:32-38) and gzip input (src/io/xopen.rs:9-20), through
sa_reads_load and the record accessor. The expectations restate what those
functions do with each input."""
import gzip
import os

import pytest

from conftest import GOLDEN


@pytest.fixture(scope="module")
def sa():
    import strobealign_amd
    strobealign_amd.load()
    return strobealign_amd


def load(sa, tmp_path, data, name="r.fq"):
    p = str(tmp_path / name)
    if name.endswith(".gz"):
        with gzip.open(p, "wb") as f:
            f.write(data)
    else:
        with open(p, "wb") as f:
            f.write(data)
    return sa.Reads(p)


def records(reads):
    return [reads.record(i) for i in range(len(reads))]


# (file bytes, records as (name, comment, seq, qual, end))
CASES = [
    # CRLF lines, a /1 suffix, a comment with spaces, lowercase bases
    (b"@r1/1 comment here\r\nacgtN\r\n+\r\nIIIII\r\n", [(b"r1", b"comment here", b"ACGTN", b"IIIII", 1)]),
    # tab separator, trailing blanks trimmed off the header only
    (b"@a\tc1 c2  \nAC\n+\nII\n@b/2\nG\n+x\n#\n", [(b"a", b"c1 c2", b"AC", b"II", 0), (b"b", None, b"G", b"#", 2)]),
    # the separator is the first blank: the comment keeps what follows it
    (b"@a  two\nA\n+\nI\n", [(b"a", b" two", b"A", b"I", 0)]),
    # a vertical tab is not ASCII whitespace for trim_ascii_end
    (b"@a\x0b\nA\n+\nI\n", [(b"a\x0b", None, b"A", b"I", 0)]),
    # the name is checked before "/1" is cut off
    (b"@/1\nA\n+\nI\n", [(b"", None, b"A", b"I", 1)]),
    # a record whose file ends after its header: empty sequence and qualities
    (b"@a\nAC\n+\nII\n@b\n", [(b"a", None, b"AC", b"II", 0), (b"b", None, b"", b"", 0)]),
    # an unterminated last line loses its last byte (here the '\r'), then no
    # second '\r' follows
    (b"@a\nAC\n+\nII\r", [(b"a", None, b"AC", b"II", 0)]),
    # FASTA reads: lines joined, trimmed at their ends, uppercased
    (b">r1 c\nac\ngt \n>r2\n\nAC\n", [(b"r1", b"c", b"ACGT", None, 0), (b"r2", None, b"AC", None, 0)]),
    (b"", []),
]


@pytest.mark.parametrize("i", range(len(CASES)))
def test_fastq_and_fasta_semantics(sa, tmp_path, i):
    data, exp = CASES[i]
    assert records(load(sa, tmp_path, data)) == exp


@pytest.mark.parametrize("data,msg", [
    (b"@a\nACGT\n+\nIIII", "lengths differ"),  # the unterminated last line loses its last character
    (b"@a\nACGT\n+\nIII\n", "lengths differ"),
    (b"a\nA\n+\nI\n", "must start with '@'"),
    (b"@a\nA\n+\nI\n\n", "must start with '@'"),  # a blank line is a record that does not start with '@'
    (b"@ x\nA\n+\nI\n", "identifier is empty"),
    (b"@a\nAC\n", "lengths differ"),  # no quality line
])
def test_fastq_errors(sa, tmp_path, data, msg):
    with pytest.raises(sa.SaError, match=msg):
        load(sa, tmp_path, data)


def test_gzip_input_multi_member(sa, tmp_path):
    """A .gz file is decompressed, also when it holds several gzip members one
    after another (MultiGzDecoder)."""
    with open(os.path.join(GOLDEN, "phix.1.fastq"), "rb") as f:
        raw = f.read()
    half = raw.index(b"\n@", len(raw) // 2) + 1
    p = str(tmp_path / "r.fq.gz")
    with open(p, "wb") as f:
        f.write(gzip.compress(raw[:half]) + gzip.compress(raw[half:]))
    plain = load(sa, tmp_path, raw, "plain.fq")
    got = sa.Reads(p)
    assert len(got) == len(plain) > 10
    assert records(got) == records(plain)
