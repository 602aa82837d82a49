"""Serving loop (sa_server_*): requests mapped with continuous batching give,
each, the output of mapping that request alone. Needs an MI355X."""
import os
import random
import sys
import threading

import pytest

from conftest import ORACLE_BIN, ROOT

pytestmark = pytest.mark.gpu

# 2x150 pairs: a chunk of 10,000 bases ends with the record that crosses it,
# i.e. every 34 pairs (34 x 300 = 10,200)
CHUNK = 10_000
PAIRS_PER_CHUNK = 34


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def sa():
    import strobealign_amd
    strobealign_amd.load()
    return strobealign_amd


@pytest.fixture(scope="module")
def pe(tmp_path_factory):
    sys.path.insert(0, ROOT)
    from strobealign_amd import synth
    return synth.make("pe100m", str(tmp_path_factory.mktemp("serve")), 3000)


def fastq_records(path):
    with open(path, "rb") as f:
        lines = f.read().split(b"\n")
    return [(lines[i][1:], lines[i + 1], lines[i + 3]) for i in range(0, len(lines) - 3, 4)]


def write_fastq(path, recs):
    with open(path, "wb") as f:
        for n, s, q in recs:
            f.write(b"@" + n + b"\n" + s + b"\n+\n" + q + b"\n")


def no_header(sam):
    return b"".join(l for l in sam.splitlines(keepends=True) if not l.startswith(b"@"))


@pytest.mark.parametrize("fused", [False, True])
def test_requests_on_chunk_boundaries_match_oracle(sa, pe, fused):
    """Requests cut at the chunk boundaries of the whole read set, submitted
    together: in order, their outputs are the oracle's SAM of the whole set
    (device batches by the plain stages, or by the fused pass, recorded for the
    batch shapes the requests make)."""
    say("case: serving, 3000 pairs in 12 requests on chunk boundaries, vs oracle, fused %s" % fused)
    import subprocess
    exp = subprocess.run([ORACLE_BIN, "-t", "8", "--no-PG", "--chunk-size", str(CHUNK), pe["ref"], pe["reads1"],
                          pe["reads2"]], check=True, capture_output=True).stdout
    sizes_in_chunks = [1, 5, 2, 20, 1, 9, 13, 3, 7, 1, 17]
    cuts = [0]
    for c in sizes_in_chunks:
        cuts.append(cuts[-1] + c * PAIRS_PER_CHUNK)
    cuts.append(3000)
    with sa.Context(0) as ctx, sa.Index(ctx, pe["ref"], read_length=150, threads=8) as ix:
        reads = sa.Reads(pe["reads1"], pe["reads2"])
        opts = sa.options(threads=8, chunk_size=CHUNK, batch_bases=60_000, slots=2, fused=fused)
        with sa.Server(ix, opts) as srv:
            tickets = [srv.submit(reads, a, b - a) for a, b in zip(cuts, cuts[1:])]
            outs = [srv.wait(t) for t in tickets]
            st = srv.stats()
        reads.close()
    say("  batches %d, requests per batch max %d" % (st["batches"], st["max_requests_per_batch"]))
    assert st["completed"] == len(tickets)
    assert st["max_requests_per_batch"] >= 2  # requests shared device batches
    assert all(lat > 0 for _, lat in outs)
    assert b"".join(o for o, _ in outs) == no_header(exp)


def test_mixed_requests_equal_alone(sa, pe, tmp_path):
    """Requests of uneven sizes, paired and single-end, submitted from several
    threads while others run: each output equals that request mapped alone by
    the plain path (sa_map_files on the same records)."""
    say("case: serving, 14 PE/SE requests of uneven size from 4 threads vs mapped alone")
    r1, r2 = fastq_records(pe["reads1"]), fastq_records(pe["reads2"])
    rnd = random.Random(5)
    reqs = []
    at = 0
    for i in range(14):
        n = rnd.choice([1, 2, 7, 40, 133, 301, 520])
        a = at % 2500
        paired = i % 3 != 2
        reqs.append((a, n, paired))
        at += n
    expected = []
    with sa.Context(0) as ctx, sa.Index(ctx, pe["ref"], read_length=150, threads=8) as ix:
        for k, (a, n, paired) in enumerate(reqs):
            p1, p2 = str(tmp_path / ("r%d_1.fq" % k)), str(tmp_path / ("r%d_2.fq" % k))
            write_fastq(p1, r1[a:a + n])
            if paired:
                write_fastq(p2, r2[a:a + n])
            expected.append(ix.map_files(p1, p2 if paired else None, threads=8, chunk_size=20_000))
        opts = sa.options(threads=4, chunk_size=20_000, batch_bases=40_000, slots=3)
        results = [None] * len(reqs)
        with sa.Server(ix, opts) as srv:
            def client(ks):
                tickets = []
                for k in ks:
                    a, n, paired = reqs[k]
                    tickets.append((k, srv.submit_records(r1[a:a + n], r2[a:a + n] if paired else None)))
                for k, t in tickets:
                    results[k] = srv.wait(t)[0]
            th = [threading.Thread(target=client, args=(list(range(c, len(reqs), 4)),)) for c in range(4)]
            for t in th:
                t.start()
            for t in th:
                t.join()
            st = srv.stats()
    say("  batches %d, requests per batch max %d" % (st["batches"], st["max_requests_per_batch"]))
    for k in range(len(reqs)):
        assert results[k] == no_header(expected[k]), reqs[k]


def test_admission_waits_for_arena(sa, pe):
    """A small device arena admits requests one after another (each waits for
    the memory of earlier ones); a request larger than the arena is refused."""
    say("case: serving with a 200 KB arena")
    with sa.Context(0) as ctx, sa.Index(ctx, pe["ref"], read_length=150, threads=8) as ix:
        reads = sa.Reads(pe["reads1"], pe["reads2"])
        opts = sa.options(threads=8, chunk_size=CHUNK, batch_bases=60_000, slots=2)
        alone = []
        with sa.Server(ix, opts, arena_bytes=200_000) as srv:
            for a in (0, 340, 680):
                alone.append(srv.wait(srv.submit(reads, a, 340))[0])
            tickets = [srv.submit(reads, a, 340) for a in (0, 340, 680)]
            mixed = [srv.wait(t)[0] for t in tickets]
            with pytest.raises(sa.SaError, match="exceed"):
                srv.submit(reads, 0, 1000)  # 300,000 bases
            st = srv.stats()
        reads.close()
    assert mixed == alone
    assert st["arena_peak"] <= 200_000


@pytest.mark.parametrize("fused", [False, True])
def test_requests_of_several_read_lengths(sa, tmp_path, fused):
    """The serving configuration's request shape: each request one read length
    (100 to 300 bp, including queries past 256 bp) on an index built for read
    length 250; every request's output equals mapping its records alone (the
    server's batches by the plain stages or by the fused pass)."""
    say("case: serving, requests of 100-300 bp pairs on an r250 index vs mapped alone, fused %s" % fused)
    sys.path.insert(0, ROOT)
    from strobealign_amd import synth
    lengths = [100, 175, 250, 275, 300]
    w = synth.make_lengths(str(tmp_path / "len"), lengths, 400)
    recs = {L: (fastq_records(w["sets"][L][0]), fastq_records(w["sets"][L][1])) for L in lengths}
    reqs = [(L, a, n) for L, a, n in ((300, 0, 400), (100, 10, 57), (275, 0, 200), (250, 100, 300), (175, 0, 1),
                                      (300, 50, 133), (100, 0, 400))]
    with sa.Context(0) as ctx, sa.Index(ctx, w["ref"], read_length=250, device_build=True) as ix:
        expected = []
        for k, (L, a, n) in enumerate(reqs):
            p1, p2 = str(tmp_path / ("q%d_1.fq" % k)), str(tmp_path / ("q%d_2.fq" % k))
            write_fastq(p1, recs[L][0][a:a + n])
            write_fastq(p2, recs[L][1][a:a + n])
            expected.append(no_header(ix.map_files(p1, p2, threads=8, chunk_size=20_000)))
        opts = sa.options(threads=8, chunk_size=20_000, batch_bases=200_000, slots=3, fused=fused)
        with sa.Server(ix, opts) as srv:
            tickets = [srv.submit_records(recs[L][0][a:a + n], recs[L][1][a:a + n]) for L, a, n in reqs]
            outs = [srv.wait(t)[0] for t in tickets]
    for k in range(len(reqs)):
        assert outs[k] == expected[k], reqs[k]
