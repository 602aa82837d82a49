"""Index hand-over on the device (through the C ABI): an index uploaded from
host arrays in the reference's in-memory layout (sa_index_upload), and .sti
files written and streamed back into HBM (sa_index_save_sti /
sa_index_load_sti, the format of src/index.rs:374-495). Each must map the reads
to the same SAM as the index built directly. Needs an MI355X."""
import os
import struct
import sys

import numpy as np
import pytest

from conftest import ROOT

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
def refs(tmp_path_factory):
    sys.path.insert(0, ROOT)
    from strobealign_amd import synth
    d = str(tmp_path_factory.mktemp("ixio"))
    small = synth.make("small", os.path.join(d, "small"), 1000)
    pe = synth.make("pe100m", os.path.join(d, "pe"), 1000)
    return {"small": (small["ref"], [small["reads1"]]), "pe": (pe["ref"], [pe["reads1"], pe["reads2"]])}


def read_fasta(path):
    names, seqs, cur = [], [], []
    with open(path, "rb") as f:
        for line in f:
            line = line.rstrip(b"\r\n")
            if line.startswith(b">"):
                if names:
                    seqs.append(b"".join(cur))
                names.append(line[1:].split()[0].decode())
                cur = []
            else:
                cur.append(line)
    seqs.append(b"".join(cur))
    return names, seqs


def pack(seqs):
    """PackedSeq (src/packed_seq.rs:12-126): base i of a contig at bits
    2(i % 32)..2(i % 32)+1 of its word i / 32, A C G T = 0 1 2 3, every contig
    starting on a word boundary. ACGT-only contigs (the synthetic references)."""
    code = np.full(256, 255, dtype=np.uint8)
    for i, c in enumerate(b"ACGT"):
        code[c] = i
    words, woff, lens = [], [0], []
    for s in seqs:
        c = code[np.frombuffer(s, dtype=np.uint8)]
        assert (c < 4).all()
        n = len(c)
        padded = np.zeros((n + 31) // 32 * 32, dtype=np.uint64)
        padded[:n] = c
        w = (padded.reshape(-1, 32) << (2 * np.arange(32, dtype=np.uint64))).sum(axis=1, dtype=np.uint64)
        words.append(w)
        woff.append(woff[-1] + len(w))
        lens.append(n)
    return np.concatenate(words).tolist(), woff, lens


@pytest.mark.parametrize("name", ["small", "pe"])
def test_index_upload_maps_like_the_built_index(sa, ctx, refs, name):
    """A host-built index downloaded and handed back through sa_index_upload
    with a reference packed here gives the same SAM."""
    fasta, reads = refs[name]
    say("case: sa_index_upload of a downloaded %s index" % name)
    with sa.Index(ctx, fasta, read_length=150, threads=8) as ix:
        exp = ix.map_files(*reads, threads=8)
        prm, n, nc = ix.params()
        ent, bs = ix.download()
    names, seqs = read_fasta(fasta)
    words, woff, lens = pack(seqs)
    assert len(lens) == nc and len(ent) == 16 * n
    with sa.Index(ctx, fasta, read_length=150, upload=(prm, ent, bs, words, woff, lens, names)) as up:
        p2, n2, _ = up.params()
        assert (n2, p2.bits, p2.filter_cutoff) == (n, prm.bits, prm.filter_cutoff)
        got = up.map_files(*reads, threads=8)
    assert got == exp and len(exp) > 10_000


@pytest.mark.parametrize("name,device_build", [("small", False), ("pe", True)])
def test_sti_round_trip(sa, ctx, refs, tmp_path, name, device_build):
    """sa_index_save_sti writes the reference's .sti layout (checked field by
    field); sa_index_load_sti streams it back into HBM and maps the same SAM."""
    fasta, reads = refs[name]
    sti = str(tmp_path / (name + ".sti"))
    say("case: .sti round trip, %s index (device build: %s)" % (name, device_build))
    with sa.Index(ctx, fasta, read_length=150, threads=8, device_build=device_build) as ix:
        exp = ix.map_files(*reads, threads=8)
        prm, n, _ = ix.params()
        ent, bs = ix.download()
        ix.save_sti(sti, 150)
    with open(sti, "rb") as f:
        raw = f.read()
    # src/index.rs:374-409
    assert raw[:4] == b"STI\x01"
    ver, rlen, rval, cutoff, bits, profile, k, s, wmin, wmax, q, md, mask = struct.unpack_from("<IQQIIIIIIIIIQ", raw, 4)
    assert (ver, rlen, rval) == (7, 8, 0)
    assert (cutoff, bits, profile) == (prm.filter_cutoff, prm.bits, 5)  # ReadLength150 (src/seeding/parameters.rs:10-19)
    assert (k, s, wmin, wmax, q, md, mask) == (prm.k, prm.s, prm.w_min, prm.w_max, prm.q, prm.max_dist,
                                                prm.main_hash_mask)
    at = 4 + struct.calcsize("<IQQIIIIIIIIIQ")
    (ne,) = struct.unpack_from("<Q", raw, at)
    assert ne == n and raw[at + 8:at + 8 + 16 * n] == ent
    at += 8 + 16 * n
    (nb,) = struct.unpack_from("<Q", raw, at)
    assert nb == (1 << prm.bits) + 1 and list(struct.unpack_from("<%dQ" % nb, raw, at + 8)) == bs
    assert len(raw) == at + 8 + 8 * nb
    with sa.Index(ctx, fasta, read_length=150, bits=prm.bits, sti=sti) as ld:
        p2, n2, _ = ld.params()
        assert (n2, p2.bits, p2.filter_cutoff) == (n, prm.bits, prm.filter_cutoff)
        assert ld.download() == (ent, bs)
        got = ld.map_files(*reads, threads=8)
    assert got == exp


def test_sti_errors(sa, ctx, refs, tmp_path):
    """The reference's read_index refusals: another profile, other bits, a
    wrong signature, another format version, a truncated file."""
    fasta, _ = refs["small"]
    sti = str(tmp_path / "s.sti")
    say("case: .sti refusals")
    with sa.Index(ctx, fasta, read_length=150) as ix:
        bits = ix.params()[0].bits
        ix.save_sti(sti, 150)
    with pytest.raises(sa.SaError, match="differ"):
        sa.Index(ctx, fasta, read_length=250, bits=bits, sti=sti)
    with pytest.raises(sa.SaError, match="differ"):
        sa.Index(ctx, fasta, read_length=150, bits=bits + 1, sti=sti)
    raw = open(sti, "rb").read()
    bad = str(tmp_path / "bad.sti")
    for blob, msg in ((b"STJ\x01" + raw[4:], "Signature"), (raw[:4] + struct.pack("<I", 6) + raw[8:], "version"),
                      (raw[:len(raw) // 2], "end of")):
        with open(bad, "wb") as f:
            f.write(blob)
        with pytest.raises(sa.SaError, match=msg):
            sa.Index(ctx, fasta, read_length=150, bits=bits, sti=bad)
