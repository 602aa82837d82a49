"""Synthetic references and reads for the benchmark and parity workloads.

Workloads (all randomness from numpy's PCG64 with the stated seeds):
  small   : 1 Mbp single contig (seed 42); 10,000 single-end 150 bp reads,
            1% substitutions, 0.1% indel events of 1-3 bp, Phred 30 qualities.
  pe100m  : 100 Mbp in 20 contigs of 5 Mbp (seed 7), about 30% of the bases
            overwritten by copies of 1-10 kbp segments with 2% divergence;
            paired-end 2x150 reads, fragment length round(Normal(300, 30))
            clipped at >= 150, FR orientation, 0.5% substitutions, 0.05% indels.
  se_var  : same reference; single-end reads of uniform length 50-500 bp,
            15% starting inside copied repeat segments.
  human   : 3.1 Gbp in 24 contigs (seed 11), about 45% of the bases overwritten
            by copies of 300 bp-50 kbp segments at a divergence drawn uniformly
            from [0, 15%] per segment; paired-end 2x250 reads, fragment length
            round(Normal(550, 50)), 3% substitutions, 0.3% indels.
Reads are generated from both strands with equal probability (single-end) or
as FR pairs (paired-end). Indel events insert or delete 1-3 bases (50/50,
uniform inserted bases); reads are cut or extended to keep their length.
"""
import argparse
import os

import numpy as np

ACGT = np.frombuffer(b"ACGT", dtype=np.uint8)
COMP = np.array([3, 2, 1, 0], dtype=np.uint8)


def random_reference(rng, n_contigs, contig_len, repeat_fraction=0.0, seg_min=1000, seg_max=10000, divergence=0.02,
                     max_divergence=None):
    """max_divergence: draw each segment's divergence uniformly from [divergence, max_divergence]."""
    contigs = [rng.integers(0, 4, contig_len, dtype=np.uint8) for _ in range(n_contigs)]
    segments = []  # (contig, start, length) of copied destinations
    target = int(repeat_fraction * n_contigs * contig_len)
    copied = 0
    while copied < target:
        L = int(rng.integers(seg_min, seg_max + 1))
        sc, dc = int(rng.integers(0, n_contigs)), int(rng.integers(0, n_contigs))
        ss, ds = int(rng.integers(0, contig_len - L)), int(rng.integers(0, contig_len - L))
        seg = contigs[sc][ss:ss + L].copy()
        d = divergence if max_divergence is None else float(rng.uniform(divergence, max_divergence))
        mut = rng.random(L) < d
        seg[mut] = (seg[mut] + rng.integers(1, 4, int(mut.sum()), dtype=np.uint8)) % 4
        contigs[dc][ds:ds + L] = seg
        segments.append((dc, ds, L))
        copied += L
    return contigs, segments


def write_fasta(path, names, contigs, width=80):
    with open(path, "wb") as f:
        for name, c in zip(names, contigs):
            f.write(b">" + name.encode() + b"\n")
            s = ACGT[c]
            full = len(s) // width
            if full:  # whole lines at once: (full, width + 1) with a newline column
                lines = np.empty((full, width + 1), dtype=np.uint8)
                lines[:, :width] = s[:full * width].reshape(full, width)
                lines[:, width] = 10
                f.write(lines.tobytes())
            if len(s) % width:
                f.write(s[full * width:].tobytes() + b"\n")


def mutate(rng, template, length, sub_rate, indel_rate):
    """Apply indel events then substitutions to a code array; returns `length` codes."""
    t = template
    if indel_rate > 0:
        events = np.nonzero(rng.random(len(t)) < indel_rate)[0]
        if len(events):
            out, last = [], 0
            for e in events:
                if e < last:
                    continue
                out.append(t[last:e])
                L = int(rng.integers(1, 4))
                if rng.random() < 0.5:
                    out.append(rng.integers(0, 4, L, dtype=np.uint8))
                    last = e
                else:
                    last = e + L
            out.append(t[last:])
            t = np.concatenate(out)
    t = t[:length].copy()
    if len(t) < length:  # extend with random bases (template ran short)
        t = np.concatenate([t, rng.integers(0, 4, length - len(t), dtype=np.uint8)])
    mut = rng.random(length) < sub_rate
    t[mut] = (t[mut] + rng.integers(1, 4, int(mut.sum()), dtype=np.uint8)) % 4
    return t


def write_fastq(f, name, codes, qual_char=b"?"):
    f.write(b"@" + name + b"\n" + ACGT[codes].tobytes() + b"\n+\n" + qual_char * len(codes) + b"\n")


def single_end(rng, contigs, n_reads, lengths, sub_rate, indel_rate, out_path, repeat_segments=None,
               repeat_share=0.0, prefix="r"):
    sizes = np.array([len(c) for c in contigs], dtype=np.float64)
    p = sizes / sizes.sum()
    margin = 12
    with open(out_path, "wb") as f:
        for i in range(n_reads):
            L = int(lengths[i])
            if repeat_segments and rng.random() < repeat_share:
                dc, ds, sl = repeat_segments[int(rng.integers(0, len(repeat_segments)))]
                c = dc
                start = ds + int(rng.integers(0, max(1, sl - 1)))
                start = min(start, len(contigs[c]) - L - margin - 1)
            else:
                c = int(rng.choice(len(contigs), p=p))
                start = int(rng.integers(0, len(contigs[c]) - L - margin))
            tmpl = contigs[c][start:start + L + margin]
            if rng.random() < 0.5:
                tmpl = COMP[tmpl[::-1]]
            write_fastq(f, ("%s%d" % (prefix, i)).encode(), mutate(rng, tmpl, L, sub_rate, indel_rate))


def _apply_errors(rng, tmpl, length, sub_rate, indel_rate):
    """Errors for a batch of templates (rows): indels row by row where they occur,
    substitutions vectorised. Returns an (n, length) code array."""
    n = tmpl.shape[0]
    out = tmpl[:, :length].copy()
    if indel_rate > 0:
        ev = rng.random(tmpl.shape) < indel_rate
        for r in np.nonzero(ev.any(axis=1))[0]:
            out[r] = _indels(rng, tmpl[r], length, np.nonzero(ev[r])[0])
    mut = rng.random((n, length)) < sub_rate
    out[mut] = (out[mut] + rng.integers(1, 4, int(mut.sum()), dtype=np.uint8)) % 4
    return out


def _indels(rng, t, length, events):
    parts, last = [], 0
    for e in events:
        if e < last:
            continue
        parts.append(t[last:e])
        L = int(rng.integers(1, 4))
        if rng.random() < 0.5:
            parts.append(rng.integers(0, 4, L, dtype=np.uint8))
            last = e
        else:
            last = e + L
    parts.append(t[last:])
    t = np.concatenate(parts)[:length]
    if len(t) < length:
        t = np.concatenate([t, rng.integers(0, 4, length - len(t), dtype=np.uint8)])
    return t


def _write_fastq_batch(f, prefix, first, codes, qual_char=b"?"):
    seqs = ACGT[codes]
    q = qual_char * codes.shape[1]
    f.write(b"".join(b"@%s%d\n%s\n+\n%s\n" % (prefix, first + i, seqs[i].tobytes(), q) for i in range(codes.shape[0])))


def paired_end(rng, contigs, n_pairs, read_len, sub_rate, indel_rate, out1, out2, mean=300, sd=30, prefix="p",
               batch=200_000):
    sizes = np.array([len(c) for c in contigs], dtype=np.int64)
    genome = np.concatenate(contigs)
    base = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    margin = 12
    L = read_len
    W = L + margin
    with open(out1, "wb") as f1, open(out2, "wb") as f2:
        for b0 in range(0, n_pairs, batch):
            n = min(batch, n_pairs - b0)
            c = rng.choice(len(contigs), size=n, p=sizes / sizes.sum())
            frag = np.maximum(L, np.round(rng.normal(mean, sd, n))).astype(np.int64)
            start = margin + (rng.random(n) * (sizes[c] - frag - 2 * margin)).astype(np.int64)
            g0 = base[c] + start
            ar = np.arange(W)
            # mate 1 reads forward from the fragment start, mate 2 is the reverse
            # complement of the fragment end; fragments on the reverse strand swap roles
            fwd_t = genome[g0[:, None] + ar[None, :]]
            rev_t = COMP[genome[(g0 + frag)[:, None] - 1 - ar[None, :]]]
            flip = rng.random(n) < 0.5
            m1 = np.where(flip[:, None], COMP[genome[(g0 + frag)[:, None] - 1 - ar[None, :]]], fwd_t)
            m2 = np.where(flip[:, None], genome[g0[:, None] + ar[None, :]], rev_t)
            r1 = _apply_errors(rng, m1, L, sub_rate, indel_rate)
            r2 = _apply_errors(rng, m2, L, sub_rate, indel_rate)
            _write_fastq_batch(f1, prefix.encode(), b0, r1)
            _write_fastq_batch(f2, prefix.encode(), b0, r2)


def make(workload, out_dir, n_reads=None):
    os.makedirs(out_dir, exist_ok=True)
    ref = os.path.join(out_dir, "ref.fa")
    if workload == "small":
        rng = np.random.default_rng(42)
        contigs, _ = random_reference(rng, 1, 1_000_000)
        write_fasta(ref, ["chr1"], contigs)
        n = n_reads or 10_000
        single_end(rng, contigs, n, np.full(n, 150), 0.01, 0.001, os.path.join(out_dir, "reads.fq"))
        return {"ref": ref, "reads1": os.path.join(out_dir, "reads.fq")}
    rng = np.random.default_rng(7)
    contigs, segs = random_reference(rng, 20, 5_000_000, repeat_fraction=0.30)
    names = ["ctg%02d" % (i + 1) for i in range(20)]
    if not os.path.exists(ref):
        write_fasta(ref, names, contigs)
    if workload == "pe100m":
        n = n_reads or 1_000_000
        r1, r2 = os.path.join(out_dir, "pe_%d_1.fq" % n), os.path.join(out_dir, "pe_%d_2.fq" % n)
        if not (os.path.exists(r1) and os.path.exists(r2)):
            paired_end(np.random.default_rng(70), contigs, n, 150, 0.005, 0.0005, r1, r2)
        return {"ref": ref, "reads1": r1, "reads2": r2}
    if workload == "se_var":
        n = n_reads or 2_000_000
        r = os.path.join(out_dir, "sevar_%d.fq" % n)
        if not os.path.exists(r):
            g = np.random.default_rng(71)
            single_end(g, contigs, n, g.integers(50, 501, n), 0.005, 0.0005, r, repeat_segments=segs,
                       repeat_share=0.15, prefix="v")
        return {"ref": ref, "reads1": r}
    raise ValueError(workload)


def make_lengths(out_dir, lengths, n_pairs):
    """Paired-end read sets of several read lengths on the pe100m reference
    (fragment Normal(max(300, 2L), 30), 0.5% substitutions, 0.05% indels);
    returns {"ref": path, "sets": {L: (reads1, reads2)}}."""
    os.makedirs(out_dir, exist_ok=True)
    ref = os.path.join(out_dir, "ref.fa")
    sets = {L: (os.path.join(out_dir, "len%d_%d_1.fq" % (L, n_pairs)), os.path.join(out_dir, "len%d_%d_2.fq" % (L, n_pairs)))
            for L in lengths}
    if os.path.exists(ref) and all(os.path.exists(x) and os.path.exists(y) for x, y in sets.values()):
        return {"ref": ref, "sets": sets}
    contigs, _ = random_reference(np.random.default_rng(7), 20, 5_000_000, repeat_fraction=0.30)
    if not os.path.exists(ref):
        write_fasta(ref, ["ctg%02d" % (i + 1) for i in range(20)], contigs)
    for L, (r1, r2) in sets.items():
        if not (os.path.exists(r1) and os.path.exists(r2)):
            paired_end(np.random.default_rng(1000 + L), contigs, n_pairs, L, 0.005, 0.0005, r1, r2,
                       mean=max(300, 2 * L), sd=30, prefix="l%d_" % L)
    return {"ref": ref, "sets": sets}


def make_human(out_dir, n_pairs=1_000_000):
    """The 3.1 Gbp workload (see the module docstring); files are cached in out_dir."""
    os.makedirs(out_dir, exist_ok=True)
    ref = os.path.join(out_dir, "human.fa")
    r1, r2 = os.path.join(out_dir, "hpe_%d_1.fq" % n_pairs), os.path.join(out_dir, "hpe_%d_2.fq" % n_pairs)
    if os.path.exists(ref) and os.path.exists(r1) and os.path.exists(r2):
        return {"ref": ref, "reads1": r1, "reads2": r2}
    rng = np.random.default_rng(11)
    contigs, _ = random_reference(rng, 24, 3_100_000_000 // 24, repeat_fraction=0.45, seg_min=300, seg_max=50_000,
                                  divergence=0.0, max_divergence=0.15)
    if not os.path.exists(ref):
        write_fasta(ref, ["chr%d" % (i + 1) for i in range(24)], contigs)
    paired_end(np.random.default_rng(110), contigs, n_pairs, 250, 0.03, 0.003, r1, r2, mean=550, sd=50, prefix="h")
    return {"ref": ref, "reads1": r1, "reads2": r2}


def make_human_small(out_dir, n_pairs=2000, total_len=50_000_000):
    """The human-scale generator's shape on a smaller reference (24 contigs of
    total_len / 24, 45% repeats of 300 bp-50 kbp at 0-15% divergence; 2x250
    pairs at 3% substitutions, 0.3% indels, fragments Normal(550, 50)), for
    tests that need its extension-heavy reads without the 3.1 Gbp reference."""
    os.makedirs(out_dir, exist_ok=True)
    ref = os.path.join(out_dir, "human_small.fa")
    r1, r2 = os.path.join(out_dir, "hs_%d_1.fq" % n_pairs), os.path.join(out_dir, "hs_%d_2.fq" % n_pairs)
    if os.path.exists(ref) and os.path.exists(r1) and os.path.exists(r2):
        return {"ref": ref, "reads1": r1, "reads2": r2}
    rng = np.random.default_rng(11)
    contigs, _ = random_reference(rng, 24, total_len // 24, repeat_fraction=0.45, seg_min=300, seg_max=50_000,
                                  divergence=0.0, max_divergence=0.15)
    write_fasta(ref, ["chr%d" % (i + 1) for i in range(24)], contigs)
    paired_end(np.random.default_rng(110), contigs, n_pairs, 250, 0.03, 0.003, r1, r2, mean=550, sd=50, prefix="h")
    return {"ref": ref, "reads1": r1, "reads2": r2}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("workload", choices=["small", "pe100m", "se_var", "human"])
    ap.add_argument("out_dir")
    ap.add_argument("-n", type=int, default=None, help="number of reads / pairs")
    a = ap.parse_args()
    print(make_human(a.out_dir, a.n or 1_000_000) if a.workload == "human" else make(a.workload, a.out_dir, a.n))


if __name__ == "__main__":
    main()
