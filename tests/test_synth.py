"""Workload generator (strobealign_amd/synth.py) and bench helpers, on the CPU."""
import os

import numpy as np

from strobealign_amd import synth


def plain_fasta(names, contigs, width=80):
    """The FASTA text written one line at a time."""
    out = []
    for name, c in zip(names, contigs):
        out.append(b">" + name.encode() + b"\n")
        s = synth.ACGT[c].tobytes()
        for i in range(0, len(s), width):
            out.append(s[i:i + width] + b"\n")
    return b"".join(out)


def test_write_fasta_matches_line_by_line(tmp_path):
    rng = np.random.default_rng(3)
    lengths = [0, 1, 79, 80, 81, 160, 1000, 12345]
    contigs = [rng.integers(0, 4, n, dtype=np.uint8) for n in lengths]
    names = ["c%d" % i for i in range(len(contigs))]
    for width in (80, 60, 1):
        p = str(tmp_path / ("w%d.fa" % width))
        synth.write_fasta(p, names, contigs, width=width)
        assert open(p, "rb").read() == plain_fasta(names, contigs, width)


def test_random_reference_divergence_range():
    """Per-segment divergence drawn from [divergence, max_divergence]: a copied
    segment differs from its source in at most about max_divergence of its bases."""
    rng = np.random.default_rng(11)
    contigs, segs = synth.random_reference(rng, 2, 200_000, repeat_fraction=0.3, seg_min=300, seg_max=3000,
                                           divergence=0.0, max_divergence=0.15)
    assert sum(L for _, _, L in segs) >= 0.3 * 400_000
    assert all(300 <= L <= 3000 for _, _, L in segs)
    assert all(len(c) == 200_000 for c in contigs)


def test_fastq_mean_length(tmp_path):
    import bench
    p = str(tmp_path / "r.fq")
    with open(p, "wb") as f:
        for n in (50, 150, 250, 10):
            f.write(b"@r\n" + b"A" * n + b"\n+\n" + b"?" * n + b"\n")
    assert bench.fastq_mean_length(p) == (50 + 150 + 250 + 10) // 4
    assert bench.fastq_mean_length(p, n=2) == 100


def test_bench_refuses_more_than_one_gpu():
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    p = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert p.returncode == 2 and "one GPU" in p.stderr
