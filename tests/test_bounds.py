"""The score bound that routes SSW jobs to the 16-bit pass first, on the CPU.

two_diagonal_score (strobealign_amd/csrc/sa_extend.hpp) must never exceed the
local alignment score: a job it routes wrongly would skip a byte pass whose
result the reference keeps (ext/ssw/ssw.c:197-588 runs the 16-bit pass only
when the byte pass saturates). bound_check.cpp compares it with an enumeration
of the alignments it stands for and with a Gotoh local alignment.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_two_diagonal_bound(tmp_path):
    exe = str(tmp_path / "bound_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "strobealign_amd", "csrc"),
                    os.path.join(ROOT, "tests", "bound_check.cpp"), "-o", exe], check=True)
    p = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    assert p.stdout.startswith("ok ")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_diagonal_ratio_without_division(tmp_path):
    """The chaining kernels' division-free diagonal-ratio test equals the
    reference's f32 division and comparison (src/chainer.rs:126-130)."""
    exe = str(tmp_path / "ratio_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "strobealign_amd", "csrc"),
                    os.path.join(ROOT, "tests", "ratio_check.cpp"), "-o", exe], check=True)
    p = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    assert p.stdout.startswith("ok ")
