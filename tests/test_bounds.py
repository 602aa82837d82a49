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


@pytest.fixture(scope="module")
def ssw_host(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("needs g++")
    import ctypes
    so = str(tmp_path_factory.mktemp("sswhost") / "libssw_host.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I", os.path.join(ROOT, "strobealign_amd", "csrc"),
                    os.path.join(ROOT, "tests", "ssw_host.cpp"), "-o", so], check=True)
    lib = ctypes.CDLL(so)
    lib.ssw_raw_host.argtypes = [ctypes.c_char_p, ctypes.c_int32, ctypes.c_char_p, ctypes.c_int32, ctypes.c_char_p,
                                 ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_uint32,
                                 ctypes.c_uint32, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_uint32),
                                 ctypes.c_uint32, ctypes.POINTER(ctypes.c_int32)]
    return lib


def _raw(lib, q, r, mat, n, score_size, go, ge, flag, filters, filterd, mask_len):
    import ctypes
    cig = (ctypes.c_uint32 * 4096)()
    out = (ctypes.c_int32 * 10)()
    lib.ssw_raw_host(q, len(q), r, len(r), mat, n, score_size, go, ge, flag, filters, filterd, mask_len, cig, 4096,
                     out)
    return list(out), [cig[i] for i in range(out[7])]


def test_device_ssw_code_on_the_host_matches_ssw_c(ssw_host):
    """The library's ssw_align_raw (the code the device SSW shim runs: striped
    passes and the banded traceback with packed direction cells) compiled for
    the host, against every ssw.c vector of tests/golden: the shim's argument
    space, the aligner's calls, and user scores."""
    import json
    g = os.path.join(ROOT, "tests", "golden")
    n = 0
    with open(os.path.join(g, "ssw_shim_vectors.json")) as f:
        for v in json.load(f):
            q = bytes(int(c) for c in v["q"])
            r = bytes(int(c) for c in v["r"])
            mat = bytes(x & 0xFF for x in v["mat"])
            o, cig = _raw(ssw_host, q, r, mat, v["n"], v["score_size"], v["gap_open"], v["gap_extend"], v["flag"],
                          v["filters"], v["filterd"], v["mask_len"])
            e = v["res"]
            if e is None:
                assert o[9] == 1, v
            else:
                assert o[9] == 0 and [o[0], o[1], o[2], o[3], o[4], o[5], o[6], o[8], cig] == e, v
            n += 1
    mat = bytes([(2 if (a == b and a < 4) else 256 - 8) for a in range(5) for b in range(5)])
    with open(os.path.join(g, "ssw_vectors.json")) as f:
        for v in json.load(f):
            q = bytes(int(c) for c in v["q"])
            r = bytes(int(c) for c in v["r"])
            o, cig = _raw(ssw_host, q, r, mat, 5, 2, 12, 1, 0x0F, 0, 2**31 - 1, max(len(q) // 2, 15))
            assert [o[0], o[2], o[3], o[4], o[5], o[8]] == v["res"], v
            if v["res"][1] >= 0:
                assert cig == v["cigar"], v
            n += 1
    with open(os.path.join(g, "ssw_score_vectors.json")) as f:
        for v in json.load(f):
            m, x, go, ge = v["sc"]
            q = bytes(int(c) for c in v["q"])
            r = bytes(int(c) for c in v["r"])
            sm = bytes(m if (a == b and a < 4) else (-x) & 0xFF for a in range(5) for b in range(5))
            o, cig = _raw(ssw_host, q, r, sm, 5, 2, go, ge, 0x0F, 0, 2**31 - 1, max(len(q) // 2, 15))
            assert ([o[0], o[2], o[3], o[4], o[5]], cig) == (v["res"], v["cigar"]), v["sc"]
            n += 1
    assert n >= 1400
