"""sa_find_nams_batch_noalloc (the entry point that allocates nothing) against
sa_find_nams_batch on the same reads: NAM lists equal field for field, on two
calls in a row, the second with fewer reads, plus its size errors. Device
buffers for the inputs and outputs come from the HIP runtime the library links
(the test allocates; the entry point does not). Needs an MI355X."""
import ctypes
import os
import sys

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
def hip(sa):
    # the runtime libsa_hip.so is linked against (already loaded: same handle)
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


class Dev:
    """A device buffer owned by the test."""

    def __init__(self, hip, nbytes):
        self.hip, self.n = hip, max(16, nbytes)
        self.p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), self.n) == 0

    def put(self, data):
        assert len(data) <= self.n
        buf = ctypes.create_string_buffer(bytes(data), len(data))
        assert self.hip.hipMemcpy(self.p, buf, len(data), 1) == 0  # host to device

    def get(self, nbytes):
        buf = ctypes.create_string_buffer(max(1, nbytes))
        assert self.hip.hipMemcpy(buf, self.p, nbytes, 2) == 0  # device to host
        return buf.raw[:nbytes]

    def free(self):
        self.hip.hipFree(self.p)


def fastq_seqs(path):
    with open(path, "rb") as f:
        lines = f.read().split(b"\n")
    return [lines[i] for i in range(1, len(lines), 4)]


def noalloc(sa, hip, ix, seqs, nam_cap, mcs=0):
    """NAM records (as bytes, 48 each) and offsets from the no-allocation entry point."""
    import struct
    blob = b"".join(seqs)
    offs = [0]
    for s in seqs:
        offs.append(offs[-1] + len(s))
    d_seq, d_off = Dev(hip, len(blob)), Dev(hip, 8 * len(offs))
    d_nams, d_noff = Dev(hip, 48 * max(1, nam_cap)), Dev(hip, 8 * len(offs))
    try:
        d_seq.put(blob)
        d_off.put(struct.pack("<%dQ" % len(offs), *offs))
        n_nams = ctypes.c_uint64(12345)
        rc = sa._lib.sa_find_nams_batch_noalloc(ix._p, d_seq.p, d_off.p, ctypes.c_uint32(len(seqs)),
                                                ctypes.c_uint32(mcs), ctypes.c_uint32(100), d_nams.p,
                                                ctypes.c_uint64(nam_cap), d_noff.p, ctypes.byref(n_nams), None)
        if rc != 0:
            return rc, sa.last_error(), None
        n = n_nams.value
        offsets = list(struct.unpack("<%dQ" % len(offs), d_noff.get(8 * len(offs))))
        return rc, offsets, d_nams.get(48 * n)
    finally:
        for d in (d_seq, d_off, d_nams, d_noff):
            d.free()


def reference_batch(sa, ix, seqs):
    """NAM records (bytes) and offsets from sa_find_nams_batch."""
    nams, offs = ix.nams_raw(seqs)
    return offs, nams


def test_noalloc_equals_find_nams_batch(sa, hip, tmp_path_factory):
    sys.path.insert(0, ROOT)
    from strobealign_amd import synth
    d = str(tmp_path_factory.mktemp("noalloc"))
    w = synth.make("pe100m", d, 3000)
    seqs = fastq_seqs(w["reads1"])[:1500] + fastq_seqs(w["reads2"])[:500]
    with sa.Context(0) as ctx, sa.Index(ctx, w["ref"], read_length=150, threads=8) as ix:
        for first, count in ((0, 2000), (1700, 300)):  # two calls in a row, the second smaller
            batch = seqs[first:first + count]
            say("case: no-allocation NAMs, %d reads from %d" % (count, first))
            exp_off, exp = reference_batch(sa, ix, batch)
            rc, offs, got = noalloc(sa, hip, ix, batch, nam_cap=len(exp) // 48 + 8)
            assert rc == 0, offs
            assert offs == exp_off
            assert got == exp
            assert len(exp) > 48 * count  # several NAMs per read: the batch exercised chaining
        say("case: no-allocation NAMs, output room one record short")
        batch = seqs[:300]
        exp_off, exp = reference_batch(sa, ix, batch)
        rc, msg, _ = noalloc(sa, hip, ix, batch, nam_cap=len(exp) // 48 - 1)
        assert rc == -1 and "NAMs" in msg
        say("case: no-allocation NAMs, batch beyond the scratch size")
        big = seqs * 12  # about 24,000 reads: far more than the scratch holds
        rc, msg, _ = noalloc(sa, hip, ix, big, nam_cap=16)
        assert rc == -1 and "scratch" in msg
        # the scratch keeps working after a refused call
        exp_off, exp = reference_batch(sa, ix, seqs[:100])
        rc, offs, got = noalloc(sa, hip, ix, seqs[:100], nam_cap=len(exp) // 48 + 1)
        assert rc == 0 and offs == exp_off and got == exp
