"""The C-ABI library loads and exports every entry point declared in
include/sa_hip.h (no device calls: this runs without a GPU)."""
import ctypes
import os
import re

from conftest import ROOT


def declared():
    src = open(os.path.join(ROOT, "include", "sa_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(sa_[a-z_0-9]+)\s*\(", src)))


def declared_ssw():
    src = open(os.path.join(ROOT, "include", "ssw.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b([a-z_]+)\s*\(", src)))


def test_library_exports_all_declared_symbols():
    import strobealign_amd
    lib = strobealign_amd.load()
    names = declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(strobealign_amd.EXPORTS)


def test_library_exports_the_ssw_link_shim():
    """ssw_init / init_destroy / ssw_align / align_destroy, the binding of
    src/ssw/raw.rs:26-51, with the s_align layout it declares (raw.rs:3-17)."""
    import strobealign_amd
    lib = strobealign_amd.load()
    assert declared_ssw() == ["align_destroy", "init_destroy", "ssw_align", "ssw_init"]
    for n in declared_ssw():
        assert hasattr(lib, n), n

    class SAlign(ctypes.Structure):
        _fields_ = [("score1", ctypes.c_uint16), ("score2", ctypes.c_uint16), ("ref_begin1", ctypes.c_int32),
                    ("ref_end1", ctypes.c_int32), ("read_begin1", ctypes.c_int32), ("read_end1", ctypes.c_int32),
                    ("ref_end2", ctypes.c_int32), ("cigar", ctypes.POINTER(ctypes.c_uint32)),
                    ("cigarLen", ctypes.c_int32), ("flag", ctypes.c_uint16)]
    # repr(C) layout of raw.rs: offsets 0, 2, 4 .. 20, pointer at 24, 32, 36; size 40
    assert [getattr(SAlign, f).offset for f, _ in SAlign._fields_] == [0, 2, 4, 8, 12, 16, 20, 24, 32, 36]
    assert ctypes.sizeof(SAlign) == 40
    hdr = open(os.path.join(ROOT, "include", "ssw.h")).read()
    body = hdr[hdr.index("typedef struct {"):hdr.index("} s_align;")]
    fields = re.findall(r"\b(uint16_t|int32_t|uint32_t\*)\s+(\w+);", body)
    assert [f for _, f in fields] == [f for f, _ in SAlign._fields_]
    # argument checks need no device: a NULL profile gives NULL
    lib.ssw_align.restype = ctypes.c_void_p
    assert lib.ssw_align(None, b"", 0, 12, 1, 15, 0, 0, 15) is None


def test_version_and_errors_without_device():
    import strobealign_amd
    lib = strobealign_amd.load()
    assert b"gfx950" in lib.sa_version()
    p = strobealign_amd.SeedingParams()
    assert lib.sa_seeding_params_for_read_length(ctypes.c_uint64(150), ctypes.byref(p)) == 0
    assert (p.k, p.s, p.w_min, p.w_max, p.max_dist) == (20, 16, 5, 11, 80)
    assert lib.sa_index_build(None, None, ctypes.c_uint64(0), 0, 1, None) == -1
    buf = ctypes.create_string_buffer(256)
    assert lib.sa_last_error(buf, ctypes.c_size_t(256)) > 0


def test_cli_exists():
    import strobealign_amd
    assert os.access(strobealign_amd.CLI_PATH, os.X_OK)
