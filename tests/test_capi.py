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


def test_library_exports_all_declared_symbols():
    import strobealign_amd
    lib = strobealign_amd.load()
    names = declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(strobealign_amd.EXPORTS)


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
