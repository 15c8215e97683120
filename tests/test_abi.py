"""The C-ABI library loads and exports every declared symbol; host-only paths
(temp-size query, argument checks, tuning table) behave like cub's (CPU only)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HIP_INVALID_VALUE = 1


def declared_symbols():
    text = (ROOT / "include" / "segsort.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(segsort_\w+)\s*\(", text, flags=re.M)))


def test_header_matches_binding(pkg):
    assert set(declared_symbols()) == set(pkg.EXPORTS)


def test_library_exports_every_symbol(pkg):
    lib = pkg.load_library()
    for name in declared_symbols():
        assert hasattr(lib, name), name
    assert lib.segsort_version() >= 100


def test_missing_library_fails_loudly(pkg, tmp_path):
    with pytest.raises(pkg.SegsortError):
        pkg.load_library(tmp_path / "nope.so")


def _query(pkg, n, s, key_type=0, vb=4, algo=0, bb=0, eb=None):
    lib = pkg.load_library()
    tb = ctypes.c_size_t(0)
    if eb is None:
        eb = 64 if key_type >= 3 else 32
    rc = lib.segsort_sort(None, ctypes.byref(tb), None, None, None, None, n, s, None, None, bb, eb, key_type, vb, 0,
                          algo, None)
    return rc, tb.value


def test_temp_query_no_gpu(pkg):
    for algo in (0, 1, 2):
        for kt in (0, 2):
            for vb in (0, 4, 8):
                rc, tb = _query(pkg, 1 << 20, 1 << 10, kt, vb, algo)
                assert rc == 0 and tb >= 256
    rc, tb = _query(pkg, 0, 0)
    assert rc == 0 and tb >= 1  # cub reports at least one byte (dispatch_radix_sort.cuh:1352)
    # batched temp grows with N (double buffers for the multi-workgroup tier)
    assert _query(pkg, 1 << 24, 1 << 16)[1] > 2 * (1 << 24) * 8


def test_argument_errors(pkg):
    assert _query(pkg, -1, 4)[0] == HIP_INVALID_VALUE
    assert _query(pkg, 4, -1)[0] == HIP_INVALID_VALUE
    assert _query(pkg, 4, 1, key_type=9)[0] == HIP_INVALID_VALUE
    assert _query(pkg, 4, 1, vb=3)[0] == HIP_INVALID_VALUE
    assert _query(pkg, 4, 1, bb=8, eb=4)[0] == HIP_INVALID_VALUE
    assert _query(pkg, 4, 1, bb=0, eb=33)[0] == HIP_INVALID_VALUE
    assert _query(pkg, 4, 1, key_type=3, algo=2)[0] == HIP_INVALID_VALUE  # fix-pass: 32-bit keys only
    assert "fix" in pkg.last_error()
    lib = pkg.load_library()
    assert lib.segsort_sort(None, None, None, None, None, None, 4, 1, None, None, 0, 32, 0, 0, 0, 0,
                            None) == HIP_INVALID_VALUE


def test_temp_too_small_is_invalid_value(pkg):
    lib = pkg.load_library()
    rc, need = _query(pkg, 1000, 10)
    small = ctypes.c_size_t(need - 1)
    dummy = ctypes.c_void_p(0x1000)  # never dereferenced: the size check comes first
    rc = lib.segsort_sort(dummy, ctypes.byref(small), dummy, dummy, dummy, dummy, 1000, 10, dummy, dummy, 0, 32, 0, 4,
                          0, 0, None)
    assert rc == HIP_INVALID_VALUE


def test_tuning_table(pkg):
    pkg.set_tuning(4, 4, 16, 0x5, 2, 0)
    assert pkg.get_tuning(4, 4) == {"wave_cap": 16, "tile_enable": 5, "grid_mult": 2, "concurrent": 0}
    with pytest.raises(pkg.SegsortError):
        pkg.set_tuning(4, 4, 513, 0xF, 1)
    with pytest.raises(pkg.SegsortError):
        pkg.set_tuning(4, 4, 32, 0xF, 1, 3)
    with pytest.raises(pkg.SegsortError):
        pkg.set_tuning(4, 4, 32, 0x800, 1)  # bit 11 and above are not defined
    pkg.set_tuning(4, 4, 512, 0x40F, 1)  # the wave tier split at 256 items
    assert pkg.get_tuning(4, 4)["tile_enable"] == 0x40F
    pkg.set_tuning(4, 4, 64, 0x3FF, 1)  # ... and the 4096-item tile without its half-depth build
    assert pkg.get_tuning(4, 4)["tile_enable"] == 0x3FF
    pkg.set_tuning(4, 4, 64, 0x1FF, 1)  # every wave-group and tile class, longest first
    assert pkg.get_tuning(4, 4)["tile_enable"] == 0x1FF
    pkg.set_tuning(4, 4, 64, 0xFF, 1)  # the same in binning order
    assert pkg.get_tuning(4, 4)["tile_enable"] == 0xFF
    assert pkg.lookup_tuning(1 << 20, 1 << 10, 4, 4)["tile_enable"] == 0xFF  # the override wins
    pkg.reset_tuning(4, 4)
    assert pkg.get_tuning(4, 4) == {"wave_cap": 256, "tile_enable": 0xEF, "grid_mult": 1, "concurrent": 0}
    assert pkg.lookup_tuning(1 << 20, 1 << 10, 4, 4)["from_table"]
    with pytest.raises(pkg.SegsortError):
        pkg.reset_tuning(5, 4)
    caps = pkg.tile_caps(4, 4)
    assert caps == sorted(caps) and caps[-1] >= 8192


def _ms_query(pkg, off, n, num_streams=4, kt=0, vb=4):
    lib = pkg.load_library()
    tb = ctypes.c_size_t(0)
    h = (ctypes.c_int * len(off))(*off)
    rc = lib.segsort_multistream_sort(None, ctypes.byref(tb), None, None, None, None, n, len(off) - 1, h, 0,
                                      64 if kt >= 3 else 32, kt, vb, 0, num_streams, 0, None)
    return rc, tb.value


def test_multistream_query_and_errors(pkg):
    off = [0, 100, 100, 5000, 70000]
    rc, one = _ms_query(pkg, off, 70000, num_streams=1)
    assert rc == 0 and one >= 256
    rc, four = _ms_query(pkg, off, 70000, num_streams=4)
    assert rc == 0 and four > one  # one temp region per stream (4 segments: 4 streams used)
    assert _ms_query(pkg, off, 70000, num_streams=32)[1] == four  # never more streams than segments
    assert _ms_query(pkg, off, 70000, num_streams=0)[0] == HIP_INVALID_VALUE
    assert _ms_query(pkg, off, 70000, num_streams=33)[0] == HIP_INVALID_VALUE
    assert _ms_query(pkg, [0, 10, 5], 10)[0] == HIP_INVALID_VALUE  # descending offsets
    assert "ascending" in pkg.last_error()
    assert _ms_query(pkg, [0, 10, 20], 15)[0] == HIP_INVALID_VALUE  # past num_items
    assert _ms_query(pkg, off, 70000, kt=7)[0] == HIP_INVALID_VALUE
