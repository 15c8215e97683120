"""The seeded input generators (drivers/gen_equal.cpp, drivers/gen_diff.cpp; the
reference's utils/gen_equal.cpp and utils/gen_diff.cpp with a seed) write the drivers'
stdin format, reproduce the Python generator's key stream, and keep each mode's
defining property (CPU only: the generators are host programs)."""
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "segmented-sorting_amd" / "bin"


@pytest.fixture(scope="module")
def gens():
    if not all((BIN / g).exists() for g in ("gen_equal", "gen_diff")):
        sys.path.insert(0, str(ROOT / "segmented-sorting_amd"))
        try:
            import build
        finally:
            sys.path.pop(0)
        build.build_generators()
    return BIN


def run(gens, *args):
    r = subprocess.run([str(gens / args[0]), *map(str, args[1:])], capture_output=True, text=True, check=True)
    tok = r.stdout.split()
    s = int(tok[0])
    off = np.array(tok[1:s + 2], dtype=np.int64)
    n = int(tok[s + 2])
    keys = np.array(tok[s + 3:s + 3 + n], dtype=np.uint64)
    assert len(tok) == s + 3 + n
    return off, keys


def test_gen_equal_rand_matches_python_stream(gens, pkg):
    off, keys = run(gens, "gen_equal", 64, 300, "--seed", 9)
    assert np.array_equal(off, np.arange(65) * 300)
    assert np.array_equal(keys.astype(np.uint32), pkg.gen.keys_u32(9, 1, 64 * 300, 12))
    _, k31 = run(gens, "gen_equal", 4, 1000, "--seed", 9, "--bits", 31)
    assert np.array_equal(k31.astype(np.uint32), pkg.gen.keys_u32(9, 1, 4000, 31))


def test_gen_equal_modes(gens):
    off, keys = run(gens, "gen_equal", 50, 40, "--mode", "randminmax", "--bits", 8, "--seed", 2)
    for a, b in zip(off[:-1], off[1:]):
        seg = keys[a:b]
        assert seg.min() == 0 and seg.max() == 255 and (seg == 0).any() and (seg == 255).any()
    _, asc = run(gens, "gen_equal", 8, 100, "--mode", "sortasc")
    _, desc = run(gens, "gen_equal", 8, 100, "--mode", "sortdesc")
    assert np.all(np.diff(asc.astype(np.int64)) >= 0) and np.all(np.diff(desc.astype(np.int64)) <= 0)
    assert np.array_equal(np.sort(asc), np.sort(desc))


def test_gen_diff(gens, pkg):
    off, keys = run(gens, "gen_diff", 100, 5000, "--seed", 4)
    assert off[0] == 0 and off[-1] == 5000 and len(off) == 101
    assert np.all(np.diff(off) > 0)  # distinct positions: no empty segment
    assert keys.max() < 1024
    assert np.array_equal(keys.astype(np.uint32), pkg.gen.keys_u32(4, 1, 5000, 10))
    off2, _ = run(gens, "gen_diff", 100, 5000, "--seed", 5)
    assert not np.array_equal(off, off2)
