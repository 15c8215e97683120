"""The reference-compatible drivers (radixseg, fixcub, mergeseg, bbsegsort) read the
reference's stdin format and print its stdout format; their output equals the
oracle's, formatted the way src/cub.out is."""
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "segmented-sorting_amd" / "bin"
DRIVERS = ("radixseg", "fixcub", "mergeseg", "bbsegsort")


def stdin_text(off, keys):
    return f"{len(off) - 1}\n" + " ".join(map(str, off)) + f"\n{len(keys)}\n" + " ".join(map(str, keys)) + "\n"


def expected_stdout(keys_sorted, signed=False):
    ks = keys_sorted.astype(np.int32) if signed else keys_sorted
    return "\n" + "".join(f"{int(k)} " for k in ks) + "\n"


def cases(pkg):
    ms = json.loads((ROOT / "tests" / "golden" / "mergesort2_inputs.json").read_text())
    for name, v in ms.items():
        yield name, np.array(v["offsets"], dtype=np.int32), np.array(v["keys"], dtype=np.uint32)
    lengths = pkg.gen.tiny_lengths(3, 300, 0, 3000)
    off = pkg.gen.offsets_from_lengths(lengths)
    yield "ragged-12bit", off, pkg.gen.keys_u32(3, 1, int(off[-1]), 12)  # the reference's rand()%4096 keys
    off = pkg.gen.equal_offsets(1 << 17, 512)
    yield "equal-31bit", off, pkg.gen.keys_u32(4, 1, 1 << 17, 31)


@pytest.mark.parametrize("drv", DRIVERS)
def test_driver_output(pkg, orc, drv):
    exe = BIN / drv
    assert exe.exists(), f"{exe} not built"
    for name, off, keys in cases(pkg):
        print(f"[case] {drv} {name} n={keys.size}", flush=True)
        ek, _ = orc.seg_sort(keys, None, offsets=off)
        r = subprocess.run([str(exe)], input=stdin_text(off, keys), capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        assert r.stdout == expected_stdout(ek, signed=(drv == "bbsegsort")), (drv, name)


def test_driver_timing_mode(pkg):
    off = pkg.gen.equal_offsets(1 << 16, 64)
    keys = pkg.gen.keys_u32(5, 1, 1 << 16)
    for drv in DRIVERS:
        r = subprocess.run([str(BIN / drv), "--time", "--executions", "3"], input=stdin_text(off, keys),
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        assert float(r.stdout.strip()) > 0


def test_cpp_facade_example():
    """Every overload of the cub-style C++ facade on cub's documented example."""
    r = subprocess.run([str(BIN / "facade_example")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "facade ok" in r.stdout, r.stdout + r.stderr
