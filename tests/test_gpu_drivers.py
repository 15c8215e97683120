"""The reference-compatible drivers (radixseg, fixcub, mergeseg, bbsegsort, nseg) read the
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
DRIVERS = ("radixseg", "fixcub", "mergeseg", "bbsegsort", "nseg")


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


def _parse_in(path):
    tok = Path(path).read_text().split()
    s = int(tok[0])
    off = np.array(tok[1:s + 2], dtype=np.int64).astype(np.int32)
    n = int(tok[s + 2])
    keys = np.array(tok[s + 3:s + 3 + n], dtype=np.int64).astype(np.uint32)
    return off, keys


def _line(a):
    return "".join(f"{int(x)} " for x in a) + "\n"


def test_golden_diff_with_values(pkg, orc, tmp_path):
    """The reference's test flow (scripts/test.sh:1-25) over inputs from the seeded
    generators (gen_equal in its four modes, gen_diff): every driver's --values output,
    keys AND values, is compared byte for byte with a golden file built from the
    oracle's stable sort; radixseg/fixcub/mergeseg carry the input index as value,
    bbsegsort the key itself (as a double)."""
    ind = tmp_path / "in"
    ind.mkdir()
    specs = [("gen_equal", 64, 500, "--mode", "rand"), ("gen_equal", 1, 4096, "--mode", "randminmax"),
             ("gen_equal", 30, 300, "--mode", "randminmax"), ("gen_equal", 16, 1000, "--mode", "sortasc"),
             ("gen_equal", 16, 1000, "--mode", "sortdesc"), ("gen_equal", 512, 256, "--mode", "rand", "--bits", 31),
             ("gen_diff", 200, 60000), ("gen_diff", 3, 40000), ("gen_diff", 1000, 1000)]
    for i, sp in enumerate(specs):
        name = f"{sp[1]}_{sp[2]}_{i}"
        r = subprocess.run([str(BIN / sp[0]), *map(str, sp[1:]), "--seed", str(100 + i)], capture_output=True,
                           text=True, check=True)
        (ind / f"{name}.in").write_text(r.stdout)
    for drv in DRIVERS:
        exp = tmp_path / f"exp_{drv}"
        exp.mkdir()
        for f in sorted(ind.glob("*.in")):
            off, keys = _parse_in(f)
            if drv == "bbsegsort":
                k = keys.view(np.int32)
                ek, ev = orc.seg_sort(k, k.astype(np.float64), offsets=off)
                text = "\n" + _line(ek) + _line(ev)
            else:
                ek, ev = orc.seg_sort(keys, np.arange(keys.size, dtype=np.uint32), offsets=off)
                text = "\n" + _line(ek) + _line(ev)
            (exp / f"{f.stem}.out").write_text(text)
        print(f"[case] golden diff {drv}", flush=True)
        r = subprocess.run(["bash", str(ROOT / "segmented-sorting_amd" / "drivers" / "test.sh"), str(BIN / drv),
                            str(ind), str(exp), str(tmp_path / f"err_{drv}"), "--values"],
                           capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout + r.stderr
        assert r.stdout.count(": ok") == len(specs), r.stdout
