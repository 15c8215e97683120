"""Build libsegsort.so (gfx950) and the C++ drivers in-tree with hipcc."""
from __future__ import annotations

import json
import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed-FP32 VALU forms stay off: the library has no float math that needs them.
COMMON = ["-O3", "-std=c++17", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", "-Wall", "-Wno-unused-result"]
SOURCES = [HERE / "csrc" / "segsort_api.hip"]
LIB = HERE / "lib" / "libsegsort.so"
DRIVERS = ["radixseg", "fixcub", "mergeseg", "bbsegsort", "nseg", "facade_example"]
GENERATORS = ["gen_equal", "gen_diff"]  # host-only input generators (g++)
CXX = os.environ.get("CXX", "g++")


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")


TABLE_JSON = HERE / "tuning.json"
TABLE_INC = HERE / "csrc" / "tuning_table.inc"


def table_source(table: dict) -> str:
    """C++ text of the library's built-in tuning table (segsort_api.hip includes it)
    from the autotuner's tuning.json: one entry per signature k<K>v<V>-n<N>-m<M>, or
    k<K>v<V>-n<N>-m<M>-x<X> for calls that bound their segment lengths by 2^X."""
    rows = []
    for sig in sorted(table):
        parts = sig.split("-")
        head, n, m = parts[:3]
        x = int(parts[3][1:]) if len(parts) > 3 else -1  # ceil(log2 max length) of bounded calls, -1: no bound
        kw, vb = head[1:].split("v")
        t = table[sig]
        rows.append(f"    {{{int(kw)}, {int(vb)}, {int(n[1:])}, {int(m[1:])}, {x}, "
                    f"{{{t['wave_cap']}, 0x{t['tile_enable']:02X}, {t['grid_mult']}, {t['concurrent']}}}}},  // {sig}")
    return ("// Generated by build.py from tuning.json (the autotuner's table); do not edit.\n"
            "// {key width, value bytes, round(log2 N), round(log2 mean length),\n"
            "//  ceil(log2 max length) of calls that bound their lengths (-1: no bound),\n"
            "//  {wave_cap, tile_enable, grid_mult, concurrent}}\n"
            f"constexpr int TUNING_TABLE_SIZE = {len(rows)};\n"
            "constexpr TableEntry TUNING_TABLE[TUNING_TABLE_SIZE > 0 ? TUNING_TABLE_SIZE : 1] = {\n"
            + "\n".join(rows) + "\n};\n")


def write_table() -> Path:
    text = table_source(json.loads(TABLE_JSON.read_text()))
    if not TABLE_INC.exists() or TABLE_INC.read_text() != text:
        TABLE_INC.write_text(text)
    return TABLE_INC


def build_library() -> Path:
    LIB.parent.mkdir(exist_ok=True)
    write_table()
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, *COMMON, "-fPIC", "-shared", "-o", str(tmp), *map(str, SOURCES)])
    os.replace(tmp, LIB)
    return LIB


def build_drivers() -> list[Path]:
    out = []
    bindir = HERE / "bin"
    bindir.mkdir(exist_ok=True)
    for name in DRIVERS:
        src = HERE / "drivers" / f"{name}.cpp"
        if not src.exists():
            continue
        exe = bindir / name
        _run([HIPCC, "-O2", "-std=c++17", f"-I{ROOT / 'include'}", "-o", str(exe), str(src), f"-L{LIB.parent}",
              "-lsegsort", f"-Wl,-rpath,{LIB.parent}"])
        out.append(exe)
    return out + build_generators()


def build_generators() -> list[Path]:
    out = []
    bindir = HERE / "bin"
    bindir.mkdir(exist_ok=True)
    for name in GENERATORS:
        exe = bindir / name
        _run([CXX, "-O2", "-std=c++17", "-Wall", "-o", str(exe), str(HERE / "drivers" / f"{name}.cpp")])
        out.append(exe)
    return out


def build_all() -> None:
    build_library()
    build_drivers()


if __name__ == "__main__":
    build_all()
    print(f"built {LIB}")
