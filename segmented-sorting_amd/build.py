"""Build libsegsort.so (gfx950) and the C++ drivers in-tree with hipcc."""
from __future__ import annotations

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
DRIVERS = ["radixseg", "fixcub", "mergeseg", "bbsegsort", "facade_example"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")


def build_library() -> Path:
    LIB.parent.mkdir(exist_ok=True)
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
    return out


def build_all() -> None:
    build_library()
    build_drivers()


if __name__ == "__main__":
    build_all()
    print(f"built {LIB}")
