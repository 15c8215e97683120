"""Writes the golden fixtures under tests/golden/ from data the reference holds.

Run in the build container (the reference tree is not on the GPU box):
    python tests/golden/make_fixtures.py /root/reference

* cub_doc_vectors.json  -- the known-answer examples printed in cub's documentation
  (lib/cub/device/device_segmented_radix_sort.cuh:104-129, :302-327, :506-526 and
  lib/cub/device/device_radix_sort.cuh:115-136, :296-317), parsed from the comment text.
  The DoubleBuffer example's value line (:236) is a known copy-paste slip of the
  non-segmented example and is not used.
* cub_out_512x256.npz   -- src/cub.out (== src/thrust.out): 131072 keys printed by the
  reference after sorting 512 segments of 256 keys.
* mergesort2_inputs.json -- src/mergesort2/in1-32 and in2-1024 in the drivers' stdin
  format (S, S+1 offsets, N, N keys).
"""
from __future__ import annotations

import json
import re
import sys
from pathlib import Path

import numpy as np

OUT = Path(__file__).resolve().parent


def _lists_after(text: str, start_line: int, n_lines: int) -> dict:
    lines = text.splitlines()[start_line - 1:start_line - 1 + n_lines]
    found = {}
    for ln in lines:
        m = re.search(r"\*\s*//\s*(d_\w+(?:\.Current\(\))?)\s*<--\s*\[([^\]]*)\]", ln)
        if m:
            found[m.group(1)] = [int(x) for x in m.group(2).split(",")]
        m = re.search(r"\*\s*int\s+\*?(\w+);\s*//\s*e\.g\.,\s*\[([0-9,\s]+)\]", ln)
        if m:
            found[m.group(1)] = [int(x) for x in m.group(2).split(",")]
        m = re.search(r"\*\s*int\s+(num_\w+);\s*//\s*e\.g\.,\s*(\d+)", ln)
        if m:
            found[m.group(1)] = int(m.group(2))
    return found


def cub_vectors(ref: Path) -> dict:
    seg = (ref / "lib/cub/device/device_segmented_radix_sort.cuh").read_text()
    flat = (ref / "lib/cub/device/device_radix_sort.cuh").read_text()
    out = {}
    a = _lists_after(seg, 100, 35)
    out["seg_pairs_ascending"] = {"source": "lib/cub/device/device_segmented_radix_sort.cuh:104-129",
                                  "offsets": a["d_offsets"], "keys": a["d_keys_in"], "values": a["d_values_in"],
                                  "keys_out": a["d_keys_out"], "values_out": a["d_values_out"]}
    d = _lists_after(seg, 298, 35)
    out["seg_pairs_descending"] = {"source": "lib/cub/device/device_segmented_radix_sort.cuh:302-327",
                                   "offsets": d["d_offsets"], "keys": d["d_keys_in"], "values": d["d_values_in"],
                                   "keys_out": d["d_keys_out"], "values_out": d["d_values_out"]}
    k = _lists_after(seg, 502, 30)
    out["seg_keys_ascending"] = {"source": "lib/cub/device/device_segmented_radix_sort.cuh:506-526",
                                 "offsets": k["d_offsets"], "keys": k["d_keys_in"], "keys_out": k["d_keys_out"]}
    kd = _lists_after(seg, 681, 30)
    out["seg_keys_descending"] = {"source": "lib/cub/device/device_segmented_radix_sort.cuh:685-705",
                                  "offsets": kd["d_offsets"], "keys": kd["d_keys_in"], "keys_out": kd["d_keys_out"]}
    f = _lists_after(flat, 112, 30)
    out["flat_pairs_ascending"] = {"source": "lib/cub/device/device_radix_sort.cuh:115-136",
                                   "offsets": [0, f["num_items"]], "keys": f["d_keys_in"],
                                   "values": f["d_values_in"], "keys_out": f["d_keys_out"],
                                   "values_out": f["d_values_out"]}
    fd = _lists_after(flat, 293, 30)
    out["flat_pairs_descending"] = {"source": "lib/cub/device/device_radix_sort.cuh:296-317",
                                    "offsets": [0, fd["num_items"]], "keys": fd["d_keys_in"],
                                    "values": fd["d_values_in"], "keys_out": fd["d_keys_out"],
                                    "values_out": fd["d_values_out"]}
    for name, v in out.items():
        for field in ("keys", "keys_out", "offsets"):
            assert field in v and v[field], (name, field)
    return out


def read_stdin_format(path: Path):
    tok = path.read_text().split()
    s = int(tok[0])
    offs = [int(x) for x in tok[1:s + 2]]
    n = int(tok[s + 2])
    keys = [int(x) for x in tok[s + 3:s + 3 + n]]
    assert len(keys) == n
    return offs, keys


def main(ref: Path) -> None:
    (OUT / "cub_doc_vectors.json").write_text(json.dumps(cub_vectors(ref), indent=1) + "\n")
    keys = np.array((ref / "src/cub.out").read_text().split(), dtype=np.uint32)
    assert keys.size == 131072, keys.size
    np.savez_compressed(OUT / "cub_out_512x256.npz", keys=keys.astype(np.uint16),
                        offsets=(np.arange(513) * 256).astype(np.int32))
    ms = {}
    for name in ("in1-32", "in2-1024"):
        offs, k = read_stdin_format(ref / "src/mergesort2" / name)
        ms[name] = {"source": f"src/mergesort2/{name}", "offsets": offs, "keys": k}
    (OUT / "mergesort2_inputs.json").write_text(json.dumps(ms) + "\n")
    print("fixtures written to", OUT)


if __name__ == "__main__":
    main(Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference"))
