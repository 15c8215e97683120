"""Autotuner of the batched segmented sort's tier cut-offs and tile classes.

The library's tuning (per key width and value width) decides which tier sorts a
range: the enabled wave-group classes (`tile_enable` bits 4..7: 64/G ranges of at
most G = 4, 8, 16, 32 items per wave), the wave tier up to `wave_cap` items, then
the enabled LDS tile classes (`tile_enable` bits 0..3, capacities from
segsort_tile_caps), and the multi-workgroup MSD tier above the largest enabled
class; `grid_mult` sizes the persistent grids.
`concurrent` (the binned ranges sorted on a second stream while the MSD tier
runs), the larger grid_mult values, `tile_enable` bit 8 (the 16-item tile
classes take their ranges longest first, from length buckets), bit 9 (the
4096-item tile class without its half-depth build) and bit 10 (a wave cap above
256 split into a class of ranges of at most 256 items and one of longer ranges,
each with its own build and launch) are tried on the finalists of the screening
stage. Candidates are timed back to back; a winner that sorts on the second stream
is stored only if it also beats the best one-stream candidate on single
synchronised calls (`keep_overlap`), where the overlap has measured slower.

`tune()` times every candidate on a resident workload and keeps the fastest one
whose output is bit-identical to the default tuning's. Results are stored in
tuning.json keyed by a signature the host knows without reading device memory:
key width, value width, round(log2 N) and round(log2 mean segment length), plus,
for calls that state a bound on their segment lengths (segsort_sort_hinted),
ceil(log2 of that bound): k4v4-n24-m8 for unbounded calls, k4v4-n24-m8-x17 for
bounded ones. `--bounded` tunes the named workloads as bounded calls (the bound
is each workload's longest segment, as a caller holding the offsets knows it).
build.py compiles the table into the library (csrc/tuning_table.inc), which picks
the entry of each call's signature (or the nearest one) itself, inside the call;
`apply()` reports that choice.

    python autotune.py --workload ragged              # tunes on the named bench workload (GPU)
    python autotune.py --workload ragged --bounded    # the same, as a call with a length bound
"""
from __future__ import annotations

import itertools
import json
import math
import sys
from pathlib import Path

TABLE = Path(__file__).resolve().parent / "tuning.json"
DEFAULT = {"wave_cap": 256, "tile_enable": 0xEF, "grid_mult": 1, "concurrent": 0}


def ceil_log2(v: int) -> int:
    """The smallest x with v <= 2^x."""
    return max(int(v) - 1, 0).bit_length()


def signature(num_items: int, num_segments: int, key_width: int, value_bytes: int, max_len: int = 0) -> str:
    n = max(num_items, 1)
    mean = max(n / max(num_segments, 1), 1.0)
    sig = f"k{key_width}v{value_bytes}-n{round(math.log2(n))}-m{round(math.log2(mean))}"
    return sig + (f"-x{ceil_log2(max_len)}" if max_len > 0 else "")


def _parse(sig: str):
    parts = sig.split("-")
    return parts[0], int(parts[1][1:]), int(parts[2][1:]), (int(parts[3][1:]) if len(parts) > 3 else -1)


# LDS tile classes (bits 0..3) and wave-group classes (bits 4..7: groups of 4, 8, 16, 32 lanes)
TILE_MASKS = (0xF, 0xB, 0x7, 0xD, 0xE, 0x3)
GROUP_MASKS = (0x00, 0xE0, 0xA0, 0x60, 0xF0)
LPT = 0x100  # bit 8: the 16-item tile classes longest first
NO_HALF = 0x200  # bit 9: the 4096-item tile class without its half-depth build
SPLIT_WAVE = 0x400  # bit 10: wave caps above 256 split into a <= 256 class and a longer one


# persistent grid sizes tried on the finalists (workgroups per resident slot): more than one
# lets a workgroup that finishes early be replaced by a fresh one, which balances the tiers'
# ranges without any per-range atomics
FINAL_GRID_MULTS = (1, 2, 3, 4, 6)


def candidates(quick: bool = False):
    wcs, tms, gms = (128, 256, 512), TILE_MASKS, GROUP_MASKS
    if quick:
        wcs, tms, gms = (128, 512), (0xF, 0xB), (0x00, 0xE0)
    for wc, tm, gm_, mult in itertools.product(wcs, tms, gms, (1,) if quick else (1, 2)):
        yield {"wave_cap": wc, "tile_enable": tm | gm_, "grid_mult": mult, "concurrent": 0}


def load_table(path: Path = TABLE) -> dict:
    if path.exists():
        return json.loads(path.read_text())
    return {}


def lookup(sig: str, table: dict | None = None) -> dict:
    """The stored tuning of the nearest signature with the same key/value widths (the
    library's rule: a bounded signature looks at the bounded entries first)."""
    table = load_table() if table is None else table
    if sig in table:
        return {k: table[sig][k] for k in DEFAULT}
    head, n, m, x = _parse(sig)
    for bounded in ((True, False) if x >= 0 else (False,)):
        best, dist = None, None
        for s, t in table.items():
            h2, n2, m2, x2 = _parse(s)
            if h2 != head or (x2 >= 0) != bounded:
                continue
            d = abs(n2 - n) + 2 * abs(m2 - m) + (2 * abs(x2 - x) if bounded else 0)
            if dist is None or d < dist:
                best, dist = t, d
        if best:
            return {k: best[k] for k in DEFAULT}
    return dict(DEFAULT)


def apply(pkg, num_items: int, num_segments: int, key_width: int, value_bytes: int, table: dict | None = None,
          max_len: int = 0) -> dict:
    """The tuning the library picks for a call of this shape (and length bound, 0:
    none). With `table` (a table other than the built-in one), install its nearest
    entry as an override."""
    if table is not None:
        t = lookup(signature(num_items, num_segments, key_width, value_bytes, max_len), table)
        pkg.set_tuning(key_width, value_bytes, t["wave_cap"], t["tile_enable"], t["grid_mult"], t["concurrent"])
        return t
    pkg.reset_tuning(key_width, value_bytes)
    t = pkg.lookup_tuning(num_items, num_segments, key_width, value_bytes, max_len)
    return {k: t[k] for k in DEFAULT}


def _time(pkg, torch, args, steps, warmup, isolated=False):
    """Milliseconds per call: back to back (the mean of `steps` calls between two
    events), or with `isolated` the median of `steps` calls each timed alone and
    followed by a device synchronisation, as the reference's drivers time a call."""
    (kin, kout, vin, vout, n, s, begin, end, kt, vb, mx) = args
    tb = [0]
    rc = pkg.DeviceSegmentedRadixSort._call(None, tb, None, None, None, None, n, s, None, None, 0, None, None, 0, kt,
                                            vb, "batched", mx)
    if rc:
        raise pkg.SegsortError(pkg.last_error())
    temp = torch.empty(tb[0], dtype=torch.uint8, device="cuda")
    st = torch.cuda.current_stream().cuda_stream

    def step():
        rc = pkg.DeviceSegmentedRadixSort._call(temp, [temp.numel()], kin, kout, vin, vout, n, s, begin, end, 0, None,
                                                st, 0, kt, vb, "batched", mx)
        if rc:
            raise pkg.SegsortError(pkg.last_error())

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    if isolated:
        times = []
        for _ in range(steps):
            e0.record()
            step()
            e1.record()
            torch.cuda.synchronize()
            times.append(e0.elapsed_time(e1))
        return sorted(times)[len(times) // 2]
    e0.record()
    for _ in range(steps):
        step()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def time_isolated(pkg, torch, keys, vals, offsets, tuning, max_len=0, calls=21):
    """Median milliseconds of single synchronised calls of the batched form with `tuning`."""
    kw, vb = keys.element_size(), (0 if vals is None else vals.element_size())
    args = (keys, torch.empty_like(keys), vals, (torch.empty_like(vals) if vals is not None else None),
            keys.numel(), offsets.numel() - 1, offsets[:-1], offsets[1:], pkg.key_type_of(keys.dtype), vb,
            int(max_len))
    pkg.set_tuning(kw, vb, **tuning)
    try:
        return _time(pkg, torch, args, calls, 3, isolated=True)
    finally:
        pkg.reset_tuning(kw, vb)


def keep_overlap(cand: str, score: dict, isolated_ms) -> str:
    """The stored choice when the best candidate `cand` (a JSON key of `score`) sorts on
    a second stream: the candidates are timed back to back, where the overlap can win
    while single synchronised calls (the reference's drivers) get slower, so it is kept
    only if it also beats the best one-stream candidate in `isolated_ms(candidate)`
    (total milliseconds over the signature's workloads); else that candidate."""
    if not json.loads(cand)["concurrent"]:
        return cand
    one = [c for c, v in score.items() if v is not None and not json.loads(c)["concurrent"]]
    if not one:
        return cand
    alt = min(one, key=lambda c: score[c])
    return cand if isolated_ms(cand) < isolated_ms(alt) else alt


def _same(torch, a, b) -> bool:
    return torch.equal(a.view(torch.uint8), b.view(torch.uint8))


def tune(pkg, torch, keys, vals, offsets, steps=5, warmup=2, verbose=True, quick=False, cands=None, max_len=0):
    """Time every candidate (default: all of `candidates(quick)`) on device-resident
    (keys, vals, offsets) tensors, as calls with the length bound max_len (0: none);
    return (best tuning, {candidate: ms}). Candidates whose output differs from the
    default tuning's are rejected (they would be a bug)."""
    kt = pkg.key_type_of(keys.dtype)
    kw = keys.element_size()
    vb = 0 if vals is None else vals.element_size()
    n, s = keys.numel(), offsets.numel() - 1
    kout = torch.empty_like(keys)
    vout = torch.empty_like(vals) if vals is not None else None
    args = (keys, kout, vals, vout, n, s, offsets[:-1], offsets[1:], kt, vb, int(max_len))
    pkg.set_tuning(kw, vb, **DEFAULT)
    _time(pkg, torch, args, 1, 0)
    ref_k, ref_v = kout.clone(), (vout.clone() if vout is not None else None)
    results, best = {}, None
    try:
        for c in (cands if cands is not None else candidates(quick)):
            pkg.set_tuning(kw, vb, **c)
            ms = _time(pkg, torch, args, steps, warmup)
            same = _same(torch, kout, ref_k) and (vout is None or _same(torch, vout, ref_v))
            key = json.dumps(c, sort_keys=True)
            results[key] = ms if same else None
            if verbose:
                print(f"[tune] {c} -> {ms:.4f} ms{'' if same else ' (output differs: rejected)'}", flush=True)
            if same and (best is None or ms < best[1]):
                best = (c, ms)
    finally:
        pkg.reset_tuning(kw, vb)
    return best, results


def largest_tile(caps, tuning: dict) -> int:
    """Longest range the tuning's classes hold (above it: the MSD tier); caps as
    segsort_tile_caps gives them for the key and value widths."""
    t = [c for i, c in enumerate(caps) if tuning["tile_enable"] & (1 << i)]
    return max(t + [min(512, max(2, tuning["wave_cap"]))])


def canonical(tuning: dict, caps, max_len: int) -> dict:
    """The tuning as stored: `concurrent` only matters when the MSD tier runs, which a
    call bounded by the largest tile never launches; there it is stored as 0."""
    t = dict(tuning)
    if 0 < max_len <= largest_tile(caps, t):
        t["concurrent"] = 0
    return t


def store(sig: str, tuning: dict, ms: float, path: Path = TABLE, note: str = "") -> None:
    table = load_table(path)
    table[sig] = dict(tuning, ms=round(ms, 4), note=note)
    path.write_text(json.dumps(table, indent=1, sort_keys=True) + "\n")


def _score(runs):
    """Sum over workloads of each candidate's time relative to that workload's best."""
    score = {}
    for results in runs:
        best = min(v for v in results.values() if v is not None)
        for cand, ms in results.items():
            if ms is None or score.get(cand, 0.0) is None:
                score[cand] = None
            else:
                score[cand] = score.get(cand, 0.0) + ms / best
    return score


def main(argv=None):
    """Two stages per signature: every candidate timed briefly on every workload of
    the signature, then the tier layouts (wave_cap, tile_enable) of the best few
    re-timed with six times the steps at each grid_mult of FINAL_GRID_MULTS, with
    `concurrent` off, on, and on for the 16-item tile classes only, the 16-item tiles in
    binning order and longest first,
    and (when the 4096-item class is on) with and without its half-depth build; the
    stored tuning minimises the sum of times relative to each workload's best."""
    import argparse

    import torch

    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", action="append", default=None)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--finalists", type=int, default=6)
    ap.add_argument("--bounded", action="store_true",
                    help="tune as calls that pass their longest segment as the length bound")
    args = ap.parse_args(argv)
    root = Path(__file__).resolve().parent.parent
    sys.path.insert(0, str(root))
    import bench

    pkg = bench.load_pkg()
    pkg.load_library()
    by_sig: dict[str, list] = {}
    for wl in args.workload or ["ragged"]:
        keys, vals, off = bench.make_workload(pkg, wl)
        kdt = {"u": {4: torch.uint32, 8: torch.uint64}, "f": {4: torch.float32, 8: torch.float64},
               "i": {4: torch.int32, 8: torch.int64}}[keys.dtype.kind][keys.dtype.itemsize]
        dev = (bench.to_dev(torch, keys).view(kdt),
               None if vals is None else bench.to_dev(torch, vals).view(torch.uint32), torch.from_numpy(off).cuda())
        mx = int((off[1:] - off[:-1]).max()) if args.bounded and off.size > 1 else 0
        best, results = tune(pkg, torch, *dev, steps=args.steps, verbose=False, max_len=mx)
        sig = signature(keys.size, off.size - 1, keys.dtype.itemsize, 0 if vals is None else vals.dtype.itemsize, mx)
        by_sig.setdefault(sig, []).append((wl, dev, results, mx))
        print(json.dumps({"workload": wl, "signature": sig, "screen_best": best[0], "ms": round(best[1], 4)}),
              flush=True)
    for sig, runs in by_sig.items():
        score = _score([r[2] for r in runs])  # r = (workload, tensors, results, length bound)
        ranked = sorted((c for c, v in score.items() if v is not None), key=lambda c: score[c])
        shapes = []  # (wave_cap, tile_enable) of the best screened candidates, in rank order
        for c in ranked:
            t = json.loads(c)
            if (t["wave_cap"], t["tile_enable"]) not in shapes:
                shapes.append((t["wave_cap"], t["tile_enable"]))
            if len(shapes) == args.finalists:
                break
        finalists = [{"wave_cap": wc, "tile_enable": te | lp | nh | sw, "grid_mult": gm, "concurrent": k}
                     for wc, te in shapes for gm in FINAL_GRID_MULTS for k in (0, 1, 2) for lp in (0, LPT)
                     for nh in ((0, NO_HALF) if te & 0x2 else (0,))
                     for sw in ((0, SPLIT_WAVE) if wc > 256 else (0,))]
        refined = []
        for wl, dev, _, mx in runs:
            _, res = tune(pkg, torch, *dev, steps=args.steps * 6, verbose=False, cands=finalists, max_len=mx)
            refined.append(res)
            print(json.dumps({"workload": wl, "finalists": {c: round(v, 4) for c, v in res.items() if v}}),
                  flush=True)
        score = _score(refined)
        cand = min((c for c, v in score.items() if v is not None), key=lambda c: score[c])
        cand = keep_overlap(cand, score, lambda c: sum(
            time_isolated(pkg, torch, *dev, json.loads(c), mx) for _, dev, _, mx in runs))
        ms = sum(r[cand] for r in refined) / len(refined)
        head = sig.split("-")[0]
        caps = pkg.tile_caps(int(head[1:].split("v")[0]), int(head.split("v")[1]))
        store(sig, canonical(json.loads(cand), caps, runs[0][3]), ms, note="tuned on " + ", ".join(r[0] for r in runs))
        print(json.dumps({"signature": sig, "chosen": json.loads(cand), "relative": round(score[cand], 4)}),
              flush=True)


if __name__ == "__main__":
    main()
