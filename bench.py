"""Benchmark of the MI355X segmented sort (one GPU).

Default workload: SortPairs of 2^24 uint32 key/value pairs (values = index) in 2^16
segments whose lengths are lognormal(mu=5, sigma=2), clipped to [0, 2^20], rescaled
to sum to 2^24, with at least 5% empty and 5% length-1 segments; keys uniform in
[0, 2^32). A step is one SortPairs call over that whole batch, inputs resident in HBM.

Prints ONE JSON line: the batched form's Gkeys/s as `value`, the plain per-class and
fix-pass forms beside it, the roofline of the dominant kernel (its algorithmic bytes
per launch over its average launch time, timed with HIP events on its stream), the
whole call against the algorithmic-bytes roofline, and the CPU baseline (the oracle's
per-segment std::stable_sort, OpenMP over the host cores).
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
PKG_DIR = ROOT / "segmented-sorting_amd"
HBM_PEAK_GBS = 8000.0
METRIC = "Gkeys/s SortPairs 2^24 u32 ragged segs, % HBM roofline, batched vs plain"
SEED = 0x5E650A7 + 2


def load_pkg():
    name = "segsort_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, PKG_DIR / "__init__.py",
                                                  submodule_search_locations=[str(PKG_DIR)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


WORKLOADS = {
    "ragged": "2^24 u32 pairs, 2^16 lognormal(5,2) segments, keys uniform u32",
    "ragged_dup": "2^24 u32 pairs, 2^16 lognormal(5,2) segments, keys uniform in [0,256)",
    "tiny_f32": "2^20 segments of U[1,32] f32 N(0,1) keys (+-0, denormals), u32 values",
    "tiny_u64": "2^20 segments of U[1,32] u64 keys, u32 values",
    "diff": "differentiable pair sort: 4096 rows of U[16,8192] f32 N(0,1), forward + backward",
    "skew_f32": "one segment of 2^24 f32 N(0,1) keys (skewed top digit), u32 values",
}


def make_workload(pkg, name):
    g = pkg.gen
    if name == "ragged":
        return g.ragged_lognormal(SEED, dup_heavy=False)
    if name == "ragged_dup":
        return g.ragged_lognormal(SEED, dup_heavy=True)
    if name.startswith("equal"):
        k = int(name.split(":")[1])
        return g.equal_segments(1 << 24, 1 << k, seed=0x5E650A7 + 1)
    if name == "keys_2e20":  # keys only: 2^20 u32 in [0, 2^31), 1024 equal segments
        return g.equal_segments(1 << 20, 1024, seed=0x5E650A7 + 0, key_bits=31, pairs=False)
    if name == "tiny_f32":
        return g.tiny_segments(0x5E650A7 + 3, key="f32")
    if name == "tiny_u64":
        return g.tiny_segments(0x5E650A7 + 3, key="u64")
    if name == "skew_f32":  # one 2^24 segment of N(0,1) f32 keys (skewed top digit), values = index
        n = 1 << 24
        return g.normal(0x5E650A7 + 14, 1, n).astype(np.float32), np.arange(n, dtype=np.uint32), \
            np.array([0, n], dtype=np.int32)
    if name == "diff_rows":  # the differentiable sort's forward pass: f32 rows, row-local indices as values
        x, off = g.diff_rows(0x5E650A7 + 4)
        local = np.concatenate([np.arange(b - a, dtype=np.uint32) for a, b in zip(off[:-1], off[1:])])
        return x, local, off
    raise SystemExit(f"unknown workload {name}")


def to_dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).copy()).cuda()


class Runner:
    """One sort form on device-resident copies of a workload. "batched_bounded" is the
    batched form called with the workload's longest segment as its length bound
    (segsort_sort_hinted), as a caller that holds the offsets on the host can."""

    def __init__(self, pkg, torch, keys, vals, off, algo):
        self.pkg, self.torch = pkg, torch
        self.bound = int(np.diff(off).max()) if algo == "batched_bounded" and off.size > 1 else 0
        if algo == "batched_bounded":
            algo = "batched"
        self.n, self.s = keys.size, off.size - 1
        self.kt = {np.dtype(np.uint32): 0, np.dtype(np.float32): 2, np.dtype(np.uint64): 3}[keys.dtype]
        self.vb = 0 if vals is None else vals.dtype.itemsize
        self.kin = to_dev(torch, keys)
        self.vin = to_dev(torch, vals) if vals is not None else None
        self.kout = torch.empty_like(self.kin)
        self.vout = torch.empty_like(self.vin) if vals is not None else None
        self.off = to_dev(torch, off.astype(np.int32))
        self.begin = self.off.data_ptr()
        self.end = self.off.data_ptr() + 4
        self.algo = algo
        self.eb = keys.dtype.itemsize * 8
        tb = [0]
        rc = pkg.DeviceSegmentedRadixSort._call(None, tb, None, None, None, None, self.n, self.s, None, None, 0,
                                                self.eb, None, 0, self.kt, self.vb, algo, self.bound)
        assert rc == 0, pkg.last_error()
        self.temp = torch.empty(tb[0], dtype=torch.uint8, device="cuda")
        self.stream = torch.cuda.current_stream().cuda_stream

    def step(self):
        tb = [self.temp.numel()]
        rc = self.pkg.DeviceSegmentedRadixSort._call(self.temp, tb, self.kin, self.kout, self.vin, self.vout, self.n,
                                                     self.s, self.begin, self.end, 0, self.eb, self.stream, 0,
                                                     self.kt, self.vb, self.algo, self.bound)
        if rc != 0:
            raise RuntimeError(f"{self.algo}: hipError {rc}: {self.pkg.last_error()}")

    def time(self, steps, warmup):
        torch = self.torch
        for _ in range(warmup):
            self.step()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record()
        for _ in range(steps):
            self.step()
        e1.record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        return e0.elapsed_time(e1) / steps, wall / steps * 1e3


class MultiRunner(Runner):
    """The per-segment multi-stream baseline (segsort_multistream_sort): one device-wide
    sort per segment, dealt over 4 streams to the least-loaded one (the reference's
    ncub.cu / newthrust.cu loops)."""

    STREAMS, SCHEDULE = 4, 1

    def __init__(self, pkg, torch, keys, vals, off, algo="multistream"):
        import ctypes
        self.pkg, self.torch = pkg, torch
        self.n, self.s = keys.size, off.size - 1
        self.kt = {np.dtype(np.uint32): 0, np.dtype(np.float32): 2, np.dtype(np.uint64): 3}[keys.dtype]
        self.vb = 0 if vals is None else vals.dtype.itemsize
        self.kin = to_dev(torch, keys)
        self.vin = to_dev(torch, vals) if vals is not None else None
        self.kout = torch.empty_like(self.kin)
        self.vout = torch.empty_like(self.vin) if vals is not None else None
        self.off = to_dev(torch, off.astype(np.int32))
        self.h_off = np.ascontiguousarray(off, dtype=np.int32)
        self.algo = algo
        self.eb = keys.dtype.itemsize * 8
        self.lib = pkg.load_library()
        self.tb = ctypes.c_size_t(0)
        rc = self.lib.segsort_multistream_sort(None, ctypes.byref(self.tb), None, None, None, None, *self._args())
        assert rc == 0, pkg.last_error()
        self.temp = torch.empty(self.tb.value, dtype=torch.uint8, device="cuda")

    def _args(self):
        st = self.torch.cuda.current_stream().cuda_stream
        return (self.n, self.s, self.h_off.ctypes.data, 0, self.eb, self.kt, self.vb, 0, self.STREAMS, self.SCHEDULE,
                st)

    def step(self):
        import ctypes
        rc = self.lib.segsort_multistream_sort(self.temp.data_ptr(), ctypes.byref(self.tb), self.kin.data_ptr(),
                                               self.kout.data_ptr(), None if self.vin is None else self.vin.data_ptr(),
                                               None if self.vout is None else self.vout.data_ptr(), *self._args())
        if rc != 0:
            raise RuntimeError(f"multistream: hipError {rc}: {self.pkg.last_error()}")


def sorted_within_segments(torch, r, key_dtype):
    """Cheap output property check on the GPU: keys non-decreasing inside every segment."""
    if key_dtype != np.uint32:
        return True
    k = r.kout.view(torch.int32).to(torch.int64) & 0xFFFFFFFF
    off = r.off.view(torch.int32).to(torch.int64)
    head = torch.zeros(r.n, dtype=torch.bool, device="cuda")
    starts = off[:-1][(off[1:] - off[:-1]) > 0]
    head[starts] = True
    ok = (k[1:] >= k[:-1]) | head[1:]
    return bool(ok.all().item())


def cpu_baseline(keys, vals, off, budget_s=12.0):
    """The oracle's per-segment std::stable_sort on the host cores (OpenMP, all threads),
    repeated over the full workload until ~budget_s of CPU wall time is spent."""
    sys.path.insert(0, str(ROOT))
    from oracle import oracle as orc  # checker / baseline only

    threads = orc.max_threads()
    reps, total = 0, 0.0
    while total < budget_s and reps < 20:
        total += orc.time_seg_sort(keys, vals, off, threads)
        reps += 1
    t = total / reps
    cpu_model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu_model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    # the sequential form of the same restatement (one core), a shorter sample
    reps1, total1 = 0, 0.0
    while total1 < budget_s / 3 and reps1 < 5:
        total1 += orc.time_seg_sort(keys, vals, off, 1)
        reps1 += 1
    out = {"value": keys.size / t / 1e9, "unit": "Gkeys/s", "cores": threads, "kind": "port",
           "sample": f"full workload ({keys.size} pairs), {reps} reps, OpenMP dynamic over segments, "
                     f"std::stable_sort per segment, {cpu_model}",
           "single_core": {"value": round(keys.size / (total1 / reps1) / 1e9, 4), "cores": 1,
                           "sample": f"full workload, {reps1} reps"}}
    # the fix-pass restated on the host (fixseq.cu:61-83, widened to a 64-bit
    # segment|key composite): one global std::stable_sort on one core
    if keys.dtype.itemsize == 4 and np.all(off[1:] >= off[:-1]):
        t0 = time.perf_counter()
        orc.fixpass_sort(keys, vals, off)
        tf = time.perf_counter() - t0
        out["fixpass_single_core"] = {"value": round(keys.size / tf / 1e9, 4), "cores": 1,
                                      "sample": "full workload, 1 rep, 64-bit composite std::stable_sort"}
    return out


def profiler_name(kernel_name):
    """The library profiler's name of a kernel, from its demangled name in a PMC trace."""
    base = kernel_name.split("(")[0].split("<")[0].split("::")[-1]
    m = re.search(r"k_tile_tier<(\d+), (\d+)", kernel_name)
    if m:
        return f"tile_{m.group(1)}x{m.group(2)}"
    m = re.search(r"k_group_tier<(\d+),", kernel_name)
    if m:
        return f"group_{m.group(1)}"
    return {"k_wave_tier": "wave_tier", "k_wave_tile_tier": "wave_tile", "k_bin": "bin", "k_init": "init",
            "k_large_hist": "large_hist",
            "k_large_scan": "large_scan", "k_large_scatter": "large_scatter",
            "k_large_local": "large_local"}.get(base)


def traffic_from_pmc(fetch_csv, write_csv, copy_mib):
    """HBM bytes per launch of each library kernel from one FETCH_SIZE and one WRITE_SIZE
    rocprofv3 pass over `bench.py --only batched --copy-calibration MIB`. Each counter
    is scaled so that the plain device copy of MIB MiB at the start of the run reads
    and writes exactly MIB MiB; the averages run over every launch of a kernel (the
    per-launch figure the in-bench profiler divides by)."""
    import csv

    def per_kernel(path, counter):
        rows = [r for r in csv.DictReader(open(path)) if r["Counter_Name"] == counter]
        copy = [float(r["Counter_Value"]) for r in rows if "copyBuffer" in r["Kernel_Name"]]
        if len(copy) != 1:
            raise SystemExit(f"{path}: expected one calibration copy dispatch, found {len(copy)}")
        scale = copy_mib * 1024.0 / copy[0]  # counter reads KiB
        acc = {}
        for r in rows:
            name = profiler_name(r["Kernel_Name"])
            if name:
                t = acc.setdefault(name, [0.0, 0])
                t[0] += float(r["Counter_Value"]) * scale * 1024.0
                t[1] += 1
        return scale, {k: v[0] / v[1] for k, v in acc.items()}

    fs, fetch = per_kernel(fetch_csv, "FETCH_SIZE")
    ws, write = per_kernel(write_csv, "WRITE_SIZE")
    return {"fetch_scale": fs, "write_scale": ws,
            "bytes": {k: int(round(fetch[k] + write.get(k, 0.0))) for k in fetch},
            "fetch": {k: int(round(v)) for k, v in fetch.items()},
            "write": {k: int(round(v)) for k, v in write.items()}}


def read_traffic(workload, kernel):
    p = ROOT / "profiles" / "traffic.json"
    if not p.exists():
        return None
    try:
        d = json.loads(p.read_text())
        return d.get(workload, {}).get(kernel)
    except (ValueError, OSError):
        return None


def bench_diff(pkg, torch, steps, warmup, cpu=False):
    """Forward + backward of the differentiable pair sort per step (4096 rows); with
    cpu, a per-row stable torch.sort on the host (forward) and the gradient scatter
    through its indices (backward) beside it."""
    x, off = pkg.gen.diff_rows(0x5E650A7 + 4)
    pkg.autotune.apply(pkg, x.size, off.size - 1, 4, 4)  # the forward's pair sort: f32 keys, int32 indices
    xd = torch.from_numpy(x).cuda().requires_grad_(True)
    od = torch.from_numpy(off).cuda()
    g = torch.from_numpy(pkg.gen.normal(98, 1, x.size).astype(np.float32)).cuda()

    def step():
        y, _ = pkg.segmented_pair_sort(xd, od, assume_cover=True)  # the generator's rows cover x
        return torch.autograd.grad(y, xd, g)[0]

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        step()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    n = x.size
    algo_bytes = 2 * 12 * n  # forward: read x, write y + perm; backward: read g + perm, write grad
    row = {"workload": WORKLOADS["diff"], "n": n, "rows": off.size - 1, "ms_fwd_bwd": round(ms, 4),
           "gkeys_s": round(n / (ms * 1e-3) / 1e9, 3), "algorithmic_bytes": algo_bytes,
           "hbm_frac": round(algo_bytes / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4)}
    if cpu:
        xc = torch.from_numpy(x)
        gc = g.cpu()
        gin = torch.empty_like(xc)
        threads = torch.get_num_threads()
        t0 = time.perf_counter()
        for a, b in zip(off[:-1].tolist(), off[1:].tolist()):
            _, idx = torch.sort(xc[a:b], stable=True)
            gin[a:b][idx] = gc[a:b]
        t = time.perf_counter() - t0
        row["cpu_baseline"] = {"value": round(n / t / 1e9, 4), "unit": "Gkeys/s", "cores": threads,
                               "kind": "port", "sample": "all 4096 rows, 1 run: torch.sort(stable) per row on the "
                                                         "host, gradient scattered through its indices"}
    return row


def sweep(pkg, torch, args):
    """One JSON line per k: ms and Gkeys/s of every form on 2^24 u32 pairs (values =
    index) in 2^k equal segments, keys uniform in [0, 2^32)."""
    rows = []
    for k in range(0, 21, 2):
        keys, vals, off = make_workload(pkg, f"equal:{k}")
        n, s = keys.size, off.size - 1
        tuning = pkg.autotune.apply(pkg, n, s, 4, 4)
        row = {"k": k, "segments": s, "segment_length": n // s, "tuning": tuning}
        for a in [a for a in args.algos.split(",") if a]:
            r = Runner(pkg, torch, keys, vals, off, a)
            ms, _ = r.time(args.steps, args.warmup)
            row[a] = {"ms": round(ms, 4), "gkeys_s": round(n / (ms * 1e-3) / 1e9, 3),
                      "hbm_frac": round((n * 16 + 8 * s) / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4),
                      "sorted": sorted_within_segments(torch, r, keys.dtype)}
            del r
            torch.cuda.empty_cache()
        if "batched" in row and "plain" in row:
            row["batched_over_plain"] = round(row["plain"]["ms"] / row["batched"]["ms"], 3)
        if "batched" in row and "fixpass" in row:
            row["batched_over_fixpass"] = round(row["fixpass"]["ms"] / row["batched"]["ms"], 3)
        print(json.dumps(row), flush=True)
        rows.append(row)
    return rows


def algo_bytes_of(keys, vals, s):
    """One read and one write of every key and value plus the begin/end offsets."""
    return keys.size * 2 * (keys.dtype.itemsize + (0 if vals is None else vals.dtype.itemsize)) + 8 * s


# The configurations of the reference's headline benchmark, each timed in every form
# that applies: name -> (workload, description)
CONFIGS = [
    ("C1", "keys_2e20", "keys only, 2^20 u32 in [0,2^31), 1024 equal segments"),
] + [(f"C2 k={k}", f"equal:{k}", f"2^24 u32 pairs in 2^{k} equal segments, values = index") for k in range(0, 21, 2)] + [
    ("C3 uniform", "ragged", WORKLOADS["ragged"]),
    ("C3 dup", "ragged_dup", WORKLOADS["ragged_dup"]),
    ("C4a", "tiny_f32", WORKLOADS["tiny_f32"]),
    ("C4b", "tiny_u64", WORKLOADS["tiny_u64"]),
    ("skew", "skew_f32", WORKLOADS["skew_f32"]),
]


MULTISTREAM_MAX_SEGMENTS = 4096  # configurations with more segments skip the per-segment baseline


def time_config(pkg, torch, orc, workload, steps, warmup, cpu=True):
    """ms, Gkeys/s and the end-to-end bytes-moved roofline of every form on one workload,
    with the oracle's per-segment stable sort (all host threads, one run) beside it."""
    keys, vals, off = make_workload(pkg, workload)
    n, s = keys.size, off.size - 1
    ab = algo_bytes_of(keys, vals, s)
    algos = ["batched", "batched_bounded", "plain"] + (["fixpass"] if keys.dtype.itemsize == 4 else [])
    vbytes = 0 if vals is None else vals.dtype.itemsize
    row = {"n": n, "segments": s, "algorithmic_bytes": ab,
           "tuning": pkg.lookup_tuning(n, s, keys.dtype.itemsize, vbytes),
           "tuning_bounded": pkg.lookup_tuning(n, s, keys.dtype.itemsize, vbytes, int(np.diff(off).max())),
           "forms": {}}
    for a in algos:
        r = Runner(pkg, torch, keys, vals, off, a)
        ms, _ = r.time(steps, warmup)
        row["forms"][a] = {"ms": round(ms, 4), "gkeys_s": round(n / (ms * 1e-3) / 1e9, 3),
                           "hbm_frac": round(ab / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4),
                           "sorted": sorted_within_segments(torch, r, keys.dtype)}
        del r
        torch.cuda.empty_cache()
    if s <= MULTISTREAM_MAX_SEGMENTS:
        # the per-segment baseline: a few steps (it launches one sort per segment)
        r = MultiRunner(pkg, torch, keys, vals, off)
        ms, _ = r.time(max(1, min(steps, 5)), 1)
        row["forms"]["multistream"] = {"ms": round(ms, 4), "gkeys_s": round(n / (ms * 1e-3) / 1e9, 3),
                                       "hbm_frac": round(ab / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4),
                                       "sorted": sorted_within_segments(torch, r, keys.dtype),
                                       "streams": r.STREAMS, "schedule": "least-loaded"}
        del r
        torch.cuda.empty_cache()
    row["batched_over_plain"] = round(row["forms"]["plain"]["ms"] / row["forms"]["batched"]["ms"], 3)
    if "fixpass" in row["forms"]:
        row["batched_over_fixpass"] = round(row["forms"]["fixpass"]["ms"] / row["forms"]["batched"]["ms"], 3)
    if cpu:
        threads = orc.max_threads()
        t = orc.time_seg_sort(keys, vals, off, threads)
        row["cpu_baseline"] = {"value": round(n / t / 1e9, 4), "unit": "Gkeys/s", "cores": threads, "kind": "port",
                               "sample": "full workload, 1 run, std::stable_sort per segment, OpenMP over segments"}
    return row


def config_suite(pkg, torch, steps, warmup, cpu=True, done=None):
    """Every configuration of CONFIGS (C3 uniform is reused from `done` when given)."""
    orc = None
    if cpu:
        sys.path.insert(0, str(ROOT))
        from oracle import oracle as orc  # CPU baseline only
    out = {}
    for name, wl, desc in CONFIGS:
        if done is not None and wl in done:
            out[name] = dict(done[wl], workload=desc)
            continue
        out[name] = dict(time_config(pkg, torch, orc, wl, steps, warmup, cpu), workload=desc)
        f = out[name]["forms"]
        print(f"[bench] {name}: " + ", ".join(f"{a} {v['ms']:.3f} ms" for a, v in f.items()), file=sys.stderr,
              flush=True)
    return out


def sweep_ns(pkg, torch, steps, warmup):
    """The reference's N x S sweep (scripts/genequalseg.sh: N = 2^15 .. 2^27, S = 1, 2, 4, ...
    while S < N and S <= 2^20): batched SortPairs of N u32 keys (uniform, drawn on the
    device) with values = index in S equal segments; one JSON line per (N, S)."""
    g = torch.Generator(device="cuda")
    g.manual_seed(0x5E650A7)
    rows = []
    for ln in range(15, 28):
        n = 1 << ln
        kin = torch.randint(0, 1 << 32, (n,), dtype=torch.int64, device="cuda", generator=g).to(torch.int32)
        vin = torch.arange(n, dtype=torch.int32, device="cuda")
        kout, vout = torch.empty_like(kin), torch.empty_like(vin)
        s = 1
        while s < n and s <= (1 << 20):
            off = (torch.arange(s + 1, dtype=torch.int64, device="cuda") * (n // s)).to(torch.int32)
            tb = [0]
            D = pkg.DeviceSegmentedRadixSort
            assert D._call(None, tb, None, None, None, None, n, s, None, None, 0, 32, None, 0, 0, 4, "batched") == 0
            temp = torch.empty(tb[0], dtype=torch.uint8, device="cuda")
            st = torch.cuda.current_stream().cuda_stream

            def step():
                rc = D._call(temp, [temp.numel()], kin, kout, vin, vout, n, s, off.data_ptr(), off.data_ptr() + 4, 0,
                             32, st, 0, 0, 4, "batched")
                if rc:
                    raise RuntimeError(pkg.last_error())

            for _ in range(warmup):
                step()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(steps):
                step()
            e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / steps
            k = kout.to(torch.int64) & 0xFFFFFFFF
            seg = torch.arange(n, device="cuda") // (n // s)
            ok = bool((((k[1:] >= k[:-1]) | (seg[1:] != seg[:-1]))).all().item())
            row = {"n": n, "segments": s, "ms": round(ms, 4), "gkeys_s": round(n / (ms * 1e-3) / 1e9, 3),
                   "hbm_frac": round((16 * n + 8 * s) / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4), "sorted": ok}
            print(json.dumps(row), flush=True)
            rows.append(row)
            del temp
            s *= 2
        del kin, vin, kout, vout
        torch.cuda.empty_cache()
    return rows


def driver_times(pkg, torch, args):
    """The workload through segmented-sorting_amd/bin/radixseg (the reference's radixseg.cu
    driver over this library): its input written in the reference's stdin format, then
    `--time` (hipEvents around the SortPairs call only, like -DELAPSED_TIME) in each mode,
    beside the library call timed back to back by this script."""
    import subprocess
    import tempfile

    keys, vals, off = make_workload(pkg, args.workload)
    if keys.dtype != np.uint32 or vals is None:
        raise SystemExit("--driver: radixseg sorts u32 pairs")
    exe = PKG_DIR / "bin" / "radixseg"
    r = Runner(pkg, torch, keys, vals, off, "batched")
    lib_ms, _ = r.time(args.steps, args.warmup)
    del r
    torch.cuda.empty_cache()
    out = {"workload": WORKLOADS.get(args.workload, args.workload), "library_batched_ms": round(lib_ms, 4), "driver": {}}
    with tempfile.NamedTemporaryFile("w", suffix=".in", delete=True) as f:
        f.write(f"{off.size - 1}\n" + " ".join(map(str, off.tolist())) + f"\n{keys.size}\n")
        f.write(" ".join(map(str, keys.tolist())) + "\n")
        f.flush()
        modes = {"reference (upload before each call, first call timed)": [],
                 "upload before each call, 3 untimed warm-up calls": ["--warmup", "3"],
                 "inputs resident, 3 warm-up calls": ["--warmup", "3", "--resident"],
                 "inputs resident, 3 warm-up calls, length bound": ["--warmup", "3", "--resident", "--bound"],
                 "plain, inputs resident, 3 warm-up calls": ["--warmup", "3", "--resident", "--algo", "plain"]}
        for name, extra in modes.items():
            with open(f.name) as inp:
                p = subprocess.run([str(exe), "--time", "--executions", str(args.steps), *extra], stdin=inp,
                                   capture_output=True, text=True, timeout=600)
            if p.returncode != 0:
                raise SystemExit(f"radixseg {extra} failed: {p.stderr}")
            out["driver"][name] = round(float(p.stdout.strip()), 4)
            print(f"[bench] radixseg {' '.join(extra)}: {p.stdout.strip()} ms", file=sys.stderr, flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="ragged", help="ragged | ragged_dup | equal:<k> | tiny_f32 | tiny_u64")
    ap.add_argument("--algos", default="batched,batched_bounded,plain,fixpass")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--only", default=None, help="time one algorithm and print its ms (profiling runs)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run one plain device copy of MIB MiB (HBM-counter calibration in profiling runs)")
    ap.add_argument("--traffic-from", nargs=2, default=None, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="reduce two rocprofv3 --pmc counter CSVs (FETCH_SIZE, WRITE_SIZE) of a --copy-calibration "
                         "run to HBM bytes per launch, store them for --workload in profiles/traffic.json (no GPU)")
    ap.add_argument("--sweep", action="store_true",
                    help="equal-segment sweep: 2^24 u32 pairs in 2^k segments, k = 0,2,...,20, every form")
    ap.add_argument("--sweep-ns", action="store_true",
                    help="N x S sweep of the batched form: N = 2^15..2^27 u32 pairs in S = 1, 2, 4, ... equal segments")
    ap.add_argument("--driver", action="store_true",
                    help="time the workload through the rebuilt radixseg driver (stdin format, --time) beside the "
                         "library call: uploads before every call as the reference does, and inputs resident")
    ap.add_argument("--no-configs", action="store_true",
                    help="headline workload only (skip the other configurations of the benchmark)")
    args = ap.parse_args()
    if args.traffic_from:
        t = traffic_from_pmc(*args.traffic_from, args.copy_calibration or 256)
        p = ROOT / "profiles" / "traffic.json"
        d = json.loads(p.read_text()) if p.exists() else {}
        d[args.workload] = t["bytes"]
        p.write_text(json.dumps(d, indent=1, sort_keys=True) + "\n")
        print(json.dumps(t, indent=1, sort_keys=True))
        return
    if args.gpus != 1:
        print(f"bench.py: measures one GPU; --gpus {args.gpus} is not supported (segments are independent, so "
              "each GPU would sort its own contiguous share of them with no exchange)", file=sys.stderr)
        sys.exit(2)

    import torch

    pkg = load_pkg()
    pkg.load_library()
    if args.sweep:
        sweep(pkg, torch, args)
        return
    if args.sweep_ns:
        sweep_ns(pkg, torch, args.steps, args.warmup)
        return
    if args.driver:
        print(json.dumps(driver_times(pkg, torch, args)), flush=True)
        return
    if args.workload == "diff":
        print(json.dumps(bench_diff(pkg, torch, args.steps, args.warmup)), flush=True)
        return
    keys, vals, off = make_workload(pkg, args.workload)
    n, s = keys.size, off.size - 1
    item_bytes = keys.dtype.itemsize + vals.dtype.itemsize
    algo_bytes = n * 2 * item_bytes + 8 * s  # one read + one write of every pair, begin+end offsets
    tuning = pkg.autotune.apply(pkg, n, s, keys.dtype.itemsize, vals.dtype.itemsize)
    if args.copy_calibration:
        a = torch.empty(args.copy_calibration << 18, dtype=torch.float32, device="cuda").fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    if args.only:
        r = Runner(pkg, torch, keys, vals, off, args.only)
        ms, wall = r.time(args.steps, args.warmup)
        print(json.dumps({"algo": args.only, "workload": args.workload, "ms": ms, "wall_ms": wall}), flush=True)
        return

    algos = [a for a in args.algos.split(",") if a]
    if keys.dtype.itemsize == 8 and "fixpass" in algos:
        algos.remove("fixpass")
    res = {}
    runners = {}
    for a in algos:
        r = Runner(pkg, torch, keys, vals, off, a)
        ms, wall = r.time(args.steps, args.warmup)
        ok = sorted_within_segments(torch, r, keys.dtype)
        res[a] = {"ms": ms, "wall_ms_per_step": wall, "gkeys_s": n / (ms * 1e-3) / 1e9,
                  "hbm_frac": algo_bytes / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), "sorted": ok}
        runners[a] = r
        print(f"[bench] {a}: {ms:.3f} ms/step, {res[a]['gkeys_s']:.1f} Gkeys/s, sorted={ok}", file=sys.stderr,
              flush=True)

    # per-kernel HIP-event timing of the batched form over the same number of steps
    rb = runners["batched"]
    pkg.profile(True)
    for _ in range(args.steps):
        rb.step()
    torch.cuda.synchronize()
    rep = pkg.profile_report()
    pkg.profile(False)
    dom = max(rep.items(), key=lambda kv: kv[1]["ms"])
    dname, d = dom
    avg_ms = d["ms"] / d["launches"]
    bytes_per_launch = d["bytes"] / d["launches"]
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    kernels = {k: {"launches": v["launches"], "avg_us": v["ms"] / v["launches"] * 1e3,
                   "share": v["ms"] / sum(x["ms"] for x in rep.values()),
                   "gb_s": (v["bytes"] / v["launches"]) / (v["ms"] / v["launches"] * 1e-3) / 1e9 if v["ms"] else 0}
               for k, v in rep.items()}

    # measured device-copy ceiling beside the nominal peak: 1 GiB device to device,
    # read + write bytes over the copy time
    a = torch.empty(1 << 28, dtype=torch.float32, device="cuda").fill_(1.0)
    c = torch.empty_like(a)
    c.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(10):
        c.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    copy_gb_s = 2 * a.numel() * 4 / (e0.elapsed_time(e1) / 10 * 1e-3) / 1e9
    del a, c

    ms = res["batched"]["ms"]
    out = {
        "metric": METRIC,
        "value": round(res["batched"]["gkeys_s"], 3),
        "unit": "Gkeys/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32" if keys.dtype == np.uint32 else str(keys.dtype),
        "data": "synthetic (seeded splitmix64)",
        "config": {"workload": WORKLOADS.get(args.workload, args.workload), "num_items": n, "num_segments": s,
                   "algorithm": "batched", "parallelism": "single GPU", "tuning": tuning},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": read_traffic(args.workload, dname),
                     "kernel": dname, "avg_launch_us": round(avg_ms * 1e3, 2),
                     "algorithmic_bytes_per_launch": int(bytes_per_launch),
                     "measured_copy_gb_s": round(copy_gb_s, 1)},
        "end_to_end": {"algorithmic_bytes": algo_bytes,
                       "achieved_gb_s": round(algo_bytes / (ms * 1e-3) / 1e9, 1),
                       "frac": round(res["batched"]["hbm_frac"], 4)},
        "forms": {a: {"ms": round(v["ms"], 4), "wall_ms": round(v["wall_ms_per_step"], 4),
                      "gkeys_s": round(v["gkeys_s"], 3), "hbm_frac": round(v["hbm_frac"], 4),
                      "sorted": v["sorted"]} for a, v in res.items()},
        "batched_over_plain": round(res["plain"]["ms"] / ms, 3) if "plain" in res else None,
        "kernels": {k: {kk: round(vv, 4) if isinstance(vv, float) else vv for kk, vv in v.items()}
                    for k, v in kernels.items()},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(keys, vals, off)
        out["cpu_baseline"]["value"] = round(out["cpu_baseline"]["value"], 4)
    if not args.no_configs and args.workload == "ragged":
        # every configuration of the benchmark, each form, steps capped for run time
        head = {"n": n, "segments": s, "algorithmic_bytes": algo_bytes, "tuning": tuning,
                "forms": {a: {"ms": round(v["ms"], 4), "gkeys_s": round(v["gkeys_s"], 3),
                              "hbm_frac": round(v["hbm_frac"], 4), "sorted": v["sorted"]} for a, v in res.items()},
                "batched_over_plain": out["batched_over_plain"]}
        if "fixpass" in res:
            head["batched_over_fixpass"] = round(res["fixpass"]["ms"] / ms, 3)
        if "cpu_baseline" in out:
            head["cpu_baseline"] = {k: out["cpu_baseline"][k] for k in ("value", "unit", "cores", "kind", "sample")}
        del runners, rb
        torch.cuda.empty_cache()
        csteps, cwarm = min(args.steps, 10), min(args.warmup, 3)
        out["configs"] = config_suite(pkg, torch, csteps, cwarm, cpu=not args.no_cpu, done={"ragged": head})
        out["configs"]["C5"] = bench_diff(pkg, torch, csteps, cwarm, cpu=not args.no_cpu)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
