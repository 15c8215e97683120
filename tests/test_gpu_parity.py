"""GPU parity: the HIP library (through its C-ABI) against the CPU oracle, bit-exact.

Small and mid sizes compare every output element with the oracle; the full-size
workloads compare against the oracle too (it finishes in seconds with OpenMP).
Every case runs the three algorithms (batched, plain per-class, fix-pass).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ALGOS = ("batched", "plain", "fixpass")
KEY_TYPES = {np.dtype(np.uint32): 0, np.dtype(np.int32): 1, np.dtype(np.float32): 2,
             np.dtype(np.uint64): 3, np.dtype(np.int64): 4, np.dtype(np.float64): 5}
FILL = 0xA5


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).copy()).cuda()


def host(t, dtype):
    return t.cpu().numpy().view(dtype)


def gpu_sort(pkg, torch, keys, vals=None, offsets=None, begin=None, end=None, algo="batched", begin_bit=0,
             end_bit=None, descending=False, max_segment_length=0):
    """Run segsort_sort (segsort_sort_hinted with a length bound) through the C-ABI;
    outputs start filled with FILL bytes so that any write outside the segments shows."""
    if begin is None:
        offsets = np.asarray(offsets, dtype=np.int32)
        begin, end = offsets[:-1], offsets[1:]
    begin = np.ascontiguousarray(begin, dtype=np.int32)
    end = np.ascontiguousarray(end, dtype=np.int32)
    kt = KEY_TYPES[keys.dtype]
    vb = 0 if vals is None else vals.dtype.itemsize
    n, s = keys.size, begin.size
    kin, bd, ed = dev(torch, keys), dev(torch, begin), dev(torch, end)
    kout = torch.full_like(kin, FILL)
    vin = dev(torch, vals) if vals is not None else None
    vout = torch.full_like(vin, FILL) if vals is not None else None
    tb = [0]
    D = pkg.DeviceSegmentedRadixSort
    eb = end_bit if end_bit is not None else keys.dtype.itemsize * 8
    rc = D._call(None, tb, None, None, None, None, n, s, None, None, begin_bit, eb, None, int(descending), kt, vb,
                 algo, max_segment_length)
    assert rc == 0, pkg.last_error()
    temp = torch.empty(max(tb[0], 1), dtype=torch.uint8, device="cuda")
    tb = [temp.numel()]
    rc = D._call(temp, tb, kin, kout, vin, vout, n, s, bd, ed, begin_bit, eb, None, int(descending), kt, vb, algo,
                 max_segment_length)
    assert rc == 0, pkg.last_error()
    torch.cuda.synchronize()
    ko = host(kout, keys.dtype)
    vo = host(vout, vals.dtype) if vals is not None else None
    return ko, vo


def oracle_sort(orc, keys, vals=None, offsets=None, begin=None, end=None, begin_bit=0, end_bit=None,
                descending=False):
    fill_k = np.full(keys.nbytes, FILL, dtype=np.uint8).view(keys.dtype)
    fill_v = np.full(vals.nbytes, FILL, dtype=np.uint8).view(vals.dtype) if vals is not None else None
    return orc.seg_sort(keys, vals, offsets=offsets, begin=begin, end=end, begin_bit=begin_bit, end_bit=end_bit,
                        descending=descending, out_keys=fill_k, out_vals=fill_v)


def check(pkg, torch, orc, keys, vals, offsets=None, begin=None, end=None, algos=ALGOS, max_segment_length=0, **kw):
    ek, ev = oracle_sort(orc, keys, vals, offsets, begin, end, **kw)
    for algo in algos:
        if algo == "fixpass" and (keys.dtype.itemsize == 8 or begin is not None):
            continue
        print(f"[case] n={keys.size} algo={algo} keys={keys.dtype} vals={None if vals is None else vals.dtype} {kw} "
              f"bound={max_segment_length}", flush=True)
        gk, gv = gpu_sort(pkg, torch, keys, vals, offsets, begin, end, algo=algo,
                          max_segment_length=max_segment_length, **kw)
        bad = np.flatnonzero(gk.view(np.uint8) != ek.view(np.uint8))
        assert np.array_equal(gk.view(np.uint8), ek.view(np.uint8)), (algo, bad[:10])
        if vals is not None:
            assert np.array_equal(gv.view(np.uint8), ev.view(np.uint8)), (algo, np.flatnonzero(gv != ev)[:10])


def test_cub_doc_vectors_gpu(pkg, torch_cuda, orc):
    import json
    from pathlib import Path
    vec = json.loads((Path(__file__).parent / "golden" / "cub_doc_vectors.json").read_text())
    for name, v in vec.items():
        keys = np.array(v["keys"], dtype=np.int32)
        vals = np.array(v["values"], dtype=np.int32) if "values" in v else None
        desc = "descending" in name
        for algo in ALGOS:
            ko, vo = gpu_sort(pkg, torch_cuda, keys, vals, v["offsets"], algo=algo, descending=desc)
            assert ko.tolist() == v["keys_out"], (name, algo)
            if vals is not None:
                assert vo.tolist() == v["values_out"], (name, algo)


def test_keys_only_2e20_1024_segments(pkg, torch_cuda, orc):
    """2^20 u32 keys uniform in [0, 2^31) in 1024 equal segments, keys only."""
    keys, _, off = pkg.gen.equal_segments(1 << 20, 1024, seed=0x5E650A7 + 0, key_bits=31, pairs=False)
    check(pkg, torch_cuda, orc, keys, None, off)


@pytest.mark.parametrize("k", [0, 4, 8, 12, 16, 20])
def test_equal_segments_pairs_reduced(pkg, torch_cuda, orc, k):
    """Equal-segment sweep at 2^20 pairs (segment length 2^(20-k))."""
    n = 1 << 20
    keys, vals, off = pkg.gen.equal_segments(n, 1 << k, seed=0x5E650A7 + 1 + k)
    check(pkg, torch_cuda, orc, keys, vals, off)


@pytest.mark.parametrize("k", list(range(0, 21, 2)))
def test_equal_segments_pairs_2e24(pkg, torch_cuda, orc, k):
    """2^24 u32 pairs (values = index) in 2^k equal segments, every form against the oracle."""
    keys, vals, off = pkg.gen.equal_segments(1 << 24, 1 << k, seed=0x5E650A7 + 1)
    check(pkg, torch_cuda, orc, keys, vals, off)


@pytest.mark.parametrize("dup", [False, True])
def test_ragged_lognormal_2e24(pkg, torch_cuda, orc, dup):
    """2^24 pairs in 2^16 lognormal segments (5% empty, 5% length 1), uniform and [0,256) keys."""
    keys, vals, off = pkg.gen.ragged_lognormal(0x5E650A7 + 2, dup_heavy=dup)
    check(pkg, torch_cuda, orc, keys, vals, off)


@pytest.mark.parametrize("order", ["table", "flipped"])
@pytest.mark.parametrize("which", ["ragged", "ragged_dup", "equal_256"])
def test_table_tuning_2e24(pkg, torch_cuda, orc, which, order):
    """The 2^24-pair workloads of the stored tuning table's headline signature, sorted with
    the tuning the table picks for them (as the benchmark runs them: several persistent
    workgroups per slot, the binned ranges on a second stream while the MSD tier runs),
    twice in a row, bit-exact against the oracle; `flipped`: the same tuning with the
    16-item tiles' order switched (longest first <-> binning order)."""
    if which == "equal_256":
        keys, vals, off = pkg.gen.equal_segments(1 << 24, 1 << 16, seed=0x5E650A7 + 1)
    else:
        keys, vals, off = pkg.gen.ragged_lognormal(0x5E650A7 + 2, dup_heavy=which == "ragged_dup")
    try:
        t = pkg.autotune.apply(pkg, keys.size, off.size - 1, 4, 4)  # the library's own table choice
        if order == "flipped":
            t = dict(t, tile_enable=t["tile_enable"] ^ 0x100)
            pkg.set_tuning(4, 4, t["wave_cap"], t["tile_enable"], t["grid_mult"], t["concurrent"])
        print(f"[case] table tuning {t}", flush=True)
        for _ in range(2):
            check(pkg, torch_cuda, orc, keys, vals, off, algos=("batched",))
    finally:
        pkg.reset_tuning(4, 4)


@pytest.mark.parametrize("key", ["f32", "u64"])
def test_tiny_segments_reduced(pkg, torch_cuda, orc, key):
    """2^16 segments of length U[1,32]: f32 N(0,1) with +-0.0 and denormals, or u64."""
    keys, vals, off = pkg.gen.tiny_segments(0x5E650A7 + 3, key=key, s=1 << 16)
    check(pkg, torch_cuda, orc, keys, vals, off)


@pytest.mark.parametrize("key", ["f32", "u64"])
def test_tiny_segments_2e20(pkg, torch_cuda, orc, key):
    """2^20 segments of length U[1,32] (about 17.3M items), every form (fix-pass: f32 only)."""
    keys, vals, off = pkg.gen.tiny_segments(0x5E650A7 + 3, key=key)
    check(pkg, torch_cuda, orc, keys, vals, off)


def test_key_types_and_value_widths(pkg, torch_cuda, orc):
    rng = np.random.default_rng(11)
    lengths = np.concatenate([rng.integers(0, 70, 200), [0, 1, 2, 300, 5000, 20000, 70000]])
    off = pkg.gen.offsets_from_lengths(lengths)
    n = int(off[-1])
    raw = pkg.gen.u64(11, 1, n)
    for dt in (np.uint32, np.int32, np.float32, np.uint64, np.int64, np.float64):
        if np.dtype(dt).itemsize == 4:
            keys = (raw >> np.uint64(32)).astype(np.uint32).view(dt)
        else:
            keys = raw.view(dt)
        keys = keys.copy()  # random bit patterns: float keys include NaNs of both signs
        keys[n // 2:] = keys[: n - n // 2]  # duplicates across and within segments
        for vdt in (None, np.uint32, np.uint64):
            vals = None if vdt is None else np.arange(n, dtype=vdt)
            check(pkg, torch_cuda, orc, keys, vals, off)


def test_bit_ranges_descending(pkg, torch_cuda, orc):
    keys, vals, off = pkg.gen.equal_segments(1 << 18, 1 << 6, seed=21)
    lengths = np.diff(off)
    for bb, eb in ((0, 32), (4, 12), (8, 32), (0, 1), (31, 32), (3, 3), (0, 16)):
        for desc in (False, True):
            check(pkg, torch_cuda, orc, keys, vals, off, begin_bit=bb, end_bit=eb, descending=desc)
    assert lengths.sum() == keys.size


def test_bit_ranges_every_tier(pkg, torch_cuda, orc):
    """Bit sub-ranges and descending order on lengths that reach every class
    (wave groups, wave tier, each LDS tile, the MSD tiers), 32- and 64-bit keys."""
    rng = np.random.default_rng(22)
    lengths = np.concatenate([rng.integers(2, 40, 400), rng.integers(40, 600, 100), rng.integers(600, 5000, 20),
                              [1024, 4096, 20000, 70000]])
    rng.shuffle(lengths)
    off = pkg.gen.offsets_from_lengths(lengths)
    n = int(off[-1])
    raw = pkg.gen.u64(22, 1, n)
    vals = np.arange(n, dtype=np.uint32)
    for keys in ((raw >> np.uint64(32)).astype(np.uint32), raw):
        w = keys.dtype.itemsize * 8
        for bb, eb in ((0, w), (4, 12), (8, w), (w - 12, w), (5, w - 7)):
            for desc in (False, True):
                check(pkg, torch_cuda, orc, keys, vals, off, begin_bit=bb, end_bit=eb, descending=desc,
                      algos=("batched", "plain"))


def _float_specials_keys(pkg, dt, n, seed):
    """n keys of dtype dt: N(0,1) values with a few percent of quiet/signalling NaNs of
    both signs and several payloads, +-inf, +-0.0, denormals of both signs, and
    repeated values (ties)."""
    rng = np.random.default_rng(seed)
    x = pkg.gen.normal(seed, 7, n).astype(dt)
    ut = np.uint32 if dt == np.float32 else np.uint64
    bits = x.view(ut).copy()
    if dt == np.float32:
        nan_pats = [0x7FC00000, 0x7F800001, 0x7FFFFFFF, 0x7FA00000, 0xFFC00000, 0xFF800001, 0xFFFFFFFF,
                    0xFFA00000]
        inf, ninf, nz, den = 0x7F800000, 0xFF800000, 0x80000000, 0x00000003
        sign = 0x80000000
    else:
        nan_pats = [0x7FF8000000000000, 0x7FF0000000000001, 0x7FFFFFFFFFFFFFFF, 0xFFF8000000000000,
                    0xFFF0000000000001, 0xFFFFFFFFFFFFFFFF]
        inf, ninf, nz, den = 0x7FF0000000000000, 0xFFF0000000000000, 0x8000000000000000, 3
        sign = 0x8000000000000000
    k = max(8, n // 40)
    idx = rng.integers(0, n, (6, k))
    bits[idx[0]] = np.array(nan_pats, dtype=ut)[rng.integers(0, len(nan_pats), k)]
    bits[idx[1]] = np.array([inf, ninf], dtype=ut)[rng.integers(0, 2, k)]
    bits[idx[2]] = np.array([0, nz], dtype=ut)[rng.integers(0, 2, k)]
    bits[idx[3]] = (rng.integers(1, 1 << 20, k).astype(ut) | np.where(rng.integers(0, 2, k) == 1, ut(sign), ut(0)))
    bits[idx[4]] = ut(den)
    bits[idx[5]] = bits[rng.integers(0, n, k)]  # ties
    return bits.view(dt)


@pytest.mark.parametrize("dt", [np.float32, np.float64], ids=["f32", "f64"])
def test_float_specials_every_tier(pkg, torch_cuda, orc, dt):
    """Float keys with NaNs (-NaN first, +NaN last, by bit pattern), +-inf, -0.0 < +0.0 and
    denormals, ascending and descending, on lengths that reach every class (wave groups,
    the wave tier, each LDS tile, the MSD tiers), pairs and keys only, every form
    (fix-pass: 32-bit keys)."""
    rng = np.random.default_rng(55)
    lengths = np.concatenate([rng.integers(2, 40, 400), rng.integers(40, 600, 100), rng.integers(600, 5000, 20),
                              [0, 1, 1024, 4096, 16384, 20000, 70000]])
    rng.shuffle(lengths)
    off = pkg.gen.offsets_from_lengths(lengths)
    n = int(off[-1])
    keys = _float_specials_keys(pkg, dt, n, 55)
    assert np.isnan(keys).any() and (np.signbit(keys) & np.isnan(keys)).any()
    vals = np.arange(n, dtype=np.uint32)
    for desc in (False, True):
        check(pkg, torch_cuda, orc, keys, vals, off, descending=desc)
        check(pkg, torch_cuda, orc, keys, None, off, descending=desc)


def test_inplace_adapters_keys_and_values(pkg, torch_cuda, orc):
    """The in-place adapters called straight through the C-ABI: bb_segsort's
    int bb_segsort(K*, T*, n, d_segs, S) (bb_segsort.h:36, returns 1) and
    mgpu::segmented_sort(keys, vals, count, segments, S) (kernel_segsort.hxx:405-416):
    begin offsets only, the last segment ends at n; sorted keys AND values equal the
    oracle's stable result, for u32 keys with u32 values, i32 keys with f64 values (the
    bb driver's double values), f32 keys with u64 values, and keys only."""
    import ctypes
    torch = torch_cuda
    lib = pkg.load_library()
    rng = np.random.default_rng(77)
    lengths = np.concatenate([rng.integers(0, 300, 500), [5000, 20000, 1, 0, 70000]])
    off = pkg.gen.offsets_from_lengths(lengths)
    n = int(off[-1])
    raw = pkg.gen.u64(77, 1, n)
    cases = [((raw >> np.uint64(52)).astype(np.uint32), np.arange(n, dtype=np.uint32)),   # 12-bit keys: ties
             ((raw >> np.uint64(44)).astype(np.int32) - (1 << 19), np.arange(n, dtype=np.float64)),
             (_float_specials_keys(pkg, np.float32, n, 78), np.arange(n, dtype=np.uint64)),
             ((raw >> np.uint64(40)).astype(np.uint32), None)]
    segs = np.ascontiguousarray(off[:-1], dtype=np.int32)
    for keys, vals in cases:
        ek, ev = orc.seg_sort(keys, vals, offsets=off)
        kt = KEY_TYPES[keys.dtype]
        vb = 0 if vals is None else vals.dtype.itemsize
        for name in ("bb", "mgpu"):
            print(f"[case] adapter {name} keys={keys.dtype} vals={None if vals is None else vals.dtype}", flush=True)
            dk, ds = dev(torch, keys), dev(torch, segs)
            dv = dev(torch, vals) if vals is not None else None
            vp = ctypes.c_void_p(dv.data_ptr()) if dv is not None else None
            if name == "bb":
                rc = lib.segsort_bb_segsort(dk.data_ptr(), vp, n, ds.data_ptr(), len(segs), kt, vb, None)
                assert rc == 1, pkg.last_error()
            else:
                rc = lib.segsort_mgpu_segmented_sort(dk.data_ptr(), vp, n, ds.data_ptr(), len(segs), kt, vb, None)
                assert rc == 0, pkg.last_error()
            torch.cuda.synchronize()
            assert np.array_equal(host(dk, keys.dtype).view(np.uint8), ek.view(np.uint8)), name
            if vals is not None:
                assert np.array_equal(host(dv, vals.dtype).view(np.uint8), ev.view(np.uint8)), name


def test_mgpu_comparator_and_indices(pkg, torch_cuda, orc):
    """mgpu::segmented_sort with less_t / greater_t (kernel_segsort.hxx:405-416) and
    mgpu::segmented_sort_indices (:418-429) through the C-ABI: keys sorted in place in
    the comparator's direction (equal keys in input order), values carried, and for the
    index form each sorted key's input position; positions before the first segment
    keep their key and get their own index."""
    import ctypes
    torch = torch_cuda
    lib = pkg.load_library()
    rng = np.random.default_rng(79)
    lengths = np.concatenate([[0, 3], rng.integers(0, 400, 300), [9000, 30000, 1]])
    off = pkg.gen.offsets_from_lengths(lengths) + 5  # 5 leading positions outside every segment
    n = int(off[-1])
    keys = (pkg.gen.u64(79, 1, n) >> np.uint64(50)).astype(np.uint32)  # 14-bit keys: ties
    segs = np.ascontiguousarray(off[:-1], dtype=np.int32)
    idx = np.arange(n, dtype=np.int32)
    for comp in (0, 1):
        ek, ev = orc.seg_sort(keys, idx, offsets=off, descending=bool(comp))
        print(f"[case] mgpu indices comp={comp}", flush=True)
        dk, ds = dev(torch, keys), dev(torch, segs)
        di = torch.full((n,), -7, dtype=torch.int32, device="cuda")
        rc = lib.segsort_mgpu_segmented_sort_indices(dk.data_ptr(), di.data_ptr(), n, ds.data_ptr(), len(segs), 0,
                                                     comp, None)
        assert rc == 0, pkg.last_error()
        torch.cuda.synchronize()
        assert np.array_equal(host(dk, np.uint32), ek), comp
        assert np.array_equal(di.cpu().numpy(), ev), comp
        print(f"[case] mgpu comparator comp={comp}", flush=True)
        vals = np.arange(n, dtype=np.uint64) * 3
        ek2, ev2 = orc.seg_sort(keys, vals, offsets=off, descending=bool(comp))
        dk, dv = dev(torch, keys), dev(torch, vals)
        rc = lib.segsort_mgpu_segmented_sort_comp(dk.data_ptr(), dv.data_ptr(), n, ds.data_ptr(), len(segs), 0, 8,
                                                  comp, None)
        assert rc == 0, pkg.last_error()
        torch.cuda.synchronize()
        assert np.array_equal(host(dk, np.uint32), ek2), comp
        assert np.array_equal(host(dv, np.uint64), ev2), comp
    assert lib.segsort_mgpu_segmented_sort_comp(dk.data_ptr(), None, n, ds.data_ptr(), len(segs), 0, 0, 2,
                                                None) != 0


@pytest.mark.parametrize("streams,schedule", [(1, 0), (4, 0), (4, 1), (32, 1)])
def test_multistream_baseline(pkg, torch_cuda, orc, streams, schedule):
    """The per-segment multi-stream baseline (ncub.cu:113-145, newthrust.cu:91-167):
    one device-wide sort per segment over 1..32 streams, round robin or greedy; empty,
    one-item, wave-, tile- and MSD-sized segments, pairs and keys only, ascending and
    descending, against the oracle."""
    torch = torch_cuda
    rng = np.random.default_rng(83 + streams + schedule)
    lengths = np.concatenate([[0, 1, 2, 17], rng.integers(0, 3000, 40), [0, 20000, 70000, 5, 300000]])
    rng.shuffle(lengths)
    off = pkg.gen.offsets_from_lengths(lengths)
    n = int(off[-1])
    cases = [(pkg.gen.u64(83, 1, n).astype(np.uint32), np.arange(n, dtype=np.uint32), False),
             (rng.standard_normal(n).astype(np.float32), None, True),
             (pkg.gen.u64(84, 1, n) >> np.uint64(20), np.arange(n, dtype=np.uint64), False)]
    for keys, vals, desc in cases:
        print(f"[case] multistream streams={streams} schedule={schedule} keys={keys.dtype} desc={desc}", flush=True)
        ek, ev = orc.seg_sort(keys, vals, offsets=off, descending=desc)
        as_torch = {np.dtype(np.uint32): torch.uint32, np.dtype(np.float32): torch.float32,
                    np.dtype(np.uint64): torch.uint64}
        kd = dev(torch, keys).view(as_torch[keys.dtype])
        vd = None if vals is None else dev(torch, vals).view(as_torch[vals.dtype])
        ko, vo = pkg.multistream_sort(kd, off, vd, num_streams=streams, schedule=bool(schedule), descending=desc)
        torch.cuda.synchronize()
        assert np.array_equal(host(ko.view(torch.uint8), keys.dtype), ek)
        if vals is not None:
            assert np.array_equal(host(vo.view(torch.uint8), vals.dtype), ev)


def _skewed_u32(rng, n, share, prefix_bits):
    """n u32 keys of which a `share` fraction have the same top `prefix_bits` bits."""
    keys = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    hit = rng.random(n) < share
    low = np.uint32((1 << (32 - prefix_bits)) - 1)
    keys[hit] = (np.uint32(0x5A5A5A5A) & ~low) | (keys[hit] & low)
    return keys


@pytest.mark.parametrize("case", ["normal_f32_2e24", "top_byte_shared_2e24"])
def test_skewed_single_segment_2e24(pkg, torch_cuda, orc, case):
    """One 2^24-item segment with skewed keys: N(0,1) f32 keys, whose top byte puts
    about 29% of the items in one bucket of the first MSD level (that child then takes
    the second multi-workgroup level), and u32 keys that all share their top byte; batched
    and plain forms against the oracle."""
    n = 1 << 24
    if case == "normal_f32_2e24":
        keys = pkg.gen.normal(0x5E650A7 + 14, 1, n).astype(np.float32)
    else:
        keys = (pkg.gen.keys_u32(0x5E650A7 + 15, 1, n, 24) | np.uint32(0x7F000000)).astype(np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    check(pkg, torch_cuda, orc, keys, vals, np.array([0, n]), algos=("batched", "plain"))


def test_second_msd_level_paths(pkg, torch_cuda, orc):
    """Children of the first multi-workgroup level with more than eight partition tiles
    go through the second level: skewed u32 and u64 keys (most items sharing their top
    bits, some sharing 16 more, so the second level's children still need the
    depth-first tier), keys only and pairs, ascending and descending, bit sub-ranges,
    beside uniform segments in the same call."""
    rng = np.random.default_rng(91)
    segs = [_skewed_u32(rng, 600000, 0.7, 8), _skewed_u32(rng, 400000, 0.9, 16),
            rng.integers(0, 1 << 32, 300000, dtype=np.uint64).astype(np.uint32), _skewed_u32(rng, 3000, 0.9, 8)]
    lengths = np.array([x.size for x in segs])
    keys = np.concatenate(segs)
    off = pkg.gen.offsets_from_lengths(lengths)
    n = keys.size
    vals = np.arange(n, dtype=np.uint32)
    for desc in (False, True):
        check(pkg, torch_cuda, orc, keys, vals, off, descending=desc, algos=("batched",))
    check(pkg, torch_cuda, orc, keys, None, off, algos=("batched",))
    check(pkg, torch_cuda, orc, keys, vals, off, begin_bit=4, end_bit=28, algos=("batched",))
    k64 = (keys.astype(np.uint64) << np.uint64(32)) | rng.integers(0, 1 << 32, n, dtype=np.uint64)
    check(pkg, torch_cuda, orc, k64, vals, off, algos=("batched",))


def test_depth_first_digit_positions(pkg, torch_cuda, orc):
    """Segments of 5 * 2^20 items, so that every child of the first MSD level is larger
    than a tile and goes to the depth-first tier, whose histogram counts the range's top
    digit position and takes the keys' OR/AND: the byte below the top varies (counted
    once), is constant (counted again one position lower), or the two bytes below it are
    (the lowest position, written straight to the output); children whose keys are all
    equal (copied); the same with 64-bit keys; ascending and descending, with values."""
    rng = np.random.default_rng(123)
    n = 5 << 20
    top = rng.integers(0, 256, n, dtype=np.uint32) << np.uint32(24)
    r24 = rng.integers(0, 1 << 24, n, dtype=np.uint32)
    segs = [top | r24,                                                      # top digit varies
            top | np.uint32(0x5A << 16) | (r24 & np.uint32(0xFFFF)),        # byte below the top constant
            top | np.uint32(0x5A5A << 8) | (r24 & np.uint32(0xFF)),         # two bytes below the top constant
            np.where(top < np.uint32(128 << 24), top | r24, top | np.uint32(0x5A5A5A))]  # half the children all equal
    keys = np.concatenate(segs).astype(np.uint32)
    off = np.arange(0, keys.size + 1, n, dtype=np.int64)
    vals = np.arange(keys.size, dtype=np.uint32)
    for desc in (False, True):
        check(pkg, torch_cuda, orc, keys, vals, off, descending=desc, algos=("batched",))
    k64 = (segs[1].astype(np.uint64) << np.uint64(32)) | rng.integers(0, 1 << 32, n, dtype=np.uint64)
    check(pkg, torch_cuda, orc, k64, np.arange(n, dtype=np.uint32), np.array([0, n]), algos=("batched",))


def test_msd_chunk_spans_cross_ranges(pkg, torch_cuda, orc):
    """The multi-workgroup MSD passes give each workgroup a contiguous span of the
    level's chunks (the chunk count over the grid, rounded up). Segments of 2, 4, 9, 2
    and 5 chunks of 16K items (150 segments, 660 chunks) make spans of three chunks that
    start and end inside ranges and cover whole short ones: keys and values against the
    oracle, both directions, keys only, and 64-bit keys."""
    rng = np.random.default_rng(77)
    lengths = np.array([16385, 50000, 131172, 20000, 70001] * 30)
    rng.shuffle(lengths)
    off = pkg.gen.offsets_from_lengths(lengths)
    n = int(off[-1])
    keys = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    for desc in (False, True):
        check(pkg, torch_cuda, orc, keys, vals, off, descending=desc, algos=("batched",))
    check(pkg, torch_cuda, orc, keys, None, off, algos=("batched",))
    k64 = rng.integers(0, 1 << 63, n, dtype=np.uint64) * np.uint64(2) + (keys & np.uint32(1)).astype(np.uint64)
    check(pkg, torch_cuda, orc, k64, vals, off, algos=("batched",))


def _runs_segment(rng, n, run_len, head_uniform):
    """n u32 keys whose top 16 bits come in runs of about run_len items (low 16 bits
    random), optionally preceded by 64 keys with uniformly random top bits."""
    tops = rng.integers(0, 1 << 16, max(1, n // max(run_len, 1)) + 1, dtype=np.uint32)
    keys = (np.repeat(tops, max(run_len, 1))[:n] << np.uint32(16)) | rng.integers(0, 1 << 16, n, dtype=np.uint32)
    rng.shuffle(keys)
    if head_uniform:
        keys[:64] = rng.integers(0, 1 << 32, 64, dtype=np.uint64).astype(np.uint32)
    return keys.astype(np.uint32)


def test_top16_finish_paths(pkg, torch_cuda, orc):
    """Tiles of more than 1024 items whose keys take each path of the top-16 finish:
    uniform keys (items alone in their run, and +-3 windows), runs of 4..16 equal top
    bits (bounded walks), runs longer than 16 behind a uniform-looking first 64 keys
    (the finish starts and falls back to the full LSD from the original keys), and
    skewed keys from the start (the sample keeps the full LSD); with ties in the low
    bits, for every tile size and descending order."""
    rng = np.random.default_rng(41)
    segs = []
    for n in (1100, 2048, 4096, 6000, 8192, 12000, 16384):
        for run_len, head in ((1, False), (3, False), (9, False), (15, True), (40, True), (300, False)):
            segs.append(_runs_segment(rng, n, run_len, head))
    ties = rng.integers(0, 1 << 16, 3000, dtype=np.uint32) << np.uint32(16)
    ties[::7] = ties[0]  # equal full keys inside a run
    segs.append(ties)
    lengths = np.array([s.size for s in segs])
    keys = np.concatenate(segs)
    off = pkg.gen.offsets_from_lengths(lengths)
    vals = np.arange(keys.size, dtype=np.uint32)
    for desc in (False, True):
        check(pkg, torch_cuda, orc, keys, vals, off, descending=desc, algos=("batched", "plain"))
    check(pkg, torch_cuda, orc, keys, None, off, algos=("batched",))  # keys only


def test_large_input_properties(pkg, torch_cuda):
    """2^26 u32 pairs (one 2^25-item segment, then 2^15 ragged segments) checked through
    size-independent properties instead of the oracle: every segment is sorted, its values
    (the input positions) are a permutation of the segment's positions that carry their
    own keys, and equal keys keep their input order."""
    n = 1 << 26
    lengths = np.concatenate([[1 << 25], pkg.gen.ragged_lognormal_lengths(61, s=1 << 15, total=1 << 25)])
    off = pkg.gen.offsets_from_lengths(lengths)
    keys = pkg.gen.keys_u32(61, 1, n, 30)  # 30-bit keys: ties inside the long segment
    vals = np.arange(n, dtype=np.uint32)
    seg = np.repeat(np.arange(lengths.size), lengths)
    for algo in ("batched", "plain"):
        print(f"[case] n={n} algo={algo} large-input properties", flush=True)
        ko, vo = gpu_sort(pkg, torch_cuda, keys, vals, off, algo=algo)
        same_seg = seg[1:] == seg[:-1]
        assert np.all(ko[1:][same_seg] >= ko[:-1][same_seg]), algo
        assert np.array_equal(np.bincount(vo, minlength=n), np.ones(n, dtype=np.int64)), algo
        assert np.array_equal(seg[vo], seg), algo
        assert np.array_equal(keys[vo], ko), algo
        ties = same_seg & (ko[1:] == ko[:-1])
        assert np.all(vo[1:][ties] > vo[:-1][ties]), algo


def test_gaps_unordered_and_empty(pkg, torch_cuda, orc):
    """Non-contiguous, unordered segments with gaps and end <= begin: positions outside
    the segments are never written."""
    n = 300000
    keys = pkg.gen.keys_u32(5, 1, n)
    vals = np.arange(n, dtype=np.uint32)
    begin = np.array([250000, 10, 100, 100, 5000, 90000, 200, 299990, 7], dtype=np.int32)
    end = np.array([299000, 90, 100, 50, 85000, 91000, 201, 300000, 9], dtype=np.int32)
    check(pkg, torch_cuda, orc, keys, vals, begin=begin, end=end, algos=("batched", "plain"))


def test_duplicate_heavy_tiles(pkg, torch_cuda, orc):
    rng = np.random.default_rng(4)
    lengths = rng.integers(1, 40000, 300)
    off = pkg.gen.offsets_from_lengths(lengths)
    n = int(off[-1])
    for bits in (1, 4, 8, 20):
        keys = pkg.gen.keys_u32(bits, 2, n, bits)
        check(pkg, torch_cuda, orc, keys, np.arange(n, dtype=np.uint32), off)
    keys = np.full(n, 7, dtype=np.uint32)
    check(pkg, torch_cuda, orc, keys, np.arange(n, dtype=np.uint32), off)


def test_edge_sizes(pkg, torch_cuda, orc):
    for lengths in ([0], [1], [2], [64], [65], [256], [257], [1024], [1025], [4096], [4097], [16384], [16385],
                    [8192], [8193], [0, 0, 0], [1, 0, 1], [33, 31, 32]):
        off = pkg.gen.offsets_from_lengths(np.array(lengths))
        n = int(off[-1])
        keys = pkg.gen.keys_u32(len(lengths) + n, 1, n)
        check(pkg, torch_cuda, orc, keys, np.arange(n, dtype=np.uint32), off)


def test_fixpass_pass_counts(pkg, torch_cuda, orc):
    """The fix-pass runs as many digit passes as the key max and the segment count
    need, and with the whole 32-bit key as the order key its last pass writes the
    output: no pass at all (one segment, every key 0: the input is copied), a
    single pass (small keys, one segment), only segment-id passes (all keys 0,
    many segments), a leading gap before the first segment, keys only and pairs,
    ascending and descending, and the index-and-gather route (bit sub-range,
    64-bit values) on the same inputs."""
    rng = np.random.default_rng(41)
    cases = []
    for n in (1, 5000, 70000):
        cases.append((np.zeros(n, np.uint32), np.array([0, n])))
        cases.append((rng.integers(0, 200, n, dtype=np.uint32), np.array([0, n])))
    lengths = rng.integers(0, 300, 900)
    off = pkg.gen.offsets_from_lengths(lengths)
    n = int(off[-1])
    cases.append((np.zeros(n, np.uint32), off))
    cases.append((rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32), off + 7))  # gap of 7 in front
    for keys, offs in cases:
        total = int(offs[-1])
        if keys.size < total:
            keys = np.concatenate([rng.integers(0, 1 << 32, total - keys.size, dtype=np.uint64).astype(np.uint32),
                                   keys])
        for vals in (None, np.arange(total, dtype=np.uint32), np.arange(total, dtype=np.uint64)):
            for desc in (False, True):
                check(pkg, torch_cuda, orc, keys, vals, offs, algos=("fixpass",), descending=desc)
        check(pkg, torch_cuda, orc, keys, np.arange(total, dtype=np.uint32), offs, algos=("fixpass",),
              begin_bit=0, end_bit=31)


TUNINGS = [
    {"wave_cap": 256, "tile_enable": 0x1F, "grid_mult": 1},  # groups of 4 only
    {"wave_cap": 256, "tile_enable": 0x2F, "grid_mult": 1},  # groups of 8
    {"wave_cap": 256, "tile_enable": 0x4F, "grid_mult": 2},  # groups of 16
    {"wave_cap": 512, "tile_enable": 0x8F, "grid_mult": 1},  # groups of 32
    {"wave_cap": 2, "tile_enable": 0xFF, "grid_mult": 1},    # every group class, no wave tier
    {"wave_cap": 64, "tile_enable": 0xF0, "grid_mult": 1},   # no LDS tiles: everything above 64 is MSD
    {"wave_cap": 2, "tile_enable": 0x00, "grid_mult": 1},    # nothing but the MSD tiers above 2 items
    {"wave_cap": 256, "tile_enable": 0x1EF, "grid_mult": 1},  # 16-item tiles longest first (length buckets)
    {"wave_cap": 128, "tile_enable": 0x10B, "grid_mult": 3},  # the same with 256x4, 256x16 and 1024x16 only
    {"wave_cap": 128, "tile_enable": 0x207, "grid_mult": 2},  # 256x16 without its half-depth build
    {"wave_cap": 256, "tile_enable": 0x30F, "grid_mult": 1},  # the same, longest first
    {"wave_cap": 512, "tile_enable": 0x40F, "grid_mult": 1},  # wave tier split: <= 256 and 257..512 items
    {"wave_cap": 512, "tile_enable": 0x583, "grid_mult": 2},  # the same with groups of 32, longest first
    {"wave_cap": 512, "tile_enable": 0x4FF, "grid_mult": 1},  # every class: too many to split, one wave class
]


@pytest.mark.parametrize("concurrent", [1, 2, 0], ids=["overlap", "overlap-large-tiles", "one-stream"])
@pytest.mark.parametrize("tuning", TUNINGS, ids=lambda t: f"wc{t['wave_cap']}-te{t['tile_enable']:03x}")
def test_tuning_variants(pkg, torch_cuda, orc, tuning, concurrent):
    """Every class layout the tuning can select sorts bit-exactly (batched and plain),
    for 32- and 64-bit keys, on lengths that hit every group, wave, tile and MSD path,
    with the binned ranges sorted on a second stream during the MSD tier or not."""
    rng = np.random.default_rng(31)
    lengths = np.concatenate([rng.integers(0, 40, 3000), rng.integers(40, 600, 200), rng.integers(600, 20000, 20),
                              [0, 1, 2, 3, 4, 5, 8, 9, 16, 17, 32, 33, 64, 65, 70000]])
    rng.shuffle(lengths)
    off = pkg.gen.offsets_from_lengths(lengths)
    n = int(off[-1])
    raw = pkg.gen.u64(31, 1, n)
    befores = ((4, 4), (8, 4), (4, 0))
    try:
        for kw, vb in befores:
            pkg.set_tuning(kw, vb, tuning["wave_cap"], tuning["tile_enable"], tuning["grid_mult"], concurrent)
        k32 = (raw >> np.uint64(40)).astype(np.uint32)  # 24-bit keys: ties inside segments
        check(pkg, torch_cuda, orc, k32, np.arange(n, dtype=np.uint32), off, algos=("batched", "plain"))
        check(pkg, torch_cuda, orc, k32.view(np.float32), None, off, algos=("batched", "plain"))
        check(pkg, torch_cuda, orc, raw, np.arange(n, dtype=np.uint32), off, algos=("batched", "plain"))
    finally:
        for kw, vb in befores:
            pkg.reset_tuning(kw, vb)


def test_length_bound_equal_segments(pkg, torch_cuda, orc):
    """segsort_sort_hinted with a true bound at most the largest tile: the batched form
    launches no MSD kernels (2^24 pairs in 2^16 segments of 256, and 2^20 keys in 1024
    segments of 1024, keys only)."""
    keys, vals, off = pkg.gen.equal_segments(1 << 24, 1 << 16, seed=0x5E650A7 + 1)
    check(pkg, torch_cuda, orc, keys, vals, off, algos=("batched", "plain"), max_segment_length=256)
    keys, _, off = pkg.gen.equal_segments(1 << 20, 1024, seed=0x5E650A7 + 0, key_bits=31, pairs=False)
    check(pkg, torch_cuda, orc, keys, None, off, algos=("batched",), max_segment_length=1024)


def test_length_bound_ragged_true_bound(pkg, torch_cuda, orc):
    """The ragged workload at 2^24 with its longest segment as the bound (above the
    largest tile: the MSD tier runs, the tuning is the bounded entry's)."""
    keys, vals, off = pkg.gen.ragged_lognormal(0x5E650A7 + 2, dup_heavy=False)
    mx = int(np.diff(off).max())
    check(pkg, torch_cuda, orc, keys, vals, off, algos=("batched",), max_segment_length=mx)


@pytest.mark.parametrize("case", ["u32-pairs", "f32-desc-pairs", "u64-keys", "gaps"])
def test_length_bound_broken(pkg, torch_cuda, orc, case):
    """A wrong bound (segments of up to ~40K items under a bound of 1000) is still
    sorted bit-exactly: the segments above the largest tile go to the one-workgroup
    fallback LSD instead of the MSD tier."""
    rng = np.random.default_rng(77)
    lengths = np.concatenate([rng.integers(0, 900, 3000), [40000, 17000, 16385, 70000, 0, 1],
                              rng.integers(0, 300, 500)]).astype(np.int64)
    rng.shuffle(lengths)
    off = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int32)
    n = int(off[-1])
    vals = np.arange(n, dtype=np.uint32)
    if case == "u32-pairs":
        keys = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
        check(pkg, torch_cuda, orc, keys, vals, off, algos=("batched",), max_segment_length=1000)
    elif case == "f32-desc-pairs":
        keys = rng.standard_normal(n).astype(np.float32)
        check(pkg, torch_cuda, orc, keys, vals, off, algos=("batched",), max_segment_length=1000, descending=True)
    elif case == "u64-keys":
        keys = rng.integers(0, 1 << 63, n, dtype=np.uint64)
        check(pkg, torch_cuda, orc, keys, None, off, algos=("batched",), max_segment_length=1000)
    else:
        # gaps between segments and unordered segments: begin/end arrays
        begin = off[:-1].copy()
        end = off[1:].copy()
        end = np.maximum(begin, end - 3).astype(np.int32)
        perm = rng.permutation(begin.size)
        keys = rng.integers(0, 256, n, dtype=np.uint32)
        check(pkg, torch_cuda, orc, keys, vals, begin=begin[perm], end=end[perm], algos=("batched",),
              max_segment_length=1000)
