"""Pins the CPU oracle before anything is checked against it (CPU only).

Each check is chosen so a plausible oracle mistake fails it: cub's own documented
answers, the reference's committed output dump, an independent numpy ordering, brute
force over all permutations of small multisets, and the float/signed closed forms.
"""
import itertools
import json
from pathlib import Path

import numpy as np
import pytest

GOLD = Path(__file__).resolve().parent / "golden"


def test_cub_doc_vectors(orc):
    """cub's documented known answers (device_segmented_radix_sort.cuh:104-129, 302-327,
    506-526, 685-705; device_radix_sort.cuh:115-136, 296-317); cub's example keys are int."""
    vec = json.loads((GOLD / "cub_doc_vectors.json").read_text())
    for name, v in vec.items():
        keys = np.array(v["keys"], dtype=np.int32)
        vals = np.array(v["values"], dtype=np.int32) if "values" in v else None
        desc = "descending" in name
        ko, vo = orc.seg_sort(keys, vals, offsets=v["offsets"], descending=desc)
        assert ko.tolist() == v["keys_out"], name
        if vals is not None:
            assert vo.tolist() == v["values_out"], name


def test_cub_out_dump(orc):
    """src/cub.out (== src/thrust.out) holds 512 sorted segments of 256 keys: every
    segment is non-decreasing, and sorting a seeded shuffle of it reproduces it."""
    z = np.load(GOLD / "cub_out_512x256.npz")
    keys = z["keys"].astype(np.uint32)
    off = z["offsets"]
    segs = keys.reshape(512, 256)
    assert (np.diff(segs.astype(np.int64), axis=1) >= 0).all()
    rng = np.random.default_rng(7)
    shuffled = np.concatenate([rng.permutation(s) for s in segs])
    ko, _ = orc.seg_sort(shuffled, None, offsets=off)
    assert np.array_equal(ko, keys)


def test_mergesort2_inputs_against_numpy(orc):
    """The reference's committed inputs (src/mergesort2/in1-32, in2-1024) sorted by the
    oracle equal numpy's stable argsort per segment (an independent implementation)."""
    ms = json.loads((GOLD / "mergesort2_inputs.json").read_text())
    for name, v in ms.items():
        keys = np.array(v["keys"], dtype=np.uint32)
        vals = np.arange(keys.size, dtype=np.uint32)
        off = v["offsets"]
        ko, vo = orc.seg_sort(keys, vals, offsets=off)
        for a, b in zip(off[:-1], off[1:]):
            idx = np.argsort(keys[a:b], kind="stable") + a
            assert np.array_equal(vo[a:b], idx.astype(np.uint32)), name
            assert np.array_equal(ko[a:b], keys[idx]), name


def _brute(keys, okey):
    order = sorted(range(len(keys)), key=lambda i: (okey(keys[i]), i))
    return order


def test_exhaustive_small_multisets(orc):
    """All orderings of small multisets with duplicates: the oracle's stable pair sort
    equals sorting by (key, original index)."""
    for ms in ([1, 1, 2, 3], [5, 5, 5, 0], [2, 0, 2, 0, 1], [7, 3, 7, 3, 7, 1]):
        for perm in set(itertools.permutations(ms)):
            keys = np.array(perm, dtype=np.uint32)
            vals = np.arange(len(keys), dtype=np.uint32)
            ko, vo = orc.seg_sort(keys, vals, offsets=[0, len(keys)])
            order = _brute(list(perm), lambda k: k)
            assert vo.tolist() == order
            assert ko.tolist() == [perm[i] for i in order]


def test_float_order_closed_form(orc):
    """Twiddled f32/f64 order: numeric order for non-NaN values with -0.0 before +0.0
    and denormals ordered by magnitude (util_type.cuh:1017-1027)."""
    specials = [0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, 1e-45, -1e-45, 1.2e-38, -1.2e-38, 3.5, -3.5, 2.0**-140,
                -(2.0**-140), np.finfo(np.float32).max, -np.finfo(np.float32).max]
    for dt in (np.float32, np.float64):
        x = np.array(specials, dtype=dt)
        ko, vo = orc.seg_sort(x, np.arange(len(x), dtype=np.uint32), offsets=[0, len(x)])
        ref = np.lexsort((~np.signbit(x), x))  # by value, then negative zero first
        assert np.array_equal(vo, ref.astype(np.uint32)), dt
        # closed forms of the twiddle itself
        bits = x.view(np.uint32 if dt == np.float32 else np.uint64)
        hi = 1 << (8 * x.itemsize - 1)
        allm = (1 << (8 * x.itemsize)) - 1
        for b in bits.tolist():
            want = (b ^ allm) if (b & hi) else (b ^ hi)
            assert orc.order_key(b, dt) == want


def _nan_specials(dt):
    """Quiet and signalling NaNs of both signs with several payloads, as raw bit patterns."""
    if dt == np.float32:
        pats = [0x7FC00000, 0x7F800001, 0x7FFFFFFF, 0x7FA00000, 0x7FC00123]
        sign = 0x80000000
        ut = np.uint32
    else:
        pats = [0x7FF8000000000000, 0x7FF0000000000001, 0x7FFFFFFFFFFFFFFF, 0x7FF4000000000000,
                0x7FF8000000000ABC]
        sign = 0x8000000000000000
        ut = np.uint64
    return np.array(pats + [p | sign for p in pats], dtype=ut)


def test_nan_order_closed_form(orc):
    """NaN keys in the radix order (util_type.cuh:1017-1027): every NaN with the sign bit
    set comes before -inf (the largest negative-NaN bit pattern first), every NaN with the
    sign bit clear after +inf (by increasing bit pattern), each NaN exactly where its
    bit pattern puts it; descending order reverses the whole sequence with ties kept in
    input order. The expected order is built from IEEE classes (numpy's isnan/signbit
    and value order for non-NaN), not from the twiddle."""
    for dt in (np.float32, np.float64):
        ut = np.uint32 if dt == np.float32 else np.uint64
        finite = np.array([0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, 1e-45, -1e-45, 2.5, -2.5], dtype=dt).view(ut)
        nans = _nan_specials(dt)
        bits = np.concatenate([finite, nans, nans[:3]])  # repeated NaNs: ties by input position
        x = bits.view(dt)
        n = len(x)
        neg_nan = [i for i in range(n) if np.isnan(x[i]) and np.signbit(x[i])]
        pos_nan = [i for i in range(n) if np.isnan(x[i]) and not np.signbit(x[i])]
        num = [i for i in range(n) if not np.isnan(x[i])]
        # negative NaNs: larger magnitude payload first; positive NaNs: smaller first
        neg_nan.sort(key=lambda i: (-int(bits[i]), i))
        pos_nan.sort(key=lambda i: (int(bits[i]), i))
        num.sort(key=lambda i: (float(x[i]), not np.signbit(x[i]), i))
        want = neg_nan + num + pos_nan
        _, vo = orc.seg_sort(x, np.arange(n, dtype=np.uint32), offsets=[0, n])
        assert vo.tolist() == want, dt
        # descending: the reverse order of distinct patterns, equal patterns still in input order
        rank = {}
        for r, i in enumerate(want):
            rank.setdefault(int(bits[i]), r)
        want_desc = sorted(range(n), key=lambda i: (-rank[int(bits[i])], i))
        _, vd = orc.seg_sort(x, np.arange(n, dtype=np.uint32), offsets=[0, n], descending=True)
        assert vd.tolist() == want_desc, dt


def test_signed_order_closed_form(orc):
    for dt in (np.int32, np.int64):
        x = np.array([0, -1, 1, np.iinfo(dt).min, np.iinfo(dt).max, -5, 5, 5, -5], dtype=dt)
        ko, vo = orc.seg_sort(x, np.arange(len(x), dtype=np.uint32), offsets=[0, len(x)])
        assert np.array_equal(vo, np.argsort(x, kind="stable").astype(np.uint32))


def test_bit_range_and_descending(orc):
    """[begin_bit, end_bit) and descending against a brute-force definition."""
    rng = np.random.default_rng(3)
    keys = rng.integers(0, 2**32, size=300, dtype=np.uint64).astype(np.uint32)
    keys[150:] = keys[:150]  # duplicates
    vals = np.arange(keys.size, dtype=np.uint32)
    off = [0, 17, 17, 100, 101, 300]
    for bb, eb in ((0, 32), (4, 12), (8, 32), (0, 1), (31, 32), (5, 5)):
        for desc in (False, True):
            ko, vo = orc.seg_sort(keys, vals, offsets=off, begin_bit=bb, end_bit=eb, descending=desc)
            mask = ((1 << (eb - bb)) - 1) if eb > bb else 0
            for a, b in zip(off[:-1], off[1:]):
                def ok(k):
                    v = (int(k) >> bb) & mask
                    return -v if desc else v
                order = sorted(range(a, b), key=lambda i: (ok(keys[i]), i))
                assert vo[a:b].tolist() == order, (bb, eb, desc)


def test_gaps_and_empty_segments_untouched(orc):
    keys = np.arange(20, 0, -1, dtype=np.uint32)
    begin = np.array([2, 10, 15, 9], dtype=np.int32)
    end = np.array([6, 14, 15, 7], dtype=np.int32)  # [15,15) and [9,7) are empty
    sentinel = np.full(20, 0xDEAD, dtype=np.uint32)
    ko, _ = orc.seg_sort(keys, None, begin=begin, end=end, out_keys=sentinel.copy())
    assert ko[2:6].tolist() == sorted(keys[2:6].tolist())
    assert ko[10:14].tolist() == sorted(keys[10:14].tolist())
    untouched = [i for i in range(20) if not (2 <= i < 6 or 10 <= i < 14)]
    assert (ko[untouched] == 0xDEAD).all()


def test_fixpass_oracle_agrees(orc, pkg):
    """The widened CPU fix-pass (fixseq.cu:61-83, 64-bit composite) and the per-segment
    stable sort are independent restatements; they agree on full-range u32 keys."""
    keys, vals, off = pkg.gen.equal_segments(1 << 14, 1 << 6, seed=5)
    k1, v1 = orc.seg_sort(keys, vals, offsets=off)
    k2, v2 = orc.fixpass_sort(keys, vals, off)
    assert np.array_equal(k1, k2) and np.array_equal(v1, v2)
    lengths = pkg.gen.tiny_lengths(9, 500, 0, 70)
    off = pkg.gen.offsets_from_lengths(lengths)
    keys = pkg.gen.keys_u32(9, 3, int(off[-1]), 6)
    vals = np.arange(keys.size, dtype=np.uint32)
    k1, v1 = orc.seg_sort(keys, vals, offsets=off)
    k2, v2 = orc.fixpass_sort(keys, vals, off)
    assert np.array_equal(k1, k2) and np.array_equal(v1, v2)


def test_cub_temp_formula(orc):
    """Temp the cub segmented path asks for (util_device.cuh:66-80): 134,217,983 B at 2^24
    u32 pairs and 4,194,559 B at 2^20 u32 keys."""
    assert orc.cub_temp_bytes(1 << 24, 4, 4) == 134217983
    assert orc.cub_temp_bytes(1 << 20, 4, 0) == 4194559


def test_perm_scatter_oracle(orc):
    off = np.array([0, 3, 3, 7], dtype=np.int32)
    perm = np.array([1, 2, 0, 2, 1, 0, 3], dtype=np.int32)
    g = np.arange(7, dtype=np.float32) + 1
    gin = orc.perm_scatter(g, perm, off)
    assert gin.tolist() == [3, 1, 2, 6, 5, 4, 7]


def test_generators(pkg):
    g = pkg.gen
    lengths = g.ragged_lognormal_lengths(0x5E650A7 + 2)
    assert lengths.sum() == 1 << 24 and lengths.size == 1 << 16 and lengths.max() <= 1 << 20
    assert (lengths == 0).mean() >= 0.05 and (lengths == 1).mean() >= 0.05
    assert (lengths > 16384).sum() > 10  # heavy tail reaches the multi-workgroup tier
    t = g.tiny_lengths(1, 1 << 16)
    assert t.min() == 1 and t.max() == 32
    x = g.f32_normal_with_specials(1, 1 << 16)
    b = x.view(np.uint32)
    assert (b == 0).any() and (b == 0x80000000).any()
    den = ((b & 0x7F800000) == 0) & ((b & 0x7FFFFF) != 0)
    assert den.sum() >= 60 and (den & (b >> 31 == 1)).any() and (den & (b >> 31 == 0)).any()
    assert not np.isnan(x).any()
    a1 = g.keys_u32(3, 1, 1000)
    assert np.array_equal(a1, g.keys_u32(3, 1, 1000)) and not np.array_equal(a1, g.keys_u32(4, 1, 1000))


def test_diff_rows_tie_free(pkg):
    x, off = pkg.gen.diff_rows(5, rows=64, lo=16, hi=2048)
    for a, b in zip(off[:-1], off[1:]):
        r = np.abs(x[a:b])
        assert np.unique(r).size == r.size


def test_planted_oracle_mistake_is_caught(orc):
    """A restatement that drops stability (plain argsort without kind='stable' on equal
    keys) disagrees with the oracle on the cub example, so the checks above can fail."""
    keys = np.array([3, 1, 3, 1, 3, 1, 2, 2] * 8, dtype=np.uint32)
    vals = np.arange(keys.size, dtype=np.uint32)
    _, vo = orc.seg_sort(keys, vals, offsets=[0, keys.size])
    unstable = np.argsort(keys, kind="quicksort")[::-1]  # a deliberately wrong restatement
    assert not np.array_equal(vo, unstable.astype(np.uint32))
