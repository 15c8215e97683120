"""Seeded synthetic inputs for the segmented-sort workloads.

The reference seeds its generators with ``srand(time(NULL))`` (utils/gen_equal.cpp:107,
utils/gen_diff.cpp:62), so its inputs cannot be reproduced. Here every array comes
from a counter-based splitmix64 stream, so the host checker and the GPU run see
identical data for a given (seed, stream) pair on any machine.

Workloads (sizes as named):
  equal_segments   n keys in s equal segments (keys-only 2^20 / 1024 segments, or
                   2^24 pairs in 2^k segments)
  ragged_lognormal 2^24 pairs in 2^16 lognormal(5, 2) segments, >=5% empty and
                   >=5% of length 1, keys uniform u32 or uniform in [0, 256)
  tiny_segments    2^20 segments of length U[1, 32], f32 N(0,1) keys with +-0.0 and
                   0.1% denormals, or u64 uniform keys
  diff_rows        4096 rows of length U[16, 8192], f32 N(0,1) without ties
"""
from __future__ import annotations

import numpy as np

_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z: np.ndarray) -> np.ndarray:
    z = z.copy()
    z ^= z >> np.uint64(30)
    z *= _M1
    z ^= z >> np.uint64(27)
    z *= _M2
    z ^= z >> np.uint64(31)
    return z


def u64(seed: int, stream: int, n: int) -> np.ndarray:
    """n uint64 words of the splitmix64 sequence keyed by (seed, stream)."""
    base = _mix(np.array([(seed * 0x1000193 + stream * 0x7F4A7C15 + 1) & 0xFFFFFFFFFFFFFFFF], dtype=np.uint64))[0]
    with np.errstate(over="ignore"):
        ctr = np.arange(1, n + 1, dtype=np.uint64) * _GOLD + base
    return _mix(ctr)


def uniform01(seed: int, stream: int, n: int) -> np.ndarray:
    """float64 in (0, 1]."""
    return ((u64(seed, stream, n) >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / 9007199254740992.0)


def normal(seed: int, stream: int, n: int) -> np.ndarray:
    """float64 N(0, 1) by Box-Muller."""
    u1 = uniform01(seed, stream, n)
    u2 = uniform01(seed, stream + 0x100, n)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def offsets_from_lengths(lengths: np.ndarray) -> np.ndarray:
    off = np.zeros(len(lengths) + 1, dtype=np.int64)
    np.cumsum(lengths, out=off[1:])
    assert off[-1] < 2**31
    return off.astype(np.int32)


def equal_offsets(n: int, s: int) -> np.ndarray:
    assert n % s == 0
    return (np.arange(s + 1, dtype=np.int64) * (n // s)).astype(np.int32)


def keys_u32(seed: int, stream: int, n: int, bits: int = 32) -> np.ndarray:
    return (u64(seed, stream, n) >> np.uint64(64 - bits)).astype(np.uint32)


def equal_segments(n: int, s: int, seed: int, key_bits: int = 32, pairs: bool = True):
    """n u32 keys uniform in [0, 2^key_bits) in s equal segments; values = index."""
    keys = keys_u32(seed, 1, n, key_bits)
    vals = np.arange(n, dtype=np.uint32) if pairs else None
    return keys, vals, equal_offsets(n, s)


def ragged_lognormal_lengths(seed: int, s: int = 1 << 16, total: int = 1 << 24, mu: float = 5.0,
                             sigma: float = 2.0, cap: int = 1 << 20) -> np.ndarray:
    """Segment lengths: lognormal(mu, sigma) clipped to [0, cap], rescaled to sum to
    exactly `total`, with ceil(5%) of the segments forced empty and ceil(5%) forced to 1."""
    n_forced = -(-s * 5 // 100)
    perm = np.argsort(u64(seed, 11, s), kind="stable")
    empty_ids, one_ids, rest = perm[:n_forced], perm[n_forced:2 * n_forced], perm[2 * n_forced:]
    raw = np.exp(mu + sigma * normal(seed, 12, s))[rest]
    raw = np.minimum(raw, cap)
    target = total - n_forced
    lengths = np.zeros(s, dtype=np.int64)
    free = np.ones(len(rest), dtype=bool)
    fixed_sum = 0
    scaled = raw.copy()
    for _ in range(64):  # rescale, pin anything over the cap, repeat
        budget = target - fixed_sum
        scale = budget / scaled[free].sum()
        trial = scaled * scale
        over = free & (trial > cap)
        if not over.any():
            break
        scaled[over] = cap
        free &= ~over
        fixed_sum = int(np.round(scaled[~free].sum()))
    vals = np.where(free, scaled * (target - fixed_sum) / scaled[free].sum(), scaled)
    base = np.floor(vals).astype(np.int64)
    base = np.minimum(base, cap)
    short = target - int(base.sum())
    frac = vals - np.floor(vals)
    order = np.argsort(-frac, kind="stable")
    i = 0
    while short > 0:  # hand out the rounding remainder, largest fraction first
        j = order[i % len(order)]
        if base[j] < cap:
            base[j] += 1
            short -= 1
        i += 1
    lengths[rest] = base
    lengths[one_ids] = 1
    lengths[empty_ids] = 0
    assert int(lengths.sum()) == total and lengths.max() <= cap
    return lengths


def ragged_lognormal(seed: int, dup_heavy: bool = False, s: int = 1 << 16, total: int = 1 << 24):
    """2^24 u32 pairs in 2^16 lognormal segments; keys uniform u32 or uniform in [0, 256)."""
    lengths = ragged_lognormal_lengths(seed, s, total)
    keys = keys_u32(seed, 2, total, 8 if dup_heavy else 32)
    vals = np.arange(total, dtype=np.uint32)
    return keys, vals, offsets_from_lengths(lengths)


def tiny_lengths(seed: int, s: int = 1 << 20, lo: int = 1, hi: int = 32) -> np.ndarray:
    span = hi - lo + 1
    return (lo + (u64(seed, 21, s) % np.uint64(span)).astype(np.int64)).astype(np.int64)


def f32_normal_with_specials(seed: int, n: int) -> np.ndarray:
    """float32 N(0,1) with a 0.1% share of denormals of both signs and explicit +-0.0; no NaN."""
    x = normal(seed, 31, n).astype(np.float32)
    bits = x.view(np.uint32).copy()
    r = u64(seed, 32, n)
    n_den = max(2, n // 1000)
    den_idx = (r[:n_den] % np.uint64(n)).astype(np.int64)
    mant = (u64(seed, 33, n_den) % np.uint64((1 << 23) - 1) + np.uint64(1)).astype(np.uint32)
    sign = ((u64(seed, 34, n_den) & np.uint64(1)).astype(np.uint32)) << np.uint32(31)
    bits[den_idx] = sign | mant
    n_zero = max(2, n // 10000)
    z_idx = (u64(seed, 35, n_zero) % np.uint64(n)).astype(np.int64)
    bits[z_idx[0::2]] = 0x00000000
    bits[z_idx[1::2]] = 0x80000000
    return bits.view(np.float32)


def tiny_segments(seed: int, key: str = "f32", s: int = 1 << 20):
    """2^20 segments of length U[1, 32] with u32 index values; key 'f32' or 'u64'."""
    lengths = tiny_lengths(seed, s)
    off = offsets_from_lengths(lengths)
    n = int(off[-1])
    if key == "f32":
        keys = f32_normal_with_specials(seed, n)
    elif key == "u64":
        keys = u64(seed, 36, n)
    else:
        raise ValueError(key)
    vals = np.arange(n, dtype=np.uint32)
    return keys, vals, off


def diff_rows(seed: int, rows: int = 4096, lo: int = 16, hi: int = 8192):
    """Rows of f32 N(0,1), no two equal values (and not both +0.0 and -0.0) in a row."""
    lengths = tiny_lengths(seed, rows, lo, hi)
    off = offsets_from_lengths(lengths)
    n = int(off[-1])
    x = normal(seed, 41, n).astype(np.float32)
    row = np.repeat(np.arange(rows), lengths)
    redraw = 0
    while True:
        mag = np.abs(x)  # +0.0 / -0.0 count as a tie too
        order = np.lexsort((mag, row))
        a, b = order[:-1], order[1:]
        tie = (row[a] == row[b]) & (mag[a] == mag[b])
        if not tie.any():
            break
        redraw += 1
        idx = b[tie]
        x[idx] = normal(seed, 1000 + redraw, len(idx)).astype(np.float32)
    return x, off
