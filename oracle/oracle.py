"""ctypes binding of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, as the
checker and the CPU baseline; the product never imports it. Functions restate the
reference's sequential path (see segsort_oracle.cpp for the file:line map).
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

KEY_KINDS = {np.dtype(np.uint32): 0, np.dtype(np.int32): 1, np.dtype(np.float32): 2,
             np.dtype(np.uint64): 3, np.dtype(np.int64): 4, np.dtype(np.float64): 5}

_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = ctypes.CDLL(str(LIB))
        vp, i32, i64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64
        L.oracle_segmented_sort.argtypes = [i32, i32, vp, vp, vp, vp, i64, i32, vp, vp, i32, i32, i32, i32]
        L.oracle_segmented_sort.restype = i32
        L.oracle_fixpass_sort.argtypes = [i32, i32, vp, vp, vp, vp, i32, vp]
        L.oracle_fixpass_sort.restype = i32
        L.oracle_order_key.argtypes = [ctypes.c_uint64, i32, i32, i32, i32]
        L.oracle_order_key.restype = ctypes.c_uint64
        L.oracle_cub_temp_bytes.argtypes = [i64, i32, i32]
        L.oracle_cub_temp_bytes.restype = i64
        L.oracle_perm_scatter_f32.argtypes = [vp, vp, vp, i32, vp]
        L.oracle_perm_scatter_f32.restype = None
        L.oracle_time_segmented_sort.argtypes = [i32, i32, vp, vp, vp, vp, i64, i32, vp, vp, i32]
        L.oracle_time_segmented_sort.restype = ctypes.c_double
        L.oracle_max_threads.argtypes = []
        L.oracle_max_threads.restype = i32
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def seg_sort(keys, vals=None, offsets=None, begin=None, end=None, begin_bit=0, end_bit=None, descending=False,
             threads=0, out_keys=None, out_vals=None):
    """Stable per-segment sort (cub SortPairs/SortKeys semantics). Returns (keys, vals).

    Positions outside every segment keep the values of out_keys/out_vals (default:
    a copy of the input)."""
    keys = np.ascontiguousarray(keys)
    kind = KEY_KINDS[keys.dtype]
    if begin is None:
        offsets = np.ascontiguousarray(offsets, dtype=np.int32)
        begin, end = offsets[:-1].copy(), offsets[1:].copy()
    begin = np.ascontiguousarray(begin, dtype=np.int32)
    end = np.ascontiguousarray(end, dtype=np.int32)
    if end_bit is None:
        end_bit = keys.dtype.itemsize * 8
    ko = keys.copy() if out_keys is None else out_keys
    vb = 0
    vo = None
    if vals is not None:
        vals = np.ascontiguousarray(vals)
        vb = vals.dtype.itemsize
        vo = vals.copy() if out_vals is None else out_vals
    lib().oracle_segmented_sort(kind, vb, _p(keys), _p(vals), _p(ko), _p(vo), keys.size, len(begin), _p(begin),
                                _p(end), int(begin_bit), int(end_bit), int(bool(descending)), int(threads))
    return ko, vo


def fixpass_sort(keys, vals, offsets):
    keys = np.ascontiguousarray(keys)
    offsets = np.ascontiguousarray(offsets, dtype=np.int32)
    ko = keys.copy()
    vo = None if vals is None else vals.copy()
    vb = 0 if vals is None else vals.dtype.itemsize
    rc = lib().oracle_fixpass_sort(KEY_KINDS[keys.dtype], vb, _p(keys), _p(vals), _p(ko), _p(vo), len(offsets) - 1,
                                   _p(offsets))
    if rc < 0:
        raise ValueError("fix-pass oracle: 32-bit keys only")
    return ko, vo


def order_key(raw: int, dtype, begin_bit=0, end_bit=None, descending=False) -> int:
    dt = np.dtype(dtype)
    if end_bit is None:
        end_bit = dt.itemsize * 8
    return int(lib().oracle_order_key(int(raw), KEY_KINDS[dt], begin_bit, end_bit, int(descending)))


def cub_temp_bytes(num_items, key_bytes, value_bytes):
    return int(lib().oracle_cub_temp_bytes(num_items, key_bytes, value_bytes))


def perm_scatter(gout, perm, offsets):
    gout = np.ascontiguousarray(gout, dtype=np.float32)
    perm = np.ascontiguousarray(perm, dtype=np.int32)
    offsets = np.ascontiguousarray(offsets, dtype=np.int32)
    gin = np.zeros_like(gout)
    lib().oracle_perm_scatter_f32(_p(gout), _p(perm), _p(offsets), len(offsets) - 1, _p(gin))
    return gin


def time_seg_sort(keys, vals, offsets, threads):
    """Seconds (steady clock) for one oracle segmented sort at `threads` OpenMP threads."""
    keys = np.ascontiguousarray(keys)
    offsets = np.ascontiguousarray(offsets, dtype=np.int32)
    begin, end = offsets[:-1].copy(), offsets[1:].copy()
    ko = keys.copy()
    vo = None if vals is None else vals.copy()
    vb = 0 if vals is None else vals.dtype.itemsize
    return float(lib().oracle_time_segmented_sort(KEY_KINDS[keys.dtype], vb, _p(keys), _p(vals), _p(ko), _p(vo),
                                                  keys.size, len(begin), _p(begin), _p(end), int(threads)))


def max_threads() -> int:
    return int(lib().oracle_max_threads())
