"""MI355X segmented sort: Python host side over the libsegsort.so C-ABI.

The product is the HIP library ``lib/libsegsort.so`` (sources in ``csrc/``), whose
C-ABI is declared in ``include/segsort.h``. This module only binds it:

* ``DeviceSegmentedRadixSort`` mirrors ``cub::DeviceSegmentedRadixSort``
  (lib/cub/device/device_segmented_radix_sort.cuh:136-855): ``SortPairs``,
  ``SortKeys``, ``SortPairsDescending``, ``SortKeysDescending`` with cub's
  argument meaning and temp-storage protocol, returning a hipError_t code.
* ``segmented_sort`` is the tensor-level convenience call.
* ``SegmentedPairSort`` is the differentiable pair sort (torch.autograd.Function)
  whose backward is the library's gradient-scatter kernel.

There is no CPU fallback: every entry point raises if the library is not built.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

from . import autotune, gen  # noqa: F401  (autotuner, seeded workload generators)

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "lib" / "libsegsort.so"

KEY_U32, KEY_I32, KEY_F32, KEY_U64, KEY_I64, KEY_F64 = range(6)
ALGO_BATCHED, ALGO_PLAIN, ALGO_FIXPASS = 0, 1, 2
ALGOS = {"batched": ALGO_BATCHED, "plain": ALGO_PLAIN, "fixpass": ALGO_FIXPASS}

# Every symbol include/segsort.h declares.
EXPORTS = (
    "segsort_sort", "segsort_sort_pairs_u32", "segsort_sort_keys_u32", "segsort_bb_segsort",
    "segsort_mgpu_segmented_sort", "segsort_row_iota", "segsort_perm_scatter_f32", "segsort_set_tuning",
    "segsort_get_tuning", "segsort_tile_caps", "segsort_last_error", "segsort_version", "segsort_profile_enable",
    "segsort_profile_report", "segsort_set_graph_replay", "segsort_reset_tuning", "segsort_lookup_tuning",
    "segsort_mgpu_segmented_sort_comp", "segsort_mgpu_segmented_sort_indices", "segsort_sort_noalloc",
    "segsort_noalloc_max_bytes", "segsort_multistream_sort", "segsort_sort_hinted", "segsort_lookup_tuning_hinted",
)

_lib = None


class SegsortError(RuntimeError):
    pass


def load_library(path: str | os.PathLike | None = None) -> ctypes.CDLL:
    """Load libsegsort.so (raises if it is not built: there is no fallback path)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise SegsortError(f"HIP library not built: {p} is missing (run __graft_entry__.build())")
    lib = ctypes.CDLL(str(p))
    vp, sz, i32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
    lib.segsort_sort.argtypes = [vp, ctypes.POINTER(sz), vp, vp, vp, vp, i32, i32, vp, vp, i32, i32, i32, i32, i32,
                                 i32, vp]
    lib.segsort_sort.restype = i32
    lib.segsort_sort_hinted.argtypes = [vp, ctypes.POINTER(sz), vp, vp, vp, vp, i32, i32, vp, vp, i32, i32, i32, i32,
                                        i32, i32, i32, vp]
    lib.segsort_sort_hinted.restype = i32
    lib.segsort_sort_noalloc.argtypes = [vp, vp, vp, vp, i32, i32, vp, vp, i32, i32, i32, i32, i32, vp]
    lib.segsort_sort_noalloc.restype = i32
    lib.segsort_noalloc_max_bytes.argtypes = []
    lib.segsort_noalloc_max_bytes.restype = sz
    lib.segsort_sort_pairs_u32.argtypes = [vp, ctypes.POINTER(sz), vp, vp, vp, vp, i32, i32, vp, vp, i32, vp]
    lib.segsort_sort_pairs_u32.restype = i32
    lib.segsort_sort_keys_u32.argtypes = [vp, ctypes.POINTER(sz), vp, vp, i32, i32, vp, vp, i32, vp]
    lib.segsort_sort_keys_u32.restype = i32
    lib.segsort_bb_segsort.argtypes = [vp, vp, i32, vp, i32, i32, i32, vp]
    lib.segsort_bb_segsort.restype = i32
    lib.segsort_mgpu_segmented_sort.argtypes = [vp, vp, i32, vp, i32, i32, i32, vp]
    lib.segsort_mgpu_segmented_sort.restype = i32
    lib.segsort_mgpu_segmented_sort_comp.argtypes = [vp, vp, i32, vp, i32, i32, i32, i32, vp]
    lib.segsort_mgpu_segmented_sort_comp.restype = i32
    lib.segsort_mgpu_segmented_sort_indices.argtypes = [vp, vp, i32, vp, i32, i32, i32, vp]
    lib.segsort_mgpu_segmented_sort_indices.restype = i32
    lib.segsort_multistream_sort.argtypes = [vp, ctypes.POINTER(sz), vp, vp, vp, vp, i32, i32, vp, i32, i32, i32, i32,
                                             i32, i32, i32, vp]
    lib.segsort_multistream_sort.restype = i32
    lib.segsort_row_iota.argtypes = [vp, vp, i32, vp, vp]
    lib.segsort_row_iota.restype = i32
    lib.segsort_perm_scatter_f32.argtypes = [vp, vp, vp, vp, i32, vp, vp]
    lib.segsort_perm_scatter_f32.restype = i32
    lib.segsort_set_tuning.argtypes = [i32, i32, i32, i32, i32, i32]
    lib.segsort_set_tuning.restype = i32
    lib.segsort_get_tuning.argtypes = [i32, i32, ctypes.POINTER(i32), ctypes.POINTER(i32), ctypes.POINTER(i32),
                                       ctypes.POINTER(i32)]
    lib.segsort_get_tuning.restype = i32
    lib.segsort_reset_tuning.argtypes = [i32, i32]
    lib.segsort_reset_tuning.restype = i32
    ip = ctypes.POINTER(i32)
    lib.segsort_lookup_tuning.argtypes = [i32, i32, i32, i32, ip, ip, ip, ip, ip]
    lib.segsort_lookup_tuning.restype = i32
    lib.segsort_lookup_tuning_hinted.argtypes = [i32, i32, i32, i32, i32, ip, ip, ip, ip, ip]
    lib.segsort_lookup_tuning_hinted.restype = i32
    lib.segsort_tile_caps.argtypes = [i32, i32, ctypes.POINTER(i32)]
    lib.segsort_tile_caps.restype = i32
    lib.segsort_last_error.argtypes = []
    lib.segsort_last_error.restype = ctypes.c_char_p
    lib.segsort_version.argtypes = []
    lib.segsort_version.restype = i32
    lib.segsort_set_graph_replay.argtypes = [i32]
    lib.segsort_set_graph_replay.restype = i32
    lib.segsort_profile_enable.argtypes = [i32]
    lib.segsort_profile_enable.restype = i32
    lib.segsort_profile_report.argtypes = [ctypes.c_char_p, sz]
    lib.segsort_profile_report.restype = sz
    if path is None:
        _lib = lib
    return lib


def last_error() -> str:
    return load_library().segsort_last_error().decode()


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise SegsortError(f"{what} failed with hipError {rc}: {last_error()}")


def _ptr(x) -> int | None:
    """Device address of a torch tensor (or an int address / None)."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    return x.data_ptr()


def _stream(stream) -> int | None:
    if stream is None:
        import torch

        return torch.cuda.current_stream().cuda_stream
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


_KEY_OF_DTYPE = None


def key_type_of(dtype) -> int:
    global _KEY_OF_DTYPE
    import torch

    if _KEY_OF_DTYPE is None:
        _KEY_OF_DTYPE = {
            torch.uint32: KEY_U32, torch.int32: KEY_I32, torch.float32: KEY_F32,
            torch.uint64: KEY_U64, torch.int64: KEY_I64, torch.float64: KEY_F64,
        }
    if dtype not in _KEY_OF_DTYPE:
        raise SegsortError(f"unsupported key dtype {dtype}")
    return _KEY_OF_DTYPE[dtype]


class DeviceSegmentedRadixSort:
    """cub::DeviceSegmentedRadixSort over libsegsort.so.

    Arguments keep cub's meaning (device_segmented_radix_sort.cuh:136-177): pass
    ``d_temp_storage=None`` to get the required size back in ``temp_storage_bytes``
    (a one-element list, cub's size_t& out-parameter). Buffers are torch tensors or
    raw device addresses; ``key_type``/``value_bytes`` must be given for raw
    addresses. Returns the hipError_t code (0 = success), as cub returns cudaError_t.
    """

    @staticmethod
    def _call(d_temp_storage, temp_storage_bytes, d_keys_in, d_keys_out, d_values_in, d_values_out, num_items,
              num_segments, d_begin_offsets, d_end_offsets, begin_bit, end_bit, stream, descending, key_type,
              value_bytes, algo, max_segment_length=0):
        lib = load_library()
        if key_type is None:
            key_type = key_type_of(d_keys_in.dtype)
        if value_bytes is None:
            value_bytes = 0 if d_values_in is None else d_values_in.element_size()
        kbits = 64 if key_type >= KEY_U64 else 32
        if end_bit is None:
            end_bit = kbits
        tb = ctypes.c_size_t(int(temp_storage_bytes[0]) if temp_storage_bytes else 0)
        args = (_ptr(d_temp_storage), ctypes.byref(tb), _ptr(d_keys_in), _ptr(d_keys_out), _ptr(d_values_in),
                _ptr(d_values_out), int(num_items), int(num_segments), _ptr(d_begin_offsets), _ptr(d_end_offsets),
                int(begin_bit), int(end_bit), int(key_type), int(value_bytes), int(descending),
                int(ALGOS.get(algo, algo)))
        st = _stream(stream) if d_temp_storage is not None else None
        if max_segment_length:
            rc = lib.segsort_sort_hinted(*args, int(max_segment_length), st)
        else:
            rc = lib.segsort_sort(*args, st)
        if temp_storage_bytes is not None and d_temp_storage is None:
            temp_storage_bytes[0] = tb.value
        return rc

    @staticmethod
    def SortPairs(d_temp_storage, temp_storage_bytes, d_keys_in, d_keys_out, d_values_in, d_values_out, num_items,
                  num_segments, d_begin_offsets, d_end_offsets, begin_bit=0, end_bit=None, stream=None,
                  key_type=None, value_bytes=None, algo="batched", max_segment_length=0):
        return DeviceSegmentedRadixSort._call(d_temp_storage, temp_storage_bytes, d_keys_in, d_keys_out, d_values_in,
                                              d_values_out, num_items, num_segments, d_begin_offsets, d_end_offsets,
                                              begin_bit, end_bit, stream, 0, key_type, value_bytes, algo,
                                              max_segment_length)

    @staticmethod
    def SortPairsDescending(d_temp_storage, temp_storage_bytes, d_keys_in, d_keys_out, d_values_in, d_values_out,
                            num_items, num_segments, d_begin_offsets, d_end_offsets, begin_bit=0, end_bit=None,
                            stream=None, key_type=None, value_bytes=None, algo="batched", max_segment_length=0):
        return DeviceSegmentedRadixSort._call(d_temp_storage, temp_storage_bytes, d_keys_in, d_keys_out, d_values_in,
                                              d_values_out, num_items, num_segments, d_begin_offsets, d_end_offsets,
                                              begin_bit, end_bit, stream, 1, key_type, value_bytes, algo,
                                              max_segment_length)

    @staticmethod
    def SortKeys(d_temp_storage, temp_storage_bytes, d_keys_in, d_keys_out, num_items, num_segments,
                 d_begin_offsets, d_end_offsets, begin_bit=0, end_bit=None, stream=None, key_type=None,
                 algo="batched", max_segment_length=0):
        return DeviceSegmentedRadixSort._call(d_temp_storage, temp_storage_bytes, d_keys_in, d_keys_out, None, None,
                                              num_items, num_segments, d_begin_offsets, d_end_offsets, begin_bit,
                                              end_bit, stream, 0, key_type, 0, algo, max_segment_length)

    @staticmethod
    def SortKeysDescending(d_temp_storage, temp_storage_bytes, d_keys_in, d_keys_out, num_items, num_segments,
                           d_begin_offsets, d_end_offsets, begin_bit=0, end_bit=None, stream=None, key_type=None,
                           algo="batched", max_segment_length=0):
        return DeviceSegmentedRadixSort._call(d_temp_storage, temp_storage_bytes, d_keys_in, d_keys_out, None, None,
                                              num_items, num_segments, d_begin_offsets, d_end_offsets, begin_bit,
                                              end_bit, stream, 1, key_type, 0, algo, max_segment_length)


def temp_bytes(num_items, num_segments, key_dtype, value_bytes=0, algo="batched", begin_bit=0, end_bit=None,
               max_segment_length=0):
    """Temp storage the call would need (host-only query, no GPU work)."""
    tb = [0]
    kt = key_type_of(key_dtype) if not isinstance(key_dtype, int) else key_dtype
    rc = DeviceSegmentedRadixSort._call(None, tb, None, None, None, None, num_items, num_segments, None, None,
                                        begin_bit, end_bit, None, 0, kt, value_bytes, algo, max_segment_length)
    _check(rc, "temp size query")
    return tb[0]


def _check_tensor(t, what, device, dtype=None):
    import torch

    if not isinstance(t, torch.Tensor):
        raise SegsortError(f"{what} must be a torch tensor")
    if t.device.type != "cuda" or t.device != device:
        raise SegsortError(f"{what} must be a CUDA tensor on {device} (got {t.device})")
    if not t.is_contiguous():
        raise SegsortError(f"{what} must be contiguous")
    if dtype is not None and t.dtype != dtype:
        raise SegsortError(f"{what} must be {dtype} (got {t.dtype}): offsets are read as int32 begin/end values")


def segmented_sort(keys, offsets=None, values=None, *, begin=None, end=None, descending=False, begin_bit=0,
                   end_bit=None, algo="batched", keys_out=None, values_out=None, temp=None, stream=None,
                   full=False, max_segment_length=0):
    """Stable per-segment sort of cuda tensors; returns (keys_out, values_out).

    Segments are [offsets[i], offsets[i+1]) or [begin[i], end[i]) (int32, contiguous, on
    the keys' device); positions outside every segment are copied from the input (the
    kernels never write them), unless full=True, which promises that the segments cover
    every position and skips that copy.

    ``stream`` (a torch.cuda.Stream; default the current stream) is the stream the sort
    runs on. The outputs and the temp storage are allocated on it and it first waits for
    the current stream, so the inputs are ready and a later reuse of the temp storage is
    ordered after the sort; the inputs are marked as used on it. The caller synchronises
    with ``stream`` before reading the outputs elsewhere, as with any torch side stream.

    ``max_segment_length`` (0: none) is the caller's bound on the segment lengths
    (segsort_sort_hinted): it selects the tuning measured for such lengths and, when no
    segment can exceed the largest LDS tile, skips the multi-workgroup MSD launches. A
    wrong bound is still sorted correctly, more slowly.
    """
    import torch

    if begin is None:
        if offsets is None:
            raise SegsortError("segmented_sort needs offsets or begin/end")
        _check_tensor(offsets, "offsets", keys.device, torch.int32)
        begin, end = offsets[:-1], offsets[1:]
    _check_tensor(keys, "keys", keys.device)
    _check_tensor(begin, "begin", keys.device, torch.int32)
    _check_tensor(end, "end", keys.device, torch.int32)
    if begin.numel() != end.numel():
        raise SegsortError("begin and end must have the same length")
    if values is not None:
        _check_tensor(values, "values", keys.device)
        if values.numel() != keys.numel():
            raise SegsortError("values must have as many elements as keys")
    for t, what in ((keys_out, "keys_out"), (values_out, "values_out")):
        if t is not None:
            _check_tensor(t, what, keys.device)
    n = keys.numel()
    s = begin.numel()
    kt = key_type_of(keys.dtype)
    vb = 0 if values is None else values.element_size()
    need = temp_bytes(n, s, kt, vb, algo, begin_bit, end_bit, max_segment_length)
    caller = torch.cuda.current_stream(keys.device)
    if stream is None:
        stream = caller
    if isinstance(stream, int):
        raise SegsortError("segmented_sort takes a torch.cuda.Stream (raw handles: DeviceSegmentedRadixSort)")
    side = stream != caller
    if side:
        stream.wait_stream(caller)  # the inputs (and caller-given outputs) are ready on the sort's stream
    with torch.cuda.stream(stream):
        if keys_out is None:
            keys_out = torch.empty_like(keys) if full else keys.clone()
        if values is not None and values_out is None:
            values_out = torch.empty_like(values) if full else values.clone()
        if temp is None or temp.numel() < need:
            temp = torch.empty(need, dtype=torch.uint8, device=keys.device)
        tb = [temp.numel()]
        rc = DeviceSegmentedRadixSort._call(temp, tb, keys, keys_out, values, values_out, n, s, begin, end,
                                            begin_bit, end_bit, stream, 1 if descending else 0, kt, vb, algo,
                                            max_segment_length)
    _check(rc, "segsort_sort")
    if side:
        # tensors made on another stream and used by this sort: their memory is not
        # reused before the sort's work on `stream` is done
        for t in (keys, values, begin, end, keys_out, values_out, temp):
            if t is not None:
                t.record_stream(stream)
    return keys_out, values_out


def multistream_sort(keys, offsets, values=None, *, num_streams=4, schedule=False, descending=False, begin_bit=0,
                     end_bit=None):
    """The per-segment multi-stream baseline (src/ncub.cu:113-145, src/newthrust.cu:91-167):
    one device-wide sort per segment over ``num_streams`` streams, round robin or, with
    ``schedule``, each segment to the stream with the fewest items so far. ``offsets`` is
    a host sequence of S+1 ascending offsets. Runs on the current stream (synchronised
    once, for the upload of the segment lengths); returns (keys_out, values_out), where
    positions outside every segment are copied from the input."""
    import numpy as np
    import torch

    _check_tensor(keys, "keys", keys.device)
    if values is not None:
        _check_tensor(values, "values", keys.device)
        if values.numel() != keys.numel():
            raise SegsortError("values must have as many elements as keys")
    off = np.ascontiguousarray(np.asarray(offsets, dtype=np.int32))
    n, s = keys.numel(), off.size - 1
    if s < 0:
        raise SegsortError("offsets needs at least one entry")
    kt = key_type_of(keys.dtype)
    vb = 0 if values is None else values.element_size()
    eb = keys.element_size() * 8 if end_bit is None else end_bit
    lib = load_library()
    hp = off.ctypes.data if s > 0 else None
    st = torch.cuda.current_stream(keys.device).cuda_stream
    tb = ctypes.c_size_t(0)
    args = (n, s, hp, begin_bit, eb, kt, vb, 1 if descending else 0, num_streams, 1 if schedule else 0, st)
    _check(lib.segsort_multistream_sort(None, ctypes.byref(tb), None, None, None, None, *args),
           "segsort_multistream_sort")
    temp = torch.empty(max(tb.value, 1), dtype=torch.uint8, device=keys.device)
    keys_out = keys.clone()
    values_out = values.clone() if values is not None else None
    _check(lib.segsort_multistream_sort(temp.data_ptr(), ctypes.byref(tb), keys.data_ptr(), keys_out.data_ptr(),
                                        _ptr(values), _ptr(values_out), *args), "segsort_multistream_sort")
    return keys_out, values_out


def set_tuning(key_width, value_bytes, wave_cap, tile_enable, grid_mult, concurrent=0):
    """Install an override used by every call with these key/value widths (instead of
    the library's built-in table) until reset_tuning."""
    _check(load_library().segsort_set_tuning(key_width, value_bytes, wave_cap, tile_enable, grid_mult, concurrent),
           "set_tuning")


def reset_tuning(key_width=0, value_bytes=0):
    """Drop the override of one slot (key_width 0: every slot): calls use the built-in table."""
    _check(load_library().segsort_reset_tuning(key_width, value_bytes), "reset_tuning")


def get_tuning(key_width, value_bytes):
    """The slot's override, or the built-in default when none is installed."""
    a, b, c, d = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _check(load_library().segsort_get_tuning(key_width, value_bytes, ctypes.byref(a), ctypes.byref(b),
                                             ctypes.byref(c), ctypes.byref(d)), "get_tuning")
    return {"wave_cap": a.value, "tile_enable": b.value, "grid_mult": c.value, "concurrent": d.value}


def lookup_tuning(num_items, num_segments, key_width, value_bytes, max_segment_length=0):
    """The tuning a call of this shape (and length bound, 0: none) runs with (the
    override, else the built-in table's entry), and whether it comes from the table."""
    v = [ctypes.c_int() for _ in range(5)]
    _check(load_library().segsort_lookup_tuning_hinted(int(num_items), int(num_segments), int(max_segment_length),
                                                       key_width, value_bytes, *[ctypes.byref(x) for x in v]),
           "lookup_tuning")
    return {"wave_cap": v[0].value, "tile_enable": v[1].value, "grid_mult": v[2].value,
            "concurrent": v[3].value, "from_table": bool(v[4].value)}


class tuning_override:
    """Context manager: calls with these widths use the given tuning inside the block;
    on exit the slot goes back to the built-in table."""

    def __init__(self, key_width, value_bytes, **tuning):
        self.kw, self.vb, self.t = key_width, value_bytes, tuning

    def __enter__(self):
        set_tuning(self.kw, self.vb, **self.t)
        return self

    def __exit__(self, *exc):
        reset_tuning(self.kw, self.vb)
        return False


def set_graph_replay(on: bool) -> bool:
    """Replay the batched form from cached hipGraphs (True) or launch directly
    (False, the default); returns the previous setting."""
    return bool(load_library().segsort_set_graph_replay(1 if on else 0))


def profile(on: bool) -> None:
    """Turn per-launch HIP-event timing on (clearing the log) or off."""
    _check(load_library().segsort_profile_enable(1 if on else 0), "profile_enable")


def profile_report() -> dict:
    """{kernel name: {"launches", "ms", "bytes"}} of the launches logged since enabling."""
    lib = load_library()
    need = lib.segsort_profile_report(None, 0)
    buf = ctypes.create_string_buffer(int(need))
    lib.segsort_profile_report(buf, need)
    out = {}
    for line in buf.value.decode().splitlines():
        name, n, ms, b = line.split("\t")
        out[name] = {"launches": int(n), "ms": float(ms), "bytes": float(b)}
    return out


def tile_caps(key_width, value_bytes):
    arr = (ctypes.c_int * 4)()
    n = load_library().segsort_tile_caps(key_width, value_bytes, arr)
    return [arr[i] for i in range(n)]


def _autograd():
    import torch

    class SegmentedPairSort(torch.autograd.Function):
        """Stable ascending per-row sort of float32 values: (sorted, perm).

        perm is row-local (int32), as torch.sort's indices along a row. The backward
        scatters the upstream gradient through perm with the library's gradient
        kernel: grad_in[begin[r] + perm[i]] = grad_sorted[i].
        """

        @staticmethod
        def forward(ctx, x, offsets, algo="batched", assume_cover=False):
            if x.dtype != torch.float32 or offsets.dtype != torch.int32:
                raise SegsortError("SegmentedPairSort expects float32 values and int32 offsets")
            x = x.contiguous()
            _check_tensor(offsets, "offsets", x.device, torch.int32)
            if offsets.numel() < 1:
                raise SegsortError("offsets needs at least one entry")
            if not assume_cover:
                # the rows must cover x exactly: outputs and gradients are written only
                # inside rows (one device-to-host read of the two end offsets)
                first, last = (int(v) for v in offsets[[0, -1]].tolist())
                if first != 0 or last != x.numel():
                    raise SegsortError(f"offsets must run from 0 to x.numel()={x.numel()} (got {first}..{last})")
            begin, end = offsets[:-1], offsets[1:]
            rows = begin.numel()
            lib = load_library()
            st = _stream(None)
            iota = torch.empty(x.numel(), dtype=torch.int32, device=x.device)
            _check(lib.segsort_row_iota(begin.data_ptr(), end.data_ptr(), rows, iota.data_ptr(), st), "row_iota")
            y, perm = segmented_sort(x, begin=begin, end=end, values=iota, algo=algo, full=True)
            ctx.save_for_backward(perm, offsets)
            ctx.mark_non_differentiable(perm)
            ctx.set_materialize_grads(False)  # no zero-filled gradient for perm (or an unused y)
            return y, perm

        @staticmethod
        def backward(ctx, grad_y, grad_perm):
            if grad_y is None:
                return None, None, None, None
            perm, offsets = ctx.saved_tensors
            grad_y = grad_y.contiguous().float()
            gin = torch.empty_like(grad_y)  # rows cover every position
            begin, end = offsets[:-1], offsets[1:]
            _check(load_library().segsort_perm_scatter_f32(grad_y.data_ptr(), perm.data_ptr(), begin.data_ptr(),
                                                           end.data_ptr(), begin.numel(), gin.data_ptr(),
                                                           _stream(None)), "perm_scatter")
            return gin, None, None, None

    return SegmentedPairSort


_SPS = None


def segmented_pair_sort(x, offsets, algo="batched", assume_cover=False):
    """Differentiable stable per-row sort: returns (sorted values, row-local perm).

    The rows [offsets[r], offsets[r+1]) must cover x exactly (offsets[0] == 0 and
    offsets[-1] == x.numel()); this is checked with one device-to-host read unless the
    caller promises it with assume_cover=True."""
    global _SPS
    if _SPS is None:
        _SPS = _autograd()
    return _SPS.apply(x, offsets, algo, assume_cover)
