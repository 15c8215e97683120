"""segsort_sort_noalloc (static device scratch, no allocation) against segsort_sort
with caller temp, through the C-ABI, bit-exact: two calls in a row, the second with
a smaller input, then a call too large for the scratch array (a clear error, nothing
run)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).copy()).cuda()


def _run_both(pkg, torch, keys, vals, off, descending=0):
    lib = pkg.load_library()
    n, s = keys.size, off.size - 1
    kt = {np.dtype(np.uint32): 0, np.dtype(np.float32): 2, np.dtype(np.uint64): 3}[keys.dtype]
    vb = 0 if vals is None else vals.dtype.itemsize
    eb = keys.dtype.itemsize * 8
    kin, od = _dev(torch, keys), _dev(torch, off.astype(np.int32))
    vin = _dev(torch, vals) if vals is not None else None
    outs = []
    for mode in ("temp", "noalloc"):
        ko = torch.full_like(kin, 0xA5)
        vo = torch.full_like(vin, 0xA5) if vin is not None else None
        vi_p = vin.data_ptr() if vin is not None else None
        vo_p = vo.data_ptr() if vo is not None else None
        if mode == "temp":
            tb = ctypes.c_size_t(0)
            assert lib.segsort_sort(None, ctypes.byref(tb), None, None, None, None, n, s, None, None, 0, eb, kt, vb,
                                    descending, 0, None) == 0
            temp = torch.empty(tb.value, dtype=torch.uint8, device="cuda")
            rc = lib.segsort_sort(temp.data_ptr(), ctypes.byref(tb), kin.data_ptr(), ko.data_ptr(), vi_p, vo_p, n, s,
                                  od.data_ptr(), od.data_ptr() + 4, 0, eb, kt, vb, descending, 0, None)
        else:
            rc = lib.segsort_sort_noalloc(kin.data_ptr(), ko.data_ptr(), vi_p, vo_p, n, s, od.data_ptr(),
                                          od.data_ptr() + 4, 0, eb, kt, vb, descending, None)
        assert rc == 0, pkg.last_error()
        torch.cuda.synchronize()
        outs.append((ko.cpu().numpy(), None if vo is None else vo.cpu().numpy()))
    (k1, v1), (k2, v2) = outs
    assert np.array_equal(k1, k2)
    if vals is not None:
        assert np.array_equal(v1, v2)
    return k2, v2


def test_noalloc_matches_caller_temp(pkg, torch_cuda, orc):
    torch = torch_cuda
    lib = pkg.load_library()
    assert lib.segsort_noalloc_max_bytes() == 64 << 20
    # first call: ragged lengths reaching the multi-workgroup tier (2^21 pairs)
    keys, vals, off = pkg.gen.ragged_lognormal(0x5E650A7 + 12, s=1 << 13, total=1 << 21)
    print("[case] noalloc 2^21 ragged pairs", flush=True)
    k, v = _run_both(pkg, torch, keys, vals, off)
    ek, ev = orc.seg_sort(keys, vals, offsets=off)
    assert np.array_equal(k.view(np.uint32), ek) and np.array_equal(v.view(np.uint32), ev)
    # second call, smaller: 2^16 tiny segments of f32 keys, keys only, descending
    keys2, _, off2 = pkg.gen.tiny_segments(0x5E650A7 + 13, key="f32", s=1 << 12)
    print("[case] noalloc small f32 keys-only descending", flush=True)
    k2, _ = _run_both(pkg, torch, keys2, None, off2, descending=1)
    ek2, _ = orc.seg_sort(keys2, None, offsets=off2, descending=True)
    assert np.array_equal(k2.view(np.float32).view(np.uint32), ek2.view(np.uint32))
    # too large for the scratch array: a clear error, and nothing is launched
    tb = ctypes.c_size_t(0)
    assert lib.segsort_sort(None, ctypes.byref(tb), None, None, None, None, 1 << 24, 1 << 16, None, None, 0, 32, 0,
                            4, 0, 0, None) == 0
    assert tb.value > lib.segsort_noalloc_max_bytes()
    dummy = ctypes.c_void_p(0x1000)  # never dereferenced: the size check comes first
    rc = lib.segsort_sort_noalloc(dummy, dummy, dummy, dummy, 1 << 24, 1 << 16, dummy, dummy, 0, 32, 0, 4, 0, None)
    assert rc == 1 and "scratch" in pkg.last_error()
