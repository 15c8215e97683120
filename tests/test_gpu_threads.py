"""Host-thread safety of the C-ABI (the contract cub callers rely on: reentrant for
distinct temp buffers): several host threads, each with its own stream, temp buffer
and input, call the batched form at once, repeatedly, with the second-stream overlap
on (each thread then also has its own side stream). Every result is checked
bit-exactly against the oracle."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

THREADS = 4
REPS = 3


def test_concurrent_host_threads(pkg, torch_cuda, orc):
    torch = torch_cuda
    cases = []
    for t in range(THREADS):
        lengths = pkg.gen.ragged_lognormal_lengths(70 + t, s=1 << 12, total=1 << 20)
        off = pkg.gen.offsets_from_lengths(lengths)
        n = int(off[-1])
        keys = pkg.gen.keys_u32(70 + t, 1, n)
        vals = np.arange(n, dtype=np.uint32)
        ek, ev = orc.seg_sort(keys, vals, offsets=off)
        cases.append((keys, vals, off, ek, ev))
    errors = []

    D = pkg.DeviceSegmentedRadixSort

    def dev(a):
        return torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).copy()).cuda()

    def worker(t):
        try:
            keys, vals, off, ek, ev = cases[t]
            n, s = keys.size, off.size - 1
            st = torch.cuda.Stream()
            with torch.cuda.stream(st):
                kin, vin = dev(keys), dev(vals)
                kout, vout = torch.empty_like(kin), torch.empty_like(vin)
                od = dev(off.astype(np.int32))
                tb = [0]
                rc = D._call(None, tb, None, None, None, None, n, s, None, None, 0, 32, None, 0, 0, 4, "batched")
                assert rc == 0, pkg.last_error()
                temp = torch.empty(tb[0], dtype=torch.uint8, device="cuda")
                for _ in range(REPS):
                    rc = D._call(temp, [temp.numel()], kin, kout, vin, vout, n, s, od.data_ptr(), od.data_ptr() + 4,
                                 0, 32, st.cuda_stream, 0, 0, 4, "batched")
                    assert rc == 0, pkg.last_error()
                    st.synchronize()
                    if not (np.array_equal(kout.cpu().numpy().view(np.uint32), ek)
                            and np.array_equal(vout.cpu().numpy().view(np.uint32), ev)):
                        errors.append(f"thread {t}: result differs from the oracle")
        except Exception as e:  # reported below, on the test's thread
            errors.append(f"thread {t}: {e!r}")

    try:
        pkg.set_tuning(4, 4, 256, 0xEB, 3, 1)  # groups, wave tier, 256x4, 256x16, 1024x16; overlap on
        ths = [threading.Thread(target=worker, args=(t,)) for t in range(THREADS)]
        for th in ths:
            th.start()
        for th in ths:
            th.join(timeout=120)
        assert not any(th.is_alive() for th in ths), "a worker thread did not finish"
    finally:
        pkg.reset_tuning(4, 4)
    assert not errors, errors
