"""The batched form under graphs: a caller may capture the call into its own graph
and replay it (with other work on the device between replays), and the library's
own graph-replay mode gives the same bytes as direct launches. Every replay is
checked against the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup(pkg, torch):
    keys, vals, off = pkg.gen.ragged_lognormal(0x5E650A7 + 9, s=1 << 12, total=1 << 20)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).copy()).cuda()
    kin, vin, od = dev(keys), dev(vals), dev(off.astype(np.int32))
    kout, vout = torch.empty_like(kin), torch.empty_like(vin)
    tb = [0]
    D = pkg.DeviceSegmentedRadixSort
    n, s = keys.size, off.size - 1
    assert D._call(None, tb, None, None, None, None, n, s, None, None, 0, 32, None, 0, 0, 4, "batched") == 0
    temp = torch.empty(tb[0], dtype=torch.uint8, device="cuda")

    def call(stream):
        rc = D._call(temp, [temp.numel()], kin, kout, vin, vout, n, s, od.data_ptr(), od.data_ptr() + 4, 0, 32,
                     stream, 0, 0, 4, "batched")
        assert rc == 0, pkg.last_error()

    return keys, vals, off, kout, vout, call


@pytest.mark.parametrize("concurrent", [0, 1], ids=["one-stream", "overlap"])
def test_caller_captures_and_replays(pkg, torch_cuda, orc, concurrent):
    """With `concurrent` on, the direct call forks a second stream (its result is checked
    too), and so does the captured call: the second stream joins the capture and the
    graph holds both branches."""
    torch = torch_cuda
    keys, vals, off, kout, vout, call = _setup(pkg, torch)
    ek, ev = orc.seg_sort(keys, vals, offsets=off)
    before = pkg.get_tuning(4, 4)
    pkg.set_tuning(4, 4, before["wave_cap"], before["tile_enable"], before["grid_mult"], concurrent)
    try:
        s = torch.cuda.Stream()
        with torch.cuda.stream(s):
            call(s.cuda_stream)  # warm-up outside the capture
        torch.cuda.synchronize()
        assert np.array_equal(kout.cpu().numpy().view(np.uint32), ek)
        assert np.array_equal(vout.cpu().numpy().view(np.uint32), ev)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            call(torch.cuda.current_stream().cuda_stream)
        for rep in range(3):
            kout.fill_(0)  # eager work on the device between replays
            vout.fill_(0)
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            assert np.array_equal(kout.cpu().numpy().view(np.uint32), ek), rep
            assert np.array_equal(vout.cpu().numpy().view(np.uint32), ev), rep
    finally:
        pkg.set_tuning(4, 4, **before)


@pytest.mark.parametrize("concurrent", [0, 1], ids=["one-stream", "overlap"])
def test_library_graph_replay_matches_direct(pkg, torch_cuda, orc, concurrent):
    torch = torch_cuda
    keys, vals, off, kout, vout, call = _setup(pkg, torch)
    ek, ev = orc.seg_sort(keys, vals, offsets=off)
    tuning = pkg.get_tuning(4, 4)
    pkg.set_tuning(4, 4, tuning["wave_cap"], tuning["tile_enable"], tuning["grid_mult"], concurrent)
    before = pkg.set_graph_replay(True)
    try:
        st = torch.cuda.current_stream().cuda_stream
        for rep in range(4):  # first call captures, the rest replay
            kout.fill_(0)
            vout.fill_(0)
            call(st)
            torch.cuda.synchronize()
            assert np.array_equal(kout.cpu().numpy().view(np.uint32), ek), rep
            assert np.array_equal(vout.cpu().numpy().view(np.uint32), ev), rep
    finally:
        pkg.set_graph_replay(before)
        pkg.set_tuning(4, 4, **tuning)
