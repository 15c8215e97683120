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
    D = pkg.DeviceSegmentedRadixSort
    n, s = keys.size, off.size - 1
    temp = []  # sized at the first call, under the tuning the test installed (temp size depends on it)

    def call(stream):
        if not temp:
            tb = [0]
            assert D._call(None, tb, None, None, None, None, n, s, None, None, 0, 32, None, 0, 0, 4, "batched") == 0
            temp.append(torch.empty(tb[0], dtype=torch.uint8, device="cuda"))
        rc = D._call(temp[0], [temp[0].numel()], kin, kout, vin, vout, n, s, od.data_ptr(), od.data_ptr() + 4, 0,
                     32, stream, 0, 0, 4, "batched")
        assert rc == 0, pkg.last_error()

    call.temp = temp
    return keys, vals, off, kout, vout, call


@pytest.mark.parametrize("concurrent", [0, 1, 2], ids=["one-stream", "overlap", "overlap-large-tiles"])
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
        pkg.reset_tuning(4, 4)


@pytest.mark.parametrize("concurrent", [0, 1, 2], ids=["one-stream", "overlap", "overlap-large-tiles"])
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
        pkg.reset_tuning(4, 4)


@pytest.mark.parametrize("concurrent", [0, 1, 2], ids=["one-stream", "overlap", "overlap-large-tiles"])
def test_graph_cache_eviction_after_replays(pkg, torch_cuda, orc, concurrent):
    """The library's graph-replay mode over 24 distinct argument sets (more than its
    16 cached graphs), three rounds with no host synchronisation inside a round: later
    captures evict graphs whose last replay was enqueued just before, so eviction has to
    wait for that replay before destroying the graph. Every output of every round is
    checked against the oracle (outputs refilled between rounds)."""
    torch = torch_cuda
    keys, vals, off = pkg.gen.ragged_lognormal(0x5E650A7 + 10, s=1 << 10, total=1 << 18)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).copy()).cuda()
    kin, vin, od = dev(keys), dev(vals), dev(off.astype(np.int32))
    n, s_all = keys.size, off.size - 1
    D = pkg.DeviceSegmentedRadixSort
    sets = 24
    seg_counts = [s_all - 7 * i for i in range(sets)]  # distinct argument sets: segment prefixes
    expect = []
    for s in seg_counts:
        fk = np.full(n, 0xA5A5A5A5, dtype=np.uint32)
        fv = np.full(n, 0xA5A5A5A5, dtype=np.uint32)
        expect.append(orc.seg_sort(keys, vals, offsets=off[:s + 1], out_keys=fk, out_vals=fv))
    outs = [(torch.empty_like(kin), torch.empty_like(vin)) for _ in range(sets)]
    tuning = pkg.get_tuning(4, 4)
    pkg.set_tuning(4, 4, tuning["wave_cap"], tuning["tile_enable"], tuning["grid_mult"], concurrent)
    tb = [0]
    assert D._call(None, tb, None, None, None, None, n, s_all, None, None, 0, 32, None, 0, 0, 4, "batched") == 0
    temps = [torch.empty(tb[0], dtype=torch.uint8, device="cuda") for _ in range(2)]
    before = pkg.set_graph_replay(True)
    try:
        st = torch.cuda.current_stream().cuda_stream
        for rnd in range(3):
            for ko, vo in outs:
                ko.fill_(0xA5)
                vo.fill_(0xA5)
            torch.cuda.synchronize()
            print(f"[case] graph-cache round {rnd} concurrent={concurrent}", flush=True)
            for i, s in enumerate(seg_counts):
                ko, vo = outs[i]
                temp = temps[i % 2]  # one stream: consecutive calls may share temp storage
                rc = D._call(temp, [temp.numel()], kin, ko, vin, vo, n, s, od.data_ptr(), od.data_ptr() + 4, 0, 32,
                             st, 0, 0, 4, "batched")
                assert rc == 0, pkg.last_error()
            torch.cuda.synchronize()
            for i in range(sets):
                ek, ev = expect[i]
                assert np.array_equal(outs[i][0].cpu().numpy().view(np.uint32), ek), (rnd, i)
                assert np.array_equal(outs[i][1].cpu().numpy().view(np.uint32), ev), (rnd, i)
    finally:
        pkg.set_graph_replay(before)
        pkg.reset_tuning(4, 4)


def test_eager_call_while_same_thread_captures(pkg, torch_cuda, orc):
    """With the overlap on, a call captured into the caller's graph forks a second stream
    that joins the capture. An uncaptured call on another stream from the same thread,
    made before that capture ends, must not use the captured second stream: it runs
    eagerly and is correct at once, and the graph's replays are correct too."""
    torch = torch_cuda
    keys, vals, off, kout, vout, call = _setup(pkg, torch)
    ek, ev = orc.seg_sort(keys, vals, offsets=off)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).copy()).cuda()
    kin2, vin2, od2 = dev(keys), dev(vals), dev(off.astype(np.int32))
    kout2, vout2 = torch.empty_like(kin2), torch.empty_like(vin2)
    D = pkg.DeviceSegmentedRadixSort
    n, s = keys.size, off.size - 1
    before = pkg.get_tuning(4, 4)
    pkg.set_tuning(4, 4, before["wave_cap"], before["tile_enable"], before["grid_mult"], 1)
    tb = [0]
    assert D._call(None, tb, None, None, None, None, n, s, None, None, 0, 32, None, 0, 0, 4, "batched") == 0
    temp2 = torch.empty(tb[0], dtype=torch.uint8, device="cuda")
    try:
        cap_stream, eager_stream = torch.cuda.Stream(), torch.cuda.Stream()
        with torch.cuda.stream(cap_stream):
            call(cap_stream.cuda_stream)  # warm-up outside the capture
        with torch.cuda.stream(eager_stream):
            rc = D._call(temp2, [temp2.numel()], kin2, kout2, vin2, vout2, n, s, od2.data_ptr(), od2.data_ptr() + 4,
                         0, 32, eager_stream.cuda_stream, 0, 0, 4, "batched")
            assert rc == 0, pkg.last_error()
        torch.cuda.synchronize()
        kout2.fill_(0)
        vout2.fill_(0)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=cap_stream, capture_error_mode="relaxed"):
            call(torch.cuda.current_stream().cuda_stream)
            # same host thread, another stream, not captured
            rc = D._call(temp2, [temp2.numel()], kin2, kout2, vin2, vout2, n, s, od2.data_ptr(), od2.data_ptr() + 4,
                         0, 32, eager_stream.cuda_stream, 0, 0, 4, "batched")
            assert rc == 0, pkg.last_error()
        eager_stream.synchronize()
        assert np.array_equal(kout2.cpu().numpy().view(np.uint32), ek)
        assert np.array_equal(vout2.cpu().numpy().view(np.uint32), ev)
        for rep in range(2):
            kout.fill_(0)
            vout.fill_(0)
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            assert np.array_equal(kout.cpu().numpy().view(np.uint32), ek), rep
            assert np.array_equal(vout.cpu().numpy().view(np.uint32), ev), rep
    finally:
        pkg.reset_tuning(4, 4)


def test_segmented_sort_on_side_stream(pkg, torch_cuda, orc):
    """segmented_sort(stream=s) with s not the current stream: the clones, outputs and
    temp storage are made on s after s waits for the current stream, so a freshly
    produced input is read complete; checked against the oracle several times with the
    input rewritten on the current stream just before each call."""
    torch = torch_cuda
    keys, vals, off = pkg.gen.ragged_lognormal(0x5E650A7 + 11, s=1 << 12, total=1 << 20)
    src_k = torch.from_numpy(keys.view(np.int32)).cuda()
    src_v = torch.from_numpy(vals.view(np.int32)).cuda()
    od = torch.from_numpy(off.astype(np.int32)).cuda()
    ek, ev = orc.seg_sort(keys, vals, offsets=off)
    side = torch.cuda.Stream()
    for rep in range(3):
        dk = torch.zeros_like(src_k)
        dv = torch.zeros_like(src_v)
        dk.copy_(src_k)  # produced on the current stream right before the call
        dv.copy_(src_v)
        ko, vo = pkg.segmented_sort(dk.view(torch.uint32), od, dv.view(torch.uint32), stream=side)
        side.synchronize()
        assert np.array_equal(ko.view(torch.int32).cpu().numpy().view(np.uint32), ek), rep
        assert np.array_equal(vo.view(torch.int32).cpu().numpy().view(np.uint32), ev), rep
        del dk, dv
    with pytest.raises(pkg.SegsortError):
        pkg.segmented_sort(src_k.view(torch.uint32), od.to(torch.int64), src_v.view(torch.uint32))



@pytest.mark.parametrize("fill", [0x00, 0xA5, 0xFF])
def test_replays_after_temp_overwritten(pkg, torch_cuda, orc, fill):
    """The call clears its own counters (no launch before the binning kernel): a graph
    of the call replayed after the caller overwrote the whole temp storage (zeros,
    0xA5 bytes, all ones), and a direct call after the same, are correct each time."""
    torch = torch_cuda
    keys, vals, off, kout, vout, call = _setup(pkg, torch)
    ek, ev = orc.seg_sort(keys, vals, offsets=off)
    st = torch.cuda.Stream()
    with torch.cuda.stream(st):
        call(st.cuda_stream)  # warm-up outside the capture; sizes the temp storage
    torch.cuda.synchronize()
    temp = call.temp[0]
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=st):
        call(torch.cuda.current_stream().cuda_stream)
    for rep in range(3):
        print(f"[case] replay {rep} after temp filled with {fill:#04x}", flush=True)
        kout.fill_(0)
        vout.fill_(0)
        temp.fill_(fill)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(kout.cpu().numpy().view(np.uint32), ek), (rep, "replay")
        assert np.array_equal(vout.cpu().numpy().view(np.uint32), ev), (rep, "replay")
        kout.fill_(0)
        vout.fill_(0)
        temp.fill_(fill)
        call(torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        assert np.array_equal(kout.cpu().numpy().view(np.uint32), ek), (rep, "direct")
        assert np.array_equal(vout.cpu().numpy().view(np.uint32), ev), (rep, "direct")
