"""Autotuner table logic (CPU) and a quick tuning run checked against the oracle (GPU)."""
import json

import numpy as np
import pytest


def test_signature_and_lookup(pkg, tmp_path):
    at = pkg.autotune
    assert at.signature(1 << 24, 1 << 16, 4, 4) == "k4v4-n24-m8"
    assert at.signature(1 << 20, 1024, 4, 0) == "k4v0-n20-m10"
    table = {"k4v4-n24-m8": {"wave_cap": 128, "tile_enable": 12, "grid_mult": 2, "concurrent": 1, "ms": 0.4},
             "k4v4-n20-m2": {"wave_cap": 512, "tile_enable": 15, "grid_mult": 1, "concurrent": 0, "ms": 0.1},
             "k8v4-n24-m4": {"wave_cap": 256, "tile_enable": 13, "grid_mult": 1, "concurrent": 0, "ms": 0.3}}
    assert at.lookup("k4v4-n24-m8", table)["wave_cap"] == 128          # exact
    assert at.lookup("k4v4-n23-m7", table)["tile_enable"] == 12        # nearest of the same widths
    assert at.lookup("k4v4-n20-m3", table)["wave_cap"] == 512
    assert at.lookup("k8v8-n24-m8", table) == at.DEFAULT                 # no entry for these widths
    # calls with a length bound: signature feature ceil(log2 bound), bounded entries first
    assert at.signature(1 << 24, 1 << 16, 4, 4, 16384) == "k4v4-n24-m8-x14"
    assert at.signature(1 << 24, 1 << 16, 4, 4, 16385) == "k4v4-n24-m8-x15"
    assert at.signature(1 << 24, 1 << 16, 4, 4, 1) == "k4v4-n24-m8-x0"
    bt = dict(table, **{"k4v4-n24-m8-x8": {"wave_cap": 256, "tile_enable": 3, "grid_mult": 6, "concurrent": 0},
                        "k4v4-n24-m8-x17": {"wave_cap": 512, "tile_enable": 11, "grid_mult": 4, "concurrent": 1}})
    assert at.lookup("k4v4-n24-m8-x8", bt)["grid_mult"] == 6            # exact bounded entry
    assert at.lookup("k4v4-n24-m8-x16", bt)["grid_mult"] == 4           # nearest bounded entry
    assert at.lookup("k4v4-n24-m8", bt)["wave_cap"] == 128              # unbounded calls: unbounded entries only
    assert at.lookup("k4v4-n20-m3-x5", table)["wave_cap"] == 512        # no bounded entry: unbounded ones
    p = tmp_path / "t.json"
    at.store("k4v0-n20-m10", at.DEFAULT, 1.5, path=p, note="x")
    assert json.loads(p.read_text())["k4v0-n20-m10"]["ms"] == 1.5


def test_keep_overlap_needs_isolated_calls_too():
    """A second-stream candidate that wins back to back is stored only if it also beats
    the best one-stream candidate on single synchronised calls."""
    import json

    from segsort_amd import autotune as at

    def key(conc, gm):
        return json.dumps({"wave_cap": 256, "tile_enable": 0x10B, "grid_mult": gm, "concurrent": conc}, sort_keys=True)

    over, one, one_b = key(2, 2), key(0, 2), key(0, 3)
    score = {over: 1.96, one: 2.01, one_b: 2.03, key(1, 2): None}
    assert at.keep_overlap(over, score, {over: 0.19, one: 0.20}.get) == over   # faster both ways
    assert at.keep_overlap(over, score, {over: 0.21, one: 0.20}.get) == one    # slower alone: best one-stream
    assert at.keep_overlap(one_b, score, lambda c: 1 / 0) == one_b             # one stream already: not timed
    assert at.keep_overlap(over, {over: 1.0}, lambda c: 1 / 0) == over          # no one-stream candidate


def test_apply_sets_library_tuning(pkg):
    table = {"k4v4-n24-m8": {"wave_cap": 128, "tile_enable": 12, "grid_mult": 2, "concurrent": 1}}
    t = pkg.autotune.apply(pkg, 1 << 24, 1 << 16, 4, 4, table)
    assert pkg.get_tuning(4, 4) == t == {"wave_cap": 128, "tile_enable": 12, "grid_mult": 2, "concurrent": 1}
    assert pkg.lookup_tuning(1 << 24, 1 << 16, 4, 4)["from_table"] is False
    pkg.reset_tuning(4, 4)
    assert pkg.lookup_tuning(1 << 24, 1 << 16, 4, 4)["from_table"] is True


def test_library_table_matches_tuning_json(pkg):
    """The table compiled into the library (csrc/tuning_table.inc, generated from
    tuning.json by build.py) gives, for every stored signature and for shapes between
    them, the same tuning as the autotuner's own nearest-signature lookup of tuning.json."""
    import sys
    sys.path.insert(0, str(pkg.HERE))
    try:
        import build
    finally:
        sys.path.pop(0)
    table = pkg.autotune.load_table()
    assert (pkg.HERE / "csrc" / "tuning_table.inc").read_text() == build.table_source(table)
    pkg.reset_tuning(0, 0)
    keys = ("wave_cap", "tile_enable", "grid_mult", "concurrent")
    shapes = [(1 << 24, 1 << k) for k in range(0, 25, 1)] + [(1 << 20, 1024), (17301504, 1 << 20), (5000, 7),
                                                             (1 << 26, 1 << 15), (16809984, 4096)]
    for kw, vb in ((4, 4), (4, 0), (8, 4), (4, 8), (8, 8), (8, 0)):
        for n, s in shapes:
            for mx in (0, 1, 32, 256, 4096, 16384, 16385, 1 << 17, n):
                want = pkg.autotune.lookup(pkg.autotune.signature(n, s, kw, vb, mx), table)
                got = pkg.lookup_tuning(n, s, kw, vb, mx)
                assert got["from_table"] and {k: got[k] for k in keys} == want, (kw, vb, n, s, mx)


@pytest.mark.gpu
def test_quick_tune_is_exact(pkg, torch_cuda, orc):
    torch = torch_cuda
    keys, vals, off = pkg.gen.ragged_lognormal(11, s=1 << 10, total=1 << 20)
    dk = torch.from_numpy(keys.view(np.int32)).cuda().view(torch.uint32)
    dv = torch.from_numpy(vals.view(np.int32)).cuda().view(torch.uint32)
    do = torch.from_numpy(off).cuda()
    best, results = pkg.autotune.tune(pkg, torch, dk, dv, do, steps=1, warmup=0, verbose=False, quick=True)
    assert best is not None and all(v is not None for v in results.values())
    pkg.set_tuning(4, 4, **best[0])
    try:
        ko, vo = pkg.segmented_sort(dk, do, dv)
        ek, ev = orc.seg_sort(keys, vals, offsets=off)
        assert np.array_equal(ko.view(torch.int32).cpu().numpy().view(np.uint32), ek)
        assert np.array_equal(vo.view(torch.int32).cpu().numpy().view(np.uint32), ev)
    finally:
        pkg.reset_tuning(4, 4)
