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
    p = tmp_path / "t.json"
    at.store("k4v0-n20-m10", at.DEFAULT, 1.5, path=p, note="x")
    assert json.loads(p.read_text())["k4v0-n20-m10"]["ms"] == 1.5


def test_apply_sets_library_tuning(pkg):
    before = pkg.get_tuning(4, 4)
    table = {"k4v4-n24-m8": {"wave_cap": 128, "tile_enable": 12, "grid_mult": 2, "concurrent": 1}}
    t = pkg.autotune.apply(pkg, 1 << 24, 1 << 16, 4, 4, table)
    assert pkg.get_tuning(4, 4) == t == {"wave_cap": 128, "tile_enable": 12, "grid_mult": 2, "concurrent": 1}
    pkg.set_tuning(4, 4, **before)


@pytest.mark.gpu
def test_quick_tune_is_exact(pkg, torch_cuda, orc):
    torch = torch_cuda
    keys, vals, off = pkg.gen.ragged_lognormal(11, s=1 << 10, total=1 << 20)
    dk = torch.from_numpy(keys.view(np.int32)).cuda().view(torch.uint32)
    dv = torch.from_numpy(vals.view(np.int32)).cuda().view(torch.uint32)
    do = torch.from_numpy(off).cuda()
    best, results = pkg.autotune.tune(pkg, torch, dk, dv, do, steps=1, warmup=0, verbose=False, quick=True)
    assert best is not None and all(v is not None for v in results.values())
    before = pkg.get_tuning(4, 4)
    pkg.set_tuning(4, 4, **best[0])
    try:
        ko, vo = pkg.segmented_sort(dk, do, dv)
        ek, ev = orc.seg_sort(keys, vals, offsets=off)
        assert np.array_equal(ko.view(torch.int32).cpu().numpy().view(np.uint32), ek)
        assert np.array_equal(vo.view(torch.int32).cpu().numpy().view(np.uint32), ev)
    finally:
        pkg.set_tuning(4, 4, **before)
