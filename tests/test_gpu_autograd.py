"""Differentiable pair sort: forward = stable per-row ascending sort returning the
sorted values and the row-local permutation; backward = the library's gradient
scatter through that permutation. Checked exactly against a per-row stable
torch.sort reference (values, permutation and gradient), against the oracle at the
full 4096-row size, and against central finite differences (eps = 1e-3,
rtol = 1e-4) on coordinates whose row neighbours are farther than 2*eps away."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPS = 1e-3
RTOL = 1e-4


def _inputs(pkg, torch, rows, lo, hi, seed=0x5E650A7 + 4):
    x, off = pkg.gen.diff_rows(seed, rows=rows, lo=lo, hi=hi)
    xd = torch.from_numpy(x).cuda().requires_grad_(True)
    od = torch.from_numpy(off).cuda()
    return x, off, xd, od


@pytest.mark.parametrize("rows,lo,hi", [(64, 16, 3000), (12, 6000, 30000)], ids=["short-rows", "long-rows"])
@pytest.mark.parametrize("algo", ["batched", "plain"])
def test_forward_backward_exact_vs_torch_sort(pkg, torch_cuda, algo, rows, lo, hi):
    """Rows up to 3000 items, and rows of 6000..30000 items on both sides of the
    gradient kernel's 8192-item LDS path."""
    torch = torch_cuda
    x, off, xd, od = _inputs(pkg, torch, rows=rows, lo=lo, hi=hi)
    y, perm = pkg.segmented_pair_sort(xd, od, algo=algo)
    g = torch.from_numpy(pkg.gen.normal(99, 1, x.size).astype(np.float32)).cuda()
    (gin,) = torch.autograd.grad(y, xd, g)
    y, perm, gin, g = y.detach().cpu(), perm.cpu(), gin.cpu(), g.cpu()
    xt = torch.from_numpy(x)
    for r in range(len(off) - 1):
        a, b = int(off[r]), int(off[r + 1])
        ys, idx = torch.sort(xt[a:b], stable=True)
        assert torch.equal(y[a:b], ys), r
        assert torch.equal(perm[a:b].long(), idx), r
        ref = torch.zeros(b - a)
        ref[idx] = g[a:b]
        assert torch.equal(gin[a:b], ref), r


def test_full_size_against_oracle(pkg, torch_cuda, orc):
    """4096 rows of U[16, 8192] float32 values (~16.8M elements)."""
    torch = torch_cuda
    x, off, xd, od = _inputs(pkg, torch, rows=4096, lo=16, hi=8192)
    y, perm = pkg.segmented_pair_sort(xd, od)
    gnp = pkg.gen.normal(98, 1, x.size).astype(np.float32)
    (gin,) = torch.autograd.grad(y, xd, torch.from_numpy(gnp).cuda())
    local = np.concatenate([np.arange(b - a, dtype=np.int32) for a, b in zip(off[:-1], off[1:])])
    ek, ev = orc.seg_sort(x, local, offsets=off)
    assert np.array_equal(y.detach().cpu().numpy().view(np.uint32), ek.view(np.uint32))
    assert np.array_equal(perm.cpu().numpy(), ev)
    assert np.array_equal(gin.cpu().numpy(), orc.perm_scatter(gnp, ev, off))


def test_central_finite_differences(pkg, torch_cuda):
    torch = torch_cuda
    x, off, xd, od = _inputs(pkg, torch, rows=32, lo=16, hi=1024, seed=0x5E650A7 + 5)
    w = pkg.gen.normal(97, 1, x.size).astype(np.float32)
    y, _ = pkg.segmented_pair_sort(xd, od)
    (gin,) = torch.autograd.grad(y, xd, torch.from_numpy(w).cuda())
    gin = gin.cpu().numpy().astype(np.float64)
    rng = np.random.default_rng(5)
    checked = 0
    for r in rng.permutation(len(off) - 1):
        a, b = int(off[r]), int(off[r + 1])
        row = x[a:b].astype(np.float64)
        srt = np.sort(row)
        for i in rng.choice(b - a, size=3, replace=False):
            v = row[i]
            k = np.searchsorted(srt, v)
            gap_lo = v - srt[k - 1] if k > 0 else np.inf
            gap_hi = srt[k + 1] - v if k + 1 < len(srt) else np.inf
            if min(gap_lo, gap_hi) <= 2 * EPS:
                continue  # a perturbation of eps could reorder this coordinate
            f = []
            for sgn in (1.0, -1.0):
                xp = x.copy()
                xp[a + i] = np.float32(x[a + i] + sgn * EPS)
                yp, _ = pkg.segmented_pair_sort(torch.from_numpy(xp).cuda(), od)
                f.append(float(np.dot(w[a:b].astype(np.float64), yp[a:b].cpu().numpy().astype(np.float64))))
            step = float(np.float32(x[a + i] + EPS)) - float(np.float32(x[a + i] - EPS))
            fd = (f[0] - f[1]) / step
            assert abs(fd - gin[a + i]) <= RTOL * abs(gin[a + i]) + 1e-6, (r, i, fd, gin[a + i])
            checked += 1
        if checked >= 24:
            break
    assert checked >= 12
