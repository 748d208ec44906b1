"""Parity at the full benchmark workloads, through the C ABI, against the
FP64 oracle.

Every workload runs at its full size on the device; the FP64 reference is
sampled (random modes / points), because an FP64 direct sum over every mode
of a 16e6-point problem is out of reach on the host:

* 64^3 modes on a 128^3 grid, w = 7, 2,000,000 uniform points: Type-1 on
  4096 random modes, Type-2 on 4096 random points.
* 100^3 modes on a 200^3 grid and 127^3 modes on a 254^3 grid (Bluestein
  axis), 16,000,000 points (70 % in 8 Gaussian clusters of sigma 0.1, 30 %
  uniform), w = 4, 6, 8: Type-1 on 512 random modes, Type-2 on 2048 points.
  A workgroup's run here spans ~30k points, so the spreading's bounded
  accumulation chains (full flushes) are exercised.
* Gradient of ||T2(f, x) - b||^2 with 64^3 modes, 1,000,000 uniform points,
  w = 6: d/dx and d/df on 256 entries each, against the closed-form FP64
  direct-sum gradient and against FP64 central differences (h = 1e-6).
* 64 Type-1 problems of 32^3 / 48^3 / 64^3 / 96^3 modes (exact grids),
  100k - 4M points each on radial lines (density ~ 1/r^2), w = 6: the
  batched call against per-problem calls and against FP64 direct sums on
  512 random modes of three problems, plus the batched call captured in a
  HIP graph and replayed against eager calls.

Tolerance: relative L2 error <= 1e-5 against the FP64 sums.
"""
import math

import numpy as np
import pytest

from helpers import TWO_PI, clustered_points, cnormal, radial_points, rel_l2, uniform_points

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _t(x, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def _say(msg):
    print(msg, flush=True)


def test_uniform_2m_points_128_grid_w7(dev, oracle_mod):
    import apollofft_amd as A
    import torch
    rng = np.random.default_rng(1234)
    n, w, M = (64, 64, 64), 7, 2_000_000
    d = (TWO_PI / 64,) * 3
    xyz = uniform_points(rng, M)
    c = cnormal(rng, M)
    f = cnormal(rng, n)
    plan = A.NufftPlan3D(*n, *d, 2, w)
    assert plan.meta["m"] == (128, 128, 128)
    tx = _t(xyz, dev)
    o1 = plan.type1(tx, _t(c, dev))
    o2 = plan.type2(tx, _t(f, dev))
    torch.cuda.synchronize()
    x64 = xyz.astype(np.float64)
    sel = rng.choice(int(np.prod(n)), size=4096, replace=False).astype(np.int64)
    r1 = oracle_mod.direct_type1_3d(x64, c.astype(np.complex128), n, d, sel)
    e1 = rel_l2(o1.cpu().numpy().ravel()[sel], r1)
    pj = rng.choice(M, size=4096, replace=False)
    r2 = oracle_mod.direct_type2_3d(f.astype(np.complex128), x64[pj], d)
    e2 = rel_l2(o2.cpu().numpy()[pj], r2)
    _say(f"uniform 2M w7: type1 {e1:.3e} type2 {e2:.3e}")
    assert e1 < TOL and e2 < TOL


@pytest.fixture(scope="module")
def clustered_16m():
    """Points and data per mode count, generated once (shared by the widths)."""
    cache = {}

    def get(nn, oracle_mod):
        if nn not in cache:
            rng = np.random.default_rng(nn)
            M = 16_000_000
            n = (nn, nn, nn)
            d = (TWO_PI / nn,) * 3
            xyz = clustered_points(rng, M)
            c = cnormal(rng, M)
            f = cnormal(rng, n)
            x64 = xyz.astype(np.float64)
            sel = rng.choice(nn ** 3, size=512, replace=False).astype(np.int64)
            r1 = oracle_mod.direct_type1_3d(x64, c.astype(np.complex128), n, d, sel)
            pj = rng.choice(M, size=2048, replace=False)
            r2 = oracle_mod.direct_type2_3d(f.astype(np.complex128), x64[pj], d)
            cache.clear()  # keep one point set in memory
            cache[nn] = (xyz, c, f, sel, r1, pj, r2)
        return cache[nn]

    return get


@pytest.mark.parametrize("nn,grid", [(100, 200), (127, 254)])
@pytest.mark.parametrize("w", [4, 6, 8])
def test_clustered_16m_points(dev, oracle_mod, clustered_16m, nn, grid, w):
    import apollofft_amd as A
    import torch
    xyz, c, f, sel, r1, pj, r2 = clustered_16m(nn, oracle_mod)
    n = (nn, nn, nn)
    d = (TWO_PI / nn,) * 3
    plan = A.NufftPlan3D(*n, *d, 2, w, grid_policy="exact")
    assert plan.meta["m"] == (grid, grid, grid)
    tx = _t(xyz, dev)
    o1 = plan.type1(tx, _t(c, dev))
    o2 = plan.type2(tx, _t(f, dev))
    torch.cuda.synchronize()
    e1 = rel_l2(o1.cpu().numpy().ravel()[sel], r1)
    e2 = rel_l2(o2.cpu().numpy()[pj], r2)
    _say(f"clustered 16M {nn}->{grid} w{w}: type1 {e1:.3e} type2 {e2:.3e}")
    del tx, o1, o2, plan
    torch.cuda.empty_cache()
    assert e1 < TOL and e2 < TOL


def test_gradient_1m_points_64_modes_w6(dev, oracle_mod):
    """L = ||T2(f, x) - b||^2 through torch.autograd (the fused Type-2
    backward).  Closed forms (torch's convention for complex leaves):
    dL/dx_j^a = 2 Re(conj(r_j) sum_k i k_a (2 pi / L_a) f_k e^{i k.x_j}),
    dL/df_k = 2 sum_j r_j e^{-i k.x_j}, with r = T2(f, x) - b in FP64."""
    import torch
    import apollofft_amd as A
    from apollofft_amd.autograd import nufft_type2
    rng = np.random.default_rng(44)
    n, w, M = (64, 64, 64), 6, 1_000_000
    d = (TWO_PI / 64,) * 3
    xyz = uniform_points(rng, M)
    f = cnormal(rng, n)
    b = cnormal(rng, M)
    plan = A.NufftPlan3D(*n, *d, 2, w)
    tx = _t(xyz, dev).requires_grad_(True)
    tf = _t(f, dev).requires_grad_(True)
    loss = (nufft_type2(tf, tx, plan) - _t(b, dev)).abs().pow(2).sum()
    loss.backward()
    torch.cuda.synchronize()
    gx = tx.grad.cpu().numpy().astype(np.float64)
    gf = tf.grad.cpu().numpy().astype(np.complex128)
    x64 = xyz.astype(np.float64)
    f64 = f.astype(np.complex128)
    b64 = b.astype(np.complex128)

    # r at every point in FP64: the device direct sum, pinned here against
    # the oracle's direct sum on 256 points
    rd = plan.direct_type2(_t(xyz, dev), _t(f, dev)).cpu().numpy()
    pk = rng.choice(M, size=256, replace=False)
    assert rel_l2(rd[pk], oracle_mod.direct_type2_3d(f64, x64[pk], d)) < 1e-12
    r = rd - b64

    # d/dx: closed form and central differences on 256 (point, axis) entries
    pj = rng.choice(M, size=256, replace=False)
    ref_x = oracle_mod.direct_type2_3d_grad_x(f64, x64[pj], d, r[pj])
    ex = rel_l2(gx[pj], ref_x)
    axis = rng.integers(0, 3, size=256)
    h = 1e-6
    xp, xm = x64[pj].copy(), x64[pj].copy()
    xp[np.arange(256), axis] += h
    xm[np.arange(256), axis] -= h
    cp = oracle_mod.direct_type2_3d(f64, xp, d)
    cm = oracle_mod.direct_type2_3d(f64, xm, d)
    fd_x = (np.abs(cp - b64[pj]) ** 2 - np.abs(cm - b64[pj]) ** 2) / (2 * h)
    efx = rel_l2(gx[pj, axis], fd_x)

    # d/df: closed form 2 T1(r) on 256 modes, and central differences of the
    # loss in Re f_k and Im f_k, summed point by point in FP64
    sel = rng.choice(int(np.prod(n)), size=256, replace=False).astype(np.int64)
    ref_f = 2.0 * oracle_mod.direct_type1_3d(x64, r, n, d, sel)
    ef = rel_l2(gf.ravel()[sel], ref_f)
    nsel = 64
    fd_f = np.empty(nsel, dtype=np.complex128)
    kk = np.stack(np.unravel_index(sel[:nsel], n), axis=1)
    kk = np.where(kk <= np.array(n) // 2, kk, kk - np.array(n))
    for i in range(nsel):
        ph = x64 @ (kk[i] * (TWO_PI / (np.array(n) * np.array(d))))
        e = np.exp(1j * ph)
        dre = np.sum(np.abs(r + h * e) ** 2 - np.abs(r - h * e) ** 2) / (2 * h)
        dim = np.sum(np.abs(r + 1j * h * e) ** 2 - np.abs(r - 1j * h * e) ** 2) / (2 * h)
        fd_f[i] = dre + 1j * dim
    eff = rel_l2(gf.ravel()[sel[:nsel]], fd_f)
    _say(f"gradient 1M w6: dx {ex:.3e} (fd {efx:.3e}) df {ef:.3e} (fd {eff:.3e})")
    assert ex < TOL and ef < TOL
    assert efx < TOL and eff < TOL


def test_batched_64_problems_radial(dev, oracle_mod):
    import torch
    import apollofft_amd as A
    rng = np.random.default_rng(55)
    nprob, w = 64, 6
    sizes = rng.choice([32, 48, 64, 96], size=nprob)
    Ms = rng.integers(100_000, 4_000_001, size=nprob)
    plans, pos, vals, host = [], [], [], []
    cache = {}
    for nn, M in zip(sizes, Ms):
        nn = int(nn)
        if nn not in cache:
            cache[nn] = A.NufftPlan3D(nn, nn, nn, *(TWO_PI / nn,) * 3, 2, w, grid_policy="exact")
        plans.append(cache[nn])
        xyz = radial_points(rng, int(M))
        c = cnormal(rng, int(M))
        host.append((xyz, c))
        pos.append(_t(xyz, dev))
        vals.append(_t(c, dev))
    _say(f"batched: {nprob} problems, {int(Ms.sum())} points")
    outs = A.execute_type1_batched(plans, pos, vals)
    seq = [p.type1(x, v) for p, x, v in zip(plans, pos, vals)]
    torch.cuda.synchronize()
    worst = max(rel_l2(o.cpu().numpy(), s.cpu().numpy()) for o, s in zip(outs, seq))
    errs = []
    for i in rng.choice(nprob, size=3, replace=False):
        nn = int(sizes[i])
        xyz, c = host[i]
        sel = rng.choice(nn ** 3, size=512, replace=False).astype(np.int64)
        ref = oracle_mod.direct_type1_3d(xyz.astype(np.float64), c.astype(np.complex128),
                                         (nn,) * 3, (TWO_PI / nn,) * 3, sel)
        errs.append(rel_l2(outs[i].cpu().numpy().ravel()[sel], ref))
    _say(f"batched: max rel diff vs per-problem {worst:.3e}, direct {max(errs):.3e}")
    # two FP32 evaluations of the same sums with different work splits (so
    # different subproblem parts and strength scales), each within about 1e-6
    # of the FP64 sums, agree to within twice that
    assert worst < 2e-6
    assert max(errs) < TOL

    # the batched call captured in a HIP graph, replayed with eager calls and
    # synchronisations between the replays
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        A.execute_type1_batched(plans, pos, vals, outs)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    gouts = [torch.empty_like(o) for o in outs]
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        A.execute_type1_batched(plans, pos, vals, gouts)
    for rep in range(3):
        for o in gouts:
            o.fill_(float("nan"))
        g.replay()
        torch.cuda.synchronize()
        eager = A.execute_type1_batched(plans, pos, vals)
        torch.cuda.synchronize()
        worst = max(rel_l2(go.cpu().numpy(), e.cpu().numpy()) for go, e in zip(gouts, eager))
        assert math.isfinite(worst) and worst < 1e-6, rep
    _say("batched: graph replays match eager")
