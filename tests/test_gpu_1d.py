"""The binned 1D path on the device: the 1D point binning (the counting
sort that replaces sort_points_1d / sort_positions_1d, transform/
dimension_1d.rs:54-99) against a CPU replay, bit for bit, and Type-1,
Type-2 and the coordinate gradient at 1,000,000 points on a prime grid
against FP64 references (the oracle's direct sums, and the closed-form
derivative of the Type-2 sum in numpy)."""
import ctypes

import numpy as np
import pytest

from helpers import TWO_PI, cnormal, rel_l2
from replay import fold_cells

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _t(x, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def _points(rng, M, case):
    if case == "uniform":
        return rng.uniform(-np.pi, np.pi, M).astype(np.float32)
    if case == "clustered":  # 70 % in three narrow clusters (dense, split tiles)
        mc = int(0.7 * M)
        centres = rng.uniform(-np.pi, np.pi, 3)
        x = np.concatenate([centres[rng.integers(0, 3, mc)] + 0.02 * rng.standard_normal(mc),
                            rng.uniform(-np.pi, np.pi, M - mc)])
        return x[rng.permutation(M)].astype(np.float32)
    vals = np.array([-np.pi, np.pi, 0.0, -1e-30, 1e-30, -1e-7, TWO_PI, -TWO_PI,
                     3 * TWO_PI + 0.1, -5 * TWO_PI - 0.2, np.nextafter(np.float32(np.pi), 0)],
                    dtype=np.float32)
    return rng.choice(vals, size=M).astype(np.float32)


@pytest.mark.parametrize("case", ["uniform", "clustered", "edges"])
def test_1d_binning_matches_cpu_replay(dev, case):
    import torch
    import apollofft_amd as A
    import apollofft_amd._native as N
    rng = np.random.default_rng(5)
    n, w = 127, 6
    plan = A.NufftPlan1D(n, TWO_PI / n, 2, w)
    meta = plan.meta
    m, b = meta["m"][0], meta["bin_size"][0]
    assert m == 254 and b == 64 - 2 * w + 1
    x = _points(rng, 50_000, case)
    M = x.size
    keys = fold_cells(x, n, TWO_PI / n, m) // b
    perm_ref = np.argsort(keys, kind="stable").astype(np.uint32)
    nb = ctypes.c_int64()
    N.check(N.lib().anufft_debug_num_bins(plan.handle, ctypes.byref(nb)))
    assert nb.value == -(-m // b)
    bs_ref = np.searchsorted(keys[perm_ref], np.arange(nb.value + 1), side="left").astype(np.uint32)
    tx = _t(x, dev)
    perm = torch.empty(M, dtype=torch.int32, device=dev)
    bstart = torch.empty(nb.value + 1, dtype=torch.int32, device=dev)
    ws = plan.workspace(M, dev)
    p = tx.data_ptr()
    N.check(N.lib().anufft_debug_sort(plan.handle, M, N.ptr_array([p, p, p]), 1, perm.data_ptr(),
                                      bstart.data_ptr(), ws.data_ptr(), ws.numel(),
                                      ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    torch.cuda.synchronize()
    assert np.array_equal(perm.cpu().numpy().view(np.uint32), perm_ref)
    assert np.array_equal(bstart.cpu().numpy().view(np.uint32), bs_ref)


@pytest.mark.parametrize("case", ["uniform", "clustered"])
def test_1d_million_points_prime_grid(dev, oracle_mod, case):
    """n = 1009 modes (prime) on m = 2018 (2 x 1009: a Bluestein axis),
    w = 6, 1,000,000 points: Type-1 on every mode and Type-2 on 4096
    points against the FP64 direct sums; the coordinate gradient
    Re(conj(u) d/dx sum_k f_k e^{+2 pi i k x / L}) on 512 points against
    the closed form in FP64."""
    import torch
    import apollofft_amd as A
    rng = np.random.default_rng(7)
    n, w, M = 1009, 6, 1_000_000
    dx = TWO_PI / n
    plan = A.NufftPlan1D(n, dx, 2, w)
    assert plan.meta["m"][0] == 2018
    x = _points(rng, M, case)
    c = cnormal(rng, M)
    f = cnormal(rng, n)
    u = cnormal(rng, M)
    tx = _t(x, dev)
    print(f"1d {case}: M {M}", flush=True)
    o1 = plan.type1(tx, _t(c, dev))
    o2 = plan.type2(tx, _t(f, dev))
    g = plan.grad_points(tx, _t(u, dev), _t(f, dev))
    torch.cuda.synchronize()
    x64 = x.astype(np.float64)
    r1 = oracle_mod.direct_type1_1d(x64, c.astype(np.complex128), n, dx)
    e1 = rel_l2(o1.cpu().numpy(), r1)
    pj = rng.choice(M, size=4096, replace=False)
    r2 = oracle_mod.direct_type2_1d(f.astype(np.complex128), x64[pj], dx)
    e2 = rel_l2(o2.cpu().numpy()[pj], r2)
    pg = rng.choice(M, size=512, replace=False)
    k = np.arange(n)
    k = np.where(k <= n // 2, k, k - n).astype(np.float64)
    om = TWO_PI * k / (n * dx)
    dsum = (np.exp(1j * np.outer(x64[pg], om)) * (1j * om)) @ f.astype(np.complex128)
    gref = np.real(np.conj(u[pg].astype(np.complex128)) * dsum)
    eg = rel_l2(g.cpu().numpy()[pg, 0], gref)
    print(f"1d {case}: type1 {e1:.3e} type2 {e2:.3e} grad {eg:.3e}", flush=True)
    assert e1 < TOL and e2 < TOL and eg < TOL
