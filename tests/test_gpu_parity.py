"""Device parity tests (HIP path through the C ABI) against the FP64 oracle.

Every test runs the HIP kernels of libapollo_nufft_hip.so on cuda:0 and
compares with the oracle (FP64 direct sums, the FP64 restatement of
apollo-nufft, or the numpy replay of the binning) on the same float32
coordinates.  Tolerance for floating point: relative L2 error <= 1e-5
against the FP64 direct sum (the target accuracy of the workload), and the
binning permutation must match the CPU replay exactly.
"""
import ctypes

import numpy as np
import pytest

from helpers import TWO_PI, clustered_points, cnormal, max_rel, rel_l2, uniform_points
from replay import replay_sort

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _t(x, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def _plan3(n, w=6, policy="reference", compat=False, bins=None, max_sub=None, method=None):
    import apollofft_amd as A
    d = [TWO_PI / k for k in n]
    return A.NufftPlan3D(n[0], n[1], n[2], d[0], d[1], d[2], 2, w, grid_policy=policy,
                         compat_reference_type2_3d_scaling=compat, bin_size=bins,
                         max_subproblem=max_sub, kernel_method=method), d


METHODS = ["valu", "mfma"]


# ---------------------------------------------------------------- binning
@pytest.mark.parametrize("case", ["uniform", "clustered", "edges"])
def test_binning_matches_cpu_replay(dev, case):
    rng = np.random.default_rng(11)
    n = (24, 20, 16)
    plan, d = _plan3(n, w=5, policy="exact")
    if case == "uniform":
        xyz = uniform_points(rng, 50000)
    elif case == "clustered":
        xyz = clustered_points(rng, 50000)
    else:  # boundary values, exact multiples of L, tiny negatives, far periods
        vals = np.array([-np.pi, np.pi, 0.0, -1e-30, 1e-30, -1e-7, TWO_PI, -TWO_PI,
                         3 * TWO_PI + 0.1, -5 * TWO_PI - 0.2, np.nextafter(np.float32(np.pi), 0)],
                        dtype=np.float32)
        xyz = rng.choice(vals, size=(3000, 3)).astype(np.float32)
    _check_binning(dev, plan, n, d, xyz)


@pytest.mark.parametrize("n,nbins", [
    ((105, 176, 176), 7680),   # subproblem list beside the scatter, 150 KiB of LDS
    ((126, 176, 176), 9216),   # list launch of its own, largest one-pass tile count
])
def test_binning_near_lds_limits_matches_cpu_replay(dev, n, nbins):
    """One-pass binning at tile counts where the scatter's LDS (count rows,
    tile starts) is close to 160 KiB, on both sides of the fused-list limit
    (w = 5 matrix-core tiles of 7 x 22 x 22 cells on 2n grids)."""
    rng = np.random.default_rng(5)
    plan, d = _plan3(n, w=5, policy="exact", method="mfma")
    assert tuple(plan.meta["bin_size"]) == (7, 22, 22)
    xyz = np.concatenate([uniform_points(rng, 150000), clustered_points(rng, 50000)])
    assert _check_binning(dev, plan, n, d, xyz) == nbins


def _check_binning(dev, plan, n, d, xyz):
    """The device permutation and tile starts equal the CPU replay's; returns
    the tile count."""
    import apollofft_amd._native as N
    import torch
    meta = plan.meta
    M = xyz.shape[0]
    keys, perm_ref, bs_ref = replay_sort(xyz, n, d, meta["m"], meta["bin_size"])
    nb = ctypes.c_int64()
    N.check(N.lib().anufft_debug_num_bins(plan.handle, ctypes.byref(nb)))
    assert nb.value + 1 == bs_ref.size
    txyz = _t(xyz, dev)
    perm = torch.empty(M, dtype=torch.int32, device=dev)
    bs = torch.empty(nb.value + 1, dtype=torch.int32, device=dev)
    ws = plan.workspace(M, dev)
    b = txyz.data_ptr()
    N.check(N.lib().anufft_debug_sort(plan.handle, M, N.ptr_array([b, b + 4, b + 8]), 3,
                                      perm.data_ptr(), bs.data_ptr(), ws.data_ptr(), ws.numel(),
                                      None))
    torch.cuda.synchronize()
    assert np.array_equal(perm.cpu().numpy().view(np.uint32), perm_ref)
    assert np.array_equal(bs.cpu().numpy().view(np.uint32), bs_ref)
    return nb.value


# ---------------------------------------------------------------- spreading
@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("w", [2, 3, 4, 5, 6, 7, 8])
def test_spread_fine_grid_matches_oracle(dev, oracle_mod, w, method):
    """The oversampled grid after spreading equals the FP64 restatement's
    (reference type1_into spread loop) for both spreading kernels."""
    import apollofft_amd._native as N
    import torch
    rng = np.random.default_rng(w)
    n = (12, 10, 16)
    plan, d = _plan3(n, w=w, method=method)
    meta = plan.meta
    assert meta["kernel_method"] == method
    xyz = uniform_points(rng, 3000)
    c = cnormal(rng, 3000)
    ref = oracle_mod.spread3d(xyz.astype(np.float64), c.astype(np.complex128), n, d, 2, w,
                              oracle_mod.POLICY_POW2)
    grid = torch.empty(tuple(meta["m"]), dtype=torch.complex64, device=dev)
    txyz, tc = _t(xyz, dev), _t(c, dev)
    ws = plan.workspace(3000, dev)
    b = txyz.data_ptr()
    N.check(N.lib().anufft_debug_spread(plan.handle, 3000, N.ptr_array([b, b + 4, b + 8]), 3,
                                        tc.data_ptr(), grid.data_ptr(), ws.data_ptr(), ws.numel(),
                                        None))
    g = grid.cpu().numpy()
    assert rel_l2(g, ref) < 1e-6


# ---------------------------------------------------------------- 3D type 1 / 2
CASES_3D = [
    ((16, 16, 16), 6, "reference"),
    ((12, 10, 16), 7, "reference"),
    ((15, 9, 10), 4, "exact"),    # m = 30, 18, 20 (radix 2,3,5)
    ((13, 7, 11), 8, "exact"),    # m = 26, 17 (max(14,17)), 22: prime radices 13, 17, 11
    ((8, 8, 8), 2, "reference"),
    ((20, 24, 18), 5, "exact"),
    ((127, 6, 10), 6, "exact"),   # m = 254 = 2 * 127: Bluestein, 512-point FFTs
    ((10, 12, 14), 3, "reference"),
]


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("n,w,policy", CASES_3D)
def test_type1_3d_vs_direct(dev, oracle_mod, n, w, policy, method):
    rng = np.random.default_rng(100 + w)
    plan, d = _plan3(n, w=w, policy=policy, method=method)
    M = 4000
    xyz = uniform_points(rng, M)
    c = cnormal(rng, M)
    out = plan.type1(_t(xyz, dev), _t(c, dev)).cpu().numpy()
    ref = oracle_mod.direct_type1_3d(xyz.astype(np.float64), c.astype(np.complex128), n, d)
    fast = oracle_mod.nufft3d_type1_fast(xyz.astype(np.float64), c.astype(np.complex128), n, d,
                                         2, w, 1 if policy == "exact" else 0)
    kb_err = rel_l2(fast, ref)  # the algorithm's own (FP64) truncation error
    err = rel_l2(out, ref)
    assert err < max(TOL, 2.0 * kb_err), (err, kb_err)


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("n,w,policy", CASES_3D)
def test_type2_3d_vs_direct(dev, oracle_mod, n, w, policy, method):
    rng = np.random.default_rng(200 + w)
    plan, d = _plan3(n, w=w, policy=policy, method=method)
    M = 3000
    xyz = uniform_points(rng, M)
    f = cnormal(rng, n)
    out = plan.type2(_t(xyz, dev), _t(f, dev)).cpu().numpy()
    ref = oracle_mod.direct_type2_3d(f.astype(np.complex128), xyz.astype(np.float64), d)
    fast = oracle_mod.nufft3d_type2_fast(f.astype(np.complex128), xyz.astype(np.float64), d, 2, w,
                                         1 if policy == "exact" else 0)
    kb_err = rel_l2(fast, ref)
    err = rel_l2(out, ref)
    assert err < max(TOL, 2.0 * kb_err), (err, kb_err)


def test_three_digit_binning_with_records(dev, oracle_mod):
    """A 512^3 grid with w = 7 has more than 65536 spreading tiles, so the
    binning takes three digit passes and the point records go out -> tmp ->
    out; Type-1 at sampled modes and Type-2 at a few points against FP64
    direct sums."""
    import apollofft_amd._native as N
    n, w = (256, 256, 256), 7
    plan, d = _plan3(n, w=w, method="mfma")
    nb = ctypes.c_int64()
    N.check(N.lib().anufft_debug_num_bins(plan.handle, ctypes.byref(nb)))
    assert nb.value > 65536
    rng = np.random.default_rng(31)
    M = 3000
    xyz = uniform_points(rng, M)
    c = cnormal(rng, M)
    out = plan.type1(_t(xyz, dev), _t(c, dev)).cpu().numpy().ravel()
    sel = rng.choice(int(np.prod(n)), size=400, replace=False).astype(np.int64)
    ref = oracle_mod.direct_type1_3d(xyz.astype(np.float64), c.astype(np.complex128), n, d, sel)
    assert rel_l2(out[sel], ref) < TOL
    f = cnormal(rng, n)
    xyz2 = xyz[:200]
    out2 = plan.type2(_t(xyz2, dev), _t(f, dev)).cpu().numpy()
    ref2 = oracle_mod.direct_type2_3d(f.astype(np.complex128), xyz2.astype(np.float64), d)
    assert rel_l2(out2, ref2) < TOL


# ---------------------------------------------------------------- direct sums
@pytest.mark.parametrize("layout", ["aos", "soa"])
def test_direct_sums_3d_vs_oracle(dev, oracle_mod, layout):
    """FP64 device direct sums (the WGPU direct backend's sums) against the
    oracle's FP64 direct sums on the same float32 coordinates."""
    rng = np.random.default_rng(41)
    n = (12, 10, 16)
    plan, d = _plan3(n, w=4)
    M = 300
    xyz = uniform_points(rng, M) * np.float32(1.7)  # some points outside one period
    c = cnormal(rng, M)
    f = cnormal(rng, n)
    pos = _t(xyz, dev) if layout == "aos" else tuple(_t(xyz[:, a].copy(), dev) for a in range(3))
    out1 = plan.direct_type1(pos, _t(c, dev)).cpu().numpy()
    ref1 = oracle_mod.direct_type1_3d(xyz.astype(np.float64), c.astype(np.complex128), n, d)
    assert out1.dtype == np.complex128 and out1.shape == n
    assert rel_l2(out1.ravel(), ref1.ravel()) < 1e-12
    out2 = plan.direct_type2(pos, _t(f, dev)).cpu().numpy()
    ref2 = oracle_mod.direct_type2_3d(f.astype(np.complex128), xyz.astype(np.float64), d)
    assert rel_l2(out2, ref2) < 1e-12


def test_direct_sums_1d_long_lines_and_empty(dev, oracle_mod):
    """1D direct sums with n = 2048: the Type-2 mode line spans two LDS
    chunks of 1024 and its negative-index run starts inside the second; no
    points: Type-1 gives zeros, Type-2 nothing."""
    import apollofft_amd as A
    rng = np.random.default_rng(42)
    n, M = 2048, 200
    dx = TWO_PI / n
    plan = A.NufftPlan1D(n, dx, 2, 6)
    x = rng.uniform(-np.pi, np.pi, M).astype(np.float32)
    c = cnormal(rng, M)
    f = cnormal(rng, n)
    out1 = plan.direct_type1(_t(x, dev), _t(c, dev)).cpu().numpy()
    ref1 = oracle_mod.direct_type1_1d(x.astype(np.float64), c.astype(np.complex128), n, dx)
    assert rel_l2(out1, ref1) < 1e-12
    out2 = plan.direct_type2(_t(x, dev), _t(f, dev)).cpu().numpy()
    ref2 = oracle_mod.direct_type2_1d(f.astype(np.complex128), x.astype(np.float64), dx)
    assert rel_l2(out2, ref2) < 1e-12
    import torch
    z = plan.direct_type1(torch.empty(0, dtype=torch.float32, device=dev),
                          torch.empty(0, dtype=torch.complex64, device=dev)).cpu().numpy()
    assert z.shape == (n,) and not np.any(z)
    e = plan.direct_type2(torch.empty(0, dtype=torch.float32, device=dev), _t(f, dev))
    assert e.numel() == 0


def test_validation_report_passes(dev):
    """The reference's five NUFFT suite cases (suite.rs:227-309) in the
    NufftReport schema (report.rs:60-73); the reference's report test asserts
    report.nufft.passed (suite.rs:2694)."""
    from apollofft_amd.validation import NUFFT_FAST_RELATIVE_LIMIT, run_nufft_suite
    r = run_nufft_suite(dev.index or 0)
    assert set(r) == {"type1_1d_max_relative_error", "type2_1d_max_relative_error",
                      "type1_3d_max_relative_error", "irrational_positions_max_relative_error",
                      "clustered_positions_max_relative_error", "passed"}
    for k, v in r.items():
        if k != "passed":
            assert 0.0 <= v <= NUFFT_FAST_RELATIVE_LIMIT, (k, v)
    assert r["passed"] is True


def test_one_plan_two_streams_concurrently(dev):
    """Type-1 and Type-2 calls on one plan from two streams with their own
    workspaces, enqueued from two host threads: the per-call fork/join of the
    plan's second stream keeps each call's results equal to a serial run."""
    import threading
    import torch
    import apollofft_amd._native as N
    rng = np.random.default_rng(51)
    n = (24, 20, 16)
    plan, d = _plan3(n, w=6)
    M = 20000
    xyz = [_t(uniform_points(rng, M), dev) for _ in range(2)]
    c = [_t(cnormal(rng, M), dev) for _ in range(2)]
    f = [_t(cnormal(rng, n), dev) for _ in range(2)]
    ref1 = [plan.type1(xyz[i], c[i]).clone() for i in range(2)]
    ref2 = [plan.type2(xyz[i], f[i]).clone() for i in range(2)]
    torch.cuda.synchronize()
    need = plan.workspace_size(M)
    ws = [torch.empty(need, dtype=torch.uint8, device=dev) for _ in range(2)]
    streams = [torch.cuda.Stream(device=dev) for _ in range(2)]
    out1 = [torch.empty(n, dtype=torch.complex64, device=dev) for _ in range(2)]
    out2 = [torch.empty(M, dtype=torch.complex64, device=dev) for _ in range(2)]
    errors = []

    def run(i):
        try:
            b = xyz[i].data_ptr()
            ptrs = N.ptr_array([b, b + 4, b + 8])
            sp = ctypes.c_void_p(streams[i].cuda_stream)
            for _ in range(5):
                N.check(N.lib().anufft_execute_type1(plan.handle, M, ptrs, 3, c[i].data_ptr(),
                                                     out1[i].data_ptr(), ws[i].data_ptr(),
                                                     ws[i].numel(), sp))
                N.check(N.lib().anufft_execute_type2(plan.handle, M, ptrs, 3, f[i].data_ptr(),
                                                     out2[i].data_ptr(), ws[i].data_ptr(),
                                                     ws[i].numel(), sp))
        except Exception as e:  # reported below
            errors.append(e)

    th = [threading.Thread(target=run, args=(i,)) for i in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    torch.cuda.synchronize()
    assert not errors, errors
    for i in range(2):
        assert rel_l2(out1[i].cpu().numpy(), ref1[i].cpu().numpy()) < 1e-6
        assert rel_l2(out2[i].cpu().numpy(), ref2[i].cpu().numpy()) < 1e-6


def test_type2_3d_compat_scaling(dev, oracle_mod):
    """compat flag reproduces the reference's direct/(mx my mz) output
    (dimension_3d.rs:549-551)."""
    rng = np.random.default_rng(5)
    n = (8, 8, 8)
    plan, d = _plan3(n, w=6, compat=True)
    xyz = uniform_points(rng, 500)
    f = cnormal(rng, n)
    out = plan.type2(_t(xyz, dev), _t(f, dev)).cpu().numpy()
    ref = oracle_mod.nufft3d_type2_fast(f.astype(np.complex128), xyz.astype(np.float64), d, 2, 6,
                                        0, compat_scaling=True)
    assert rel_l2(out, ref) < TOL


@pytest.mark.parametrize("method", METHODS)
def test_type1_clustered_and_tiny_subproblems(dev, oracle_mod, method):
    """Clustered points force many subproblems per tile (max_subproblem=64)."""
    rng = np.random.default_rng(7)
    n = (16, 16, 16)
    plan, d = _plan3(n, w=6, max_sub=64, method=method)
    xyz = clustered_points(rng, 20000, sigma=0.05)
    c = cnormal(rng, 20000)
    out = plan.type1(_t(xyz, dev), _t(c, dev)).cpu().numpy()
    ref = oracle_mod.direct_type1_3d(xyz.astype(np.float64), c.astype(np.complex128), n, d)
    assert rel_l2(out, ref) < TOL


@pytest.mark.parametrize("max_sub,sub_work,per_cu", [(100, 0, 2), (1000, 1, 3), (16384, 384, 1)])
def test_work_shares_with_odd_settings(dev, oracle_mod, max_sub, sub_work, per_cu):
    """The persistent launches split subproblems at the boundaries of the
    workgroups' work shares; with an odd max_subproblem, no or a large fixed
    cost per subproblem and one to three workgroups per CU the shares still
    cover every point once (Type-1 and Type-2 against FP64 direct sums)."""
    import apollofft_amd as A
    rng = np.random.default_rng(17)
    n, w = (20, 16, 24), 6
    d = [TWO_PI / k for k in n]
    plan = A.NufftPlan3D(*n, *d, 2, w, kernel_method="mfma", max_subproblem=max_sub,
                         sub_work=sub_work, blocks_per_cu=per_cu)
    M = 30000
    xyz = clustered_points(rng, M)
    c = cnormal(rng, M)
    f = cnormal(rng, n)
    tx = _t(xyz, dev)
    x64 = xyz.astype(np.float64)
    out1 = plan.type1(tx, _t(c, dev)).cpu().numpy()
    sel = rng.choice(int(np.prod(n)), size=512, replace=False).astype(np.int64)
    ref1 = oracle_mod.direct_type1_3d(x64, c.astype(np.complex128), n, d, sel)
    assert rel_l2(out1.ravel()[sel], ref1) < TOL
    out2 = plan.type2(tx, _t(f, dev)).cpu().numpy()
    pj = rng.choice(M, size=500, replace=False)
    ref2 = oracle_mod.direct_type2_3d(f.astype(np.complex128), x64[pj], d)
    assert rel_l2(out2[pj], ref2) < TOL


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("scale", [1e-30, 1e-6, 1.0, 1e8, 1e30])
def test_type1_strength_range(dev, oracle_mod, method, scale):
    """Strengths spanning six decades inside one problem, at overall scales
    from 1e-30 to 1e30: the MFMA kernel's per-subproblem power-of-two
    normalisation keeps its FP16 operands in range."""
    rng = np.random.default_rng(11)
    n = (16, 12, 10)
    plan, d = _plan3(n, w=7, method=method)
    M = 3000
    xyz = uniform_points(rng, M)
    mag = 10.0 ** rng.uniform(-3, 3, M)
    c = (cnormal(rng, M).astype(np.complex128) * mag * scale).astype(np.complex64)
    out = plan.type1(_t(xyz, dev), _t(c, dev)).cpu().numpy().astype(np.complex128)
    ref = oracle_mod.direct_type1_3d(xyz.astype(np.float64), c.astype(np.complex128), n, d)
    assert np.all(np.isfinite(out))
    assert rel_l2(out, ref) < TOL


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("scale", [1e-30, 1.0, 1e30])
def test_type2_mode_range(dev, oracle_mod, method, scale):
    """Modes spanning six decades at overall scales 1e-30 .. 1e30: the MFMA
    interpolation's per-tile power-of-two normalisation."""
    rng = np.random.default_rng(12)
    n = (16, 12, 10)
    plan, d = _plan3(n, w=7, method=method)
    M = 2000
    xyz = uniform_points(rng, M)
    f = cnormal(rng, n).astype(np.complex128) * 10.0 ** rng.uniform(-3, 3, n) * scale
    f = f.astype(np.complex64)
    out = plan.type2(_t(xyz, dev), _t(f, dev)).cpu().numpy().astype(np.complex128)
    ref = oracle_mod.direct_type2_3d(f.astype(np.complex128), xyz.astype(np.float64), d)
    assert np.all(np.isfinite(out))
    assert rel_l2(out, ref) < TOL


def test_empty_inputs(dev):
    import torch
    plan, d = _plan3((8, 8, 8), w=4)
    out = plan.type1(torch.empty((0, 3), dtype=torch.float32, device=dev),
                     torch.empty(0, dtype=torch.complex64, device=dev))
    assert torch.count_nonzero(out).item() == 0
    f = torch.ones((8, 8, 8), dtype=torch.complex64, device=dev)
    out2 = plan.type2(torch.empty((0, 3), dtype=torch.float32, device=dev), f)
    assert out2.numel() == 0


def test_soa_and_aos_coordinates_agree(dev):
    import torch
    rng = np.random.default_rng(9)
    plan, d = _plan3((16, 16, 16), w=6)
    xyz = uniform_points(rng, 5000)
    c = _t(cnormal(rng, 5000), dev)
    a = plan.type1(_t(xyz, dev), c)
    soa = tuple(_t(np.ascontiguousarray(xyz[:, k]), dev) for k in range(3))
    b = plan.type1(soa, c)
    torch.cuda.synchronize()
    assert rel_l2(a.cpu().numpy(), b.cpu().numpy()) < 1e-6


# ---------------------------------------------------------------- 1D
def test_type1_1d_c1_config(dev, oracle_mod):
    """1D Type-1: M=4096 uniform in [-pi, pi), N=256, sigma=2, w=6 (m=512)."""
    import apollofft_amd as A
    rng = np.random.default_rng(1)
    x = rng.uniform(-np.pi, np.pi, 4096).astype(np.float32)
    c = cnormal(rng, 4096)
    plan = A.NufftPlan1D(256, TWO_PI / 256, 2, 6)
    assert plan.meta["m"][0] == 512
    out = plan.type1(_t(x, dev), _t(c, dev)).cpu().numpy()
    ref = oracle_mod.direct_type1_1d(x.astype(np.float64), c.astype(np.complex128), 256, TWO_PI / 256)
    assert rel_l2(out, ref) < TOL


@pytest.mark.parametrize("n,w", [(32, 6), (33, 4), (8, 6), (4, 6), (100, 8), (37, 6), (127, 7)])
def test_1d_type1_type2_vs_direct(dev, oracle_mod, n, w):
    import apollofft_amd as A
    rng = np.random.default_rng(n)
    dx = 0.05
    L = n * dx
    x = rng.uniform(-L, 2 * L, 700).astype(np.float32)
    c = cnormal(rng, 700)
    plan = A.NufftPlan1D(n, dx, 2, w)
    out = plan.type1(_t(x, dev), _t(c, dev)).cpu().numpy()
    ref = oracle_mod.direct_type1_1d(x.astype(np.float64), c.astype(np.complex128), n, dx)
    fast = oracle_mod.nufft1d_type1_fast(x.astype(np.float64), c.astype(np.complex128), n, dx, 2, w)
    assert rel_l2(out, ref) < max(TOL, 2 * rel_l2(fast, ref))
    f = cnormal(rng, n)
    out2 = plan.type2(_t(x, dev), _t(f, dev)).cpu().numpy()
    ref2 = oracle_mod.direct_type2_1d(f.astype(np.complex128), x.astype(np.float64), dx)
    fast2 = oracle_mod.nufft1d_type2_fast(f.astype(np.complex128), x.astype(np.float64), dx, 2, w)
    assert rel_l2(out2, ref2) < max(TOL, 2 * rel_l2(fast2, ref2))


# ---------------------------------------------------------------- gradient
@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("w", [4, 5, 6, 7, 8])
def test_grad_points_type2_vs_direct(dev, oracle_mod, method, w):
    """d/dx of L = ||T2(f,x) - b||^2 against the closed-form direct sum
    (the windowed MFMA kernel with 3 to 6 y-offset classes; w = 8 runs the
    LDS-tile gradient kernel under both methods)."""
    import torch
    rng = np.random.default_rng(3)
    n = (12, 12, 12)
    plan, d = _plan3(n, w=w, method=method)
    M = 1500
    xyz = uniform_points(rng, M)
    f = cnormal(rng, n)
    b = cnormal(rng, M)
    x64 = xyz.astype(np.float64)
    cvals = oracle_mod.direct_type2_3d(f.astype(np.complex128), x64, d)
    r = cvals - b.astype(np.complex128)
    ref = oracle_mod.direct_type2_3d_grad_x(f.astype(np.complex128), x64, d, r)
    g = (2.0 * r).astype(np.complex64)  # torch's upstream gradient of ||.||^2
    out = plan.grad_points(_t(xyz, dev), _t(g, dev), _t(f, dev))
    torch.cuda.synchronize()
    assert rel_l2(out.cpu().numpy(), ref) < TOL


@pytest.mark.parametrize("w", [2, 3])
def test_grad_points_narrow_kernels(dev, oracle_mod, w):
    """At w = 2 and 3 the kernel's own truncation error exceeds 1e-5, so the
    gradient is pinned to the derivative of the same approximation: FP64
    central differences (h = 1e-7) of the oracle's fast Type-2 restatement,
    at points more than 1e-4 cells from a cell boundary on every axis (the
    tap set changes only at cell boundaries, and inside a cell the tap
    weights are smooth in the coordinate).  The windowed MFMA kernel (3
    classes of 13 / 11 offsets) and the LDS-tile vector-ALU kernel must both
    match it, and each other, to 1e-5."""
    import torch
    rng = np.random.default_rng(5 + w)
    n = (12, 12, 12)
    pm, d = _plan3(n, w=w, method="mfma")
    pv, _ = _plan3(n, w=w, method="valu")
    m = pm.meta["m"]
    M = 1500
    xyz = uniform_points(rng, M)
    f = cnormal(rng, n)
    u = cnormal(rng, M)
    tx, tf, tu = _t(xyz, dev), _t(f, dev), _t(u, dev)
    gm = pm.grad_points(tx, tu, tf).cpu().numpy()
    gv = pv.grad_points(tx, tu, tf).cpu().numpy()
    torch.cuda.synchronize()
    assert rel_l2(gm, gv) < TOL
    x64 = xyz.astype(np.float64)
    f64 = f.astype(np.complex128)
    t = x64 * (np.array(m, dtype=np.float64) / (np.array(n) * np.array(d)))
    fr = t - np.floor(t)
    keep = np.all(np.minimum(fr, 1.0 - fr) > 1e-4, axis=1)
    assert keep.sum() > 0.99 * M
    h = 1e-7
    fd = np.empty((M, 3))
    for a in range(3):
        xp, xm = x64.copy(), x64.copy()
        xp[:, a] += h
        xm[:, a] -= h
        dt = (oracle_mod.nufft3d_type2_fast(f64, xp, d, 2, w, 0) -
              oracle_mod.nufft3d_type2_fast(f64, xm, d, 2, w, 0)) / (2 * h)
        fd[:, a] = np.real(np.conj(u.astype(np.complex128)) * dt)
    assert rel_l2(gm[keep], fd[keep]) < TOL
    assert rel_l2(gv[keep], fd[keep]) < TOL


@pytest.mark.parametrize("w", [4, 6, 7])
def test_windowed_kernels_large_subproblems(dev, oracle_mod, w):
    """Type-2 and the coordinate gradient with one dense tile of ~40,000
    points and max_subproblem = 16384: the windowed kernels list a
    subproblem's points by y-offset class in pieces of 4096, so this runs
    several pieces per subproblem and several subproblems per tile.  A
    point's sum does not depend on which points share its group or piece,
    so the results equal those of 64-point subproblems bit for bit; both
    are checked against FP64 direct sums on sampled points."""
    import torch
    rng = np.random.default_rng(91)
    n = (24, 20, 28)
    big, d = _plan3(n, w=w, method="mfma", max_sub=16384)
    small, _ = _plan3(n, w=w, method="mfma", max_sub=64)
    M = 60000
    xyz = uniform_points(rng, M)
    mc = int(0.7 * M)
    xyz[:mc] = (0.4 + 0.01 * rng.standard_normal((mc, 3))).astype(np.float32)
    f = cnormal(rng, n)
    u = cnormal(rng, M)
    tx, tf, tu = _t(xyz, dev), _t(f, dev), _t(u, dev)
    print(f"windowed w={w}: M {M}", flush=True)
    o_big = big.type2(tx, tf).cpu().numpy()
    o_small = small.type2(tx, tf).cpu().numpy()
    g_big = big.grad_points(tx, tu, tf).cpu().numpy()
    g_small = small.grad_points(tx, tu, tf).cpu().numpy()
    torch.cuda.synchronize()
    assert np.array_equal(o_big, o_small)
    assert np.array_equal(g_big, g_small)
    x64 = xyz.astype(np.float64)
    pj = np.concatenate([rng.choice(mc, 600, replace=False), mc + rng.choice(M - mc, 200, replace=False)])
    ref = oracle_mod.direct_type2_3d(f.astype(np.complex128), x64[pj], d)
    assert rel_l2(o_big[pj], ref) < TOL
    ref_x = oracle_mod.direct_type2_3d_grad_x(f.astype(np.complex128), x64[pj], d,
                                              u[pj].astype(np.complex128) / 2.0)
    assert rel_l2(g_big[pj], ref_x) < TOL


@pytest.mark.parametrize("w", [4, 7])
def test_type2_backward_matches_separate_calls(dev, oracle_mod, w):
    """The fused Type-2 backward (one binning) against the Type-1 of g and the
    coordinate gradient computed by separate calls, and against FP64 direct
    sums; an odd grid, clustered points and small subproblems (several rounds
    and subproblems per tile); M = 0; and the unsupported vector-ALU plan."""
    import torch
    import apollofft_amd._native as N
    rng = np.random.default_rng(61)
    n = (20, 13, 16)
    plan, d = _plan3(n, w=w, method="mfma", max_sub=300)
    M = 6000
    xyz = uniform_points(rng, M)
    xyz[: M // 2] = (0.3 + 0.05 * rng.standard_normal((M // 2, 3))).astype(np.float32)
    f = cnormal(rng, n)
    g = cnormal(rng, M)
    tx, tg, tf = _t(xyz, dev), _t(g, dev), _t(f, dev)
    assert plan.fused_backward
    gm, gx = plan.type2_backward(tx, tg, tf)
    gm_sep = plan.type1(tx, tg)
    gx_sep = plan.grad_points(tx, tg, tf)
    torch.cuda.synchronize()
    assert rel_l2(gm.cpu().numpy(), gm_sep.cpu().numpy()) < 1e-6
    assert rel_l2(gx.cpu().numpy(), gx_sep.cpu().numpy()) < 1e-6
    x64 = xyz.astype(np.float64)
    ref_m = oracle_mod.direct_type1_3d(x64, g.astype(np.complex128), n, d)
    assert rel_l2(gm.cpu().numpy(), ref_m) < TOL
    pj = rng.choice(M, size=300, replace=False)
    ref_x = oracle_mod.direct_type2_3d_grad_x(f.astype(np.complex128), x64[pj], d,
                                              g[pj].astype(np.complex128) / 2.0)
    assert rel_l2(gx.cpu().numpy()[pj], ref_x) < TOL
    gm0, gx0 = plan.type2_backward(tx[:0], tg[:0], tf)
    torch.cuda.synchronize()
    assert gx0.shape == (0, 3) and float(gm0.abs().max()) == 0.0
    pv, _ = _plan3(n, w=w, method="valu")
    assert not pv.fused_backward
    with pytest.raises(N.Unsupported):
        pv.type2_backward(tx, tg, tf)


@pytest.mark.parametrize("w", [4, 7])
def test_type1_backward_matches_separate_calls(dev, oracle_mod, w):
    """The fused Type-1 backward (one FFT of G, one binning) against T2(G)
    and the coordinate gradient from separate calls and FP64 direct sums;
    clustered points with small subproblems; M = 0; the vector-ALU plan."""
    import torch
    import apollofft_amd._native as N
    rng = np.random.default_rng(62)
    n = (18, 14, 11)
    plan, d = _plan3(n, w=w, method="mfma", max_sub=300)
    M = 5000
    xyz = uniform_points(rng, M)
    xyz[: M // 2] = (-1.0 + 0.05 * rng.standard_normal((M // 2, 3))).astype(np.float32)
    c = cnormal(rng, M)
    G = cnormal(rng, n)
    tx, tc, tG = _t(xyz, dev), _t(c, dev), _t(G, dev)
    assert plan.fused_backward
    gv, gx = plan.type1_backward(tx, tc, tG)
    gv_sep = plan.type2(tx, tG)
    gx_sep = plan.grad_points(tx, tc, tG)
    torch.cuda.synchronize()
    assert rel_l2(gv.cpu().numpy(), gv_sep.cpu().numpy()) < 1e-6
    assert rel_l2(gx.cpu().numpy(), gx_sep.cpu().numpy()) < 1e-6
    x64 = xyz.astype(np.float64)
    ref_v = oracle_mod.direct_type2_3d(G.astype(np.complex128), x64, d)
    assert rel_l2(gv.cpu().numpy(), ref_v) < TOL
    pj = rng.choice(M, size=300, replace=False)
    ref_x = oracle_mod.direct_type2_3d_grad_x(G.astype(np.complex128), x64[pj], d,
                                              c[pj].astype(np.complex128) / 2.0)
    assert rel_l2(gx.cpu().numpy()[pj], ref_x) < TOL
    gv0, gx0 = plan.type1_backward(tx[:0], tc[:0], tG)
    torch.cuda.synchronize()
    assert gv0.shape == (0,) and gx0.shape == (0, 3)
    pv, _ = _plan3(n, w=w, method="valu")
    with pytest.raises(N.Unsupported):
        pv.type1_backward(tx, tc, tG)


# ---------------------------------------------------------------- host drop-in API
def test_host_api_matches_direct(dev, oracle_mod):
    import apollofft_amd as A
    rng = np.random.default_rng(21)
    pos = rng.uniform(-np.pi, np.pi, (2000, 3))
    vals = rng.standard_normal(2000) + 1j * rng.standard_normal(2000)
    n = 16
    out = A.nufft_type1_3d_fast(pos, vals, n, n, n, TWO_PI / n, TWO_PI / n, TWO_PI / n)
    ref = oracle_mod.direct_type1_3d(pos.astype(np.float32).astype(np.float64), vals, (n,) * 3,
                                     (TWO_PI / n,) * 3)
    assert out.dtype == np.complex128 and out.shape == (n, n, n)
    assert rel_l2(out, ref) < TOL
    modes = rng.standard_normal((n, n, n)) + 1j * rng.standard_normal((n, n, n))
    g = A.nufft_type2_3d_fast(pos, modes, TWO_PI / n, TWO_PI / n, TWO_PI / n)
    refg = oracle_mod.direct_type2_3d(modes, pos.astype(np.float32).astype(np.float64),
                                      (TWO_PI / n,) * 3)
    assert rel_l2(g, refg) < TOL


def test_host_profile_mismatch_and_f16(dev):
    import apollofft_amd as A
    import apollofft_amd._native as N
    plan = A.NufftPlan1D(8, 0.125, 2, 6)
    pos = np.array([0.0, 0.09, 0.21, 0.47])
    v64 = np.array([1 + 0.25j, -0.5 + 0.75j, 0.25 - 0.1j, 0.125 + 0.5j])
    with pytest.raises(A.ProfileMismatch):
        plan.type1_host(pos, v64.astype(np.complex64), storage=N.STORAGE_C64,
                        profile=N.PROFILE_HIGH_ACCURACY_F64)
    expected = plan.type1_host(pos, v64)
    v16 = np.stack([v64.real, v64.imag], axis=-1).astype(np.float16)
    out16 = plan.type1_host(pos, v16, storage=N.STORAGE_F16X2)
    e16 = plan.type1_host(pos, v16[..., 0].astype(np.float64) + 1j * v16[..., 1].astype(np.float64))
    # f16 output is the f32 result quantized once (dimension_1d.rs:388-401)
    bound_re = np.abs(e16.real) * 2.0 ** -10 + 2.0 ** -14
    bound_im = np.abs(e16.imag) * 2.0 ** -10 + 2.0 ** -14
    assert np.all(np.abs(out16[:, 0].astype(np.float64) - e16.real) <= bound_re + 1e-6)
    assert np.all(np.abs(out16[:, 1].astype(np.float64) - e16.imag) <= bound_im + 1e-6)
    assert expected.shape == (8,)


# ---------------------------------------------------------------- batched
@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("mixed_w", [False, True])
def test_batched_type1_matches_single(dev, oracle_mod, mixed_w, method):
    """Batched Type-1 over uneven problems (fused persistent spreading when
    the widths agree, back-to-back otherwise) equals per-problem Type-1 and
    the FP64 direct sum."""
    import apollofft_amd as A
    import torch
    rng = np.random.default_rng(31)
    shapes = [(8, 8, 8), (12, 10, 16), (16, 16, 16), (6, 9, 5), (16, 12, 8)]
    Ms = [0, 900, 5000, 37, 2500]
    plans, pos, vals, refs = [], [], [], []
    for i, (n, M) in enumerate(zip(shapes, Ms)):
        w = 5 + (i % 2) if mixed_w else 6
        d = [TWO_PI / k for k in n]
        plans.append(A.NufftPlan3D(*n, *d, 2, w, grid_policy="exact", kernel_method=method))
        xyz = clustered_points(rng, M, sigma=0.2) if M > 100 else uniform_points(rng, M)
        c = cnormal(rng, M)
        pos.append(_t(xyz.reshape(-1, 3), dev))
        vals.append(_t(c, dev))
        refs.append(oracle_mod.direct_type1_3d(xyz.astype(np.float64).reshape(-1, 3),
                                               c.astype(np.complex128), n, d))
    outs = A.execute_type1_batched(plans, pos, vals)
    torch.cuda.synchronize()
    for p, x, v, o, r, M in zip(plans, pos, vals, outs, refs, Ms):
        single = p.type1(x, v)
        torch.cuda.synchronize()
        if M == 0:
            assert torch.count_nonzero(o).item() == 0
            continue
        assert rel_l2(o.cpu().numpy(), single.cpu().numpy()) < 1e-6
        assert rel_l2(o.cpu().numpy(), r) < TOL


def test_graph_replays_match_eager(dev, oracle_mod):
    """Type-1 then Type-2 captured in one HIP graph (torch.cuda.graph): every
    replay, with fresh inputs copied in and eager work and synchronisations
    between replays, matches the eager calls; the execute path records only
    kernels, so cleared buffers start at zero on every replay."""
    import torch
    rng = np.random.default_rng(77)
    n, w, M = (16, 12, 20), 6, 3000
    plan, d = _plan3(n, w=w, policy="exact")
    sx = torch.empty((M, 3), dtype=torch.float32, device=dev)
    sc = torch.empty(M, dtype=torch.complex64, device=dev)
    sf = torch.empty(n, dtype=torch.complex64, device=dev)

    def fill(seed):
        r = np.random.default_rng(seed)
        xyz, c, f = uniform_points(r, M), cnormal(r, M), cnormal(r, n)
        sx.copy_(_t(xyz, dev))
        sc.copy_(_t(c, dev))
        sf.copy_(_t(f, dev))
        return xyz, c, f

    fill(0)
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):  # warm-up sizes the workspace outside capture
        for _ in range(2):
            plan.type1(sx, sc)
            plan.type2(sx, sf)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        o1 = plan.type1(sx, sc)
        o2 = plan.type2(sx, sf)
    for rep in range(4):
        xyz, c, f = fill(100 + rep)
        g.replay()
        torch.cuda.synchronize()
        e1 = plan.type1(sx, sc)
        e2 = plan.type2(sx, sf)
        torch.cuda.synchronize()
        # Type-1 flushes tiles with float atomics (summation order varies);
        # Type-2 is gather-only and reproduces bit for bit
        assert rel_l2(o1.cpu().numpy(), e1.cpu().numpy()) < 1e-6, rep
        assert torch.equal(o2, e2), rep
        if rep == 0:
            x64 = xyz.astype(np.float64)
            assert rel_l2(o1.cpu().numpy(), oracle_mod.direct_type1_3d(x64, c.astype(np.complex128), n, d)) < TOL
            assert rel_l2(o2.cpu().numpy(), oracle_mod.direct_type2_3d(f.astype(np.complex128), x64, d)) < TOL


def test_batched_falls_back_to_per_problem_binning(dev, oracle_mod):
    """A batch with one plan above the one-pass tile limit (150^3 modes on a
    300^3 grid, w = 7: 28,900 tiles, LSD binning) bins per problem and still
    spreads in the shared launch: every output equals its single call and
    the FP64 direct sum on sampled modes."""
    import apollofft_amd as A
    import apollofft_amd._native as N
    import ctypes
    import torch
    rng = np.random.default_rng(41)
    shapes, Ms = [(150, 150, 150), (16, 16, 16)], [20_000, 3000]
    plans, pos, vals, host = [], [], [], []
    for n, M in zip(shapes, Ms):
        d = [TWO_PI / k for k in n]
        plans.append(A.NufftPlan3D(*n, *d, 2, 7, grid_policy="exact"))
        xyz, c = uniform_points(rng, M), cnormal(rng, M)
        pos.append(_t(xyz, dev))
        vals.append(_t(c, dev))
        host.append((xyz, c, n, d))
    nb = ctypes.c_int64()
    N.check(N.lib().anufft_debug_num_bins(plans[0].handle, ctypes.byref(nb)))
    assert nb.value > 9216
    outs = A.execute_type1_batched(plans, pos, vals)
    torch.cuda.synchronize()
    for p, x, v, o, (xyz, c, n, d) in zip(plans, pos, vals, outs, host):
        single = p.type1(x, v)
        torch.cuda.synchronize()
        assert rel_l2(o.cpu().numpy(), single.cpu().numpy()) < 1e-6
        sel = rng.choice(int(np.prod(n)), size=300, replace=False).astype(np.int64)
        ref = oracle_mod.direct_type1_3d(xyz.astype(np.float64), c.astype(np.complex128), n, d, sel)
        assert rel_l2(o.cpu().numpy().ravel()[sel], ref) < TOL
