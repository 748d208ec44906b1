"""anufft_execute_type1_noalloc (Type-1 on the library's fixed per-GPU
scratch, no workspace argument, no allocation) against anufft_execute_type1
with a caller workspace, through the C ABI.

Two calls in a row on one plan (64^3 modes -> 128^3 grid, w = 7), the
second with fewer points, each against the workspace entry point (same
kernels; float atomics may add in another order: relative L2 <= 1e-6) and
against the FP64 direct sum on 1024 random modes (relative L2 <= 1e-5);
then a 1D plan, and a point count beyond the scratch capacity, which must
return ANUFFT_CAPACITY_EXCEEDED.
"""
import ctypes

import numpy as np
import pytest

from helpers import TWO_PI, cnormal, rel_l2, uniform_points

pytestmark = pytest.mark.gpu


def _t(x, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def _type1_both(plan, xyz, c, dev, dim):
    """(no-alloc output, workspace output) of the same Type-1 call."""
    import torch
    import apollofft_amd._native as N
    L = N.lib()
    M = xyz.shape[0]
    tx = _t(xyz, dev)
    tc = _t(c, dev)
    b = tx.data_ptr()
    coords = N.ptr_array([b, b + 4, b + 8] if dim == 3 else [b, b, b])
    stride = 3 if dim == 3 else 1
    shape = plan.meta["n"][:dim]
    out_na = torch.full(shape, float("nan"), dtype=torch.complex64, device=dev)
    out_ws = torch.empty(shape, dtype=torch.complex64, device=dev)
    nbytes = ctypes.c_size_t()
    N.check(L.anufft_workspace_size(plan.handle, M, ctypes.byref(nbytes)))
    ws = torch.empty(nbytes.value, dtype=torch.uint8, device=dev)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    print(f"noalloc: dim {dim} M {M}", flush=True)
    N.check(L.anufft_execute_type1_noalloc(plan.handle, M, coords, stride, tc.data_ptr(),
                                           out_na.data_ptr(), stream), "noalloc")
    N.check(L.anufft_execute_type1(plan.handle, M, coords, stride, tc.data_ptr(),
                                   out_ws.data_ptr(), ws.data_ptr(), nbytes.value, stream),
            "workspace")
    torch.cuda.synchronize()
    return out_na.cpu().numpy(), out_ws.cpu().numpy()


def test_noalloc_type1_matches_workspace_call(dev, oracle_mod):
    import apollofft_amd as A
    rng = np.random.default_rng(71)
    n = (64, 64, 64)
    d = (TWO_PI / 64,) * 3
    plan = A.NufftPlan3D(*n, *d, 2, 7)
    assert plan.meta["m"] == (128, 128, 128)
    for M in (400_000, 150_000):  # second call smaller than the first
        xyz = uniform_points(rng, M)
        c = cnormal(rng, M)
        na, ws = _type1_both(plan, xyz, c, dev, 3)
        assert np.all(np.isfinite(na))
        assert rel_l2(na, ws) < 1e-6, M
        sel = rng.choice(int(np.prod(n)), size=1024, replace=False).astype(np.int64)
        ref = oracle_mod.direct_type1_3d(xyz.astype(np.float64), c.astype(np.complex128), n, d, sel)
        assert rel_l2(na.ravel()[sel], ref) < 1e-5, M


def test_noalloc_type1_1d_and_capacity(dev, oracle_mod):
    import apollofft_amd as A
    import apollofft_amd._native as N
    rng = np.random.default_rng(72)
    p1 = A.NufftPlan1D(256, TWO_PI / 256, 2, 6)
    for M in (20_000, 4096):
        x = uniform_points(rng, M, dim=1)
        c = cnormal(rng, M)
        na, ws = _type1_both(p1, x, c, dev, 1)
        assert rel_l2(na, ws) < 1e-6, M
        ref = oracle_mod.direct_type1_1d(x.astype(np.float64), c.astype(np.complex128), 256,
                                         TWO_PI / 256)
        assert rel_l2(na, ref) < 1e-5, M
    # beyond the fixed scratch: a clear capacity error, nothing launched
    L = N.lib()
    cap = ctypes.c_size_t()
    N.check(L.anufft_scratch_capacity(ctypes.byref(cap)))
    p3 = A.NufftPlan3D(64, 64, 64, *(TWO_PI / 64,) * 3, 2, 7)
    big = 2_000_000
    need = ctypes.c_size_t()
    N.check(L.anufft_workspace_size(p3.handle, big, ctypes.byref(need)))
    assert need.value > cap.value
    buf = _t(np.zeros(16, dtype=np.float32), dev).data_ptr()
    st = L.anufft_execute_type1_noalloc(p3.handle, big, N.ptr_array([buf, buf, buf]), 1, buf, buf,
                                        None)
    assert st == N.CAPACITY_EXCEEDED
    assert "scratch" in L.anufft_last_error().decode()
