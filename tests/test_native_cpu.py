"""C-ABI checks that need no GPU: the library loads, exports every symbol the
header declares, plan metadata matches the oracle's restatement of
NufftPlan{1D,3D}::new, validation errors map to the reference's error kinds,
and host-side argument checks fire before any device work."""
import ctypes
import os
import re

import numpy as np
import pytest

import apollofft_amd as A
import apollofft_amd._native as N
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_symbols():
    with open(os.path.join(ROOT, "include", "apollo_nufft.h")) as fh:
        text = fh.read()
    return sorted(set(re.findall(r"\b(anufft_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_header_symbol():
    lib = N.lib()
    syms = _header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(N.EXPORTS)
    assert lib.anufft_version().decode().startswith("apollo-nufft-hip")


def test_rust_binding_covers_the_header():
    """The Rust extern block of the integration guide declares exactly the
    header's entry points."""
    with open(os.path.join(ROOT, "INTEGRATION.md")) as fh:
        text = fh.read()
    block = text[text.index("extern \"C\" {"):]
    block = block[:block.index("\n}\n")]
    rust = sorted(set(re.findall(r"pub fn (anufft_[a-z0-9_]+)\s*\(", block)))
    assert rust == _header_symbols()


def _d(n):
    return 2 * np.pi / n


@pytest.mark.parametrize("n,w,policy", [((64, 64, 64), 7, "reference"), ((100, 100, 100), 8, "exact"),
                                        ((100, 100, 100), 4, "reference"), ((127, 96, 48), 6, "exact"),
                                        ((5, 3, 2), 2, "reference")])
def test_plan3d_metadata_matches_oracle(n, w, policy):
    p = A.NufftPlan3D(*n, _d(n[0]), _d(n[1]), _d(n[2]), 2, w, grid_policy=policy)
    m = p.meta
    pol = O.POLICY_EXACT if policy == "exact" else O.POLICY_POW2
    assert m["m"] == tuple(O.oversampled_len(k, 2, w, pol) for k in n)
    assert m["beta"] == O.beta(2, w)
    assert m["i0_beta"] == O.i0(O.beta(2, w))
    assert m["kernel_width"] == w and m["sigma"] == 2 and m["dim"] == 3
    for a in range(3):
        ref = O.axis_deconv(n[a], m["m"][a], w, m["beta"], m["i0_beta"])
        assert np.allclose(p.deconv(a), ref, rtol=1e-15, atol=0)
    assert m["poly_max_err"] < 5e-8
    # tiles fit the gfx950 workgroup LDS and cover the grid
    for a in range(3):
        assert 1 <= m["bin_size"][a] <= m["m"][a]


def test_plan1d_metadata_matches_oracle():
    p = A.NufftPlan1D(256, _d(256), 2, 6)
    m = p.meta
    assert m["m"][0] == O.oversampled_len(256, 2, 6, O.POLICY_1D) == 512
    assert np.allclose(p.deconv(0), O.axis_deconv(256, 512, 6, m["beta"], m["i0_beta"]), rtol=1e-15)
    # sigma n without the 2w+1 floor (dimension_1d.rs:131)
    assert A.NufftPlan1D(4, 0.25, 2, 6).meta["m"][0] == 8


@pytest.mark.parametrize("w", range(2, 9))
def test_kernel_polynomial_fit_accuracy(w):
    p = A.NufftPlan3D(16, 16, 16, 0.1, 0.1, 0.1, 2, w)
    assert p.meta["poly_max_err"] < 5e-8
    assert p.meta["poly_degree"] == 10


def test_validation_errors():
    with pytest.raises(A.InvalidArgument, match="nx"):
        A.NufftPlan3D(0, 4, 4, 0.1, 0.1, 0.1)
    with pytest.raises(A.InvalidArgument, match="dy"):
        A.NufftPlan3D(4, 4, 4, 0.1, -0.1, 0.1)
    with pytest.raises(A.InvalidArgument, match="dz"):
        A.NufftPlan3D(4, 4, 4, 0.1, 0.1, float("nan"))
    with pytest.raises(A.InvalidArgument, match="sigma"):
        A.NufftPlan3D(4, 4, 4, 0.1, 0.1, 0.1, 1, 6)
    with pytest.raises(A.InvalidArgument, match="kernel_width"):
        A.NufftPlan1D(8, 0.1, 2, 1)
    with pytest.raises(A.InvalidArgument, match="field n"):
        A.NufftPlan1D(0, 0.1)
    with pytest.raises(A.InvalidArgument, match="dx"):
        A.NufftPlan1D(8, float("inf"))
    with pytest.raises(A.Unsupported):
        A.NufftPlan3D(4, 4, 4, 0.1, 0.1, 0.1, 2, 9)
    with pytest.raises(A.Unsupported):
        A.NufftPlan1D(4096, 0.1, 2, 6)  # m = 8192 exceeds the in-LDS FFT line
    with pytest.raises(A.InvalidArgument):
        A.NufftPlan3D(64, 64, 64, 0.1, 0.1, 0.1, 2, 7, bin_size=32)  # LDS overflow


def test_workspace_grows_with_points():
    p = A.NufftPlan3D(32, 32, 32, _d(32), _d(32), _d(32), 2, 6)
    sizes = [p.workspace_size(M) for M in (0, 1, 1000, 10 ** 6)]
    assert sizes == sorted(sizes) and sizes[0] > 0
    # the fine grid alone is 64^3 complex64
    assert sizes[0] >= 64 ** 3 * 8


def _has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except ImportError:
        return False


@pytest.mark.skipif(_has_gpu(), reason="checks the device-less error path")
def test_execute_without_device_reports_unavailable():
    p = A.NufftPlan1D(8, 0.125)
    with pytest.raises(A.DeviceUnavailable):
        p.type1_host(np.zeros(3), np.zeros(3, complex))
    lib = N.lib()
    st = lib.anufft_execute_type1(p.handle, 0, N.ptr_array([None] * 3), 1, None, None, None, 0, None)
    assert st in (N.INVALID_ARG, N.DEVICE_UNAVAILABLE)


def test_c_abi_null_and_capacity_errors():
    lib = N.lib()
    assert lib.anufft_plan_get_meta(None, None) == N.INVALID_ARG
    assert "NULL" in lib.anufft_last_error().decode()
    h = ctypes.c_void_p()
    o = N.default_opts()
    assert lib.anufft_plan3d_create(4, 4, 4, 0.1, 0.1, 0.1, 2, 6, ctypes.byref(o), None) == N.INVALID_ARG
    assert lib.anufft_plan3d_create(4, 4, 4, 0.1, 0.1, 0.1, 2, 6, ctypes.byref(o), ctypes.byref(h)) == N.OK
    m = N.Meta()
    assert lib.anufft_plan_get_meta(h, ctypes.byref(m)) == N.OK
    out = np.empty(4)
    assert lib.anufft_plan_get_deconv(h, 3, out.ctypes.data) == N.INVALID_ARG
    b = ctypes.c_size_t()
    assert lib.anufft_workspace_size(h, -1, ctypes.byref(b)) == N.INVALID_ARG
    assert lib.anufft_plan_destroy(h) == N.OK


def test_python_api_argument_checks():
    pos = np.zeros((10, 3))
    with pytest.raises(ValueError, match="C-contiguous"):
        A.nufft_type1_3d_fast(pos[::2], np.zeros(5, complex), 4, 4, 4, 0.1, 0.1, 0.1)
    with pytest.raises(ValueError, match="shape"):
        A.nufft_type1_3d_fast(np.zeros((10, 2)), np.zeros(10, complex), 4, 4, 4, 0.1, 0.1, 0.1)
    with pytest.raises(ValueError, match="mismatch"):
        A.nufft_type1_3d_fast(pos, np.zeros(9, complex), 4, 4, 4, 0.1, 0.1, 0.1)
    with pytest.raises(ValueError, match="mismatch"):
        A.nufft_type1_1d_fast(np.zeros(4), np.zeros(3, complex), 0.1)


def test_direct_sum_argument_checks():
    """anufft_direct_type1/2 reject bad arguments before touching a device."""
    lib = N.lib()
    nul = N.ptr_array([None] * 3)
    assert lib.anufft_direct_type1(None, 0, nul, 1, None, None, None) == N.INVALID_ARG
    p = A.NufftPlan3D(4, 4, 4, 0.1, 0.1, 0.1)
    assert lib.anufft_direct_type1(p.handle, -1, nul, 1, None, None, None) == N.INVALID_ARG
    assert lib.anufft_direct_type2(p.handle, 3, nul, 1, None, None, None) == N.INVALID_ARG
    buf = np.zeros(64, np.float32)
    ptrs = N.ptr_array([buf.ctypes.data] * 3)
    assert lib.anufft_direct_type1(p.handle, 3, ptrs, 0, buf.ctypes.data, buf.ctypes.data,
                                   None) == N.INVALID_ARG  # stride 0
    assert lib.anufft_direct_type2(p.handle, 3, ptrs, 3, None, buf.ctypes.data,
                                   None) == N.INVALID_ARG  # no modes
    assert lib.anufft_direct_type1(p.handle, 0, nul, 1, None, None, None) == N.INVALID_ARG  # no out


def test_validation_metric_is_the_reference_one():
    """relative_complex_error = max |e - f| / max(|e|, 1) (suite.rs:2665-2674)."""
    from apollofft_amd.validation import relative_complex_error
    e = np.array([2.0, 0.5j, 0.0])
    f = np.array([2.3, 0.5j + 0.1, 0.0])
    assert relative_complex_error(e, f) == pytest.approx(0.15)
    assert relative_complex_error(e, e) == 0.0
    assert relative_complex_error([], []) == 0.0


def test_point_limit_and_stride_errors():
    """More than 2^31 - 1 points is ANUFFT_U32_OVERFLOW and a coordinate
    stride below 1 is ANUFFT_INVALID_ARG, for every entry point that takes
    points, before any device work (no GPU needed)."""
    lib = N.lib()
    p = A.NufftPlan3D(8, 8, 8, 0.1, 0.1, 0.1)
    buf = np.zeros(64, np.float32)
    ptrs = N.ptr_array([buf.ctypes.data] * 3)
    b = buf.ctypes.data
    big = 2 ** 31
    assert lib.anufft_execute_type1(p.handle, big, ptrs, 1, b, b, b, 1 << 40, None) == N.U32_OVERFLOW
    assert lib.anufft_execute_type2(p.handle, big, ptrs, 1, b, b, b, 1 << 40, None) == N.U32_OVERFLOW
    assert lib.anufft_execute_grad_points(p.handle, big, ptrs, 1, b, b, N.ptr_array([b] * 3), b,
                                          1 << 40, None) == N.U32_OVERFLOW
    assert lib.anufft_direct_type1(p.handle, big, ptrs, 1, b, b, None) == N.U32_OVERFLOW
    for stride in (0, -3):
        assert lib.anufft_execute_type1(p.handle, 3, ptrs, stride, b, b, b, 1 << 40, None) == N.INVALID_ARG
        assert "stride" in lib.anufft_last_error().decode()
        assert lib.anufft_execute_type2(p.handle, 3, ptrs, stride, b, b, b, 1 << 40, None) == N.INVALID_ARG
        assert lib.anufft_execute_type2_backward(p.handle, 3, ptrs, stride, b, b, b,
                                                 N.ptr_array([b] * 3), b, 1 << 40, None) == N.INVALID_ARG
        assert lib.anufft_execute_type1_backward(p.handle, 3, ptrs, stride, b, b, b,
                                                 N.ptr_array([b] * 3), b, 1 << 40, None) == N.INVALID_ARG
        dptrs = N.ptr_array([np.zeros(8).ctypes.data] * 3)
        assert lib.anufft_execute_type1_host(p.handle, 0, 0, 3, dptrs, stride, b, b) == N.INVALID_ARG
        assert lib.anufft_direct_type1_host(p.handle, 3, dptrs, stride, b, b) == N.INVALID_ARG
        hs = N.ptr_array([p.handle.value])
        Ms = (ctypes.c_int64 * 1)(3)
        assert lib.anufft_execute_type1_batched(hs, 1, Ms, ptrs, stride, N.ptr_array([b]),
                                                N.ptr_array([b]), b, 1 << 40, None) == N.INVALID_ARG
    hs = N.ptr_array([p.handle.value])
    Ms = (ctypes.c_int64 * 1)(big)
    assert lib.anufft_execute_type1_batched(hs, 1, Ms, ptrs, 3, N.ptr_array([b]), N.ptr_array([b]),
                                            b, 1 << 62, None) == N.U32_OVERFLOW


def test_python_wrappers_reject_bad_tensors():
    """Tensor checks of the torch wrappers (dtype, contiguity, shape, device)
    run before any pointer reaches the library."""
    torch = pytest.importorskip("torch")
    p = A.NufftPlan3D(4, 4, 4, 0.1, 0.1, 0.1)
    pos = torch.zeros((5, 3), dtype=torch.float32)
    with pytest.raises(ValueError, match="GPU"):
        p.type1(pos, torch.zeros(5, dtype=torch.complex64))
    from apollofft_amd.plan import _check_tensor
    dev = torch.device("cpu")
    with pytest.raises(ValueError, match="complex64"):
        _check_tensor(torch.zeros(5), "values", torch.complex64, dev, numel=5)
    with pytest.raises(ValueError, match="contiguous"):
        _check_tensor(torch.zeros(10, dtype=torch.complex64)[::2], "values", torch.complex64, dev)
    with pytest.raises(ValueError, match="elements"):
        _check_tensor(torch.zeros(4, dtype=torch.complex64), "values", torch.complex64, dev, numel=5)
    with pytest.raises(ValueError, match="shape"):
        _check_tensor(torch.zeros((4, 4), dtype=torch.complex64), "out", torch.complex64, dev,
                      shape=(4, 4, 4))
    with pytest.raises(ValueError, match="is on"):
        _check_tensor(torch.zeros(5, dtype=torch.complex64), "out", torch.complex64,
                      torch.device("cuda:0"), numel=5)


def test_noalloc_entry_point_errors():
    """anufft_execute_type1_noalloc rejects a NULL plan, bad point counts and
    strides, and calls that need more than its fixed scratch (capacity
    error naming the scratch), all before any device work (no GPU needed)."""
    import ctypes
    lib = N.lib()
    cap = ctypes.c_size_t()
    assert lib.anufft_scratch_capacity(ctypes.byref(cap)) == N.OK
    assert cap.value == 64 << 20
    assert lib.anufft_scratch_capacity(None) == N.INVALID_ARG
    p = A.NufftPlan3D(64, 64, 64, 0.1, 0.1, 0.1, 2, 7)
    buf = np.zeros(64, np.float32)
    b = buf.ctypes.data
    ptrs = N.ptr_array([b] * 3)
    assert lib.anufft_execute_type1_noalloc(None, 3, ptrs, 1, b, b, None) == N.INVALID_ARG
    assert lib.anufft_execute_type1_noalloc(p.handle, 2 ** 31, ptrs, 1, b, b, None) == N.U32_OVERFLOW
    assert lib.anufft_execute_type1_noalloc(p.handle, 3, ptrs, 0, b, b, None) == N.INVALID_ARG
    need = ctypes.c_size_t()
    assert lib.anufft_workspace_size(p.handle, 600_000, ctypes.byref(need)) == N.OK
    assert need.value > cap.value
    assert lib.anufft_execute_type1_noalloc(p.handle, 600_000, ptrs, 1, b, b, None) == N.CAPACITY_EXCEEDED
    assert "scratch" in lib.anufft_last_error().decode()


def test_fp32_cell_division_is_exact():
    """The binning divides cells by tile sizes as floor((cell + 1/2) *
    fp32(1/bin)) in FP32 (no integer division); for every cell and tile
    size below 2^13 (grids are at most 4096 cells per axis) that equals the
    integer quotient."""
    cells = np.arange(8192, dtype=np.int64)
    cf = cells.astype(np.float32) + np.float32(0.5)
    for b in range(1, 8192):
        q = (cf * (np.float32(1.0) / np.float32(b))).astype(np.int64)
        assert np.array_equal(q, cells // b), b


def test_launch_settings_and_autotune_argument_checks():
    """Launch settings are validated at plan creation and reported by the
    metadata; anufft_autotune3d rejects bad arguments before device work."""
    import ctypes
    p = A.NufftPlan3D(16, 16, 16, 0.1, 0.1, 0.1, 2, 6, sub_work=16, blocks_per_cu=3, fft_lines=4)
    m = p.meta
    assert (m["sub_work"], m["blocks_per_cu"], m["fft_lines"]) == (16, 3, 4)
    assert A.NufftPlan3D(16, 16, 16, 0.1, 0.1, 0.1).meta["sub_work"] == 48  # automatic
    for kw in ({"sub_work": -1}, {"blocks_per_cu": 17}, {"fft_lines": 5000}):
        with pytest.raises(A.InvalidArgument, match="launch settings"):
            A.NufftPlan3D(16, 16, 16, 0.1, 0.1, 0.1, **kw)
    lib = N.lib()
    o = N.default_opts()
    out = N.Opts()
    ms = ctypes.c_double()
    args = (16, 16, 16, 0.1, 0.1, 0.1, 2, 6, ctypes.byref(o))
    assert lib.anufft_autotune3d(*args, 1000, None, 0, 3, None, None, None) == N.INVALID_ARG
    assert lib.anufft_autotune3d(*args, 0, None, 0, 3, None, ctypes.byref(out), None) == N.INVALID_ARG
    assert lib.anufft_autotune3d(*args, 1000, None, 0, 4, None, ctypes.byref(out), None) == N.INVALID_ARG
    assert lib.anufft_autotune3d(0, 16, 16, 0.1, 0.1, 0.1, 2, 6, ctypes.byref(o), 1000, None, 0, 3,
                                 None, ctypes.byref(out), None) == N.INVALID_ARG
    assert "nx" in lib.anufft_last_error().decode()
    bad = N.ptr_array([None] * 3)
    assert lib.anufft_autotune3d(*args, 1000, bad, 3, 3, None, ctypes.byref(out),
                                 ctypes.byref(ms)) == N.INVALID_ARG
