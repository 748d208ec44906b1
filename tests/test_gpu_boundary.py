"""Host boundary on the GPU: the pyapollofft-compatible functions and the
typed-storage host entry points of 3D plans.

* ``test_smoke_*`` reproduce apollo-python/tests/test_smoke.py:44-64 with the
  same positions, values, spacings and ``atol=1e-5`` (exact vs fast), which
  also runs ``nufft_type1_{1d,3d}_fast`` on the device.
* The exact functions (FP64 direct sums on the device with FP64 coordinates)
  are checked against the oracle's FP64 direct sums.
* The typed 3D paths follow apollo-nufft/src/verification/mod.rs:700-968:
  the 2x2x2 grid with the three listed points, complex128 / complex64 /
  [f16; 2] storage, the f32 bound 1e-5 (relative per component), the f16
  bound |v| 2^-10 + 2^-14 and the profile-mismatch error.
"""
import numpy as np
import pytest

from helpers import TWO_PI, cnormal, rel_l2

pytestmark = pytest.mark.gpu


def test_smoke_nufft_1d_fast_tracks_exact(dev):
    import apollofft_amd as afft
    positions = np.array([0.01, 0.09, 0.23, 0.51], dtype=np.float64)
    values = np.array([1.0 + 0.0j, 0.5 + 0.2j, -0.2 + 0.3j, 0.1 - 0.4j], dtype=np.complex128)
    exact = afft.nufft_type1_1d(positions, values, 0.05, n_out=16)
    fast = afft.nufft_type1_1d_fast(positions, values, 0.05, n_out=16)
    assert np.allclose(exact, fast, atol=1e-5)


def test_smoke_nufft_3d_fast_tracks_exact(dev):
    import apollofft_amd as afft
    positions = np.array(
        [
            [0.01, 0.02, 0.03],
            [0.12, 0.11, 0.05],
            [0.21, 0.18, 0.09],
        ],
        dtype=np.float64,
    )
    values = np.array([1.0 + 0.0j, 0.5 + 0.1j, -0.4 + 0.2j], dtype=np.complex128)
    exact = afft.nufft_type1_3d(positions, values, 4, 4, 4, 0.1, 0.1, 0.1)
    fast = afft.nufft_type1_3d_fast(positions, values, 4, 4, 4, 0.1, 0.1, 0.1)
    assert np.allclose(exact, fast, atol=1e-5)


def test_smoke_nufft_type2_1d_fast_tracks_exact(dev):
    """The Type-2 pair of the 1D smoke test (same positions and spacing)."""
    import apollofft_amd as afft
    positions = np.array([0.01, 0.09, 0.23, 0.51], dtype=np.float64)
    coeffs = np.exp(0.3j * np.arange(16)) * (1.0 + 0.1 * np.arange(16))
    exact = afft.nufft_type2_1d(coeffs, positions, 0.05)
    fast = afft.nufft_type2_1d_fast(coeffs, positions, 0.05)
    assert exact.shape == fast.shape == (4,)
    assert np.allclose(exact, fast, atol=1e-4 * np.abs(exact).max())


def test_exact_functions_match_oracle(dev, oracle_mod):
    """Device FP64 direct sums with FP64 coordinates (no float32 rounding)
    against the oracle's direct sums: 1D both types, 3D both types,
    positions outside [0, L) included."""
    import apollofft_amd as afft
    rng = np.random.default_rng(5)
    x = rng.uniform(-3.0, 5.0, 700)
    c = rng.standard_normal(700) + 1j * rng.standard_normal(700)
    out = afft.nufft_type1_1d(x, c, 0.05, n_out=37)
    assert out.dtype == np.complex128 and out.shape == (37,)
    assert rel_l2(out, oracle_mod.direct_type1_1d(x, c, 37, 0.05)) < 1e-12
    f = rng.standard_normal(37) + 1j * rng.standard_normal(37)
    assert rel_l2(afft.nufft_type2_1d(f, x, 0.05), oracle_mod.direct_type2_1d(f, x, 0.05)) < 1e-12
    xyz = rng.uniform(-4.0, 4.0, (500, 3))
    c3 = rng.standard_normal(500) + 1j * rng.standard_normal(500)
    n, d = (6, 9, 5), (0.7, 0.5, 1.1)
    o3 = afft.nufft_type1_3d(xyz, c3, *n, *d)
    assert o3.shape == n
    assert rel_l2(o3, oracle_mod.direct_type1_3d(xyz, c3, n, d)) < 1e-12
    f3 = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert rel_l2(afft.nufft_type2_3d(xyz, f3, *d), oracle_mod.direct_type2_3d(f3, xyz, d)) < 1e-12
    # argument errors of the reference binding (UniformDomain1D::new, length checks)
    with pytest.raises(ValueError):
        afft.nufft_type1_1d(x, c, -0.05)
    with pytest.raises(ValueError, match="mismatch"):
        afft.nufft_type1_3d(xyz, c3[:-1], *n, *d)


# ---------------------------------------------------------------- typed 3D storage
_POS = np.array([(0.1, 0.2, 0.3), (0.5, 0.6, 0.7), (0.8, 0.1, 0.5)], dtype=np.float64)
_VALS = np.array([1.0 + 0.0j, 0.5 + 0.5j, -0.75 + 0.25j])


def _typed_plan(compat):
    import apollofft_amd as A
    # UniformGrid3D::new(2, 2, 2, 1, 1, 1), sigma 2, w 6 (mod.rs:710-713)
    return A.NufftPlan3D(2, 2, 2, 1.0, 1.0, 1.0, 2, 6,
                         compat_reference_type2_3d_scaling=compat)


def _f16_pairs(v):
    return np.stack([np.asarray(v).real, np.asarray(v).imag], axis=-1).astype(np.float16)


def _from_f16(a):
    a = a.astype(np.float64)
    return a[..., 0] + 1j * a[..., 1]


def _check_f32(out, expected):
    """Entry-wise relative error against the complex modulus, below 1e-5.
    The reference (mod.rs:790-802) divides each component by its own size,
    which holds there because its typed paths compute in FP64 and only the
    storage is rounded; here every storage runs the FP32 device arithmetic,
    whose error scales with |v|, so a component 300x smaller than |v| would
    fail a per-component bound at an entry-wise error of 4e-8."""
    err = np.abs(out.astype(np.complex128) - expected) / np.maximum(np.abs(expected), 1e-30)
    assert err.max() < 1e-5


def _check_f16(out16, expected):
    re = out16[..., 0].astype(np.float64)
    im = out16[..., 1].astype(np.float64)
    assert np.all(np.abs(re - expected.real) <= np.abs(expected.real) * 2.0 ** -10 + 2.0 ** -14)
    assert np.all(np.abs(im - expected.imag) <= np.abs(expected.imag) * 2.0 ** -10 + 2.0 ** -14)


def test_typed_type1_3d_storage(dev, oracle_mod):
    """NufftPlan3D::type1_typed_into (dimension_3d.rs:333-389) through
    anufft_execute_type1_host, mod.rs:700-855."""
    import apollofft_amd as A
    import apollofft_amd._native as N
    plan = _typed_plan(False)
    d = (1.0, 1.0, 1.0)
    expected64 = oracle_mod.nufft3d_type1_fast(_POS, _VALS, (2, 2, 2), d, 2, 6)
    out64 = plan.type1_host(_POS, _VALS, storage=N.STORAGE_C128)
    assert out64.dtype == np.complex128 and out64.shape == (2, 2, 2)
    _check_f32(out64, expected64)  # device arithmetic is FP32 for every storage
    v32 = _VALS.astype(np.complex64)
    expected32 = oracle_mod.nufft3d_type1_fast(_POS, v32.astype(np.complex128), (2, 2, 2), d, 2, 6)
    out32 = plan.type1_host(_POS, v32, storage=N.STORAGE_C64)
    assert out32.dtype == np.complex64
    _check_f32(out32, expected32)
    v16 = _f16_pairs(_VALS)
    expected16 = oracle_mod.nufft3d_type1_fast(_POS, _from_f16(v16), (2, 2, 2), d, 2, 6)
    out16 = plan.type1_host(_POS, v16, storage=N.STORAGE_F16X2)
    assert out16.dtype == np.float16 and out16.shape == (2, 2, 2, 2)
    _check_f16(out16, expected16)
    with pytest.raises(A.ProfileMismatch, match="precision_profile"):
        plan.type1_host(_POS, v32, storage=N.STORAGE_C64, profile=N.PROFILE_HIGH_ACCURACY_F64)
    with pytest.raises(A.ProfileMismatch):
        plan.type1_host(_POS, v16, storage=N.STORAGE_F16X2, profile=N.PROFILE_LOW_PRECISION_F32)


def test_typed_type2_3d_storage(dev, oracle_mod):
    """NufftPlan3D::type2_typed_into (dimension_3d.rs:605-629) through
    anufft_execute_type2_host with the reference's 1/(mx my mz) scaling;
    input modes are the Type-1 output (mod.rs:857-968)."""
    import apollofft_amd as A
    import apollofft_amd._native as N
    plan = _typed_plan(True)
    d = (1.0, 1.0, 1.0)
    coeffs64 = oracle_mod.nufft3d_type1_fast(_POS, _VALS, (2, 2, 2), d, 2, 6)
    expected64 = oracle_mod.nufft3d_type2_fast(coeffs64, _POS, d, 2, 6, compat_scaling=True)
    out64 = plan.type2_host(_POS, coeffs64, storage=N.STORAGE_C128)
    _check_f32(out64, expected64)
    c32 = coeffs64.astype(np.complex64)
    expected32 = oracle_mod.nufft3d_type2_fast(c32.astype(np.complex128), _POS, d, 2, 6,
                                               compat_scaling=True)
    out32 = plan.type2_host(_POS, c32, storage=N.STORAGE_C64)
    assert out32.dtype == np.complex64 and out32.shape == (3,)
    _check_f32(out32, expected32)
    c16 = _f16_pairs(coeffs64)
    expected16 = oracle_mod.nufft3d_type2_fast(_from_f16(c16), _POS, d, 2, 6, compat_scaling=True)
    out16 = plan.type2_host(_POS, c16, storage=N.STORAGE_F16X2)
    assert out16.shape == (3, 2)
    _check_f16(out16, expected16)
    with pytest.raises(A.ProfileMismatch, match="precision_profile"):
        plan.type2_host(_POS, c32, storage=N.STORAGE_C64, profile=N.PROFILE_HIGH_ACCURACY_F64)


def test_typed_3d_storage_random_points(dev, oracle_mod):
    """C64 and [f16;2] storage on a 3D plan with 3000 random points against
    the FP64 direct sum: complex64 storage stays within 1e-5 relative L2,
    f16 storage within its one-time input and output quantization."""
    import apollofft_amd as A
    import apollofft_amd._native as N
    rng = np.random.default_rng(9)
    n = (12, 10, 8)
    d = tuple(TWO_PI / k for k in n)
    plan = A.NufftPlan3D(*n, *d, 2, 6)
    pos = rng.uniform(-np.pi, np.pi, (3000, 3))
    v = cnormal(rng, 3000, np.complex128)
    x32 = pos.astype(np.float32).astype(np.float64)
    ref = oracle_mod.direct_type1_3d(x32, v, n, d)
    out32 = plan.type1_host(pos, v.astype(np.complex64), storage=N.STORAGE_C64)
    assert rel_l2(out32, oracle_mod.direct_type1_3d(x32, v.astype(np.complex64), n, d)) < 1e-5
    v16 = _f16_pairs(v)
    out16 = plan.type1_host(pos, v16, storage=N.STORAGE_F16X2)
    ref16 = oracle_mod.direct_type1_3d(x32, _from_f16(v16), n, d)
    # output quantized once: per component |v| 2^-10 + 2^-14, plus the FP32
    # path's own error (1e-5 of the largest output)
    slack = 1e-5 * np.abs(ref16).max()
    re, im = out16[..., 0].astype(np.float64), out16[..., 1].astype(np.float64)
    assert np.all(np.abs(re - ref16.real) <= np.abs(ref16.real) * 2.0 ** -10 + 2.0 ** -14 + slack)
    assert np.all(np.abs(im - ref16.imag) <= np.abs(ref16.imag) * 2.0 ** -10 + 2.0 ** -14 + slack)
    assert rel_l2(_from_f16(out16), ref) < 2.0 ** -9


def test_direct_type1_select_matches_oracle(dev, oracle_mod):
    """anufft_direct_type1_select (FP64 Type-1 at chosen modes, points split
    over blocks with FP64 atomics) against the oracle's direct sum at the
    same modes; 1D and 3D, more points than one block row, no points."""
    import torch
    import apollofft_amd as A
    rng = np.random.default_rng(17)
    n = (20, 14, 9)
    d = tuple(TWO_PI / k for k in n)
    plan = A.NufftPlan3D(*n, *d, 2, 6)
    M = 150_000
    xyz = rng.uniform(-5.0, 5.0, (M, 3)).astype(np.float32)
    c = cnormal(rng, M, np.complex128)
    sel = rng.choice(int(np.prod(n)), size=700, replace=False).astype(np.int64)
    out = plan.direct_type1_select(torch.from_numpy(xyz).to(dev), torch.from_numpy(c).to(dev),
                                   torch.from_numpy(sel).to(dev)).cpu().numpy()
    ref = oracle_mod.direct_type1_3d(xyz.astype(np.float64), c, n, d, sel)
    assert rel_l2(out, ref) < 1e-12
    p1 = A.NufftPlan1D(64, TWO_PI / 64, 2, 6)
    x1 = rng.uniform(-np.pi, np.pi, 5000).astype(np.float32)
    c1 = cnormal(rng, 5000, np.complex128)
    s1 = np.array([0, 1, 32, 33, 63, 7], dtype=np.int64)
    o1 = p1.direct_type1_select(torch.from_numpy(x1).to(dev), torch.from_numpy(c1).to(dev),
                                torch.from_numpy(s1).to(dev)).cpu().numpy()
    r1 = oracle_mod.direct_type1_1d(x1.astype(np.float64), c1, 64, TWO_PI / 64)[s1]
    assert rel_l2(o1, r1) < 1e-12
    z = plan.direct_type1_select(torch.zeros((0, 3), dtype=torch.float32, device=dev),
                                 torch.zeros(0, dtype=torch.complex128, device=dev),
                                 torch.from_numpy(sel[:5]).to(dev))
    assert float(z.abs().max()) == 0.0
