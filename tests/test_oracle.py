"""Pins the FP64 oracle (CPU only).

Each check is independent of the oracle's own formulas: the reference's
fixture values and known answers (apollo-nufft kaiser_bessel.rs tests,
dimension_{1d,3d}.rs tests, verification/mod.rs, apollo-validation
suite.rs), scipy/numpy routines, numerical quadrature, brute force on tiny
inputs, and finite differences.  A dropped term, a wrong sign or index, or
a transposed operand in the oracle fails at least one of them.
"""
import numpy as np
import pytest
from scipy import integrate, special

from helpers import TWO_PI, max_rel, rel_l2

import oracle as O


# ---- I0, phi, phi_hat, signed index (kaiser_bessel.rs:166-256) --------------
def test_i0_known_values():
    assert abs(O.i0(0.0) - 1.0) < 1e-14
    assert abs(O.i0(1.0) - 1.2660658777520082) < 1e-15 * 2
    for x in (0.5, 1.0, 2.0, 3.0):
        assert abs(O.i0(x) - O.i0(-x)) < 1e-14


@pytest.mark.parametrize("z", [0.0, 0.3, 1.0, 5.0, 9.42477796076938, 18.84955592153876,
                               28.274333882308138, 32.98672286269283, 37.69911184307752])
def test_i0_matches_scipy(z):
    assert abs(O.i0(z) - special.i0(z)) <= 2e-15 * special.i0(z)


def test_kb_kernel_peak_nonneg_and_support():
    beta, w = 8.0, 4
    i0b = O.i0(beta)
    assert abs(O.kb_kernel(0.0, w, beta, i0b) - 1.0) < 1e-14
    for x in (-1.0, -0.8, -0.5, -0.2, 0.0, 0.2, 0.5, 0.8, 1.0):
        v = O.kb_kernel(x, w, beta, i0b)
        assert 0.0 <= v <= 1.0 + 1e-14
    assert O.kb_kernel(w + 0.1, w, beta, i0b) == 0.0
    assert O.kb_kernel(-w, w, beta, i0b) == 0.0


def test_kb_kernel_ft_at_zero_closed_form():
    beta, w = 8.0, 4
    i0b = O.i0(beta)
    expected = 2.0 * w * np.sinh(beta) / (beta * i0b)
    assert abs(O.kb_kernel_ft(0.0, w, beta, i0b) - expected) / expected < 1e-10


@pytest.mark.parametrize("w,xi", [(4, 0.0), (4, 0.13), (6, 0.25), (7, 0.2), (7, 0.4), (8, 0.1)])
def test_kb_kernel_ft_is_fourier_transform_of_phi(w, xi):
    """phi_hat(xi) = integral of phi(x) exp(-2 pi i xi x) over [-w, w], by quadrature."""
    beta = O.beta(2, w)
    i0b = O.i0(beta)
    f = lambda x: O.kb_kernel(x, w, beta, i0b) * np.cos(2 * np.pi * xi * x)
    val, _ = integrate.quad(f, -w, w, limit=200, epsabs=1e-14, epsrel=1e-12)
    assert abs(O.kb_kernel_ft(xi, w, beta, i0b) - val) <= 1e-9 * abs(val)


def test_signed_index_and_buckets():
    assert [O.fft_signed_index(k, 8) for k in (0, 4, 5, 7)] == [0, 4, -3, -1]
    assert O.bucket_count(128, 7) == 9
    assert O.bucket_count(0, 2) == 1


def test_beta_and_oversampled_lengths():
    assert abs(O.beta(2, 6) - 9 * np.pi) < 1e-12   # kaiser_bessel.rs doc: W=6, sigma=2 -> 9 pi
    assert O.oversampled_len(64, 2, 7, O.POLICY_POW2) == 128
    assert O.oversampled_len(100, 2, 7, O.POLICY_POW2) == 256
    assert O.oversampled_len(100, 2, 7, O.POLICY_EXACT) == 200
    assert O.oversampled_len(127, 2, 6, O.POLICY_EXACT) == 254
    assert O.oversampled_len(2, 2, 7, O.POLICY_EXACT) == 15
    assert O.oversampled_len(4, 2, 6, O.POLICY_1D) == 8


def test_axis_deconv_is_reciprocal_ft():
    beta, w = O.beta(2, 6), 6
    i0b = O.i0(beta)
    d = O.axis_deconv(10, 20, w, beta, i0b)
    for k in range(10):
        ks = k if k <= 5 else k - 10
        assert abs(d[k] * O.kb_kernel_ft(ks / 20, w, beta, i0b) - 1.0) < 1e-14


# ---- FFT --------------------------------------------------------------------
@pytest.mark.parametrize("n", [1, 2, 3, 8, 12, 15, 17, 96, 127, 128, 200, 254])
def test_fft_matches_numpy(n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert np.abs(O.fft(a, -1) - np.fft.fft(a)).max() < 1e-12 * max(1, n)
    assert np.abs(O.fft(a, +1) - np.fft.ifft(a) * n).max() < 1e-12 * max(1, n)


def test_fft3_matches_numpy():
    rng = np.random.default_rng(0)
    g = rng.standard_normal((12, 10, 6)) + 1j * rng.standard_normal((12, 10, 6))
    assert np.abs(O.fft3(g, -1) - np.fft.fftn(g)).max() < 1e-11
    assert np.abs(O.fft3(g, +1) - np.fft.ifftn(g) * g.size).max() < 1e-11


# ---- direct sums: reference fixtures and brute force --------------------------
def test_published_fixtures():
    """suite.rs:851-866 (x=0 -> all ones) and :931-950 (x=L/4 -> [1,-i,-1,i])."""
    d = np.pi / 2
    assert np.abs(O.direct_type1_1d([0.0], [1 + 0j], 4, d) - 1.0).max() < 1e-12
    out = O.direct_type1_1d([np.pi / 2], [1 + 0j], 4, d)
    assert np.abs(out - np.array([1, -1j, -1, 1j])).max() < 1e-12


def test_dc_identities():
    """dimension_1d.rs:506-530, dimension_3d.rs:778-801, verification/mod.rs:46-90,256-282."""
    out = O.direct_type1_1d([0.0, 0.125, 0.25, 0.375],
                            [1.0, 0.5 - 0.25j, -0.5 + 0.75j, 0.25 + 0.1j], 8, 0.125)
    assert abs(out[0] - (1.25 + 0.6j)) < 1e-12
    out3 = O.direct_type1_3d([(0.0, 0.0, 0.0), (0.1, 0.2, 0.3), (0.25, 0.15, 0.05)],
                             [1.0, 0.5 - 0.25j, -0.2 + 0.75j], (4, 4, 4), (0.1, 0.1, 0.1))
    assert abs(out3[0, 0, 0] - (1.3 + 0.5j)) < 1e-12
    out3b = O.direct_type1_3d([(0.1, 0.2, 0.3), (0.5, 0.6, 0.7), (0.8, 0.1, 0.5)],
                              [1.0, 0.5 + 0.5j, -0.75 + 0.25j], (2, 2, 2), (1.0, 1.0, 1.0))
    assert abs(out3b[0, 0, 0] - (0.75 + 0.75j)) < 1e-10


def test_type2_3d_origin_is_sum_of_modes():
    """dimension_3d.rs:806-823."""
    kx, ky, kz = np.meshgrid(np.arange(3), np.arange(2), np.arange(2), indexing="ij")
    modes = (0.25 + 0.1 * kx - 0.05 * ky + 0.03 * kz) + 1j * (-0.4 + 0.07 * kx + 0.11 * ky - 0.02 * kz)
    out = O.direct_type2_3d(modes, [(0.0, 0.0, 0.0)], (0.5, 0.75, 1.0))
    assert abs(out[0] - modes.sum()) < 1e-12


def test_adjoint_fixture():
    """suite.rs:2085-2110: N=2, pos [0, 0.5], c=[1,2], f=[3,4] -> LHS = RHS = 5."""
    F = O.direct_type1_1d([0.0, 0.5], [1.0, 2.0], 2, 0.5)
    g = O.direct_type2_1d([3.0, 4.0], [0.0, 0.5], 0.5)
    lhs = np.real(np.vdot(F, [3.0, 4.0]))
    rhs = np.real(np.vdot([1.0, 2.0], g))
    assert abs(lhs - 5.0) < 1e-12 and abs(rhs - 5.0) < 1e-12


def test_uniform_positions_give_dft():
    """verification/mod.rs:405-427."""
    n = 8
    x = np.arange(n) * 0.125
    v = np.sin(np.arange(n) * 0.3) + 1j * np.cos(np.arange(n) * 0.17)
    assert np.abs(O.direct_type1_1d(x, v, n, 0.125) - np.fft.fft(v)).max() < 1e-10


def test_direct_3d_matches_python_brute_force():
    rng = np.random.default_rng(3)
    n, d = (3, 4, 5), (0.7, 0.4, 0.3)
    L = np.array(n) * np.array(d)
    pts = rng.uniform(-2, 2, (6, 3))
    c = rng.standard_normal(6) + 1j * rng.standard_normal(6)
    f = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ks = [np.array([k if k <= nn // 2 else k - nn for k in range(nn)]) for nn in n]
    ref1 = np.zeros(n, complex)
    for a in range(n[0]):
        for b in range(n[1]):
            for cc in range(n[2]):
                k = np.array([ks[0][a], ks[1][b], ks[2][cc]])
                ref1[a, b, cc] = sum(c[j] * np.exp(-2j * np.pi * np.sum(k * pts[j] / L)) for j in range(6))
    assert np.abs(O.direct_type1_3d(pts, c, n, d) - ref1).max() < 1e-12
    ref2 = np.array([sum(f[a, b, cc] * np.exp(2j * np.pi * np.sum(
        np.array([ks[0][a], ks[1][b], ks[2][cc]]) * pts[j] / L))
        for a in range(n[0]) for b in range(n[1]) for cc in range(n[2])) for j in range(6)])
    assert np.abs(O.direct_type2_3d(f, pts, d) - ref2).max() < 1e-12
    sel = np.array([0, 7, 59], dtype=np.int64)
    assert np.abs(O.direct_type1_3d(pts, c, n, d, sel) - ref1.ravel()[sel]).max() < 1e-12


def test_grad_closed_form_matches_finite_differences():
    """d/dx ||T2(f,x) - b||^2 by central differences (h = 1e-6) on the direct sum."""
    rng = np.random.default_rng(4)
    n, d = (4, 5, 3), (TWO_PI / 4, TWO_PI / 5, TWO_PI / 3)
    pts = rng.uniform(-np.pi, np.pi, (5, 3))
    f = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    b = rng.standard_normal(5) + 1j * rng.standard_normal(5)
    r = O.direct_type2_3d(f, pts, d) - b
    g = O.direct_type2_3d_grad_x(f, pts, d, r)
    loss = lambda p: np.sum(np.abs(O.direct_type2_3d(f, p, d) - b) ** 2)
    h = 1e-6
    for j in range(5):
        for a in range(3):
            pp, pm = pts.copy(), pts.copy()
            pp[j, a] += h
            pm[j, a] -= h
            fd = (loss(pp) - loss(pm)) / (2 * h)
            assert abs(fd - g[j, a]) < 1e-6 * max(1.0, abs(fd))


# ---- the restated fast algorithm ---------------------------------------------------
def test_fine_grid_matches_brute_force():
    """Fine grid after spreading equals sum_j c_j phi(g - t_j) with periodic
    images, by brute force in Python (n=4 per axis -> m=16, w=6)."""
    rng = np.random.default_rng(5)
    n, d, w = (4, 4, 4), (TWO_PI / 4,) * 3, 6
    pts = rng.uniform(-np.pi, np.pi, (5, 3))
    c = rng.standard_normal(5) + 1j * rng.standard_normal(5)
    grid = O.spread3d(pts, c, n, d, 2, w, O.POLICY_POW2)
    m = grid.shape[0]
    beta = O.beta(2, w)
    i0b = O.i0(beta)
    t = np.mod(pts, TWO_PI) * m / TWO_PI
    ref = np.zeros((m, m, m), complex)
    for j in range(5):
        wts = []
        for a in range(3):
            wa = np.zeros(m)
            for gidx in range(m):
                for img in (-1, 0, 1):
                    wa[gidx] += O.kb_kernel(gidx + img * m - t[j, a], w, beta, i0b)
            wts.append(wa)
        ref += c[j] * wts[0][:, None, None] * wts[1][None, :, None] * wts[2][None, None, :]
    assert np.abs(grid - ref).max() < 1e-13


def test_suite_cases_fast_vs_exact():
    """apollo-validation run_nufft_suite (suite.rs:227-309): all <= 1e-5."""
    n, dx = 32, 0.05
    L = n * dx
    pos = np.array([(i * 0.137) % L for i in range(20)])
    vals = np.array([complex(np.cos(0.3 * i), np.sin(0.2 * i)) for i in range(20)])
    assert max_rel(O.nufft1d_type1_fast(pos, vals, n, dx), O.direct_type1_1d(pos, vals, n, dx)) <= 1e-5
    coef = np.array([complex(np.cos(0.4 * k), -np.sin(0.25 * k)) for k in range(n)])
    assert max_rel(O.nufft1d_type2_fast(coef, pos, dx), O.direct_type2_1d(coef, pos, dx)) <= 1e-5
    pts = np.array([((0.071 * i) % 1, (0.113 * i) % 1, (0.157 * i) % 1) for i in range(12)])
    e3 = O.direct_type1_3d(pts, vals[:12], (8, 8, 8), (0.125,) * 3)
    f3 = O.nufft3d_type1_fast(pts, vals[:12], (8, 8, 8), (0.125,) * 3)
    assert max_rel(f3, e3) <= 1e-5
    irr = np.array([(np.sqrt(2.0) * i * dx) % L for i in range(20)])
    assert max_rel(O.nufft1d_type1_fast(irr, vals, n, dx), O.direct_type1_1d(irr, vals, n, dx)) <= 1e-5
    clu = np.array([(L - 1e-6 * (i + 1)) % L for i in range(20)])
    assert max_rel(O.nufft1d_type1_fast(clu, vals, n, dx), O.direct_type1_1d(clu, vals, n, dx)) <= 1e-5


def test_error_by_width_and_sigma():
    """verification/mod.rs:177-208 and :347-386."""
    dom_n, dx = 16, 0.125
    pos = np.array([0.1, 0.5, 0.9, 1.4, 1.7])
    vals = np.array([1.0, -0.5 + 0.5j, 0.75 - 0.25j, -0.25 + 0.75j, 0.5 - 0.5j])
    exact = O.direct_type1_1d(pos, vals, dom_n, dx)
    for sigma, w, tol in ((2, 4, 1e-4), (2, 6, 1e-6), (4, 6, 1e-8)):
        fast = O.nufft1d_type1_fast(pos, vals, dom_n, dx, sigma, w)
        assert np.abs(fast - exact).max() <= tol
    pos8 = np.arange(8.0)
    vals8 = np.exp(1j * np.arange(8) * np.pi / 4)
    exact8 = O.direct_type1_1d(pos8, vals8, 32, 0.25)
    errs = [np.abs(O.nufft1d_type1_fast(pos8, vals8, 32, 0.25, 2, w) - exact8).max() for w in (2, 4, 6)]
    assert errs[0] > errs[1] > errs[2] and errs[2] < 1e-5


def test_type2_3d_reference_scaling_closed_form():
    """The reference 3D Type-2 returns direct / (mx my mz) (dimension_3d.rs:549-551)."""
    rng = np.random.default_rng(6)
    n, d = (6, 5, 4), (TWO_PI / 6, TWO_PI / 5, TWO_PI / 4)
    f = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    pts = rng.uniform(-np.pi, np.pi, (40, 3))
    direct = O.direct_type2_3d(f, pts, d)
    m = [O.oversampled_len(k, 2, 6, O.POLICY_POW2) for k in n]
    compat = O.nufft3d_type2_fast(f, pts, d, 2, 6, O.POLICY_POW2, compat_scaling=True)
    fixed = O.nufft3d_type2_fast(f, pts, d, 2, 6, O.POLICY_POW2)
    assert rel_l2(compat * np.prod(m), direct) < 1e-9
    assert rel_l2(fixed, direct) < 1e-9


@pytest.mark.parametrize("w,policy", [(4, 0), (6, 1), (8, 1)])
def test_fast_3d_vs_direct_random(w, policy):
    rng = np.random.default_rng(w)
    n, d = (10, 9, 8), (TWO_PI / 10, TWO_PI / 9, TWO_PI / 8)
    pts = rng.uniform(-np.pi, np.pi, (300, 3))
    c = rng.standard_normal(300) + 1j * rng.standard_normal(300)
    bound = {4: 1e-5, 6: 1e-8, 8: 1e-10}[w]
    assert rel_l2(O.nufft3d_type1_fast(pts, c, n, d, 2, w, policy),
                  O.direct_type1_3d(pts, c, n, d)) < bound
    f = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert rel_l2(O.nufft3d_type2_fast(f, pts, d, 2, w, policy), O.direct_type2_3d(f, pts, d)) < bound


def test_linearity_and_adjoint_of_fast_path():
    rng = np.random.default_rng(8)
    n, d = (8, 8, 8), (TWO_PI / 8,) * 3
    pts = rng.uniform(-np.pi, np.pi, (50, 3))
    c1 = rng.standard_normal(50) + 1j * rng.standard_normal(50)
    c2 = rng.standard_normal(50) + 1j * rng.standard_normal(50)
    a = O.nufft3d_type1_fast(pts, c1 + 2.5 * c2, n, d)
    b = O.nufft3d_type1_fast(pts, c1, n, d) + 2.5 * O.nufft3d_type1_fast(pts, c2, n, d)
    assert rel_l2(a, b) < 1e-13
    f = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    lhs = np.vdot(O.nufft3d_type1_fast(pts, c1, n, d), f)
    rhs = np.vdot(c1, O.nufft3d_type2_fast(f, pts, d))
    assert abs(lhs - rhs) < 1e-9 * abs(lhs)
