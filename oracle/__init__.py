"""FP64 oracle for the apollo KB NUFFT path -- TEST INFRASTRUCTURE ONLY.

numpy wrappers over ``liboracle.so`` (built from ``nufft_oracle.c`` by the
Makefile beside it).  The C file restates apollo-nufft's algorithm
(``crates/apollo-nufft/src/infrastructure/kernel/kaiser_bessel.rs``,
``.../application/execution/transform/dimension_{1d,3d}.rs``) in FP64 and
adds FP64 direct sums.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` may import this package, and only as the
checker or the timed CPU baseline -- never as the device path.

Every wrapper names the reference lines it follows; the C file carries the
per-function citations.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

POLICY_POW2 = 0   # dimension_3d.rs:48-54 next_pow2(max(sigma n, 2w+1))
POLICY_EXACT = 1  # max(sigma n, 2w+1), e.g. 200, 254 (mixed radix)
POLICY_1D = 2     # dimension_1d.rs:131 m = sigma n


def build():
    """Compile liboracle.so with the committed Makefile."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        d, i64, vp, i = ctypes.c_double, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int
        sig = {
            "or_i0": (d, [d]),
            "or_kb_kernel": (d, [d, d, d, d]),
            "or_kb_kernel_ft": (d, [d, i64, d, d]),
            "or_axis_deconv": (None, [i64, i64, i64, d, d, vp]),
            "or_fft_signed_index": (i64, [i64, i64]),
            "or_bucket_count": (i64, [i64, i64]),
            "or_beta": (d, [i64, i64]),
            "or_oversampled_len": (i64, [i64, i64, i64, i]),
            "or_fft_inplace": (None, [vp, i64, i]),
            "or_fft3_inplace": (None, [vp, vp, i]),
            "or_nufft1d_type1_fast": (i, [i64, d, i64, i64, i64, vp, vp, vp]),
            "or_nufft1d_type2_fast": (i, [i64, d, i64, i64, vp, i64, vp, vp]),
            "or_spread3d": (i, [vp, vp, i64, i64, i, i64, vp, vp, vp]),
            "or_nufft3d_type1_fast": (i, [vp, vp, i64, i64, i, i64, vp, vp, vp]),
            "or_nufft3d_type2_fast": (i, [vp, vp, i64, i64, i, i, vp, i64, vp, vp]),
            "or_direct_type1_1d": (None, [i64, d, i64, vp, vp, vp]),
            "or_direct_type2_1d": (None, [i64, d, vp, i64, vp, vp]),
            "or_direct_type1_3d": (None, [vp, vp, i64, vp, vp, i64, vp, vp]),
            "or_direct_type2_3d": (None, [vp, vp, vp, i64, vp, vp]),
            "or_direct_type2_3d_grad_x": (None, [vp, vp, vp, i64, vp, vp, vp]),
            "or_num_threads": (i, []),
            "or_set_fast_threads": (None, [i]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i64x3(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.int64).reshape(3))


def _f64x3(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(3))


# --- KB primitives (kaiser_bessel.rs:55-164) --------------------------------
def i0(z):
    return lib().or_i0(float(z))


def kb_kernel(x, w, beta, i0_beta):
    return lib().or_kb_kernel(float(x), float(w), float(beta), float(i0_beta))


def kb_kernel_ft(xi, w, beta, i0_beta):
    return lib().or_kb_kernel_ft(float(xi), int(w), float(beta), float(i0_beta))


def axis_deconv(n, m, w, beta, i0_beta):
    out = np.empty(n, dtype=np.float64)
    lib().or_axis_deconv(n, m, w, beta, i0_beta, _p(out))
    return out


def fft_signed_index(k, n):
    return lib().or_fft_signed_index(int(k), int(n))


def bucket_count(length, w):
    return lib().or_bucket_count(int(length), int(w))


def beta(sigma, w):
    return lib().or_beta(int(sigma), int(w))


def oversampled_len(n, sigma, w, policy):
    return lib().or_oversampled_len(int(n), int(sigma), int(w), int(policy))


# --- FP64 FFT ------------------------------------------------------------------
def fft(x, sign=-1):
    a = _c128(x).copy()
    lib().or_fft_inplace(_p(a), a.size, int(sign))
    return a


def fft3(g, sign=-1):
    a = _c128(g).copy()
    m = _i64x3(a.shape)
    lib().or_fft3_inplace(_p(a), _p(m), int(sign))
    return a


# --- fast restatements -----------------------------------------------------------
def nufft1d_type1_fast(x, c, n, dx, sigma=2, w=6):
    """NufftPlan1D::type1 (dimension_1d.rs:158-207)."""
    x, c = _f64(x), _c128(c)
    out = np.empty(n, dtype=np.complex128)
    rc = lib().or_nufft1d_type1_fast(n, dx, sigma, w, x.size, _p(x), _p(c), _p(out))
    if rc:
        raise ValueError("invalid 1D plan parameters")
    return out


def nufft1d_type2_fast(f, x, dx, sigma=2, w=6):
    """NufftPlan1D::type2 (dimension_1d.rs:211-276)."""
    f, x = _c128(f), _f64(x)
    out = np.empty(x.size, dtype=np.complex128)
    rc = lib().or_nufft1d_type2_fast(f.size, dx, sigma, w, _p(f), x.size, _p(x), _p(out))
    if rc:
        raise ValueError("invalid 1D plan parameters")
    return out


def spread3d(xyz, c, n, d, sigma=2, w=6, policy=POLICY_POW2):
    """Fine grid after the T1 spreading loop (dimension_3d.rs:277-327)."""
    xyz, c = _f64(xyz).reshape(-1, 3), _c128(c)
    n3, d3 = _i64x3(n), _f64x3(d)
    m = [oversampled_len(int(n3[a]), sigma, w, policy) for a in range(3)]
    out = np.empty(m, dtype=np.complex128)
    rc = lib().or_spread3d(_p(n3), _p(d3), sigma, w, policy, c.size, _p(xyz), _p(c), _p(out))
    if rc:
        raise ValueError("invalid 3D plan parameters")
    return out


def nufft3d_type1_fast(xyz, c, n, d, sigma=2, w=6, policy=POLICY_POW2):
    """NufftPlan3D::type1 (dimension_3d.rs:236-330, 391-414)."""
    xyz, c = _f64(xyz).reshape(-1, 3), _c128(c)
    n3, d3 = _i64x3(n), _f64x3(d)
    out = np.empty(tuple(n3), dtype=np.complex128)
    rc = lib().or_nufft3d_type1_fast(_p(n3), _p(d3), sigma, w, policy, c.size, _p(xyz),
                                     _p(c), _p(out))
    if rc:
        raise ValueError("invalid 3D plan parameters")
    return out


def nufft3d_type2_fast(f, xyz, d, sigma=2, w=6, policy=POLICY_POW2, compat_scaling=False):
    """NufftPlan3D::type2 (dimension_3d.rs:473-602).  compat_scaling=True
    reproduces the reference's direct/(mx my mz) output."""
    f = _c128(f)
    xyz = _f64(xyz).reshape(-1, 3)
    n3, d3 = _i64x3(f.shape), _f64x3(d)
    out = np.empty(xyz.shape[0], dtype=np.complex128)
    rc = lib().or_nufft3d_type2_fast(_p(n3), _p(d3), sigma, w, policy, int(bool(compat_scaling)),
                                     _p(f), xyz.shape[0], _p(xyz), _p(out))
    if rc:
        raise ValueError("invalid 3D plan parameters")
    return out


# --- FP64 direct sums --------------------------------------------------------------
def direct_type1_1d(x, c, n, dx):
    """nufft_type1_1d (dimension_1d.rs:429-450)."""
    x, c = _f64(x), _c128(c)
    out = np.empty(n, dtype=np.complex128)
    lib().or_direct_type1_1d(n, dx, x.size, _p(x), _p(c), _p(out))
    return out


def direct_type2_1d(f, x, dx):
    """nufft_type2_1d (dimension_1d.rs:454-472)."""
    f, x = _c128(f), _f64(x)
    out = np.empty(x.size, dtype=np.complex128)
    lib().or_direct_type2_1d(f.size, dx, _p(f), x.size, _p(x), _p(out))
    return out


def direct_type1_3d(xyz, c, n, d, sel=None):
    """nufft_type1_3d (dimension_3d.rs:679-708) at flat mode indices ``sel``
    (all modes when None)."""
    xyz, c = _f64(xyz).reshape(-1, 3), _c128(c)
    n3, d3 = _i64x3(n), _f64x3(d)
    full = sel is None
    sel = np.arange(int(np.prod(n3)), dtype=np.int64) if full else np.ascontiguousarray(sel, dtype=np.int64)
    out = np.empty(sel.size, dtype=np.complex128)
    lib().or_direct_type1_3d(_p(n3), _p(d3), c.size, _p(xyz), _p(c), sel.size, _p(sel), _p(out))
    return out.reshape(tuple(n3)) if full else out


def direct_type2_3d(f, xyz, d):
    """nufft_type2_3d (dimension_3d.rs:712-740)."""
    f = _c128(f)
    xyz = _f64(xyz).reshape(-1, 3)
    n3, d3 = _i64x3(f.shape), _f64x3(d)
    out = np.empty(xyz.shape[0], dtype=np.complex128)
    lib().or_direct_type2_3d(_p(n3), _p(d3), _p(f), xyz.shape[0], _p(xyz), _p(out))
    return out


def direct_type2_3d_grad_x(f, xyz, d, r):
    """Closed-form dL/dx for L = ||T2(f,x) - b||^2 given residual r."""
    f = _c128(f)
    xyz = _f64(xyz).reshape(-1, 3)
    r = _c128(r)
    n3, d3 = _i64x3(f.shape), _f64x3(d)
    out = np.empty((xyz.shape[0], 3), dtype=np.float64)
    lib().or_direct_type2_3d_grad_x(_p(n3), _p(d3), _p(f), xyz.shape[0], _p(xyz), _p(r), _p(out))
    return out


def num_threads():
    return lib().or_num_threads()


def set_fast_threads(n):
    """Threads of the 3D fast restatement (1: the reference's single-thread
    path; more: the multi-thread CPU baseline column)."""
    lib().or_set_fast_threads(int(n))
