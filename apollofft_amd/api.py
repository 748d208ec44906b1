"""pyapollofft-compatible NUFFT functions running on the GPU.

Same names, positional arguments and keywords as the reference's Python
binding (crates/apollo-python/src/lib.rs:754-951), numpy in / numpy out
(float64 positions, complex128 values), plus ``device=`` (HIP device id) and
the Type-2 3D functions the reference binding lacks.  The exact functions
(``nufft_type1_1d``, ``nufft_type2_1d``, ``nufft_type1_3d``) are FP64 direct
sums on the device with FP64 coordinates; the ``_fast`` functions run the
KB path in FP32/complex64 through the C ABI host entry points, positions
rounded once to float32.  Plans are cached per (shape, spacing, width,
device) instead of being rebuilt per call as in the reference wrappers
(dimension_3d.rs:744-766).
"""
import functools

import numpy as np

from . import _native as N
from .plan import NufftPlan1D, NufftPlan3D

DEFAULT_NUFFT_KERNEL_WIDTH = 6   # apollo-nufft/src/lib.rs:27
DEFAULT_NUFFT_OVERSAMPLING = 2   # apollo-nufft/src/lib.rs:30


def _require_contiguous(a, what):
    # require_contiguous_1d / _2d of the reference binding
    if not isinstance(a, np.ndarray) or not a.flags.c_contiguous:
        raise ValueError(f"{what} must be a C-contiguous numpy array")


@functools.lru_cache(maxsize=64)
def _plan1d(n, dx, w, device):
    return NufftPlan1D(n, dx, DEFAULT_NUFFT_OVERSAMPLING, w, device_id=device)


@functools.lru_cache(maxsize=64)
def _plan3d(nx, ny, nz, dx, dy, dz, w, device):
    return NufftPlan3D(nx, ny, nz, dx, dy, dz, DEFAULT_NUFFT_OVERSAMPLING, w, device_id=device)


def nufft_type1_1d(positions, values, dx, n_out=None, device=None):
    """apollo-python lib.rs:754-780 (nufft_type1_1d): the exact direct sum
    (dimension_1d.rs:429-450), FP64 on the GPU."""
    _require_contiguous(positions, "nufft_type1_1d positions")
    _require_contiguous(values, "nufft_type1_1d values")
    if positions.ndim != 1 or values.ndim != 1:
        raise ValueError("nufft_type1_1d takes 1-D arrays")
    if positions.size != values.size:
        raise ValueError("nufft_type1_1d positions/value length mismatch")
    n = values.size if n_out is None else int(n_out)
    plan = _plan1d(n, float(dx), DEFAULT_NUFFT_KERNEL_WIDTH, device)
    return plan.direct_type1_host(positions.astype(np.float64, copy=False),
                                  values.astype(np.complex128, copy=False))


def nufft_type2_1d(fourier_coeffs, positions, dx, device=None):
    """apollo-python lib.rs:782-804 (nufft_type2_1d): the exact direct sum
    (dimension_1d.rs:452-472), FP64 on the GPU."""
    _require_contiguous(fourier_coeffs, "nufft_type2_1d fourier_coeffs")
    _require_contiguous(positions, "nufft_type2_1d positions")
    if fourier_coeffs.ndim != 1 or positions.ndim != 1:
        raise ValueError("nufft_type2_1d takes 1-D arrays")
    plan = _plan1d(fourier_coeffs.size, float(dx), DEFAULT_NUFFT_KERNEL_WIDTH, device)
    return plan.direct_type2_host(positions.astype(np.float64, copy=False),
                                  fourier_coeffs.astype(np.complex128, copy=False))


def nufft_type1_3d(positions, values, nx, ny, nz, dx, dy, dz, device=None):
    """apollo-python lib.rs:806-836 (nufft_type1_3d): the exact direct sum
    (dimension_3d.rs:679-708), FP64 on the GPU."""
    _require_contiguous(positions, "nufft_type1_3d positions")
    _require_contiguous(values, "nufft_type1_3d values")
    if positions.ndim != 2 or positions.shape[1] != 3:
        raise ValueError("nufft_type1_3d positions must have shape (n_samples, 3)")
    if positions.shape[0] != values.size:
        raise ValueError("nufft_type1_3d positions/value length mismatch")
    plan = _plan3d(int(nx), int(ny), int(nz), float(dx), float(dy), float(dz),
                   DEFAULT_NUFFT_KERNEL_WIDTH, device)
    return plan.direct_type1_host(positions.astype(np.float64, copy=False),
                                  values.astype(np.complex128, copy=False))


def nufft_type2_3d(positions, modes, dx, dy, dz, device=None):
    """Exact Type-2 3D sum (nufft_type2_3d, dimension_3d.rs:710-740), which
    the reference binding does not expose; FP64 on the GPU."""
    _require_contiguous(positions, "nufft_type2_3d positions")
    _require_contiguous(modes, "nufft_type2_3d modes")
    if positions.ndim != 2 or positions.shape[1] != 3:
        raise ValueError("nufft_type2_3d positions must have shape (n_samples, 3)")
    if modes.ndim != 3:
        raise ValueError("nufft_type2_3d modes must be a 3-D array")
    nx, ny, nz = modes.shape
    plan = _plan3d(nx, ny, nz, float(dx), float(dy), float(dz), DEFAULT_NUFFT_KERNEL_WIDTH, device)
    return plan.direct_type2_host(positions.astype(np.float64, copy=False),
                                  modes.astype(np.complex128, copy=False))


def nufft_type1_1d_fast(positions, values, dx, n_out=None, kernel_width=DEFAULT_NUFFT_KERNEL_WIDTH,
                        device=None):
    """apollo-python lib.rs:839-867 (nufft_type1_1d_fast) on the GPU."""
    _require_contiguous(positions, "nufft_type1_1d_fast positions")
    _require_contiguous(values, "nufft_type1_1d_fast values")
    if positions.ndim != 1 or values.ndim != 1:
        raise ValueError("nufft_type1_1d_fast takes 1-D arrays")
    if positions.size != values.size:
        raise ValueError("nufft_type1_1d_fast positions/value length mismatch")
    n = values.size if n_out is None else int(n_out)
    plan = _plan1d(n, float(dx), int(kernel_width), device)
    return plan.type1_host(positions.astype(np.float64, copy=False),
                           values.astype(np.complex128, copy=False))


def nufft_type2_1d_fast(fourier_coeffs, positions, dx, kernel_width=DEFAULT_NUFFT_KERNEL_WIDTH,
                        device=None):
    """apollo-python lib.rs:869-898 (nufft_type2_1d_fast) on the GPU."""
    _require_contiguous(fourier_coeffs, "nufft_type2_1d_fast fourier_coeffs")
    _require_contiguous(positions, "nufft_type2_1d_fast positions")
    plan = _plan1d(fourier_coeffs.size, float(dx), int(kernel_width), device)
    return plan.type2_host(positions.astype(np.float64, copy=False),
                           fourier_coeffs.astype(np.complex128, copy=False))


def nufft_type1_3d_fast(positions, values, nx, ny, nz, dx, dy, dz,
                        kernel_width=DEFAULT_NUFFT_KERNEL_WIDTH, device=None):
    """apollo-python lib.rs:900-951 (nufft_type1_3d_fast) on the GPU."""
    _require_contiguous(positions, "nufft_type1_3d_fast positions")
    _require_contiguous(values, "nufft_type1_3d_fast values")
    if positions.ndim != 2 or positions.shape[1] != 3:
        raise ValueError("nufft_type1_3d_fast positions must have shape (n_samples, 3)")
    if positions.shape[0] != values.size:
        raise ValueError("nufft_type1_3d_fast positions/value length mismatch")
    plan = _plan3d(int(nx), int(ny), int(nz), float(dx), float(dy), float(dz),
                   int(kernel_width), device)
    return plan.type1_host(positions.astype(np.float64, copy=False),
                           values.astype(np.complex128, copy=False))


def nufft_type2_3d_fast(positions, modes, dx, dy, dz, kernel_width=DEFAULT_NUFFT_KERNEL_WIDTH,
                        device=None):
    """Type-2 3D (NufftPlan3D::type2, dimension_3d.rs:473-602), which the
    reference binding does not expose.  Scaled to match the direct sum."""
    _require_contiguous(positions, "nufft_type2_3d_fast positions")
    _require_contiguous(modes, "nufft_type2_3d_fast modes")
    if positions.ndim != 2 or positions.shape[1] != 3:
        raise ValueError("nufft_type2_3d_fast positions must have shape (n_samples, 3)")
    if modes.ndim != 3:
        raise ValueError("nufft_type2_3d_fast modes must be a 3-D array")
    nx, ny, nz = modes.shape
    plan = _plan3d(nx, ny, nz, float(dx), float(dy), float(dz), int(kernel_width), device)
    return plan.type2_host(positions.astype(np.float64, copy=False),
                           modes.astype(np.complex128, copy=False))


__all__ = [
    "DEFAULT_NUFFT_KERNEL_WIDTH", "DEFAULT_NUFFT_OVERSAMPLING",
    "nufft_type1_1d", "nufft_type2_1d", "nufft_type1_3d", "nufft_type2_3d",
    "nufft_type1_1d_fast", "nufft_type2_1d_fast", "nufft_type1_3d_fast", "nufft_type2_3d_fast",
    "N",
]
