"""NUFFT validation report in the reference's ``NufftReport`` schema.

``run_nufft_suite`` evaluates the five cases of apollo-validation's
``run_nufft_suite`` (crates/apollo-validation/src/application/suite.rs:
227-309) -- fast 1D Type-1, fast 1D Type-2, fast 3D Type-1 on an 8^3 grid,
irrational positions, positions clustered just below the period -- with the
device fast path, and compares each with the exact direct sum computed on
the device in FP64 (``anufft_direct_type1/2``).  Both sides see the same
inputs: positions rounded once to float32 and values to complex64, as the
device path receives them, so the error is that of the fast algorithm.  The
metric is the reference's ``relative_complex_error`` (suite.rs:2665-2674),
max_k |e_k - f_k| / max(|e_k|, 1), and a case passes at or below 1e-5
(``NUFFT_FAST_RELATIVE_LIMIT``, suite.rs:53).  The report has the fields of
``NufftReport`` (crates/apollo-validation/src/domain/report.rs:60-73).

``python -m apollofft_amd.validation`` prints ``{"nufft": report}`` as JSON.
"""
import json
import math

import numpy as np

from .api import DEFAULT_NUFFT_KERNEL_WIDTH, DEFAULT_NUFFT_OVERSAMPLING
from .plan import NufftPlan1D, NufftPlan3D

NUFFT_FAST_RELATIVE_LIMIT = 1.0e-5  # suite.rs:53


def relative_complex_error(exact, fast):
    """max_k |e_k - f_k| / max(|e_k|, 1) (suite.rs:2665-2674)."""
    e = np.asarray(exact, dtype=np.complex128).ravel()
    f = np.asarray(fast, dtype=np.complex128).ravel()
    if e.size == 0:
        return 0.0
    return float(np.max(np.abs(e - f) / np.maximum(np.abs(e), 1.0)))


def _suite_values(count):
    i = np.arange(count, dtype=np.float64)
    return np.cos(0.3 * i) + 1j * np.sin(0.2 * i)


def _fast_vs_exact(plan, positions, data, kind, dev):
    import torch
    pos = torch.from_numpy(np.ascontiguousarray(positions, dtype=np.float32)).to(dev)
    val = torch.from_numpy(np.ascontiguousarray(data, dtype=np.complex64)).to(dev)
    if kind == 1:
        fast, exact = plan.type1(pos, val), plan.direct_type1(pos, val)
    else:
        fast, exact = plan.type2(pos, val), plan.direct_type2(pos, val)
    return relative_complex_error(exact.cpu().numpy(), fast.cpu().numpy())


def run_nufft_suite(device=0, kernel_width=DEFAULT_NUFFT_KERNEL_WIDTH):
    """The five NUFFT suite cases on HIP device ``device``; returns a dict
    with the NufftReport fields."""
    import torch
    dev = torch.device("cuda", device)
    w, s = int(kernel_width), DEFAULT_NUFFT_OVERSAMPLING
    # UniformDomain1D::new(32, 0.05): n = 32, L = 1.6
    n, dx = 32, 0.05
    L = n * dx
    plan1 = NufftPlan1D(n, dx, s, w, device_id=device)
    values = _suite_values(20)
    positions = np.mod(np.arange(20) * 0.137, L)
    t1_1d = _fast_vs_exact(plan1, positions, values, 1, dev)
    k = np.arange(n, dtype=np.float64)
    coeffs = np.cos(0.4 * k) - 1j * np.sin(0.25 * k)
    t2_1d = _fast_vs_exact(plan1, positions, coeffs, 2, dev)
    # UniformGrid3D::new(8, 8, 8, 0.125, 0.125, 0.125): L = 1 per axis
    plan3 = NufftPlan3D(8, 8, 8, 0.125, 0.125, 0.125, s, w, device_id=device)
    i = np.arange(12, dtype=np.float64)
    points = np.stack([np.mod(0.071 * i, 1.0), np.mod(0.113 * i, 1.0), np.mod(0.157 * i, 1.0)],
                      axis=1)
    t1_3d = _fast_vs_exact(plan3, points, values[:12], 1, dev)
    irrational = np.mod(math.sqrt(2.0) * np.arange(20) * dx, L)
    irr = _fast_vs_exact(plan1, irrational, values, 1, dev)
    clustered = np.mod(L - 1.0e-6 * (np.arange(20) + 1.0), L)
    clu = _fast_vs_exact(plan1, clustered, values, 1, dev)
    errs = [t1_1d, t2_1d, t1_3d, irr, clu]
    return {
        "type1_1d_max_relative_error": t1_1d,
        "type2_1d_max_relative_error": t2_1d,
        "type1_3d_max_relative_error": t1_3d,
        "irrational_positions_max_relative_error": irr,
        "clustered_positions_max_relative_error": clu,
        "passed": bool(all(e <= NUFFT_FAST_RELATIVE_LIMIT for e in errs)),
    }


def main():
    print(json.dumps({"nufft": run_nufft_suite()}))


if __name__ == "__main__":
    main()
