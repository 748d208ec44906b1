"""Writes the golden fixtures of tests/golden/ (run from the repo root:
`python tests/golden/make_golden.py`).

Inputs are seeded; coordinates are drawn in float64 and rounded to float32
(stored as float64 that float32 represents exactly), so the FP64 direct sums
stored here are exact references for the device path, which reads float32
coordinates.  Outputs come from the FP64 oracle (oracle/), whose functions
are pinned by tests/test_oracle.py; I0 values are cross-checked against
scipy.special.i0 before writing.
"""
import json
import os
import sys

import numpy as np
from scipy import special

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle as O  # noqa: E402

TWO_PI = 2.0 * np.pi


def f32(a):
    return np.asarray(a, dtype=np.float64).astype(np.float32).astype(np.float64)


def cn(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def kb_meta():
    out = {"sigma": 2, "widths": {}}
    for w in range(2, 9):
        beta = O.beta(2, w)
        i0b = O.i0(beta)
        assert abs(i0b - special.i0(beta)) <= 2e-15 * i0b
        out["widths"][str(w)] = {"beta": beta, "i0_beta": i0b, "i0_beta_scipy": float(special.i0(beta))}
    out["deconv_w6"] = {}
    beta, i0b = O.beta(2, 6), O.i0(O.beta(2, 6))
    for n, m in ((8, 16), (64, 128), (100, 200), (127, 254), (256, 512)):
        out["deconv_w6"][f"{n}_{m}"] = O.axis_deconv(n, m, 6, beta, i0b).tolist()
    with open(os.path.join(HERE, "kb_meta.json"), "w") as fh:
        json.dump(out, fh, indent=1)


def main():
    kb_meta()
    rng = np.random.default_rng(1)
    # 1D Type-1, M=4096 uniform in [-pi, pi), N=256, dx=2 pi/256
    x = f32(rng.uniform(-np.pi, np.pi, 4096))
    c = cn(rng, 4096)
    np.savez(os.path.join(HERE, "type1_1d_n256_m4096.npz"), x=x, c=c, n=256, dx=TWO_PI / 256,
             direct=O.direct_type1_1d(x, c, 256, TWO_PI / 256))
    # 3D Type-1, 16^3 modes, M=2000
    rng = np.random.default_rng(2)
    n = (16, 16, 16)
    d = (TWO_PI / 16,) * 3
    xyz = f32(rng.uniform(-np.pi, np.pi, (2000, 3)))
    c = cn(rng, 2000)
    np.savez(os.path.join(HERE, "type1_3d_16_m2000.npz"), xyz=xyz, c=c, n=n, d=d,
             direct=O.direct_type1_3d(xyz, c, n, d))
    # 3D Type-2, 16^3 modes, M=2000
    rng = np.random.default_rng(3)
    xyz = f32(rng.uniform(-np.pi, np.pi, (2000, 3)))
    f = cn(rng, n)
    np.savez(os.path.join(HERE, "type2_3d_16_m2000.npz"), xyz=xyz, f=f, d=d,
             direct=O.direct_type2_3d(f, xyz, d))
    # gradient of ||T2(f,x) - b||^2 w.r.t. x and f, 12^3 modes, M=1500
    rng = np.random.default_rng(4)
    n = (12, 12, 12)
    d = (TWO_PI / 12,) * 3
    xyz = f32(rng.uniform(-np.pi, np.pi, (1500, 3)))
    f = cn(rng, n)
    b = cn(rng, 1500)
    r = O.direct_type2_3d(f, xyz, d) - b
    np.savez(os.path.join(HERE, "grad_type2_3d_12_m1500.npz"), xyz=xyz, f=f, b=b, d=d,
             grad_x=O.direct_type2_3d_grad_x(f, xyz, d, r),
             grad_f=2.0 * O.direct_type1_3d(xyz, r, n, d))


if __name__ == "__main__":
    main()
