"""Input generators and error metrics shared by the tests."""
import numpy as np

TWO_PI = 2.0 * np.pi


def rel_l2(approx, exact):
    approx = np.asarray(approx).ravel()
    exact = np.asarray(exact).ravel()
    return float(np.linalg.norm(approx - exact) / max(np.linalg.norm(exact), 1e-300))


def max_rel(approx, exact):
    """apollo-validation relative_complex_error (suite.rs:2665-2674)."""
    approx = np.asarray(approx).ravel()
    exact = np.asarray(exact).ravel()
    return float(np.max(np.abs(exact - approx) / np.maximum(np.abs(exact), 1.0)))


def uniform_points(rng, M, dim=3, lo=-np.pi, hi=np.pi):
    """Coordinates drawn in float64 then rounded to float32 (the device
    input); returns the float32 array."""
    return rng.uniform(lo, hi, size=(M, dim) if dim > 1 else M).astype(np.float32)


def clustered_points(rng, M, nclusters=8, sigma=0.1, frac=0.7):
    """70% from Gaussian clusters (wrapped), 30% uniform, like the precision
    sweep workload."""
    mc = int(M * frac)
    centres = rng.uniform(-np.pi, np.pi, size=(nclusters, 3))
    which = rng.integers(0, nclusters, size=mc)
    pts = centres[which] + sigma * rng.standard_normal((mc, 3))
    pts = (pts + np.pi) % TWO_PI - np.pi
    uni = rng.uniform(-np.pi, np.pi, size=(M - mc, 3))
    allp = np.concatenate([pts, uni])
    rng.shuffle(allp)
    return allp.astype(np.float32)


def cnormal(rng, shape, dtype=np.complex64):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dtype)


def radial_points(rng, M, per_line=512):
    """Points on ceil(M / per_line) lines through the origin, directions
    uniform on the sphere, radius uniform in [-pi, pi) along each line: a 3D
    density proportional to 1/r^2.  float32, shape (M, 3)."""
    lines = -(-M // per_line)
    dirs = rng.standard_normal((lines, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    r = rng.uniform(-np.pi, np.pi, size=(lines, per_line))
    pts = (r[:, :, None] * dirs[:, None, :]).reshape(-1, 3)[:M]
    return np.ascontiguousarray(pts, dtype=np.float32)
