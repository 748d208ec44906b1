"""CPU replay of the device point binning (numpy, IEEE double).

The device computes, per axis, t = float64(x) * (m / L); t -= floor(t * (1/m)) m;
t += m if t < 0, else t -= m if t >= m; cell = min(floor(t), m - 1);
tile = cell // bin; key =
(tx * nby + ty) * nbz + tz, then a stable counting sort by key.  numpy never
fuses a*b+c, and q*m is exact for integer q, so this replay reproduces the
device keys and permutation bit for bit.
"""
import numpy as np


def fold_cells(xf32, n, d, m):
    scale = float(m) / (float(n) * float(d))
    t = np.asarray(xf32, dtype=np.float32).astype(np.float64) * scale
    q = np.floor(t * (1.0 / float(m)))
    t = t - q * float(m)
    t = np.where(t < 0.0, t + float(m), np.where(t >= float(m), t - float(m), t))
    c = np.floor(t).astype(np.int64)
    return np.clip(c, 0, m - 1)


def replay_sort(xyz_f32, n, d, m, bins):
    """Return (keys, perm, bin_start) for (M,3) float32 coordinates."""
    xyz_f32 = np.asarray(xyz_f32, dtype=np.float32)
    nb = [(m[a] + bins[a] - 1) // bins[a] for a in range(3)]
    cells = [fold_cells(xyz_f32[:, a], n[a], d[a], m[a]) for a in range(3)]
    t = [cells[a] // bins[a] for a in range(3)]
    keys = (t[0] * nb[1] + t[1]) * nb[2] + t[2]
    perm = np.argsort(keys, kind="stable")
    nbins = nb[0] * nb[1] * nb[2]
    bin_start = np.searchsorted(keys[perm], np.arange(nbins + 1), side="left")
    return keys, perm.astype(np.uint32), bin_start.astype(np.uint32)
