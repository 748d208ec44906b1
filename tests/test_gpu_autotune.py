"""The autotuner (anufft_autotune3d) on the device: picks settings from its
candidate sets, keeps them in a JSON cache and reads them back, and plans
built with tuned or forced settings give the results of the automatic ones
(only the summation order may differ) within the FP64 direct-sum bound."""
import json

import numpy as np
import pytest

from helpers import TWO_PI, cnormal, radial_points, rel_l2, uniform_points

pytestmark = pytest.mark.gpu

CANDIDATES = {"kernel_method": {1, 2}, "max_subproblem": {1024, 2048, 4096, 8192, 16384},
              "sub_work": {8, 16, 48, 128, 384}, "blocks_per_cu": {0, 1, 2, 3, 4},
              "fft_lines": {0, 1, 2, 4, 8, 16, 32}}


def _t(x, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def _outputs(plan, tx, tc, tf):
    import torch
    o1 = plan.type1(tx, tc)
    o2 = plan.type2(tx, tf)
    torch.cuda.synchronize()
    return o1.cpu().numpy(), o2.cpu().numpy()


def test_autotune_cache_and_parity(dev, oracle_mod, tmp_path):
    import apollofft_amd as A
    n, w, M = (32, 32, 32), 6, 200_000
    d = (TWO_PI / 32,) * 3
    cache = tmp_path / "tune.json"
    t = A.autotune3d(*n, *d, 2, w, num_points=M, cache=cache)
    print(f"autotune uniform: {t}", flush=True)
    for k, vals in CANDIDATES.items():
        assert t[k] in vals, (k, t[k])
    assert t["ms"] > 0
    doc = json.loads(cache.read_text())
    assert len(doc) == 1
    key = next(iter(doc))
    assert "|w=6|" in key and "points=uniform" in key
    # a second call reads the cache
    t2 = A.autotune3d(*n, *d, 2, w, num_points=M, cache=cache)
    assert t2 == {**t, "ms": pytest.approx(t["ms"], abs=2e-6)}

    rng = np.random.default_rng(31)
    xyz = uniform_points(rng, M)
    c = cnormal(rng, M)
    f = cnormal(rng, n)
    tx, tc, tf = _t(xyz, dev), _t(c, dev), _t(f, dev)
    base = _outputs(A.NufftPlan3D(*n, *d, 2, w), tx, tc, tf)
    tuned = _outputs(A.NufftPlan3D(*n, *d, 2, w, tuned=t), tx, tc, tf)
    assert rel_l2(tuned[0], base[0]) < 1e-6 and rel_l2(tuned[1], base[1]) < 1e-6
    sel = rng.choice(int(np.prod(n)), size=512, replace=False).astype(np.int64)
    x64 = xyz.astype(np.float64)
    r1 = oracle_mod.direct_type1_3d(x64, c.astype(np.complex128), n, d, sel)
    pj = rng.choice(M, size=512, replace=False)
    r2 = oracle_mod.direct_type2_3d(f.astype(np.complex128), x64[pj], d)
    assert rel_l2(tuned[0].ravel()[sel], r1) < 1e-5
    assert rel_l2(tuned[1][pj], r2) < 1e-5


def test_autotune_on_caller_points(dev, tmp_path):
    """Tuning on the caller's (radially clustered) points is a separate
    cache entry; the settings come from the candidate sets."""
    import apollofft_amd as A
    rng = np.random.default_rng(32)
    n, w = (48, 48, 48), 6
    d = (TWO_PI / 48,) * 3
    pts = _t(radial_points(rng, 300_000), dev)
    cache = tmp_path / "tune.json"
    A.autotune3d(*n, *d, 2, w, num_points=300_000, cache=cache, grid_policy="exact")
    t = A.autotune3d(*n, *d, 2, w, positions=pts, workload="type1", cache=cache,
                     grid_policy="exact")
    print(f"autotune radial: {t}", flush=True)
    for k, vals in CANDIDATES.items():
        assert t[k] in vals, (k, t[k])
    doc = json.loads(cache.read_text())
    assert len(doc) == 2 and any("points=caller" in k and "workload=1" in k for k in doc)
    # more entries: the file stays valid JSON and earlier entries keep their values
    A.autotune3d(*n, *d, 2, w, positions=pts, workload="type2", cache=cache, grid_policy="exact")
    A.autotune3d(*n, *d, 2, w, positions=pts, workload="both", cache=cache, grid_policy="exact")
    doc2 = json.loads(cache.read_text())
    assert len(doc2) == 4
    for k, v in doc.items():
        assert doc2[k] == v


@pytest.mark.parametrize("settings", [
    {"max_subproblem": 1024, "sub_work": 8, "blocks_per_cu": 1, "fft_lines": 1},
    {"max_subproblem": 16384, "sub_work": 384, "blocks_per_cu": 4, "fft_lines": 32},
    {"kernel_method": "valu", "blocks_per_cu": 3, "fft_lines": 2},
])
def test_forced_launch_settings_keep_results(dev, oracle_mod, settings):
    import apollofft_amd as A
    rng = np.random.default_rng(33)
    n, w, M = (40, 36, 32), 7, 120_000
    d = tuple(TWO_PI / k for k in n)
    xyz = uniform_points(rng, M)
    c = cnormal(rng, M)
    f = cnormal(rng, n)
    tx, tc, tf = _t(xyz, dev), _t(c, dev), _t(f, dev)
    base = _outputs(A.NufftPlan3D(*n, *d, 2, w, grid_policy="exact"), tx, tc, tf)
    got = _outputs(A.NufftPlan3D(*n, *d, 2, w, grid_policy="exact", **settings), tx, tc, tf)
    assert rel_l2(got[0], base[0]) < 1e-6 and rel_l2(got[1], base[1]) < 1e-6
    sel = rng.choice(int(np.prod(n)), size=256, replace=False).astype(np.int64)
    r1 = oracle_mod.direct_type1_3d(xyz.astype(np.float64), c.astype(np.complex128), n, d, sel)
    assert rel_l2(got[0].ravel()[sel], r1) < 1e-5
