"""Golden fixtures (tests/golden/) against the oracle and the plan metadata
(CPU), and against the device path (GPU)."""
import json
import os

import numpy as np
import pytest

from helpers import rel_l2

import apollofft_amd as A
import oracle as O

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(G, name), allow_pickle=False)


def test_kb_meta_fixture_matches_plans():
    meta = json.load(open(os.path.join(G, "kb_meta.json")))
    for w, v in meta["widths"].items():
        p = A.NufftPlan3D(8, 8, 8, 0.1, 0.1, 0.1, 2, int(w))
        m = p.meta
        assert m["beta"] == v["beta"]
        assert abs(m["i0_beta"] - v["i0_beta_scipy"]) <= 2e-15 * v["i0_beta_scipy"]
    for key, d in meta["deconv_w6"].items():
        n, m = (int(t) for t in key.split("_"))
        p = A.NufftPlan1D(n, 2 * np.pi / n, 2, 6)
        assert p.meta["m"][0] == m
        assert np.allclose(p.deconv(0), d, rtol=1e-15, atol=0)


def test_oracle_reproduces_golden_direct_sums():
    z = _load("type1_1d_n256_m4096.npz")
    assert np.allclose(O.direct_type1_1d(z["x"], z["c"], int(z["n"]), float(z["dx"])), z["direct"],
                       rtol=1e-12, atol=1e-11)
    z = _load("type1_3d_16_m2000.npz")
    out = O.direct_type1_3d(z["xyz"], z["c"], tuple(z["n"]), tuple(z["d"]))
    assert rel_l2(out, z["direct"]) < 1e-13
    z = _load("type2_3d_16_m2000.npz")
    assert rel_l2(O.direct_type2_3d(z["f"], z["xyz"], tuple(z["d"])), z["direct"]) < 1e-13
    # fast restatement within the KB bound of the golden direct sums
    z = _load("type1_1d_n256_m4096.npz")
    fast = O.nufft1d_type1_fast(z["x"], z["c"], int(z["n"]), float(z["dx"]), 2, 6)
    assert rel_l2(fast, z["direct"]) < 1e-9


def _t(a, dev, dtype=None):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a))
    return (t.to(dtype) if dtype is not None else t).to(dev)


@pytest.mark.gpu
def test_gpu_type1_1d_golden(dev):
    import torch
    z = _load("type1_1d_n256_m4096.npz")
    plan = A.NufftPlan1D(int(z["n"]), float(z["dx"]), 2, 6)
    out = plan.type1(_t(z["x"], dev, torch.float32), _t(z["c"], dev, torch.complex64))
    assert rel_l2(out.cpu().numpy(), z["direct"]) < 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("w", [4, 6, 7, 8])
def test_gpu_type1_type2_3d_golden(dev, w):
    import torch
    z = _load("type1_3d_16_m2000.npz")
    d = tuple(float(v) for v in z["d"])
    plan = A.NufftPlan3D(16, 16, 16, *d, 2, w)
    out = plan.type1(_t(z["xyz"], dev, torch.float32), _t(z["c"], dev, torch.complex64))
    assert rel_l2(out.cpu().numpy(), z["direct"]) < (1e-5 if w > 4 else 3e-5)
    z = _load("type2_3d_16_m2000.npz")
    out = plan.type2(_t(z["xyz"], dev, torch.float32), _t(z["f"], dev, torch.complex64))
    assert rel_l2(out.cpu().numpy(), z["direct"]) < (1e-5 if w > 4 else 3e-5)


@pytest.mark.gpu
def test_gpu_gradients_golden(dev):
    """Coordinate gradient and mode gradient (opposite-type NUFFT) of
    ||T2(f,x) - b||^2 against the golden direct-sum gradients."""
    import torch
    z = _load("grad_type2_3d_12_m1500.npz")
    d = tuple(float(v) for v in z["d"])
    plan = A.NufftPlan3D(12, 12, 12, *d, 2, 7)
    xyz = _t(z["xyz"], dev, torch.float32)
    f = _t(z["f"], dev, torch.complex64)
    c = plan.type2(xyz, f)
    r = c - _t(z["b"], dev, torch.complex64)
    g = 2.0 * r
    gx = plan.grad_points(xyz, g.contiguous(), f)
    gf = plan.type1(xyz, g.contiguous())
    torch.cuda.synchronize()
    assert rel_l2(gx.cpu().numpy(), z["grad_x"]) < 1e-5
    assert rel_l2(gf.cpu().numpy(), z["grad_f"]) < 1e-5
