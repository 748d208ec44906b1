#!/usr/bin/env python3
"""Benchmark of the gfx950 KB NUFFT: 3D Type-1/2 Mpts/s at rel. err <= 1e-5.

Default workload (one MI355X): 3D Type-1 and Type-2 NUFFT, 64^3 modes on a
128^3 oversampled grid, KB half-width w=7, M = 2,000,000 points uniform in
[-pi, pi)^3 (float32), strengths/modes with re, im ~ N(0,1) (complex64).
One step = one Type-1 and one Type-2 transform of all M points, inputs
resident in HBM.  value = 2 M steps / elapsed (point-transforms per second,
in Mpts/s).  After timing, both outputs are checked against an FP64 direct
sum on 4096 random modes (Type-1) and 4096 random points (Type-2).

Printed: ONE JSON line with the metric, the roofline of the dominant kernel
(its average duration from HIP events recorded on its own stream during the
timed steps) and the CPU baseline (the FP64 C restatement of apollo-nufft's
algorithm, one thread, on a bounded sample of the same workload).

Other forms (not the default line): --config grad (Type-2 forward plus the
gradient w.r.t. modes and coordinates, 64^3 modes, M = 1e6, w = 6),
--config sweep (100^3 -> 200^3 and 127^3 -> 254^3, M = 16e6 clustered,
--w 4/6/8), --config batched (64 uneven problems).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3   # MI355X published FP32 peak
PEAK_HBM_GBS = 8000.0      # MI355X HBM3E
TWO_PI = 2.0 * np.pi


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def make_inputs(rng, M, n):
    xyz = rng.uniform(-np.pi, np.pi, size=(M, 3)).astype(np.float32)
    c = (rng.standard_normal(M) + 1j * rng.standard_normal(M)).astype(np.complex64)
    f = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    return xyz, c, f


def load_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary (FETCH_SIZE / WRITE_SIZE scaled by a copy-kernel calibration)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as fh:
            t = json.load(fh)
        k = t["kernels"][kernel]
        return float(k["fetch_kib"] * 1024.0 * t["fetch_scale"] + k["write_kib"] * 1024.0 * t["write_scale"])
    except (OSError, KeyError, ValueError, TypeError):
        return None


def cpu_baseline(n, w, M_sample, seed=7):
    """apollo-nufft algorithm restated in C (FP64, one thread): Type-1 +
    Type-2 on a sample of the workload; returns Mpts/s and the sample text."""
    import oracle
    rng = np.random.default_rng(seed)
    xyz, c, f = make_inputs(rng, M_sample, n)
    d = (TWO_PI / n[0], TWO_PI / n[1], TWO_PI / n[2])
    x64 = xyz.astype(np.float64)
    t0 = time.perf_counter()
    oracle.nufft3d_type1_fast(x64, c.astype(np.complex128), n, d, 2, w, oracle.POLICY_POW2)
    oracle.nufft3d_type2_fast(f.astype(np.complex128), x64, d, 2, w, oracle.POLICY_POW2)
    el = time.perf_counter() - t0
    return 2.0 * M_sample / el / 1e6, el


def bench_plain(args, torch, A, dev):
    n = (64, 64, 64)
    w = args.w or 7
    M = args.points or 2_000_000
    rng = np.random.default_rng(1234)
    xyz, c, f = make_inputs(rng, M, n)
    d = (TWO_PI / n[0],) * 3
    plan = A.NufftPlan3D(*n, *d, 2, w)
    meta = plan.meta
    txyz = torch.from_numpy(xyz).to(dev)
    tc = torch.from_numpy(c).to(dev)
    tf = torch.from_numpy(f).to(dev)
    out1 = torch.empty(n, dtype=torch.complex64, device=dev)
    out2 = torch.empty(M, dtype=torch.complex64, device=dev)
    plan.workspace(M, dev)

    def step():
        plan.type1(txyz, tc, out=out1)
        plan.type2(txyz, tf, out=out2)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    plan.timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    stages = plan.timing_collect()
    plan.timing(False)

    # per-type rates from separate timed loops of the same length
    per_type = {}
    for name, fn in (("type1", lambda: plan.type1(txyz, tc, out=out1)),
                     ("type2", lambda: plan.type2(txyz, tf, out=out2))):
        torch.cuda.synchronize()
        ta = time.perf_counter()
        for _ in range(args.steps):
            fn()
        torch.cuda.synchronize()
        per_type[name] = M * args.steps / (time.perf_counter() - ta) / 1e6
    res = {"elapsed": el, "M": M, "n": n, "w": w, "meta": meta, "stages": stages,
           "per_type": per_type}
    # accuracy vs FP64 direct sums
    if not args.no_check:
        import oracle
        sel = rng.choice(int(np.prod(n)), size=4096, replace=False).astype(np.int64)
        o1 = out1.cpu().numpy().ravel()[sel]
        r1 = oracle.direct_type1_3d(xyz.astype(np.float64), c.astype(np.complex128), n, d, sel)
        pj = rng.choice(M, size=4096, replace=False)
        o2 = out2.cpu().numpy()[pj]
        r2 = oracle.direct_type2_3d(f.astype(np.complex128), xyz[pj].astype(np.float64), d)
        res["err"] = {"type1": rel_l2(o1, r1), "type2": rel_l2(o2, r2)}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="plain", choices=["plain"])
    ap.add_argument("--w", type=int, default=0)
    ap.add_argument("--points", type=int, default=0)
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=400_000)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run device-to-device copies of MIB MiB (HBM counter calibration)")
    args = ap.parse_args()

    if args.calibrate_copy:
        import torch
        nbytes = args.calibrate_copy << 20
        src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": nbytes, "copies": 5}))
        return

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this NUFFT build runs on one GPU; multi-GPU sharding is "
                                   "designed but not built (DESIGN.md), use --gpus 1"}))
        sys.exit(2)

    import torch
    import apollofft_amd as A
    dev = torch.device("cuda:0")
    r = bench_plain(args, torch, A, dev)
    M, K, el = r["M"], args.steps, r["elapsed"]
    st = r["stages"]
    per_step = {k: v[0] / K for k, v in st.items()}
    value = 2.0 * M * K / el / 1e6
    # dominant kernel: spreading or interpolation, whichever takes longer
    dom = "spread" if per_step["spread"] >= per_step["interp"] else "interp"
    dom_ms = st[dom][0] / max(st[dom][1], 1)
    T = 2 * r["w"]
    flops = 4.0 * M * T ** 3
    achieved = flops / (dom_ms * 1e-3) / 1e12
    kname = "spread3d_kernel" if dom == "spread" else "interp3d_kernel"
    roof = {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
            "traffic": load_traffic(kname), "kernel": kname,
            "kernel_ms": round(dom_ms, 4), "flops_per_launch": flops,
            "compute": "fp32 vector (VALU) FMAs priced at the published FP32 peak"}
    line = {
        "metric": "3D Type-1/2 NUFFT Mpts/s at rel err<=1e-5 vs FP64 direct; plain/grad/batched",
        "value": round(value, 2), "unit": "Mpts/s", "n_gpus": 1, "steps": K,
        "warmup": args.warmup, "ms_per_step": round(el / K * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "3D Type-1 + Type-2 NUFFT, 64^3 modes -> 128^3 grid, KB w=7, "
                               "M=2,000,000 uniform points, complex64",
                   "modes": list(r["n"]), "grid": list(r["meta"]["m"]), "kernel_width": r["w"],
                   "points": M, "bin_size": list(r["meta"]["bin_size"])},
        "mpts_per_s_by_type": {k: round(v, 2) for k, v in r["per_type"].items()},
        "stage_ms_per_step": {k: round(v, 4) for k, v in per_step.items()},
        "rel_err_vs_fp64_direct": r.get("err"),
        "roofline": roof,
    }
    if r.get("err"):
        line["accuracy_ok"] = bool(max(r["err"].values()) <= 1e-5)
    if not args.no_cpu:
        v, el_cpu = cpu_baseline(r["n"], r["w"], args.cpu_sample)
        line["cpu_baseline"] = {
            "value": round(v, 4), "unit": "Mpts/s", "cores": 1, "kind": "port",
            "sample": f"Type-1 + Type-2 of {args.cpu_sample} of the {M} points on the full "
                      f"128^3 grid, w={r['w']}, FP64 C restatement of apollo-nufft, 1 thread, "
                      f"{el_cpu:.1f} s"}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
