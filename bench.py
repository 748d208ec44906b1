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
PEAK_FP16_TFLOPS = 2500.0  # MI355X published dense FP16 MFMA peak
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


def clustered_inputs(rng, M, n):
    """70% of the points from 8 Gaussian clusters (sigma 0.1, wrapped into
    [-pi, pi)), 30% uniform; values and modes re/im ~ N(0,1)."""
    mc = int(0.7 * M)
    centres = rng.uniform(-np.pi, np.pi, size=(8, 3))
    pts = np.empty((M, 3), dtype=np.float32)
    which = rng.integers(0, 8, size=mc)
    pts[:mc] = ((centres[which] + 0.1 * rng.standard_normal((mc, 3)) + np.pi) % TWO_PI - np.pi)
    pts[mc:] = rng.uniform(-np.pi, np.pi, size=(M - mc, 3))
    pts = pts[rng.permutation(M)]
    c = (rng.standard_normal(M) + 1j * rng.standard_normal(M)).astype(np.complex64)
    f = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    return pts, c, f


def time_pair(torch, plan, txyz, tc, tf, out1, out2, steps, warmup):
    """Time Type-1 and Type-2 separately; returns (t1_s, t2_s) per call and
    per-stage ms per call."""
    for _ in range(warmup):
        plan.type1(txyz, tc, out=out1)
        plan.type2(txyz, tf, out=out2)
    res = {}
    plan.timing(True)
    for name, fn in (("type1", lambda: plan.type1(txyz, tc, out=out1)),
                     ("type2", lambda: plan.type2(txyz, tf, out=out2))):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        res[name] = (time.perf_counter() - t0) / steps
    st = plan.timing_collect()
    plan.timing(False)
    return res, {k: v[0] / steps for k, v in st.items()}


def check_errors(rng, xyz, c, f, n, d, out1, out2, nsel=4096, nsel_modes=4096):
    import oracle
    sel = rng.choice(int(np.prod(n)), size=nsel_modes, replace=False).astype(np.int64)
    o1 = out1.cpu().numpy().ravel()[sel]
    r1 = oracle.direct_type1_3d(xyz.astype(np.float64), c.astype(np.complex128), n, d, sel)
    pj = rng.choice(xyz.shape[0], size=nsel, replace=False)
    o2 = out2.cpu().numpy()[pj]
    r2 = oracle.direct_type2_3d(f.astype(np.complex128), xyz[pj].astype(np.float64), d)
    return {"type1": rel_l2(o1, r1), "type2": rel_l2(o2, r2)}


def bench_sweep(args, torch, A, dev):
    """Precision sweep: 100^3 -> 200^3 and 127^3 -> 254^3 grids, M = 16e6
    clustered points, KB w in {4, 6, 8}; Type-1 and Type-2 Mpts/s and rel.
    L2 error vs the FP64 direct sum on 4096 random modes / points."""
    M = args.points or 16_000_000
    widths = [args.w] if args.w else [4, 6, 8]
    rows = []
    for nn in (100, 127):
        n = (nn, nn, nn)
        d = (TWO_PI / nn,) * 3
        rng = np.random.default_rng(nn)
        xyz, c, f = clustered_inputs(rng, M, n)
        txyz = torch.from_numpy(xyz).to(dev)
        tc = torch.from_numpy(c).to(dev)
        tf = torch.from_numpy(f).to(dev)
        out1 = torch.empty(n, dtype=torch.complex64, device=dev)
        out2 = torch.empty(M, dtype=torch.complex64, device=dev)
        for w in widths:
            plan = A.NufftPlan3D(*n, *d, 2, w, grid_policy="exact")
            plan.workspace(M, dev)
            per, st = time_pair(torch, plan, txyz, tc, tf, out1, out2, args.steps, args.warmup)
            row = {"modes": nn, "grid": plan.meta["m"][0], "w": w, "points": M,
                   "type1_mpts_per_s": round(M / per["type1"] / 1e6, 1),
                   "type2_mpts_per_s": round(M / per["type2"] / 1e6, 1),
                   "stage_ms": {k: round(v, 3) for k, v in st.items() if v}}
            if not args.no_check:
                # 1024 modes for Type-1: each FP64 direct mode sums all 16e6 points
                row["rel_err_vs_fp64_direct"] = check_errors(rng, xyz, c, f, n, d, out1, out2,
                                                             nsel_modes=1024)
            rows.append(row)
            print(json.dumps(row), file=sys.stderr, flush=True)
            del plan
            torch.cuda.empty_cache()
    return rows


def bench_grad(args, torch, A, dev):
    """Gradient form: L = ||T2(f, x) - b||^2 with 64^3 modes, M = 1e6 uniform
    points, w = 6; one step = forward Type-2 + backward (Type-1 for df and
    the kernel-derivative interpolation for dx)."""
    from apollofft_amd.autograd import nufft_type2
    n = (64, 64, 64)
    M = args.points or 1_000_000
    w = args.w or 6
    d = (TWO_PI / 64,) * 3
    rng = np.random.default_rng(44)
    xyz, _, f = make_inputs(rng, M, n)
    b = (rng.standard_normal(M) + 1j * rng.standard_normal(M)).astype(np.complex64)
    plan = A.NufftPlan3D(*n, *d, 2, w)
    tx = torch.from_numpy(xyz).to(dev).requires_grad_(True)
    tf = torch.from_numpy(f).to(dev).requires_grad_(True)
    tb = torch.from_numpy(b).to(dev)

    def step():
        tx.grad = None
        tf.grad = None
        loss = (nufft_type2(tf, tx, plan) - tb).abs().pow(2).sum()
        loss.backward()

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / args.steps
    row = {"modes": 64, "w": w, "points": M, "ms_per_step": round(el * 1e3, 3),
           "mpts_per_s": round(M / el / 1e6, 1)}
    if not args.no_check:
        import oracle
        gx = tx.grad.cpu().numpy().astype(np.float64)
        gf = tf.grad.cpu().numpy()
        pj = rng.choice(M, size=256, replace=False)
        x64 = xyz.astype(np.float64)
        f64 = f.astype(np.complex128)
        r = oracle.direct_type2_3d(f64, x64[pj], d) - b[pj]
        ref = oracle.direct_type2_3d_grad_x(f64, x64[pj], d, r)
        row["rel_err_grad_x_vs_fp64_direct"] = rel_l2(gx[pj], ref)
        sel = rng.choice(int(np.prod(n)), size=256, replace=False).astype(np.int64)
        rfull = oracle.direct_type2_3d(f64, x64, d) - b if M <= 200_000 else None
        if rfull is not None:
            refg = 2.0 * oracle.direct_type1_3d(x64, rfull, n, d, sel)
            row["rel_err_grad_modes_vs_fp64_direct"] = rel_l2(gf.ravel()[sel], refg)
    return row


def radial_points(rng, M):
    """Points on ceil(M/512) lines through the origin with directions uniform
    on the sphere and r uniform in [-pi, pi) (3D density ~ 1/r^2)."""
    lines = -(-M // 512)
    dirs = rng.standard_normal((lines, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    r = rng.uniform(-np.pi, np.pi, size=(lines, 512))
    pts = (r[:, :, None] * dirs[:, None, :]).reshape(-1, 3)[:M]
    return np.ascontiguousarray(pts, dtype=np.float32)


def bench_batched(args, torch, A, dev):
    """Batched Type-1: 64 independent problems, modes per problem drawn from
    {32^3, 48^3, 64^3, 96^3} (grids 64^3, 96^3, 128^3, 192^3), M uniform in
    [100k, 4M] on radial lines, w = 6.  One batched call (fused persistent
    spreading, captured in a HIP graph) against 64 sequential plain calls."""
    nprob = args.problems or 64
    w = args.w or 6
    rng = np.random.default_rng(55)
    sizes = rng.choice([32, 48, 64, 96], size=nprob)
    Ms = rng.integers(100_000, 4_000_001, size=nprob)
    if args.points:
        Ms = np.minimum(Ms, args.points)
    plans, pos, vals, outs, host = [], [], [], [], []
    plan_cache = {}
    for nn, M in zip(sizes, Ms):
        nn = int(nn)
        if nn not in plan_cache:
            plan_cache[nn] = A.NufftPlan3D(nn, nn, nn, *(TWO_PI / nn,) * 3, 2, w, grid_policy="exact")
        plans.append(plan_cache[nn])
        xyz = radial_points(rng, int(M))
        c = (rng.standard_normal(int(M)) + 1j * rng.standard_normal(int(M))).astype(np.complex64)
        host.append((xyz, c))
        pos.append(torch.from_numpy(xyz).to(dev))
        vals.append(torch.from_numpy(c).to(dev))
        outs.append(torch.empty((nn, nn, nn), dtype=torch.complex64, device=dev))
    total = int(Ms.sum())
    # sequential plain launches (one workspace per distinct plan, as a user would)
    seq_outs = [torch.empty_like(o) for o in outs]

    def sequential():
        for p, x, v, o in zip(plans, pos, vals, seq_outs):
            p.type1(x, v, out=o)

    for _ in range(args.warmup):
        sequential()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sequential()
    torch.cuda.synchronize()
    t_seq = (time.perf_counter() - t0) / args.steps
    # batched: one call, captured once into a HIP graph and replayed
    A.execute_type1_batched(plans, pos, vals, outs)
    torch.cuda.synchronize()
    graph = None
    if not args.no_graph:
        graph = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            A.execute_type1_batched(plans, pos, vals, outs)
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        with torch.cuda.graph(graph):
            A.execute_type1_batched(plans, pos, vals, outs)
        torch.cuda.synchronize()

    def batched():
        if graph is not None:
            graph.replay()
        else:
            A.execute_type1_batched(plans, pos, vals, outs)

    for _ in range(args.warmup):
        batched()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batched()
    torch.cuda.synchronize()
    t_bat = (time.perf_counter() - t0) / args.steps
    row = {"problems": nprob, "points": total, "w": w, "graph": graph is not None,
           "batched_ms": round(t_bat * 1e3, 3), "sequential_ms": round(t_seq * 1e3, 3),
           "batched_mpts_per_s": round(total / t_bat / 1e6, 1),
           "sequential_mpts_per_s": round(total / t_seq / 1e6, 1),
           "speedup_vs_sequential": round(t_seq / t_bat, 4)}
    diffs = [rel_l2(o.cpu().numpy(), so.cpu().numpy()) for o, so in zip(outs, seq_outs)]
    row["max_rel_diff_batched_vs_sequential"] = max(diffs)
    if not args.no_check:
        import oracle
        errs = []
        for i in rng.choice(nprob, size=2, replace=False):
            nn = int(sizes[i])
            xyz, c = host[i]
            sel = rng.choice(nn ** 3, size=512, replace=False).astype(np.int64)
            ref = oracle.direct_type1_3d(xyz.astype(np.float64), c.astype(np.complex128),
                                         (nn,) * 3, (TWO_PI / nn,) * 3, sel)
            errs.append(rel_l2(outs[i].cpu().numpy().ravel()[sel], ref))
        row["rel_err_vs_fp64_direct_sampled"] = max(errs)
    return row


def bench_plain(args, torch, A, dev):
    n = (64, 64, 64)
    w = args.w or 7
    M = args.points or 2_000_000
    rng = np.random.default_rng(1234)
    xyz, c, f = make_inputs(rng, M, n)
    d = (TWO_PI / n[0],) * 3
    plan = A.NufftPlan3D(*n, *d, 2, w, kernel_method=args.method)
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
    ap.add_argument("--config", default="plain", choices=["plain", "sweep", "grad", "batched"])
    ap.add_argument("--problems", type=int, default=0)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--w", type=int, default=0)
    ap.add_argument("--points", type=int, default=0)
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--method", default="auto", choices=["auto", "valu", "mfma"],
                    help="3D spreading / interpolation kernels")
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
                                   "designed but not built here, use --gpus 1"}))
        sys.exit(2)

    import torch
    import apollofft_amd as A
    dev = torch.device("cuda:0")
    if args.config == "sweep":
        print(json.dumps({"config": "sweep", "rows": bench_sweep(args, torch, A, dev)}), flush=True)
        return
    if args.config == "batched":
        print(json.dumps({"config": "batched", **bench_batched(args, torch, A, dev)}), flush=True)
        return
    if args.config == "grad":
        print(json.dumps({"config": "grad", **bench_grad(args, torch, A, dev)}), flush=True)
        return
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
    mfma = r["meta"]["kernel_method"] == "mfma"
    kname = ("spread3d_mfma_kernel" if dom == "spread" else "interp3d_mfma_kernel") if mfma else \
        ("spread3d_kernel" if dom == "spread" else "interp3d_kernel")
    roof = {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
            "traffic": load_traffic(kname), "kernel": kname,
            "kernel_ms": round(dom_ms, 4), "flops_per_launch": flops}
    if mfma:
        # executed: per point and tile x cell, 6 (spread, 16 points) or 12 (interp,
        # 32 points) 32x32x16 FP16 MFMAs of 32768 flops each
        px = r["meta"]["bin_size"][0] + T - 1
        ex_flops = 12288.0 * px * M
        ex = ex_flops / (dom_ms * 1e-3) / 1e12
        roof["compute"] = ("FP32-accurate products as FP16 hi/lo split MFMAs (3 per product); "
                           "achieved = algorithmic FP32 flops, priced at the FP32 peak")
        roof["mfma_executed"] = {"flops_per_launch": ex_flops, "tflops": round(ex, 2),
                                 "peak_fp16_dense": PEAK_FP16_TFLOPS,
                                 "frac": round(ex / PEAK_FP16_TFLOPS, 4)}
    else:
        roof["compute"] = "fp32 vector (VALU) FMAs priced at the published FP32 peak"
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
