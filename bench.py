#!/usr/bin/env python3
"""Benchmark of the gfx950 KB NUFFT: 3D Type-1/2 Mpts/s at rel. err <= 1e-5.

Default run (one MI355X), three forms of the metric:
  plain    (the headline `value`): 3D Type-1 and Type-2 NUFFT, 64^3 modes on
           a 128^3 oversampled grid, KB half-width w = 7, M = 2,000,000 points
           uniform in [-pi, pi)^3 (float32), strengths / modes with re, im ~
           N(0, 1) (complex64).  One step = one Type-1 and one Type-2 of all
           M points, inputs resident in HBM; value = 2 M steps / elapsed.
  grad     (`forms.grad`): L = ||T2(f, x) - b||^2 with 64^3 modes, 1,000,000
           uniform points, w = 6; one step = forward Type-2 plus the backward
           (gradients w.r.t. modes and coordinates) through torch.autograd.
  batched  (`forms.batched`): 64 Type-1 problems of 32^3 / 48^3 / 64^3 /
           96^3 modes (exact 64^3 / 96^3 / 128^3 / 192^3 grids), 100k - 4M
           points each on radial lines (density ~ 1/r^2), w = 6; the batched
           call captured in a HIP graph against 64 sequential calls.
Every form states its relative L2 error against FP64 direct sums, computed
on the device (anufft_direct_type1_select / anufft_direct_type2 /
anufft_direct_type2_host, FP64) on sampled modes / points after timing.

Printed: ONE JSON line with the metric, the pipeline roofline fraction and
HBM fraction of the plain step, the roofline of its dominant kernel and the
CPU baseline (the FP64 C restatement of apollo-nufft's algorithm, one thread
and all host cores, on a bounded sample of the same workload).  The rate
comes from K plain steps; the same K steps run once before them with HIP
events recorded around every stage on the stream the stage runs on, which
give the per-stage times and the dominant kernel's average duration (the
event records themselves add a few microseconds between kernels, so they
stay out of the rate).

Other runs: --config sweep (100^3 -> 200^3 and 127^3 -> 254^3, M = 16e6
clustered, --w 4/6/8), --config plain|grad|batched (one form only),
--method valu (vector-ALU kernels).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3   # MI355X published FP32 vector peak
PEAK_FP16_TFLOPS = 2500.0  # MI355X published dense FP16 MFMA peak
PEAK_HBM_GBS = 8000.0      # MI355X HBM3E
TWO_PI = 2.0 * np.pi
METRIC = "3D Type-1/2 NUFFT Mpts/s at rel err<=1e-5 vs FP64 direct; plain/grad/batched"


def rel_l2(a, b):
    a = np.ravel(a)
    b = np.ravel(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def make_inputs(rng, M, n):
    xyz = rng.uniform(-np.pi, np.pi, size=(M, 3)).astype(np.float32)
    c = (rng.standard_normal(M) + 1j * rng.standard_normal(M)).astype(np.complex64)
    f = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    return xyz, c, f


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


def radial_points(rng, M):
    """Points on ceil(M/512) lines through the origin with directions uniform
    on the sphere and r uniform in [-pi, pi) (3D density ~ 1/r^2)."""
    lines = -(-M // 512)
    dirs = rng.standard_normal((lines, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    r = rng.uniform(-np.pi, np.pi, size=(lines, 512))
    pts = (r[:, :, None] * dirs[:, None, :]).reshape(-1, 3)[:M]
    return np.ascontiguousarray(pts, dtype=np.float32)


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


TUNE_CACHE = None


def tune_cache():
    """The autotuner's JSON cache: --tune-cache PATH (entries found there are
    reused without timing runs), else a fresh file, so nothing carries over
    between runs."""
    if TUNE_CACHE:
        return TUNE_CACHE
    import tempfile
    return os.path.join(tempfile.mkdtemp(prefix="anufft_tune_"), "tune.json")


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


# ---------------------------------------------------------------- device FP64 references
def device_errors_t1(torch, plan, txyz, tc, out1, n, rng, nsel):
    """Rel. L2 of Type-1 output on nsel random modes vs the device FP64 direct sum."""
    sel = rng.choice(int(np.prod(n)), size=nsel, replace=False).astype(np.int64)
    tsel = torch.from_numpy(sel).to(txyz.device)
    ref = plan.direct_type1_select(txyz, tc.to(torch.complex128), tsel).cpu().numpy()
    return rel_l2(out1.reshape(-1)[tsel].cpu().numpy(), ref)


def device_errors_t2(torch, plan, txyz, tf, out2, rng, npts):
    """Rel. L2 of Type-2 output on npts random points vs the device FP64 direct sum."""
    M = txyz.shape[0]
    pj = torch.from_numpy(rng.choice(M, size=npts, replace=False).astype(np.int64)).to(txyz.device)
    ref = plan.direct_type2(txyz[pj].contiguous(), tf).cpu().numpy()
    return rel_l2(out2[pj].cpu().numpy(), ref)


# ---------------------------------------------------------------- forms
def bench_plain(args, torch, A, dev, steps, warmup):
    n = (64, 64, 64)
    w = args.w or 7
    M = args.points or 2_000_000
    rng = np.random.default_rng(1234)
    xyz, c, f = make_inputs(rng, M, n)
    d = (TWO_PI / n[0],) * 3
    txyz = torch.from_numpy(xyz).to(dev)
    tc = torch.from_numpy(c).to(dev)
    tf = torch.from_numpy(f).to(dev)
    # launch settings from the autotuner, timed on these points (untimed here)
    tuned = None if args.no_autotune else A.autotune3d(
        *n, *d, 2, w, positions=txyz, workload="both", cache=tune_cache(),
        kernel_method=None if args.method == "auto" else args.method)
    plan = A.NufftPlan3D(*n, *d, 2, w, kernel_method=args.method, tuned=tuned)
    meta = plan.meta
    out1 = torch.empty(n, dtype=torch.complex64, device=dev)
    out2 = torch.empty(M, dtype=torch.complex64, device=dev)
    plan.workspace(M, dev)

    def step():
        plan.type1(txyz, tc, out=out1)
        plan.type2(txyz, tf, out=out2)

    for _ in range(warmup):
        step()
    # K steps with HIP events around every stage (on the stream each stage
    # runs on): stage and kernel times; the event records add a few
    # microseconds between kernels, so the rate comes from the K plain steps
    # that follow
    torch.cuda.synchronize()
    plan.timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    el_instr = time.perf_counter() - t0
    stages = plan.timing_collect()
    plan.timing(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0

    # per-type rates from separate timed loops of the same length
    per_type = {}
    for name, fn in (("type1", lambda: plan.type1(txyz, tc, out=out1)),
                     ("type2", lambda: plan.type2(txyz, tf, out=out2))):
        torch.cuda.synchronize()
        ta = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        per_type[name] = M * steps / (time.perf_counter() - ta) / 1e6
    res = {"elapsed": el, "elapsed_instrumented": el_instr, "M": M, "n": n, "w": w, "meta": meta,
           "stages": stages, "per_type": per_type, "tuned": tuned}
    if not args.no_check:
        res["err"] = {"type1": device_errors_t1(torch, plan, txyz, tc, out1, n, rng, 4096),
                      "type2": device_errors_t2(torch, plan, txyz, tf, out2, rng, 4096)}
    return res


def bench_grad(args, torch, A, dev, steps, warmup):
    """Gradient form: L = ||T2(f, x) - b||^2 with 64^3 modes, M = 1e6 uniform
    points, w = 6; one step = forward Type-2 + backward (the fused backward
    gives T1 for df and the kernel-derivative interpolation for dx)."""
    from apollofft_amd.autograd import nufft_type2
    n = (64, 64, 64)
    M = args.grad_points or 1_000_000
    w = 6
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

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / steps
    row = {"workload": "L = ||T2(f,x) - b||^2, 64^3 modes -> 128^3 grid, KB w=6, "
                       f"M={M:,} uniform points; forward + backward (df and dx)",
           "points": M, "kernel_width": w, "ms_per_step": round(el * 1e3, 4),
           "mpts_per_s": round(M / el / 1e6, 2)}
    if not args.no_check:
        gx = tx.grad.detach().cpu().numpy().astype(np.float64)
        gf = tf.grad.detach().cpu().numpy().astype(np.complex128)
        # r = T2(f, x) - b in FP64 at every point (device direct sum)
        txd = tx.detach()
        tfd = tf.detach()
        r = plan.direct_type2(txd, tfd) - tb.to(torch.complex128)
        # dL/df = 2 T1(r): FP64 direct sum at 256 random modes
        sel = rng.choice(int(np.prod(n)), size=256, replace=False).astype(np.int64)
        ref_f = 2.0 * plan.direct_type1_select(txd, r, torch.from_numpy(sel).to(dev)).cpu().numpy()
        row["rel_err_grad_modes_vs_fp64_direct"] = rel_l2(gf.ravel()[sel], ref_f)
        # dL/dx: FP64 central differences (h = 1e-6) of the loss on 256
        # (point, axis) entries; only term j moves with x_j
        pj = rng.choice(M, size=256, replace=False)
        axis = rng.integers(0, 3, size=256)
        h = 1e-6
        x64 = xyz[pj].astype(np.float64)
        xp, xm = x64.copy(), x64.copy()
        xp[np.arange(256), axis] += h
        xm[np.arange(256), axis] -= h
        f64 = f.astype(np.complex128)
        cp = plan.direct_type2_host(xp, f64)
        cm = plan.direct_type2_host(xm, f64)
        bj = b[pj].astype(np.complex128)
        fd = (np.abs(cp - bj) ** 2 - np.abs(cm - bj) ** 2) / (2.0 * h)
        row["rel_err_grad_x_vs_fp64_central_diff"] = rel_l2(gx[pj, axis], fd)
        row["accuracy_ok"] = bool(max(row["rel_err_grad_modes_vs_fp64_direct"],
                                      row["rel_err_grad_x_vs_fp64_central_diff"]) <= 1e-5)
    del plan
    return row


def bench_batched(args, torch, A, dev, steps, warmup):
    """Batched Type-1: 64 independent problems; one batched call captured in
    a HIP graph against 64 sequential plain calls on the same GPU."""
    nprob = args.problems or 64
    w = 6
    rng = np.random.default_rng(55)
    sizes = rng.choice([32, 48, 64, 96], size=nprob)
    Ms = rng.integers(100_000, 4_000_001, size=nprob)
    if args.batch_points:
        Ms = np.minimum(Ms, args.batch_points)
    plans, pos, vals, outs = [], [], [], []
    for nn, M in zip(sizes, Ms):
        nn = int(nn)
        xyz = radial_points(rng, int(M))
        c = (rng.standard_normal(int(M)) + 1j * rng.standard_normal(int(M))).astype(np.complex64)
        pos.append(torch.from_numpy(xyz).to(dev))
        vals.append(torch.from_numpy(c).to(dev))
        outs.append(torch.empty((nn, nn, nn), dtype=torch.complex64, device=dev))
    # one plan per mode count; launch settings from the autotuner, timed on
    # the largest problem of that size (untimed here)
    plan_cache, tuned = {}, {}
    cache = tune_cache()
    for nn in sorted(set(int(v) for v in sizes)):
        d = (TWO_PI / nn,) * 3
        t = None
        if not args.no_autotune:
            big = max((i for i in range(nprob) if int(sizes[i]) == nn), key=lambda i: int(Ms[i]))
            t = A.autotune3d(nn, nn, nn, *d, 2, w, positions=pos[big], workload="type1",
                             cache=cache, grid_policy="exact")
            tuned[str(nn)] = t
        plan_cache[nn] = A.NufftPlan3D(nn, nn, nn, *d, 2, w, grid_policy="exact", tuned=t)
    plans = [plan_cache[int(nn)] for nn in sizes]
    total = int(Ms.sum())
    seq_outs = [torch.empty_like(o) for o in outs]

    def sequential():
        for p, x, v, o in zip(plans, pos, vals, seq_outs):
            p.type1(x, v, out=o)

    for _ in range(warmup):
        sequential()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        sequential()
    torch.cuda.synchronize()
    t_seq = (time.perf_counter() - t0) / steps
    # batched: captured once into a HIP graph and replayed
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

    for _ in range(warmup):
        batched()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        batched()
    torch.cuda.synchronize()
    t_bat = (time.perf_counter() - t0) / steps
    # stage times of one eager batched call (HIP events on its streams)
    plans[0].timing(True)
    A.execute_type1_batched(plans, pos, vals, outs)
    torch.cuda.synchronize()
    bst = plans[0].timing_collect()
    plans[0].timing(False)
    row = {"workload": f"{nprob} Type-1 problems, modes 32^3/48^3/64^3/96^3 on exact grids, "
                       f"KB w={w}, {total:,} points on radial lines (~1/r^2)",
           "problems": nprob, "points": total, "kernel_width": w, "graph": graph is not None,
           "batched_ms": round(t_bat * 1e3, 3), "sequential_ms": round(t_seq * 1e3, 3),
           "batched_mpts_per_s": round(total / t_bat / 1e6, 1),
           "sequential_mpts_per_s": round(total / t_seq / 1e6, 1),
           "speedup_vs_sequential": round(t_seq / t_bat, 4), "tuned": tuned or None,
           "batched_stage_ms": {k: round(v[0], 3) for k, v in bst.items() if v[0]}}
    diffs = [rel_l2(o.cpu().numpy(), so.cpu().numpy()) for o, so in zip(outs, seq_outs)]
    row["max_rel_diff_batched_vs_sequential"] = max(diffs)
    if not args.no_check:
        errs = []
        for i in rng.choice(nprob, size=3, replace=False):
            nn = int(sizes[i])
            errs.append(device_errors_t1(torch, plans[i], pos[i], vals[i], outs[i], (nn,) * 3,
                                         rng, 512))
        row["rel_err_vs_fp64_direct_sampled"] = max(errs)
        row["accuracy_ok"] = bool(max(errs) <= 1e-5)
    del graph
    return row


def bench_sweep(args, torch, A, dev):
    """Precision sweep: 100^3 -> 200^3 and 127^3 -> 254^3 grids, M = 16e6
    clustered points, KB w in {4, 6, 8}; Type-1 and Type-2 Mpts/s and rel.
    L2 error vs FP64 direct sums on random modes / points."""
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
            plan = A.NufftPlan3D(*n, *d, 2, w, grid_policy="exact", kernel_method=args.method)
            plan.workspace(M, dev)
            for _ in range(args.warmup):
                plan.type1(txyz, tc, out=out1)
                plan.type2(txyz, tf, out=out2)
            plan.timing(True)
            per = {}
            for name, fn in (("type1", lambda: plan.type1(txyz, tc, out=out1)),
                             ("type2", lambda: plan.type2(txyz, tf, out=out2))):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for _ in range(args.steps):
                    fn()
                torch.cuda.synchronize()
                per[name] = (time.perf_counter() - t0) / args.steps
            st = plan.timing_collect()
            plan.timing(False)
            row = {"modes": nn, "grid": plan.meta["m"][0], "w": w, "points": M,
                   "type1_mpts_per_s": round(M / per["type1"] / 1e6, 1),
                   "type2_mpts_per_s": round(M / per["type2"] / 1e6, 1),
                   "stage_ms": {k: round(v[0] / args.steps, 3) for k, v in st.items() if v[0]}}
            if not args.no_check:
                row["rel_err_vs_fp64_direct"] = {
                    "type1": device_errors_t1(torch, plan, txyz, tc, out1, n, rng, 1024),
                    "type2": device_errors_t2(torch, plan, txyz, tf, out2, rng, 4096)}
            rows.append(row)
            print(json.dumps(row), file=sys.stderr, flush=True)
            del plan
            torch.cuda.empty_cache()
    return rows


# ---------------------------------------------------------------- roofline
def pipeline_roofline(M, n, m, w, stage_ms_per_step, step_ms):
    """t_roof = sum over the stages of one plain step (Type-1 + Type-2) of
    max(bytes / 8 TB/s, flops / 157.3 TFLOP/s), with the algorithmic bytes
    and flops per stage: binning 24 B/point; spreading and interpolation
    24 B/point + 8 B/cell and 4 (2w)^3 flops/point; each 3D FFT 32 B/cell
    (two read+write sweeps) and 5 G log2 G flops; crop/deconvolution 16 B
    per mode (Type-1), pad/deconvolution 8 B per mode + 8 B per cell."""
    G = float(np.prod(m))
    N = float(np.prod(n))
    T = 2 * w
    stages = {
        "bin": (2 * 24.0 * M, 0.0),
        "spread": (24.0 * M + 8.0 * G, 4.0 * M * T ** 3),
        "interp": (24.0 * M + 8.0 * G, 4.0 * M * T ** 3),
        "fft": (2 * 32.0 * G + 16.0 * N + 8.0 * N + 8.0 * G, 2 * 5.0 * G * math.log2(G)),
    }
    t_roof = 0.0
    total_bytes = 0.0
    per = {}
    for k, (by, fl) in stages.items():
        t = max(by / (PEAK_HBM_GBS * 1e9), fl / (PEAK_FP32_TFLOPS * 1e12))
        per[k] = round(t * 1e3, 4)
        t_roof += t
        total_bytes += by
    return {"t_roof_ms": round(t_roof * 1e3, 4), "t_measured_ms": round(step_ms, 4),
            "roofline_fraction": round(t_roof * 1e3 / step_ms, 4),
            "hbm_fraction": round(total_bytes / (step_ms * 1e-3) / (PEAK_HBM_GBS * 1e9), 4),
            "algorithmic_bytes_per_step": total_bytes, "stage_roof_ms": per}


def cpu_baseline(n, w, M_sample, seed=7):
    """apollo-nufft algorithm restated in C (FP64): Type-1 + Type-2 on a
    sample of the workload, one thread (the reference path) and all host
    cores (OpenMP); returns {threads: (Mpts/s, seconds)}."""
    import oracle
    rng = np.random.default_rng(seed)
    xyz, c, f = make_inputs(rng, M_sample, n)
    d = (TWO_PI / n[0], TWO_PI / n[1], TWO_PI / n[2])
    x64 = xyz.astype(np.float64)
    c128 = c.astype(np.complex128)
    f128 = f.astype(np.complex128)
    res = {}
    for threads in (1, host_cores()):
        oracle.set_fast_threads(threads)
        t0 = time.perf_counter()
        oracle.nufft3d_type1_fast(x64, c128, n, d, 2, w, oracle.POLICY_POW2)
        oracle.nufft3d_type2_fast(f128, x64, d, 2, w, oracle.POLICY_POW2)
        el = time.perf_counter() - t0
        res[threads] = (2.0 * M_sample / el / 1e6, el)
    oracle.set_fast_threads(1)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="all", choices=["all", "plain", "sweep", "grad", "batched"])
    ap.add_argument("--problems", type=int, default=0)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--w", type=int, default=0)
    ap.add_argument("--points", type=int, default=0)
    ap.add_argument("--grad-points", type=int, default=0)
    ap.add_argument("--batch-points", type=int, default=0,
                    help="cap the points of each batched problem")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="automatic launch settings instead of the autotuner's")
    ap.add_argument("--tune-cache", default="",
                    help="autotuner cache file to read and extend (default: a fresh one)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--method", default="auto", choices=["auto", "valu", "mfma"],
                    help="3D spreading / interpolation kernels")
    ap.add_argument("--cpu-sample", type=int, default=400_000)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run device-to-device copies of MIB MiB (HBM counter calibration)")
    args = ap.parse_args()
    global TUNE_CACHE
    TUNE_CACHE = args.tune_cache or None

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
        print(json.dumps({"config": "batched",
                          **bench_batched(args, torch, A, dev, args.steps, args.warmup)}), flush=True)
        return
    if args.config == "grad":
        print(json.dumps({"config": "grad",
                          **bench_grad(args, torch, A, dev, args.steps, args.warmup)}), flush=True)
        return

    r = bench_plain(args, torch, A, dev, args.steps, args.warmup)
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
    kname = ("spread3d_mfma_kernel" if dom == "spread" else "interp3d_win_kernel") if mfma else \
        ("spread3d_kernel" if dom == "spread" else "interp3d_kernel")
    roof = {"bound": "mfma" if mfma else "valu", "achieved": round(achieved, 3),
            "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
            "traffic": load_traffic(kname), "kernel": kname,
            "kernel_ms": round(dom_ms, 4), "flops_per_launch": flops}
    if mfma:
        # executed per point and tile x cell: spreading 6 32x32x16 FP16 MFMAs
        # of 32768 flops per 16 points; interpolation 6 16x16x32 MFMAs of
        # 16384 flops per 16 points (one 16-row window, pad slots not counted)
        px = r["meta"]["bin_size"][0] + T - 1
        ex_flops = (12288.0 if dom == "spread" else 6144.0) * px * M
        ex = ex_flops / (dom_ms * 1e-3) / 1e12
        roof["compute"] = ("FP32-accurate products as FP16 hi/lo split MFMAs (3 per product); "
                           "achieved = algorithmic FP32 flops (4 (2w)^3 per point), priced at "
                           "the FP32 peak")
        roof["mfma_executed"] = {"flops_per_launch": ex_flops, "tflops": round(ex, 2),
                                 "peak_fp16_dense": PEAK_FP16_TFLOPS,
                                 "frac": round(ex / PEAK_FP16_TFLOPS, 4)}
    else:
        roof["compute"] = "FP32 vector-ALU FMAs (4 (2w)^3 flops per point) priced at the FP32 peak"
    line = {
        "metric": METRIC,
        "value": round(value, 2), "unit": "Mpts/s", "n_gpus": 1, "steps": K,
        "warmup": args.warmup, "ms_per_step": round(el / K * 1e3, 4),
        "ms_per_step_stage_timed": round(r["elapsed_instrumented"] / K * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"3D Type-1 + Type-2 NUFFT, 64^3 modes -> 128^3 grid, KB w={r['w']}, "
                               f"M={M:,} uniform points, complex64",
                   "modes": list(r["n"]), "grid": list(r["meta"]["m"]), "kernel_width": r["w"],
                   "points": M, "bin_size": list(r["meta"]["bin_size"]),
                   "launch_settings": r["tuned"]},
        "mpts_per_s_by_type": {k: round(v, 2) for k, v in r["per_type"].items()},
        "stage_ms_per_step": {k: round(v, 4) for k, v in per_step.items()},
        "rel_err_vs_fp64_direct": r.get("err"),
        "pipeline": pipeline_roofline(M, r["n"], r["meta"]["m"], r["w"], per_step, el / K * 1e3),
        "roofline": roof,
    }
    line["roofline_fraction"] = line["pipeline"]["roofline_fraction"]
    line["hbm_fraction"] = line["pipeline"]["hbm_fraction"]
    if r.get("err"):
        line["accuracy_ok"] = bool(max(r["err"].values()) <= 1e-5)
    del r
    torch.cuda.empty_cache()
    if args.config == "all":
        forms = {}
        forms["grad"] = bench_grad(args, torch, A, dev, args.steps, args.warmup)
        torch.cuda.empty_cache()
        forms["batched"] = bench_batched(args, torch, A, dev, args.steps, args.warmup)
        torch.cuda.empty_cache()
        line["forms"] = forms
    if not args.no_cpu:
        base = cpu_baseline(tuple(line["config"]["modes"]), line["config"]["kernel_width"],
                            args.cpu_sample)
        v1, el1 = base[1]
        cores = host_cores()
        vm, elm = base[cores]
        line["cpu_baseline"] = {
            "value": round(v1, 4), "unit": "Mpts/s", "cores": 1, "kind": "port",
            "sample": f"Type-1 + Type-2 of {args.cpu_sample} of the {M} points on the full "
                      f"128^3 grid, w={line['config']['kernel_width']}, FP64 C restatement of "
                      f"apollo-nufft, 1 thread, {el1:.1f} s",
            "host_cores": cores,
            "multi_thread": {"value": round(vm, 4), "unit": "Mpts/s", "threads": cores,
                             "seconds": round(elm, 2)}}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
