#!/usr/bin/env python3
"""Benchmark of the ConvStencil hot path on one MI355X.

Workload (default): box2d1r (9-point) stencil, FP64, 10240 x 10240 interior
with a zero halo of width 1, T = 64 time steps, seeded synthetic input (interior
uniform in [-1, 1), weights uniform in (0, 1) normalised to sum 1).  One bench
"step" is one full 64-time-step run over the grid, with the grid resident in
HBM; GStencil/s counts interior points x time steps / s, as the reference does
(src/2d/gpu.cu:329).

Prints one JSON line.  `roofline` describes the dominant kernel: algorithmic
bytes (8 B read + 8 B written per interior point per launch) or flops (2 per tap
per point-step) per launch over its mean launch time measured with HIP events
on the plan stream; `traffic` is the measured HBM bytes per launch from the
committed rocprofv3 PMC summary (profiles/), or null.  `cpu_baseline` is the
CPU oracle (the reference's loop, restated) timed on this host's cores on a
bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6

WORKLOADS = {
    # name: (shape, interior sizes, halo, steps)
    "box2d1r": ("box2d1r", [10240, 10240], 1, 64),
    "star2d3r": ("star2d3r", [8192, 8192], 3, 64),
    "box2d3r": ("box2d3r", [8192, 8192], 3, 64),
    "box3d1r": ("box3d1r", [512, 512, 512], 1, 32),
    "star3d1r": ("star3d1r", [512, 512, 512], 1, 32),
    "1d1r": ("1d1r", [1 << 20], 1, 64),
}


def load_traffic(kernel_key):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            d = json.load(f)
        v = d.get(kernel_key)
        return None if v is None else float(v["bytes_per_launch"])
    except (OSError, ValueError, KeyError, TypeError):
        return None


def cpu_baseline(layout, offs, w, seed, budget_s):
    """Times the oracle (OpenMP over all host cores) on the full grid for as many
    whole time steps as fit in about budget_s seconds (at least one).  Two untimed
    steps first (page faults of the fresh arrays), then a 2-step calibration."""
    from oracle import oracle
    threads = oracle.max_threads()
    x = oracle.fill_data(layout, seed)
    oracle.run(layout, offs, w, x, 2, threads=threads)
    t0 = time.perf_counter()
    oracle.run(layout, offs, w, x, 2, threads=threads)
    one = (time.perf_counter() - t0) / 2
    steps = max(1, int(budget_s / max(one, 1e-6)))
    t0 = time.perf_counter()
    oracle.run(layout, offs, w, x, steps, threads=threads)
    el = time.perf_counter() - t0
    pts = layout.points
    res = {"value": pts * steps / el / 1e9, "unit": "GStencil/s", "cores": threads, "kind": "port",
           "sample": f"full grid, {steps} time step(s) of the same stencil, {el:.1f} s, OpenMP {threads} threads"}
    # single thread, as the reference's loop runs (no OpenMP in the reference): 2 steps
    t0 = time.perf_counter()
    oracle.run(layout, offs, w, x, 2, threads=1)
    el1 = time.perf_counter() - t0
    res["single_thread"] = {"value": pts * 2 / el1 / 1e9, "cores": 1, "sample": f"full grid, 2 steps, {el1:.1f} s"}
    return res


def unfused_hbm(L, r, dense, T, seed, graph):
    """The single-step (fusion depth 1) kernel on the same grid: its HBM throughput is the
    streaming roofline figure (16 algorithmic bytes per point per step)."""
    from convstencil_amd import Plan
    with Plan(L, r, dense, fusion_depth=1, use_graph=graph) as p:
        p.fill_random(seed)
        rep = p.time(T, warmup=1, reps=3)
        kernel = p.describe().split()[1].split("=", 1)[1]
    ms_launch = rep.elapsed_ms / rep.launches
    gbs = 16.0 * L.points / (ms_launch * 1e-3) / 1e9
    return {"kernel": kernel, "gstencil_per_s": round(L.points * T / (rep.elapsed_ms * 1e-3) / 1e9, 2),
            "hbm_gbs_alg": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}


def split_comparison(seed, graph, shape="box2d3r", sizes=(8192, 8192), T=64):
    """One workload (default box2d3r 8192^2, halo 3, T = 64) on the FP64 path and the
    split forms (BF16x3; FP16 with three weight slices, fp16w3; FP16x2 with compensated
    weights), each with the
    autotuner's choice, on the same seeded grid; each split result's relative L2 and max
    error against the FP64 result (the FP64 path matches the CPU oracle to 1e-12
    relative max in the test suite)."""
    import convstencil_amd as cs
    from convstencil_amd import Layout, Plan, shape_offsets, dense_weights, SHAPES
    ndim, r, kind = SHAPES[shape]
    L = Layout.with_halo(list(sizes), r)
    offs = shape_offsets(shape)
    w = cs.fill_weights_host(len(offs), seed)
    dense = dense_weights(ndim, r, offs, w)
    res, out = {}, {}
    splits = ("bf16x3", "fp16w3", "fp16x2")
    for prec in ("fp64",) + splits:
        if ndim == 3 and prec == "fp16x2":
            continue  # 3-D split forms: BF16x3 and FP16 w3
        with Plan(L, r, dense, precision=prec, use_graph=graph, autotune=1) as p:
            p.fill_random(seed)
            rep = p.time(T, warmup=1, reps=3)
            p.fill_random(seed)
            p.run(T)
            out[prec] = L.interior(p.download()).copy()
            res[prec] = {"gstencil_per_s": round(L.points * T / (rep.elapsed_ms * 1e-3) / 1e9, 2),
                         "kernel": p.describe().split()[1].split("=", 1)[1]}
    for prec in [sp for sp in splits if sp in out]:
        d = out[prec] - out["fp64"]
        res[prec]["rel_l2_vs_fp64"] = float(np.sqrt(np.sum(d * d) / np.sum(out["fp64"] ** 2)))
        res[prec]["rel_max_vs_fp64"] = float(np.max(np.abs(d)) / np.max(np.abs(out["fp64"])))
        res[f"speedup_{prec}_over_fp64"] = round(res[prec]["gstencil_per_s"] / res["fp64"]["gstencil_per_s"], 3)
    res["workload"] = f"{shape} {'x'.join(map(str, sizes))} halo {r} T={T}"
    return res


def batch_workload(args):
    """256 independent 2-D grids, sides log-uniform in [128, 2048] (uneven, non-square),
    stencil drawn from {box2d1r, star2d3r}, own seeded weights, zero halo of the radius,
    T = 100; all grids advance in one persistent graph-captured launch per run, FP64 and
    BF16x3.  Reports aggregate GStencil/s and, per grid, the error against the CPU oracle
    (FP64: relative max; BF16x3: relative L2 and max)."""
    import convstencil_amd as cs
    from convstencil_amd import Batch, Layout, shape_offsets, dense_weights
    from oracle import oracle
    rng = np.random.default_rng(args.seed)
    G, T = args.batch_grids, 100
    sides = np.exp(rng.uniform(np.log(128), np.log(2048), (G, 2))).astype(int)
    names = rng.choice(["box2d1r", "star2d3r"], G)
    grids = []
    for i in range(G):
        r = 1 if names[i] == "box2d1r" else 3
        L = Layout.with_halo([int(sides[i, 0]), int(sides[i, 1])], r)
        offs = shape_offsets(names[i])
        w = cs.fill_weights_host(len(offs), args.seed + i)
        grids.append((L, r, offs, w, dense_weights(2, r, offs, w)))
    pts = sum(g[0].points for g in grids)
    seeds = [args.seed + 1000 + i for i in range(G)]
    res = {}
    finals = {}
    precs = ("fp64", "bf16x3", "fp16x2")
    for prec in precs:
        with Batch([g[0] for g in grids], [g[1] for g in grids], [g[4] for g in grids], precision=prec,
                   use_graph=True) as b:
            b.fill_random(seeds)
            rep = b.time(T, warmup=args.warmup, reps=args.steps)
            b.fill_random(seeds)
            b.run(T)
            finals[prec] = [b.download(i) for i in range(G)]
            res[prec] = {"gstencil_per_s": round(pts * T / (rep.elapsed_ms * 1e-3) / 1e9, 2),
                         "ms_per_run": round(rep.elapsed_ms, 3), "describe": b.describe()}
    threads = oracle.max_threads()
    errs = {p: [] for p in precs}
    t0 = time.perf_counter()
    for i, (L, r, offs, w, dense) in enumerate(grids):
        want = L.interior(oracle.run(L, offs, w, oracle.fill_data(L, seeds[i]), T, threads=threads))
        for prec in precs:
            got = L.interior(finals[prec][i])
            d = got - want
            errs[prec].append((float(np.max(np.abs(d)) / np.max(np.abs(want))),
                               float(np.sqrt(np.sum(d * d) / np.sum(want * want)))))
    oracle_s = time.perf_counter() - t0
    res["fp64"]["max_rel_max_err"] = max(e[0] for e in errs["fp64"])
    for prec in precs[1:]:
        res[prec]["max_rel_max_err"] = max(e[0] for e in errs[prec])
        res[prec]["max_rel_l2_err"] = max(e[1] for e in errs[prec])
        # the split bar: relative L2 <= 1e-6 (and relative max <= 5e-6) on every grid
        res[prec]["within_bar"] = bool(res[prec]["max_rel_l2_err"] <= 1e-6 and res[prec]["max_rel_max_err"] <= 5e-6)
    res["fp64"]["within_bar"] = bool(res["fp64"]["max_rel_max_err"] <= 1e-12)
    out = {
        "metric": "GStencil/s at 1 GPU: FP64 MFMA vs split-precision at stated error; % roofline",
        "value": res["fp64"]["gstencil_per_s"], "unit": "GStencil/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": res["fp64"]["ms_per_run"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic (SplitMix64 seeded grids and weights)",
        "config": {"workload": f"batch of {G} 2-D grids, sides log-uniform 128..2048, box2d1r/star2d3r, T={T}",
                   "points": int(pts), "box2d1r_grids": int(np.sum(names == "box2d1r"))},
        "batch": res,
        "oracle_check": {"grids": G, "threads": threads, "seconds": round(oracle_s, 1)},
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="box2d1r", choices=sorted(WORKLOADS) + ["batch"])
    ap.add_argument("--batch-grids", type=int, default=256)
    ap.add_argument("--fusion", type=int, default=0, help="fusion depth (0 = library default)")
    ap.add_argument("--variant", type=int, default=0)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="library defaults instead of the autotuner's choice")
    ap.add_argument("--no-extras", action="store_true", help="skip the unfused and split-precision lines")
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "bf16x3", "fp16x2", "fp16w3", "fp32x2"])
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="N",
                    help="only run N plain device copies of 2^27 doubles (reference for HBM byte counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import ctypes
        from convstencil_amd import _lib
        lib = _lib.load()
        lib.cs_debug_copy_bench.restype = ctypes.c_int32
        lib.cs_debug_copy_bench.argtypes = [ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]
        ms = ctypes.c_double(0)
        rc = lib.cs_debug_copy_bench(1 << 27, args.copy_calibration, ctypes.byref(ms))
        print(json.dumps({"copy_doubles": 1 << 27, "copy_bytes_read": 8 << 27, "copy_bytes_written": 8 << 27,
                          "ms_per_copy": ms.value, "rc": rc}), flush=True)
        sys.exit(rc)

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this build runs on exactly one GPU; multi-GPU runs are not part of it "
              "(the sharding design is described in DESIGN.md).", file=sys.stderr)
        sys.exit(2)

    if args.workload == "batch":
        batch_workload(args)
        return

    import convstencil_amd as cs
    from convstencil_amd import Layout, Plan, shape_offsets, dense_weights, SHAPES

    shape, sizes, halo, T = WORKLOADS[args.workload]
    ndim, r, kind = SHAPES[shape]
    L = Layout.with_halo(sizes, halo)
    offs = shape_offsets(shape)
    w = cs.fill_weights_host(len(offs), args.seed)
    dense = dense_weights(ndim, r, offs, w)

    tune = 0 if (args.variant or args.fusion or args.no_tune) else 1
    plan = Plan(L, r, dense, fusion_depth=args.fusion, variant=args.variant, use_graph=not args.no_graph,
                precision=args.precision, autotune=tune)
    plan.fill_random(args.seed)
    desc = plan.describe()
    # warmup + timed region: HIP events on the plan stream around exactly K runs of T steps
    t0 = time.perf_counter()
    rep = plan.time(T, warmup=args.warmup, reps=args.steps)
    wall = time.perf_counter() - t0
    ms = rep.elapsed_ms
    pts = L.points
    value = pts * T / (ms * 1e-3) / 1e9
    launches = rep.launches
    ms_launch = ms / launches
    kf = rep.fusion_depth
    alg_bytes = 16.0 * pts
    alg_flops = 2.0 * len(offs) * pts * kf
    hbm = alg_bytes / (ms_launch * 1e-3) / 1e9
    tfl = alg_flops / (ms_launch * 1e-3) / 1e12
    frac_hbm, frac_fl = hbm / HBM_PEAK_GBS, tfl / FP64_PEAK_TFLOPS
    kernel = desc.split()[1].split("=", 1)[1]
    traffic = load_traffic(f"{args.workload}:{kernel}")
    if frac_hbm >= frac_fl:
        roof = {"bound": "hbm", "achieved": round(hbm, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(frac_hbm, 4), "traffic": traffic}
    else:
        roof = {"bound": "mfma", "achieved": round(tfl, 2), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(frac_fl, 4), "traffic": traffic}
    roof.update({"kernel": kernel, "launches_per_step": launches,
                 "ms_per_launch": round(ms_launch, 5), "hbm_gbs_alg": round(hbm, 1), "fp64_tflops_alg": round(tfl, 2),
                 "alg_bytes_per_launch": alg_bytes, "alg_flops_per_launch": alg_flops})
    plan.close()

    extras = {}
    if not args.no_extras and args.workload == "box2d1r" and args.precision == "fp64":
        extras["unfused"] = unfused_hbm(L, r, dense, T, args.seed, not args.no_graph)
        extras["split_box2d3r"] = split_comparison(args.seed, not args.no_graph)
    if not args.no_extras and ndim == 3 and args.precision == "fp64":
        extras[f"split_{shape}"] = split_comparison(args.seed, not args.no_graph, shape, sizes, T)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(L, offs, w, args.seed, args.cpu_seconds)

    out = {
        "metric": "GStencil/s at 1 GPU: FP64 MFMA vs split-precision at stated error; % roofline",
        "value": round(value, 2),
        "unit": "GStencil/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else f"f64 state, {args.precision} products, f32 accumulation",
        "data": "synthetic (SplitMix64 seeded grid and weights)",
        "config": {"workload": f"{shape} {'x'.join(map(str, sizes))} halo {halo} T={T} {args.precision}",
                   "time_steps": T,
                   "fusion_depth": kf, "plan": desc, "wall_s": round(wall, 3)},
        "roofline": roof,
        "cpu_baseline": cpu,
        "extras": extras,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
