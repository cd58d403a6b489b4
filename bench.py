#!/usr/bin/env python3
"""Benchmark of the ConvStencil hot path on one MI355X.

Workload (default): box2d1r (9-point) stencil, FP64, 10240 x 10240 interior
with a zero halo of width 1, T = 64 time steps, seeded synthetic input (interior
uniform in [-1, 1), weights uniform in (0, 1) normalised to sum 1).  One bench
"step" is one full 64-time-step run over the grid, with the grid resident in
HBM; GStencil/s counts interior points x time steps / s, as the reference does
(src/2d/gpu.cu:329).

Prints one JSON line.  `roofline` describes the dominant kernel: algorithmic
bytes (8 B read + 8 B written per interior point per launch) over its mean
launch time, measured with HIP events on the plan stream around a run made of
that kernel's launches only; beside it the executed MFMA work (or, for the
vector-FMA kernels, the executed FP64 VALU work) per launch from the committed
rocprofv3 counter summary (profiles/counters.json) over the same launch time,
against the FP64 peak, and `traffic`, the measured HBM bytes
per launch (profiles/traffic.json), or null.  `cpu_baseline` is the CPU oracle
(the reference's loop, restated) timed on this host's cores on a bounded sample
of the same workload.  `extras` carries every other configuration of the
workload set (star2d3r / box2d3r 8192^2, box3d1r / star3d1r 512^3, 1d1r 2^20,
the batch of 256 grids), each FP64 and its split forms, with GStencil/s, the
dominant kernel's HBM fraction and the error against the CPU oracle (on
windows whose light cone the oracle recomputes exactly, or whole grids).
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
MFMA_PEAK_TFLOPS = {"fp64": 78.6, "bf16x3": 2500.0, "fp16x2": 2500.0, "fp16w3": 2500.0, "fp32x2": 157.3}
METRIC = "GStencil/s at 1 GPU: FP64 MFMA vs split-precision at stated error; % roofline"

WORKLOADS = {
    # name: (shape, interior sizes, halo, steps)
    "box2d1r": ("box2d1r", [10240, 10240], 1, 64),
    "star2d3r": ("star2d3r", [8192, 8192], 3, 64),
    "box2d3r": ("box2d3r", [8192, 8192], 3, 64),
    "box3d1r": ("box3d1r", [512, 512, 512], 1, 32),
    "star3d1r": ("star3d1r", [512, 512, 512], 1, 32),
    "1d1r": ("1d1r", [1 << 20], 1, 64),
}


def _profile_table(name):
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return {}
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def load_traffic(kernel_key):
    v = _profile_table("traffic.json").get(kernel_key)
    try:
        return None if v is None else float(v["bytes_per_launch"])
    except (KeyError, TypeError, ValueError):
        return None


def load_counters(kernel_key):
    """Per-launch counter figures of a kernel from profiles/counters.json (rocprofv3 --pmc
    runs of the same command), or None."""
    v = _profile_table("counters.json").get(kernel_key)
    return v if isinstance(v, dict) else None


def kernel_of(desc):
    return desc.split()[1].split("=", 1)[1]


# ------------------------------------------------------------------ errors against the oracle

def _windows(L, T, r, size):
    """Window origins (interior coordinates): the corners, an edge and the inside."""
    n = [L.n[d] for d in range(3 - L.ndim, 3)]
    size = min(size, *n)
    lo = [0] * len(n)
    hi = [x - size for x in n]
    mid = [x // 2 - size // 2 + 3 for x in n]
    mid = [min(max(m, 0), h) for m, h in zip(mid, hi)]
    if L.ndim == 2:
        return size, [(lo[0], lo[1]), (hi[0], hi[1]), (lo[0], mid[1]), (mid[0], mid[1])]
    return size, [tuple(lo), tuple(hi), (mid[0], lo[1], mid[2]), tuple(mid)]


def window_errors(L, offs, w, x, got, T, r, size=64, threads=0):
    """Relative max and L2 error of `got` (the final padded array) against the oracle on
    windows whose light cone (T * r cells each side) the oracle recomputes exactly; the
    errors are taken over all windows together.  1-D grids are checked whole."""
    from oracle import oracle
    from convstencil_amd import Layout
    if L.ndim == 1:
        want = L.interior(oracle.run(L, offs, w, x, T, threads=threads))
        g = L.interior(got)
        d = g - want
        return float(np.max(np.abs(d)) / np.max(np.abs(want))), float(np.sqrt(np.sum(d * d) / np.sum(want * want)))
    size, origins = _windows(L, T, r, size)
    nd = L.ndim
    n = [L.n[d] for d in range(3 - nd, 3)]
    xi = L.interior(x).reshape(n)
    gi = L.interior(got).reshape(n)
    num2 = den2 = 0.0
    dmax = wmax = 0.0
    cone = T * r
    for o in origins:
        a0 = [max(0, o[d] - cone) for d in range(nd)]
        a1 = [min(n[d], o[d] + size + cone) for d in range(nd)]
        sub = Layout.with_halo([a1[d] - a0[d] for d in range(nd)], r)
        xs = np.zeros(sub.cells)
        sl = tuple(slice(a0[d], a1[d]) for d in range(nd))
        sub.interior(xs)[...] = xi[sl].reshape(sub.interior(xs).shape)
        ys = sub.interior(oracle.run(sub, offs, w, xs, T, threads=threads)).reshape([a1[d] - a0[d] for d in range(nd)])
        ws = tuple(slice(o[d] - a0[d], o[d] - a0[d] + size) for d in range(nd))
        want = ys[ws]
        g = gi[tuple(slice(o[d], o[d] + size) for d in range(nd))]
        d = g - want
        num2 += float(np.sum(d * d))
        den2 += float(np.sum(want * want))
        dmax = max(dmax, float(np.max(np.abs(d))))
        wmax = max(wmax, float(np.max(np.abs(want))))
    return dmax / wmax, float(np.sqrt(num2 / den2))


# ------------------------------------------------------------------ one configuration

def measure(name, prec, seed, reps, warmup, graph=True, autotune=1, variant=0, fusion=0, check=True,
            dominant=True):
    """One workload in one precision: GStencil/s over the whole T-step run (HIP events on the
    plan stream), the dominant kernel's own launch time (a run of its launches only) and HBM
    fraction, and the error of the final state against the CPU oracle."""
    import convstencil_amd as cs
    from convstencil_amd import Layout, Plan, shape_offsets, dense_weights, SHAPES
    shape, sizes, halo, T = WORKLOADS[name]
    ndim, r, kind = SHAPES[shape]
    L = Layout.with_halo(sizes, halo)
    offs = shape_offsets(shape)
    w = cs.fill_weights_host(len(offs), seed)
    dense = dense_weights(ndim, r, offs, w)
    res = {}
    with Plan(L, r, dense, precision=prec, use_graph=graph, autotune=autotune, variant=variant,
              fusion_depth=fusion) as p:
        p.fill_random(seed)
        desc = p.describe()
        t0 = time.perf_counter()
        rep = p.time(T, warmup=warmup, reps=reps)
        res["wall_s"] = round(time.perf_counter() - t0, 3)
        kf = rep.fusion_depth
        res.update({"gstencil_per_s": round(L.points * T / (rep.elapsed_ms * 1e-3) / 1e9, 2),
                    "ms_per_run": round(rep.elapsed_ms, 4), "launches_per_run": rep.launches,
                    "kernel": kernel_of(desc), "fusion_depth": kf, "plan": desc})
        if dominant:
            # the dominant kernel alone: a run of kf * m steps is m launches of it
            m = max(1, T // kf)
            dr = p.time(kf * m, warmup=1, reps=reps)
            ms_launch = dr.elapsed_ms / dr.launches
            gbs = 16.0 * L.points / (ms_launch * 1e-3) / 1e9
            res.update({"ms_per_launch": round(ms_launch, 5), "hbm_gbs_alg": round(gbs, 1),
                        "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)})
        if check:
            from convstencil_amd import fill_random_host
            p.fill_random(seed)
            p.run(T)
            got = p.download()
            x = fill_random_host(L, seed)
            t0 = time.perf_counter()
            rm, rl2 = window_errors(L, offs, w, x, got, T, r)
            res.update({"rel_max_err": rm, "rel_l2_err": rl2, "oracle_s": round(time.perf_counter() - t0, 2),
                        "checked_on": "whole grid" if ndim == 1 else "light-cone windows (64 cells, 4 windows)"})
            bar = (rm <= 1e-12) if prec == "fp64" else (rl2 <= 1e-6 and rm <= 5e-6)
            res["within_bar"] = bool(bar)
    return res, (L, r, offs, w, dense, T)


def with_speedups(entry):
    base = entry.get("fp64", {}).get("gstencil_per_s")
    for prec, v in list(entry.items()):
        if prec != "fp64" and isinstance(v, dict) and base:
            v["speedup_over_fp64"] = round(v["gstencil_per_s"] / base, 3)
    return entry


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


# ------------------------------------------------------------------ the batched workload

def batch_grids(seed, G):
    import convstencil_amd as cs
    from convstencil_amd import Layout, shape_offsets, dense_weights
    rng = np.random.default_rng(seed)
    sides = np.exp(rng.uniform(np.log(128), np.log(2048), (G, 2))).astype(int)
    names = rng.choice(["box2d1r", "star2d3r"], G)
    grids = []
    for i in range(G):
        r = 1 if names[i] == "box2d1r" else 3
        L = Layout.with_halo([int(sides[i, 0]), int(sides[i, 1])], r)
        offs = shape_offsets(names[i])
        w = cs.fill_weights_host(len(offs), seed + i)
        grids.append((L, r, offs, w, dense_weights(2, r, offs, w)))
    return grids, names


def batch_measure(seed, G, reps, warmup, precs=("fp64", "bf16x3", "fp16x2"), check_every=1):
    """256 independent 2-D grids, sides log-uniform in [128, 2048] (uneven, non-square),
    stencil drawn from {box2d1r, star2d3r}, own seeded weights, zero halo of the radius,
    T = 100; all grids advance in one persistent graph-captured launch per run.  Aggregate
    GStencil/s and, per grid (every `check_every`-th grid), the error against the oracle."""
    from convstencil_amd import Batch, fill_random_host
    from oracle import oracle
    T = 100
    grids, names = batch_grids(seed, G)
    pts = sum(g[0].points for g in grids)
    seeds = [seed + 1000 + i for i in range(G)]
    res, finals = {}, {}
    for prec in precs:
        with Batch([g[0] for g in grids], [g[1] for g in grids], [g[4] for g in grids], precision=prec,
                   use_graph=True) as b:
            b.fill_random(seeds)
            rep = b.time(T, warmup=warmup, reps=reps)
            b.fill_random(seeds)
            b.run(T)
            finals[prec] = {i: b.download(i) for i in range(0, G, check_every)}
            res[prec] = {"gstencil_per_s": round(pts * T / (rep.elapsed_ms * 1e-3) / 1e9, 2),
                         "ms_per_run": round(rep.elapsed_ms, 3), "describe": b.describe()}
    threads = oracle.max_threads()
    errs = {p: [] for p in precs}
    t0 = time.perf_counter()
    for i in range(0, G, check_every):
        L, r, offs, w, dense = grids[i]
        want = L.interior(oracle.run(L, offs, w, fill_random_host(L, seeds[i]), T, threads=threads))
        for prec in precs:
            d = L.interior(finals[prec][i]) - want
            errs[prec].append((float(np.max(np.abs(d)) / np.max(np.abs(want))),
                               float(np.sqrt(np.sum(d * d) / np.sum(want * want)))))
    for prec in precs:
        res[prec]["max_rel_max_err"] = max(e[0] for e in errs[prec])
        res[prec]["max_rel_l2_err"] = max(e[1] for e in errs[prec])
        if prec == "fp64":
            res[prec]["within_bar"] = bool(res[prec]["max_rel_max_err"] <= 1e-12)
        else:  # the split bar: relative L2 <= 1e-6 (and relative max <= 5e-6) on every grid
            res[prec]["within_bar"] = bool(res[prec]["max_rel_l2_err"] <= 1e-6 and res[prec]["max_rel_max_err"] <= 5e-6)
    info = {"grids": G, "points": int(pts), "box2d1r_grids": int(np.sum(names == "box2d1r")), "time_steps": T,
            "oracle_checked_grids": len(range(0, G, check_every)), "oracle_threads": threads,
            "oracle_s": round(time.perf_counter() - t0, 1)}
    return with_speedups(res), info


def batch_workload(args):
    res, info = batch_measure(args.seed, args.batch_grids, args.steps, args.warmup)
    out = {
        "metric": METRIC, "value": res["fp64"]["gstencil_per_s"], "unit": "GStencil/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": res["fp64"]["ms_per_run"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic (SplitMix64 seeded grids and weights)",
        "config": {"workload": f"batch of {info['grids']} 2-D grids, sides log-uniform 128..2048, box2d1r/star2d3r, "
                               f"T={info['time_steps']}", **info},
        "batch": res,
    }
    print(json.dumps(out), flush=True)


# ------------------------------------------------------------------ the other configurations

def extras_all(args):
    """Every other configuration of the workload set, bounded in time: the single-step
    kernel of the headline grid (the streaming roofline), box2d3r and star2d3r 8192^2 in
    FP64 and the split forms, box3d1r / star3d1r 512^3 in FP64 and BF16x3 / FP16 w3,
    1d1r 2^20, and the batch of 256 grids (errors checked on every 8th grid).  FP64 and
    box2d3r's forms are autotuned (as the headline); the rest run the library's defaults."""
    ex = {}
    g = not args.no_graph
    u, _ = measure("box2d1r", "fp64", args.seed, reps=5, warmup=2, graph=g, autotune=0, fusion=1, check=False)
    ex["box2d1r_single_step"] = {k: u[k] for k in ("kernel", "gstencil_per_s", "ms_per_launch", "hbm_gbs_alg",
                                                   "hbm_frac")}
    for name, precs, tune in (("box2d3r", ("fp64", "bf16x3", "fp16w3", "fp16x2", "fp32x2"), 1),
                              ("star2d3r", ("fp64", "bf16x3", "fp16x2"), 0),
                              ("box3d1r", ("fp64", "bf16x3", "fp16w3"), 0),
                              ("star3d1r", ("fp64", "bf16x3", "fp16w3"), 0),
                              ("1d1r", ("fp64",), 0)):
        entry = {}
        for prec in precs:
            r, _ = measure(name, prec, args.seed, reps=5, warmup=2, graph=g,
                           autotune=tune if prec != "fp32x2" else 0)
            entry[prec] = {k: r[k] for k in ("gstencil_per_s", "kernel", "fusion_depth", "ms_per_launch", "hbm_frac",
                                             "rel_max_err", "rel_l2_err", "within_bar")}
        shape, sizes, halo, T = WORKLOADS[name]
        entry["workload"] = f"{shape} {'x'.join(map(str, sizes))} halo {halo} T={T}"
        ex[name] = with_speedups(entry)
    br, info = batch_measure(args.seed, args.batch_grids, reps=2, warmup=1, check_every=8)
    ex["batch"] = {**{p: {k: v[k] for k in ("gstencil_per_s", "max_rel_max_err", "max_rel_l2_err", "within_bar",
                                            "speedup_over_fp64") if k in v} for p, v in br.items()}, **info}
    return ex


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="box2d1r", choices=sorted(WORKLOADS) + ["batch"])
    ap.add_argument("--batch-grids", type=int, default=256)
    ap.add_argument("--fusion", type=int, default=0, help="fusion depth (0 = library default)")
    ap.add_argument("--variant", type=int, default=0)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="library defaults instead of the autotuner's choice")
    ap.add_argument("--no-extras", action="store_true", help="only the headline workload")
    ap.add_argument("--no-check", action="store_true", help="skip the oracle check of the headline workload")
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

    tune = 0 if (args.variant or args.fusion or args.no_tune) else 1
    res, (L, r, offs, w, dense, T) = measure(args.workload, args.precision, args.seed, reps=args.steps,
                                             warmup=args.warmup, graph=not args.no_graph, autotune=tune,
                                             variant=args.variant, fusion=args.fusion, check=not args.no_check)
    kernel, kf = res["kernel"], res["fusion_depth"]
    ms_launch = res["ms_per_launch"]
    hbm = res["hbm_gbs_alg"]
    alg_flops = 2.0 * len(offs) * L.points * kf
    roof = {"bound": "hbm", "achieved": hbm, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": res["hbm_frac"],
            "traffic": load_traffic(f"{args.workload}:{kernel}"), "kernel": kernel,
            "ms_per_launch": ms_launch, "alg_bytes_per_launch": 16.0 * L.points,
            "alg_flops_per_launch": alg_flops,
            "alg_tflops": round(alg_flops / (ms_launch * 1e-3) / 1e12, 2)}
    cnt = load_counters(f"{args.workload}:{kernel}")
    if cnt and cnt.get("mfma_flops_per_launch"):
        peak = MFMA_PEAK_TFLOPS[args.precision]
        ach = float(cnt["mfma_flops_per_launch"]) / (ms_launch * 1e-3) / 1e12
        roof["mfma"] = {"executed_tflops": round(ach, 2), "peak_tflops": peak, "frac": round(ach / peak, 4),
                        "executed_flops_per_launch": cnt["mfma_flops_per_launch"], "source": cnt.get("source")}
        for k in ("lds_bank_conflict_frac", "valu_insts_per_launch", "l2_hit_frac"):
            if k in cnt:
                roof[k] = cnt[k]
    if cnt and cnt.get("fp64_valu_flops_per_launch"):
        # vector-FMA kernels: executed FP64 VALU work against the FP64 vector peak
        ach = float(cnt["fp64_valu_flops_per_launch"]) / (ms_launch * 1e-3) / 1e12
        roof["fp64_valu"] = {"executed_tflops": round(ach, 2), "peak_tflops": FP64_PEAK_TFLOPS,
                             "frac": round(ach / FP64_PEAK_TFLOPS, 4),
                             "executed_flops_per_launch": cnt["fp64_valu_flops_per_launch"], "source": cnt.get("source")}
        for k in ("valu_insts_per_launch", "l2_hit_frac"):
            if k in cnt:
                roof[k] = cnt[k]

    extras = {}
    if not args.no_extras and args.workload == "box2d1r" and args.precision == "fp64":
        extras = extras_all(args)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(L, offs, w, args.seed, args.cpu_seconds)

    shape, sizes, halo, T = WORKLOADS[args.workload]
    out = {
        "metric": METRIC,
        "value": res["gstencil_per_s"],
        "unit": "GStencil/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["ms_per_run"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else f"f64 state, {args.precision} products, f32 accumulation",
        "data": "synthetic (SplitMix64 seeded grid and weights)",
        "config": {"workload": f"{shape} {'x'.join(map(str, sizes))} halo {halo} T={T} {args.precision}",
                   "time_steps": T, "fusion_depth": kf, "launches_per_step": res["launches_per_run"],
                   "plan": res["plan"], "wall_s": res["wall_s"]},
        "roofline": roof,
        "parity": {k: res[k] for k in ("rel_max_err", "rel_l2_err", "within_bar", "checked_on") if k in res},
        "cpu_baseline": cpu,
        "extras": extras,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
