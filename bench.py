"""Benchmark: point-evaluations per second of the Smolyak interpolant on one MI355X.

Default workload (the headline configuration): d_in = 1000, d_out = 1, Leja nodes on [-1, 1],
k_j = log((2 + j) / log 2), t tuned so |Lambda| ~ 1e4 nodes, f(x) = exp(-sum_j x_j / (j+1)^2),
16384 points uniform on [-1, 1]^1000 from seed 0.  One step = one evaluation of all 16384 points
(inputs already resident in HBM).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config NAME] [--no-cpu]

Prints one JSON line.  `value` is the headline path's throughput; the plain FP64 barycentric path
is measured beside it, with the error of every path against the FP64 CPU oracle on a sample.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="c2_exp_d1000")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-points", type=int, default=1024)
    ap.add_argument("--oracle-points", type=int, default=256)
    return ap.parse_args()


def timed_loop(fn, steps, warmup, torch):
    """Warm up, then time `steps` calls; also per-call HIP event durations on the current stream."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for a, b in evs:
        a.record(stream)
        fn()
        b.record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ev_ms = [a.elapsed_time(b) for a, b in evs]
    return (t1 - t0), float(np.mean(ev_ms)), float(np.median(ev_ms))


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    from oracle import indices as OI
    from oracle import nodes as ON
    from oracle import smolyak as OS
    from oracle import cpu_baseline as CB
    from smolyax_amd import nodes as SN
    from smolyax_amd import workloads as W
    from smolyax_amd.interpolation import SmolyakBarycentricInterpolator
    from smolyax_amd import indices as SI

    cfg = W.CONFIGS[args.config]
    d, d_out, m, P = cfg["d_in"], cfg["d_out"], cfg["m"], cfg["P"]
    f = cfg["f"]()
    k = W.k_rule(d)
    ts = time.perf_counter()
    t = SI.find_approximate_threshold(k, m, True)
    ip = SmolyakBarycentricInterpolator(d_out=d_out, node_gen=SN.Leja(dim=d), k=k, t=t)
    ip.set_f(f=lambda X: f(X), vectorized=True)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - ts
    L = ip.layout
    x_np = W.points_uniform(P, d, seed=0)
    x = torch.as_tensor(x_np, device="cuda")
    y = torch.empty((P, d_out), dtype=torch.float64, device="cuda")

    from smolyax_amd import _native as N
    import ctypes

    stream_ptr = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def make_step(fn_name, plan):
        fn = N.fn[fn_name]

        def step():
            rc = fn(plan, x.data_ptr(), P, x.stride(0), y.data_ptr(), stream_ptr)
            if rc != 0:
                N.check(rc, fn_name)
        return step

    paths = {"mixed": ("sbi_newton_eval_mixed", ip._nplan),
             "newton-fp64": ("sbi_newton_eval_f64", ip._nplan),
             "fp64": ("sbi_bary_eval_f64", ip._plan)}
    results = {}
    outputs = {}
    for name, (fn_name, plan) in paths.items():
        wall, ev_mean_ms, ev_med_ms = timed_loop(make_step(fn_name, plan), args.steps, args.warmup,
                                                 torch)
        outputs[name] = y.cpu().numpy().copy()
        results[name] = {"value": P * args.steps / wall, "ms_per_step": wall / args.steps * 1e3,
                         "kernel_ms": ev_mean_ms, "kernel_ms_median": ev_med_ms}

    # accuracy against the FP64 CPU oracle on a sample of the same points, and of every path
    # against the FP64 GPU path on all points
    no = min(args.oracle_points, P)
    ref = OS.Interpolant(d_out=d_out, node_gen=ON.leja(dim=d), k=k, t=t, f=f)
    y_ref = ref(x_np[:no])

    def errs(a, b):
        diff = a - b
        return {"max_rel": float(np.max(np.abs(diff)) / np.max(np.abs(b))),
                "rel_rmse": float(np.sqrt(np.mean(diff ** 2) / np.mean(b ** 2)))}

    accuracy = {}
    for name in paths:
        accuracy[name] = {"vs_oracle": errs(outputs[name][:no], y_ref),
                          "vs_gpu_fp64_all_points": errs(outputs[name], outputs["fp64"])}
    accuracy["oracle_points"] = int(no)

    # rooflines.  Mixed (headline) kernel: HBM bytes per launch = x (FP64) + surpluses (FP32)
    # + parent rows (int32) + y; the FP64 barycentric kernel: 2 * d_out * K_raw flops per point.
    NL = ip._nlayout
    t_mixed = results["mixed"]["kernel_ms"] * 1e-3
    bytes_mixed = 8.0 * P * d + (4.0 + 4.0) * NL.n_nodes + 8.0 * P * d_out
    roofline = {"bound": "hbm", "achieved": bytes_mixed / t_mixed / 1e9, "peak": PEAK_HBM_GBS,
                "unit": "GB/s", "frac": bytes_mixed / t_mixed / 1e9 / PEAK_HBM_GBS, "traffic": None,
                "kernel": "newton_eval_kernel<float>", "bytes_per_launch": bytes_mixed,
                "avg_launch_ms": results["mixed"]["kernel_ms"]}
    t_b = results["fp64"]["kernel_ms"] * 1e-3
    flops = 2.0 * d_out * L.k_raw * P
    roof_fp64 = {"bound": "valu", "kernel": "bary_eval_kernel<1>", "achieved": flops / t_b / 1e12,
                 "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                 "frac": flops / t_b / 1e12 / PEAK_FP64_TFLOPS, "avg_launch_ms": results["fp64"]["kernel_ms"]}

    cpu = None
    if not args.no_cpu:
        import torch as _t

        _t.set_num_threads(CB.threads())
        tc = CB.TorchCPUInterpolant(ref)
        nc = min(args.cpu_points, P)
        tc(x_np[:64])
        reps = 3
        c0 = time.perf_counter()
        for _ in range(reps):
            tc(x_np[:nc])
        c1 = time.perf_counter()
        cpu = {"value": nc * reps / (c1 - c0), "unit": "point-evals/s", "cores": CB.threads(),
               "kind": "port",
               "sample": f"{nc} of the {P} points x {reps} calls, FP64 torch-CPU restatement of the "
                         f"reference's padded-group einsum evaluation"}

    value = results["mixed"]["value"]
    out = {
        "metric": "point-evals/s, d_in=1e3 n=1e4 (mixed vs FP64, <=1e-6 rel); serve p50/p99 ms",
        "value": value,
        "unit": "point-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": results["mixed"]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 products / f64 accumulation",
        "data": "synthetic",
        "config": {"workload": args.config, "d_in": d, "d_out": d_out, "points": P,
                   "nodes": int(ip.n_f_evals), "t": t, "k_raw": int(L.k_raw),
                   "blocks": int(len(L.blk_n)), "internal_nodes": int(NL.n_internal),
                   "path": "mixed (hierarchical Newton form)"},
        "roofline": roofline,
        "roofline_fp64": roof_fp64,
        "cpu_baseline": cpu,
        "paths": results,
        "speedup_mixed_vs_fp64": results["mixed"]["value"] / results["fp64"]["value"],
        "accuracy": accuracy,
        "setup_s": setup_s,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
