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
    ap.add_argument("--cpu-points", type=int, default=8192)
    ap.add_argument("--oracle-points", type=int, default=256)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--no-legs", action="store_true",
                    help="skip the c3_sin_d100_o4096 and c4_tanh_d200_o256 legs of the default run")
    ap.add_argument("--serve-requests", type=int, default=600)
    return ap.parse_args()


TRAFFIC_FILES = {(k, "c2_exp_d1000"): "profiles/newton_mixed_mfma_c2_d1000_traffic.md"
                 for k in ("newton_mixed_mfma_kernel", "bary_eval_prog_kernel",
                           "newton_eval_kernel<double>", "newton_f64_mfma_kernel")}


def measured_traffic(kernel, config):
    """HBM bytes per launch of `kernel` on `config` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (copy-calibrated), or None when no summary covers it."""
    rel = TRAFFIC_FILES.get((kernel, config))
    path = os.path.join(ROOT, rel) if rel else None
    if not path or not os.path.exists(path):
        return None
    key = f"traffic_bytes_per_launch {kernel}:"
    with open(path) as fh:
        for line in fh:
            if line.startswith(key):
                return float(line.split(":")[1])
    return None


def timed_loop(fn, steps, warmup, torch):
    """Warm up, then time `steps` calls between two synchronisations; HIP events recorded on the
    launch stream at the start and the end of the timed region give the average device time per
    launch (no event between launches, so the launches queue back to back)."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    return (t1 - t0), e0.elapsed_time(e1) / steps


PEAK_FP32_TFLOPS = 157.3
PEAK_BF16_TFLOPS = 2500.0


def contraction_times(ip, x, torch, steps):
    """Average device time of the contraction kernel alone (HIP events on the launch stream over
    `steps` back-to-back launches) for each GEMM path, on basis matrices computed once:
    mixed = sbi_contract_split_k4 (FP64 head + FP32 rows), bf16 = sbi_contract_bf16_k8 (FP64 head
    + BF16 rows), newton-fp64 = sbi_contract (FP64)."""
    import ctypes

    from smolyax_amd import _native as N

    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    NL = ip._nlayout
    P, n, Nn, nh = x.shape[0], ip.d_out, NL.basis_n, NL.n_head
    y = torch.empty((P, n), dtype=torch.float64, device="cuda")
    H64 = torch.empty((max(nh, 1), P), dtype=torch.float64, device="cuda")
    N.check(N.fn["sbi_newton_rows_f64"](ip._nplan, x.data_ptr(), P, x.stride(0),
                                        ip._head_fac.data_ptr(), nh, H64.data_ptr(), P, st), "rows")
    S64 = ip._surpluses("f64", gemm=True)
    out = {}
    for name, SQ, q, dt, code, fn in (
            ("mixed", ip._surpluses_k4(nh), 4, torch.float32, 3, "sbi_contract_split_k4"),
            ("bf16", ip._surpluses_k8(nh), 8, torch.bfloat16, 4, "sbi_contract_bf16_k8")):
        KQ = SQ.shape[0] * q
        HT = torch.empty((KQ // q, P, q), dtype=dt, device="cuda")
        N.check(N.fn["sbi_newton_basis_rows"](ip._nplan, x.data_ptr(), P, x.stride(0), nh, Nn,
                                              HT.data_ptr(), P, code, st), "basis")

        def run(fn=fn, HT=HT, SQ=SQ, KQ=KQ):
            N.check(N.fn[fn](H64.data_ptr(), P, S64.data_ptr(), n, nh, HT.data_ptr(), P,
                             SQ.data_ptr(), n, KQ, None, y.data_ptr(), n, P, n, st), fn)
        _, ms = timed_loop(run, steps, 2, torch)
        out[name] = {"kernel": fn, "ms": ms, "k_rows": KQ + nh}
        del HT
    HT = torch.empty((Nn, P), dtype=torch.float64, device="cuda")
    N.check(N.fn["sbi_newton_basis_rows"](ip._nplan, x.data_ptr(), P, x.stride(0), 0, Nn,
                                          HT.data_ptr(), P, 0, st), "basis")

    def run64():
        ip._contract(HT, S64, None, y)
    _, ms = timed_loop(run64, max(2, steps // 2), 1, torch)
    out["newton-fp64"] = {"kernel": "sbi_contract (FP64)", "ms": ms, "k_rows": Nn}
    return out


def bench_gemm_leg(args, torch, W, name, cpu=True, log=None):
    """A large-d_out workload (c3_sin_d100_o4096 or c4_tanh_d200_o256): every path is a basis kernel + one contraction on
    the matrix cores.  Reports each path's throughput (inputs resident, whole step), the
    contraction kernel's roofline on the matrix cores, accuracy against the FP64 oracle and the
    exact interpolant (the oracle's extended precision) on a sample of points, the CPU baseline
    on a sample, and for c4_tanh_d200_o256 the adjoint (vjp) in mixed and FP64 with its own accuracy."""
    from oracle import cpu_baseline as CB
    from oracle import nodes as ON
    from oracle import smolyak as OS
    from smolyax_amd import indices as SI
    from smolyax_amd import nodes as SN
    from smolyax_amd.interpolation import SmolyakBarycentricInterpolator

    log = log or _log
    cfg = W.CONFIGS[name]
    d, d_out, m, P = cfg["d_in"], cfg["d_out"], cfg["m"], cfg["P"]
    f = cfg["f"]()
    k = W.k_rule(d)
    ts = time.perf_counter()
    t = SI.find_approximate_threshold(k, m, True)
    ip = SmolyakBarycentricInterpolator(d_out=d_out, node_gen=SN.Leja(dim=d), k=k, t=t,
                                        memory_limit=16.0)
    ip.set_f(f=lambda X: f(X), vectorized=True)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - ts
    log(f"[bench] {name}: set up in {setup_s:.1f} s")
    x_np = W.points_uniform(P, d, seed=0)
    x = torch.as_tensor(x_np, device="cuda")
    results, outputs = {}, {}
    for pname in ("bf16", "mixed", "newton-fp64", "fp64"):
        steps = args.steps if pname != "fp64" else max(2, args.steps // 4)
        box = {}

        def step():
            box["y"] = ip.evaluate(x, precision=pname)

        wall, ev_ms = timed_loop(step, steps, min(args.warmup, 2), torch)
        outputs[pname] = box["y"]
        results[pname] = {"value": P * steps / wall, "ms_per_step": wall / steps * 1e3,
                          "device_ms_per_step": ev_ms}
        log(f"[bench] {name} {pname}: {results[pname]['ms_per_step']:.3f} ms per step")
    L, NL = ip.layout, ip._nlayout
    # contraction kernels alone, on the matrix cores: FLOP on the Newton terms (2 d_out |Lambda|
    # per point, the product's own work) and on the reference's terms (2 d_out K_raw per point)
    ct = contraction_times(ip, x, torch, args.steps)
    flop_newton = 2.0 * d_out * NL.basis_n * P
    flop_kraw = 2.0 * d_out * L.k_raw * P
    peaks = {"mixed": PEAK_FP32_TFLOPS, "bf16": PEAK_BF16_TFLOPS, "newton-fp64": PEAK_FP64_TFLOPS}
    roof = {}
    for pname, c in ct.items():
        ach = flop_newton / (c["ms"] * 1e-3) / 1e12
        roof[pname] = {"bound": "mfma", "kernel": c["kernel"], "avg_launch_ms": c["ms"],
                       "achieved": ach, "peak": peaks[pname], "unit": "TFLOP/s",
                       "frac": ach / peaks[pname], "flop_per_launch": flop_newton,
                       "flop_basis": "2 d_out |Lambda| per point (Newton terms)"}
    # accuracy on a sample: FP64 oracle (the reference's algorithm) and the exact interpolant
    no = 16
    cols = np.arange(d_out) if d_out <= 256 else \
        np.sort(np.random.default_rng(5).choice(d_out, 64, replace=False))
    fc = f if len(cols) == d_out else (lambda X: np.asarray(f(X))[..., cols])
    ref = OS.Interpolant(d_out=len(cols), node_gen=ON.leja(dim=d), k=k, t=t, f=fc)
    y64 = ref(x_np[:no])
    yex = np.asarray(ref(x_np[:no], dtype=np.longdouble), dtype=np.float64)

    def errs(a, b):
        diff = a - b
        return {"max_rel": float(np.max(np.abs(diff)) / np.max(np.abs(b))),
                "rel_rmse": float(np.sqrt(np.mean(diff ** 2) / np.mean(b ** 2)))}

    accuracy = {"points": no, "outputs": int(len(cols)),
                "oracle_fp64_vs_exact": errs(y64, yex)}
    for pname, yv in outputs.items():
        a = yv[:no].cpu().numpy()[:, cols]
        accuracy[pname] = {"vs_oracle": errs(a, y64), "vs_exact": errs(a, yex),
                           "vs_gpu_fp64_all_points": errs(yv.cpu().numpy(),
                                                          outputs["fp64"].cpu().numpy())}
    leg = {"config": {"workload": name, "d_in": d, "d_out": d_out, "points": P,
                      "nodes": int(ip.n_f_evals), "k_raw": int(L.k_raw), "t": t},
           "setup_s": setup_s, "paths": results,
           "speedup_mixed_vs_fp64": results["mixed"]["value"] / results["fp64"]["value"],
        "speedup_mixed_vs_newton_fp64": results["mixed"]["value"] / results["newton-fp64"]["value"],
           "speedup_mixed_vs_newton_fp64": results["mixed"]["value"]
           / results["newton-fp64"]["value"],
           "contraction_roofline": roof, "accuracy": accuracy,
           "path_tflops_kraw_basis": {pn: flop_kraw / (r["ms_per_step"] * 1e-3) / 1e12
                                      for pn, r in results.items()}}
    if cfg.get("vjp"):
        w = np.random.default_rng(2).standard_normal(d_out)
        v = torch.as_tensor(np.tile(w, (P, 1)), device="cuda")
        nov = 8
        Jex = np.asarray(ref.gradient(x_np[:nov], dtype=np.longdouble), dtype=np.float64)
        gex = np.einsum("m,pmj->pj", w, Jex)
        g64 = np.einsum("m,pmj->pj", w, ref.gradient(x_np[:nov]))
        vj = {}
        for pname in ("mixed", "newton-fp64"):
            box = {}

            def vstep():
                box["g"] = ip.vjp(x, v, precision=pname)

            wall, ev_ms = timed_loop(vstep, args.steps, min(args.warmup, 2), torch)
            g = box["g"][:nov].cpu().numpy()
            vj[pname] = {"value": P * args.steps / wall, "ms_per_step": wall / args.steps * 1e3,
                         "vs_exact": errs(g, gex), "vs_oracle": errs(g, g64)}
            log(f"[bench] {name} vjp {pname}: {vj[pname]['ms_per_step']:.3f} ms per step")
        vj["oracle_fp64_vs_exact"] = errs(g64, gex)
        vj["points_checked"] = nov
        leg["vjp"] = vj
    leg["gradient_integral"] = bench_gradient_integral(ip, x, torch)
    if cpu:
        import torch as _t

        _t.set_num_threads(CB.threads())
        full = ref if len(cols) == d_out else OS.Interpolant(
            d_out=d_out, node_gen=ON.leja(dim=d), k=k, t=t, f=lambda X: f(X))
        tc = CB.TorchCPUInterpolant(full)
        nc = 16 if d_out > 1024 else 512
        tc(x_np[:2])
        c0 = time.perf_counter()
        tc(x_np[:nc])
        c1 = time.perf_counter()
        leg["cpu_baseline"] = {
            "value": nc / (c1 - c0), "unit": "point-evals/s", "cores": CB.threads(),
            "kind": "port",
            "sample": f"{nc} of the {P} points, one call ({c1 - c0:.1f} s), FP64 torch-CPU "
                      f"restatement of the reference's padded-group einsum evaluation"}
        log(f"[bench] {name} cpu baseline: {leg['cpu_baseline']['value']:.3g} point-evals/s")
    else:
        leg["cpu_baseline"] = None
    del ip
    torch.cuda.empty_cache()
    return leg


def bench_gradient_integral(ip, x, torch, steps=10):
    """gradient (dense Jacobian, FP64) and integral (quadrature sum over F) through the C ABI.

    The Jacobian runs the path gradient() takes (the Newton form's sbi_newton_jacobian_f64 when
    the plan has one, else the barycentric kernel) on at most 4096 of the points, fewer when J
    would pass 1 GiB. Algorithmic work per point: J bytes 8 * d_out * d_in, and 2 * d_out FLOP
    per derivative-basis term (Newton: one term per (node, factor) pair; barycentric:
    sum_b n_b * nent_b); the roofline is the bound with the longer ideal time. The integral reads
    F (K_raw * d_out * 8 bytes) once: HBM roofline."""
    import ctypes

    import numpy as np

    from smolyax_amd import _native as N

    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    L = ip.layout
    d_in, d_out = ip.d_in, ip.d_out
    per_point = 8 * d_out * d_in
    Pg = max(64, min(int(x.shape[0]), 4096, (1 << 30) // per_point) // 64 * 64)
    Pg = min(Pg, int(x.shape[0]))
    xg = x[:Pg]
    J = torch.empty((Pg, d_out, d_in), dtype=torch.float64, device="cuda")
    bary_terms = float(np.dot(L.blk_n.astype(np.float64), np.diff(L.blk_foff).astype(np.float64)))
    if ip._nplan is not None:
        kernel = "sbi_newton_jacobian_f64"
        terms = float(ip._nlayout.n_vjp_ent)
        S = ip._S["f64"]
        nb = int(N.fn["sbi_newton_jacobian_workspace"](ip._nplan, Pg, d_out))
        ws = torch.empty(max(1, (nb + 7) // 8), dtype=torch.float64, device="cuda")

        def jac():
            N.check(N.fn["sbi_newton_jacobian_f64"](ip._nplan, xg.data_ptr(), Pg, xg.stride(0),
                                                    S.data_ptr(), S.stride(0), d_out, J.data_ptr(),
                                                    ws.data_ptr(), nb, st), "jacobian")
    else:
        kernel = "sbi_bary_jacobian_f64"
        terms = bary_terms

        def jac():
            N.check(N.fn["sbi_bary_jacobian_f64"](ip._plan, xg.data_ptr(), Pg, xg.stride(0),
                                                  J.data_ptr(), st), "jacobian")

    ws_i = int(N.fn["sbi_bary_integral_workspace"](ip._plan))
    part = torch.empty(ws_i, dtype=torch.float64, device="cuda")
    Q = torch.empty(d_out, dtype=torch.float64, device="cuda")

    def integ():
        N.check(N.fn["sbi_bary_integral_f64"](ip._plan, Q.data_ptr(), part.data_ptr(), st),
                "integral")

    out = {}
    try:
        t0 = time.perf_counter()
        jac()
        torch.cuda.synchronize()
        once = time.perf_counter() - t0
        js = max(1, min(steps, int(5.0 / max(once, 1e-6))))
        _log(f"[bench] jacobian ({kernel}): {Pg} points, first call {once * 1e3:.1f} ms, "
             f"{js} timed calls")
        _, jm = timed_loop(jac, js, 1, torch)
        flops = 2.0 * d_out * terms * Pg
        jbytes = 8.0 * Pg * d_out * d_in + 8.0 * Pg * d_in
        t_s = jm * 1e-3
        if jbytes / (PEAK_HBM_GBS * 1e9) >= flops / (PEAK_FP64_TFLOPS * 1e12):
            roof = {"bound": "hbm", "achieved": jbytes / t_s / 1e9, "peak": PEAK_HBM_GBS,
                    "unit": "GB/s"}
        else:
            roof = {"bound": "fp64", "achieved": flops / t_s / 1e12, "peak": PEAK_FP64_TFLOPS,
                    "unit": "TFLOP/s"}
        roof["frac"] = roof["achieved"] / roof["peak"]
        roof.update({"flop_per_point": 2.0 * d_out * terms, "bytes_per_point": jbytes / Pg})
        out["gradient"] = {"kernel": kernel, "points": Pg, "ms": jm,
                           "point_evals_per_s": Pg / t_s, "roofline": roof,
                           "barycentric_equivalent_tflops": 2.0 * d_out * bary_terms * Pg / t_s
                           / 1e12}
    except Exception as e:  # a documented limit of the Jacobian kernels
        out["gradient"] = {"error": str(e)}
    _log("[bench] integral")
    _, im = timed_loop(integ, steps, 2, torch)
    ib = 8.0 * L.k_raw * d_out
    out["integral"] = {"kernel": "bary_integral_partial_kernel + final", "ms": im,
                       "roofline": {"bound": "hbm", "achieved": ib / (im * 1e-3) / 1e9,
                                    "peak": PEAK_HBM_GBS, "unit": "GB/s",
                                    "frac": ib / (im * 1e-3) / 1e9 / PEAK_HBM_GBS}}
    return out


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


SERVE_QPS = 10000.0  # fixed Poisson arrival rate of the headline latency figures (requests/s)


def bench_serving(torch, W, n_requests=600, qps=SERVE_QPS, seed=0):
    """Serving workload: 32 resident Gauss-Hermite interpolants, Poisson arrivals, FP32.

    For each scheduler (Python: direct launches, HIP graphs, adaptive; the C++ scheduler of
    sbi_server_*): one untimed warm-up pass, the capacity (all requests submitted at once), then
    p50/p99 under Poisson arrivals at `qps`.  The adaptive and the C++ scheduler also report
    p50/p99 at 25 %, 50 % and 70 % of their capacity (latency against load), each after one
    untimed pass at that load (the first pass allocates buffers for its batch sizes).  The
    headline figures are the C++ scheduler's."""
    from smolyax_amd import serving

    ts = time.perf_counter()
    ips = W.build_serving_interpolants(W.serving_specs(32, seed=seed), progress=_log)
    names = sorted(ips)
    for n in names:  # first touch of every interpolant outside the timed run
        ips[n].evaluate(torch.zeros((4, ips[n].d_in), dtype=torch.float64, device="cuda"))
        torch.cuda.synchronize()
        _log(f"serve: first evaluation of {n} done")
    setup_s = time.perf_counter() - ts
    inputs = None

    def poisson(srv, q):
        nonlocal inputs
        srv.completed = []
        reqs, inputs = serving.run_poisson(srv, names, qps=q, n_requests=n_requests, seed=seed,
                                           inputs=inputs, time_limit_s=120.0, progress=_log)
        return srv.stats(reqs)

    def run(graphs, curve=False):
        if graphs == "native":
            srv = serving.NativeServer(ips, n_streams=4)
        else:
            srv = serving.Server(ips, precision="mixed", n_streams=4, graphs=graphs)
        if graphs not in (False, "native"):
            t0 = time.perf_counter()
            srv.prewarm()
            _log(f"serve: {sum(len(sl.graphs) for sl in srv.slots)} graphs captured in "
                 f"{time.perf_counter() - t0:.1f} s")
        # the first pass of a server pays one-time costs (device allocations of the batch buffers,
        # first replays) that a long-running server does not
        poisson(srv, 1e9)
        cap = poisson(srv, 1e9)  # closed loop: every request queued at once
        cap_qps = cap["requests"] / cap["span_s"]
        _log(f"serve (graphs={graphs}): capacity {cap_qps:.0f} requests/s")
        st = poisson(srv, qps)
        _log(f"serve (graphs={graphs}): {qps:.0f} requests/s: p50 {st['p50_ms']:.3f} ms "
             f"p99 {st['p99_ms']:.3f} ms")
        out = {"capacity_qps": cap_qps, "capacity_point_evals_per_s": cap["point_evals_per_s"],
               "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "mean_ms": st["mean_ms"],
               "point_evals_per_s": st["point_evals_per_s"], "requests": st["requests"]}
        if curve:
            out["load_curve"] = []
            for frac in (0.25, 0.5, 0.7):
                poisson(srv, frac * cap_qps)  # untimed: batch sizes of this load seen once
                c = poisson(srv, frac * cap_qps)
                _log(f"serve (graphs={graphs}): {frac:.0%} of capacity: p50 {c['p50_ms']:.3f} ms "
                     f"p99 {c['p99_ms']:.3f} ms")
                out["load_curve"].append({"qps": frac * cap_qps, "p50_ms": c["p50_ms"],
                                          "p99_ms": c["p99_ms"]})
        return out

    e = run(False)
    g = run(True)
    a = run("auto", curve=True)
    n = run("native", curve=True)
    res = {"qps": qps, "p50_ms": n["p50_ms"], "p99_ms": n["p99_ms"], "mean_ms": n["mean_ms"],
           "point_evals_per_s": n["point_evals_per_s"], "requests": n["requests"],
           "capacity_qps": n["capacity_qps"],
           "capacity_point_evals_per_s": n["capacity_point_evals_per_s"],
           "load_curve": n["load_curve"],
           "scheduler": "C++ (sbi_server_*), direct launches",
           "python_auto": {k: a[k] for k in ("p50_ms", "p99_ms", "capacity_qps", "load_curve")},
           "eager": {k: e[k] for k in ("p50_ms", "p99_ms", "capacity_qps")},
           "graphs": {k: g[k] for k in ("p50_ms", "p99_ms", "capacity_qps")},
           "precision": "mixed (FP32)", "interpolants": len(names), "setup_s": setup_s,
           "workload": "32 GH interpolants, d_in logU[8,500], nodes logU[1e2,2e4], d_out "
                       "logU[1,1024], batch logU[1,4096], Poisson arrivals at qps requests/s"}
    return res


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "smolyax_amd runs on one GPU; --gpus must be 1 (points are independent: several GPUs would each take a share of them)",
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

    if args.config == "c5_serve":
        out = {"metric": "serve p50/p99 ms (point-evals/s)", "config": {"workload": "c5_serve"},
               "serve": bench_serving(torch, W, n_requests=args.serve_requests), "n_gpus": 1}
        out["value"] = out["serve"]["point_evals_per_s"]
        print(json.dumps(out), flush=True)
        return
    cfg = W.CONFIGS[args.config]
    if cfg["d_out"] > 1:  # one large-d_out workload on its own
        leg = bench_gemm_leg(args, torch, W, args.config, cpu=not args.no_cpu)
        out = {"metric": "point-evals/s (mixed path) of the named workload", "unit": "point-evals/s",
               "value": leg["paths"]["mixed"]["value"], "n_gpus": 1, "steps": args.steps,
               "warmup": args.warmup, "ms_per_step": leg["paths"]["mixed"]["ms_per_step"],
               "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
               "dtype": "f32 products / f64 accumulation", "data": "synthetic",
               "config": leg["config"], "roofline": leg["contraction_roofline"]["mixed"],
               "cpu_baseline": leg["cpu_baseline"], "leg": leg}
        print(json.dumps(out), flush=True)
        return
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
        wall, ev_mean_ms = timed_loop(make_step(fn_name, plan), args.steps, args.warmup,
                                                 torch)
        outputs[name] = y.cpu().numpy().copy()
        results[name] = {"value": P * args.steps / wall, "ms_per_step": wall / args.steps * 1e3,
                         "kernel_ms": ev_mean_ms}
    # the drop-in API (validation, output allocation, dispatch) on the same points
    box = {}

    def api_step():
        box["y"] = ip.evaluate(x, precision="mixed")

    wall, ev_mean_ms = timed_loop(api_step, args.steps, args.warmup, torch)
    assert np.array_equal(box["y"].cpu().numpy(), outputs["mixed"])
    results["mixed-api"] = {"value": P * args.steps / wall, "ms_per_step": wall / args.steps * 1e3,
                            "device_ms_per_step": ev_mean_ms, "call": "ip.evaluate(x, 'mixed')"}

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

    # rooflines on the algorithmic work of the reference's evaluation: bytes per launch =
    # x (8 B per point and dimension) + F read once (4 B per entry on the FP32 path, d_out * K_raw
    # entries) + y (8 B per point and output); flops = 2 * d_out * K_raw per point.
    NL = ip._nlayout
    t_mixed = results["mixed"]["kernel_ms"] * 1e-3
    bytes_mixed = 8.0 * P * d + 4.0 * d_out * L.k_raw + 8.0 * P * d_out
    kname = {2: "newton_mixed_mfma_kernel", 1: "newton_mixed_staged_kernel"}.get(
        int(N.fn["sbi_newton_mixed_staged"](ip._nplan)), "newton_eval_kernel<float>")
    roofline = {"bound": "hbm", "achieved": bytes_mixed / t_mixed / 1e9, "peak": PEAK_HBM_GBS,
                "unit": "GB/s", "frac": bytes_mixed / t_mixed / 1e9 / PEAK_HBM_GBS,
                "traffic": measured_traffic(kname, args.config), "kernel": kname,
                "bytes_per_launch": bytes_mixed, "avg_launch_ms": results["mixed"]["kernel_ms"]}
    t_b = results["fp64"]["kernel_ms"] * 1e-3
    flops = 2.0 * d_out * L.k_raw * P
    kname_b = "bary_eval_prog_kernel" if ip._prog is not None else "bary_eval_kernel<1>"
    roof_fp64 = {"bound": "valu", "kernel": kname_b, "achieved": flops / t_b / 1e12,
                 "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                 "frac": flops / t_b / 1e12 / PEAK_FP64_TFLOPS, "avg_launch_ms": results["fp64"]["kernel_ms"],
                 "traffic": measured_traffic(kname_b, args.config),
                 "flop_basis": "2 d_out K_raw per point (the reference's terms)"}
    # the FP64 Newton kernel on the same algorithmic basis, and on its own terms (one FMA per
    # node and point plus the per-group FP64 work: 2 |Lambda| + 4 groups per point)
    t_n = results["newton-fp64"]["kernel_ms"] * 1e-3
    flop_nt = (2.0 * NL.n_nodes + 4.0 * NL.n_groups) * P
    mc64 = int(N.fn["sbi_newton_f64_matrix_cores"](ip._nplan)) == 1
    kname_n = "newton_f64_mfma_kernel" if mc64 else "newton_eval_kernel<double>"
    roof_newton = {"bound": "mfma" if mc64 else "valu", "kernel": kname_n,
                   "achieved": flops / t_n / 1e12, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                   "frac": flops / t_n / 1e12 / PEAK_FP64_TFLOPS,
                   "flop_basis": "2 d_out K_raw per point (the reference's terms)",
                   "own_terms_tflops": flop_nt / t_n / 1e12,
                   "traffic": measured_traffic(kname_n, args.config),
                   "avg_launch_ms": results["newton-fp64"]["kernel_ms"]}

    rows = bench_gradient_integral(ip, x, torch)

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
               "sample": f"{nc} of the {P} points x {reps} calls ({c1 - c0:.1f} s), FP64 torch-CPU "
                         f"restatement of the reference's padded-group einsum evaluation"}

    n_nodes = int(ip.n_f_evals)
    legs = {}
    if not args.no_legs:
        del ip, x, y
        torch.cuda.empty_cache()
        for lname in ("c3_sin_d100_o4096", "c4_tanh_d200_o256"):
            legs[lname] = bench_gemm_leg(args, torch, W, lname, cpu=not args.no_cpu)
    serve = None if args.no_serve else bench_serving(torch, W, n_requests=args.serve_requests)
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
                   "nodes": n_nodes, "t": t, "k_raw": int(L.k_raw),
                   "blocks": int(len(L.blk_n)), "internal_nodes": int(NL.n_internal),
                   "path": "mixed (hierarchical Newton form)"},
        "roofline": roofline,
        "roofline_fp64": roof_fp64,
        "roofline_newton_fp64": roof_newton,
        "cpu_baseline": cpu,
        "paths": results,
        "speedup_mixed_vs_fp64": results["mixed"]["value"] / results["fp64"]["value"],
        "speedup_mixed_vs_newton_fp64": results["mixed"]["value"] / results["newton-fp64"]["value"],
        "accuracy": accuracy,
        "gradient_integral": rows,
        "legs": legs,
        "serve": serve,
        "setup_s": setup_s,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
