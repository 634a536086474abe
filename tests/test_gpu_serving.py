"""GPU: continuous-batching server over 4 streams.  Every request's result is identical, bit for
bit, to the same request evaluated alone (the basis kernels and sbi_contract compute each point
independently of its batch), and within 1e-4 of max |y| of the FP64 path (FP32 operands, FP64
accumulation); admission respects the HBM budget; overlapping streams complete every request."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def small_specs():
    from smolyax_amd import workloads as W

    specs = W.serving_specs(n=6, seed=3)
    for sp in specs:  # keep the test quick
        sp["d_in"] = min(sp["d_in"], 40)
        sp["nodes"] = min(sp["nodes"], 800)
        sp["d_out"] = min(sp["d_out"], 64)
    return specs


def test_results_same_alone_and_in_mix(gpu):
    from smolyax_amd import serving
    from smolyax_amd import workloads as W

    ips = W.build_serving_interpolants(small_specs())
    names = sorted(ips)
    srv = serving.Server(ips, precision="mixed", n_streams=4, max_batch_points=2048)
    reqs, inputs = serving.run_poisson(srv, names, qps=20000.0, n_requests=120, batch_hi=512, seed=1)
    assert len(reqs) == 120 and all(r.done is not None for r in reqs)
    assert any(r.batch_points > r.x.shape[0] for r in reqs), "no request was batched with others"
    torch.cuda.synchronize()
    for r in reqs:
        alone = srv.ips[r.name].evaluate(r.x, precision="mixed")
        ref64 = srv.ips[r.name].evaluate(r.x, precision="fp64")
        scale = float(ref64.abs().max()) + 1e-30
        assert torch.equal(r.y, alone), (r.rid, r.name, r.batch_points)
        assert float((r.y - ref64).abs().max()) / scale < 1e-4
    st = srv.stats(reqs)
    assert st["p99_ms"] >= st["p50_ms"] > 0


def test_large_requests_cut_into_pieces(gpu):
    """With a small work cap per batch, requests above their interpolant's cap are cut into
    pieces served in separate batches (other requests between them); every result equals the
    request evaluated alone bit for bit, also in graph mode."""
    from smolyax_amd import serving
    from smolyax_amd import workloads as W

    ips = W.build_serving_interpolants(small_specs()[:4])
    names = sorted(ips)
    for graphs in (False, True):
        srv = serving.Server(ips, precision="mixed", n_streams=4, max_batch_work=2e6,
                             graphs=graphs)
        caps = {n: srv._cap[n] for n in names}
        reqs, _ = serving.run_poisson(srv, names, qps=20000.0, n_requests=60, batch_hi=1500,
                                      seed=7, time_limit_s=60.0)
        torch.cuda.synchronize()
        assert len(reqs) == 60 and all(r.done is not None for r in reqs)
        assert any(r.n_pieces > 1 for r in reqs), caps
        for r in reqs:
            assert r.n_pieces == -(-r.x.shape[0] // caps[r.name]) or r.n_pieces >= 1
            alone = srv.ips[r.name].evaluate(r.x, precision="mixed")
            assert torch.equal(r.y, alone), (graphs, r.rid, r.name, r.n_pieces)


def test_native_server_same_as_alone(gpu):
    """The C++ scheduler (sbi_server_*): every request of a Poisson run, including requests cut
    into pieces by a small work cap and batches of several requests (gathered and scattered by
    the copy kernel), equals the request evaluated alone bit for bit, and is within 1e-4 of the
    FP64 path; points produced on another stream are read after their producer."""
    from smolyax_amd import serving
    from smolyax_amd import workloads as W

    ips = W.build_serving_interpolants(small_specs())
    names = sorted(ips)
    for work in (1e10, 2e6):
        print(f"native server, max_batch_work={work:.0e}", flush=True)
        srv = serving.NativeServer(ips, n_streams=4, max_batch_points=2048, max_batch_work=work)
        reqs, _ = serving.run_poisson(srv, names, qps=20000.0, n_requests=150, batch_hi=1500,
                                      seed=5, time_limit_s=60.0)
        assert len(reqs) == 150 and all(r.done is not None for r in reqs)
        if work < 1e9:
            assert any(r.x.shape[0] > srv._cap[r.name] for r in reqs), "no request was cut"
        torch.cuda.synchronize()
        for r in reqs:
            alone = ips[r.name].evaluate(r.x, precision="mixed")
            assert torch.equal(r.y, alone), (r.rid, r.name, r.x.shape[0])
            ref64 = ips[r.name].evaluate(r.x, precision="fp64")
            scale = float(ref64.abs().max()) + 1e-30
            assert float((r.y - ref64).abs().max()) / scale < 1e-4
        srv.close()
    # row views of wider arrays (row stride above d_in), alone and batched with other requests
    srv = serving.NativeServer(ips, n_streams=4)
    views = []
    for i, name in enumerate(names):
        wide = torch.randn((300 + 50 * i, ips[name].d_in + 3), dtype=torch.float64, device="cuda")
        views.append((name, wide[:, 1:1 + ips[name].d_in]))
    reqs = [srv.submit(name, v) for name, v in views + views]
    srv.drain()
    torch.cuda.synchronize()
    for r in reqs:
        assert torch.equal(r.y, ips[r.name].evaluate(r.x, precision="mixed")), (r.rid, r.name)
    srv.close()
    # points written by a kernel still pending on another stream
    srv = serving.NativeServer(ips, n_streams=2)
    side = torch.cuda.Stream()
    name = names[0]
    d_in = ips[name].d_in
    with torch.cuda.stream(side):
        base = torch.randn((3000, d_in), dtype=torch.float64, device="cuda")
        torch.cuda._sleep(2_000_000)
        x = base * 0.5
        r = srv.submit(name, x)
    srv.drain()
    torch.cuda.synchronize()
    assert torch.equal(r.y, ips[name].evaluate(x, precision="mixed"))
    srv.close()


def test_admission_by_hbm_budget(gpu):
    from smolyax_amd import serving
    from smolyax_amd import workloads as W

    ips = W.build_serving_interpolants(small_specs()[:2])
    names = sorted(ips)
    one = max(ips[n].layout.k_raw * 4 + ips[n].d_out * 8 + ips[n].d_in * 8 for n in names)
    srv = serving.Server(ips, precision="mixed", n_streams=4, hbm_budget_bytes=300 * one)
    with pytest.raises(ValueError):
        srv.submit(names[0], torch.zeros((1000, ips[names[0]].d_in), dtype=torch.float64, device="cuda"))
    xs = [torch.randn((200, ips[names[i % 2]].d_in), dtype=torch.float64, device="cuda") for i in range(8)]
    reqs = [srv.submit(names[i % 2], xs[i]) for i in range(8)]
    peak = 0
    while srv.step():
        peak = max(peak, srv.reserved)
        assert srv.reserved <= srv.budget
    assert peak <= 300 * one
    for r in reqs:
        alone = srv.ips[r.name].evaluate(r.x, precision="mixed")
        assert torch.equal(r.y, alone)


def test_bf16_served_on_several_streams(gpu):
    """BF16 runs on the library's own MFMA kernel (no workgroup waits on another), so the server
    overlaps BF16 batches on 4 streams; results are the same bits alone and in the mix."""
    from smolyax_amd import serving
    from smolyax_amd import workloads as W

    ips = W.build_serving_interpolants(small_specs()[:4])
    names = sorted(ips)
    srv = serving.Server(ips, precision="bf16", n_streams=4, max_batch_points=2048)
    reqs, _ = serving.run_poisson(srv, names, qps=30000.0, n_requests=80, batch_hi=512, seed=4,
                                  time_limit_s=60.0)
    torch.cuda.synchronize()
    assert len(reqs) == 80 and all(r.done is not None for r in reqs)
    assert any(r.batch_points > r.x.shape[0] for r in reqs)
    for r in reqs:
        alone = srv.ips[r.name].evaluate(r.x, precision="bf16")
        ref64 = srv.ips[r.name].evaluate(r.x, precision="fp64")
        assert torch.equal(r.y, alone), (r.rid, r.name, r.batch_points)
        scale = float(ref64.abs().max()) + 1e-30
        assert float((r.y - ref64).abs().max()) / scale < 3e-3


def test_graph_replays_equal_eager(gpu):
    """Captured batches (HIP graphs, replayed many times with other work and synchronisations in
    between) give every request the same bits as evaluating it alone without graphs."""
    from smolyax_amd import serving
    from smolyax_amd import workloads as W

    ips = W.build_serving_interpolants(small_specs()[:4])
    names = sorted(ips)
    srv = serving.Server(ips, precision="mixed", n_streams=4, max_batch_points=1024, graphs=True)
    srv.prewarm(buckets=[64, 128, 256, 512, 1024])
    auto = serving.Server(ips, precision="mixed", n_streams=4, max_batch_points=1024,
                          graphs="auto")
    for rnd, server in enumerate([srv, srv, auto]):
        reqs, _ = serving.run_poisson(server, names, qps=30000.0, n_requests=60, batch_hi=400,
                                      seed=10 + rnd)
        torch.cuda.synchronize()
        assert len(reqs) == 60
        for r in reqs:
            alone = srv.ips[r.name].evaluate(r.x, precision="mixed")
            assert torch.equal(r.y, alone), (rnd, r.rid, r.name, r.batch_points)
    assert any(len(sl.graphs) > 0 for sl in srv.slots)


def test_points_from_pending_producer(gpu):
    """Points written by kernels still pending on the submitting stream are read only after that
    stream reaches them: a long spin kernel is queued before the copy that produces x, and x is
    submitted before the copy has run."""
    from smolyax_amd import serving
    from smolyax_amd import workloads as W

    ips = W.build_serving_interpolants(small_specs()[:2])
    names = sorted(ips)
    srv = serving.Server(ips, precision="mixed", n_streams=4)
    torch.manual_seed(0)
    bases = [torch.randn((300, ips[names[i % 2]].d_in), dtype=torch.float64, device="cuda")
             for i in range(6)]
    torch.cuda.synchronize()
    reqs = []
    for i, b in enumerate(bases):
        x = torch.full_like(b, float("nan"))
        torch.cuda._sleep(20_000_000)  # bounded spin (~10 ms) on the current stream
        x.copy_(b)
        reqs.append(srv.submit(names[i % 2], x))
        del x
    srv.drain()
    torch.cuda.synchronize()
    for i, r in enumerate(reqs):
        alone = srv.ips[r.name].evaluate(bases[i], precision="mixed")
        assert torch.equal(r.y, alone), i


def test_serving_workload_at_size(gpu):
    """The serving workload's extremes at their full size: d_in 500 and d_in 8 at 2e4 nodes with
    d_out 1024, d_in 500 at 1e2 nodes with d_out 1, and the largest interpolant of the 32-spec set
    (by nodes x d_out), beside two small ones.  Served results (FP32 operands, FP64 accumulation)
    equal each request evaluated alone bit for bit (Python and C++ schedulers); on a sample of 8
    points per interpolant the FP64 path is held to the exact interpolant of the FP64 data (the
    oracle in extended precision) within 4x the FP64 oracle's own distance from it, and the served
    values to 1e-6 of max |y| where the Newton form serves 'mixed' (1e-4 on the FP32 barycentric
    form)."""
    from oracle import nodes as ON
    from oracle import smolyak as OS
    from smolyax_amd import serving
    from smolyax_amd import workloads as W

    base = W.serving_specs(32, seed=0)
    largest = max(base, key=lambda sp: sp["nodes"] * sp["d_out"])
    specs = [dict(name="xl500", d_in=500, nodes=20000, d_out=1024, fseed=7),
             dict(name="xl8", d_in=8, nodes=20000, d_out=1024, fseed=8),
             dict(name="xs500", d_in=500, nodes=100, d_out=1, fseed=9),
             dict(largest, name="big")] + small_specs()[:2]
    ips = W.build_serving_interpolants(specs)
    names = sorted(ips)
    for n in names:
        ip = ips[n]
        print(f"serving at size: {n} d_in={ip.d_in} d_out={ip.d_out} k_raw={ip.layout.k_raw} "
              f"mixed via {ip.mixed_path} ({ip.mixed_head_rows} FP64 head rows, "
              f"{ip.mixed_error} from FP64 on sample points)", flush=True)
    srv = serving.Server(ips, precision="mixed", n_streams=4, max_batch_points=4096)
    reqs, _ = serving.run_poisson(srv, names, qps=5000.0, n_requests=90, batch_hi=2048, seed=5,
                                  time_limit_s=120.0)
    torch.cuda.synchronize()
    assert len(reqs) == 90 and all(r.done is not None for r in reqs)
    for r in reqs:
        alone = srv.ips[r.name].evaluate(r.x, precision="mixed")
        assert torch.equal(r.y, alone), (r.rid, r.name, r.batch_points)
    # the C++ scheduler on the same workload (its interpolants all have a Newton form here)
    if all(ip._nplan is not None for ip in ips.values()):
        nsrv = serving.NativeServer(ips, n_streams=4, max_batch_points=4096)
        nreqs, _ = serving.run_poisson(nsrv, names, qps=5000.0, n_requests=90, batch_hi=2048,
                                       seed=6, time_limit_s=120.0)
        torch.cuda.synchronize()
        assert len(nreqs) == 90 and all(r.done is not None for r in nreqs)
        for r in nreqs:
            assert torch.equal(r.y, ips[r.name].evaluate(r.x, precision="mixed")), (r.rid, r.name)
        nsrv.close()
    by_name = {}
    for r in reqs:
        by_name.setdefault(r.name, r)
    sp_of = {sp["name"]: sp for sp in specs}
    bad = []
    for n, r in sorted(by_name.items()):
        sp = sp_of[n]
        ip = srv.ips[n]
        d = sp["d_in"]
        f = W.f_sin(d, sp["d_out"], seed=sp["fseed"])
        ref = OS.Interpolant(d_out=sp["d_out"], node_gen=ON.gauss_hermite(
            scaling=np.full(d, np.sqrt(2.0))), k=W.k_rule(d), t=ip._t,
            f=lambda x, f=f: f(np.asarray(x)[None, :])[0])
        xs = r.x[:8].cpu().numpy()
        y_ex = ref(xs, dtype=np.longdouble).astype(np.float64)
        scale = np.max(np.abs(y_ex))
        e_ora = np.max(np.abs(ref(xs) - y_ex)) / scale
        e64 = np.max(np.abs(ip.evaluate(r.x[:8], precision="fp64").cpu().numpy() - y_ex)) / scale
        e_srv = np.max(np.abs(r.y[:8].cpu().numpy() - y_ex)) / scale
        print(f"{n}: served {e_srv:.1e}, FP64 path {e64:.1e}, FP64 oracle {e_ora:.1e} from the "
              f"exact values ({ip.mixed_path})", flush=True)
        if e64 > max(4.0 * e_ora, 1e-14) or e_srv > (1e-6 if ip.mixed_path == "newton" else 1e-4):
            bad.append(n)
    assert not bad, bad
