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
