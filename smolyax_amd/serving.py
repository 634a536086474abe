"""Continuous-batching evaluation server for many resident interpolants on one GPU.

Two schedulers with the same admission rule: `Server` (below, in Python: HIP graphs under a
backlog, an HBM budget, every precision) and `NativeServer` (the library's C++ scheduler,
sbi_server_*, mixed path: Python only submits requests and reads completions, so the host work
per batch is a few HIP calls; the lower latency under load).

Requests (interpolant name, points already on the device) arrive at any time.  A scheduler loop
on one host thread keeps up to `n_streams` batches in flight, one per HIP stream, so their work
overlaps on the device.  Whenever a stream's batch completes (polled through HIP events), new
work is admitted in arrival order: the oldest pending request picks the interpolant, and every
pending request for the same interpolant joins its batch up to `max_batch_points` points and the
free HBM budget (the batch's basis matrix, output and inputs are counted before launch).  The
basis kernels and the matrix-core contraction (sbi_contract) compute every point independently of
the batch it is in, so a request's result is the same bit for bit alone and in the mix, and their
workgroups never wait on each other, so batches on different streams may overlap freely.  Every
precision runs on the library's own kernels (the BF16 one on sbi_contract_bf16_k8), none on a
library GEMM whose workgroups may wait on each other.

With `graphs="auto"` the server replays graphs only while requests are queued (at least
`n_streams` pending): a captured launch costs the host less, so it raises throughput under a
backlog, but on this stack a replay takes longer to complete than the same few kernels launched
directly, so a short queue is served eagerly for latency.

With `graphs=True` each (interpolant, batch bucket, stream) evaluation is captured once in a
HIP graph (torch.cuda.CUDAGraph) on static input/output buffers: a batch is copied into the
static input of the smallest power-of-two bucket that holds it, the graph is replayed on the
stream, and the batch's rows of the static output are copied out.  The captured work is the basis
kernel and the contraction (no memset nodes); rows past the batch hold earlier inputs and are
ignored, and every point is computed independently, so results equal eager evaluation bit for bit.

Large requests: a batch is capped at `max_batch_work` multiply-adds of contraction (basis rows x
outputs per point), so a request above its interpolant's cap is cut into pieces that take
separate batches and short requests of other interpolants run between them; its output is
allocated at submit() and each piece copies its rows into it (every point is computed on its own,
so the result is the same bits as evaluating the request alone).

Stream ordering: submit() records an event on the caller's current stream, and the batch that
takes the request waits for it on its own stream before reading the points, so points produced by
kernels still pending on the caller's stream are read finished.  The points are marked as used on
the server's stream (record_stream), so the caller may drop them right after submit().  A result
`r.y` is written on the server's stream: it is final once `r.done` is set; a caller that hands it
to work on another stream without waiting for that, or frees it there, records that stream on it.

Latency = completion (event observed) - arrival; throughput = point evaluations per second.
"""

import collections
import ctypes
import gc
import time
from dataclasses import dataclass, field

import numpy as np
import torch


@dataclass
class Request:
    rid: int
    name: str
    x: torch.Tensor
    arrival: float
    y: torch.Tensor = None
    done: float = None
    batch_points: int = 0
    taken: bool = False             # every point is in a launched batch
    ready: torch.cuda.Event = None  # recorded on the submitting stream after x was produced
    pos: int = 0                    # points already in launched batches
    parts: int = 0                  # launched pieces not completed yet
    n_pieces: int = 0               # batches the request was cut into


@dataclass
class _Slot:
    stream: torch.cuda.Stream
    event: torch.cuda.Event = None               # the batch in flight (None: free)
    done_ev: torch.cuda.Event = None             # this slot's completion event, re-recorded
    reqs: list = field(default_factory=list)
    reserved: int = 0
    keep: list = field(default_factory=list)
    pool: object = None                          # graph memory pool of this stream
    graphs: dict = field(default_factory=dict)   # (name, bucket) -> (graph, x_static, y_static)


MIN_BUCKET = 64


class Server:
    def __init__(self, interpolants, precision="mixed", n_streams=4, hbm_budget_bytes=8 << 30,
                 max_batch_points=8192, graphs=False, max_batch_work=1e10):
        if graphs not in (False, True, "auto"):
            raise ValueError("graphs must be False, True or 'auto'")
        self.graph_mode = graphs
        self.graphs = graphs is not False
        self.ips = dict(interpolants)
        self.precision = precision
        self.max_batch_points = int(max_batch_points)
        self.max_batch_work = float(max_batch_work)
        self.budget = int(hbm_budget_bytes)
        self.reserved = 0
        # arrival order (entries already batched are skipped lazily) and per-interpolant queues
        self.pending = collections.deque()
        self.by_name = {name: collections.deque() for name in self.ips}
        self.n_pending = 0
        self._bpp = {name: self.footprint(name, 1) for name in self.ips}
        # points per batch within max_batch_work (multiply-adds: basis rows x outputs per point)
        self._cap = {name: max(MIN_BUCKET, min(self.max_batch_points,
                                                int(self.max_batch_work // max(self.work(name), 1))))
                     for name in self.ips}
        self.slots = [_Slot(torch.cuda.Stream(), done_ev=torch.cuda.Event())
                      for _ in range(n_streams)]
        self._ev_pool = []  # ready events of requests already waited for (re-recorded later)
        if self.graphs:
            for sl in self.slots:
                sl.pool = torch.cuda.graph_pool_handle()
        self.completed = []
        self._next = 0

    # ------------------------------------------------------------------ footprint / admission
    @staticmethod
    def bucket(n_points):
        """Captured batch size for n_points: the smallest of MIN_BUCKET * 2^(i/2), rounded up to
        a multiple of 32, that holds them (at most ~1.41 x padding, ~1.2 x on average)."""
        i = 0
        while True:
            b = -(-int(MIN_BUCKET * 2.0 ** (i / 2)) // 32) * 32
            if b >= n_points:
                return b
            i += 1

    def _graph(self, slot, name, bucket):
        """The captured evaluation of `name` on `bucket` points on this slot's stream."""
        key = (name, bucket)
        entry = slot.graphs.get(key)
        if entry is None:
            ip = self.ips[name]
            xs = torch.zeros((bucket, ip.d_in), dtype=torch.float64, device="cuda")
            with torch.cuda.stream(slot.stream):
                ip.evaluate(xs, precision=self.precision)  # warm-up outside the capture
            slot.stream.synchronize()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, pool=slot.pool, stream=slot.stream):
                ys = ip.evaluate(xs, precision=self.precision)
            entry = slot.graphs[key] = (g, xs, ys)
        return entry

    def prewarm(self, buckets=None):
        """Capture every (interpolant, bucket, stream) graph ahead of serving."""
        if not self.graphs:
            return
        if buckets is None:
            buckets, b = [], 1
            while b <= self.max_batch_points:
                b = self.bucket(b)
                buckets.append(b)
                b += 1
        for sl in self.slots:
            for name in self.ips:
                for b in sorted(set(buckets)):
                    if b <= self.bucket(self.max_batch_points):
                        self._graph(sl, name, b)
        torch.cuda.synchronize()

    def footprint(self, name, n_points):
        """Device bytes a batch needs while in flight: basis matrix + output + input copy."""
        ip = self.ips[name]
        rows = ip.layout.k_raw if ip._nlayout is None or self.precision == "bary-fp32" else \
            ip._nlayout.basis_n
        item = {"fp64": 8, "newton-fp64": 8, "bf16": 2}.get(self.precision, 4)
        return n_points * (rows * item + ip.d_out * 8 + ip.d_in * 8)

    def work(self, name):
        """Multiply-adds of one point of `name`'s contraction (basis rows x outputs)."""
        ip = self.ips[name]
        rows = ip.layout.k_raw if ip._nlayout is None or self.precision == "bary-fp32" else \
            ip._nlayout.basis_n
        return rows * ip.d_out

    def submit(self, name, x, arrival=None):
        if name not in self.ips:
            raise KeyError(name)
        r = Request(self._next, name, x, time.perf_counter() if arrival is None else arrival)
        if self.footprint(name, x.shape[0]) > self.budget:
            raise ValueError("request larger than the HBM budget of the server")
        if x.shape[0] > self._cap[name]:  # cut into pieces: its output is allocated here, on the
            # submitting stream before `ready`, so every piece's stream is ordered after it
            r.y = torch.empty((x.shape[0], self.ips[name].d_out), dtype=torch.float64,
                              device=x.device)
        r.ready = self._ev_pool.pop() if self._ev_pool else torch.cuda.Event()
        r.ready.record(torch.cuda.current_stream(x.device))
        self._next += 1
        self.pending.append(r)
        self.by_name[name].append(r)
        self.n_pending += 1
        return r

    def _poll(self):
        now = time.perf_counter()
        freed = 0
        for s in self.slots:
            if s.event is not None and s.event.query():
                for r in s.reqs:
                    r.parts -= 1
                    if r.taken and r.parts == 0:
                        r.done = now
                        self.completed.append(r)
                self.reserved -= s.reserved
                s.event, s.reqs, s.reserved, s.keep = None, [], 0, []
                freed += 1
        return freed

    def _admit(self):
        """Oldest pending request picks the interpolant; its queue (arrival order) fills the batch
        up to the interpolant's point cap (max_batch_points, and max_batch_work of device work)
        and the free HBM budget.  A request larger than the cap is cut into pieces, one per
        batch, so that short requests of other interpolants interleave with it; its result is
        assembled from the pieces (each point is computed on its own, so the bits are those of
        evaluating it alone).  O(batch) per launch."""
        launched = 0
        for s in self.slots:
            if s.event is not None:
                continue
            while self.pending and self.pending[0].taken:
                self.pending.popleft()
            if not self.pending:
                break
            head = self.pending[0]
            q = self.by_name[head.name]
            bpp = self._bpp[head.name]
            cap = self._cap[head.name]
            pieces, pts = [], 0  # (request, first point, points)
            while q and pts < cap:
                r = q[0]
                n = min(r.x.shape[0] - r.pos, cap - pts)
                if pieces and r.pos == 0 and n < r.x.shape[0]:
                    break  # a request joins a batch whole, or starts one
                if self.reserved + bpp * (pts + n) > self.budget:
                    break
                pieces.append((r, r.pos, n))
                r.pos += n
                r.parts += 1
                r.n_pieces += 1
                pts += n
                if r.pos == r.x.shape[0]:
                    r.taken = True
                    q.popleft()
                    self.n_pending -= 1
            if not pieces:
                break  # wait for memory to free up
            need = bpp * pts
            self.reserved += need
            ip = self.ips[head.name]
            use_graph = self.graphs and (self.graph_mode is True or
                                         self.n_pending >= len(self.slots))
            for r, _, _ in pieces:  # the points are read on this stream once their producer is done
                s.stream.wait_event(r.ready)
                r.x.record_stream(s.stream)
                if r.taken:  # no later piece waits for it: the event may be recorded again
                    self._ev_pool.append(r.ready)
                    r.ready = None
            xin = [r.x if (a0 == 0 and n == r.x.shape[0]) else r.x[a0:a0 + n]
                   for r, a0, n in pieces]
            with torch.cuda.stream(s.stream):
                if use_graph:
                    g, x_st, y_st = self._graph(s, head.name, self.bucket(pts))
                    if len(xin) > 1:  # one copy kernel for the whole batch
                        torch.cat(xin, out=x_st[:pts])
                    else:
                        x_st[:pts].copy_(xin[0])
                    g.replay()
                    y = y_st[:pts].clone()
                    xs = x_st
                else:
                    xs = torch.cat(xin) if len(xin) > 1 else xin[0]
                    y = ip.evaluate(xs, precision=self.precision)
                a = 0
                for r, a0, n in pieces:
                    if a0 == 0 and n == r.x.shape[0]:
                        r.y = y[a:a + n]
                    else:  # a piece of a larger request: into the request's own output
                        r.y[a0:a0 + n].copy_(y[a:a + n])
                        r.y.record_stream(s.stream)
                    r.batch_points = max(r.batch_points, pts)
                    a += n
                ev = s.done_ev  # free slot: its previous batch was observed complete
                ev.record(s.stream)
            s.event, s.reqs, s.reserved, s.keep = ev, [r for r, _, _ in pieces], need, [xs, y]
            launched += 1
        return launched

    def step(self):
        """One scheduler iteration: observe completions, admit work.  Returns True if busy."""
        self._poll()
        self._admit()
        return self.n_pending > 0 or any(s.event is not None for s in self.slots)

    def drain(self):
        while self.step():
            time.sleep(20e-6)

    # ------------------------------------------------------------------ statistics
    def stats(self, reqs=None):
        reqs = self.completed if reqs is None else reqs
        lat = np.array([(r.done - r.arrival) * 1e3 for r in reqs])
        pts = sum(r.x.shape[0] for r in reqs)
        span = max(r.done for r in reqs) - min(r.arrival for r in reqs)
        return {"requests": len(reqs), "p50_ms": float(np.percentile(lat, 50)),
                "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()),
                "point_evals_per_s": pts / span, "span_s": span}


class NativeServer:
    """The same scheduling as Server in the library's C++ scheduler (sbi_server_*): arrival
    order, per-interpolant batches of whole requests up to the interpolant's point cap (at most
    max_batch_points and max_batch_work multiply-adds, at most 64 requests), requests above the
    cap cut into pieces, up to `n_streams` batches in flight on HIP streams of their own.  Python
    only allocates each request's output and reads completions, so the host work per batch is a
    few HIP calls instead of the Python scheduler's tensor operations.  Mixed path (FP32
    products, FP64 accumulation) of every interpolant, which must have a Newton form.

    Buffers: a request keeps its points and its output alive until it is reported done; the
    output is allocated on the caller's current stream and written on a server stream, and is
    final once `r.done` is set (the host has seen the batch finish)."""

    def __init__(self, interpolants, n_streams=4, max_batch_points=8192, max_batch_work=5e9):
        from . import _native as N

        self._N = N
        self.ips = dict(interpolants)
        self.max_batch_points = int(max_batch_points)
        self.max_batch_work = float(max_batch_work)
        srv = ctypes.c_void_p()
        N.check(N.fn["sbi_server_create"](int(n_streams), ctypes.byref(srv)), "server_create")
        self._srv = srv
        self._id = {}
        self._cap = {}
        for name, ip in self.ips.items():
            if ip._nplan is None:
                raise ValueError(f"{name}: the native server needs a Newton form ({ip._newton_off})")
            nl = ip._nlayout
            work = max(nl.basis_n * ip.d_out, 1)
            cap = max(MIN_BUCKET, min(self.max_batch_points, int(self.max_batch_work // work)))
            self._cap[name] = cap
            i = ctypes.c_int32()
            N.check(N.fn["sbi_server_add"](srv, ip._nplan, cap, ctypes.byref(i)), "server_add")
            self._id[name] = i.value
        self._live = {}   # rid -> Request
        self._done_buf = (ctypes.c_int64 * 256)()
        self._n_done, self._busy = ctypes.c_int32(), ctypes.c_int32()
        self._rid = ctypes.c_int64()
        self.completed = []
        self.n_pending = 0

    def submit(self, name, x, arrival=None):
        ip = self.ips[name]
        if x.dim() != 2 or x.shape[1] != ip.d_in or x.dtype != torch.float64 or not x.is_cuda:
            raise ValueError(f"{name}: points must be a (P, {ip.d_in}) float64 device tensor")
        if x.shape[0] == 0:
            raise ValueError("empty request")
        if x.stride(1) != 1:
            x = x.contiguous()
        r = Request(-1, name, x, time.perf_counter() if arrival is None else arrival)
        r.y = torch.empty((x.shape[0], ip.d_out), dtype=torch.float64, device=x.device)
        st = torch.cuda.current_stream(x.device).cuda_stream
        self._N.check(self._N.fn["sbi_server_submit"](
            self._srv, self._id[name], x.data_ptr(), x.shape[0], max(x.stride(0), ip.d_in),
            r.y.data_ptr(), ip.d_out, st, ctypes.byref(self._rid)), "server_submit")
        r.rid = self._rid.value
        self._live[r.rid] = r
        self.n_pending += 1
        return r

    def step(self):
        """Observe finished batches and launch new ones.  Returns True while busy."""
        N = self._N
        N.check(N.fn["sbi_server_step"](self._srv, self._done_buf, len(self._done_buf),
                                        ctypes.byref(self._n_done), ctypes.byref(self._busy)),
                "server_step")
        k = self._n_done.value
        if k:
            now = time.perf_counter()
            for j in range(k):
                r = self._live.pop(self._done_buf[j])
                r.done = now
                r.taken = True
                self.completed.append(r)
            self.n_pending -= k
        return bool(self._busy.value)

    def drain(self):
        while self.step():
            time.sleep(20e-6)

    def close(self):
        if self._srv is not None:
            self._N.check(self._N.fn["sbi_server_destroy"](self._srv), "server_destroy")
            self._srv = None
            self._live.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    stats = Server.stats


def run_poisson(server, names, qps, n_requests, batch_lo=1, batch_hi=4096, seed=0, inputs=None,
                time_limit_s=None, progress=None):
    """Open-loop Poisson arrivals at `qps`; each request picks an interpolant uniformly and a
    batch size log-uniform in [batch_lo, batch_hi].  Inputs are drawn on the device beforehand
    (points ~ N(0, I)) so arrivals only enqueue.  With `time_limit_s`, a run that has not
    completed every request by then raises RuntimeError naming what is still pending.  Python's
    cyclic garbage collector is paused during the run (objects that exist before it are frozen
    out of later collections): a full collection over the requests of earlier runs stalled the
    scheduling loop for tens of milliseconds, as it would any long-running serving process that
    does not manage it."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / qps, n_requests)
    picks = rng.integers(0, len(names), n_requests)
    sizes = np.exp(rng.uniform(np.log(batch_lo), np.log(batch_hi + 1), n_requests)).astype(int)
    sizes = np.clip(sizes, batch_lo, batch_hi)
    if inputs is None:
        gen = torch.Generator(device="cuda").manual_seed(seed)
        inputs = [torch.randn((int(b), server.ips[names[p]].d_in), dtype=torch.float64,
                              device="cuda", generator=gen) for p, b in zip(picks, sizes)]
    torch.cuda.synchronize()
    reqs = []
    gc_was = gc.isenabled()
    gc.collect()
    gc.freeze()
    gc.disable()
    try:
        _poisson_loop(server, names, n_requests, picks, inputs, gaps, reqs, time_limit_s, progress)
    finally:
        gc.unfreeze()
        if gc_was:
            gc.enable()
    return reqs, inputs


def _poisson_loop(server, names, n_requests, picks, inputs, gaps, reqs, time_limit_s, progress):
    t0 = time.perf_counter()
    arrivals = t0 + np.cumsum(gaps)
    i = 0
    last_report = t0
    while True:
        now = time.perf_counter()
        if time_limit_s is not None and now - t0 > time_limit_s:
            raise RuntimeError(
                f"serving run exceeded {time_limit_s} s: {i} of {n_requests} submitted, "
                f"{len(server.completed)} completed, {server.n_pending} pending, "
                f"{sum(s.event is not None for s in getattr(server, 'slots', []))} batches of the "
                f"Python scheduler in flight")
        if progress is not None and now - last_report > 10.0:
            progress(f"serve: {i} submitted, {len(server.completed)} completed")
            last_report = now
        while i < n_requests and now >= arrivals[i]:
            reqs.append(server.submit(names[picks[i]], inputs[i], arrival=float(arrivals[i])))
            i += 1
        busy = server.step()
        if i >= n_requests and not busy:
            break
        if not busy and i < n_requests:
            time.sleep(max(0.0, min(50e-6, arrivals[i] - time.perf_counter())))
