"""Continuous-batching evaluation server for many resident interpolants on one GPU.

Requests (interpolant name, points already on the device) arrive at any time.  A scheduler loop
on one host thread keeps up to `n_streams` batches in flight, one per HIP stream, so their work
overlaps on the device.  Whenever a stream's batch completes (polled through HIP events), new
work is admitted in arrival order: the oldest pending request picks the interpolant, and every
pending request for the same interpolant joins its batch up to `max_batch_points` points and the
free HBM budget (the batch's basis matrix, output and inputs are counted before launch).  The
basis kernels and the matrix-core contraction (sbi_contract) compute every point independently of
the batch it is in, so a request's result is the same bit for bit alone and in the mix, and their
workgroups never wait on each other, so batches on different streams may overlap freely.  The BF16
precision runs on the library GEMM instead, whose kernels may rely on all workgroups of a launch
being resident (stream-K): two such launches overlapping on different streams can deadlock, so
that precision is served on one stream.

Latency = completion (event observed) - arrival; throughput = point evaluations per second.
"""

import collections
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


@dataclass
class _Slot:
    stream: torch.cuda.Stream
    event: torch.cuda.Event = None
    reqs: list = field(default_factory=list)
    reserved: int = 0
    keep: list = field(default_factory=list)


class Server:
    def __init__(self, interpolants, precision="mixed", n_streams=4, hbm_budget_bytes=8 << 30,
                 max_batch_points=8192):
        if precision == "bf16" and n_streams > 1:
            raise ValueError("precision 'bf16' uses the library GEMM: serve it on one stream")
        self.ips = dict(interpolants)
        self.precision = precision
        self.max_batch_points = int(max_batch_points)
        self.budget = int(hbm_budget_bytes)
        self.reserved = 0
        self.pending = collections.deque()
        self.slots = [_Slot(torch.cuda.Stream()) for _ in range(n_streams)]
        self.completed = []
        self._next = 0

    # ------------------------------------------------------------------ footprint / admission
    def footprint(self, name, n_points):
        """Device bytes a batch needs while in flight: basis matrix + output + input copy."""
        ip = self.ips[name]
        rows = ip.layout.k_raw if ip._nlayout is None or self.precision == "bary-fp32" else \
            ip._nlayout.basis_n
        item = 8 if self.precision in ("fp64", "newton-fp64") else 4
        return n_points * (rows * item + ip.d_out * 8 + ip.d_in * 8)

    def submit(self, name, x, arrival=None):
        if name not in self.ips:
            raise KeyError(name)
        r = Request(self._next, name, x, time.perf_counter() if arrival is None else arrival)
        if self.footprint(name, x.shape[0]) > self.budget:
            raise ValueError("request larger than the HBM budget of the server")
        self._next += 1
        self.pending.append(r)
        return r

    def _poll(self):
        now = time.perf_counter()
        freed = 0
        for s in self.slots:
            if s.event is not None and s.event.query():
                for r in s.reqs:
                    r.done = now
                    self.completed.append(r)
                self.reserved -= s.reserved
                s.event, s.reqs, s.reserved, s.keep = None, [], 0, []
                freed += 1
        return freed

    def _admit(self):
        launched = 0
        for s in self.slots:
            if s.event is not None or not self.pending:
                continue
            head = self.pending[0]
            batch, pts = [], 0
            for r in list(self.pending):
                if r.name != head.name:
                    continue
                n = r.x.shape[0]
                if batch and pts + n > self.max_batch_points:
                    break
                if self.reserved + self.footprint(head.name, pts + n) > self.budget:
                    break
                batch.append(r)
                pts += n
            if not batch:
                break  # wait for memory to free up
            for r in batch:
                self.pending.remove(r)
            need = self.footprint(head.name, pts)
            self.reserved += need
            ip = self.ips[head.name]
            with torch.cuda.stream(s.stream):
                xs = torch.cat([r.x for r in batch]) if len(batch) > 1 else batch[0].x
                y = ip.evaluate(xs, precision=self.precision)
                ev = torch.cuda.Event()
                ev.record(s.stream)
            a = 0
            for r in batch:
                n = r.x.shape[0]
                r.y = y[a:a + n]
                r.batch_points = pts
                a += n
            s.event, s.reqs, s.reserved, s.keep = ev, batch, need, [xs, y]
            launched += 1
        return launched

    def step(self):
        """One scheduler iteration: observe completions, admit work.  Returns True if busy."""
        self._poll()
        self._admit()
        return bool(self.pending) or any(s.event is not None for s in self.slots)

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


def run_poisson(server, names, qps, n_requests, batch_lo=1, batch_hi=4096, seed=0, inputs=None,
                time_limit_s=None, progress=None):
    """Open-loop Poisson arrivals at `qps`; each request picks an interpolant uniformly and a
    batch size log-uniform in [batch_lo, batch_hi].  Inputs are drawn on the device beforehand
    (points ~ N(0, I)) so arrivals only enqueue.  With `time_limit_s`, a run that has not
    completed every request by then raises RuntimeError naming what is still pending."""
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
    t0 = time.perf_counter()
    arrivals = t0 + np.cumsum(gaps)
    i = 0
    last_report = t0
    while True:
        now = time.perf_counter()
        if time_limit_s is not None and now - t0 > time_limit_s:
            raise RuntimeError(
                f"serving run exceeded {time_limit_s} s: {i} of {n_requests} submitted, "
                f"{len(server.completed)} completed, {len(server.pending)} pending, "
                f"{sum(s.event is not None for s in server.slots)} batches in flight")
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
    return reqs, inputs
