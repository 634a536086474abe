"""CPU baseline: the reference evaluation algorithm restated with torch.einsum on all host cores.

Same padded groups as the oracle (interpolation.py:168-235), same per-group einsum structure
(barycentric.py:69-123, jnp.einsum -> torch.einsum) and the same memory_limit summand batching
(interpolation.py:283-302).  Used only by bench.py's cpu_baseline leg and tests; never by the
product.  Threads: torch.set_num_threads(len(os.sched_getaffinity(0))).
"""

import os
import string

import numpy as np
import torch


def threads():
    return len(os.sched_getaffinity(0))


class TorchCPUInterpolant:
    def __init__(self, oracle_interpolant):
        ip = oracle_interpolant
        self.d_in, self.d_out = ip.d_in, ip.d_out
        self.memory_limit = ip.memory_limit
        self.offset = torch.as_tensor(np.broadcast_to(ip.offset, (ip.d_out,)).astype(np.float64))
        self.groups = []
        for n, g in ip.groups.items():
            self.groups.append(dict(
                n=n, tau=g["tau"], F=torch.as_tensor(g["F"]), nodes=torch.as_tensor(g["nodes"]),
                w=torch.as_tensor(g["weights"]), dims=torch.as_tensor(g["dims"]),
                degs=torch.as_tensor(g["degs"]), z=torch.as_tensor(g["zetas"])))

    def __call__(self, x):
        x = torch.as_tensor(np.asarray(x, dtype=np.float64))
        P = x.shape[0]
        y = self.offset.expand(P, self.d_out).clone()
        for g in self.groups:
            n = g["n"]
            fa = string.ascii_lowercase[:n]
            subs = "zo" + fa + "," + ",".join("zp" + c for c in fa) + "->zpo"
            nn = g["F"].shape[0]
            mem = P * self.d_out * int(np.prod(g["F"].shape[3:])) * 8 / 1024**3
            bs = max(1, int(np.floor(self.memory_limit / mem)))
            for a in range(0, nn, bs):
                b = min(a + bs, nn)
                bases = []
                for s, m in enumerate(g["tau"]):
                    m = m + 1
                    xs = x[:, g["dims"][a:b, s]].T[:, :, None]
                    xi = g["nodes"][a:b, s, None, :m]
                    w = g["w"][a:b, s, None, :m]
                    nu = g["degs"][a:b, s][:, None]
                    diffs = xs - xi
                    mask = torch.arange(m) <= nu[..., None]
                    zero = diffs == 0
                    mz = (zero & mask).any(-1, keepdim=True)
                    bb = torch.where(mz, zero.double(), w / diffs)
                    bb = torch.where(mask, bb, torch.zeros((), dtype=torch.float64))
                    bases.append(bb / bb.sum(-1, keepdim=True))
                r = torch.einsum(subs, g["F"][a:b], *bases)
                y += torch.einsum("z,zpo->po", g["z"][a:b], r)
        return y.numpy()
