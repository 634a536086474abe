r"""Drop-in `SmolyakBarycentricInterpolator` (src/smolyax/interpolation.py:15-390) on MI355X.

Same keyword-only constructor, `set_f`, `__call__`, `gradient`, `integral` and properties as the
reference.  Evaluation runs in hand-written HIP kernels (libsmolyax_amd.so) on the current torch
device; inputs may be numpy arrays or torch tensors, results are float64 torch tensors on the GPU
(the reference returns device-resident jax arrays).  There is no CPU fallback.
"""

import ctypes
from typing import Callable, Sequence

import numpy as np
import torch

from . import _native as N
from . import indices, newton, packing
from .nodes import Generator

PRECISIONS = ("fp64", "mixed", "newton-fp64")


def _dev_tensor(a, dtype, device):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).to(device)


class SmolyakBarycentricInterpolator:
    r"""Smolyak operator $I^{\Lambda_{k,t}}[f] = \sum_{\nu} \zeta_{\Lambda,\nu} I^\nu[f]$."""

    @property
    def d_in(self) -> int:
        return self._d_in

    @property
    def d_out(self) -> int:
        return self._d_out

    @property
    def n_f_evals(self) -> int:
        return self._n_f_evals

    @property
    def n_f_evals_new(self) -> int:
        return self._n_f_evals_new

    def __init__(self, *, d_out: int, node_gen: Generator, k: Sequence[float], t: float,
                 f: Callable = None, n_inputs: int = None, memory_limit: float = 4.0,
                 device=None, precision: str = "fp64") -> None:
        """Keyword-only, as interpolation.py:51-113.  Extensions: `device` (torch device) and
        `precision`, the default evaluation path of __call__:
          "fp64"        barycentric combination technique in FP64 (the reference's algorithm);
          "mixed"       hierarchical Newton form, FP32 surpluses / products / group sums with FP64
                        node differences and accumulation (nested nodes, d_out == 1);
          "newton-fp64" hierarchical Newton form in FP64 (nested nodes, d_out == 1).
        """
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {PRECISIONS}")
        self._precision = precision
        self._d_in = len(k)
        self._d_out = int(d_out)
        self._node_gen = node_gen
        self._is_nested = node_gen.is_nested
        self._k = np.asarray(k, dtype=np.float64)
        self._t = float(t)
        self._memory_limit = memory_limit
        self._n_inputs = n_inputs
        self._n_f_evals = indices.nodeset_cardinality(self._k, self._t, nested=self._is_nested)
        self._n_f_evals_new = 0
        self._device = torch.device(device) if device is not None else None
        self._plan = None
        self._nplan = None
        self._layout = None
        if f is not None:
            self.set_f(f=f)

    # ------------------------------------------------------------------------------ set-up
    def set_f(self, *, f: Callable, f_evals: dict = None, vectorized: bool = False) -> dict:
        """Evaluate (or reuse) f at the interpolation nodes and upload the packed layout.

        `f_evals` has the reference's format (interpolation.py:115-139).  With vectorized=True,
        f is called once on an (n_new, d_in) array and must return (n_new, d_out).
        """
        if f_evals is None:
            f_evals = {}
        blocks = packing.nonzero_blocks(self._k, self._t)
        L = packing.build_layout(blocks, self._node_gen, self._d_in, self._d_out)
        F, offset, n_new = packing.assemble_values(L, blocks, self._node_gen, f, f_evals, vectorized)
        self._n_f_evals_new += n_new
        L.slot_q = packing.quadrature_table(L, self._node_gen)
        self._layout = L
        self._F_host = F
        self._offset_host = offset
        self._nlayout = None
        if self._is_nested and self._d_out == 1:
            nf = newton.NewtonForm(self._k, self._t, self._node_gen, self._d_out)
            vals = np.array([np.broadcast_to(np.asarray(f_evals[nu], dtype=np.float64).reshape(-1),
                                             (self._d_out,)) for nu in nf.nus])
            nf.set_values(vals)
            self._nlayout = newton.build_newton_layout(nf)
        elif self._precision != "fp64":
            raise ValueError(f"precision {self._precision!r} needs nested nodes and d_out == 1")
        self._upload()
        if self._n_inputs is not None:
            x0 = torch.rand((self._n_inputs, self._d_in), dtype=torch.float64, device=self._device)
            self(x0)
        return f_evals

    def _upload(self):
        if not torch.cuda.is_available():
            raise RuntimeError("smolyax_amd needs a HIP device (MI355X); none is visible")
        if self._device is None:
            self._device = torch.device("cuda", torch.cuda.current_device())
        L, dev = self._layout, self._device
        i32, i64, f64 = torch.int32, torch.int64, torch.float64
        t = {}
        t["slot_dim"] = _dev_tensor(L.slot_dim if len(L.slot_dim) else [0], i32, dev)
        t["slot_deg"] = _dev_tensor(L.slot_deg if len(L.slot_deg) else [0], i32, dev)
        t["slot_nodes"] = _dev_tensor(L.slot_nodes, f64, dev)
        t["slot_w"] = _dev_tensor(L.slot_w, f64, dev)
        t["slot_q"] = _dev_tensor(L.slot_q, f64, dev)
        t["prefix_slot"] = _dev_tensor(L.prefix_slot if len(L.prefix_slot) else [0], i32, dev)
        t["prefix_off"] = _dev_tensor(L.prefix_off if len(L.prefix_off) else [0], i32, dev)
        t["run_slot"] = _dev_tensor(L.run_slot if len(L.run_slot) else [0], i32, dev)
        t["run_start"] = _dev_tensor(L.run_start, i32, dev)
        t["blk_n"] = _dev_tensor(L.blk_n if len(L.blk_n) else [0], i32, dev)
        t["blk_fac"] = _dev_tensor(L.blk_fac, i32, dev)
        t["fac_row"] = _dev_tensor(L.fac_row, i32, dev)
        t["fac_deg"] = _dev_tensor(L.fac_deg, i32, dev)
        t["fac_slot"] = _dev_tensor(L.fac_slot, i32, dev)
        t["blk_last_slot"] = _dev_tensor(L.blk_last_slot, i32, dev)
        t["blk_foff"] = _dev_tensor(L.blk_foff, i64, dev)
        t["F"] = _dev_tensor(self._F_host if self._F_host.size else np.zeros((1, self._d_out)), f64, dev)
        t["offset"] = _dev_tensor(self._offset_host, f64, dev)
        t["blk_desc"] = _dev_tensor(L.blk_desc, i32, dev)
        t["seg_off"] = _dev_tensor(L.seg_off, i32, dev)
        t["seg_slot"] = _dev_tensor(L.seg_slot, i32, dev)
        t["seg_b0"] = _dev_tensor(L.seg_b0, i32, dev)
        t["seg_b1"] = _dev_tensor(L.seg_b1, i32, dev)
        t["dim_blk"] = _dev_tensor(L.dim_blk, i32, dev)
        t["dim_blk_idx"] = _dev_tensor(L.dim_blk_idx, i32, dev)
        t["dim_blk_pos"] = _dev_tensor(L.dim_blk_pos, i32, dev)
        self._dev = t
        d = N.BaryDesc()
        d.d_in, d.d_out = self._d_in, self._d_out
        d.n_slots = len(L.slot_key)
        d.dmax1 = L.dmax1
        d.n_prefix_vals = L.n_prefix_vals
        d.n_prefix_slots = len(L.prefix_slot)
        d.n_runs = len(L.run_slot)
        d.n_blocks = len(L.blk_n)
        d.max_n = L.max_n
        d.k_raw = L.k_raw
        for name, ten in t.items():
            setattr(d, name, ten.data_ptr())
        d.n_split = packing.N_WAVES
        d.n_seg = int(L.seg_off[-1])
        self._desc = d
        if self._plan is not None:
            N.fn["sbi_plan_destroy"](self._plan)
            self._plan = None
        plan = ctypes.c_void_p()
        N.check(N.fn["sbi_bary_plan_create"](ctypes.byref(d), ctypes.byref(plan)), "plan_create")
        self._plan = plan
        if self._nlayout is not None:
            self._upload_newton()

    def _upload_newton(self):
        NL, dev = self._nlayout, self._device
        i32, f32, f64 = torch.int32, torch.float32, torch.float64
        t = {
            "s_root": _dev_tensor(NL.s_root, f64, dev),
            "int_parent": _dev_tensor(NL.int_parent, i32, dev),
            "int_dim": _dev_tensor(NL.int_dim, i32, dev),
            "int_deg": _dev_tensor(NL.int_deg, i32, dev),
            "level_off": _dev_tensor(NL.level_off, i32, dev),
            "dim_xi": _dev_tensor(NL.dim_xi, f64, dev),
            "dim_c": _dev_tensor(NL.dim_c, f64, dev),
            "grp": _dev_tensor(NL.grp, i32, dev),
            "grp_next_dim": _dev_tensor(NL.grp_next_dim, i32, dev),
            "wave_grp": _dev_tensor(NL.wave_grp, i32, dev),
            "node_s32": _dev_tensor(NL.node_s32, f32, dev),
            "node_s64": _dev_tensor(NL.node_s64, f64, dev),
            "node_q": _dev_tensor(NL.node_q, i32, dev),
            "node_sq": _dev_tensor(NL.node_sq, i32, dev),
            "grp_rec": torch.as_tensor(NL.grp_rec.view(np.uint8).copy()).to(dev),
            "grp_j": _dev_tensor(NL.grp_j, i32, dev),
            "hs_dim": _dev_tensor(NL.hs_dim, i32, dev),
            "hs_kmax": _dev_tensor(NL.hs_kmax, i32, dev),
            "hs_row": _dev_tensor(NL.hs_row, i32, dev),
            "int_rec": _dev_tensor(NL.int_rec, i32, dev),
        }
        self._ndev = t
        d = N.NewtonDesc()
        d.d_in, d.d_out = self._d_in, self._d_out
        d.n_nodes, d.n_internal, d.n_groups = NL.n_nodes, NL.n_internal, NL.n_groups
        d.n_split, d.n_levels, d.kmax1 = newton.N_WAVES, NL.n_levels, NL.kmax1
        d.n_nodes_pad, d.n_hs_dims, d.n_hs_rows = NL.n_nodes_pad, NL.n_hs_dims, NL.n_hs_rows
        for name, ten in t.items():
            setattr(d, name, ten.data_ptr())
        self._ndesc = d
        if self._nplan is not None:
            N.fn["sbi_newton_plan_destroy"](self._nplan)
            self._nplan = None
        plan = ctypes.c_void_p()
        N.check(N.fn["sbi_newton_plan_create"](ctypes.byref(d), ctypes.byref(plan)), "newton_plan")
        self._nplan = plan

    def __del__(self):
        for attr, fn in (("_plan", "sbi_plan_destroy"), ("_nplan", "sbi_newton_plan_destroy")):
            plan = getattr(self, attr, None)
            if plan is not None:
                try:
                    N.fn[fn](plan)
                except Exception:
                    pass

    # ------------------------------------------------------------------------------ inputs
    def _validate_input(self, x) -> torch.Tensor:
        """interpolation.py:253-262."""
        assert self._plan is not None, \
            "The operator has not yet been set up for a target function via `set_f`."
        if isinstance(x, torch.Tensor):
            x = x.to(device=self._device, dtype=torch.float64)
        else:
            x = torch.as_tensor(np.asarray(x, dtype=np.float64), device=self._device)
        if tuple(x.shape) == (self._d_in,):
            x = x[None, :]
        assert x.shape[1] == self._d_in, f"{x.shape[1]} != {self._d_in}"
        if x.stride(1) != 1:
            x = x.contiguous()
        self._n_inputs = int(x.shape[0])
        return x

    @staticmethod
    def _stream():
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    # ------------------------------------------------------------------------------ evaluation
    def __call__(self, x) -> torch.Tensor:
        """interpolation.py:264-304: y[P, d_out] (float64 tensor) on the default precision path."""
        return self.evaluate(x, precision=self._precision)

    def evaluate(self, x, precision: str = None) -> torch.Tensor:
        """Evaluate on a chosen path (see the constructor's `precision`)."""
        precision = precision or self._precision
        x = self._validate_input(x)
        P = x.shape[0]
        y = torch.empty((P, self._d_out), dtype=torch.float64, device=self._device)
        if precision == "fp64":
            N.check(N.fn["sbi_bary_eval_f64"](self._plan, x.data_ptr(), P, x.stride(0), y.data_ptr(),
                                              self._stream()), "eval")
        elif precision in ("mixed", "newton-fp64"):
            if self._nplan is None:
                raise ValueError(f"precision {precision!r} needs nested nodes and d_out == 1")
            fn = "sbi_newton_eval_mixed" if precision == "mixed" else "sbi_newton_eval_f64"
            N.check(N.fn[fn](self._nplan, x.data_ptr(), P, x.stride(0), y.data_ptr(),
                             self._stream()), precision)
        else:
            raise ValueError(f"precision must be one of {PRECISIONS}")
        return y

    def gradient(self, x) -> torch.Tensor:
        """interpolation.py:306-345: Jacobian J[P, d_out, d_in] in float64."""
        x = self._validate_input(x)
        P = x.shape[0]
        J = torch.empty((P, self._d_out, self._d_in), dtype=torch.float64, device=self._device)
        N.check(N.fn["sbi_bary_jacobian_f64"](self._plan, x.data_ptr(), P, x.stride(0), J.data_ptr(),
                                              self._stream()), "jacobian")
        return J

    def integral(self) -> torch.Tensor:
        """interpolation.py:347-390: Q[d_out]; synchronises like the reference's block_until_ready."""
        assert self._plan is not None, \
            "The operator has not yet been set up for a target function via `set_f`."
        ws = int(N.fn["sbi_bary_integral_workspace"](self._plan))
        part = torch.empty(ws, dtype=torch.float64, device=self._device)
        Q = torch.empty(self._d_out, dtype=torch.float64, device=self._device)
        N.check(N.fn["sbi_bary_integral_f64"](self._plan, Q.data_ptr(), part.data_ptr(),
                                              self._stream()), "integral")
        torch.cuda.current_stream().synchronize()
        return Q

    # ------------------------------------------------------------------------------ info
    @property
    def layout(self):
        return self._layout
