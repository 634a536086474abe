r"""Drop-in `SmolyakBarycentricInterpolator` (src/smolyax/interpolation.py:15-390) on MI355X.

Same keyword-only constructor, `set_f`, `__call__`, `gradient`, `integral` and properties as the
reference.  Evaluation runs in hand-written HIP kernels (libsmolyax_amd.so) on the current torch
device; inputs may be numpy arrays or torch tensors, results are float64 torch tensors on the GPU
(the reference returns device-resident jax arrays).  There is no CPU fallback.
"""

import ctypes
import warnings
from typing import Callable, Sequence

import numpy as np
import torch

from . import _native as N
from . import indices, newton, packing
from .nodes import Generator

PRECISIONS = ("fp64", "mixed", "newton-fp64", "bf16", "bary-fp32")
GEMM_MIN_DOUT = 16  # from this d_out on, the d_out contraction is one dense GEMM
MAX_CALL_POINTS = 65535 * 64  # points per call of sbi_contract / _jacobian_f64 / sbi_newton_vjp
# Newton form of a non-nested interpolant on a nested proxy grid (see _newton_proxy): its largest
# relative distance from the FP64 barycentric path on sample points decides what it may serve.
PROXY_TOL_FP64 = 1e-12   # up to this: every path (newton-fp64, gradient, adjoint in FP64)
PROXY_TOL_MIXED = 1e-7   # up to this: the FP32-accurate paths (mixed, bf16) only
PROXY_SAMPLES = 256
# "mixed" on the GEMM path (d_out > 1): while sample points are further than this from the FP64
# path (relative to max |y|), 4x more of the heaviest Newton basis rows move into the FP64 head
MIXED_TARGET = 2e-7


class SmolyaxPrecisionWarning(UserWarning):
    """A requested precision path is unavailable for an interpolant and a less accurate one
    serves the call (stated in the message)."""


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
          "mixed"       hierarchical Newton form with FP32 surpluses and products and FP64 node
                        differences; d_out == 1: FP64 accumulation across groups; larger d_out:
                        the heaviest rows in FP64, the rest FP32 on MFMA with FP64 accumulation;
          "newton-fp64" hierarchical Newton form in FP64;
          "bf16"        hierarchical Newton form: the heaviest rows in FP64, the rest with BF16
                        basis and surpluses on the BF16 matrix cores (FP32 partial sums added into
                        FP64);
          "bary-fp32"   barycentric form with FP64 node differences, FP32 entry weights and FP32
                        contraction on MFMA.
        The Newton form is built on the interpolant's own nodes when they are nested (Leja) and on
        a nested proxy grid otherwise (Gauss-Hermite: the same polynomial, see _newton_proxy);
        when the index set exceeds the Newton layout's limits (more than 6 factors in a node, or
        a factor table beyond 65534 rows), "mixed" is served by "bary-fp32" and a
        SmolyaxPrecisionWarning says so (`mixed_path` names the path in use); "newton-fp64" and
        "bf16" raise ValueError.
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
        self._newton_off = "set_f has not been called"
        self._proxy_err = None      # non-nested: proxy Newton form vs the FP64 barycentric path
        self._newton_fp64_ok = False
        self._warned_mixed = False
        self._mixed_err = None      # mixed GEMM path vs the FP64 path on sample points (d_out > 1)
        self._maxdeg = None         # per dim: highest degree in Lambda (sample points)
        self._layout = None
        self._prog = None           # record program of the d_out = 1 FP64 eval kernel
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
        vt = packing.gather_values(L, blocks, self._node_gen, f, f_evals, vectorized)
        self._n_f_evals_new += vt.n_new
        L.slot_q = packing.quadrature_table(L, self._node_gen)
        self._layout = L
        self._prog = None
        self._values = vt          # host value table, packed into F on the device by _upload
        self._nlayout = None
        self._nform = None
        self._newton_off = None  # why the Newton form is unavailable (None: available)
        self._newton_fp64_ok = self._is_nested
        self._proxy_err = None
        self._upload()
        self._values = None
        try:
            if self._is_nested:
                nf = newton.NewtonForm(self._k, self._t, self._node_gen, self._d_out)
                self._set_surpluses(nf, packing.stack_values([f_evals[nu] for nu in nf.nus],
                                                             self._d_out))
                self._nform = nf
            else:  # Newton form on a nested proxy grid (needs the FP64 plan)
                self._nform = self._newton_proxy()
            self._nlayout = newton.build_newton_layout(self._nform)
            self._upload_newton()
        except newton.CapacityError as e:
            self._nform = self._nlayout = None
            self._newton_off = str(e)
        if not self._is_nested and self._nplan is not None:
            self._check_proxy()
        self._calibrate_mixed()
        if self._newton_off and self._precision in ("newton-fp64", "bf16"):
            raise ValueError(f"precision {self._precision!r}: {self._newton_off}")
        if self._n_inputs is not None:
            x0 = torch.rand((self._n_inputs, self._d_in), dtype=torch.float64, device=self._device)
            self(x0)
        return f_evals

    def _newton_proxy(self):
        """Newton form of a non-nested (Gauss-Hermite) interpolant.

        The Smolyak interpolant is a polynomial in P_Lambda = span{x^nu : nu in Lambda} for any
        node family, and the sparse grid of a nested sequence is unisolvent for P_Lambda.  So the
        interpolant is evaluated (FP64 barycentric path) at the |Lambda| grid points of Leja
        sequences spanning each dim's outermost node, and hierarchised: the same polynomial with
        |Lambda| terms instead of K_raw, without the combination technique's cancellation.
        """
        from . import nodes as _nodes

        d = self._d_in
        maxdeg = np.zeros(d, dtype=np.int64)
        for nu in indices.indexset(self._k, self._t):
            for j, v in nu:
                maxdeg[j] = max(maxdeg[j], v)
        self._maxdeg = maxdeg
        dom = np.zeros((d, 2))
        for j, g in enumerate(self._node_gen):
            pts = np.asarray(g(int(maxdeg[j])), dtype=np.float64)
            lo, hi = float(pts.min()), float(pts.max())
            if hi - lo <= 1e-12 * max(1.0, abs(hi)):  # one node: a unit interval around it
                lo, hi = lo - 1.0, hi + 1.0
            dom[j] = (lo, hi)
        nf = newton.NewtonForm(self._k, self._t, _nodes.Leja(domains=dom), self._d_out)
        zero = np.array([g(0)[0] for g in nf.node_gen], dtype=np.float64)
        X = np.stack([nf.grid_point(nu, zero) for nu in nf.nus])
        keep = self._n_inputs
        Y = self.evaluate(torch.as_tensor(X, device=self._device), precision="fp64")
        self._n_inputs = keep
        self._set_surpluses(nf, Y)
        return nf

    def _set_surpluses(self, nf, values):
        """nf.s from the grid values at nf.nus (rows; host array or device tensor), hierarchised
        on the device by sbi_hierarchize_f64 (the host program newton.hierarchy_steps)."""
        dev, i32 = self._device, torch.int32
        S = torch.as_tensor(values, dtype=torch.float64, device=dev).reshape(len(nf.nus), -1)
        S = S.clone().contiguous()
        if S.shape[1] != self._d_out:
            S = S.expand(len(nf.nus), self._d_out).contiguous()
        steps = newton.hierarchy_steps(nf.nus, nf.node_gen)
        a = [_dev_tensor(steps.step_off, i32, dev), _dev_tensor(steps.tgt if len(steps.tgt) else [0], i32, dev),
             _dev_tensor(steps.src_off, i32, dev), _dev_tensor(steps.src if len(steps.src) else [0], i32, dev),
             _dev_tensor(steps.coef if len(steps.coef) else [0.0], torch.float64, dev)]
        st = torch.cuda.current_stream(dev)
        N.check(N.fn["sbi_hierarchize_f64"](S.data_ptr(), S.shape[0], self._d_out, self._d_out,
                                            a[0].data_ptr(), steps.n_steps, a[1].data_ptr(),
                                            a[2].data_ptr(), a[3].data_ptr(), a[4].data_ptr(),
                                            ctypes.c_void_p(st.cuda_stream)), "hierarchize")
        nf.s = S.cpu().numpy()

    def _calibrate_mixed(self):
        """The mixed GEMM path (d_out > 1) contracts the Newton basis rows with the largest terms
        in FP64 (the head) and the others with FP32 operands.  Its error is measured against the
        FP64 path on sample points; while it exceeds MIXED_TARGET of max |y|, 4x more rows move
        into the head (at most all of them, which is the FP64 contraction)."""
        self._mixed_err = None
        if self._nplan is None or self._d_out == 1:
            return
        # points from the nodes' measure only: far out, the polynomial's own size sets the scale
        X = torch.as_tensor(self._proxy_samples()[: PROXY_SAMPLES // 2], device=self._device)
        keep, warned = self._n_inputs, self._warned_mixed
        ref = self.evaluate(X, precision="newton-fp64" if self._newton_fp64_ok else "fp64")
        scale = float(ref.abs().max()) or 1.0
        NL = self._nlayout
        while True:
            err = float((self.evaluate(X, precision="mixed") - ref).abs().max()) / scale
            self._mixed_err = err
            if err <= MIXED_TARGET or NL.n_head >= NL.basis_n:
                break
            newton.set_head(self._nform, NL, min(NL.basis_n, 4 * NL.n_head))
            self._upload_newton()
        self._n_inputs, self._warned_mixed = keep, warned

    def _proxy_samples(self):
        """Sample points: half from the nodes' measure (Gauss-Hermite: mean + scaling * N(0, 1/2)
        per dim; Leja: uniform on each dim's domain), half uniform over the span of each dim's
        nodes."""
        rng = np.random.default_rng(12345)
        d, n = self._d_in, PROXY_SAMPLES // 2
        if getattr(self, "_maxdeg", None) is None or len(self._maxdeg) != d:
            self._maxdeg = np.zeros(d, dtype=np.int64)
            for nu in indices.indexset(self._k, self._t):
                for j, v in nu:
                    self._maxdeg[j] = max(self._maxdeg[j], v)
        # Leja: uniform on the reference interval [-1, 1], the domain's measure
        Z = rng.uniform(-1.0, 1.0, (n, d)) if self._is_nested else \
            rng.standard_normal((n, d)) / np.sqrt(2.0)
        A = np.stack([np.asarray(g.scale(Z[:, j]), dtype=np.float64)
                      for j, g in enumerate(self._node_gen)], axis=1)
        B = np.empty((n, d))
        for j, g in enumerate(self._node_gen):
            pts = np.asarray(g(int(self._maxdeg[j])), dtype=np.float64)
            B[:, j] = rng.uniform(pts.min(), pts.max(), n) if pts.size > 1 else pts[0]
        return np.concatenate([A, B])

    def _check_proxy(self):
        """Measure the proxy Newton form against the FP64 barycentric path (the reference's
        algorithm) on sample points and keep it only for the paths whose accuracy it meets: all
        of them up to PROXY_TOL_FP64, the FP32-accurate ones up to PROXY_TOL_MIXED, none above.
        The proxy's error grows with the degree (its hierarchisation runs on nodes spanning the
        Gauss-Hermite ones; at degree ~30 it reaches ~1e-8)."""
        X = torch.as_tensor(self._proxy_samples(), device=self._device)
        keep = self._n_inputs
        y64 = self.evaluate(X, precision="fp64")
        self._newton_fp64_ok = True  # for this measurement only
        yn = self.evaluate(X, precision="newton-fp64")
        self._n_inputs = keep
        scale = float(y64.abs().max())
        err = float((yn - y64).abs().max()) / scale if scale > 0 else float((yn - y64).abs().max())
        self._proxy_err = err
        self._newton_fp64_ok = err <= PROXY_TOL_FP64
        if err > PROXY_TOL_MIXED:
            self._newton_off = (f"the Newton form on the proxy grid is {err:.1e} from the FP64 "
                                f"barycentric path (above {PROXY_TOL_MIXED:.0e})")
            for attr, fn in (("_nplan", "sbi_newton_plan_destroy"),):
                N.fn[fn](getattr(self, attr))
            self._nplan = None
            self._nform = self._nlayout = None

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
        eq = packing.entry_quadrature(L)
        t["ent_q"] = _dev_tensor(eq if eq.size else [0.0], f64, dev)
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
        vt = self._values
        st = torch.cuda.current_stream(dev)
        V = _dev_tensor(vt.V if vt.V.size else np.zeros((1, self._d_out)), f64, dev)
        rn = _dev_tensor(vt.row_node if vt.row_node.size else [-1], i32, dev)
        rs = _dev_tensor(vt.row_scale if vt.row_scale.size else [0.0], f64, dev)
        t["F"] = torch.empty((max(L.k_raw, 1), self._d_out), dtype=f64, device=dev)
        N.check(N.fn["sbi_pack_rows_f64"](V.data_ptr(), len(vt.V), self._d_out, rn.data_ptr(),
                                          rs.data_ptr(), max(L.k_raw, 1), t["F"].data_ptr(),
                                          ctypes.c_void_p(st.cuda_stream)), "pack_rows")
        t["offset"] = _dev_tensor(vt.offset, f64, dev)
        t["blk_desc"] = _dev_tensor(L.blk_desc, i32, dev)
        t["seg_off"] = _dev_tensor(L.seg_off, i32, dev)
        t["seg_slot"] = _dev_tensor(L.seg_slot, i32, dev)
        t["seg_b0"] = _dev_tensor(L.seg_b0, i32, dev)
        t["seg_b1"] = _dev_tensor(L.seg_b1, i32, dev)
        t["dim_blk"] = _dev_tensor(L.dim_blk, i32, dev)
        t["dim_blk_idx"] = _dev_tensor(L.dim_blk_idx, i32, dev)
        t["dim_blk_pos"] = _dev_tensor(L.dim_blk_pos, i32, dev)
        # record program of the d_out = 1 eval kernel
        prog = packing.build_eval_program(L) if self._d_out == 1 and self._prog is None else self._prog
        self._prog = prog
        if prog is not None:
            t["pseg_off"] = _dev_tensor(prog.seg_off, i32, dev)
            t["pseg"] = _dev_tensor(prog.seg, i32, dev)
            t["pbasis"] = _dev_tensor(prog.basis, f64, dev)
            t["prow"] = _dev_tensor(prog.rows, i32, dev)
            pn, ps = packing.program_rows(prog, vt.row_node, vt.row_scale)
            pn, ps = _dev_tensor(pn, i32, dev), _dev_tensor(ps, f64, dev)
            t["pF"] = torch.empty(len(prog.f_index), dtype=f64, device=dev)
            N.check(N.fn["sbi_pack_rows_f64"](V.data_ptr(), len(vt.V), 1, pn.data_ptr(),
                                              ps.data_ptr(), len(prog.f_index), t["pF"].data_ptr(),
                                              ctypes.c_void_p(st.cuda_stream)), "pack_rows")
        st.synchronize()  # F is read on other streams (serving) once set_f returns
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
        if prog is not None:
            d.n_pseg = int(prog.seg_off[-1])
            d.ones_row = prog.ones_row
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
            # one spare element so that 16-byte copies of its rows never pass its end
            "dim_xi": _dev_tensor(np.append(NL.dim_xi.reshape(-1), 0.0), f64, dev),
            "dim_c": _dev_tensor(NL.dim_c, f64, dev),
            "grp": _dev_tensor(NL.grp, i32, dev),
            "grp_next_dim": _dev_tensor(NL.grp_next_dim, i32, dev),
            "wave_grp": _dev_tensor(NL.wave_grp, i32, dev),
            "node_s32": _dev_tensor(NL.node_s32, f32, dev),
            "node_s64": _dev_tensor(NL.node_s64, f64, dev),
            "node_q": _dev_tensor(NL.node_q, i32, dev),
            "grp_rec": torch.as_tensor(NL.grp_rec.view(np.uint8).copy()).to(dev),
            "grp_j": _dev_tensor(NL.grp_j, i32, dev),
            "hs_dim": _dev_tensor(NL.hs_dim, i32, dev),
            "hs_kmax": _dev_tensor(NL.hs_kmax, i32, dev),
            "hs_row": _dev_tensor(NL.hs_row, i32, dev),
            "int_rec": _dev_tensor(NL.int_rec, i32, dev),
            "basis_desc": torch.as_tensor(NL.basis_desc.copy()).to(dev),
            "tab_dim": _dev_tensor(NL.tab_dim, i32, dev),
            "tab_row": _dev_tensor(NL.tab_row, i32, dev),
            "tab_kmax": _dev_tensor(NL.tab_kmax, i32, dev),
            "vjp_slot_off": _dev_tensor(NL.vjp_slot_off, i32, dev),
            "vjp_ent": torch.as_tensor(NL.vjp_ent.copy()).to(dev),
            "vjp_dim_off": _dev_tensor(NL.vjp_dim_off, i32, dev),
            "jac_unit": _dev_tensor(NL.jac_unit, i32, dev),
            "jac_red": _dev_tensor(NL.jac_red, i32, dev),
            "vjp_ent_row": _dev_tensor(NL.vjp_ent_row, i32, dev),
            "row_tdim": _dev_tensor(NL.row_tdim, i32, dev),
        }
        sing_off, sing_x = self._singular_coordinates()
        # the Jacobian kernels mark NaNs while writing the dims that have factors: every dim with
        # singular coordinates is one of them (a block's dims are dims of Lambda's nodes)
        sing_dims = np.nonzero(np.diff(sing_off))[0]
        if not np.isin(sing_dims, NL.tab_dim[: NL.n_tab_dims]).all():
            raise RuntimeError("singular coordinates in a dim without factors")
        t["sing_off"], t["sing_x"] = _dev_tensor(sing_off, i32, dev), _dev_tensor(sing_x, f64, dev)
        if NL.node_sq is not None:  # staged kernel tables (else the unstaged kernel runs)
            t["node_sq"] = _dev_tensor(NL.node_sq, i32, dev)
        if getattr(NL, "mt_a", None) is not None:  # matrix-core kernel tables
            t["rint_rec"] = _dev_tensor(NL.rint_rec, i32, dev)
            t["rint_h"] = _dev_tensor(NL.rint_h, i32, dev)
            t["mt_a"] = _dev_tensor(NL.mt_a, f32, dev)
            t["mt_a64"] = _dev_tensor(NL.mt_a64, f64, dev)
            t["mt_grp"] = torch.as_tensor(NL.mt_grp.view(np.uint8).copy()).to(dev)
            t["mi_rec"] = _dev_tensor(NL.mi_rec, i32, dev)
            t["mwave_item"] = _dev_tensor(NL.mwave_item, i32, dev)
        # surpluses in nf.nus order (adjoint), and in the basis-row order of the GEMM paths (the
        # FP64 head rows first)
        self._S = {"f64": _dev_tensor(self._nform.s, f64, dev)}
        self._Sg = {"f64": _dev_tensor(self._nform.s[NL.basis_order], f64, dev)}
        self._head_fac = _dev_tensor(NL.head_fac, i32, dev)
        # the mixed GEMM path's operands, for sbi_newton_eval_mixed_gemm
        nh = NL.n_head
        t["head_fac"] = self._head_fac
        t["gemm_s64"] = self._Sg["f64"][:nh]
        t["gemm_sq"] = self._surpluses_k4(nh)
        self._ndev = t
        d = N.NewtonDesc()
        d.d_in, d.d_out = self._d_in, self._d_out
        d.n_nodes, d.n_internal, d.n_groups = NL.n_nodes, NL.n_internal, NL.n_groups
        d.n_split, d.n_levels, d.kmax1 = newton.N_WAVES, NL.n_levels, NL.kmax1
        d.n_nodes_pad, d.n_hs_dims, d.n_hs_rows = NL.n_nodes_pad, NL.n_hs_dims, NL.n_hs_rows
        d.basis_n, d.n_tab_dims, d.n_tab_rows = NL.basis_n, NL.n_tab_dims, NL.n_tab_rows
        d.n_vjp_ent = NL.n_vjp_ent
        d.n_jac_units, d.n_jac_partials, d.n_jac_red = NL.n_jac_units, NL.n_jac_partials, NL.n_jac_red
        d.gemm_head, d.gemm_kq = nh, t["gemm_sq"].shape[0] * 4
        d.ld_s64 = max(t["gemm_s64"].stride(0), self._d_out)
        if "mt_a" in t:
            d.rw_n = NL.rw_n
            d.mt_n, d.mt_steps, d.mi_n = NL.mt_n, NL.mt_steps, NL.mi_n
            d.n_xi_dims = NL.n_xi_dims
        for name, ten in t.items():
            setattr(d, name, ten.data_ptr())
        self._ndesc = d
        if self._nplan is not None:
            N.fn["sbi_newton_plan_destroy"](self._nplan)
            self._nplan = None
        plan = ctypes.c_void_p()
        rc = N.fn["sbi_newton_plan_create"](ctypes.byref(d), ctypes.byref(plan))
        if rc == N.SBI_ERR_CAPACITY:  # too many internal nodes for the on-chip table
            self._newton_off = N.fn["sbi_last_error"]().decode(errors="replace")
            self._nform = self._nlayout = None
            return
        N.check(rc, "newton_plan")
        self._nplan = plan

    def _singular_coordinates(self):
        """CSR over input dims of the coordinates at which the reference's gradient is NaN: the
        nodes of every (dim, degree) factor of the barycentric blocks (barycentric.py:152-154
        masks (x - xi_mu)^2 == 0 for mu <= nu_i)."""
        L = self._layout
        per_dim = [[] for _ in range(self._d_in)]
        for s, (dim, deg) in enumerate(L.slot_key):
            per_dim[dim].extend(L.slot_nodes[s, : deg + 1].tolist())
        vals = [np.unique(np.asarray(v, dtype=np.float64)) for v in per_dim]
        off = np.zeros(self._d_in + 1, dtype=np.int32)
        off[1:] = np.cumsum([len(v) for v in vals])
        x = np.concatenate(vals) if off[-1] else np.zeros(1)
        return off, x.astype(np.float64)

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
            if x.dtype != torch.float64 or x.device != self._device:
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
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {PRECISIONS}")
        if precision == "mixed" and self._nplan is None:
            if not self._warned_mixed:
                warnings.warn(f"precision 'mixed': the Newton form is unavailable "
                              f"({self._newton_off}); evaluating with 'bary-fp32' (FP32 "
                              f"barycentric, about 1e-4 relative)", SmolyaxPrecisionWarning,
                              stacklevel=2)
                self._warned_mixed = True
            precision = "bary-fp32"
        if precision == "newton-fp64" and self._nplan is not None and not self._newton_fp64_ok:
            raise ValueError(f"precision 'newton-fp64': the proxy Newton form of this non-nested "
                             f"interpolant is {self._proxy_err:.1e} from the FP64 barycentric "
                             f"path (above {PROXY_TOL_FP64:.0e}); use 'fp64'")
        x = self._validate_input(x)
        P = x.shape[0]
        if P == 0:  # empty batch: nothing to launch (an empty tensor may have no data pointer)
            return torch.empty((0, self._d_out), dtype=torch.float64, device=self._device)
        if precision == "bary-fp32":
            return self._eval_bary_gemm(x, fp32=True)
        if precision != "fp64" and self._nplan is None:
            raise ValueError(f"precision {precision!r}: {self._newton_off}")
        # the Newton streaming kernels take d_out == 1; the barycentric one small d_out
        gemm = (self._d_out >= GEMM_MIN_DOUT or precision == "bf16"
                or (precision != "fp64" and self._d_out > 1))
        if not gemm:
            y = torch.empty((P, self._d_out), dtype=torch.float64, device=self._device)
            if precision == "fp64":
                rc = N.fn["sbi_bary_eval_f64"](self._plan, x.data_ptr(), P, x.stride(0),
                                               y.data_ptr(), self._stream())
                if rc == N.SBI_ERR_CAPACITY:  # prefix table too large for LDS: GEMM form
                    return self._eval_bary_gemm(x)
                N.check(rc, "eval")
            else:
                fn = "sbi_newton_eval_mixed" if precision == "mixed" else "sbi_newton_eval_f64"
                rc = N.fn[fn](self._nplan, x.data_ptr(), P, x.stride(0), y.data_ptr(),
                              self._stream())
                if rc == N.SBI_ERR_CAPACITY:  # internal-node table beyond LDS: GEMM form
                    return self._eval_newton_gemm(x, precision)
                N.check(rc, precision)
            return y
        if precision == "fp64":
            return self._eval_bary_gemm(x)
        return self._eval_newton_gemm(x, precision)

    def _point_chunk(self, rows, itemsize, P):
        """Points per chunk so that the basis matrix stays within memory_limit GiB, and within
        the 65535 * 64 points one call of the contraction, Jacobian and adjoint entry points
        takes."""
        per_point = max(1, rows) * itemsize
        chunk = int(self._memory_limit * 1024**3 // per_point) // 64 * 64 or 64
        return max(64, min(P, chunk, MAX_CALL_POINTS))

    @staticmethod
    def _ld(t):
        """Row stride of a row-major matrix (a single row may carry any stride in torch)."""
        return t.stride(0) if t.shape[0] > 1 else max(t.shape[1], 1)

    def _contract(self, A, B, offset, y):
        """y[P, n] = offset + A^T B on the matrix cores (sbi_contract: FP64 or FP32 MFMA with
        FP64 accumulation; rows independent of the batch)."""
        code = {torch.float64: 0, torch.float32: 1}[A.dtype]
        ld = self._ld
        N.check(N.fn["sbi_contract"](A.data_ptr(), ld(A), B.data_ptr(), ld(B),
                                     None if offset is None else offset.data_ptr(), y.data_ptr(),
                                     ld(y), y.shape[0], y.shape[1], A.shape[0], code,
                                     self._stream()), "contract")

    def _eval_bary_gemm(self, x, fp32=False, out=None):
        """y = offset + W F: barycentric entry weights (HIP kernel), then the contraction with the
        zeta-scaled values on the matrix cores (FP64, or FP32 operands with FP64 accumulation)."""
        P = x.shape[0]
        K = self._layout.k_raw
        if fp32:
            F = self._dev.get("F32")
            if F is None:
                F = self._dev["F"][:K].to(torch.float32)
                self._dev["F32"] = F
            dt, fn = torch.float32, "sbi_bary_basis_f32"
        else:
            F, dt, fn = self._dev["F"][:K], torch.float64, "sbi_bary_basis_f64"
        y = out if out is not None else torch.empty((P, self._d_out), dtype=torch.float64,
                                                     device=self._device)
        pc = self._point_chunk(K, 4 if fp32 else 8, P)
        for a in range(0, P, pc):
            b = min(P, a + pc)
            xc = x[a:b]
            WT = torch.empty((K, b - a), dtype=dt, device=self._device)
            N.check(N.fn[fn](self._plan, xc.data_ptr(), b - a, xc.stride(0), WT.data_ptr(), b - a,
                             self._stream()), "bary_basis")
            self._contract(WT, F, self._dev["offset"], y[a:b])
        return y

    def _surpluses(self, kind, gemm=False):
        """Surpluses in dtype `kind` ("f64", "f32", "bf16"): nf.nus row order, or with gemm=True
        the basis-row order of the GEMM paths."""
        store = self._Sg if gemm else self._S
        S = store.get(kind)
        if S is None:
            dt = {"f32": torch.float32, "bf16": torch.bfloat16}[kind]
            S = store["f64"].to(dt)
            store[kind] = S
        return S

    def _surpluses_t(self, kind):
        """Transposed surpluses S^T (d_out x nodes, nf.nus order) in dtype `kind` ("f64" or
        "f32"): the A operand of the adjoint's contraction G^T = S V^T."""
        key = kind + "T"
        ST = self._S.get(key)
        if ST is None:
            ST = self._surpluses(kind).t().contiguous()
            self._S[key] = ST
        return ST

    def _surpluses_k4(self, nh):
        """FP32 surpluses of basis rows nh.. in the k-interleaved layout of sbi_contract_split_k4:
        (K32 / 4, d_out, 4) with K32 = rows rounded up to a multiple of 16 (zero rows)."""
        S = self._Sg.get("f32k4")
        if S is None:
            rows = self._surpluses("f32", gemm=True)[nh:]
            k32 = -(-rows.shape[0] // 16) * 16
            pad = torch.zeros((k32, self._d_out), dtype=torch.float32, device=self._device)
            pad[: rows.shape[0]] = rows
            S = pad.view(k32 // 4, 4, self._d_out).permute(0, 2, 1).contiguous()
            self._Sg["f32k4"] = S
        return S

    def _surpluses_k8(self, nh):
        """BF16 surpluses of basis rows nh.. in the k-interleaved layout of sbi_contract_bf16_k8:
        (K16 / 8, d_out, 8) with K16 = rows rounded up to a multiple of 32 (zero rows); values
        rounded FP64 -> FP32 -> BF16 like the basis kernel's."""
        S = self._Sg.get("bf16k8")
        if S is None:
            rows = self._surpluses("f32", gemm=True)[nh:]
            k16 = -(-rows.shape[0] // 32) * 32
            pad = torch.zeros((k16, self._d_out), dtype=torch.float32, device=self._device)
            pad[: rows.shape[0]] = rows
            S = pad.view(k16 // 8, 8, self._d_out).permute(0, 2, 1).contiguous().to(torch.bfloat16)
            self._Sg["bf16k8"] = S
        return S

    def _eval_newton_gemm(self, x, precision):
        """y = H S: Newton basis matrix (HIP kernel), then the contraction with the surpluses on
        the matrix cores: sbi_contract for FP64; for "mixed" (FP32) and "bf16" the heaviest rows
        in FP64 (sbi_newton_rows_f64, FP64 MFMA) and the rest in FP32 or BF16, k-interleaved
        operands, FP64 accumulation ("mixed": all three steps in one call,
        sbi_newton_eval_mixed_gemm, which splits k over more workgroups for small batches)."""
        P = x.shape[0]
        Nn = self._nlayout.basis_n
        kind = {"newton-fp64": "f64", "mixed": "f32", "bf16": "bf16"}[precision]
        st = self._stream()
        y = torch.empty((P, self._d_out), dtype=torch.float64, device=self._device)
        if kind == "f64":
            S = self._surpluses("f64", gemm=True)
            pc = self._point_chunk(Nn, 8, P)
            for a in range(0, P, pc):
                b = min(P, a + pc)
                xc = x[a:b]
                HT = torch.empty((Nn, b - a), dtype=torch.float64, device=self._device)
                N.check(N.fn["sbi_newton_basis_rows"](self._nplan, xc.data_ptr(), b - a,
                                                      xc.stride(0), 0, Nn, HT.data_ptr(), b - a, 0,
                                                      st), "basis")
                self._contract(HT, S, None, y[a:b])
            return y
        if kind == "f32":  # one call per chunk: head rows, FP32 rows, contraction
            pc = self._point_chunk(self._ndesc.gemm_kq + 2 * self._ndesc.gemm_head, 4, P)
            for a in range(0, P, pc):
                b = min(P, a + pc)
                xc, yc = x[a:b], y[a:b]
                ws = N.fn["sbi_newton_eval_mixed_gemm_workspace"](self._nplan, b - a)
                W = torch.empty(max(ws, 1), dtype=torch.uint8, device=self._device)
                N.check(N.fn["sbi_newton_eval_mixed_gemm"](self._nplan, xc.data_ptr(), b - a,
                                                           xc.stride(0), yc.data_ptr(),
                                                           yc.stride(0), W.data_ptr(), ws, st),
                        "mixed")
            return y
        # FP64 head rows, BF16 remainder in the k-interleaved layout (16-byte operand loads)
        nh = self._nlayout.n_head
        S64 = self._surpluses("f64", gemm=True)[:nh]
        SQ = self._surpluses_k8(nh)
        KQ = SQ.shape[0] * 8
        pc = self._point_chunk(KQ + 2 * nh, 2, P)
        for a in range(0, P, pc):
            b = min(P, a + pc)
            xc = x[a:b]
            HT = torch.empty((KQ // 8, b - a, 8), dtype=torch.bfloat16, device=self._device)
            if Nn > nh:
                N.check(N.fn["sbi_newton_basis_rows"](self._nplan, xc.data_ptr(), b - a,
                                                      xc.stride(0), nh, Nn, HT.data_ptr(), b - a,
                                                      4, st), "basis")
            H64 = torch.empty((max(nh, 1), b - a), dtype=torch.float64, device=self._device)
            if nh:
                N.check(N.fn["sbi_newton_rows_f64"](self._nplan, xc.data_ptr(), b - a,
                                                    xc.stride(0), self._head_fac.data_ptr(),
                                                    nh, H64.data_ptr(), b - a,
                                                    st), "rows")
            yc = y[a:b]
            N.check(N.fn["sbi_contract_bf16_k8"](
                H64.data_ptr(), H64.stride(0), S64.data_ptr(), max(S64.stride(0), self._d_out),
                nh, HT.data_ptr() if KQ else None, b - a, SQ.data_ptr() if KQ else None,
                self._d_out, KQ, None, yc.data_ptr(), yc.stride(0), b - a, self._d_out,
                st), "contract")
        return y

    def vjp(self, x, v, precision: str = None) -> torch.Tensor:
        """Vector-Jacobian product g[p, j] = sum_m v[p, m] dy_m/dx_j (FP64 result).

        Nested nodes: hand-written adjoint of the Newton form after G^T = S V^T on the matrix
        cores (sbi_contract; "mixed" and "bf16" take FP32 operands with FP64 accumulation).
        Without a Newton form that meets FP64 accuracy: the barycentric Jacobian contracted with
        v (interpolation.py:306-345).
        """
        precision = precision or self._precision
        x = self._validate_input(x)
        P = x.shape[0]
        v = torch.as_tensor(v, device=self._device).reshape(P, self._d_out)
        if P == 0:
            return torch.empty((0, self._d_in), dtype=torch.float64, device=self._device)
        kind, code = ("f32", 1) if precision in ("mixed", "bf16") else ("f64", 0)
        if self._nplan is None or (kind == "f64" and not self._newton_fp64_ok):
            # barycentric Jacobian contracted with v over the outputs, in point chunks of at
            # most memory_limit GiB of Jacobian
            gx = torch.empty((P, self._d_in), dtype=torch.float64, device=self._device)
            chunk = self._point_chunk(self._d_out * self._d_in, 8, P)
            for p0 in range(0, P, chunk):
                p1 = min(P, p0 + chunk)
                J = self.gradient(x[p0:p1])
                gx[p0:p1] = (v[p0:p1].to(torch.float64).unsqueeze(2) * J).sum(dim=1)
            return gx
        ST = self._surpluses_t(kind)
        dt = ST.dtype
        n_nodes = ST.shape[1]
        gx = torch.empty((P, self._d_in), dtype=torch.float64, device=self._device)
        # point chunks: G^T (nodes x points, FP64) and the workspace within memory_limit
        ws1 = int(N.fn["sbi_newton_vjp_workspace"](self._nplan, 64))
        chunk = self._point_chunk(n_nodes * 8 + max(0, ws1) // 64 + 1, 1, P)
        for p0 in range(0, P, chunk):
            p1 = min(P, p0 + chunk)
            xc = x[p0:p1]
            # G^T[nu, p] = sum_m S[nu, m] v[p, m] on the matrix cores (sbi_contract with
            # A = S^T and B = v^T; FP64, or FP32 operands with FP64 accumulation)
            VT = v[p0:p1].t().to(dt).contiguous()
            GT = torch.empty((n_nodes, p1 - p0), dtype=torch.float64, device=self._device)
            self._contract(ST, VT, None, GT)
            nb = int(N.fn["sbi_newton_vjp_workspace"](self._nplan, p1 - p0))
            ws = torch.empty(max(1, (nb + 7) // 8), dtype=torch.float64, device=self._device)
            N.check(N.fn["sbi_newton_vjp"](self._nplan, xc.data_ptr(), p1 - p0, xc.stride(0),
                                           GT.data_ptr(), GT.stride(0), 0,
                                           gx[p0:p1].data_ptr(), ws.data_ptr(), nb,
                                           self._stream()), "vjp")
        return gx

    def differentiable(self, x: torch.Tensor, precision: str = None) -> torch.Tensor:
        """__call__ as a torch.autograd.Function: backward runs the hand-written adjoint."""
        from .autograd import SmolyakEval

        return SmolyakEval.apply(x, self, precision or self._precision)

    def gradient(self, x) -> torch.Tensor:
        """interpolation.py:306-345: Jacobian J[P, d_out, d_in] in float64."""
        x = self._validate_input(x)
        P = x.shape[0]
        J = torch.empty((P, self._d_out, self._d_in), dtype=torch.float64, device=self._device)
        if P == 0:
            return J
        if self._nplan is not None and self._newton_fp64_ok:
            self._jacobian_newton(x, J)
        else:
            N.check(N.fn["sbi_bary_jacobian_f64"](self._plan, x.data_ptr(), P, x.stride(0),
                                                  J.data_ptr(), self._stream()), "jacobian")
        return J

    def _jacobian_newton(self, x, J):
        """J of the Newton form (same polynomial, 6-8x fewer terms than the barycentric blocks):
        sbi_newton_jacobian_f64 in point chunks whose workspace stays within memory_limit."""
        P, n = x.shape[0], self._d_out
        ws1 = int(N.fn["sbi_newton_jacobian_workspace"](self._nplan, 64, n))
        chunk = self._point_chunk(max(1, ws1 // 64), 1, P)
        S = self._S["f64"]
        for p0 in range(0, P, chunk):
            p1 = min(P, p0 + chunk)
            xc = x[p0:p1]
            nb = int(N.fn["sbi_newton_jacobian_workspace"](self._nplan, p1 - p0, n))
            ws = torch.empty(max(1, (nb + 7) // 8), dtype=torch.float64, device=self._device)
            N.check(N.fn["sbi_newton_jacobian_f64"](self._nplan, xc.data_ptr(), p1 - p0,
                                                    xc.stride(0), S.data_ptr(), S.stride(0), n,
                                                    J[p0:p1].data_ptr(), ws.data_ptr(), nb,
                                                    self._stream()), "jacobian")

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

    @property
    def mixed_path(self) -> str:
        """The path that serves precision 'mixed': "newton" or "bary-fp32" (Newton form
        unavailable, see the constructor)."""
        return "newton" if self._nplan is not None else "bary-fp32"

    @property
    def mixed_error(self):
        """d_out > 1 with a Newton form: the mixed GEMM path's largest relative distance from the
        FP64 path on sample points, after its FP64 head was sized (None otherwise)."""
        return self._mixed_err

    @property
    def mixed_head_rows(self):
        """Newton basis rows the mixed GEMM path contracts in FP64 (None without a Newton form)."""
        return None if self._nlayout is None else int(self._nlayout.n_head)

    @property
    def proxy_error(self):
        """Non-nested interpolants: largest relative distance of the proxy Newton form from the
        FP64 barycentric path on sample points (None for nested nodes or without a proxy)."""
        return self._proxy_err
