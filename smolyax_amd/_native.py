"""ctypes binding of libsmolyax_amd.so (the C ABI declared in include/smolyax_amd.h).

The library is loaded after torch so that its HIP runtime dependency resolves to the runtime torch
already mapped.  There is no fallback: a missing library raises at import time.
"""

import ctypes
import os

import torch  # noqa: F401  (maps torch's HIP runtime first)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsmolyax_amd.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: build it with `python -m smolyax_amd.build` "
        "(hipcc --offload-arch=gfx950); smolyax_amd has no CPU fallback."
    )

lib = ctypes.CDLL(LIB_PATH)

c_i32, c_i64, c_f64, c_vp = ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p
P_f64 = ctypes.POINTER(ctypes.c_double)
P_i64 = ctypes.POINTER(ctypes.c_int64)
P_i32 = ctypes.POINTER(ctypes.c_int32)


class BaryDesc(ctypes.Structure):
    _fields_ = [
        ("d_in", c_i64), ("d_out", c_i64),
        ("n_slots", c_i32), ("dmax1", c_i32), ("n_prefix_vals", c_i32), ("n_prefix_slots", c_i32),
        ("n_runs", c_i32), ("n_blocks", c_i32), ("max_n", c_i32), ("_pad", c_i32),
        ("k_raw", c_i64),
        ("slot_dim", c_vp), ("slot_deg", c_vp), ("slot_nodes", c_vp), ("slot_w", c_vp),
        ("slot_q", c_vp), ("prefix_slot", c_vp), ("prefix_off", c_vp), ("run_slot", c_vp),
        ("run_start", c_vp), ("blk_n", c_vp), ("blk_fac", c_vp), ("fac_row", c_vp),
        ("fac_deg", c_vp), ("fac_slot", c_vp), ("blk_last_slot", c_vp), ("blk_foff", c_vp),
        ("F", c_vp), ("offset", c_vp), ("blk_desc", c_vp), ("seg_off", c_vp), ("seg_slot", c_vp),
        ("seg_b0", c_vp), ("seg_b1", c_vp), ("n_split", c_i32), ("n_seg", c_i32),
        ("dim_blk", c_vp), ("dim_blk_idx", c_vp), ("dim_blk_pos", c_vp), ("ent_q", c_vp),
        ("pseg_off", c_vp), ("pseg", c_vp), ("pbasis", c_vp), ("prow", c_vp), ("pF", c_vp),
        ("n_pseg", c_i32),
        ("ones_row", c_i32),
    ]


class NewtonDesc(ctypes.Structure):
    _fields_ = [
        ("d_in", c_i64), ("d_out", c_i64),
        ("n_nodes", c_i32), ("n_internal", c_i32), ("n_groups", c_i32), ("n_split", c_i32),
        ("n_levels", c_i32), ("kmax1", c_i32),
        ("s_root", c_vp), ("int_parent", c_vp), ("int_dim", c_vp), ("int_deg", c_vp),
        ("level_off", c_vp), ("dim_xi", c_vp), ("dim_c", c_vp), ("grp", c_vp),
        ("grp_next_dim", c_vp), ("wave_grp", c_vp), ("node_s32", c_vp), ("node_s64", c_vp),
        ("node_q", c_vp),
        ("n_nodes_pad", c_i32), ("n_hs_dims", c_i32), ("n_hs_rows", c_i32), ("_pad", c_i32),
        ("node_sq", c_vp), ("grp_rec", c_vp), ("grp_j", c_vp),
        ("hs_dim", c_vp), ("hs_kmax", c_vp), ("hs_row", c_vp), ("int_rec", c_vp),
        ("basis_n", c_i32), ("n_tab_dims", c_i32), ("n_tab_rows", c_i32), ("n_vjp_ent", c_i32),
        ("basis_desc", c_vp), ("tab_dim", c_vp), ("tab_row", c_vp), ("tab_kmax", c_vp),
        ("vjp_slot_off", c_vp), ("vjp_ent", c_vp),
        ("vjp_dim_off", c_vp), ("sing_off", c_vp), ("sing_x", c_vp),
        ("n_jac_units", c_i32), ("n_jac_partials", c_i32), ("n_jac_red", c_i32), ("_pad4", c_i32),
        ("jac_unit", c_vp), ("jac_red", c_vp), ("vjp_ent_row", c_vp), ("row_tdim", c_vp),
        ("rw_n", c_i32), ("_pad6", c_i32), ("rint_rec", c_vp), ("rint_h", c_vp),
        ("mt_n", c_i32), ("mt_steps", c_i32), ("mi_n", c_i32), ("n_xi_dims", c_i32),
        ("mt_a", c_vp), ("mt_grp", c_vp), ("mi_rec", c_vp), ("mwave_item", c_vp),
        ("gemm_head", c_i32), ("gemm_kq", c_i32), ("head_fac", c_vp), ("gemm_s64", c_vp),
        ("ld_s64", c_i64), ("gemm_sq", c_vp), ("mt_a64", c_vp),
    ]


def _sig(name, restype, argtypes):
    fn = getattr(lib, name)
    fn.restype = restype
    fn.argtypes = argtypes
    return fn


# name -> (restype, argtypes); every symbol declared in include/smolyax_amd.h
SIGNATURES = {
    "sbi_last_error": (ctypes.c_char_p, []),
    "sbi_version": (ctypes.c_char_p, []),
    "sbi_indexset_cardinality": (c_i64, [P_f64, c_i64, c_f64]),
    "sbi_smolyak_coefficient": (c_i64, [P_f64, c_i64, c_f64, c_i32]),
    "sbi_nodeset_cardinality": (c_i64, [P_f64, c_i64, c_f64, c_i32]),
    "sbi_find_approximate_threshold": (c_f64, [P_f64, c_i64, c_i64, c_i32, c_i32, c_f64]),
    "sbi_indexset": (c_i32, [P_f64, c_i64, c_f64, P_i64, P_i64, P_i64, P_i32, P_i32]),
    "sbi_nonzero_indices_and_zetas": (c_i32, [P_f64, c_i64, c_f64, P_i64, P_i64, P_i64, P_i32,
                                             P_i32, P_i64]),
    "sbi_bary_plan_create": (c_i32, [ctypes.POINTER(BaryDesc), ctypes.POINTER(c_vp)]),
    "sbi_plan_destroy": (c_i32, [c_vp]),
    "sbi_bary_eval_f64": (c_i32, [c_vp, c_vp, c_i64, c_i64, c_vp, c_vp]),
    "sbi_bary_jacobian_f64": (c_i32, [c_vp, c_vp, c_i64, c_i64, c_vp, c_vp]),
    "sbi_bary_integral_workspace": (c_i64, [c_vp]),
    "sbi_bary_integral_f64": (c_i32, [c_vp, c_vp, c_vp, c_vp]),
    "sbi_bary_integral_f64_noalloc": (c_i32, [c_vp, c_vp, c_vp]),
    "sbi_newton_plan_create": (c_i32, [ctypes.POINTER(NewtonDesc), ctypes.POINTER(c_vp)]),
    "sbi_newton_plan_destroy": (c_i32, [c_vp]),
    "sbi_newton_eval_mixed": (c_i32, [c_vp, c_vp, c_i64, c_i64, c_vp, c_vp]),
    "sbi_newton_eval_f64": (c_i32, [c_vp, c_vp, c_i64, c_i64, c_vp, c_vp]),
    "sbi_newton_basis": (c_i32, [c_vp, c_vp, c_i64, c_i64, c_vp, c_i64, c_i32, c_vp]),
    "sbi_newton_vjp_workspace": (c_i64, [c_vp, c_i64]),
    "sbi_newton_vjp": (c_i32, [c_vp, c_vp, c_i64, c_i64, c_vp, c_i64, c_i32, c_vp, c_vp, c_i64,
                               c_vp]),
    "sbi_newton_jacobian_workspace": (c_i64, [c_vp, c_i64, c_i64]),
    "sbi_newton_jacobian_f64": (c_i32, [c_vp, c_vp, c_i64, c_i64, c_vp, c_i64, c_i64, c_vp, c_vp,
                                        c_i64, c_vp]),
    "sbi_bary_basis_f64": (c_i32, [c_vp, c_vp, c_i64, c_i64, c_vp, c_i64, c_vp]),
    "sbi_bary_basis_f32": (c_i32, [c_vp, c_vp, c_i64, c_i64, c_vp, c_i64, c_vp]),
    "sbi_newton_mixed_staged": (c_i32, [c_vp]),
    "sbi_newton_f64_matrix_cores": (c_i32, [c_vp]),
    "sbi_newton_basis_rows": (c_i32, [c_vp, c_vp, c_i64, c_i64, c_i32, c_i32, c_vp, c_i64, c_i32,
                                      c_vp]),
    "sbi_newton_rows_f64": (c_i32, [c_vp, c_vp, c_i64, c_i64, c_vp, c_i32, c_vp, c_i64, c_vp]),
    "sbi_contract_split_k4": (c_i32, [c_vp, c_i64, c_vp, c_i64, c_i64, c_vp, c_i64, c_vp, c_i64,
                                      c_i64, c_vp, c_vp, c_i64, c_i64, c_i64, c_vp]),
    "sbi_contract_split_k4_workspace": (c_i64, [c_i64, c_i64, c_i64, c_i64]),
    "sbi_newton_eval_mixed_gemm_workspace": (c_i64, [c_vp, c_i64]),
    "sbi_newton_eval_mixed_gemm": (c_i32, [c_vp, c_vp, c_i64, c_i64, c_vp, c_i64, c_vp, c_i64,
                                           c_vp]),
    "sbi_contract_split_k4_ws": (c_i32, [c_vp, c_i64, c_vp, c_i64, c_i64, c_vp, c_i64, c_vp, c_i64,
                                         c_i64, c_vp, c_vp, c_i64, c_i64, c_i64, c_vp, c_i64,
                                         c_vp]),
    "sbi_contract_bf16_k8": (c_i32, [c_vp, c_i64, c_vp, c_i64, c_i64, c_vp, c_i64, c_vp, c_i64,
                                     c_i64, c_vp, c_vp, c_i64, c_i64, c_i64, c_vp]),
    "sbi_contract_split": (c_i32, [c_vp, c_i64, c_vp, c_i64, c_i64, c_vp, c_i64, c_vp, c_i64, c_i64,
                                   c_vp, c_vp, c_i64, c_i64, c_i64, c_vp]),
    "sbi_contract": (c_i32, [c_vp, c_i64, c_vp, c_i64, c_vp, c_vp, c_i64, c_i64, c_i64, c_i64,
                             c_i32, c_vp]),
    "sbi_pack_rows_f64": (c_i32, [c_vp, c_i64, c_i64, c_vp, c_vp, c_i64, c_vp, c_vp]),
    "sbi_hierarchize_f64": (c_i32, [c_vp, c_i64, c_i64, c_i64, c_vp, c_i32, c_vp, c_vp, c_vp, c_vp,
                                    c_vp]),
    "sbi_server_create": (c_i32, [c_i32, c_vp]),
    "sbi_server_add": (c_i32, [c_vp, c_vp, c_i64, c_vp]),
    "sbi_server_submit": (c_i32, [c_vp, c_i32, c_vp, c_i64, c_i64, c_vp, c_i64, c_vp, c_vp]),
    "sbi_server_step": (c_i32, [c_vp, c_vp, c_i32, c_vp, c_vp]),
    "sbi_server_destroy": (c_i32, [c_vp]),
    "sbi_bary_plan_desc": (c_i32, [c_vp, c_vp]),
    "sbi_newton_plan_desc": (c_i32, [c_vp, c_vp]),
}

fn = {name: _sig(name, r, a) for name, (r, a) in SIGNATURES.items()}


SBI_OK, SBI_ERR_INVALID, SBI_ERR_HIP, SBI_ERR_UNSUPPORTED, SBI_ERR_CAPACITY = 0, -1, -2, -3, -4


class NativeError(RuntimeError):
    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


def check(rc, what=""):
    if rc != 0:
        msg = fn["sbi_last_error"]().decode(errors="replace")
        raise NativeError(f"{what} failed (code {rc}): {msg}", rc)


def f64_ptr(a):
    """numpy float64 array -> ctypes pointer (array must stay alive)."""
    return a.ctypes.data_as(P_f64)


def i64_ptr(a):
    return a.ctypes.data_as(P_i64)


def i32_ptr(a):
    return a.ctypes.data_as(P_i32)
