/*
 * smolyax_amd — C ABI of the MI355X evaluation engine for barycentric Smolyak interpolants.
 *
 * The reference library (JoWestermann/smolyax) has no FFI; its boundary is the Python class
 * SmolyakBarycentricInterpolator (src/smolyax/interpolation.py:15-390) and the per-group operator
 * it jit/vmaps (interpolation.py:241-248).  Every entry point below names the reference function
 * whose work it replaces.  Plain pointers and sizes only; device pointers are borrowed (the caller,
 * e.g. torch, owns device memory); no entry point allocates device memory or synchronises the
 * stream, except the serving scheduler (sbi_server_*), which owns its streams and staging buffers;
 * return value 0 = success, < 0 = error (message via sbi_last_error(), thread-local).
 */
#ifndef SMOLYAX_AMD_H
#define SMOLYAX_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ----------------------------------------------------------------------------------------------
 * Error codes and diagnostics
 * -------------------------------------------------------------------------------------------- */
enum {
  SBI_OK = 0,
  SBI_ERR_INVALID = -1,   /* bad argument / descriptor */
  SBI_ERR_HIP = -2,       /* a HIP runtime call failed */
  SBI_ERR_UNSUPPORTED = -3,
  SBI_ERR_CAPACITY = -4   /* a size exceeds a documented limit */
};

/* Message of the last failing call on this thread ("" if none). */
const char* sbi_last_error(void);
/* Library version string. */
const char* sbi_version(void);

/* ----------------------------------------------------------------------------------------------
 * Multi-index sets (host, exact).  Bit-identical restatements of the numba kernels of
 * src/smolyax/indices.py: same traversal order, same sequence of IEEE double operations.
 * `k` is the anisotropy vector (length d), `t` the threshold.  A NULL k with d > 0 or a negative
 * d is an error: -1 from the integer functions, NaN from the threshold search, SBI_ERR_INVALID
 * from the two set builders (sbi_last_error() says why).
 * -------------------------------------------------------------------------------------------- */

/* indices.py:72-118  indexset_cardinality(k, t) */
int64_t sbi_indexset_cardinality(const double* k, int64_t d, double t);
/* indices.py:121-168  smolyak_coefficient(k, d, rem_t, parity) */
int64_t sbi_smolyak_coefficient(const double* k, int64_t d, double rem_t, int32_t parity);
/* indices.py:262-289  nodeset_cardinality(k, t, nested) (non-nested branch: indices.py:175-211) */
int64_t sbi_nodeset_cardinality(const double* k, int64_t d, double t, int32_t nested);
/* indices.py:292-355  find_approximate_threshold(k, m, nested, max_iter, accuracy) */
double sbi_find_approximate_threshold(const double* k, int64_t d, int64_t m, int32_t nested,
                                      int32_t max_iter, double accuracy);

/* indices.py:20-69  indexset(k, t).  Two-call pattern: with dims == NULL only the sizes are
 * returned (*n_nu = |Lambda|, *n_entries = total number of (dim, deg) pairs).  With buffers,
 * fills offsets[n_nu+1] (CSR into dims/degs) in the reference's DFS output order. */
int sbi_indexset(const double* k, int64_t d, double t, int64_t* n_nu, int64_t* n_entries,
                 int64_t* offsets, int32_t* dims, int32_t* degs);

/* indices.py:214-259  non_zero_indices_and_zetas(k, t): multi-indices with zeta != 0 in the
 * reference's discovery order (its per-n lists are this order filtered by n).  Same two-call
 * pattern; zetas[n_nu] receives the Smolyak coefficients. */
int sbi_nonzero_indices_and_zetas(const double* k, int64_t d, double t, int64_t* n_nu,
                                  int64_t* n_entries, int64_t* offsets, int32_t* dims,
                                  int32_t* degs, int64_t* zetas);

/* ----------------------------------------------------------------------------------------------
 * Evaluation plans.  A plan is the packed device layout of one interpolant (built by the host
 * from set_f's packing, interpolation.py:140-235).  All pointers in the descriptor are DEVICE
 * pointers owned by the caller and must outlive the plan; the plan copies only scalars.
 * -------------------------------------------------------------------------------------------- */

/* Barycentric combination-technique layout (the reference's algorithm, ragged instead of padded).
 * Blocks are the multi-indices nu != 0 with zeta_nu != 0; each block's factors are its active
 * dims in ascending order; the LAST factor of a block is "streamed", the others are "prefix"
 * factors whose 1D bases live in an on-chip table.  Blocks are sorted by last slot (runs). */
typedef struct sbi_bary_desc {
  int64_t d_in, d_out;
  int32_t n_slots;          /* distinct (dim, deg) pairs */
  int32_t dmax1;            /* max deg + 1 over slots (row stride of slot_nodes / slot_w) */
  int32_t n_prefix_vals;    /* sum over prefix slots of (deg + 1): on-chip table rows */
  int32_t n_prefix_slots;
  int32_t n_runs;           /* runs of consecutive blocks sharing their last slot */
  int32_t n_blocks;
  int32_t max_n;            /* max active dims per block */
  int32_t _pad;
  int64_t k_raw;            /* sum over blocks of prod(deg_i + 1) */
  const int32_t* slot_dim;       /* [n_slots] */
  const int32_t* slot_deg;       /* [n_slots] */
  const double* slot_nodes;      /* [n_slots * dmax1] node values xi_mu */
  const double* slot_w;          /* [n_slots * dmax1] first-form weights 1/prod_{nu!=mu}(xi_mu-xi_nu) */
  const double* slot_q;          /* [n_slots * dmax1] quadrature weights (integral) */
  const int32_t* prefix_slot;    /* [n_prefix_slots] slot id of each prefix slot */
  const int32_t* prefix_off;     /* [n_prefix_slots] first table row of each prefix slot */
  const int32_t* run_slot;       /* [n_runs] last slot of the run */
  const int32_t* run_start;      /* [n_runs + 1] first block of each run */
  const int32_t* blk_n;          /* [n_blocks] active dims */
  const int32_t* blk_fac;        /* [n_blocks + 1] CSR start into fac_* (prefix factors only) */
  const int32_t* fac_row;        /* prefix factor: first table row of its slot */
  const int32_t* fac_deg;        /* prefix factor degree */
  const int32_t* fac_slot;       /* prefix factor slot id */
  const int32_t* blk_last_slot;  /* [n_blocks] streamed (last) slot of each block, -1 if none */
  const int64_t* blk_foff;       /* [n_blocks + 1] offset of the block's F rows (rows of d_out) */
  const double* F;               /* [k_raw * d_out] zeta-scaled values, row-major, block entries
                                    in C order over (prefix factors..., last factor) */
  const double* offset;          /* [d_out] zeta_0 * f(xi_0) */
  const int32_t* blk_desc;       /* [4 * n_blocks] packed block record (eval kernel):
                                    [0] F row offset; [1] npre | dl<<3 | d1<<8 | d2<<13 | d3<<18
                                    (npre = 7: more than 3 prefix factors, read fac_*);
                                    [2] row1 | row2<<16; [3] row3 | (streamed slot + 1)<<16 */
  const int32_t* seg_off;        /* [n_split + 1] per wave: range of its segments */
  const int32_t* seg_slot;       /* [n_seg] streamed slot of a segment (-1: none) */
  const int32_t* seg_b0;         /* [n_seg] first block of a segment */
  const int32_t* seg_b1;         /* [n_seg] one past the last block of a segment */
  int32_t n_split;               /* waves per point tile (16); segments = runs cut at wave
                                    boundaries, cost-balanced over the waves */
  int32_t n_seg;
  const int32_t* dim_blk;        /* [d_in + 1] CSR: blocks containing each dim (jacobian) */
  const int32_t* dim_blk_idx;    /* block ids, and the factor position in dim_blk_pos */
  const int32_t* dim_blk_pos;    /*   (0..n-1, n-1 = last factor) */
  const double* ent_q;           /* [k_raw] quadrature weight of every entry: the product of its
                                    factors' slot_q values (integral; optional) */
  /* Record program of the d_out = 1 eval kernel (optional; pseg NULL: per-block records).  A
   * record holds 1..3 table rows (one value of each prefix factor, the last one's values r and
   * r + 1) and the 2 * DL values of F they multiply; a program segment is a run of records of one
   * streamed slot grouped by their number of rows, then blocks left to the per-block path. */
  const int32_t* pseg_off;       /* [n_split + 1] per wave: range of its program segments */
  const int32_t* pseg;           /* [n_pseg * 8] segment headers: slot, first F value, first row,
                                    n1 | n2 << 16, n3 | DL << 16 | n_gen << 24, first per-block
                                    block, streamed dim of the wave's next segment (-1: none),
                                    streamed dim of this segment (n_k = records with k rows;
                                    DL = 0: none) */
  const double* pbasis;          /* [n_pseg * 8] nodes[4], first-form weights[4] of each
                                    segment's streamed slot (DL <= 4) */
  const int32_t* prow;           /* table rows of the records */
  const double* pF;              /* F values of the records, 2 * DL per record */
  int32_t n_pseg;
  int32_t ones_row;              /* first of two table rows of 1.0 after the prefix rows */
} sbi_bary_desc;

typedef struct sbi_plan sbi_plan;

/* Create / destroy a barycentric plan (host metadata only, validates the descriptor). */
int sbi_bary_plan_create(const sbi_bary_desc* desc, sbi_plan** out);
int sbi_plan_destroy(sbi_plan* plan);

/* interpolation.py:264-304 __call__ with barycentric.py:69-123 per block:
 * y[P, d_out] (FP64) = offset + sum_blocks zeta * tensor-product interpolant at x[P, ldx].
 * x, y device pointers; stream = hipStream_t (NULL = default stream).  Returns SBI_ERR_CAPACITY
 * when the prefix basis table of a 64-point tile exceeds LDS (more than 320 rows); the GEMM form
 * (sbi_bary_basis_f64 + a dense GEMM) evaluates such plans. */
int sbi_bary_eval_f64(const sbi_plan* plan, const double* x, int64_t P, int64_t ldx, double* y,
                      void* stream);

/* interpolation.py:306-345 gradient with barycentric.py:126-229: dense Jacobian
 * J[P, d_out, d_in] (FP64), NaN where x_j equals a node of a slot of dim j (reference
 * semantics, barycentric.py:152-154). */
int sbi_bary_jacobian_f64(const sbi_plan* plan, const double* x, int64_t P, int64_t ldx,
                          double* J, void* stream);

/* interpolation.py:347-390 integral: Q[d_out] = offset + sum_blocks zeta <F_nu, (x) q>, computed
 * as offset + ent_q^T F (zeta is folded into F; ent_q holds each entry's product of quadrature
 * weights).  F is read once; sums run in a fixed order.  `partial` is caller-provided device
 * workspace of sbi_bary_integral_workspace(plan) doubles. */
int64_t sbi_bary_integral_workspace(const sbi_plan* plan);
int sbi_bary_integral_f64(const sbi_plan* plan, double* Q, double* partial, void* stream);

/* sbi_bary_integral_f64 without a workspace argument and without any allocation: the partial sums
 * live in a device array of the library (SBI_INTEGRAL_SCRATCH_DOUBLES doubles = 4.5 MiB, one copy
 * per device, held until the process ends).  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  Same kernels and summation order as sbi_bary_integral_f64, so the same Q bit
 * for bit.  SBI_ERR_CAPACITY when the plan's partial sums exceed the array (never for d_out up to
 * 65536).  Calls that overlap on one device (several streams or threads) share the array and
 * race: serialise them. */
#define SBI_INTEGRAL_SCRATCH_DOUBLES 589824
int sbi_bary_integral_f64_noalloc(const sbi_plan* plan, double* Q, void* stream);


/* ----------------------------------------------------------------------------------------------
 * Hierarchical (Newton-surplus) form, nested nodes only (Leja).  Same interpolant as the
 * combination technique of interpolation.py:264-304, written as
 *   y(x) = s_root + sum_{nu != 0} s_nu prod_{(j,k) in nu} h^{(j)}_k(x_j),
 *   h^{(j)}_k(x) = c_k prod_{r<k} (x - xi_r).
 * It has no counterpart function in the reference: it is the mixed-precision path of
 * __call__ (FP32 surpluses / products / group sums, FP64 node differences and accumulation).
 * -------------------------------------------------------------------------------------------- */
typedef struct sbi_newton_desc {
  int64_t d_in, d_out;        /* d_out must be 1 for this kernel */
  int32_t n_nodes;            /* nodes nu != 0, in (dim, deg) group order */
  int32_t n_internal;         /* nodes with children: rows 1..n_internal of the on-chip table */
  int32_t n_groups;           /* (dim j, deg k) groups of nodes whose last factor is (j, k) */
  int32_t n_split;            /* waves per point tile (16) */
  int32_t n_levels;           /* depth levels of the internal nodes */
  int32_t kmax1;              /* row stride of dim_xi / dim_c (max deg + 1) */
  const double* s_root;       /* [d_out] surplus of nu = 0 */
  const int32_t* int_parent;  /* [n_internal] table row of the parent (0 = root) */
  const int32_t* int_dim;     /* [n_internal] dim of the node's last factor */
  const int32_t* int_deg;     /* [n_internal] deg of the node's last factor */
  const int32_t* level_off;   /* [n_levels + 1] internal nodes grouped by depth */
  const double* dim_xi;       /* [d_in * kmax1] nodes of each dim's sequence */
  const double* dim_c;        /* [d_in * kmax1] Newton constants c_k = 1/prod_{r<k}(xi_k-xi_r) */
  const int32_t* grp;         /* [4 * n_groups] first node, end node, dim, deg */
  const int32_t* grp_next_dim;/* [n_groups] next different dim in the same wave (-1: none) */
  const int32_t* wave_grp;    /* [n_split + 1] group ranges per wave (cost-balanced) */
  const float* node_s32;      /* [n_nodes] surpluses, FP32 (mixed path) */
  const double* node_s64;     /* [n_nodes] surpluses, FP64 */
  const int32_t* node_q;      /* [n_nodes] table row of the node's parent */
  /* staged mixed kernel (all optional: NULL selects the unstaged kernel) */
  int32_t n_nodes_pad;        /* nodes with every group padded to an even count */
  int32_t n_hs_dims;          /* dims of the internal nodes' factors */
  int32_t n_hs_rows;          /* rows of the per-tile h table */
  int32_t _pad;
  const int32_t* node_sq;     /* [2 * n_nodes_pad] pairs {c_deg * surplus as FP32 bits, byte offset
                                 of the parent's table row (row * 256)}; groups padded to even
                                 counts with zero pairs */
  const void* grp_rec;        /* [n_groups] 16-byte records {int32 n0 | n1 << 16, dim;
                                 double xi_{deg-1}} over the padded node ranges */
  const int32_t* grp_j;       /* [n_groups] dim of each group */
  const int32_t* hs_dim;      /* [n_hs_dims] */
  const int32_t* hs_kmax;     /* [n_hs_dims] max deg of the dim among internal factors */
  const int32_t* hs_row;      /* [n_hs_dims] first h-table row of the dim */
  const int32_t* int_rec;     /* [n_internal] parent row | (h-table row << 16) */
  /* basis matrix and adjoint (GEMM-shaped paths, any d_out; optional) */
  int32_t basis_n;            /* rows of H^T: all nodes of Lambda including the root */
  int32_t n_tab_dims;         /* dims with a factor in some node */
  int32_t n_tab_rows;         /* rows of the per-tile factor table: sum over dims of max deg */
  int32_t n_vjp_ent;          /* adjoint entries: one per (node nu, factor (j, k) of nu) */
  const void* basis_desc;     /* [basis_n] 16 bytes: 8 x u16 factor-table rows (0xffff: none) */
  const int32_t* tab_dim;     /* [n_tab_dims] */
  const int32_t* tab_row;     /* [n_tab_dims] first table row of the dim (row of k is row+k-1) */
  const int32_t* tab_kmax;    /* [n_tab_dims] */
  const int32_t* vjp_slot_off;/* [n_tab_rows + 1] CSR over factor rows of the adjoint entries */
  const void* vjp_ent;        /* 16 bytes per entry: int32 node row of G^T, 6 x u16 table rows
                                 of the node's other factors (0xffff: none) */
  /* dense Jacobian (optional) */
  const int32_t* vjp_dim_off; /* [d_in + 1] adjoint entries of each input dim (entries are stored
                                 by dim; empty range: the dim has no factor) */
  const int32_t* sing_off;    /* [d_in + 1] CSR over dims of sing_x */
  const double* sing_x;       /* coordinates of dim j at which dy/dx_j is NaN: the nodes of every
                                 (dim j, degree) factor of the barycentric blocks */
  int32_t n_jac_units;        /* work units of the register pass (d_out <= 8) */
  int32_t n_jac_partials;     /* units whose dim is split over several units */
  int32_t n_jac_red;          /* dims split over several units */
  int32_t _pad4;
  const int32_t* jac_unit;    /* [4 * n_jac_units] {table dim index, first entry, end entry,
                                 partial slot or -1 (the unit covers its dim)}; units of one dim
                                 are consecutive */
  const int32_t* jac_red;     /* [3 * n_jac_red] {dim, first partial slot, slot count} */
  const int32_t* vjp_ent_row; /* [n_vjp_ent] factor-table row of every adjoint entry */
  const int32_t* row_tdim;    /* [n_tab_rows] table dim index of every factor-table row */
  /* matrix-core mixed kernel (d_out = 1; optional: NULL mt_a selects another kernel).  The
   * internal nodes and the root are numbered by increasing cost sum_j k_j nu_j (rows of W in that
   * order); each group (j, k) keeps a dense vector over the rows 0 .. (its last parent) of
   * c_k * surplus (FP32, zero where a row is not a parent). */
  int32_t rw_n;               /* rows of W (internal nodes + root) */
  int32_t _pad6;
  const int32_t* rint_rec;    /* [n_internal] W row | parent W row << 16, in depth order (the
                                 depth levels of level_off) */
  const int32_t* rint_h;      /* [n_internal] h-table row of the node's last factor */
  /* Groups in tiles of 16; a tile's dense surplus vectors form one operand of
   * v_mfma_f32_16x16x4 against 4 rows of W for 16 points, one 64-float step per 4 rows (lane l:
   * group l % 16, row l / 16 of the step). */
  int32_t mt_n;               /* tiles */
  int32_t mt_steps;           /* A steps */
  int32_t mi_n;               /* work items: step ranges of one tile */
  int32_t n_xi_dims;          /* dims 0 .. n_xi_dims - 1 hold every group of degree >= 2: their
                                 rows of dim_xi are staged on chip */
  const float* mt_a;          /* [64 * mt_steps] */
  const void* mt_grp;         /* [16 * mt_n] 16-byte group descriptors {int32 dim, int32 degree;
                                 double xi_0 of the dim} (dummies: dim 0, degree 1, zero A) */
  const int32_t* mi_rec;      /* [4 * mi_n] {first A step, first step within the tile (W rows
                                 4 s ..), step count, tile} */
  const int32_t* mwave_item;  /* [5] work-item ranges of the 4 wave groups of a point tile (a
                                 group of 4 waves takes 16 of the 64 points each) */
  /* mixed GEMM path (sbi_newton_eval_mixed_gemm; optional: gemm_kq = gemm_head = 0 disables it).
   * The basis rows in GEMM order: rows 0 .. gemm_head - 1 (the heaviest) in FP64, the rest FP32. */
  int32_t gemm_head;          /* FP64 head rows */
  int32_t gemm_kq;            /* FP32 rows basis_n - gemm_head rounded up to a multiple of 16 */
  const int32_t* head_fac;    /* [16 * gemm_head] factors of the head rows (sbi_newton_rows_f64) */
  const double* gemm_s64;     /* [gemm_head, ld_s64] FP64 surpluses of the head rows */
  int64_t ld_s64;
  const float* gemm_sq;       /* FP32 surpluses of the other rows, k-interleaved as in
                                 sbi_contract_split_k4 ([gemm_kq / 4, d_out, 4], zero rows pad) */
  /* FP64 matrix-core kernel of sbi_newton_eval_f64 (optional: NULL selects the streaming one):
   * the A steps of mt_a in FP64 ([64 * mt_steps], read from global memory) */
  const double* mt_a64;
} sbi_newton_desc;

typedef struct sbi_newton_plan sbi_newton_plan;

int sbi_newton_plan_create(const sbi_newton_desc* desc, sbi_newton_plan** out);
int sbi_newton_plan_destroy(sbi_newton_plan* plan);
/* y[P] (FP64) at x[P, ldx]: FP32 surpluses, products and per-group sums; FP64 node differences
 * x - xi, Newton factors and accumulation across groups.  SBI_ERR_CAPACITY when the tile's table
 * of internal-node products exceeds LDS (n_internal + 1 > 640; FP64: > 320): the plan's basis,
 * Jacobian and adjoint entry points still work, and y = H S through sbi_newton_basis_rows and
 * sbi_contract gives the same polynomial. */
int sbi_newton_eval_mixed(const sbi_newton_plan* plan, const double* x, int64_t P, int64_t ldx,
                          double* y, void* stream);
/* Which kernel sbi_newton_eval_mixed runs for this plan: 2 the matrix-core kernel (the groups'
 * sums on the FP32 matrix cores), 1 the LDS-staged kernel (node tables fit in LDS), 0 the
 * streaming kernel that reads them from global memory. */
int sbi_newton_mixed_staged(const sbi_newton_plan* plan);
/* interpolation.py:264-304 __call__ for any d_out on the mixed path, in one call: the FP64 head
 * rows of the basis (sbi_newton_rows_f64), the other rows in FP32 k-interleaved
 * (sbi_newton_basis_rows dtype 3), then sbi_contract_split_k4_ws against the plan's gemm_s64 /
 * gemm_sq: y[P, ldy] (FP64), the same bits as those three calls.  workspace: device memory of
 * sbi_newton_eval_mixed_gemm_workspace(plan, P) bytes, 256-byte aligned, not shared with a
 * concurrent call (basis matrices and, for small batches, the split contraction's partial tiles).
 * P <= 65535 * 64 per call. */
int64_t sbi_newton_eval_mixed_gemm_workspace(const sbi_newton_plan* plan, int64_t P);
int sbi_newton_eval_mixed_gemm(const sbi_newton_plan* plan, const double* x, int64_t P,
                               int64_t ldx, double* y, int64_t ldy, void* workspace,
                               int64_t ws_bytes, void* stream);
/* Same in FP64 throughout: the matrix-core kernel in FP64 (W and the h table in FP64, the A
 * steps of mt_a64 from global memory, v_mfma_f64_16x16x4) when the plan has mt_a64 and its FP64
 * tables fit LDS, else the streaming kernel; sbi_newton_f64_matrix_cores(plan) says which (1 / 0).
 * The two differ by FP64 rounding only. */
int sbi_newton_f64_matrix_cores(const sbi_newton_plan* plan);
int sbi_newton_eval_f64(const sbi_newton_plan* plan, const double* x, int64_t P, int64_t ldx,
                        double* y, void* stream);

/* Basis matrix H^T[basis_n, ldh] of the Newton form at x[P, ldx] (dtype 0: FP64, 1: FP32,
 * 2: BF16 rounded from FP32 values; 3 and 4: the k-interleaved layouts of sbi_newton_basis_rows);
 * y = H S is then one dense GEMM (large d_out). */
int sbi_newton_basis(const sbi_newton_plan* plan, const double* x, int64_t P, int64_t ldx, void* HT,
                     int64_t ldh, int32_t dtype, void* stream);
/* Rows [row0, row1) of the same basis matrix into HT rows 0 .. row1 - row0 - 1.  dtype 3: FP32 in
 * the k-interleaved layout of sbi_contract_split_k4 (HT[((r / 4) * ldh + p) * 4 + r % 4], rows
 * padded with zeros to a multiple of 16, HT 16-byte aligned); dtype 4: BF16 in the layout of
 * sbi_contract_bf16_k8 (HT[((r / 8) * ldh + p) * 8 + r % 8], rows padded to a multiple of 32). */
int sbi_newton_basis_rows(const sbi_newton_plan* plan, const double* x, int64_t P, int64_t ldx,
                          int32_t row0, int32_t row1, void* HT, int64_t ldh, int32_t dtype,
                          void* stream);
/* n_rows rows of the basis matrix computed directly in FP64 from the coordinates (no factor
 * table), row r = prod over q < 8 of h_{fac[16r+2q+1]}(x_{fac[16r+2q]}) (fac: device int32
 * [16 * n_rows], dim < 0 = no factor; (dim, deg) outside the plan's tables contribute 1).  For the
 * few heaviest rows, which the mixed contraction keeps in FP64 (sbi_contract_split). */
int sbi_newton_rows_f64(const sbi_newton_plan* plan, const double* x, int64_t P, int64_t ldx,
                        const int32_t* fac, int32_t n_rows, double* HT, int64_t ldh, void* stream);
/* Adjoint of the Newton form (the VJP of __call__, gradient path of interpolation.py:306-345):
 * gx[P, d_in] = sum_nu G[p, nu] d H[p, nu] / dx  with G^T[basis_n, ldg] = S V^T given
 * (dtype 0: FP64, 1: FP32), rows in the adjoint's node order.  Work units of at most 128 adjoint
 * entries of one dim; a dim split over several units adds their partial sums in unit order, so
 * the result is deterministic.  workspace: sbi_newton_vjp_workspace(plan, P) bytes (partial sums,
 * and the factor table when it takes more than 40 KB per 64 points). */
int64_t sbi_newton_vjp_workspace(const sbi_newton_plan* plan, int64_t P);
int sbi_newton_vjp(const sbi_newton_plan* plan, const double* x, int64_t P, int64_t ldx,
                   const void* GT, int64_t ldg, int32_t dtype, double* gx, void* workspace,
                   int64_t ws_bytes, void* stream);
/* Bytes of device workspace sbi_newton_jacobian_f64 needs for P points and n outputs: the factor
 * table when it takes more than 40 KB per 64 points; for n <= 8 the partial sums of split dims;
 * for n > 8 the derivative basis D^T (adjoint entries x P). */
int64_t sbi_newton_jacobian_workspace(const sbi_newton_plan* plan, int64_t P, int64_t n);
/* Dense Jacobian of the Newton form, J[P, n, d_in] (FP64, contiguous), the reference's gradient
 * (interpolation.py:306-345, barycentric.py:126-229) of the same polynomial:
 *   J[p, o, j] = sum over nodes nu with a factor (j, k):  S[nu, o] h'_k(x_j) prod_{other (j',k')} h_k'(x_j')
 * S: surpluses [basis_n, lds], rows in the node order of the adjoint (the rows of G^T of
 * sbi_newton_vjp).  n <= 8: one pass over work units of at most 128 entries, sums in registers
 * (a dim split over several units adds their partial sums in unit order);
 * n > 8: D^T is written to the workspace and contracted on the FP64 matrix cores with the
 * entries' surplus rows, one group of rows per input dim, eight dims per workgroup.
 * J[p, :, j] = NaN where x_j is one of sing_x of dim j (the reference's result at a node
 * coordinate).  Deterministic. */
int sbi_newton_jacobian_f64(const sbi_newton_plan* plan, const double* x, int64_t P, int64_t ldx,
                            const double* S, int64_t lds, int64_t n, double* J, void* workspace,
                            int64_t ws_bytes, void* stream);
/* Barycentric entry weights W^T[k_raw, ldw] (FP64) at x[P, ldx]: y = offset + W F, the
 * reference's per-block contraction (barycentric.py:69-123) as one dense GEMM (large d_out). */
int sbi_bary_basis_f64(const sbi_plan* plan, const double* x, int64_t P, int64_t ldx, double* WT,
                       int64_t ldw, void* stream);
/* Same weights (FP64 node differences and products) stored in FP32 for the FP32 contraction. */
int sbi_bary_basis_f32(const sbi_plan* plan, const double* x, int64_t P, int64_t ldx, float* WT,
                       int64_t ldw, void* stream);

/* Coefficient-by-basis contraction on the matrix cores for many outputs:
 *   Y[p, j] = offset[j] + sum_k A[k, p] B[k, j]   (p < P, j < n, k < K; offset may be NULL)
 * with A the basis matrix written by sbi_newton_basis / sbi_bary_basis_* (K x P, row stride lda)
 * and B the coefficients (K x n, row stride ldb), both of dtype 0 (FP64, FP64 MFMA) or 1 (FP32,
 * FP32 MFMA with the partial sums added into FP64 every 256 terms); Y is FP64 (row stride ldy).
 * Replaces the dense GEMM that follows the per-point weights in the reference's evaluation
 * (barycentric.py:69-123 contracts F with the weight tensors block by block).  Each row of Y
 * depends only on its own column of A, summed in a fixed order, so a point's result is the same
 * bit for bit whatever else is in the batch; launches are independent and may run concurrently on
 * several streams. */
int sbi_contract(const void* A, int64_t lda, const void* B, int64_t ldb, const double* offset,
                 double* Y, int64_t ldy, int64_t P, int64_t n, int64_t K, int32_t dtype,
                 void* stream);

/* Mixed contraction: Y = offset + A64^T B64 (K64 rows, FP64 MFMA) + A32^T B32 (K32 rows, FP32
 * MFMA, partial sums added into FP64 every 256 terms), with the same layouts as sbi_contract.
 * The few rows that carry most of the result (the constant and low-degree terms) go in the FP64
 * head so that their rounding does not dominate the FP32 remainder.  Same determinism and
 * concurrency guarantees as sbi_contract. */
int sbi_contract_split(const double* A64, int64_t lda64, const double* B64, int64_t ldb64,
                       int64_t K64, const float* A32, int64_t lda32, const float* B32,
                       int64_t ldb32, int64_t K32, const double* offset, double* Y, int64_t ldy,
                       int64_t P, int64_t n, void* stream);

/* sbi_contract_split with the FP32 operands in the k-interleaved layout: element (k, p) of A32 at
 * A32[((k / 4) * lda32 + p) * 4 + k % 4], element (k, c) of B32 at B32[((k / 4) * ldb32 + c) * 4
 * + k % 4], K32 a multiple of 16 (zero rows pad it), both 16-byte aligned.  A lane then loads the
 * four k of its MFMA steps as one 16-byte vector.  Same results bit for bit as sbi_contract_split
 * on the same values in the plain layout.  No counterpart in the reference (a layout of the
 * contraction in interpolation.py:264-304). */
int sbi_contract_split_k4(const double* A64, int64_t lda64, const double* B64, int64_t ldb64,
                          int64_t K64, const float* A32, int64_t lda32, const float* B32,
                          int64_t ldb32, int64_t K32, const double* offset, double* Y,
                          int64_t ldy, int64_t P, int64_t n, void* stream);

/* sbi_contract_split_k4 for small batches (the same contraction of interpolation.py:264-304, the
 * same bits): when its P x n tiles are too few to fill the GPU, the k range is split over extra
 * workgroups that write each 256-k partial tile (and the FP64 head) to `workspace`, and a second
 * kernel adds them in sbi_contract_split_k4's order.  sbi_contract_split_k4_workspace gives the
 * bytes it needs for a shape (0: no split, the workspace is not used; -1: invalid); with less
 * (or NULL) the call runs sbi_contract_split_k4.  workspace: device memory, 256-byte aligned,
 * not shared with a concurrent call. */
int64_t sbi_contract_split_k4_workspace(int64_t P, int64_t n, int64_t K64, int64_t K32);
int sbi_contract_split_k4_ws(const double* A64, int64_t lda64, const double* B64, int64_t ldb64,
                             int64_t K64, const float* A32, int64_t lda32, const float* B32,
                             int64_t ldb32, int64_t K32, const double* offset, double* Y,
                             int64_t ldy, int64_t P, int64_t n, void* workspace, int64_t ws_bytes,
                             void* stream);

/* BF16 form of sbi_contract_split_k4 (gfx950 BF16 matrix cores): the K16 remainder rows of A and
 * B in BF16, k-interleaved by 8 (element (k, p) of A16 at A16[((k / 8) * lda16 + p) * 8 + k % 8],
 * element (k, c) of B16 at B16[((k / 8) * ldb16 + c) * 8 + k % 8], K16 a multiple of 32 with zero
 * rows padding it, both 16-byte aligned); the K64 head rows in FP64 as in sbi_contract_split.  The
 * BF16 products are exact in FP32; FP32 partial sums are added into FP64 every 1024 terms.  Same
 * determinism and concurrency guarantees as sbi_contract (no inter-workgroup waits).  Replaces
 * the BF16 form of the contraction of interpolation.py:264-304 (barycentric.py:123). */
int sbi_contract_bf16_k8(const double* A64, int64_t lda64, const double* B64, int64_t ldb64,
                         int64_t K64, const void* A16, int64_t lda16, const void* B16,
                         int64_t ldb16, int64_t K16, const double* offset, double* Y, int64_t ldy,
                         int64_t P, int64_t n, void* stream);

/* ----------------------------------------------------------------------------------------------
 * Set-up on the device (all pointers device memory, row-major, stream = hipStream_t).
 * ------------------------------------------------------------------------------------------- */

/* interpolation.py:199-230 (set_f fills each group's F from the f values at the grid nodes):
 * F[r, :] = scale[r] * V[node[r], :] for r < n_rows, with V[n_nodes, d_out] the f values at the
 * distinct grid nodes, node[r] the grid node of F's row r and scale[r] the Smolyak coefficient of
 * its block (this layout folds zeta into F); rows whose node lies outside [0, n_nodes) are zero
 * (padding).  One product per element, so F equals the host packing's V[node] * scale bit for bit. */
int sbi_pack_rows_f64(const double* V, int64_t n_nodes, int64_t d_out, const int32_t* node,
                      const double* scale, int64_t n_rows, double* F, void* stream);

/* No counterpart in the reference (its evaluation is barycentric; the Newton form is this
 * library's own): grid values -> Newton surpluses in place, s[n_rows, ld] with d_out columns.
 * Steps run in order; step k updates the targets tgt[step_off[k] .. step_off[k+1]) and target i
 * takes the sources src[src_off[i] .. src_off[i+1]) in order:
 *   s[tgt_i] = (...((s[tgt_i] - coef_q0 * s[src_q0]) - coef_q1 * s[src_q1]) ...),
 * each product and difference rounded on its own (the host restatement's order and rounding, so
 * the same bits).  A step's targets must not be sources of the same step.  Rows outside
 * [0, n_rows) are skipped. */
int sbi_hierarchize_f64(double* s, int64_t n_rows, int64_t d_out, int64_t ld,
                        const int32_t* step_off, int32_t n_steps, const int32_t* tgt,
                        const int32_t* src_off, const int32_t* src, const double* coef,
                        void* stream);

/* ----------------------------------------------------------------------------------------------
 * Serving: a continuous-batching scheduler over many resident interpolants (mixed path of
 * __call__, interpolation.py:264-304, for many interpolants and requests at once; no counterpart
 * in the reference, which evaluates one interpolant per call).  Unlike the plans, the server
 * owns device resources: n_streams HIP streams (default priority), one completion event per
 * stream, pooled events for requests, and per-stream staging buffers sized at sbi_server_add
 * (cap_points x (d_in + d_out) doubles plus the mixed GEMM workspace, the largest over the
 * interpolants).  One host thread drives a server, on the device that was current at
 * sbi_server_create (plans and buffers of that device); after a failed sbi_server_add, _submit
 * or _step every later call but sbi_server_destroy is refused.
 * ------------------------------------------------------------------------------------------- */
typedef struct sbi_server sbi_server;
int sbi_server_create(int32_t n_streams, sbi_server** out);
/* Register an interpolant (its Newton plan must outlive the server): d_out = 1 runs the
 * streaming mixed kernels (sbi_newton_eval_mixed; sbi_newton_eval_mixed_gemm if those return
 * SBI_ERR_CAPACITY), other d_out sbi_newton_eval_mixed_gemm; a batch holds at most cap_points
 * points.  Only before the first submit.  *id numbers the interpolants from 0. */
int sbi_server_add(sbi_server* server, const sbi_newton_plan* plan, int64_t cap_points,
                   int32_t* id);
/* Queue a request: P points x[P, ldx] (device) of interpolant `id`, results into y[P, ldy]
 * (device, FP64).  An event is recorded on `stream` (the stream that produced x; NULL = default
 * stream) and every batch that reads x waits for it.  x and y must stay allocated until the
 * request is reported done.  *rid numbers the requests from 0 in submission order. */
int sbi_server_submit(sbi_server* server, int32_t id, const double* x, int64_t P, int64_t ldx,
                      double* y, int64_t ldy, void* stream, int64_t* rid);
/* One scheduler iteration: observe finished batches (a request is done when every batch holding
 * a piece of it has finished: its rid goes to done[], at most max_done (>= 64) per call, *n_done
 * of them), then launch batches on the free streams.  *busy = 1 while requests are queued or
 * batches are in flight. */
int sbi_server_step(sbi_server* server, int64_t* done, int32_t max_done, int32_t* n_done,
                    int32_t* busy);
/* Wait for the batches in flight, release the server's streams, events and buffers. */
int sbi_server_destroy(sbi_server* server);

/* Descriptor accessors of the plan objects (shared by the library's translation units). */
int sbi_bary_plan_desc(const sbi_plan* plan, const sbi_bary_desc** out);
int sbi_newton_plan_desc(const sbi_newton_plan* plan, const sbi_newton_desc** out);

#ifdef __cplusplus
}
#endif
#endif /* SMOLYAX_AMD_H */
