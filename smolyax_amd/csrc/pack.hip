// Set-up kernels: packing of function values into the device layout and hierarchisation of grid
// values into Newton surpluses, both on the device so that set_f moves the node values (one row
// per distinct grid node) over PCIe instead of the packed F (one row per block entry), and the
// surpluses are not computed on the host.
//
// Both kernels reproduce the host formulas operation for operation (each product and difference
// rounded on its own: this file is compiled with -ffp-contract=off), so their results equal the
// host packing's bit for bit.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <type_traits>

#include "../../include/smolyax_amd.h"
#include "common.h"

namespace {

constexpr int kPackThreads = 256;
constexpr int kPackMaxBlocks = 8192;
constexpr int kHierThreads = 1024;

// F[r, :] = scale[r] * V[node[r], :]; a row whose node lies outside [0, n_nodes) is zero.
// Rows are spread over the threads of a workgroup by 2^tpr_log2 threads per row, each thread
// covering the row's VEC-wide vectors at a stride of that many threads.
template <int VEC>
__global__ __launch_bounds__(kPackThreads) void pack_rows_kernel(
    const double* __restrict__ V, int64_t n_nodes, int64_t d_out, const int32_t* __restrict__ node,
    const double* __restrict__ scale, int64_t n_rows, double* __restrict__ F, int tpr_log2) {
  using vec_t = typename std::conditional<VEC == 2, double2, double>::type;
  const int tpr = 1 << tpr_log2;
  const int rpb = kPackThreads >> tpr_log2;
  const int lr = (int)threadIdx.x >> tpr_log2;
  const int lc = (int)threadIdx.x & (tpr - 1);
  const int64_t nv = d_out / VEC;
  for (int64_t r = (int64_t)blockIdx.x * rpb + lr; r < n_rows; r += (int64_t)gridDim.x * rpb) {
    const int32_t nd = node[r];
    const double sc = scale[r];
    const bool ok = nd >= 0 && (int64_t)nd < n_nodes;
    const vec_t* src = reinterpret_cast<const vec_t*>(V + (ok ? (int64_t)nd : 0) * d_out);
    vec_t* dst = reinterpret_cast<vec_t*>(F + r * d_out);
    for (int64_t c = lc; c < nv; c += tpr) {
      vec_t o;
      if constexpr (VEC == 2) {
        if (ok) {
          const double2 v = src[c];
          o = make_double2(sc * v.x, sc * v.y);
        } else {
          o = make_double2(0.0, 0.0);
        }
      } else {
        o = ok ? sc * src[c] : 0.0;
      }
      dst[c] = o;
    }
  }
}

// In place: for each step in order, for each target row i of the step,
//   s[tgt_i, c] = (...((s[tgt_i, c] - coef_q0 * s[src_q0, c]) - coef_q1 * s[src_q1, c]) ...)
// over the target's sources q in order.  One workgroup per chunk of 2^cw_log2 columns runs all
// steps, with a workgroup barrier between steps (the targets of a step are not sources of it, and
// the columns of different workgroups never meet).  Rows outside [0, n_rows) are skipped.
__global__ __launch_bounds__(kHierThreads) void hierarchize_kernel(
    double* s, int64_t n_rows, int64_t d_out, int64_t ld, const int32_t* __restrict__ step_off,
    int32_t n_steps, const int32_t* __restrict__ tgt, const int32_t* __restrict__ src_off,
    const int32_t* __restrict__ src, const double* __restrict__ coef, int cw_log2) {
  const int cw = 1 << cw_log2;
  const int tpp = kHierThreads >> cw_log2;  // targets per pass
  const int64_t c = (int64_t)blockIdx.x * cw + ((int)threadIdx.x & (cw - 1));
  const int ti = (int)threadIdx.x >> cw_log2;
  const bool col_ok = c < d_out;
  for (int32_t st = 0; st < n_steps; ++st) {
    const int32_t t0 = step_off[st], t1 = step_off[st + 1];
    if (col_ok) {
      for (int32_t i = t0 + ti; i < t1; i += tpp) {
        const int32_t r = tgt[i];
        if (r < 0 || (int64_t)r >= n_rows) continue;
        double v = s[(int64_t)r * ld + c];
        const int32_t q1 = src_off[i + 1];
        for (int32_t q = src_off[i]; q < q1; ++q) {
          const int32_t sr = src[q];
          if (sr < 0 || (int64_t)sr >= n_rows) continue;
          const double p = coef[q] * s[(int64_t)sr * ld + c];
          v = v - p;
        }
        s[(int64_t)r * ld + c] = v;
      }
    }
    __syncthreads();
  }
}

int ilog2_ceil(int64_t v) {
  int l = 0;
  while ((int64_t(1) << l) < v) ++l;
  return l;
}

}  // namespace

extern "C" int sbi_pack_rows_f64(const double* V, int64_t n_nodes, int64_t d_out,
                                 const int32_t* node, const double* scale, int64_t n_rows,
                                 double* F, void* stream) {
  if (d_out < 1 || n_rows < 0 || n_nodes < 0 || (n_rows > 0 && (!node || !scale || !F)) ||
      (n_nodes > 0 && !V) || d_out > (int64_t(1) << 40) / (n_rows > 0 ? n_rows : 1)) {
    sbi::set_error("sbi_pack_rows_f64: invalid argument");
    return SBI_ERR_INVALID;
  }
  if (n_rows == 0) return SBI_OK;
  hipStream_t st = (hipStream_t)stream;
  const bool vec2 = d_out % 2 == 0 && (uintptr_t)V % 16 == 0 && (uintptr_t)F % 16 == 0;
  const int64_t nv = vec2 ? d_out / 2 : d_out;
  const int tpr_log2 = ilog2_ceil(nv < kPackThreads ? nv : kPackThreads);
  const int64_t rpb = kPackThreads >> tpr_log2;
  const int64_t blocks = (n_rows + rpb - 1) / rpb;
  const dim3 grid((unsigned)(blocks < kPackMaxBlocks ? blocks : kPackMaxBlocks));
  if (vec2)
    hipLaunchKernelGGL(pack_rows_kernel<2>, grid, dim3(kPackThreads), 0, st, V, n_nodes, d_out,
                       node, scale, n_rows, F, tpr_log2);
  else
    hipLaunchKernelGGL(pack_rows_kernel<1>, grid, dim3(kPackThreads), 0, st, V, n_nodes, d_out,
                       node, scale, n_rows, F, tpr_log2);
  SBI_HIP_CHECK(hipGetLastError());
  return SBI_OK;
}

extern "C" int sbi_hierarchize_f64(double* s, int64_t n_rows, int64_t d_out, int64_t ld,
                                   const int32_t* step_off, int32_t n_steps, const int32_t* tgt,
                                   const int32_t* src_off, const int32_t* src, const double* coef,
                                   void* stream) {
  if (!s || n_rows < 0 || d_out < 1 || ld < d_out || n_steps < 0 ||
      (n_steps > 0 && (!step_off || !tgt || !src_off || !src || !coef)) ||
      n_rows > (int64_t(1) << 31) - 1) {
    sbi::set_error("sbi_hierarchize_f64: invalid argument");
    return SBI_ERR_INVALID;
  }
  if (n_steps == 0 || n_rows == 0) return SBI_OK;
  // columns per workgroup: up to 64, fewer while that leaves fewer than 256 workgroups (one per
  // CU), at least 8 (a 64-byte run of each row per target)
  int cw_log2 = ilog2_ceil(d_out < 64 ? d_out : 64);
  while (cw_log2 > 3 && (d_out + (int64_t(1) << cw_log2) - 1) >> cw_log2 < 256) --cw_log2;
  const int64_t groups = (d_out + (int64_t(1) << cw_log2) - 1) >> cw_log2;
  if (groups > 0x7fffffff) {
    sbi::set_error("sbi_hierarchize_f64: d_out too large");
    return SBI_ERR_CAPACITY;
  }
  hipLaunchKernelGGL(hierarchize_kernel, dim3((unsigned)groups), dim3(kHierThreads), 0,
                     (hipStream_t)stream, s, n_rows, d_out, ld, step_off, n_steps, tgt, src_off,
                     src, coef, cw_log2);
  SBI_HIP_CHECK(hipGetLastError());
  return SBI_OK;
}
