// Hierarchical (Newton-surplus) evaluation of a Smolyak interpolant on nested nodes, gfx950.
//
// y(x) = s_root + sum_{nu in Lambda, nu != 0} s_nu prod_{(j,k) in nu} h^{(j)}_k(x_j),
// h^{(j)}_k(x) = c_k prod_{r<k} (x - xi_r).  The same polynomial as the reference's combination
// technique (interpolation.py:264-304) written with |Lambda| terms instead of K_raw.
//
// Every node nu != 0 is a child of its parent (nu without its highest-dim factor).  Nodes that
// have children ("internal", few) get a row of the on-chip table W: W[q] = prod over the node's
// factors, computed per tile in depth order.  The remaining work is streamed per (dim j, deg k)
// group: t = sum_{children c of (j,k)} s_c * W[parent(c)], y += h_k(x_j) * t.
//
// Schedule: one tile = 64 points, one point per lane; 16 waves per workgroup share the tile's W
// table and split the groups (cost-balanced on the host).  Node metadata (surplus, parent row) is
// fetched 64 nodes at a time with one coalesced vector load per array (one node per lane), one
// chunk ahead of use, and broadcast with v_readlane: the dependent-load latency never sits on the
// critical path.  Precision T = float (mixed: FP32 surpluses, table and group sums; the node
// differences x - xi, h and the accumulation of y in FP64) or T = double.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/smolyax_amd.h"
#include "common.h"

#define SBI_HIP_CHECK(expr)                                                              \
  do {                                                                                   \
    hipError_t _e = (expr);                                                              \
    if (_e != hipSuccess) {                                                              \
      sbi::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));                 \
      return SBI_ERR_HIP;                                                                \
    }                                                                                    \
  } while (0)

namespace {

constexpr int kTile = 64;
constexpr int kWaves = 16;
constexpr int kMaxLds = 160 * 1024;

template <typename T>
using cptr = const __attribute__((address_space(4))) T*;
template <typename T>
__device__ __forceinline__ cptr<T> C(const T* p) {
  return (cptr<T>)(p);
}

template <typename T>
struct lds_of;
template <>
struct lds_of<float> {
  typedef __attribute__((address_space(3))) float type;
};
template <>
struct lds_of<double> {
  typedef __attribute__((address_space(3))) double type;
};

__device__ __forceinline__ float bcast(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double bcast(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// h_k(x) = c_k prod_{r<k} (x - xi_r) in FP64
__device__ __forceinline__ double newton_h(double x, cptr<double> xi, cptr<double> c, int k) {
  double p = c[k];
  for (int r = 0; r < k; ++r) p *= (x - xi[r]);
  return p;
}

template <typename T>
__global__ __launch_bounds__(kTile* kWaves) void newton_eval_kernel(sbi_newton_desc d,
                                                                    const double* __restrict__ x,
                                                                    int64_t P, int64_t ldx,
                                                                    double* __restrict__ y) {
  extern __shared__ double lds_raw[];
  typedef typename lds_of<T>::type LT;
  LT* W = (LT*)lds_raw;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t p_raw = (int64_t)blockIdx.x * kTile + lane;
  const bool valid = p_raw < P;
  const int64_t p = valid ? p_raw : (P - 1);
  const double* __restrict__ xp = x + p * ldx;

  cptr<int32_t> int_parent = C(d.int_parent), int_dim = C(d.int_dim), int_deg = C(d.int_deg),
                level_off = C(d.level_off), grp = C(d.grp), grp_next_dim = C(d.grp_next_dim),
                wave_grp = C(d.wave_grp);
  cptr<double> dim_xi = C(d.dim_xi), dim_c = C(d.dim_c);
  const T* __restrict__ node_s =
      (const T*)(sizeof(T) == 4 ? (const void*)d.node_s32 : (const void*)d.node_s64);
  const int32_t* __restrict__ node_q = d.node_q;

  // first streamed coordinate of this wave, requested before the table work
  const int g_begin = wave_grp[wave];
  const int g_end = wave_grp[wave + 1];
  double x_next = 0.0;
  if (g_begin < g_end) x_next = xp[grp[4 * g_begin + 2]];

  // 1. table of internal-node products, level by level (row 0 = root = 1)
  if (wave == 0) W[lane] = (T)1;
  __syncthreads();
  for (int lv = 0; lv < d.n_levels; ++lv) {
    const int a = level_off[lv], b = level_off[lv + 1];
    for (int q = a + wave; q < b; q += kWaves) {
      const int j = int_dim[q];
      const double h = newton_h(xp[j], dim_xi + (int64_t)j * d.kmax1, dim_c + (int64_t)j * d.kmax1,
                                int_deg[q]);
      const double par = (double)W[int_parent[q] * kTile + lane];
      W[(q + 1) * kTile + lane] = (T)(par * h);
    }
    __syncthreads();
  }

  // 2. streamed groups
  double acc = 0.0;
  if (g_begin < g_end) {
    const int n_first = grp[4 * g_begin];
    int chunk = n_first >> 6;
    // current and next metadata chunks (one node per lane)
    const int nn = d.n_nodes;
    int i0 = chunk * 64 + lane;
    T s_cur = (i0 < nn) ? node_s[i0] : (T)0;
    int q_cur = (i0 < nn) ? node_q[i0] : 0;
    int i1 = i0 + 64;
    T s_nxt = (i1 < nn) ? node_s[i1] : (T)0;
    int q_nxt = (i1 < nn) ? node_q[i1] : 0;

    double xj = 0.0, pr = 1.0;
    int jcur = -1, kdone = 0;  // pr = prod_{r < kdone} (x_j - xi_r)
    for (int g = g_begin; g < g_end; ++g) {
      const int n0 = grp[4 * g], n1 = grp[4 * g + 1], j = grp[4 * g + 2], k = grp[4 * g + 3];
      cptr<double> xi = dim_xi + (int64_t)j * d.kmax1;
      if (j != jcur) {
        jcur = j;
        xj = x_next;
        const int jn = grp_next_dim[g];
        if (jn >= 0) x_next = xp[jn];
        pr = 1.0;
        kdone = 0;
      }
      // groups of one dim come in increasing k; a wave may start inside a dim's run of groups
      for (; kdone < k; ++kdone) pr *= (xj - xi[kdone]);
      const double h = pr * dim_c[(int64_t)j * d.kmax1 + k];
      T t = (T)0;
      int i = n0;
      while (i < n1) {
        if ((i >> 6) != chunk) {
          // advance one chunk (groups are visited in node order, so this is the next one)
          chunk += 1;
          s_cur = s_nxt;
          q_cur = q_nxt;
          const int in = (chunk + 1) * 64 + lane;
          s_nxt = (in < nn) ? node_s[in] : (T)0;
          q_nxt = (in < nn) ? node_q[in] : 0;
        }
        const int lim = min(n1, (chunk + 1) * 64);
        // batches of 4 nodes: broadcast metadata, gather the parents' table entries, accumulate
        for (; i + 4 <= lim; i += 4) {
          const int l = i & 63;
          const T s0 = bcast(s_cur, l), s1 = bcast(s_cur, l + 1), s2 = bcast(s_cur, l + 2),
                  s3 = bcast(s_cur, l + 3);
          const int q0 = __builtin_amdgcn_readlane(q_cur, l);
          const int q1 = __builtin_amdgcn_readlane(q_cur, l + 1);
          const int q2 = __builtin_amdgcn_readlane(q_cur, l + 2);
          const int q3 = __builtin_amdgcn_readlane(q_cur, l + 3);
          const T w0 = W[q0 * kTile + lane], w1 = W[q1 * kTile + lane];
          const T w2 = W[q2 * kTile + lane], w3 = W[q3 * kTile + lane];
          t = fma(s0, w0, t);
          t = fma(s1, w1, t);
          t = fma(s2, w2, t);
          t = fma(s3, w3, t);
        }
        for (; i < lim; ++i) {
          const int l = i & 63;
          const T s0 = bcast(s_cur, l);
          const int q0 = __builtin_amdgcn_readlane(q_cur, l);
          t = fma(s0, W[q0 * kTile + lane], t);
        }
      }
      acc = fma(h, (double)t, acc);
    }
  }

  // 3. deterministic cross-wave reduction in FP64
  __syncthreads();
  typedef __attribute__((address_space(3))) double lds_f64;
  lds_f64* R = (lds_f64*)lds_raw;
  R[wave * kTile + lane] = acc;
  __syncthreads();
  if (wave == 0 && valid) {
    double s = d.s_root[0];
    for (int w = 0; w < kWaves; ++w) s += R[w * kTile + lane];
    y[p] = s;
  }
}

}  // namespace

struct sbi_newton_plan {
  sbi_newton_desc d;
  size_t lds_f32, lds_f64;
};

extern "C" int sbi_newton_plan_create(const sbi_newton_desc* desc, sbi_newton_plan** out) {
  if (!desc || !out) {
    sbi::set_error("sbi_newton_plan_create: null argument");
    return SBI_ERR_INVALID;
  }
  const sbi_newton_desc& d = *desc;
  if (d.d_in <= 0 || d.d_out != 1 || d.n_nodes < 0 || d.n_internal < 0 || d.n_groups < 0 ||
      !d.s_root || !d.wave_grp || !d.grp || !d.node_q || !d.node_s32 || !d.node_s64 ||
      d.kmax1 <= 0) {
    sbi::set_error("sbi_newton_plan_create: invalid descriptor (this kernel takes d_out == 1)");
    return SBI_ERR_INVALID;
  }
  if (d.n_split != kWaves) {
    sbi::set_error("sbi_newton_plan_create: n_split must be 16");
    return SBI_ERR_INVALID;
  }
  const size_t red = (size_t)kWaves * kTile * sizeof(double);
  const size_t t32 = (size_t)(d.n_internal + 1) * kTile * sizeof(float);
  const size_t t64 = (size_t)(d.n_internal + 1) * kTile * sizeof(double);
  sbi_newton_plan* p = new sbi_newton_plan();
  p->d = d;
  p->lds_f32 = t32 > red ? t32 : red;
  p->lds_f64 = t64 > red ? t64 : red;
  if (p->lds_f32 > (size_t)kMaxLds) {
    delete p;
    sbi::set_error("sbi_newton_plan_create: too many internal nodes for the on-chip table");
    return SBI_ERR_CAPACITY;
  }
  *out = p;
  return SBI_OK;
}

extern "C" int sbi_newton_plan_destroy(sbi_newton_plan* plan) {
  delete plan;
  return SBI_OK;
}

template <typename T>
static int newton_launch(const sbi_newton_plan* plan, const double* x, int64_t P, int64_t ldx,
                         double* y, void* stream) {
  if (!plan || !x || !y || P < 0 || ldx < plan->d.d_in) {
    sbi::set_error("sbi_newton_eval: invalid argument");
    return SBI_ERR_INVALID;
  }
  const size_t lds = sizeof(T) == 4 ? plan->lds_f32 : plan->lds_f64;
  if (lds > (size_t)kMaxLds) {
    sbi::set_error("sbi_newton_eval: FP64 table exceeds LDS");
    return SBI_ERR_CAPACITY;
  }
  if (P == 0) return SBI_OK;
  static bool attr = false;
  if (!attr) {
    (void)hipFuncSetAttribute((const void*)newton_eval_kernel<float>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, kMaxLds);
    (void)hipFuncSetAttribute((const void*)newton_eval_kernel<double>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, kMaxLds);
    attr = true;
  }
  const sbi_newton_desc& d = plan->d;
  dim3 grid((unsigned)((P + kTile - 1) / kTile));
  hipLaunchKernelGGL(newton_eval_kernel<T>, grid, dim3(kTile * kWaves), lds, (hipStream_t)stream, d,
                     x, P, ldx, y);
  SBI_HIP_CHECK(hipGetLastError());
  return SBI_OK;
}

extern "C" int sbi_newton_eval_mixed(const sbi_newton_plan* plan, const double* x, int64_t P,
                                     int64_t ldx, double* y, void* stream) {
  return newton_launch<float>(plan, x, P, ldx, y, stream);
}

extern "C" int sbi_newton_eval_f64(const sbi_newton_plan* plan, const double* x, int64_t P,
                                   int64_t ldx, double* y, void* stream) {
  return newton_launch<double>(plan, x, P, ldx, y, stream);
}
