// Basis matrices and the adjoint for the GEMM-shaped paths (large d_out), gfx950.
//
// Newton form (nested nodes): y[p, :] = sum_nu H[p, nu] S[nu, :] with
//   H[p, nu] = prod_{(j,k) in nu} h^{(j)}_k(x_{p,j})     (H[p, root] = 1).
// newton_basis_kernel writes H^T (nodes x points, points contiguous) so that the d_out
// contraction Y = H S is one dense GEMM (done by the caller's BLAS on MFMA in FP64, FP32 or
// BF16 with FP32 accumulation).
//
// Adjoint (vector-Jacobian product) of the same form: with G = V S^T (a GEMM, V = dL/dy),
//   dL/dx_{p,j} = sum_k h'^{(j)}_k(x_{p,j}) A[p, (j,k)],
//   A[p, (j,k)] = sum_{nu containing (j,k)} G[p, nu] prod_{other factors (j',k') of nu} h_{k'}(x_{p,j'}).
// newton_jac_kernel<1, kG, T> evaluates it per tile from G^T over work units of at most 128
// adjoint entries of one dim (no atomics: each (p, j) is written once, a dim split over several
// units adds their partial sums in unit order).
//
// Barycentric form (the reference's algorithm): bary_basis_kernel writes the per-entry weights
// W^T[e, p] = prod_i Lhat_i(x_{p, s_i})[mu_i] of every block entry e (K_raw rows) so that
// y = offset + W F (zeta folded into F) is a dense FP64 GEMM.
//
// Tile = 64 points, one per lane; the 1D factors h_k(x_j) of the tile live in an LDS table.

#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <type_traits>

#include "../../include/smolyax_amd.h"
#include "common.h"
#include "launch.h"
#include "gemm.h"

namespace {

constexpr int kTile = 64;
constexpr int kBWaves = 4;   // waves per workgroup of the Jacobian / adjoint passes
constexpr int kHWaves = 16;  // waves per workgroup of the basis kernel (one LDS table per tile)
constexpr int kMaxLds = 160 * 1024;

template <typename T>
using cptr = const __attribute__((address_space(4))) T*;
template <typename T>
__device__ __forceinline__ cptr<T> C(const T* p) {
  return (cptr<T>)(p);
}

template <typename T>
struct lds_t;
template <>
struct lds_t<float> {
  typedef __attribute__((address_space(3))) float type;
};
template <>
struct lds_t<double> {
  typedef __attribute__((address_space(3))) double type;
};

// Table of h_k(x_j) (and optionally h'_k(x_j)) for all table dims of the tile.  Row of (j, k):
// tab_row[i] + k - 1 for table dim i.  FP64 arithmetic, stored as T.
template <typename T>
__device__ void build_h_table(const sbi_newton_desc& d, const double* __restrict__ xp,
                              typename lds_t<T>::type* tab, typename lds_t<T>::type* dtab,
                              int wave, int lane, int nwaves) {
  cptr<int32_t> tdim = C(d.tab_dim), trow = C(d.tab_row), tk = C(d.tab_kmax);
  cptr<double> dim_xi = C(d.dim_xi), dim_c = C(d.dim_c);
  for (int i = wave; i < d.n_tab_dims; i += nwaves) {
    const int j = tdim[i], row = trow[i], km = tk[i];
    const double xj = xp[j];
    cptr<double> xi = dim_xi + (int64_t)j * d.kmax1;
    cptr<double> c = dim_c + (int64_t)j * d.kmax1;
    // pr = prod_{r<k} (x - xi_r);  dpr = d/dx pr
    double pr = 1.0, dpr = 0.0;
    for (int k = 1; k <= km; ++k) {
      const double dx = xj - xi[k - 1];
      dpr = dpr * dx + pr;
      pr *= dx;
      tab[(row + k - 1) * kTile + lane] = (T)(pr * c[k]);
      if (dtab) dtab[(row + k - 1) * kTile + lane] = (T)(dpr * c[k]);
    }
  }
}

__device__ __forceinline__ void store_out(double* p, double v) { *p = v; }
__device__ __forceinline__ void store_out(float* p, double v) { *p = (float)v; }
__device__ __forceinline__ void store_out(__hip_bfloat16* p, double v) {
  *p = __float2bfloat16((float)v);
}

// Number of factor rows of each lane's node descriptor (8 x u16 rows packed first, 0xffff after).
__device__ __forceinline__ int desc_factor_count(const uint4& dv) {
  int c = 0;
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    const uint32_t w = f == 0 ? dv.x : f == 1 ? dv.y : f == 2 ? dv.z : dv.w;
    c += (int)((w & 0xffff) != 0xffff) + (int)((w >> 16) != 0xffff);
  }
  return c;
}

// Product of the lane's table values at the nf factor rows of node i of the wave's 64, in the
// stored order (the same products as multiplying row by row from 1).
template <typename T, typename LT>
__device__ __forceinline__ T node_product(const LT* tab, const uint4& dv, int i, int nf,
                                          int lane) {
  T v = (T)1;
  if (nf == 0) return v;
  const uint32_t w0 = (uint32_t)__builtin_amdgcn_readlane((int)dv.x, i);
  v *= tab[(w0 & 0xffff) * kTile + lane];
  if (nf == 1) return v;
  v *= tab[(w0 >> 16) * kTile + lane];
  if (nf == 2) return v;
  const uint32_t w1 = (uint32_t)__builtin_amdgcn_readlane((int)dv.y, i);
  v *= tab[(w1 & 0xffff) * kTile + lane];
  if (nf == 3) return v;
  v *= tab[(w1 >> 16) * kTile + lane];
  if (nf == 4) return v;
  const uint32_t w2 = (uint32_t)__builtin_amdgcn_readlane((int)dv.z, i);
  v *= tab[(w2 & 0xffff) * kTile + lane];
  if (nf == 5) return v;
  v *= tab[(w2 >> 16) * kTile + lane];
  if (nf == 6) return v;
  const uint32_t w3 = (uint32_t)__builtin_amdgcn_readlane((int)dv.w, i);
  v *= tab[(w3 & 0xffff) * kTile + lane];
  if (nf == 7) return v;
  return v * tab[(w3 >> 16) * kTile + lane];
}

// k-interleaved outputs for the contraction (kQ = 4: FP32, kQ = 8: BF16): row r = row - row0 of
// point p at HT[((r / kQ) * ldh + p) * kQ + r % kQ], rows padded with zeros to a multiple of 16
// (FP32) or 32 (BF16, one MFMA k step); each lane stores the kQ rows of a group as one 16-byte
// vector.
template <int kQ>
struct kq_traits {
  static constexpr int pad = 1;
};
template <>
struct kq_traits<4> {
  static constexpr int pad = 16;
};
template <>
struct kq_traits<8> {
  static constexpr int pad = 32;
};

__device__ __forceinline__ uint32_t bf16_bits(float v) {
  const __hip_bfloat16 b = __float2bfloat16(v);
  return (uint32_t)__bfloat16_as_ushort(b);
}

template <int kQ>
__device__ __forceinline__ void store_group(void* HT, int64_t ldh, int r, int64_t p,
                                            const float* q) {
  if constexpr (kQ == 4) {
    typedef float v4f_t __attribute__((ext_vector_type(4)));
    *reinterpret_cast<v4f_t*>((float*)HT + ((int64_t)(r >> 2) * ldh + p) * 4) =
        v4f_t{q[0], q[1], q[2], q[3]};
  } else {
    uint4 v;
    v.x = bf16_bits(q[0]) | (bf16_bits(q[1]) << 16);
    v.y = bf16_bits(q[2]) | (bf16_bits(q[3]) << 16);
    v.z = bf16_bits(q[4]) | (bf16_bits(q[5]) << 16);
    v.w = bf16_bits(q[6]) | (bf16_bits(q[7]) << 16);
    reinterpret_cast<uint4*>(HT)[(int64_t)(r >> 3) * ldh + p] = v;
  }
}

// H^T[nu, p] for a chunk of nodes.  Node descriptor: 8 x u16 table rows (0xffff = none).
// kQ = 0: plain layout HT[r * ldh + p] of type OUT; kQ = 4 / 8: k-interleaved (see above).
// W waves per workgroup share the tile's table: 16 for large batches, 4 (with 256-row chunks)
// when 16-wave workgroups would be too few to fill the GPU; each element is computed the same way.
template <typename T, typename OUT, int kQ = 0, int W = kHWaves>
__global__ __launch_bounds__(kTile* W) void newton_basis_kernel(sbi_newton_desc d,
                                                                      const double* __restrict__ x,
                                                                      int64_t P, int64_t ldx,
                                                                      OUT* __restrict__ HT,
                                                                      int64_t ldh, int chunk,
                                                                      int row0, int row1) {
  extern __shared__ double lds_raw[];
  typedef typename lds_t<T>::type LT;
  LT* tab = (LT*)lds_raw;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t p_raw = (int64_t)blockIdx.x * kTile + lane;
  const bool valid = p_raw < P;
  const int64_t p = valid ? p_raw : (P - 1);
  build_h_table<T>(d, x + p * ldx, tab, nullptr, wave, lane, W);
  __syncthreads();

  const int n0 = row0 + blockIdx.y * chunk;
  constexpr int kPad = kq_traits<kQ>::pad;
  const int rend = kQ ? row0 + ((row1 - row0 + kPad - 1) & ~(kPad - 1)) : row1;
  const int n1 = min(rend, n0 + chunk);
  const uint4* __restrict__ desc = reinterpret_cast<const uint4*>(d.basis_desc);
  // each wave: 64-node groups, descriptors one per lane, broadcast with v_readlane
  for (int base = n0 + wave * 64; base < n1; base += W * 64) {
    const int mine = base + lane;
    const uint4 dv = (mine < min(n1, row1)) ? desc[mine]
                                            : make_uint4(0xffffffffu, 0xffffffffu, 0xffffffffu,
                                                         0xffffffffu);
    const int nfl = desc_factor_count(dv);
    const int cnt = min(64, n1 - base);
    if constexpr (kQ != 0) {
      // cnt is a multiple of kQ (row padding to kPad, 64-row groups)
      for (int i = 0; i < cnt; i += kQ) {
        float q[kQ];
#pragma unroll
        for (int u = 0; u < kQ; ++u) {
          const T v = node_product<T>(tab, dv, i + u,
                                      __builtin_amdgcn_readlane(nfl, i + u), lane);
          q[u] = (base + i + u < row1) ? (float)v : 0.0f;
        }
        if (valid) store_group<kQ>(HT, ldh, base + i - row0, p, q);
      }
      continue;
    }
    for (int i = 0; i < cnt; ++i) {
      const T v = node_product<T>(tab, dv, i, __builtin_amdgcn_readlane(nfl, i), lane);
      if (valid) store_out(HT + (int64_t)(base + i - row0) * ldh + p, (double)v);
    }
  }
}

// Factor-table value of row r (table dim i = row_tdim[r], dim j, degree k = r - tab_row[i] + 1)
// computed from the coordinate, with the same operations as build_h_table.
template <typename T>
__device__ __forceinline__ T factor_value(const sbi_newton_desc& d, const double* __restrict__ xp,
                                          int r) {
  const int i = C(d.row_tdim)[r];
  const int j = C(d.tab_dim)[i], k = r - C(d.tab_row)[i] + 1;
  cptr<double> xi = C(d.dim_xi) + (int64_t)j * d.kmax1;
  const double xj = xp[j];
  double pr = 1.0;
  for (int m = 0; m < k; ++m) pr *= (xj - xi[m]);
  return (T)(pr * C(d.dim_c)[(int64_t)j * d.kmax1 + k]);
}

// H^T like newton_basis_kernel when the tile's factor table does not fit in LDS: each factor
// value is computed from the coordinate where the node needs it (same values, same products).
template <typename T, typename OUT, int kQ = 0>
__global__ __launch_bounds__(kTile* kHWaves) void newton_basis_direct_kernel(
    sbi_newton_desc d, const double* __restrict__ x, int64_t P, int64_t ldx,
    OUT* __restrict__ HT, int64_t ldh, int chunk, int row0, int row1) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t p_raw = (int64_t)blockIdx.x * kTile + lane;
  const bool valid = p_raw < P;
  const int64_t p = valid ? p_raw : (P - 1);
  const double* __restrict__ xp = x + p * ldx;
  const int n0 = row0 + blockIdx.y * chunk;
  constexpr int kPad = kq_traits<kQ>::pad;
  const int rend = kQ ? row0 + ((row1 - row0 + kPad - 1) & ~(kPad - 1)) : row1;
  const int n1 = min(rend, n0 + chunk);
  const uint4* __restrict__ desc = reinterpret_cast<const uint4*>(d.basis_desc);
  auto product = [&](const uint4& dv, int i) {
    T v = (T)1;
    const uint32_t w[4] = {(uint32_t)__builtin_amdgcn_readlane((int)dv.x, i),
                           (uint32_t)__builtin_amdgcn_readlane((int)dv.y, i),
                           (uint32_t)__builtin_amdgcn_readlane((int)dv.z, i),
                           (uint32_t)__builtin_amdgcn_readlane((int)dv.w, i)};
    for (int f = 0; f < 8; ++f) {
      const uint32_t r = (w[f >> 1] >> ((f & 1) * 16)) & 0xffff;
      if (r == 0xffff) break;  // rows are packed first
      v *= factor_value<T>(d, xp, (int)r);
    }
    return v;
  };
  for (int base = n0 + wave * 64; base < n1; base += kHWaves * 64) {
    const int mine = base + lane;
    const uint4 dv = (mine < min(n1, row1)) ? desc[mine]
                                            : make_uint4(0xffffffffu, 0xffffffffu, 0xffffffffu,
                                                         0xffffffffu);
    const int cnt = min(64, n1 - base);
    if constexpr (kQ != 0) {
      for (int i = 0; i < cnt; i += kQ) {
        float q[kQ];
#pragma unroll
        for (int u = 0; u < kQ; ++u) {
          const T v = product(dv, i + u);
          q[u] = (base + i + u < row1) ? (float)v : 0.0f;
        }
        if (valid) store_group<kQ>(HT, ldh, base + i - row0, p, q);
      }
      continue;
    }
    for (int i = 0; i < cnt; ++i) {
      const T v = product(dv, i);
      if (valid) store_out(HT + (int64_t)(base + i - row0) * ldh + p, (double)v);
    }
  }
}

// Rows of H^T computed directly in FP64 from the coordinates (no factor table): each row is a
// product of at most 8 factors h_k(x_j), given as (dim, deg) pairs (dim < 0: none).  Used for the
// few heaviest rows, which the mixed contraction keeps in FP64.  A workgroup covers 64 points
// (one per lane) and kRowsPerBlock rows, one wave per fourth row, so that small batches still
// spread over many workgroups; each element is computed on its own (same bits for any P).
constexpr int kRowsPerBlock = 16;
__global__ __launch_bounds__(256) void newton_rows_f64_kernel(sbi_newton_desc d,
                                                             const double* __restrict__ x,
                                                             int64_t P, int64_t ldx,
                                                             const int32_t* __restrict__ fac,
                                                             int n_rows, double* __restrict__ HT,
                                                             int64_t ldh) {
  const int64_t p = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
  const int wave = threadIdx.x >> 6;
  if (p >= P) return;
  const double* __restrict__ xp = x + p * ldx;
  cptr<int32_t> f = C(fac);
  cptr<double> dim_xi = C(d.dim_xi), dim_c = C(d.dim_c);
  const int r1 = min(n_rows, (int)(blockIdx.y + 1) * kRowsPerBlock);
  for (int r = (int)blockIdx.y * kRowsPerBlock + wave; r < r1; r += 4) {
    double v = 1.0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int j = f[16 * r + 2 * q], k = f[16 * r + 2 * q + 1];
      if (j >= 0 && j < d.d_in && k >= 0 && k < d.kmax1) {  // (dim, deg) in the plan's tables
        const double xj = xp[j];
        cptr<double> xi = dim_xi + (int64_t)j * d.kmax1;
        double h = dim_c[(int64_t)j * d.kmax1 + k];
        for (int m = 0; m < k; ++m) h *= (xj - xi[m]);
        v *= h;
      }
    }
    HT[(int64_t)r * ldh + p] = v;
  }
}

// ------------------------------------------------------------------------------------------
// Dense Jacobian of the Newton form.  Adjoint entry e = (node nu, factor (j, k) of nu) has the
// derivative-basis value D[p, e] = h'_k(x_j) prod_{other factors (j',k') of nu} h_k'(x_j'), and
//   J[p, o, j] = sum over the entries e of dim j of D[p, e] S[nu_e, o].
// The factor table of the tile sits in LDS when it fits, else in a global table written first by
// newton_htab_kernel, tile by tile ([tile][row][64 points]: one tile's table is contiguous, so a
// wave's reads stay within a few pages).
// ------------------------------------------------------------------------------------------
template <bool kG>
struct HTab {
  typename lds_t<double>::type* lds;
  const double* __restrict__ g;
  int64_t ldt;  // rows per tile of the global table
  __device__ __forceinline__ double at(int r, int lane, int64_t p) const {
    if constexpr (kG) {
      return g[((int64_t)blockIdx.x * ldt + r) * kTile + lane];
    } else {
      return lds[r * kTile + lane];
    }
  }
};

__global__ __launch_bounds__(256) void newton_htab_kernel(sbi_newton_desc d,
                                                         const double* __restrict__ x, int64_t P,
                                                         int64_t ldx, double* __restrict__ tab,
                                                         int64_t ldt) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= P) return;
  cptr<int32_t> tdim = C(d.tab_dim), trow = C(d.tab_row), tk = C(d.tab_kmax);
  cptr<double> dim_xi = C(d.dim_xi), dim_c = C(d.dim_c);
  for (int i = blockIdx.y; i < d.n_tab_dims; i += gridDim.y) {
    const int j = tdim[i], row = trow[i], km = tk[i];
    const double xj = x[p * ldx + j];
    cptr<double> xi = dim_xi + (int64_t)j * d.kmax1;
    cptr<double> c = dim_c + (int64_t)j * d.kmax1;
    double pr = 1.0;
    for (int k = 1; k <= km; ++k) {
      pr *= (xj - xi[k - 1]);
      tab[((p / kTile) * ldt + row + k - 1) * kTile + (p % kTile)] = pr * c[k];
    }
  }
}

// Other factor rows of an adjoint entry: 16-bit rows packed first in ev.y, ev.z, ev.w, 0xffff
// after the last.  other_factor_count gives each lane's count (once per 64 entries);
// times_other_factors multiplies v by the lane's table values at the nf rows of entry q of the
// wave's 64, in the stored order, with the nf loads issued together.
__device__ __forceinline__ int other_factor_count(const uint4& ev) {
  return (int)((ev.y & 0xffff) != 0xffff) + (int)((ev.y >> 16) != 0xffff) +
         (int)((ev.z & 0xffff) != 0xffff) + (int)((ev.z >> 16) != 0xffff) +
         (int)((ev.w & 0xffff) != 0xffff) + (int)((ev.w >> 16) != 0xffff);
}

template <bool kG>
__device__ __forceinline__ double times_other_factors(const HTab<kG>& H, const uint4& ev, int q,
                                                      int nf, int lane, int64_t p, double v) {
  if (nf == 0) return v;
  const uint32_t w0 = (uint32_t)__builtin_amdgcn_readlane((int)ev.y, q);
  const double a = H.at((int)(w0 & 0xffff), lane, p);
  if (nf == 1) return v * a;
  const double b = H.at((int)(w0 >> 16), lane, p);
  if (nf == 2) return (v * a) * b;
  const uint32_t w1 = (uint32_t)__builtin_amdgcn_readlane((int)ev.z, q);
  const double c = H.at((int)(w1 & 0xffff), lane, p);
  if (nf == 3) return ((v * a) * b) * c;
  const double e = H.at((int)(w1 >> 16), lane, p);
  if (nf == 4) return (((v * a) * b) * c) * e;
  const uint32_t w2 = (uint32_t)__builtin_amdgcn_readlane((int)ev.w, q);
  const double f = H.at((int)(w2 & 0xffff), lane, p);
  if (nf == 5) return ((((v * a) * b) * c) * e) * f;
  return (((((v * a) * b) * c) * e) * f) * H.at((int)(w2 >> 16), lane, p);
}

// x_j is a node coordinate of dim j: the reference's singular mask (x - xi)^2 == 0 over the nodes
// of the dim's (dim, degree) factors (barycentric.py:152-154); the gradient is NaN there.
__device__ __forceinline__ bool on_node(const sbi_newton_desc& d, double xv, int j) {
  cptr<int32_t> so = C(d.sing_off);
  cptr<double> sx = C(d.sing_x);
  bool hit = false;
  for (int i = so[j]; i < so[j + 1]; ++i) {
    const double dv = xv - sx[i];
    hit |= (dv * dv == 0.0);
  }
  return hit;
}

// h'_k(x_j) of factor row s = (j, k) of table dim i (row s = trow[i] + k - 1).
__device__ __forceinline__ double row_derivative(const sbi_newton_desc& d, const double* xp, int j,
                                                 int k) {
  cptr<double> xi = C(d.dim_xi) + (int64_t)j * d.kmax1;
  const double xj = xp[j];
  double pr = 1.0, dpr = 0.0;
  for (int r = 0; r < k; ++r) {
    const double dx = xj - xi[r];
    dpr = dpr * dx + pr;
    pr *= dx;
  }
  return dpr * C(d.dim_c)[(int64_t)j * d.kmax1 + k];
}

// One pass over n <= NO outputs, sums in registers.  Waves take work units (runs of at most 128
// adjoint entries of one dim) round-robin over grid.y workgroups per tile; a unit that covers
// its dim writes J, the others write partial sums that jac_reduce_kernel adds in unit order.
//
// TG = void: the coefficients are the surplus rows S[node, o] (the same for every point, scalar
// loads).  TG = float / double: the adjoint (VJP), one output, the coefficient of point p is
// G^T[node, p] = (S V^T)[node, p] (a per-lane load); J is then gx[P, d_in].
template <int NO, bool kG, typename TG = void>
__global__ __launch_bounds__(kTile* kBWaves) void newton_jac_kernel(
    sbi_newton_desc d, const double* __restrict__ x, int64_t P, int64_t ldx,
    const double* __restrict__ S, int64_t lds, int n, const double* __restrict__ gtab,
    int64_t ldt, double* __restrict__ J, double* __restrict__ part,
    const void* __restrict__ GTv = nullptr, int64_t ldg = 0) {
  extern __shared__ double lds_raw[];
  typedef typename lds_t<double>::type LT;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t p_raw = (int64_t)blockIdx.x * kTile + lane;
  const bool valid = p_raw < P;
  const int64_t p = valid ? p_raw : (P - 1);
  const double* __restrict__ xp = x + p * ldx;
  HTab<kG> H{(LT*)lds_raw, gtab, ldt};
  if constexpr (!kG) {
    build_h_table<double>(d, xp, H.lds, nullptr, wave, lane, kBWaves);
    __syncthreads();
  }
  cptr<int32_t> tdim = C(d.tab_dim), trow = C(d.tab_row),
                unit = C(d.jac_unit);
  const uint4* __restrict__ ent = reinterpret_cast<const uint4*>(d.vjp_ent);
  const int nw = (int)gridDim.y * kBWaves;
  for (int u = (int)blockIdx.y * kBWaves + wave; u < d.n_jac_units; u += nw) {
    const int ti = unit[4 * u], e0 = unit[4 * u + 1], e1 = unit[4 * u + 2],
              slot = unit[4 * u + 3];
    const int j = tdim[ti], row0 = trow[ti];
    int s_done = -1;
    double dh = 0.0;
    double g[NO];
#pragma unroll
    for (int o = 0; o < NO; ++o) g[o] = 0.0;
    for (int base = e0; base < e1; base += 64) {
      const int mine = base + lane;
      const uint4 ev = (mine < e1) ? ent[mine]
                                   : make_uint4(0u, 0xffffffffu, 0xffffffffu, 0xffffffffu);
      const int rl = d.vjp_ent_row[min(mine, e1 - 1)];
      const int nfl = other_factor_count(ev);
      const int cnt = min(64, e1 - base);
      // one output: each lane fetches the surplus of its entry, broadcast per entry below
      double sl = 0.0;
      if constexpr (std::is_void_v<TG> && NO == 1) sl = S[(int64_t)ev.x * lds];
      // adjoint: the per-point coefficients of kPre entries are requested together, ahead of
      // their use (independent loads in flight instead of one at a time)
      constexpr int kPre = std::is_void_v<TG> ? 1 : 8;
      for (int q0 = 0; q0 < cnt; q0 += kPre) {
        double cv[kPre];
        (void)cv;
        if constexpr (!std::is_void_v<TG>) {
          const TG* __restrict__ GT = reinterpret_cast<const TG*>(GTv);
#pragma unroll
          for (int u = 0; u < kPre; ++u) {
            const int nq = __builtin_amdgcn_readlane((int)ev.x, min(q0 + u, cnt - 1));
            cv[u] = (double)GT[(int64_t)nq * ldg + p];
          }
        }
#pragma unroll
        for (int u = 0; u < kPre; ++u) {
          const int q = q0 + u;
          if (q >= cnt) break;
          const int s = __builtin_amdgcn_readlane(rl, q);
          if (s != s_done) {  // next factor row of the dim
            dh = row_derivative(d, xp, j, s - row0 + 1);
            s_done = s;
          }
          const double v =
              times_other_factors(H, ev, q, __builtin_amdgcn_readlane(nfl, q), lane, p, dh);
          if constexpr (std::is_void_v<TG> && NO == 1) {
            const long long sb = __double_as_longlong(sl);
            const int lo = __builtin_amdgcn_readlane((int)(sb & 0xffffffff), q);
            const int hi = __builtin_amdgcn_readlane((int)(sb >> 32), q);
            g[0] = fma(v, __longlong_as_double(((long long)hi << 32) | (unsigned int)lo), g[0]);
          } else if constexpr (std::is_void_v<TG>) {
            const int node = __builtin_amdgcn_readlane((int)ev.x, q);
            cptr<double> Sn = C(S) + (int64_t)node * lds;
#pragma unroll
            for (int o = 0; o < NO; ++o)
              if (o < n) g[o] = fma(v, Sn[o], g[o]);
          } else {
            g[0] = fma(v, cv[u], g[0]);
          }
        }
      }
    }
    if (valid) {
      if (slot < 0) {
        if constexpr (std::is_void_v<TG>) {  // Jacobian: NaN at node coordinates
          if (on_node(d, xp[j], j)) {
#pragma unroll
            for (int o = 0; o < NO; ++o) g[o] = NAN;
          }
        }
#pragma unroll
        for (int o = 0; o < NO; ++o)
          if (o < n) J[(p * n + o) * d.d_in + j] = g[o];
      } else {
#pragma unroll
        for (int o = 0; o < NO; ++o)
          if (o < n) part[((int64_t)slot * P + p) * n + o] = g[o];
      }
    }
  }
}

// J[p, o, j] = sum of the partial sums of dim j's units, in unit order (with x: the Jacobian,
// NaN at node coordinates; without: the adjoint).
__global__ __launch_bounds__(256) void jac_reduce_kernel(sbi_newton_desc d, int64_t P, int64_t n,
                                                        const double* __restrict__ part,
                                                        double* __restrict__ J,
                                                        const double* __restrict__ x, int64_t ldx) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= P) return;
  cptr<int32_t> red = C(d.jac_red);
  const int r = blockIdx.y;
  const int j = red[3 * r], s0 = red[3 * r + 1], ns = red[3 * r + 2];
  const bool nan_here = x && on_node(d, x[p * ldx + j], j);
  for (int64_t o = 0; o < n; ++o) {
    double acc = 0.0;
    for (int u = 0; u < ns; ++u) acc += part[((int64_t)(s0 + u) * P + p) * n + o];
    J[(p * n + o) * d.d_in + j] = nan_here ? NAN : acc;
  }
}

// D^T[e, p] for all adjoint entries, in groups of 64 entries per wave; the workgroups of a tile
// (grid.y of them, kDTWaves waves each) share out the groups.  The factor row of each entry and
// its table dim are read per lane from the plan's tables; h'_k(x_j) is recomputed when the row
// changes.
constexpr int kDTWaves = 8;

template <bool kG>
__global__ __launch_bounds__(kTile* kDTWaves) void newton_jac_basis_kernel(
    sbi_newton_desc d, const double* __restrict__ x, int64_t P, int64_t ldx,
    const double* __restrict__ gtab, int64_t ldt, double* __restrict__ DT, int64_t ldd) {
  extern __shared__ double lds_raw[];
  typedef typename lds_t<double>::type LT;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t p_raw = (int64_t)blockIdx.x * kTile + lane;
  const bool valid = p_raw < P;
  const int64_t p = valid ? p_raw : (P - 1);
  const double* __restrict__ xp = x + p * ldx;
  HTab<kG> H{(LT*)lds_raw, gtab, ldt};
  if constexpr (!kG) {
    build_h_table<double>(d, xp, H.lds, nullptr, wave, lane, kDTWaves);
    __syncthreads();
  }
  cptr<int32_t> tdim = C(d.tab_dim);
  const uint4* __restrict__ ent = reinterpret_cast<const uint4*>(d.vjp_ent);
  const int n_ent = d.n_vjp_ent;
  const int ngrp = (n_ent + 63) / 64;
  const int nw = (int)gridDim.y * kDTWaves;
  for (int gi = (int)blockIdx.y * kDTWaves + wave; gi < ngrp; gi += nw) {
    const int base = gi * 64;
    const int mine = min(base + lane, n_ent - 1);
    const uint4 ev = ent[mine];
    const int rl = d.vjp_ent_row[mine];
    const int nfl = other_factor_count(ev);
    const int il = d.row_tdim[rl];  // table dim of the row
    const int kl = rl - d.tab_row[il] + 1;
    const int cnt = min(64, n_ent - base);
    int s_done = -1;
    double dh = 0.0;
    for (int q = 0; q < cnt; ++q) {
      const int e = base + q;
      const int s = __builtin_amdgcn_readlane(rl, q);
      if (s != s_done) {
        const int i = __builtin_amdgcn_readlane(il, q), k = __builtin_amdgcn_readlane(kl, q);
        dh = row_derivative(d, xp, tdim[i], k);
        s_done = s;
      }
      const double v =
          times_other_factors(H, ev, q, __builtin_amdgcn_readlane(nfl, q), lane, p, dh);
      if (valid) DT[(int64_t)e * ldd + p] = v;
    }
  }
}

// J[p, :, j] = NaN where x_j is a node coordinate of dim j (the reference's singular mask
// (x - xi)^2 == 0 of barycentric.py:152-154); one thread per (p, j).
__global__ __launch_bounds__(256) void jac_mark_nodes_kernel(sbi_newton_desc d,
                                                            const double* __restrict__ x,
                                                            int64_t P, int64_t ldx, int64_t n,
                                                            double* __restrict__ J) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= P * d.d_in) return;
  const int64_t p = t / d.d_in;
  const int64_t j = t - p * d.d_in;
  const double xv = x[p * ldx + j];
  bool hit = false;
  for (int i = d.sing_off[j]; i < d.sing_off[j + 1]; ++i) {
    const double dv = xv - d.sing_x[i];
    hit |= (dv * dv == 0.0);
  }
  if (hit)
    for (int64_t o = 0; o < n; ++o) J[(p * n + o) * d.d_in + j] = NAN;
}

// Zero the gradient of dims that no factor touches (written separately so every entry of gx is
// defined).
__global__ void fill_zero_kernel(double* __restrict__ g, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) g[i] = 0.0;
}

// W^T[e, p] for the entries of a chunk of barycentric blocks.  Factor order on the device:
// prefix factors, then the streamed one; entries in C order (last factor fastest).  A factor
// value Lhat_mu(x) = w_mu prod_{nu != mu} (x - xi_nu) (exact one-hot at nodes) is formed once per
// block for the last factor and once per index of the other factors for those.
__device__ __forceinline__ double lagrange_first_form(double xv, cptr<double> xi, cptr<double> w,
                                                      int deg, int mu) {
  double v = w[mu];
  bool hit_other = false, hit_self = false;
  for (int q = 0; q <= deg; ++q) {
    const double dq = xv - xi[q];
    if (q == mu) {
      hit_self = (dq == 0.0);
    } else {
      hit_other |= (dq == 0.0);
      v *= dq;
    }
  }
  if (hit_self) return 1.0;
  if (hit_other) return 0.0;
  return v;
}

constexpr int kBLMax = 12;  // last-factor basis values held in registers (degree < 12)

template <typename OUT>
__global__ __launch_bounds__(kTile) void bary_basis_kernel(sbi_bary_desc d,
                                                          const double* __restrict__ x, int64_t P,
                                                          int64_t ldx, OUT* __restrict__ WT,
                                                          int64_t ldw, int chunk) {
  const int lane = threadIdx.x;
  const int64_t p_raw = (int64_t)blockIdx.x * kTile + lane;
  const bool valid = p_raw < P;
  const int64_t p = valid ? p_raw : (P - 1);
  const double* __restrict__ xp = x + p * ldx;
  cptr<int32_t> blk_fac = C(d.blk_fac), fac_slot = C(d.fac_slot), last = C(d.blk_last_slot),
                slot_dim = C(d.slot_dim), slot_deg = C(d.slot_deg);
  cptr<int64_t> foff = C(d.blk_foff);
  cptr<double> nodes = C(d.slot_nodes), wts = C(d.slot_w);
  const int b0 = blockIdx.y * chunk;
  const int b1 = min(d.n_blocks, b0 + chunk);
  for (int b = b0; b < b1; ++b) {
    const int f0 = blk_fac[b], npre = blk_fac[b + 1] - f0, ls = last[b];
    const int nf = npre + (ls >= 0 ? 1 : 0);
    const int64_t r0 = foff[b];
    if (nf == 0) {  // the constant block: one entry of weight 1
      if (valid) WT[r0 * ldw + p] = (OUT)1.0;
      continue;
    }
    // entries e = q * (dl + 1) + m in C order: m indexes the last factor, q the others (the
    // second-to-last fastest).  The last factor's basis is computed once per block, the product
    // of the others once per q.
    auto slot_of = [&](int i) { return (i < npre) ? fac_slot[f0 + i] : ls; };
    const int sl = slot_of(nf - 1);
    const int dl = slot_deg[sl];
    const double xl = xp[slot_dim[sl]];
    cptr<double> nl = nodes + (int64_t)sl * d.dmax1, wl = wts + (int64_t)sl * d.dmax1;
    double Ll[kBLMax];
#pragma unroll
    for (int m = 0; m < kBLMax; ++m)
      Ll[m] = (m <= dl) ? lagrange_first_form(xl, nl, wl, dl, m) : 0.0;
    int nq = 1;
    for (int i = 0; i < nf - 1; ++i) nq *= slot_deg[slot_of(i)] + 1;
    for (int q = 0; q < nq; ++q) {
      double pre = 1.0;
      int rem = q;
      for (int i = nf - 2; i >= 0; --i) {
        const int si = slot_of(i);
        const int di1 = slot_deg[si] + 1;
        const int mu = rem % di1;
        rem /= di1;
        pre *= lagrange_first_form(xp[slot_dim[si]], nodes + (int64_t)si * d.dmax1,
                                   wts + (int64_t)si * d.dmax1, di1 - 1, mu);
      }
      OUT* __restrict__ out = WT + (r0 + (int64_t)q * (dl + 1)) * ldw + p;
      if (dl < kBLMax) {
#pragma unroll
        for (int m = 0; m < kBLMax; ++m)
          if (m <= dl && valid) out[(int64_t)m * ldw] = (OUT)(Ll[m] * pre);
      } else {  // degree beyond the register table: the last factor's basis entry by entry
        for (int m = 0; m <= dl; ++m)
          if (valid) out[(int64_t)m * ldw] = (OUT)(lagrange_first_form(xl, nl, wl, dl, m) * pre);
      }
    }
  }
}

}  // namespace

extern "C" int sbi_newton_plan_desc(const sbi_newton_plan* plan, const sbi_newton_desc** out);

template <typename T, typename OUT, int kQ = 0>
static int basis_launch(const sbi_newton_desc& d, const double* x, int64_t P, int64_t ldx, OUT* HT,
                        int64_t ldh, int row0, int row1, hipStream_t st) {
  const size_t lds = (size_t)d.n_tab_rows * kTile * sizeof(T);
  const bool direct = lds > (size_t)kMaxLds;  // factor values computed where needed
  if (direct && !d.row_tdim) {
    sbi::set_error("sbi_newton_basis: the tile's factor table exceeds LDS and the plan has no "
                   "row_tdim table");
    return SBI_ERR_CAPACITY;
  }
  const int chunk = 2048;
  if (row1 <= row0) return SBI_OK;
  constexpr int kPad = kq_traits<kQ>::pad;
  const int rows = kQ ? ((row1 - row0 + kPad - 1) & ~(kPad - 1)) : row1 - row0;
  dim3 grid((unsigned)((P + kTile - 1) / kTile), (unsigned)((rows + chunk - 1) / chunk));
  if (direct) {
    hipLaunchKernelGGL((newton_basis_direct_kernel<T, OUT, kQ>), grid, dim3(kTile * kHWaves), 0,
                       st, d, x, P, ldx, HT, ldh, chunk, row0, row1);
    SBI_HIP_CHECK(hipGetLastError());
    return SBI_OK;
  }
  const int64_t tiles = (P + kTile - 1) / kTile;
  if (tiles * ((rows + chunk - 1) / chunk) < 256) {  // small batch: 4-wave workgroups, 256 rows
    constexpr int kSmall = 256;
    const dim3 g2((unsigned)tiles, (unsigned)((rows + kSmall - 1) / kSmall));
    sbi::allow_lds((const void*)newton_basis_kernel<T, OUT, kQ, 4>);
    hipLaunchKernelGGL((newton_basis_kernel<T, OUT, kQ, 4>), g2, dim3(kTile * 4), lds, st, d, x,
                       P, ldx, HT, ldh, kSmall, row0, row1);
    SBI_HIP_CHECK(hipGetLastError());
    return SBI_OK;
  }
  sbi::allow_lds((const void*)newton_basis_kernel<T, OUT, kQ>);
  hipLaunchKernelGGL((newton_basis_kernel<T, OUT, kQ>), grid, dim3(kTile * kHWaves), lds, st, d, x,
                     P, ldx, HT, ldh, chunk, row0, row1);
  SBI_HIP_CHECK(hipGetLastError());
  return SBI_OK;
}

extern "C" int sbi_newton_basis_rows(const sbi_newton_plan* plan, const double* x, int64_t P,
                                     int64_t ldx, int32_t row0, int32_t row1, void* HT,
                                     int64_t ldh, int32_t dtype, void* stream) {
  const sbi_newton_desc* dp = nullptr;
  if (sbi_newton_plan_desc(plan, &dp) != SBI_OK || !x || !HT || P < 0 || ldh < P) {
    sbi::set_error("sbi_newton_basis: invalid argument");
    return SBI_ERR_INVALID;
  }
  const sbi_newton_desc& d = *dp;
  if (!d.basis_desc || !d.tab_dim || ldx < d.d_in) {
    sbi::set_error("sbi_newton_basis: plan has no basis tables");
    return SBI_ERR_INVALID;
  }
  if (row0 < 0 || row1 > d.basis_n || row0 > row1) {
    sbi::set_error("sbi_newton_basis: row range outside [0, basis_n]");
    return SBI_ERR_INVALID;
  }
  if (P == 0 || row0 == row1) return SBI_OK;
  hipStream_t st = (hipStream_t)stream;
  switch (dtype) {
    case 0: return basis_launch<double, double>(d, x, P, ldx, (double*)HT, ldh, row0, row1, st);
    case 1: return basis_launch<float, float>(d, x, P, ldx, (float*)HT, ldh, row0, row1, st);
    case 2:
      return basis_launch<float, __hip_bfloat16>(d, x, P, ldx, (__hip_bfloat16*)HT, ldh, row0,
                                                 row1, st);
    case 3:
    case 4:
      if (((uintptr_t)HT & 15) != 0) {
        sbi::set_error("sbi_newton_basis: the k-interleaved layouts need a 16-byte aligned HT");
        return SBI_ERR_INVALID;
      }
      if (dtype == 3)
        return basis_launch<float, float, 4>(d, x, P, ldx, (float*)HT, ldh, row0, row1, st);
      return basis_launch<float, __hip_bfloat16, 8>(d, x, P, ldx, (__hip_bfloat16*)HT, ldh, row0,
                                                     row1, st);
    default:
      sbi::set_error("sbi_newton_basis: dtype must be 0 (f64), 1 (f32), 2 (bf16), 3 (f32, "
                     "k-interleaved by 4) or 4 (bf16, k-interleaved by 8)");
      return SBI_ERR_INVALID;
  }
}

extern "C" int sbi_newton_basis(const sbi_newton_plan* plan, const double* x, int64_t P,
                                int64_t ldx, void* HT, int64_t ldh, int32_t dtype, void* stream) {
  const sbi_newton_desc* dp = nullptr;
  if (sbi_newton_plan_desc(plan, &dp) != SBI_OK) {
    sbi::set_error("sbi_newton_basis: invalid argument");
    return SBI_ERR_INVALID;
  }
  return sbi_newton_basis_rows(plan, x, P, ldx, 0, dp->basis_n, HT, ldh, dtype, stream);
}

extern "C" int sbi_newton_rows_f64(const sbi_newton_plan* plan, const double* x, int64_t P,
                                   int64_t ldx, const int32_t* fac, int32_t n_rows, double* HT,
                                   int64_t ldh, void* stream) {
  const sbi_newton_desc* dp = nullptr;
  if (sbi_newton_plan_desc(plan, &dp) != SBI_OK || !x || (n_rows > 0 && (!fac || !HT)) ||
      P < 0 || ldh < P || n_rows < 0 || n_rows > 65535 * kRowsPerBlock || ldx < dp->d_in) {
    sbi::set_error("sbi_newton_rows_f64: invalid argument");
    return SBI_ERR_INVALID;
  }
  if (P == 0 || n_rows == 0) return SBI_OK;
  if ((P + 63) / 64 > 0x7fffffff) {
    sbi::set_error("sbi_newton_rows_f64: too many points for one call");
    return SBI_ERR_INVALID;
  }
  const dim3 grid((unsigned)((P + 63) / 64), (unsigned)((n_rows + kRowsPerBlock - 1) / kRowsPerBlock));
  hipLaunchKernelGGL(newton_rows_f64_kernel, grid, dim3(256), 0, (hipStream_t)stream, *dp, x, P,
                     ldx, fac, n_rows, HT, ldh);
  SBI_HIP_CHECK(hipGetLastError());
  return SBI_OK;
}

static int64_t align256(int64_t v) { return (v + 255) / 256 * 256; }

extern "C" int64_t sbi_newton_eval_mixed_gemm_workspace(const sbi_newton_plan* plan, int64_t P) {
  const sbi_newton_desc* dp = nullptr;
  if (sbi_newton_plan_desc(plan, &dp) != SBI_OK || P < 0 || dp->gemm_head + dp->gemm_kq <= 0) {
    sbi::set_error("sbi_newton_eval_mixed_gemm_workspace: invalid argument (or the plan has no "
                   "mixed GEMM fields)");
    return -1;
  }
  const int64_t sk = sbi_contract_split_k4_workspace(P, dp->d_out, dp->gemm_head, dp->gemm_kq);
  if (sk < 0) return -1;
  return align256((int64_t)dp->gemm_kq * P * 4) +
         align256((int64_t)(dp->gemm_head > 0 ? dp->gemm_head : 1) * P * 8) + sk;
}

extern "C" int sbi_newton_eval_mixed_gemm(const sbi_newton_plan* plan, const double* x, int64_t P,
                                          int64_t ldx, double* y, int64_t ldy, void* workspace,
                                          int64_t ws_bytes, void* stream) {
  const sbi_newton_desc* dp = nullptr;
  if (sbi_newton_plan_desc(plan, &dp) != SBI_OK || !x || !y || P < 0 || ldx < dp->d_in ||
      ldy < dp->d_out || dp->gemm_head + dp->gemm_kq <= 0 || !workspace ||
      ((uintptr_t)workspace & 255) || P > (int64_t)65535 * 64) {
    sbi::set_error("sbi_newton_eval_mixed_gemm: invalid argument (or the plan has no mixed GEMM "
                   "fields, or the workspace is not 256-byte aligned)");
    return SBI_ERR_INVALID;
  }
  if (P == 0) return SBI_OK;
  const sbi_newton_desc& d = *dp;
  const int64_t ht = align256((int64_t)d.gemm_kq * P * 4);
  const int64_t h64 = align256((int64_t)(d.gemm_head > 0 ? d.gemm_head : 1) * P * 8);
  if (ws_bytes < ht + h64) {
    sbi::set_error("sbi_newton_eval_mixed_gemm: workspace smaller than "
                   "sbi_newton_eval_mixed_gemm_workspace");
    return SBI_ERR_INVALID;
  }
  char* w = (char*)workspace;
  float* HT = (float*)w;
  double* H64 = (double*)(w + ht);
  int rc = SBI_OK;
  if (d.basis_n > d.gemm_head) {
    rc = sbi_newton_basis_rows(plan, x, P, ldx, d.gemm_head, d.basis_n, HT, P, 3, stream);
    if (rc != SBI_OK) return rc;
  }
  if (d.gemm_head > 0) {
    rc = sbi_newton_rows_f64(plan, x, P, ldx, d.head_fac, d.gemm_head, H64, P, stream);
    if (rc != SBI_OK) return rc;
  }
  return sbi_contract_split_k4_ws(H64, P, d.gemm_s64, d.ld_s64, d.gemm_head,
                                  d.gemm_kq ? HT : nullptr, P, d.gemm_kq ? d.gemm_sq : nullptr,
                                  d.d_out, d.gemm_kq, nullptr, y, ldy, P, d.d_out,
                                  w + ht + h64, ws_bytes - ht - h64, stream);
}

static bool jac_global_table(const sbi_newton_desc& d);

extern "C" int64_t sbi_newton_vjp_workspace(const sbi_newton_plan* plan, int64_t P) {
  const sbi_newton_desc* dp = nullptr;
  if (sbi_newton_plan_desc(plan, &dp) != SBI_OK || P < 0) {
    sbi::set_error("sbi_newton_vjp_workspace: null plan or negative P");
    return -1;
  }
  const int64_t P64 = (P + kTile - 1) / kTile * kTile;
  int64_t bytes = (int64_t)dp->n_jac_partials * P * (int64_t)sizeof(double);
  if (jac_global_table(*dp)) bytes += (int64_t)dp->n_tab_rows * P64 * (int64_t)sizeof(double);
  return bytes;
}

template <typename T>
static int vjp_launch(const sbi_newton_desc& d, const double* x, int64_t P, int64_t ldx,
                      const T* GT, int64_t ldg, double* gx, void* workspace, hipStream_t st) {
  const bool g = jac_global_table(d);
  const int64_t P64 = (P + kTile - 1) / kTile * kTile;
  double* gtab = g ? (double*)workspace : nullptr;
  double* part = (double*)workspace + (g ? (int64_t)d.n_tab_rows * P64 : 0);
  if (g && d.n_tab_dims > 0) {
    const int64_t ny = std::min<int64_t>(d.n_tab_dims, 64);
    hipLaunchKernelGGL(newton_htab_kernel, dim3((unsigned)((P + 255) / 256), (unsigned)ny),
                       dim3(256), 0, st, d, x, P, ldx, gtab, (int64_t)d.n_tab_rows);
    SBI_HIP_CHECK(hipGetLastError());
  }
  if (d.n_tab_dims < d.d_in) {  // dims without a factor keep 0
    const int64_t n = P * d.d_in;
    hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, gx,
                       n);
    SBI_HIP_CHECK(hipGetLastError());
  }
  if (d.n_jac_units > 0) {
    const size_t lds_bytes = g ? 0 : (size_t)d.n_tab_rows * kTile * sizeof(double);
    const int64_t tiles = (P + kTile - 1) / kTile;
    const int64_t ny = std::max<int64_t>(
        1, std::min<int64_t>((d.n_jac_units + kBWaves - 1) / kBWaves, (4096 + tiles - 1) / tiles));
    const dim3 grid((unsigned)tiles, (unsigned)ny);
    if (g) {
      sbi::allow_lds((const void*)newton_jac_kernel<1, true, T>);
      hipLaunchKernelGGL((newton_jac_kernel<1, true, T>), grid, dim3(kTile * kBWaves), lds_bytes,
                         st, d, x, P, ldx, nullptr, (int64_t)0, 1, gtab, (int64_t)d.n_tab_rows,
                         gx, part, (const void*)GT, ldg);
    } else {
      sbi::allow_lds((const void*)newton_jac_kernel<1, false, T>);
      hipLaunchKernelGGL((newton_jac_kernel<1, false, T>), grid, dim3(kTile * kBWaves), lds_bytes,
                         st, d, x, P, ldx, nullptr, (int64_t)0, 1, gtab, (int64_t)d.n_tab_rows,
                         gx, part, (const void*)GT, ldg);
    }
    SBI_HIP_CHECK(hipGetLastError());
  }
  if (d.n_jac_red > 0) {
    hipLaunchKernelGGL(jac_reduce_kernel, dim3((unsigned)((P + 255) / 256), (unsigned)d.n_jac_red),
                       dim3(256), 0, st, d, P, (int64_t)1, part, gx, (const double*)nullptr,
                       (int64_t)0);
    SBI_HIP_CHECK(hipGetLastError());
  }
  return SBI_OK;
}

extern "C" int sbi_newton_vjp(const sbi_newton_plan* plan, const double* x, int64_t P, int64_t ldx,
                              const void* GT, int64_t ldg, int32_t dtype, double* gx,
                              void* workspace, int64_t ws_bytes, void* stream) {
  const sbi_newton_desc* dp = nullptr;
  if (sbi_newton_plan_desc(plan, &dp) != SBI_OK || !x || !GT || !gx || P < 0 || ldg < P ||
      ws_bytes < 0) {
    sbi::set_error("sbi_newton_vjp: invalid argument");
    return SBI_ERR_INVALID;
  }
  const sbi_newton_desc& d = *dp;
  if (!d.vjp_ent || !d.vjp_slot_off || !d.tab_dim || !d.jac_unit || !d.jac_red ||
      !d.vjp_ent_row || ldx < d.d_in) {
    sbi::set_error("sbi_newton_vjp: plan has no adjoint tables");
    return SBI_ERR_INVALID;
  }
  if (P == 0) return SBI_OK;
  if (P > (int64_t)65535 * kTile) {
    sbi::set_error("sbi_newton_vjp: at most 65535*64 points per call");
    return SBI_ERR_INVALID;
  }
  const int64_t need = sbi_newton_vjp_workspace(plan, P);
  if (ws_bytes < need || (need > 0 && !workspace)) {
    sbi::set_error("sbi_newton_vjp: workspace smaller than sbi_newton_vjp_workspace()");
    return SBI_ERR_INVALID;
  }
  hipStream_t st = (hipStream_t)stream;
  if (dtype == 0) return vjp_launch<double>(d, x, P, ldx, (const double*)GT, ldg, gx, workspace, st);
  if (dtype == 1) return vjp_launch<float>(d, x, P, ldx, (const float*)GT, ldg, gx, workspace, st);
  sbi::set_error("sbi_newton_vjp: dtype must be 0 (f64) or 1 (f32)");
  return SBI_ERR_INVALID;
}

extern "C" int sbi_bary_plan_desc(const sbi_plan* plan, const sbi_bary_desc** out);

template <typename OUT>
static int bary_basis_launch(const sbi_plan* plan, const double* x, int64_t P, int64_t ldx, OUT* WT,
                             int64_t ldw, void* stream) {
  const sbi_bary_desc* dp = nullptr;
  if (sbi_bary_plan_desc(plan, &dp) != SBI_OK || !x || !WT || P < 0 || ldw < P) {
    sbi::set_error("sbi_bary_basis: invalid argument");
    return SBI_ERR_INVALID;
  }
  const sbi_bary_desc& d = *dp;
  if (ldx < d.d_in) {
    sbi::set_error("sbi_bary_basis: invalid argument (ldx < d_in)");
    return SBI_ERR_INVALID;
  }
  if (P == 0 || d.n_blocks == 0) return SBI_OK;
  const int chunk = 64;
  dim3 grid((unsigned)((P + kTile - 1) / kTile), (unsigned)((d.n_blocks + chunk - 1) / chunk));
  hipLaunchKernelGGL(bary_basis_kernel<OUT>, grid, dim3(kTile), 0, (hipStream_t)stream, d, x, P,
                     ldx, WT, ldw, chunk);
  SBI_HIP_CHECK(hipGetLastError());
  return SBI_OK;
}

extern "C" int sbi_bary_basis_f64(const sbi_plan* plan, const double* x, int64_t P, int64_t ldx,
                                  double* WT, int64_t ldw, void* stream) {
  return bary_basis_launch<double>(plan, x, P, ldx, WT, ldw, stream);
}

extern "C" int sbi_bary_basis_f32(const sbi_plan* plan, const double* x, int64_t P, int64_t ldx,
                                  float* WT, int64_t ldw, void* stream) {
  return bary_basis_launch<float>(plan, x, P, ldx, WT, ldw, stream);
}

// ------------------------------------------------------------------------------------------
// Dense Jacobian of the Newton form
// ------------------------------------------------------------------------------------------
constexpr int kJacDirectMax = 8;  // outputs handled in one register pass

// The Jacobian and adjoint passes are bound by latency per entry, so they need many waves per
// CU: a factor table above kJacLdsMax (which would leave room for too few workgroups per CU) goes
// to global memory, written once by newton_htab_kernel and read through the caches.
constexpr size_t kJacLdsMax = 40 * 1024;
static bool jac_global_table(const sbi_newton_desc& d) {
  return (size_t)d.n_tab_rows * kTile * sizeof(double) > kJacLdsMax;
}

extern "C" int64_t sbi_newton_jacobian_workspace(const sbi_newton_plan* plan, int64_t P,
                                                 int64_t n) {
  const sbi_newton_desc* dp = nullptr;
  if (sbi_newton_plan_desc(plan, &dp) != SBI_OK || P < 0 || n < 0) {
    sbi::set_error("sbi_newton_jacobian_workspace: null plan or negative P or n");
    return -1;
  }
  int64_t bytes = 0;
  const int64_t P64 = (P + kTile - 1) / kTile * kTile;
  if (jac_global_table(*dp)) bytes += (int64_t)dp->n_tab_rows * P64 * (int64_t)sizeof(double);
  if (n > kJacDirectMax) bytes += (int64_t)dp->n_vjp_ent * P * (int64_t)sizeof(double);
  else bytes += (int64_t)dp->n_jac_partials * P * n * (int64_t)sizeof(double);
  return bytes;
}

template <int NO, bool kG>
static void jac_direct_launch(const sbi_newton_desc& d, const double* x, int64_t P, int64_t ldx,
                              const double* S, int64_t lds, int n, const double* gtab,
                              double* J, double* part, hipStream_t st) {
  const size_t lds_bytes = kG ? 0 : (size_t)d.n_tab_rows * kTile * sizeof(double);
  sbi::allow_lds((const void*)newton_jac_kernel<NO, kG>);
  const int64_t tiles = (P + kTile - 1) / kTile;
  // workgroups per tile: enough for ~4096 workgroups, at most one wave per work unit
  const int64_t ny = std::max<int64_t>(
      1, std::min<int64_t>((d.n_jac_units + kBWaves - 1) / kBWaves, (4096 + tiles - 1) / tiles));
  hipLaunchKernelGGL((newton_jac_kernel<NO, kG>), dim3((unsigned)tiles, (unsigned)ny),
                     dim3(kTile * kBWaves), lds_bytes, st, d, x, P, ldx, S, lds, n, gtab,
                     (int64_t)d.n_tab_rows, J, part);
}

template <bool kG>
static void jac_basis_launch(const sbi_newton_desc& d, const double* x, int64_t P, int64_t ldx,
                             const double* gtab, double* DT, hipStream_t st) {
  const size_t lds_bytes = kG ? 0 : (size_t)d.n_tab_rows * kTile * sizeof(double);
  sbi::allow_lds((const void*)newton_jac_basis_kernel<kG>);
  const int64_t tiles = (P + kTile - 1) / kTile;
  const int64_t ngrp = (d.n_vjp_ent + 63) / 64;
  // the factor table is built once per workgroup: ~1024 workgroups, each wave looping over groups
  const int64_t ny = std::max<int64_t>(
      1, std::min<int64_t>((ngrp + kDTWaves - 1) / kDTWaves, (1024 + tiles - 1) / tiles));
  hipLaunchKernelGGL(newton_jac_basis_kernel<kG>, dim3((unsigned)tiles, (unsigned)ny),
                     dim3(kTile * kDTWaves), lds_bytes, st, d, x, P, ldx, gtab, (int64_t)d.n_tab_rows,
                     DT, P);
}

extern "C" int sbi_newton_jacobian_f64(const sbi_newton_plan* plan, const double* x, int64_t P,
                                       int64_t ldx, const double* S, int64_t lds, int64_t n,
                                       double* J, void* workspace, int64_t ws_bytes,
                                       void* stream) {
  const sbi_newton_desc* dp = nullptr;
  if (sbi_newton_plan_desc(plan, &dp) != SBI_OK || !x || !S || !J || P < 0 || n < 1 || lds < n ||
      ws_bytes < 0) {
    sbi::set_error("sbi_newton_jacobian_f64: invalid argument");
    return SBI_ERR_INVALID;
  }
  const sbi_newton_desc& d = *dp;
  if (!d.vjp_ent || !d.vjp_slot_off || !d.tab_dim || !d.vjp_dim_off || !d.sing_off ||
      !d.sing_x || !d.jac_unit || !d.jac_red || !d.vjp_ent_row || !d.row_tdim || ldx < d.d_in) {
    sbi::set_error("sbi_newton_jacobian_f64: plan has no Jacobian tables");
    return SBI_ERR_INVALID;
  }
  if (P == 0) return SBI_OK;
  if (d.d_in > 65535 || P > (int64_t)65535 * kTile) {
    sbi::set_error("sbi_newton_jacobian_f64: at most 65535 input dims and 65535*64 points");
    return SBI_ERR_INVALID;
  }
  const int64_t need = sbi_newton_jacobian_workspace(plan, P, n);
  if (ws_bytes < need || (need > 0 && !workspace)) {
    sbi::set_error("sbi_newton_jacobian_f64: workspace smaller than "
                   "sbi_newton_jacobian_workspace()");
    return SBI_ERR_INVALID;
  }
  hipStream_t st = (hipStream_t)stream;
  const bool g = jac_global_table(d);
  double* gtab = g ? (double*)workspace : nullptr;
  const int64_t P64 = (P + kTile - 1) / kTile * kTile;
  double* DT = (double*)workspace + (g ? (int64_t)d.n_tab_rows * P64 : 0);
  if (g && d.n_tab_dims > 0) {
    const int64_t ny = std::min<int64_t>(d.n_tab_dims, 64);
    hipLaunchKernelGGL(newton_htab_kernel, dim3((unsigned)((P + 255) / 256), (unsigned)ny),
                       dim3(256), 0, st, d, x, P, ldx, gtab, (int64_t)d.n_tab_rows);
    SBI_HIP_CHECK(hipGetLastError());
  }
  if (n <= kJacDirectMax) {
    if (d.n_tab_dims < d.d_in) {  // dims without a factor keep 0
      const int64_t tot = P * n * d.d_in;
      hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st,
                         J, tot);
      SBI_HIP_CHECK(hipGetLastError());
    }
    const int ni = (int)n;
    double* part = DT;  // partial sums of split dims: [n_jac_partials][P][n]
    if (d.n_jac_units > 0) {
      if (g) {
        if (n == 1) jac_direct_launch<1, true>(d, x, P, ldx, S, lds, ni, gtab, J, part, st);
        else if (n <= 2) jac_direct_launch<2, true>(d, x, P, ldx, S, lds, ni, gtab, J, part, st);
        else if (n <= 4) jac_direct_launch<4, true>(d, x, P, ldx, S, lds, ni, gtab, J, part, st);
        else jac_direct_launch<8, true>(d, x, P, ldx, S, lds, ni, gtab, J, part, st);
      } else {
        if (n == 1) jac_direct_launch<1, false>(d, x, P, ldx, S, lds, ni, gtab, J, part, st);
        else if (n <= 2) jac_direct_launch<2, false>(d, x, P, ldx, S, lds, ni, gtab, J, part, st);
        else if (n <= 4) jac_direct_launch<4, false>(d, x, P, ldx, S, lds, ni, gtab, J, part, st);
        else jac_direct_launch<8, false>(d, x, P, ldx, S, lds, ni, gtab, J, part, st);
      }
      SBI_HIP_CHECK(hipGetLastError());
    }
    if (d.n_jac_red > 0) {
      hipLaunchKernelGGL(jac_reduce_kernel, dim3((unsigned)((P + 255) / 256), (unsigned)d.n_jac_red),
                         dim3(256), 0, st, d, P, n, part, J, x, ldx);
      SBI_HIP_CHECK(hipGetLastError());
    }
    return SBI_OK;  // the register pass and the reduction wrote the NaNs at node coordinates
  } else {
    if (d.n_vjp_ent > 0) {
      if (g) jac_basis_launch<true>(d, x, P, ldx, gtab, DT, st);
      else jac_basis_launch<false>(d, x, P, ldx, gtab, DT, st);
      SBI_HIP_CHECK(hipGetLastError());
    }
    // one group of D^T rows per input dim (dims without entries get zeros); the surplus row of
    // an entry is its node, the first int32 of its 16-byte record
    const int rc = sbi::contract_jacobian_f64(DT, P, S, lds, (const int32_t*)d.vjp_ent, 4,
                                              d.vjp_dim_off, (int)d.d_in, J, P, n, st);
    if (rc != SBI_OK) return rc;
  }
  const int64_t tot = P * d.d_in;
  hipLaunchKernelGGL(jac_mark_nodes_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st,
                     d, x, P, ldx, n, J);
  SBI_HIP_CHECK(hipGetLastError());
  return SBI_OK;
}
