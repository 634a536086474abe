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
#include "launch.h"

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

// A window of 8 consecutive coordinates of this lane's point, fetched with independent loads
// and addressed only with compile-time indices (so it stays in registers).
constexpr int kXW = 8;
struct XWin {
  double v[kXW];
};
__device__ __forceinline__ XWin load_window(const double* __restrict__ xp, int j0, int d_in) {
  // unconditional loads (index clamped into the row): a fixed number of loads in flight lets the
  // compiler wait for exactly the older window instead of draining every outstanding load
  XWin w;
#pragma unroll
  for (int i = 0; i < kXW; ++i) w.v[i] = xp[min(j0 + i, d_in - 1)];
  return w;
}
// Same window with two coordinates per 16-byte load (row stride even, x 16-byte aligned, d_in
// even and >= 2): half the load instructions, each of which touches one cache line per lane.
// Pairs past the row end re-read the last pair (index d_in - 2, in bounds); those values are
// unused.  With an odd d_in the last coordinate has no partner inside the row, so the launcher
// takes the 8-byte form for such plans.
typedef double xv2d __attribute__((ext_vector_type(2)));
__device__ __forceinline__ XWin load_window2(const double* __restrict__ xp, int j0, int d_in) {
  const int jmax = d_in - 2;
  XWin w;
#pragma unroll
  for (int q = 0; q < kXW / 2; ++q) {
    const xv2d v = *(const xv2d*)(xp + min(j0 + 2 * q, jmax));
    w.v[2 * q] = v.x;
    w.v[2 * q + 1] = v.y;
  }
  return w;
}
template <bool kPairs>
__device__ __forceinline__ XWin load_win(const double* __restrict__ xp, int j0, int d_in) {
  if constexpr (kPairs) {
    return load_window2(xp, j0, d_in);
  } else {
    return load_window(xp, j0, d_in);
  }
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
                level_off = C(d.level_off), grp = C(d.grp), wave_grp = C(d.wave_grp);
  cptr<double> dim_xi = C(d.dim_xi), dim_c = C(d.dim_c);
  const T* __restrict__ node_s =
      (const T*)(sizeof(T) == 4 ? (const void*)d.node_s32 : (const void*)d.node_s64);
  const int32_t* __restrict__ node_q = d.node_q;

  // the wave streams its dims in increasing order through two windows of coordinates (current,
  // next); both are requested before the table work so their latency overlaps it
  const int g_begin = wave_grp[wave];
  const int g_end = wave_grp[wave + 1];
  const int d_in = (int)d.d_in;
  int wj0 = (g_begin < g_end) ? (grp[4 * g_begin + 2] & ~(kXW - 1)) : 0;
  XWin xa = load_window(xp, wj0, d_in);
  XWin xb = load_window(xp, wj0 + kXW, d_in);

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

  // 2. streamed groups, in increasing dim order, consumed window by window
  double acc = 0.0;
  if (g_begin < g_end) {
    const int nn = d.n_nodes;
    int chunk = grp[4 * g_begin] >> 6;
    // current and next metadata chunks (one node per lane)
    const int i0 = chunk * 64 + lane;
    T s_cur = (i0 < nn) ? node_s[i0] : (T)0;
    int q_cur = (i0 < nn) ? node_q[i0] : 0;
    T s_nxt = (i0 + 64 < nn) ? node_s[i0 + 64] : (T)0;
    int q_nxt = (i0 + 64 < nn) ? node_q[i0 + 64] : 0;

    int g = g_begin;
    while (g < g_end) {
#pragma unroll
      for (int u = 0; u < kXW; ++u) {
        const int j = wj0 + u;
        const double xj = xa.v[u];
        double pr = 1.0;
        int kdone = 0;  // pr = prod_{r < kdone} (x_j - xi_r)
        cptr<double> xi = dim_xi + (int64_t)j * d.kmax1;
        cptr<double> cj = dim_c + (int64_t)j * d.kmax1;
        // the groups of dim j come in increasing deg k
        while (g < g_end && grp[4 * g + 2] == j) {
          const int n0 = grp[4 * g], n1 = grp[4 * g + 1], k = grp[4 * g + 3];
          for (; kdone < k; ++kdone) pr *= (xj - xi[kdone]);
          const double h = pr * cj[k];
          T t = (T)0;
          int i = n0;
          while (i < n1) {
            if ((i >> 6) != chunk) {
              // groups are visited in node order, so this is always the next chunk
              chunk += 1;
              s_cur = s_nxt;
              q_cur = q_nxt;
              const int in = (chunk + 1) * 64 + lane;
              s_nxt = (in < nn) ? node_s[in] : (T)0;
              q_nxt = (in < nn) ? node_q[in] : 0;
            }
            const int lim = min(n1, (chunk + 1) * 64);
            // batches of 4 nodes: broadcast metadata, gather parents' table entries, accumulate
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
              t = fma(bcast(s_cur, l), W[__builtin_amdgcn_readlane(q_cur, l) * kTile + lane], t);
            }
          }
          acc = fma(h, (double)t, acc);
          ++g;
        }
      }
      // next window: slide (requesting the one after) or jump over a gap in the dims
      if (g < g_end) {
        const int jn = grp[4 * g + 2];
        if (jn >= wj0 + 2 * kXW) {
          wj0 = jn & ~(kXW - 1);
          xa = load_window(xp, wj0, d_in);
          xb = load_window(xp, wj0 + kXW, d_in);
        } else {
          wj0 += kXW;
          xa = xb;
          xb = load_window(xp, wj0 + kXW, d_in);
        }
      }
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


// ------------------------------------------------------------------------------------------
// Staged mixed-precision kernel.  Everything the streaming loop reads is placed in LDS once per
// tile by asynchronous global->LDS copies (global_load_lds_dwordx4), so the loop's only memory
// operations are LDS reads (in order, batched by the compiler) and the per-lane coordinate
// windows, which are requested one window ahead.  LDS layout:
//   [W table: (n_internal+1) x 64 floats][region R]
//   region R, phase 1: h table of the internal nodes' factors (n_hs_rows x 64 floats)
//   region R, phase 2: node pairs (8 bytes per node: c_deg * surplus as FP32 bits, byte offset of
//                      the parent's W row) | group records (16 bytes: first node | end node << 16,
//                      dim, a = xi_{deg-1} as two words)
// Every group starts and ends on an even node (padding nodes have surplus 0), so two nodes are
// one aligned 16-byte broadcast read.
// ------------------------------------------------------------------------------------------
typedef int v4i __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));
typedef unsigned int v2u __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) float lds_f32;
typedef __attribute__((address_space(3))) double lds_f64;
typedef __attribute__((address_space(3))) void lds_void;

__device__ __forceinline__ void stage_to_lds(const void* src, lds_void* dst, int nbytes, int wave,
                                             int lane) {
  // nbytes is a multiple of 16; each wave copies 1 KiB chunks, one 16-byte piece per lane
  const int nchunks = (nbytes + 1023) >> 10;
  for (int c = wave; c < nchunks; c += kWaves) {
    const int off = c * 1024 + lane * 16;
    if (off < nbytes)
      __builtin_amdgcn_global_load_lds((const char*)src + off,
                                       (lds_void*)((__attribute__((address_space(3))) char*)dst +
                                                   c * 1024),
                                       16, 0, 0);
  }
}

template <bool kPairs>
__global__ __launch_bounds__(kTile* kWaves) void newton_mixed_staged_kernel(
    sbi_newton_desc d, const double* __restrict__ x, int64_t P, int64_t ldx, double* __restrict__ y) {
  extern __shared__ double lds_raw[];
  lds_f32* W = (lds_f32*)lds_raw;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t p_raw = (int64_t)blockIdx.x * kTile + lane;
  const bool valid = p_raw < P;
  const int64_t p = valid ? p_raw : (P - 1);
  const double* __restrict__ xp = x + p * ldx;
  const int d_in = (int)d.d_in;
  const int n_w = d.n_internal + 1;
  lds_f32* R = W + n_w * kTile;  // region R (16-byte aligned: n_w * 256 bytes)

  cptr<int32_t> hs_dim = C(d.hs_dim), hs_kmax = C(d.hs_kmax), hs_row = C(d.hs_row),
                int_rec = C(d.int_rec), level_off = C(d.level_off), wave_grp = C(d.wave_grp),
                grp_j = C(d.grp_j);
  cptr<double> dim_xi = C(d.dim_xi), dim_c = C(d.dim_c);

  // coordinate windows of the wave's first dims, requested before the table work
  const int g_begin = wave_grp[wave];
  const int g_end = wave_grp[wave + 1];
  int wj0 = (g_begin < g_end) ? (grp_j[g_begin] & ~(kXW - 1)) : 0;
  const int wlast = (g_end > g_begin) ? (grp_j[g_end - 1] & ~(kXW - 1)) : 0;
  XWin xa = load_win<kPairs>(xp, wj0, d_in);
  XWin xb = load_win<kPairs>(xp, min(wj0 + kXW, wlast), d_in);

  // 1a. h values of the internal nodes' factors: h_k(x_j) for k = 1..kmax(j), FP64 -> FP32
  double xh[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const int i = wave + m * kWaves;
    xh[m] = (i < d.n_hs_dims) ? xp[hs_dim[i]] : 0.0;
  }
  for (int i = wave; i < d.n_hs_dims; i += kWaves) {
    const int j = hs_dim[i], km = hs_kmax[i], row = hs_row[i];
    const int m = (i - wave) / kWaves;
    double xj = xh[0];
#pragma unroll
    for (int mm = 1; mm < 4; ++mm)
      if (m == mm) xj = xh[mm];
    if (m >= 4) xj = xp[j];
    cptr<double> xi = dim_xi + (int64_t)j * d.kmax1;
    cptr<double> cj = dim_c + (int64_t)j * d.kmax1;
    double pr = 1.0;
    for (int k = 1; k <= km; ++k) {
      pr *= (xj - xi[k - 1]);
      R[(row + k - 1) * kTile + lane] = (float)(pr * cj[k]);
    }
  }
  if (wave == 0) W[lane] = 1.0f;
  __syncthreads();
  // 1b. table of internal-node products, level by level: W[q] = W[parent] * h
  for (int lv = 0; lv < d.n_levels; ++lv) {
    const int a = level_off[lv], b = level_off[lv + 1];
    for (int q = a + wave; q < b; q += kWaves) {
      const int rec = int_rec[q];
      const int par = rec & 0xffff, hr = (rec >> 16) & 0xffff;
      W[(q + 1) * kTile + lane] = W[par * kTile + lane] * R[hr * kTile + lane];
    }
    __syncthreads();
  }

  // 2. stage node pairs {c_deg * surplus (FP32 bits), parent byte offset} and group records
  const int n_pad = d.n_nodes_pad;  // even
  lds_void* SQ = (lds_void*)R;
  lds_void* G = (lds_void*)((__attribute__((address_space(3))) char*)SQ + (size_t)n_pad * 8);
  stage_to_lds(d.node_sq, SQ, n_pad * 8, wave, lane);
  stage_to_lds(d.grp_rec, G, d.n_groups * 16, wave, lane);
  __builtin_amdgcn_s_waitcnt(0);  // this wave's copies have landed
  __syncthreads();
  const __attribute__((address_space(3))) v4i* GI = (const __attribute__((address_space(3))) v4i*)G;
  const __attribute__((address_space(3))) v4i* N2 = (const __attribute__((address_space(3))) v4i*)SQ;
  const __attribute__((address_space(3))) char* Wb = (const __attribute__((address_space(3))) char*)W;
  const int lane4 = lane * 4;

  // 3. streamed groups in increasing dim order, window by window (static window slots).  Two
  // coordinate windows alternate roles (A is consumed while B is in flight and vice versa), so
  // no register copy ever waits on a load issued for a later window.  The next group's record
  // is read while the current group is processed; nodes are handled in batches of 8.
  double acc = 0.0;
  int g = g_begin;
  // group record: {n0 | n1 << 16, dim, lo(a), hi(a)}, a = xi_{deg-1} (FP64)
  v4i ri = (g < g_end) ? GI[g] : v4i{0, -1, 0, 0};

  auto wld = [&](int off) { return *(const __attribute__((address_space(3))) float*)(Wb + off + lane4); };
  // after_first runs once the window's first dim is done: the refill of the other window is
  // issued there, so no wait at the start of a window covers a load issued just before it
  auto consume = [&](const XWin& xw, const int w0, auto&& after_first) {
#pragma unroll
    for (int u = 0; u < kXW; ++u) {
      if (u == 1) after_first();
      const int j = w0 + u;
      const double xj = xw.v[u];
      double pr = 1.0;
      while (g < g_end && ri.y == j) {
        const int gn = g + 1;
        const v4i ri_n = (gn < g_end) ? GI[gn] : v4i{0, -1, 0, 0};
        const double a = __hiloint2double(ri.w, ri.z);
        pr *= (xj - a);
        float t0 = 0.0f, t1 = 0.0f;
        // node pairs: {s_a, off_a, s_b, off_b}; groups start and end on even nodes
        const int e2 = ((unsigned)ri.x >> 16) >> 1;
        int i2 = (ri.x & 0xffff) >> 1;
        for (; i2 + 4 <= e2; i2 += 4) {
          const v4i p0 = N2[i2], p1 = N2[i2 + 1], p2 = N2[i2 + 2], p3 = N2[i2 + 3];
          t0 = fmaf(__int_as_float(p0.x), wld(p0.y), t0);
          t1 = fmaf(__int_as_float(p0.z), wld(p0.w), t1);
          t0 = fmaf(__int_as_float(p1.x), wld(p1.y), t0);
          t1 = fmaf(__int_as_float(p1.z), wld(p1.w), t1);
          t0 = fmaf(__int_as_float(p2.x), wld(p2.y), t0);
          t1 = fmaf(__int_as_float(p2.z), wld(p2.w), t1);
          t0 = fmaf(__int_as_float(p3.x), wld(p3.y), t0);
          t1 = fmaf(__int_as_float(p3.z), wld(p3.w), t1);
        }
        if (i2 + 2 <= e2) {  // tail: two pairs together, then at most one
          const v4i p0 = N2[i2], p1 = N2[i2 + 1];
          t0 = fmaf(__int_as_float(p0.x), wld(p0.y), t0);
          t1 = fmaf(__int_as_float(p0.z), wld(p0.w), t1);
          t0 = fmaf(__int_as_float(p1.x), wld(p1.y), t0);
          t1 = fmaf(__int_as_float(p1.z), wld(p1.w), t1);
          i2 += 2;
        }
        if (i2 < e2) {
          const v4i p0 = N2[i2];
          t0 = fmaf(__int_as_float(p0.x), wld(p0.y), t0);
          t1 = fmaf(__int_as_float(p0.z), wld(p0.w), t1);
        }
        acc = fma(pr, (double)(t0 + t1), acc);
        g = gn;
        // the record read at the start of this group goes to scalar registers only now, so its
        // LDS latency overlaps the group's work instead of stalling at the read
        int r0 = __builtin_amdgcn_readfirstlane(ri_n.x), r1 = __builtin_amdgcn_readfirstlane(ri_n.y);
        int r2 = __builtin_amdgcn_readfirstlane(ri_n.z), r3 = __builtin_amdgcn_readfirstlane(ri_n.w);
        asm volatile("" : "+s"(r0), "+s"(r1), "+s"(r2), "+s"(r3));
        ri = v4i{r0, r1, r2, r3};
      }
    }
  };

  // windows advance sequentially over the wave's dim range (a wave's dims are contiguous runs of
  // active dims; a window without groups costs one refill).  No other path loads a window, so
  // the two slots keep their registers and no LDS read waits on an outstanding coordinate load.
  // Refills past the wave's last window re-read that window (cache hits) instead of fetching the
  // next wave's coordinates.
  int wa = wj0, wb = wj0 + kXW;
  consume(xa, wa, [] {});
  while (g < g_end) {
    consume(xb, wb, [&] {  // A was consumed: refill it two windows ahead, in place
      wa = wb + kXW;
      xa = load_win<kPairs>(xp, min(wa, wlast), d_in);
    });
    if (g >= g_end) break;
    consume(xa, wa, [&] {  // refill B, in place
      wb = wa + kXW;
      xb = load_win<kPairs>(xp, min(wb, wlast), d_in);
    });
  }

  // 4. deterministic cross-wave reduction in FP64 (W area reused)
  __syncthreads();
  lds_f64* Rd = (lds_f64*)lds_raw;
  Rd[wave * kTile + lane] = acc;
  __syncthreads();
  if (wave == 0 && valid) {
    double sacc = d.s_root[0];
    for (int w = 0; w < kWaves; ++w) sacc += Rd[w * kTile + lane];
    y[p] = sacc;
  }
}

// ------------------------------------------------------------------------------------------
// Matrix-core mixed-precision kernel (d_out = 1).  The rows of W (internal nodes and the root)
// are numbered by increasing cost, so a group (j, k) sums t = sum_q S_g[q] W[q] over a dense
// vector S_g of its parents' rows 0 .. len_g - 1 (zeros in the gaps) and needs no parent index.
// The groups' dense vectors, in tiles of 16 groups, are contracted with W on the FP32 matrix cores: per 4 rows, v_mfma_f32_16x16x4 takes A = the tile's surpluses (lane l: group
// l % 16, row l / 16) and B = W for 16 points (lane l: row l / 16, point l % 16), four times for
// the tile's 64 points.  The result T[group, point] (lane l, register r: group 4 (l / 16) + r,
// point l % 16 of the subset) is multiplied in FP64 by the Newton factor
// h(x) = prod_{r < k} (x_j - xi_r) of the group (c_k is folded into the surpluses) and summed.
// Every LDS read of the loop is a 256-byte lane-ordered vector: no broadcast reads.
//   LDS: [W: rw_pad rows x 64 floats, element (q, p) at ((q / 4) * 64 + p) * 4 + q % 4][region R]
//   region R, phase 1: h table of the internal nodes' factors
//   region R, phase 2: A steps (256 bytes each) | group descriptors (16 bytes) | work items
// 16 waves per 64-point tile (4 wave groups x 4 subsets of 16 points) take work items (step
// ranges of one tile; y is linear in the partial sums) balanced on the host.  (The operand roles
// are swapped in the loop: see step 3.)
// ------------------------------------------------------------------------------------------
constexpr int kMWaves = 16;   // waves per tile: kMGroups wave groups x 4 point subsets
constexpr int kHsBatch = 4;   // table dims per wave whose coordinates are requested together
constexpr int kLvBatch = 8;   // internal-node records per wave read together
constexpr int kMGroups = 4;   // wave groups (the host balances the work items over them)

__device__ __forceinline__ int widx(int q, int p) { return ((q >> 2) * kTile + p) * 4 + (q & 3); }

__global__ __launch_bounds__(kTile* kMWaves) void newton_mixed_mfma_kernel(
    sbi_newton_desc d, const double* __restrict__ x, int64_t P, int64_t ldx, double* __restrict__ y) {
  extern __shared__ double lds_raw[];
  lds_f32* W = (lds_f32*)lds_raw;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t tile0 = (int64_t)blockIdx.x * kTile;
  const int64_t p_raw = tile0 + lane;
  const bool valid = p_raw < P;
  const int64_t p = valid ? p_raw : (P - 1);
  const double* __restrict__ xp = x + p * ldx;
  const int rw_pad = (d.rw_n + 3) & ~3;
  lds_f32* R = W + rw_pad * kTile;  // region R (16-byte aligned: rw_pad * 256 bytes)

  cptr<int32_t> hs_dim = C(d.hs_dim), hs_kmax = C(d.hs_kmax), hs_row = C(d.hs_row),
                rint_rec = C(d.rint_rec), rint_h = C(d.rint_h), level_off = C(d.level_off),
                wave_item = C(d.mwave_item), mi_rec = C(d.mi_rec);
  cptr<double> dim_xi = C(d.dim_xi), dim_c = C(d.dim_c);

  // Work items (step 3): wave w = 4 g + s takes the items of wave group g for the 16 points
  // 16 s .. 16 s + 15 of the tile.  Lane l owns group l % 16 of an item's tile and the points
  // 16 s + 4 (l / 16) + r (r < 4).  Coordinates are read four items ahead of their use; the first
  // four items' are requested here, before the tables, from the descriptors in global memory.
  const int li = lane & 15, qd = lane >> 4;
  const int sub = wave & 3, wgrp = wave >> 2;
  const double* xr[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) xr[r] = x + min(tile0 + 16 * sub + 4 * qd + r, P - 1) * ldx;
  const int i_begin = wave_item[wgrp], i_end = wave_item[wgrp + 1];
  const v4i* __restrict__ gdesc = (const v4i*)d.mt_grp;
  struct Pre {
    int j, k;
    double a, xv[4];
  };
  auto fetch = [&](const v4i gdsc) {
    Pre q;
    q.j = gdsc.x;
    q.k = gdsc.y;
    q.a = __hiloint2double(gdsc.w, gdsc.z);
#pragma unroll
    for (int r = 0; r < 4; ++r) q.xv[r] = xr[r][q.j];
    return q;
  };
  // (past the group's items: item 0's descriptor, loaded but never used; mi_n >= 1)
  auto first = [&](int i) { return gdesc[mi_rec[4 * (i < i_end ? i : 0) + 3] * 16 + li]; };
  // The internal-node records of this wave (record q = wave + 16 u in lane u) and the level
  // offsets (lane l), by vector loads issued before anything else: the products of 1b then need
  // no scalar round trip per level (records beyond 64 per wave take the scalar path).
  const int nint = d.n_internal, nlev = d.n_levels;
  const bool rec_v_ok = nint <= kTile * kMWaves && nlev < kTile;
  const int qv = min(wave + kMWaves * lane, max(nint - 1, 0));
  const int rec_v = d.rint_rec[qv], hrow_v = d.rint_h[qv];
  const int lo_v = d.level_off[min(lane, nlev)];
  // The four descriptors first, then the coordinates of the first kHsBatch table dims of this
  // wave (1a), then the four items' coordinates: vector loads complete in order, so each wait
  // below covers exactly the loads it needs (every such load is unconditional, indices clamped).
  const v4i f1 = first(i_begin), f2 = first(i_begin + 1), f3 = first(i_begin + 2),
            f4 = first(i_begin + 3);
  const int nhs = d.n_hs_dims;
  int hj[kHsBatch];
  double hx[kHsBatch];
#pragma unroll
  for (int u = 0; u < kHsBatch; ++u) hj[u] = hs_dim[min(wave + u * kMWaves, max(nhs - 1, 0))];
#pragma unroll
  for (int u = 0; u < kHsBatch; ++u) hx[u] = xp[hj[u]];

  // 1a. h values of the internal nodes' factors (lane = point), FP64 -> FP32
  auto h_rows = [&](int i, int j, double xj) {
    const int km = hs_kmax[i], row = hs_row[i];
    cptr<double> xi = dim_xi + (int64_t)j * d.kmax1;
    cptr<double> cj = dim_c + (int64_t)j * d.kmax1;
    double pr = 1.0;
    for (int k = 1; k <= km; ++k) {
      pr *= (xj - xi[k - 1]);
      R[(row + k - 1) * kTile + lane] = (float)(pr * cj[k]);
    }
  };
#pragma unroll
  for (int u = 0; u < kHsBatch; ++u)
    if (wave + u * kMWaves < nhs) h_rows(wave + u * kMWaves, hj[u], hx[u]);
  for (int i = wave + kHsBatch * kMWaves; i < nhs; i += kMWaves) h_rows(i, hs_dim[i], xp[hs_dim[i]]);
  for (int q = wave; q < rw_pad; q += kMWaves)
    if (q == 0 || q >= d.rw_n) W[widx(q, lane)] = (q == 0) ? 1.0f : 0.0f;
  __syncthreads();
  // 1b. products of the internal nodes, depth by depth: W[row] = W[parent] * h; a wave's records
  // of a level are read kLvBatch at a time (one scalar round trip per batch)
  for (int lv = 0; rec_v_ok && lv < nlev; ++lv) {
    const int a = __builtin_amdgcn_readlane(lo_v, lv), b = __builtin_amdgcn_readlane(lo_v, lv + 1);
    // this wave's records q in [a, b), q = wave (mod 16): lanes u0 .. u1 - 1
    const int u0 = a <= wave ? 0 : (a - wave + kMWaves - 1) / kMWaves;
    const int u1 = b <= wave ? 0 : (b - wave + kMWaves - 1) / kMWaves;
    for (int u = u0; u < u1; ++u) {
      const uint32_t rec = (uint32_t)__builtin_amdgcn_readlane(rec_v, u);
      const int hr = __builtin_amdgcn_readlane(hrow_v, u);
      W[widx(rec & 0xffff, lane)] = W[widx(rec >> 16, lane)] * R[hr * kTile + lane];
    }
    __syncthreads();
  }
  for (int lv = 0; !rec_v_ok && lv < nlev; ++lv) {
    const int a = level_off[lv], b = level_off[lv + 1];
    for (int q0 = a + wave; q0 < b; q0 += kLvBatch * kMWaves) {
      uint32_t rec[kLvBatch];
      int hr[kLvBatch];
#pragma unroll
      for (int u = 0; u < kLvBatch; ++u) {
        const int q = min(q0 + u * kMWaves, b - 1);
        rec[u] = (uint32_t)rint_rec[q];
        hr[u] = rint_h[q];
      }
#pragma unroll
      for (int u = 0; u < kLvBatch; ++u)
        if (q0 + u * kMWaves < b)
          W[widx(rec[u] & 0xffff, lane)] = W[widx(rec[u] >> 16, lane)] * R[hr[u] * kTile + lane];
    }
    __syncthreads();
  }

  // 2. stage the A steps, the group descriptors and the work items (region R)
  typedef __attribute__((address_space(3))) char lds_char;
  lds_char* AS = (lds_char*)R;
  lds_char* GD = AS + (size_t)d.mt_steps * 256;
  lds_char* IT = GD + (size_t)d.mt_n * 256;
  lds_char* XI = IT + (size_t)d.mi_n * 16;  // dim_xi rows of dims < n_xi_dims
  auto stage = [&](const void* src, lds_char* dst, int nbytes) {
    for (int c = wave; c * 1024 < nbytes; c += kMWaves) {
      const int off = c * 1024 + lane * 16;
      if (off < nbytes)
        __builtin_amdgcn_global_load_lds((const char*)src + off, (lds_void*)(dst + c * 1024), 16, 0,
                                         0);
    }
  };
  stage(d.mt_a, AS, d.mt_steps * 256);
  stage(d.mt_grp, GD, d.mt_n * 256);
  stage(d.mi_rec, IT, d.mi_n * 16);
  stage(d.dim_xi, XI, d.n_xi_dims * d.kmax1 * 8);
  // the first four items' coordinates (after the tables: their burst of loads would hold up the
  // tables' scalar loads)
  Pre n1 = fetch(f1), n2 = fetch(f2), n3 = fetch(f3), n4 = fetch(f4);
  __builtin_amdgcn_s_waitcnt(0);  // this wave's copies have landed
  __syncthreads();
  const lds_f32* A = (const lds_f32*)AS;
  const __attribute__((address_space(3))) v4i* GI = (const __attribute__((address_space(3))) v4i*)GD;
  const __attribute__((address_space(3))) v4i* II = (const __attribute__((address_space(3))) v4i*)IT;

  // 3. work items.  v_mfma_f32_16x16x4 with A = W (lane l: point 16 s + l % 16 of subset s,
  // row l / 16 of the step) and B = the tile's surpluses (lane l: row l / 16, group l % 16): the
  // result T[point, group] sits in lane l, register r at point 16 s + 4 (l / 16) + r, group l % 16.
  // So each lane owns one group of the tile (its dim j) and 16 points: its coordinates are x_j of
  // 16 rows, and the 16 lanes of a lane row read 16 dims of one row (for the consecutive dims of
  // the prefix tiles, one 128-byte line).
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  // one work item; its coordinates are in `slot`, which is refilled with those of item it + 4
  // once they are used.  The loop below is unrolled over the four slots so that each keeps its
  // registers: a refill lands where it is read, never copied (a copy would wait for its loads).
  auto process = [&](int it, Pre& slot) {
    const v4i ir = II[it];
    const int a0 = __builtin_amdgcn_readfirstlane(ir.x), w0 = __builtin_amdgcn_readfirstlane(ir.y);
    const int ns = __builtin_amdgcn_readfirstlane(ir.z);
    const v4i gnext = GI[__builtin_amdgcn_readfirstlane(II[it + 4 < i_end ? it + 4 : it].w) * 16 + li];
    const Pre& cur = slot;
    v4f D = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    const int wb = sub * 64 + li * 4 + qd;
    for (int st = 0; st < ns; ++st)
      D = __builtin_amdgcn_mfma_f32_16x16x4f32(W[(w0 + st) * kTile * 4 + wb],
                                               A[(a0 + st) * kTile + lane], D, 0, 0, 0);
    // node values xi_1 .. of the group's dim from the on-chip copy (a global load here would
    // wait for every coordinate load in flight)
    const lds_f64* xi = (const lds_f64*)XI + cur.j * d.kmax1;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      double pr = cur.xv[r] - cur.a;
      for (int k = 1; k < cur.k; ++k) pr *= (cur.xv[r] - xi[k]);
      acc[r] = fma(pr, (double)D[r], acc[r]);
    }
    slot = fetch(gnext);  // unconditional (past the end: this item's dims again, never used), so
                          // that no branch hides the loads from the wait-count bookkeeping
  };
  for (int it = i_begin; it < i_end; it += 4) {
    process(it, n1);
    if (it + 1 >= i_end) break;
    process(it + 1, n2);
    if (it + 2 >= i_end) break;
    process(it + 2, n3);
    if (it + 3 >= i_end) break;
    process(it + 3, n4);
  }

  // 4. deterministic reduction in FP64: each point's sums over the 16 lanes of its lane row (one
  // group each) by a fixed butterfly, then over the wave groups in order
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int m = 8; m > 0; m >>= 1) acc[r] += __shfl_xor(acc[r], m, 16);
  __syncthreads();
  lds_f64* Rd = (lds_f64*)lds_raw;
  if (li == 0) {
#pragma unroll
    for (int r = 0; r < 4; ++r) Rd[wgrp * kTile + 16 * sub + 4 * qd + r] = acc[r];
  }
  __syncthreads();
  if (wave == 0 && valid) {
    double sacc = d.s_root[0];
    for (int w = 0; w < kMGroups; ++w) sacc += Rd[w * kTile + lane];
    y[p] = sacc;
  }
}

// ------------------------------------------------------------------------------------------
// FP64 form of the matrix-core kernel (sbi_newton_eval_f64 when the plan has mt_a64 and the FP64
// W table fits LDS): the same work items and tables, W and the h table in FP64, the A steps
// (c_k * surplus in FP64) read from global memory per item (they do not fit in LDS beside the
// FP64 W), v_mfma_f64_16x16x4, whose result holds point 16 s + l / 16 + 4 r of the subset in
// register r (the coordinates and the reduction use that order).
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kTile* kMWaves) void newton_f64_mfma_kernel(
    sbi_newton_desc d, const double* __restrict__ x, int64_t P, int64_t ldx, double* __restrict__ y) {
  extern __shared__ double lds_raw[];
  lds_f64* W = (lds_f64*)lds_raw;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t tile0 = (int64_t)blockIdx.x * kTile;
  const int64_t p_raw = tile0 + lane;
  const bool valid = p_raw < P;
  const int64_t p = valid ? p_raw : (P - 1);
  const double* __restrict__ xp = x + p * ldx;
  const int rw_pad = (d.rw_n + 3) & ~3;
  lds_f64* R = W + rw_pad * kTile;  // region R (16-byte aligned: rw_pad * 512 bytes)

  cptr<int32_t> hs_dim = C(d.hs_dim), hs_kmax = C(d.hs_kmax), hs_row = C(d.hs_row),
                rint_rec = C(d.rint_rec), rint_h = C(d.rint_h), level_off = C(d.level_off),
                wave_item = C(d.mwave_item), mi_rec = C(d.mi_rec);
  cptr<double> dim_xi = C(d.dim_xi), dim_c = C(d.dim_c);

  // Work items (step 3): wave w = 4 g + s takes the items of wave group g for the 16 points
  // 16 s .. 16 s + 15 of the tile.  Lane l owns group l % 16 of an item's tile and the points
  // 16 s + 4 (l / 16) + r (r < 4).  Coordinates are read four items ahead of their use; the first
  // four items' are requested here, before the tables, from the descriptors in global memory.
  const int li = lane & 15, qd = lane >> 4;
  const int sub = wave & 3, wgrp = wave >> 2;
  const double* xr[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) xr[r] = x + min(tile0 + 16 * sub + qd + 4 * r, P - 1) * ldx;
  const int i_begin = wave_item[wgrp], i_end = wave_item[wgrp + 1];
  const v4i* __restrict__ gdesc = (const v4i*)d.mt_grp;
  struct Pre {
    int j, k;
    double a, xv[4];
  };
  auto fetch = [&](const v4i gdsc) {
    Pre q;
    q.j = gdsc.x;
    q.k = gdsc.y;
    q.a = __hiloint2double(gdsc.w, gdsc.z);
#pragma unroll
    for (int r = 0; r < 4; ++r) q.xv[r] = xr[r][q.j];
    return q;
  };
  // (past the group's items: item 0's descriptor, loaded but never used; mi_n >= 1)
  auto first = [&](int i) { return gdesc[mi_rec[4 * (i < i_end ? i : 0) + 3] * 16 + li]; };
  // The internal-node records of this wave (record q = wave + 16 u in lane u) and the level
  // offsets (lane l), by vector loads issued before anything else: the products of 1b then need
  // no scalar round trip per level (records beyond 64 per wave take the scalar path).
  const int nint = d.n_internal, nlev = d.n_levels;
  const bool rec_v_ok = nint <= kTile * kMWaves && nlev < kTile;
  const int qv = min(wave + kMWaves * lane, max(nint - 1, 0));
  const int rec_v = d.rint_rec[qv], hrow_v = d.rint_h[qv];
  const int lo_v = d.level_off[min(lane, nlev)];
  // The four descriptors first, then the coordinates of the first kHsBatch table dims of this
  // wave (1a), then the four items' coordinates: vector loads complete in order, so each wait
  // below covers exactly the loads it needs (every such load is unconditional, indices clamped).
  const v4i f1 = first(i_begin), f2 = first(i_begin + 1), f3 = first(i_begin + 2),
            f4 = first(i_begin + 3);
  const int nhs = d.n_hs_dims;
  int hj[kHsBatch];
  double hx[kHsBatch];
#pragma unroll
  for (int u = 0; u < kHsBatch; ++u) hj[u] = hs_dim[min(wave + u * kMWaves, max(nhs - 1, 0))];
#pragma unroll
  for (int u = 0; u < kHsBatch; ++u) hx[u] = xp[hj[u]];

  // 1a. h values of the internal nodes' factors (lane = point), FP64
  auto h_rows = [&](int i, int j, double xj) {
    const int km = hs_kmax[i], row = hs_row[i];
    cptr<double> xi = dim_xi + (int64_t)j * d.kmax1;
    cptr<double> cj = dim_c + (int64_t)j * d.kmax1;
    double pr = 1.0;
    for (int k = 1; k <= km; ++k) {
      pr *= (xj - xi[k - 1]);
      R[(row + k - 1) * kTile + lane] = pr * cj[k];
    }
  };
#pragma unroll
  for (int u = 0; u < kHsBatch; ++u)
    if (wave + u * kMWaves < nhs) h_rows(wave + u * kMWaves, hj[u], hx[u]);
  for (int i = wave + kHsBatch * kMWaves; i < nhs; i += kMWaves) h_rows(i, hs_dim[i], xp[hs_dim[i]]);
  for (int q = wave; q < rw_pad; q += kMWaves)
    if (q == 0 || q >= d.rw_n) W[widx(q, lane)] = (q == 0) ? 1.0 : 0.0;
  __syncthreads();
  // 1b. products of the internal nodes, depth by depth: W[row] = W[parent] * h; a wave's records
  // of a level are read kLvBatch at a time (one scalar round trip per batch)
  for (int lv = 0; rec_v_ok && lv < nlev; ++lv) {
    const int a = __builtin_amdgcn_readlane(lo_v, lv), b = __builtin_amdgcn_readlane(lo_v, lv + 1);
    // this wave's records q in [a, b), q = wave (mod 16): lanes u0 .. u1 - 1
    const int u0 = a <= wave ? 0 : (a - wave + kMWaves - 1) / kMWaves;
    const int u1 = b <= wave ? 0 : (b - wave + kMWaves - 1) / kMWaves;
    for (int u = u0; u < u1; ++u) {
      const uint32_t rec = (uint32_t)__builtin_amdgcn_readlane(rec_v, u);
      const int hr = __builtin_amdgcn_readlane(hrow_v, u);
      W[widx(rec & 0xffff, lane)] = W[widx(rec >> 16, lane)] * R[hr * kTile + lane];
    }
    __syncthreads();
  }
  for (int lv = 0; !rec_v_ok && lv < nlev; ++lv) {
    const int a = level_off[lv], b = level_off[lv + 1];
    for (int q0 = a + wave; q0 < b; q0 += kLvBatch * kMWaves) {
      uint32_t rec[kLvBatch];
      int hr[kLvBatch];
#pragma unroll
      for (int u = 0; u < kLvBatch; ++u) {
        const int q = min(q0 + u * kMWaves, b - 1);
        rec[u] = (uint32_t)rint_rec[q];
        hr[u] = rint_h[q];
      }
#pragma unroll
      for (int u = 0; u < kLvBatch; ++u)
        if (q0 + u * kMWaves < b)
          W[widx(rec[u] & 0xffff, lane)] = W[widx(rec[u] >> 16, lane)] * R[hr[u] * kTile + lane];
    }
    __syncthreads();
  }

  // 2. stage the A steps, the group descriptors and the work items (region R)
  typedef __attribute__((address_space(3))) char lds_char;
  lds_char* GD = (lds_char*)R;  // (the A steps stay in global memory)
  lds_char* IT = GD + (size_t)d.mt_n * 256;
  lds_char* XI = IT + (size_t)d.mi_n * 16;  // dim_xi rows of dims < n_xi_dims
  auto stage = [&](const void* src, lds_char* dst, int nbytes) {
    for (int c = wave; c * 1024 < nbytes; c += kMWaves) {
      const int off = c * 1024 + lane * 16;
      if (off < nbytes)
        __builtin_amdgcn_global_load_lds((const char*)src + off, (lds_void*)(dst + c * 1024), 16, 0,
                                         0);
    }
  };
  stage(d.mt_grp, GD, d.mt_n * 256);
  stage(d.mi_rec, IT, d.mi_n * 16);
  stage(d.dim_xi, XI, d.n_xi_dims * d.kmax1 * 8);
  // the first four items' coordinates (after the tables: their burst of loads would hold up the
  // tables' scalar loads)
  Pre n1 = fetch(f1), n2 = fetch(f2), n3 = fetch(f3), n4 = fetch(f4);
  __builtin_amdgcn_s_waitcnt(0);  // this wave's copies have landed
  __syncthreads();
  const double* __restrict__ A64 = d.mt_a64;
  const __attribute__((address_space(3))) v4i* GI = (const __attribute__((address_space(3))) v4i*)GD;
  const __attribute__((address_space(3))) v4i* II = (const __attribute__((address_space(3))) v4i*)IT;

  // 3. work items.  v_mfma_f64_16x16x4 with A = W (lane l: point 16 s + l % 16 of subset s,
  // row l / 16 of the step) and B = the tile's surpluses (lane l: row l / 16, group l % 16): the
  // result T[point, group] sits in lane l, register r at point 16 s + l / 16 + 4 r, group l % 16.
  // So each lane owns one group of the tile (its dim j) and 16 points: its coordinates are x_j of
  // 16 rows, and the 16 lanes of a lane row read 16 dims of one row (for the consecutive dims of
  // the prefix tiles, one 128-byte line).
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  // one work item; its coordinates are in `slot`, which is refilled with those of item it + 4
  // once they are used.  The loop below is unrolled over the four slots so that each keeps its
  // registers: a refill lands where it is read, never copied (a copy would wait for its loads).
  auto process = [&](int it, Pre& slot) {
    const v4i ir = II[it];
    const int a0 = __builtin_amdgcn_readfirstlane(ir.x), w0 = __builtin_amdgcn_readfirstlane(ir.y);
    const int ns = __builtin_amdgcn_readfirstlane(ir.z);
    const v4i gnext = GI[__builtin_amdgcn_readfirstlane(II[it + 4 < i_end ? it + 4 : it].w) * 16 + li];
    const Pre& cur = slot;
    v4d D = v4d{0.0, 0.0, 0.0, 0.0};
    const int wb = sub * 64 + li * 4 + qd;
    // A steps from global memory, four at a time (clamped to the item's last step, unconditional
    // loads), then four FP64 MFMAs
    const double* __restrict__ Ai = A64 + (int64_t)a0 * kTile + lane;
    for (int st0 = 0; st0 < ns; st0 += 4) {
      double bv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) bv[u] = Ai[(int64_t)min(st0 + u, ns - 1) * kTile];
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (st0 + u < ns)
          D = __builtin_amdgcn_mfma_f64_16x16x4f64(W[(w0 + st0 + u) * kTile * 4 + wb], bv[u], D,
                                                   0, 0, 0);
    }
    // node values xi_1 .. of the group's dim from the on-chip copy (a global load here would
    // wait for every coordinate load in flight)
    const lds_f64* xi = (const lds_f64*)XI + cur.j * d.kmax1;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      double pr = cur.xv[r] - cur.a;
      for (int k = 1; k < cur.k; ++k) pr *= (cur.xv[r] - xi[k]);
      acc[r] = fma(pr, D[r], acc[r]);
    }
    slot = fetch(gnext);  // unconditional (past the end: this item's dims again, never used), so
                          // that no branch hides the loads from the wait-count bookkeeping
  };
  for (int it = i_begin; it < i_end; it += 4) {
    process(it, n1);
    if (it + 1 >= i_end) break;
    process(it + 1, n2);
    if (it + 2 >= i_end) break;
    process(it + 2, n3);
    if (it + 3 >= i_end) break;
    process(it + 3, n4);
  }

  // 4. deterministic reduction in FP64: each point's sums over the 16 lanes of its lane row (one
  // group each) by a fixed butterfly, then over the wave groups in order
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int m = 8; m > 0; m >>= 1) acc[r] += __shfl_xor(acc[r], m, 16);
  __syncthreads();
  lds_f64* Rd = (lds_f64*)lds_raw;
  if (li == 0) {
#pragma unroll
    for (int r = 0; r < 4; ++r) Rd[wgrp * kTile + 16 * sub + qd + 4 * r] = acc[r];
  }
  __syncthreads();
  if (wave == 0 && valid) {
    double sacc = d.s_root[0];
    for (int w = 0; w < kMGroups; ++w) sacc += Rd[w * kTile + lane];
    y[p] = sacc;
  }
}

}  // namespace

struct sbi_newton_plan {
  sbi_newton_desc d;
  size_t lds_f32, lds_f64, lds_staged, lds_mfma, lds_mfma64;
};

extern "C" int sbi_newton_plan_create(const sbi_newton_desc* desc, sbi_newton_plan** out) {
  if (!desc || !out) {
    sbi::set_error("sbi_newton_plan_create: null argument");
    return SBI_ERR_INVALID;
  }
  const sbi_newton_desc& d = *desc;
  if (d.d_in <= 0 || d.d_out < 1 || d.n_nodes < 0 || d.n_internal < 0 || d.n_groups < 0 ||
      !d.s_root || !d.wave_grp || !d.grp || !d.node_q || !d.node_s32 || !d.node_s64 ||
      d.kmax1 <= 0) {
    sbi::set_error("sbi_newton_plan_create: invalid descriptor");
    return SBI_ERR_INVALID;
  }
  if (d.n_split != kWaves) {
    sbi::set_error("sbi_newton_plan_create: n_split must be 16");
    return SBI_ERR_INVALID;
  }
  if (d.gemm_head < 0 || d.gemm_kq < 0 || d.gemm_kq % 16 != 0 ||
      (d.gemm_kq > 0 && !d.gemm_sq) ||
      (d.gemm_head > 0 && (!d.head_fac || !d.gemm_s64 || d.ld_s64 < d.d_out)) ||
      (d.gemm_head + d.gemm_kq > 0 && (d.gemm_head > d.basis_n ||
                                        d.gemm_kq < d.basis_n - d.gemm_head ||
                                        d.gemm_kq >= d.basis_n - d.gemm_head + 16))) {
    sbi::set_error("sbi_newton_plan_create: invalid mixed GEMM fields (gemm_kq must be basis_n - "
                   "gemm_head rounded up to a multiple of 16, with gemm_sq; gemm_head rows need "
                   "head_fac and gemm_s64)");
    return SBI_ERR_INVALID;
  }
  const size_t red = (size_t)kWaves * kTile * sizeof(double);
  const size_t t32 = (size_t)(d.n_internal + 1) * kTile * sizeof(float);
  const size_t t64 = (size_t)(d.n_internal + 1) * kTile * sizeof(double);
  sbi_newton_plan* p = new sbi_newton_plan();
  p->d = d;
  p->lds_f32 = t32 > red ? t32 : red;
  p->lds_f64 = t64 > red ? t64 : red;
  // staged kernel: W table + max(h table, surpluses + parent rows + group records)
  {
    const size_t w = (size_t)(d.n_internal + 1) * kTile * 4;
    const size_t hs = (size_t)d.n_hs_rows * kTile * 4;
    const size_t st = (size_t)d.n_nodes_pad * 8 + (size_t)d.n_groups * 16;  // pairs + records
    size_t tot = w + (hs > st ? hs : st);
    if (tot < red) tot = red;
    p->lds_staged = (d.node_sq && d.grp_rec && d.hs_dim && d.n_nodes_pad % 2 == 0)
                        ? tot
                        : (size_t)-1;
  }
  // matrix-core kernel: W table + max(h table, A steps + descriptors + work items), at least the
  // reduction's 16 x 4 x 64 doubles
  {
    const size_t w = (size_t)((d.rw_n + 3) & ~3) * kTile * 4;
    const size_t hs = (size_t)d.n_hs_rows * kTile * 4;
    const size_t st = (size_t)d.mt_steps * 256 + (size_t)d.mt_n * 256 + (size_t)d.mi_n * 16 +
                      (((size_t)d.n_xi_dims * d.kmax1 * 8 + 15) & ~(size_t)15);
    size_t tot = w + (hs > st ? hs : st);
    const size_t red = (size_t)kMWaves * 4 * kTile * sizeof(double);
    if (tot < red) tot = red;
    p->lds_mfma = (d.mt_a && d.mt_grp && d.mi_rec && d.mwave_item && d.rint_rec && d.rint_h &&
                   d.hs_dim && d.rw_n == d.n_internal + 1 && d.mt_n > 0 && d.mi_n > 0 &&
                   d.n_xi_dims >= 0 && d.n_xi_dims <= d.d_in)
                      ? tot
                      : (size_t)-1;
    // its FP64 form: W and the h table in FP64, the A steps in global memory
    const size_t w64 = 2 * w, hs64 = 2 * hs;
    const size_t st64 = (size_t)d.mt_n * 256 + (size_t)d.mi_n * 16 +
                        (((size_t)d.n_xi_dims * d.kmax1 * 8 + 15) & ~(size_t)15);
    size_t tot64 = w64 + (hs64 > st64 ? hs64 : st64);
    if (tot64 < red) tot64 = red;
    p->lds_mfma64 = (p->lds_mfma != (size_t)-1 && d.mt_a64) ? tot64 : (size_t)-1;
  }
  // a node table beyond LDS leaves the plan to the basis, Jacobian and adjoint entry points: the
  // streaming kernels then return SBI_ERR_CAPACITY (the caller takes the GEMM form)
  *out = p;
  return SBI_OK;
}

extern "C" int sbi_newton_mixed_staged(const sbi_newton_plan* plan) {
  if (!plan) {
    sbi::set_error("sbi_newton_mixed_staged: null plan");
    return SBI_ERR_INVALID;
  }
  if (plan->lds_mfma <= (size_t)kMaxLds) return 2;
  return plan->lds_staged <= (size_t)kMaxLds ? 1 : 0;
}

extern "C" int sbi_newton_f64_matrix_cores(const sbi_newton_plan* plan) {
  if (!plan) {
    sbi::set_error("sbi_newton_f64_matrix_cores: null plan");
    return SBI_ERR_INVALID;
  }
  return plan->lds_mfma64 <= (size_t)kMaxLds ? 1 : 0;
}

extern "C" int sbi_newton_plan_desc(const sbi_newton_plan* plan, const sbi_newton_desc** out) {
  if (!plan || !out) {
    sbi::set_error("sbi_newton_plan_desc: null argument");
    return SBI_ERR_INVALID;
  }
  *out = &plan->d;
  return SBI_OK;
}

extern "C" int sbi_newton_plan_destroy(sbi_newton_plan* plan) {
  delete plan;
  return SBI_OK;
}

template <typename T>
static int newton_launch(const sbi_newton_plan* plan, const double* x, int64_t P, int64_t ldx,
                         double* y, void* stream) {
  if (!plan || !x || !y || P < 0 || ldx < plan->d.d_in || plan->d.d_out != 1) {
    sbi::set_error("sbi_newton_eval: invalid argument (the streaming kernels take d_out == 1)");
    return SBI_ERR_INVALID;
  }
  const size_t lds = sizeof(T) == 4 ? plan->lds_f32 : plan->lds_f64;
  if (lds > (size_t)kMaxLds) {
    sbi::set_error(sizeof(T) == 4 ? "sbi_newton_eval: the internal-node table exceeds LDS"
                                  : "sbi_newton_eval: the FP64 internal-node table exceeds LDS");
    return SBI_ERR_CAPACITY;
  }
  if (P == 0) return SBI_OK;
  sbi::allow_lds((const void*)newton_eval_kernel<float>);
  sbi::allow_lds((const void*)newton_eval_kernel<double>);
  const sbi_newton_desc& d = plan->d;
  dim3 grid((unsigned)((P + kTile - 1) / kTile));
  hipLaunchKernelGGL(newton_eval_kernel<T>, grid, dim3(kTile * kWaves), lds, (hipStream_t)stream, d,
                     x, P, ldx, y);
  SBI_HIP_CHECK(hipGetLastError());
  return SBI_OK;
}

extern "C" int sbi_newton_eval_mixed(const sbi_newton_plan* plan, const double* x, int64_t P,
                                     int64_t ldx, double* y, void* stream) {
  if (!plan || !x || !y || P < 0 || ldx < plan->d.d_in || plan->d.d_out != 1) {
    sbi::set_error("sbi_newton_eval_mixed: invalid argument (the streaming kernel takes d_out == 1)");
    return SBI_ERR_INVALID;
  }
  if (plan->lds_mfma <= (size_t)kMaxLds) {
    if (P == 0) return SBI_OK;
    sbi::allow_lds((const void*)newton_mixed_mfma_kernel);
    hipLaunchKernelGGL(newton_mixed_mfma_kernel, dim3((unsigned)((P + kTile - 1) / kTile)),
                       dim3(kTile * kMWaves), plan->lds_mfma, (hipStream_t)stream, plan->d, x, P,
                       ldx, y);
    SBI_HIP_CHECK(hipGetLastError());
    return SBI_OK;
  }
  if (plan->lds_staged > (size_t)kMaxLds) return newton_launch<float>(plan, x, P, ldx, y, stream);
  if (P == 0) return SBI_OK;
  dim3 grid((unsigned)((P + kTile - 1) / kTile));
  // 16-byte coordinate loads need an even row stride, a 16-byte aligned x and an even d_in >= 2
  // (every pair {2i, 2i+1} with 2i < d_in then lies inside the row)
  const bool pairs = (ldx % 2 == 0) && ((uintptr_t)x % 16 == 0) && plan->d.d_in >= 2 &&
                     plan->d.d_in % 2 == 0;
  sbi::allow_lds((const void*)newton_mixed_staged_kernel<true>);
  sbi::allow_lds((const void*)newton_mixed_staged_kernel<false>);
  if (pairs)
    hipLaunchKernelGGL(newton_mixed_staged_kernel<true>, grid, dim3(kTile * kWaves),
                       plan->lds_staged, (hipStream_t)stream, plan->d, x, P, ldx, y);
  else
    hipLaunchKernelGGL(newton_mixed_staged_kernel<false>, grid, dim3(kTile * kWaves),
                       plan->lds_staged, (hipStream_t)stream, plan->d, x, P, ldx, y);
  SBI_HIP_CHECK(hipGetLastError());
  return SBI_OK;
}

extern "C" int sbi_newton_eval_f64(const sbi_newton_plan* plan, const double* x, int64_t P,
                                   int64_t ldx, double* y, void* stream) {
  if (plan && x && y && P >= 0 && ldx >= plan->d.d_in && plan->d.d_out == 1 &&
      plan->lds_mfma64 <= (size_t)kMaxLds) {
    if (P == 0) return SBI_OK;
    sbi::allow_lds((const void*)newton_f64_mfma_kernel);
    hipLaunchKernelGGL(newton_f64_mfma_kernel, dim3((unsigned)((P + kTile - 1) / kTile)),
                       dim3(kTile * kMWaves), plan->lds_mfma64, (hipStream_t)stream, plan->d, x,
                       P, ldx, y);
    SBI_HIP_CHECK(hipGetLastError());
    return SBI_OK;
  }
  return newton_launch<double>(plan, x, P, ldx, y, stream);
}
