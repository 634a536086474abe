// Barycentric combination-technique kernels (FP64) for gfx950.
//
// Math (reference: src/smolyax/barycentric.py:34-123, interpolation.py:264-304):
//   y(x) = offset + sum_{blocks b} zeta_b sum_{mu <= nu_b} F_b[mu] prod_i Lhat^{(i)}_{mu_i}(x_{s_i})
// where Lhat is the normalised 1D barycentric basis of a slot (dim, deg).  The normalised basis is
// evaluated in the division-free first (modified Lagrange) form
//   Lhat_mu(x) = w_mu * prod_{nu != mu} (x - xi_nu),   w_mu = 1 / prod_{nu != mu} (xi_mu - xi_nu),
// which equals b_mu / sum(b) of the reference up to rounding; when x hits a node exactly the
// result is the exact one-hot vector, as in barycentric.py:62-65.
//
// Layout and schedule (one tile = 64 points = one point per lane of a wave64):
//   * every block's factors are its active dims in ascending order; the last (highest-dim)
//     factor is "streamed", the others are "prefix" factors;
//   * the 1D bases of all prefix slots for the tile's 64 points are computed once into LDS
//     (row-major [table row][64 lanes], conflict-free ds_read_b64);
//   * blocks are sorted by their last slot, so each run of blocks shares one streamed basis that
//     is computed once per run into registers; all block metadata and F values are wave-uniform
//     and arrive through scalar loads, so 64 points share every F read;
//   * the 16 waves of a workgroup split the runs (cost-balanced on the host) and reduce their
//     partial sums through LDS in a fixed order (deterministic, no atomics).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/smolyax_amd.h"
#include "common.h"
#include "launch.h"

namespace sbi {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

}  // namespace sbi

namespace {

constexpr int kTile = 64;       // points per tile (one per lane)
constexpr int kWaves = 16;      // waves per workgroup sharing one tile
constexpr int kLMax = 12;       // max (deg + 1) of a streamed factor held in registers
constexpr int kMaxLds = 160 * 1024;

// Read-only tables are read through the constant address space: a load with a wave-uniform
// address then becomes a scalar (SMEM) load into SGPRs and the control flow that depends on it
// stays uniform.  LDS is addressed through address space 3 explicitly.
template <typename T>
using cptr = const __attribute__((address_space(4))) T*;
template <typename T>
__device__ __forceinline__ cptr<T> C(const T* p) {
  return (cptr<T>)(p);
}
typedef __attribute__((address_space(3))) double lds_f64;

// ------------------------------------------------------------------------------------------
// 1D basis helpers
// ------------------------------------------------------------------------------------------

// Streamed factor: first-form basis of degree `deg` (< kLMax) into registers L[0..kLMax).
// Entries beyond deg are zero.  Exact one-hot at nodes.
__device__ __forceinline__ void basis_regs(double x, cptr<double> xi, cptr<double> w, int deg,
                                           double (&L)[kLMax]) {
  double dlt[kLMax];
#pragma unroll
  for (int m = 0; m < kLMax; ++m) dlt[m] = (m <= deg) ? (x - xi[m]) : 1.0;
  double suf[kLMax];
  double s = 1.0;
#pragma unroll
  for (int m = kLMax - 1; m >= 0; --m) {
    suf[m] = s;
    s *= dlt[m];
  }
  double pre = 1.0;
#pragma unroll
  for (int m = 0; m < kLMax; ++m) {
    double v = (m <= deg) ? w[m] * pre * suf[m] : 0.0;
    if (m <= deg && dlt[m] == 0.0) v = 1.0;
    L[m] = v;
    pre *= dlt[m];
  }
}

// Prefix slot: first-form basis of degree `deg` (any) written to LDS rows row0..row0+deg for this
// lane.  Two passes: suffix products stored, then the forward pass finishes each value.
__device__ __forceinline__ void basis_lds(double x, cptr<double> xi, cptr<double> w, int deg,
                                          lds_f64* tab, int row0, int lane) {
  double s = 1.0;
  for (int m = deg; m >= 0; --m) {
    tab[(row0 + m) * kTile + lane] = s;
    s *= (x - xi[m]);
  }
  double pre = 1.0;
  int hit = -1;
  for (int m = 0; m <= deg; ++m) {
    const double dm = x - xi[m];
    if (dm == 0.0) hit = m;
    tab[(row0 + m) * kTile + lane] *= w[m] * pre;
    pre *= dm;
  }
  if (hit >= 0) {
    for (int m = 0; m <= deg; ++m) tab[(row0 + m) * kTile + lane] = (m == hit) ? 1.0 : 0.0;
  }
}

// ------------------------------------------------------------------------------------------
// Evaluation kernel
// ------------------------------------------------------------------------------------------

// Inner product over the streamed factor for OC outputs, DL = deg + 1 known at compile time:
//   out[o] = sum_{m < DL} F[m * d_out + o] * Ll[m]
template <int DL, int OC>
__device__ __forceinline__ void row_dot(cptr<double> Frow, int64_t d_out, int oc_n,
                                        const double (&Ll)[kLMax], double (&out)[OC]) {
  if constexpr (OC == 1) {
    // d_out == 1: the streamed entries are contiguous (wide scalar loads)
    double s = 0.0;
#pragma unroll
    for (int m = 0; m < DL; ++m) s = fma(Frow[m], Ll[m], s);
    out[0] = s;
  } else {
#pragma unroll
    for (int o = 0; o < OC; ++o) {
      double s = 0.0;
      if (o < oc_n) {
#pragma unroll
        for (int m = 0; m < DL; ++m) s = fma(Frow[m * d_out + o], Ll[m], s);
      }
      out[o] = s;
    }
  }
}

// Contraction of one block for a streamed degree known at compile time (DL = deg + 1).
// Prefix factors (up to three fast paths, any number generically) come from the LDS table.
template <int DL, int OC>
__device__ __forceinline__ void contract_block(const sbi_bary_desc& d, int b, cptr<double> Fb,
                                               int oc_n, uint32_t w1, uint32_t w2, uint32_t w3,
                                               const double (&Ll)[kLMax], const lds_f64* tab,
                                               int lane, double (&acc)[OC]) {
  const int64_t dout = (OC == 1) ? 1 : d.d_out;
  const int64_t rs = (int64_t)DL * dout;
  const int npre = w1 & 7;
  double inner[OC];
  if (npre == 0) {
    row_dot<DL, OC>(Fb, dout, oc_n, Ll, inner);
#pragma unroll
    for (int o = 0; o < OC; ++o) acc[o] += inner[o];
  } else if (npre == 1) {
    // two prefix entries per step: both F rows are requested before either is used (one scalar
    // load round trip per pair), the accumulation order is unchanged
    const int d1 = (w1 >> 8) & 31, r1 = w2 & 0xffff;
    int m1 = 0;
    for (; m1 + 1 <= d1; m1 += 2) {
      double inner2[OC];
      row_dot<DL, OC>(Fb + m1 * rs, dout, oc_n, Ll, inner);
      row_dot<DL, OC>(Fb + (m1 + 1) * rs, dout, oc_n, Ll, inner2);
      const double p1 = tab[(r1 + m1) * kTile + lane], q1 = tab[(r1 + m1 + 1) * kTile + lane];
#pragma unroll
      for (int o = 0; o < OC; ++o) acc[o] = fma(q1, inner2[o], fma(p1, inner[o], acc[o]));
    }
    if (m1 <= d1) {
      const double p1 = tab[(r1 + m1) * kTile + lane];
      row_dot<DL, OC>(Fb + m1 * rs, dout, oc_n, Ll, inner);
#pragma unroll
      for (int o = 0; o < OC; ++o) acc[o] = fma(p1, inner[o], acc[o]);
    }
  } else if (npre == 2) {
    const int d1 = (w1 >> 8) & 31, d2 = (w1 >> 13) & 31;
    const int r1 = w2 & 0xffff, r2 = w2 >> 16;
    cptr<double> Fe = Fb;
    for (int m1 = 0; m1 <= d1; ++m1) {
      const double p1 = tab[(r1 + m1) * kTile + lane];
      int m2 = 0;
      for (; m2 + 1 <= d2; m2 += 2, Fe += 2 * rs) {  // two entries per step, as above
        double inner2[OC];
        row_dot<DL, OC>(Fe, dout, oc_n, Ll, inner);
        row_dot<DL, OC>(Fe + rs, dout, oc_n, Ll, inner2);
        const double p2 = p1 * tab[(r2 + m2) * kTile + lane];
        const double q2 = p1 * tab[(r2 + m2 + 1) * kTile + lane];
#pragma unroll
        for (int o = 0; o < OC; ++o) acc[o] = fma(q2, inner2[o], fma(p2, inner[o], acc[o]));
      }
      if (m2 <= d2) {
        const double p2 = p1 * tab[(r2 + m2) * kTile + lane];
        row_dot<DL, OC>(Fe, dout, oc_n, Ll, inner);
#pragma unroll
        for (int o = 0; o < OC; ++o) acc[o] = fma(p2, inner[o], acc[o]);
        Fe += rs;
      }
    }
  } else if (npre == 3) {
    const int d1 = (w1 >> 8) & 31, d2 = (w1 >> 13) & 31, d3 = (w1 >> 18) & 31;
    const int r1 = w2 & 0xffff, r2 = w2 >> 16, r3 = w3 & 0xffff;
    cptr<double> Fe = Fb;
    for (int m1 = 0; m1 <= d1; ++m1) {
      const double p1 = tab[(r1 + m1) * kTile + lane];
      for (int m2 = 0; m2 <= d2; ++m2) {
        const double p2 = p1 * tab[(r2 + m2) * kTile + lane];
        for (int m3 = 0; m3 <= d3; ++m3, Fe += rs) {
          const double p3 = p2 * tab[(r3 + m3) * kTile + lane];
          row_dot<DL, OC>(Fe, dout, oc_n, Ll, inner);
#pragma unroll
          for (int o = 0; o < OC; ++o) acc[o] = fma(p3, inner[o], acc[o]);
        }
      }
    }
  } else {
    // generic: any number of prefix factors, product recomputed per prefix entry
    cptr<int32_t> blk_fac = C(d.blk_fac), fac_deg = C(d.fac_deg), fac_row = C(d.fac_row);
    const int f0 = blk_fac[b];
    const int np = blk_fac[b + 1] - f0;
    int64_t nent = 1;
    for (int i = 0; i < np; ++i) nent *= (fac_deg[f0 + i] + 1);
    for (int64_t e = 0; e < nent; ++e) {
      int64_t rem = e;
      double pw = 1.0;
      for (int i = np - 1; i >= 0; --i) {
        const int di1 = fac_deg[f0 + i] + 1;
        const int mi = (int)(rem % di1);
        rem /= di1;
        pw *= tab[(fac_row[f0 + i] + mi) * kTile + lane];
      }
      row_dot<DL, OC>(Fb + e * rs, dout, oc_n, Ll, inner);
#pragma unroll
      for (int o = 0; o < OC; ++o) acc[o] = fma(pw, inner[o], acc[o]);
    }
  }
}

template <int OC>
__device__ __forceinline__ void contract_dispatch(const sbi_bary_desc& d, int b, cptr<double> Fb,
                                                  int oc_n, uint32_t w1, uint32_t w2, uint32_t w3,
                                                  const double (&Ll)[kLMax], const lds_f64* tab,
                                                  int lane, double (&acc)[OC]) {
  switch ((w1 >> 3) & 31) {
    case 0: contract_block<1, OC>(d, b, Fb, oc_n, w1, w2, w3, Ll, tab, lane, acc); break;
    case 1: contract_block<2, OC>(d, b, Fb, oc_n, w1, w2, w3, Ll, tab, lane, acc); break;
    case 2: contract_block<3, OC>(d, b, Fb, oc_n, w1, w2, w3, Ll, tab, lane, acc); break;
    case 3: contract_block<4, OC>(d, b, Fb, oc_n, w1, w2, w3, Ll, tab, lane, acc); break;
    case 4: contract_block<5, OC>(d, b, Fb, oc_n, w1, w2, w3, Ll, tab, lane, acc); break;
    case 5: contract_block<6, OC>(d, b, Fb, oc_n, w1, w2, w3, Ll, tab, lane, acc); break;
    case 6: contract_block<7, OC>(d, b, Fb, oc_n, w1, w2, w3, Ll, tab, lane, acc); break;
    case 7: contract_block<8, OC>(d, b, Fb, oc_n, w1, w2, w3, Ll, tab, lane, acc); break;
    case 8: contract_block<9, OC>(d, b, Fb, oc_n, w1, w2, w3, Ll, tab, lane, acc); break;
    case 9: contract_block<10, OC>(d, b, Fb, oc_n, w1, w2, w3, Ll, tab, lane, acc); break;
    case 10: contract_block<11, OC>(d, b, Fb, oc_n, w1, w2, w3, Ll, tab, lane, acc); break;
    default: contract_block<12, OC>(d, b, Fb, oc_n, w1, w2, w3, Ll, tab, lane, acc); break;
  }
}

template <int OC>
__global__ __launch_bounds__(kTile* kWaves) void bary_eval_kernel(
    sbi_bary_desc d, const double* __restrict__ x, int64_t P, int64_t ldx,
    double* __restrict__ y) {
  extern __shared__ double lds_raw[];
  lds_f64* lds = (lds_f64*)lds_raw;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t p_raw = (int64_t)blockIdx.x * kTile + lane;
  const bool valid = p_raw < P;
  const int64_t p = valid ? p_raw : (P - 1);
  const int oc0 = blockIdx.y * OC;
  const int oc_n = (int)min<int64_t>(OC, d.d_out - oc0);
  const double* __restrict__ xp = x + p * ldx;
  cptr<int32_t> seg_off = C(d.seg_off), seg_slot = C(d.seg_slot), seg_b0 = C(d.seg_b0),
                seg_b1 = C(d.seg_b1), slot_dim = C(d.slot_dim), slot_deg = C(d.slot_deg),
                prefix_slot = C(d.prefix_slot), prefix_off = C(d.prefix_off);
  cptr<double> slot_nodes = C(d.slot_nodes), slot_w = C(d.slot_w), F = C(d.F);
  cptr<int32_t> desc = C(d.blk_desc);

  // segments of this wave; the first streamed coordinate is requested before the table work
  const int g_begin = seg_off[wave];
  const int g_end = seg_off[wave + 1];
  double x_next = 0.0;
  if (g_begin < g_end) {
    const int s0 = seg_slot[g_begin];
    if (s0 >= 0) x_next = xp[slot_dim[s0]];
  }

  // 1. prefix table (all waves cooperate)
  for (int s = wave; s < d.n_prefix_slots; s += kWaves) {
    const int slot = prefix_slot[s];
    basis_lds(xp[slot_dim[slot]], slot_nodes + (int64_t)slot * d.dmax1,
              slot_w + (int64_t)slot * d.dmax1, slot_deg[slot], lds, prefix_off[s], lane);
  }
  __syncthreads();

  double acc[OC];
#pragma unroll
  for (int o = 0; o < OC; ++o) acc[o] = 0.0;

  for (int g = g_begin; g < g_end; ++g) {
    const int ls = seg_slot[g];
    const double x_cur = x_next;
    if (g + 1 < g_end) {  // prefetch the next segment's streamed coordinate
      const int sn = seg_slot[g + 1];
      if (sn >= 0) x_next = xp[slot_dim[sn]];
    }
    double Ll[kLMax];
    if (ls >= 0) {
      basis_regs(x_cur, slot_nodes + (int64_t)ls * d.dmax1, slot_w + (int64_t)ls * d.dmax1,
                 slot_deg[ls], Ll);
    } else {
#pragma unroll
      for (int m = 0; m < kLMax; ++m) Ll[m] = (m == 0) ? 1.0 : 0.0;
    }
    const int b_end = seg_b1[g];
    for (int b = seg_b0[g]; b < b_end; ++b) {
      const int rx = desc[4 * b];
      const uint32_t ry = (uint32_t)desc[4 * b + 1], rz = (uint32_t)desc[4 * b + 2],
                     rw = (uint32_t)desc[4 * b + 3];
      cptr<double> Fb = F + (OC == 1 ? (int64_t)rx : (int64_t)rx * d.d_out + oc0);
      contract_dispatch<OC>(d, b, Fb, oc_n, ry, rz, rw, Ll, lds, lane, acc);
    }
  }

  // 2. deterministic cross-wave reduction (table area is reused)
  __syncthreads();
#pragma unroll
  for (int o = 0; o < OC; ++o) lds[(wave * OC + o) * kTile + lane] = acc[o];
  __syncthreads();
  if (wave == 0 && valid) {
#pragma unroll
    for (int o = 0; o < OC; ++o) {
      if (o < oc_n) {
        double s = d.offset[oc0 + o];
        for (int w = 0; w < kWaves; ++w) s += lds[(w * OC + o) * kTile + lane];
        y[p * d.d_out + oc0 + o] = s;
      }
    }
  }
}

// ------------------------------------------------------------------------------------------
// Record-program evaluation kernel (d_out = 1)
// ------------------------------------------------------------------------------------------
// A program segment covers blocks of one streamed slot, so the streamed basis Ll is fixed over
// it.  Its blocks are cut into records: a record fixes one value of every prefix factor but the
// last and two consecutive values of the last one, so it reads its table values with one
// ds_read_b64 per leading factor and one ds_read2st64_b64 for the pair, and multiplies 2 DL
// consecutive values of F.  The kernel accumulates V[m] += P * F[., m] over the records and adds
// V . Ll once per segment: one FMA per entry.  Records with the same number of rows are read U
// at a time, every scalar load of the chunk issued before the first is used, so a memory round
// trip serves U records (scalar loads complete out of order, so any wait drains them all).
// Blocks the records do not cover (more than three prefix factors, streamed sizes outside
// kProgDL) take the per-block path of bary_eval_kernel.

constexpr int kProgHdr = 8;     // int32 per program segment header
constexpr int kProgBasis = 8;   // doubles per segment: nodes[4], weights[4] of the streamed slot
constexpr int kProgSgpr = 48;   // SGPRs per record chunk for rows and F values
constexpr int kProgTab = 8;     // prefix slots per wave whose coordinates are requested together
constexpr int kXGroup = 4;      // segments per group of streamed coordinates requested together

template <int DL, int NP>
struct RecChunk {
  static constexpr int kPer = 4 * DL + NP;  // SGPRs per record
  static constexpr int U = kProgSgpr / kPer < 1 ? 1 : (kProgSgpr / kPer > 12 ? 12 : kProgSgpr / kPer);
  static constexpr int U2 = (U + 1) / 2;
};

// U records from record i on; TAIL: only records < n count (loads clamped to record n - 1).
template <int DL, int NP, int U, bool TAIL>
__device__ __forceinline__ void rec_chunk(int i, int n, cptr<int32_t> rows, cptr<double> Fs,
                                          const lds_f64* tab, int lane, double (&V)[DL]) {
  constexpr int FS = 2 * DL;
  int r[U][NP];
  double f[U][FS];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int k = TAIL ? min(i + u, n - 1) : i + u;
#pragma unroll
    for (int q = 0; q < NP; ++q) r[u][q] = rows[k * NP + q];
#pragma unroll
    for (int e = 0; e < FS; ++e) f[u][e] = Fs[(int64_t)k * FS + e];
  }
  double t[U][NP + 1];
#pragma unroll
  for (int u = 0; u < U; ++u) {
#pragma unroll
    for (int q = 0; q + 1 < NP; ++q) t[u][q] = tab[r[u][q] * kTile + lane];
    t[u][NP - 1] = tab[r[u][NP - 1] * kTile + lane];
    t[u][NP] = tab[(r[u][NP - 1] + 1) * kTile + lane];
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    if (!TAIL || i + u < n) {
      double p0 = t[u][NP - 1], p1 = t[u][NP];
      if constexpr (NP >= 2) {
        double pre = t[u][0];
        if constexpr (NP == 3) pre *= t[u][1];
        p0 *= pre;
        p1 *= pre;
      }
#pragma unroll
      for (int m = 0; m < DL; ++m) V[m] = fma(p0, f[u][m], V[m]);
#pragma unroll
      for (int m = 0; m < DL; ++m) V[m] = fma(p1, f[u][DL + m], V[m]);
    }
  }
}

template <int DL, int NP>
__device__ __forceinline__ void rec_run(int n, cptr<int32_t>& rows, cptr<double>& Fs,
                                        const lds_f64* tab, int lane, double (&V)[DL]) {
  using S = RecChunk<DL, NP>;
  int i = 0;
  for (; i + S::U <= n; i += S::U) rec_chunk<DL, NP, S::U, false>(i, n, rows, Fs, tab, lane, V);
  if (i < n) {
    if (n - i <= S::U2)
      rec_chunk<DL, NP, S::U2, true>(i, n, rows, Fs, tab, lane, V);
    else
      rec_chunk<DL, NP, S::U, true>(i, n, rows, Fs, tab, lane, V);
  }
  rows += n * NP;
  Fs += (int64_t)n * 2 * DL;
}

__device__ __forceinline__ double readlane_f64(double v, int l) {
  const int64_t b = __builtin_bit_cast(int64_t, v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __builtin_bit_cast(double, ((int64_t)hi << 32) | (uint32_t)lo);
}

// Streamed basis of DL entries from the segment's nodes and weights, held in lanes
// l0 + m (nodes) and l0 + 4 + m (weights) of bv: the arithmetic of basis_regs (same products in
// the same order), exact one-hot at a node.
template <int DL>
__device__ __forceinline__ void basis_small(double x, double bv, int l0, double (&L)[DL]) {
  double xi[DL], w[DL];
#pragma unroll
  for (int m = 0; m < DL; ++m) {
    xi[m] = readlane_f64(bv, l0 + m);
    w[m] = readlane_f64(bv, l0 + 4 + m);
  }
  double dlt[DL];
#pragma unroll
  for (int m = 0; m < DL; ++m) dlt[m] = x - xi[m];
  double suf[DL];
  double s = 1.0;
#pragma unroll
  for (int m = DL - 1; m >= 0; --m) {
    suf[m] = s;
    s *= dlt[m];
  }
  double pre = 1.0;
#pragma unroll
  for (int m = 0; m < DL; ++m) {
    double v = w[m] * pre * suf[m];
    if (dlt[m] == 0.0) v = 1.0;
    L[m] = v;
    pre *= dlt[m];
  }
}

// The records of one program segment (streamed size DL; counts n1 | n2 << 16 and n3 in the low
// half of h4): V over the records, then V . Ll with the basis from the segment's own nodes and
// weights (lanes l0.. of bv).  rows / Fs run on over the wave's segments.
template <int DL>
__device__ __forceinline__ double prog_records(uint32_t h3, uint32_t h4, double bv, int l0,
                                               double x_cur, cptr<int32_t>& rows,
                                               cptr<double>& Fs, const lds_f64* tab, int lane,
                                               double acc) {
  double V[DL];
#pragma unroll
  for (int m = 0; m < DL; ++m) V[m] = 0.0;
  rec_run<DL, 1>((int)(h3 & 0xffff), rows, Fs, tab, lane, V);
  rec_run<DL, 2>((int)(h3 >> 16), rows, Fs, tab, lane, V);
  rec_run<DL, 3>((int)(h4 & 0xffff), rows, Fs, tab, lane, V);
  double Lb[DL];
  basis_small<DL>(x_cur, bv, l0, Lb);
#pragma unroll
  for (int m = 0; m < DL; ++m) acc = fma(V[m], Lb[m], acc);
  return acc;
}

__global__ __launch_bounds__(kTile* kWaves) void bary_eval_prog_kernel(
    sbi_bary_desc d, const double* __restrict__ x, int64_t P, int64_t ldx,
    double* __restrict__ y) {
  extern __shared__ double lds_raw[];
  lds_f64* lds = (lds_f64*)lds_raw;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t p_raw = (int64_t)blockIdx.x * kTile + lane;
  const bool valid = p_raw < P;
  const int64_t p = valid ? p_raw : (P - 1);
  const double* __restrict__ xp = x + p * ldx;
  cptr<int32_t> slot_dim = C(d.slot_dim), slot_deg = C(d.slot_deg),
                prefix_slot = C(d.prefix_slot), prefix_off = C(d.prefix_off);
  cptr<double> slot_nodes = C(d.slot_nodes), slot_w = C(d.slot_w), F = C(d.F);
  cptr<int32_t> desc = C(d.blk_desc), pseg = C(d.pseg), prow = C(d.prow);
  cptr<double> pF = C(d.pF);

  const int g_begin = C(d.pseg_off)[wave];
  const int g_end = C(d.pseg_off)[wave + 1];
  // Segments are processed in groups of kXGroup.  Per-segment data arrive in vector registers,
  // lane-distributed (lane l: segment l / 8 of a group, word or double l % 8), through vector
  // loads, whose in-order counter lets every wait cover exactly what it needs: the headers of
  // the group after next, the basis nodes and weights of the next group and the next group's
  // streamed coordinates are requested at the start of a group, so each has a whole group of
  // segments to arrive.  Every such load is unconditional with a clamped index; a segment without
  // a streamed factor has dim -1 and reads dim 0, unused.  The wave's record streams start at
  // its first segment's offsets and run on from there.
  const int g_last = g_begin < g_end ? g_end - 1 : 0;
  const int* __restrict__ pseg_v = d.pseg;
  const double* __restrict__ pbasis_v = d.pbasis;
  const int lseg = (lane >> 3) & (kXGroup - 1), lword = lane & 7;
  int hv_cur = pseg_v[min(g_begin + lseg, g_last) * kProgHdr + lword];
  int hv_next = pseg_v[min(g_begin + kXGroup + lseg, g_last) * kProgHdr + lword];
  double bv_cur = pbasis_v[min(g_begin + lseg, g_last) * kProgBasis + lword];
  double xg[kXGroup];
#pragma unroll
  for (int k = 0; k < kXGroup; ++k)
    xg[k] = xp[max(pseg[min(g_begin + k, g_last) * kProgHdr + 7], 0)];
  cptr<int32_t> rows = prow + pseg[min(g_begin, g_last) * kProgHdr + 2];
  cptr<double> Fs = pF + pseg[min(g_begin, g_last) * kProgHdr + 1];

  // 1. prefix table (all waves cooperate; each wave requests the coordinates of up to kProgTab
  //    slots before computing any) and the two rows of ones after it
  const int nps = d.n_prefix_slots;
  for (int s0 = wave; s0 < nps; s0 += kProgTab * kWaves) {
    int sl[kProgTab];
    double xv[kProgTab];
#pragma unroll
    for (int u = 0; u < kProgTab; ++u) sl[u] = prefix_slot[min(s0 + u * kWaves, nps - 1)];
#pragma unroll
    for (int u = 0; u < kProgTab; ++u) xv[u] = xp[slot_dim[sl[u]]];
#pragma unroll
    for (int u = 0; u < kProgTab; ++u) {
      const int s = s0 + u * kWaves;
      if (s < nps)
        basis_lds(xv[u], slot_nodes + (int64_t)sl[u] * d.dmax1, slot_w + (int64_t)sl[u] * d.dmax1,
                  slot_deg[sl[u]], lds, prefix_off[s], lane);
    }
  }
  if (wave == kWaves - 1) {
    lds[d.ones_row * kTile + lane] = 1.0;
    lds[(d.ones_row + 1) * kTile + lane] = 1.0;
  }
  __syncthreads();

  double acc = 0.0;
  for (int g0 = g_begin; g0 < g_end; g0 += kXGroup) {
    const int hv_next2 = pseg_v[min(g0 + 2 * kXGroup + lseg, g_last) * kProgHdr + lword];
    const double bv_next = pbasis_v[min(g0 + kXGroup + lseg, g_last) * kProgBasis + lword];
    double xn[kXGroup];
#pragma unroll
    for (int k = 0; k < kXGroup; ++k)
      xn[k] = xp[max(__builtin_amdgcn_readlane(hv_next, 8 * k + 7), 0)];
    const int g1 = min(g0 + kXGroup, g_end);
    if (__builtin_amdgcn_readlane(hv_cur, 6) & 1) {
      // a group of kXGroup segments with one record each (DL = 2, one prefix row pair, nothing
      // on the per-block path): their records are consecutive, so one chunk serves the group;
      // the arithmetic is that of the per-segment path
      int r[kXGroup];
      double f[kXGroup][4];
#pragma unroll
      for (int k = 0; k < kXGroup; ++k) {
        r[k] = rows[k];
#pragma unroll
        for (int e = 0; e < 4; ++e) f[k][e] = Fs[4 * k + e];
      }
      double t0[kXGroup], t1[kXGroup];
#pragma unroll
      for (int k = 0; k < kXGroup; ++k) {
        t0[k] = lds[r[k] * kTile + lane];
        t1[k] = lds[(r[k] + 1) * kTile + lane];
      }
#pragma unroll
      for (int k = 0; k < kXGroup; ++k) {
        double V[2];
#pragma unroll
        for (int m = 0; m < 2; ++m) V[m] = fma(t0[k], f[k][m], 0.0);
#pragma unroll
        for (int m = 0; m < 2; ++m) V[m] = fma(t1[k], f[k][2 + m], V[m]);
        double Lb[2];
        basis_small<2>(xg[k], bv_cur, 8 * k, Lb);
#pragma unroll
        for (int m = 0; m < 2; ++m) acc = fma(V[m], Lb[m], acc);
      }
      rows += kXGroup;
      Fs += 4 * kXGroup;
    } else {
    for (int g = g0; g < g1; ++g) {
      const int k = g - g0;
      double x_cur = xg[0];
#pragma unroll
      for (int q = 1; q < kXGroup; ++q) x_cur = (k == q) ? xg[q] : x_cur;
      const uint32_t h3 = (uint32_t)__builtin_amdgcn_readlane(hv_cur, 8 * k + 3);
      const uint32_t h4 = (uint32_t)__builtin_amdgcn_readlane(hv_cur, 8 * k + 4);
      switch ((h4 >> 16) & 0xff) {
        case 2: acc = prog_records<2>(h3, h4, bv_cur, 8 * k, x_cur, rows, Fs, lds, lane, acc); break;
        case 3: acc = prog_records<3>(h3, h4, bv_cur, 8 * k, x_cur, rows, Fs, lds, lane, acc); break;
        case 4: acc = prog_records<4>(h3, h4, bv_cur, 8 * k, x_cur, rows, Fs, lds, lane, acc); break;
        default: break;
      }
      // blocks on the per-block path (the streamed basis again in the general form: same values)
      const int n_gen = (int)(h4 >> 24);
      if (n_gen > 0) {
        const int ls = __builtin_amdgcn_readlane(hv_cur, 8 * k);
        const int gb0 = __builtin_amdgcn_readlane(hv_cur, 8 * k + 5);
        double Ll[kLMax];
        if (ls >= 0) {
          basis_regs(x_cur, slot_nodes + (int64_t)ls * d.dmax1, slot_w + (int64_t)ls * d.dmax1,
                     slot_deg[ls], Ll);
        } else {
#pragma unroll
          for (int m = 0; m < kLMax; ++m) Ll[m] = (m == 0) ? 1.0 : 0.0;
        }
        double a1[1] = {acc};
        for (int b = gb0; b < gb0 + n_gen; ++b)
          contract_dispatch<1>(d, b, F + desc[4 * b], 1, (uint32_t)desc[4 * b + 1],
                               (uint32_t)desc[4 * b + 2], (uint32_t)desc[4 * b + 3], Ll, lds, lane,
                               a1);
        acc = a1[0];
      }
    }
    }
#pragma unroll
    for (int k = 0; k < kXGroup; ++k) xg[k] = xn[k];
    hv_cur = hv_next;
    hv_next = hv_next2;
    bv_cur = bv_next;
  }

  // 2. deterministic cross-wave reduction (table area is reused)
  __syncthreads();
  lds[wave * kTile + lane] = acc;
  __syncthreads();
  if (wave == 0 && valid) {
    double s = d.offset[0];
    for (int w = 0; w < kWaves; ++w) s += lds[w * kTile + lane];
    y[p] = s;
  }
}

// ------------------------------------------------------------------------------------------
// Jacobian kernel (parity path of `gradient`): one wave per (tile, dim j); loops over the blocks
// that contain j; every (p, o, j) entry is written by exactly one lane.
// ------------------------------------------------------------------------------------------
constexpr int kJMaxN = 8;    // max active dims per block supported by the jacobian kernel
constexpr int kJDm = 32;     // max (deg + 1) per factor in the jacobian kernel

// Basis (first form) L_mu = w_mu prod_{nu != mu} (x - xi_nu) and, for the differentiated factor,
// its derivative by the product rule (running pair (p, p') over the factors, O(deg^2) per point)
// into LDS rows [f][m][lane].  At x == xi_mu the basis is one-hot; the derivative is NaN where
// (x - xi_nu)^2 == 0 for some nu <= deg, the reference's singular mask (barycentric.py:152-154).
__device__ __forceinline__ void basis_and_deriv_lds(double x, const double* __restrict__ xi,
                                                    const double* __restrict__ w, int deg,
                                                    bool want_d, double* __restrict__ Lrow,
                                                    double* __restrict__ dLrow, int lane) {
  int hit = -1;
  bool sing = false;
  for (int m = 0; m <= deg; ++m) {
    const double dx = x - xi[m];
    if (dx == 0.0) hit = m;
    sing |= (dx * dx == 0.0);
  }
  for (int m = 0; m <= deg; ++m) {
    double v = w[m], dv = 0.0;
    for (int q = 0; q <= deg; ++q) {
      if (q == m) continue;
      const double a = x - xi[q];
      dv = dv * a + v;
      v *= a;
    }
    Lrow[m * kTile + lane] = (hit >= 0) ? (m == hit ? 1.0 : 0.0) : v;
    if (want_d) dLrow[m * kTile + lane] = sing ? NAN : dv;
  }
}

constexpr int kJO = 8;  // outputs accumulated in registers per pass of the jacobian kernel

// One workgroup = one dim j for 64 points (one per lane).  For each pass over kJO outputs, every
// block containing j contributes sum_e d(prod_i Lhat_i)/dx_j [e] * F_b[e, o]: the factors' bases
// (and the derivative of factor j) are computed per lane into LDS, entries are walked with an
// odometer (last factor innermost, the others' product hoisted), F is read through scalar loads
// (the same for all lanes) and the sums stay in registers until one write per output.
__global__ __launch_bounds__(64) void bary_jacobian_kernel(sbi_bary_desc d,
                                                           const double* __restrict__ x, int64_t P,
                                                           int64_t ldx, double* __restrict__ J) {
  // LDS (sized by the plan): max_n factor rows of dmax1 basis values per lane, then the
  // derivative row of factor j
  extern __shared__ double jl_raw[];
  const int dm = d.dmax1;
  double* Ls = jl_raw;
  double* dLs = jl_raw + (int64_t)d.max_n * dm * kTile;
  const int lane = threadIdx.x;
  const int j = blockIdx.y;
  const int64_t p_raw = (int64_t)blockIdx.x * kTile + lane;
  const bool valid = p_raw < P;
  const int64_t p = valid ? p_raw : (P - 1);
  const double* __restrict__ xp = x + p * ldx;
  const int64_t dout = d.d_out;
  cptr<int32_t> dim_blk = C(d.dim_blk), dbi = C(d.dim_blk_idx), dbp = C(d.dim_blk_pos),
                blk_fac = C(d.blk_fac), blast = C(d.blk_last_slot), fac_slot = C(d.fac_slot),
                slot_deg = C(d.slot_deg), slot_dim = C(d.slot_dim);
  cptr<int64_t> foff = C(d.blk_foff);
  cptr<double> F = C(d.F);
  const int q0 = dim_blk[j], q1 = dim_blk[j + 1];

  for (int64_t o0 = (int64_t)blockIdx.z * kJO; o0 < dout; o0 += (int64_t)gridDim.z * kJO) {
    const int no = (int)min((int64_t)kJO, dout - o0);
    double acc[kJO];
#pragma unroll
    for (int oi = 0; oi < kJO; ++oi) acc[oi] = 0.0;
    for (int q = q0; q < q1; ++q) {
      const int b = dbi[q], pos = dbp[q];
      const int f0 = blk_fac[b], npre = blk_fac[b + 1] - f0, ls = blast[b];
      const int nf = npre + (ls >= 0 ? 1 : 0);
      int deg1[kJMaxN], idx[kJMaxN];
#pragma unroll
      for (int i = 0; i < kJMaxN; ++i) {
        deg1[i] = 1;
        idx[i] = 0;
        if (i < nf) {
          const int s = (i < npre) ? fac_slot[f0 + i] : ls;
          const int dg = slot_deg[s];
          deg1[i] = dg + 1;
          basis_and_deriv_lds(xp[slot_dim[s]], d.slot_nodes + (int64_t)s * d.dmax1,
                              d.slot_w + (int64_t)s * d.dmax1, dg, i == pos,
                              Ls + i * dm * kTile, dLs, lane);
        }
      }
      const double* lrow = (nf - 1 == pos) ? dLs : Ls + (nf - 1) * dm * kTile;
      int nl = 1;
#pragma unroll
      for (int i = 0; i < kJMaxN; ++i)
        if (i == nf - 1) nl = deg1[i];
      int64_t e = foff[b];
      while (true) {
        double pre = 1.0;
#pragma unroll
        for (int i = 0; i < kJMaxN - 1; ++i)
          if (i < nf - 1)
            pre *= (i == pos) ? dLs[idx[i] * kTile + lane] : Ls[(i * dm + idx[i]) * kTile + lane];
        for (int m = 0; m < nl; ++m, ++e) {
          const double wgt = pre * lrow[m * kTile + lane];
          cptr<double> Fe = F + e * dout + o0;
#pragma unroll
          for (int oi = 0; oi < kJO; ++oi)
            if (oi < no) acc[oi] = fma(wgt, Fe[oi], acc[oi]);
        }
        bool carry = true;  // odometer over the factors before the last, last-but-one fastest
#pragma unroll
        for (int i = kJMaxN - 2; i >= 0; --i) {
          if (carry && i < nf - 1) {
            if (++idx[i] < deg1[i]) {
              carry = false;
            } else {
              idx[i] = 0;
            }
          }
        }
        if (carry) break;
      }
    }
    if (valid) {
#pragma unroll
      for (int oi = 0; oi < kJO; ++oi)
        if (oi < no) J[(p * dout + o0 + oi) * d.d_in + j] = acc[oi];
    }
  }
}

// ------------------------------------------------------------------------------------------
// Integral: Q[o] = offset[o] + sum_e ent_q[e] F[e, o], a GEMV over F (streamed once).
// A workgroup of 256 threads = kW outputs x (256 / kW) entry lanes takes one chunk of entries for
// one tile of kW outputs; the entry lanes' sums are added in LDS in a fixed tree order, the
// chunks' partial sums in chunk order by the final kernel: the result does not depend on timing.
// kW follows d_out (1 .. 64) so that small d_out still uses every lane and reads F coalesced.
// ------------------------------------------------------------------------------------------
constexpr int kIntThreads = 256;
constexpr int kIntMaxChunks = 1024;

template <int kW>
__global__ __launch_bounds__(kIntThreads) void bary_integral_partial_kernel(
    const double* __restrict__ F, const double* __restrict__ q, int64_t K, int64_t d_out,
    int64_t chunk, double* __restrict__ partial) {
  constexpr int kE = kIntThreads / kW;  // entry lanes
  __shared__ double red[kIntThreads];
  const int tx = threadIdx.x % kW, ty = threadIdx.x / kW;
  const int c = blockIdx.x;
  const int64_t o = (int64_t)blockIdx.y * kW + tx;
  const int64_t e0 = (int64_t)c * chunk, e1 = min(K, e0 + chunk);
  double acc = 0.0;
  if (o < d_out) {
    int64_t e = e0 + ty;
    for (; e + 7 * kE < e1; e += 8 * kE) {  // eight independent loads in flight
      double f[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) f[u] = F[(e + u * kE) * d_out + o];
#pragma unroll
      for (int u = 0; u < 8; ++u) acc = fma(q[e + u * kE], f[u], acc);
    }
    for (; e < e1; e += kE) acc = fma(q[e], F[e * d_out + o], acc);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int h = kE / 2; h > 0; h >>= 1) {  // fixed-order tree over the entry lanes
    if (ty < h) red[threadIdx.x] += red[threadIdx.x + h * kW];
    __syncthreads();
  }
  if (ty == 0 && o < d_out) partial[(int64_t)c * d_out + o] = red[tx];
}

// Q[o] = offset[o] + sum over chunks: 16 outputs x 16 chunk lanes per workgroup, lane l adding
// chunks l, l + 16, ... in order, then a fixed tree over the lanes.
constexpr int kFinO = 16;

__global__ __launch_bounds__(kIntThreads) void bary_integral_final_kernel(
    int64_t d_out, int nc, const double* __restrict__ offset, const double* __restrict__ partial,
    double* __restrict__ Q) {
  constexpr int kL = kIntThreads / kFinO;
  __shared__ double red[kIntThreads];
  const int tx = threadIdx.x % kFinO, ty = threadIdx.x / kFinO;
  const int64_t o = (int64_t)blockIdx.x * kFinO + tx;
  double s = 0.0;
  if (o < d_out)
    for (int c = ty; c < nc; c += kL) s += partial[(int64_t)c * d_out + o];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = kL / 2; h > 0; h >>= 1) {
    if (ty < h) red[threadIdx.x] += red[threadIdx.x + h * kFinO];
    __syncthreads();
  }
  if (ty == 0 && o < d_out) Q[o] = offset[o] + red[tx];
}

}  // namespace

// Device scratch of sbi_bary_integral_f64_noalloc: the chunks' partial sums (at most
// kIntScratchDoubles doubles = 4.5 MiB on every device the library runs on, held until the process
// ends).  Host code never passes `scratch` itself to a kernel: in host code the name is a host-side
// stand-in; the device address of each device's copy comes from hipGetSymbolAddress.
constexpr int64_t kIntScratchDoubles = SBI_INTEGRAL_SCRATCH_DOUBLES;
static __device__ __align__(256) double scratch[kIntScratchDoubles];

// ------------------------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------------------------
struct sbi_plan {
  int kind;  // 1 = barycentric
  sbi_bary_desc bary;
  size_t eval_lds_bytes;
  size_t prog_lds_bytes;  // record-program kernel (0: no program)
};

extern "C" {

const char* sbi_last_error(void) { return sbi::g_err.c_str(); }
const char* sbi_version(void) { return "smolyax_amd 0.1 (gfx950)"; }

int sbi_bary_plan_create(const sbi_bary_desc* desc, sbi_plan** out) {
  if (!desc || !out) {
    sbi::set_error("sbi_bary_plan_create: null argument");
    return SBI_ERR_INVALID;
  }
  const sbi_bary_desc& d = *desc;
  if (d.d_in <= 0 || d.d_out <= 0 || d.n_slots < 0 || d.n_blocks < 0 || d.n_runs < 0 ||
      d.dmax1 <= 0 || !d.F || !d.offset || !d.blk_desc || !d.seg_off || !d.run_start || !d.blk_fac ||
      !d.blk_foff) {
    sbi::set_error("sbi_bary_plan_create: invalid descriptor");
    return SBI_ERR_INVALID;
  }
  if (d.n_split != kWaves) {
    sbi::set_error("sbi_bary_plan_create: n_split must equal the waves per workgroup (16)");
    return SBI_ERR_INVALID;
  }
  if (d.dmax1 < kLMax) {
    sbi::set_error("sbi_bary_plan_create: slot rows must be padded to at least 12 entries");
    return SBI_ERR_INVALID;
  }
  const size_t table = (size_t)d.n_prefix_vals * kTile * sizeof(double);
  const size_t red = (size_t)kWaves * 4 * kTile * sizeof(double);
  const size_t lds = table > red ? table : red;
  // a plan whose prefix table exceeds LDS is still valid for the basis / jacobian / integral
  // entry points; sbi_bary_eval_f64 then reports SBI_ERR_CAPACITY
  size_t prog_lds = 0;
  if (d.pseg) {
    if (d.d_out != 1 || !d.pseg_off || !d.prow || !d.pF || !d.pbasis || d.n_pseg <= 0 ||
        d.ones_row != d.n_prefix_vals) {
      sbi::set_error("sbi_bary_plan_create: invalid record program (d_out must be 1 and the rows "
                     "of ones follow the prefix rows)");
      return SBI_ERR_INVALID;
    }
    const size_t ptab = (size_t)(d.ones_row + 2) * kTile * sizeof(double);
    const size_t pred = (size_t)kWaves * kTile * sizeof(double);
    prog_lds = ptab > pred ? ptab : pred;
    if (prog_lds > (size_t)kMaxLds) {
      sbi::set_error("sbi_bary_plan_create: record program table exceeds LDS (318 prefix rows)");
      return SBI_ERR_CAPACITY;
    }
  }
  sbi_plan* p = new sbi_plan();
  p->kind = 1;
  p->bary = d;
  p->eval_lds_bytes = lds;
  p->prog_lds_bytes = prog_lds;
  *out = p;
  return SBI_OK;
}

int sbi_bary_plan_desc(const sbi_plan* plan, const sbi_bary_desc** out) {
  if (!plan || !out || plan->kind != 1) {
    sbi::set_error("sbi_bary_plan_desc: null argument or not a barycentric plan");
    return SBI_ERR_INVALID;
  }
  *out = &plan->bary;
  return SBI_OK;
}

int sbi_plan_destroy(sbi_plan* plan) {
  delete plan;
  return SBI_OK;
}

}  // extern "C"

template <int OC>
static int launch_eval(const sbi_plan* plan, const double* x, int64_t P, int64_t ldx, double* y,
                       hipStream_t st) {
  const sbi_bary_desc& d = plan->bary;
  dim3 grid((unsigned)((P + kTile - 1) / kTile), (unsigned)((d.d_out + OC - 1) / OC));
  sbi::allow_lds((const void*)bary_eval_kernel<OC>);
  hipLaunchKernelGGL(bary_eval_kernel<OC>, grid, dim3(kTile * kWaves), plan->eval_lds_bytes, st, d,
                     x, P, ldx, y);
  SBI_HIP_CHECK(hipGetLastError());
  return SBI_OK;
}

extern "C" int sbi_bary_eval_f64(const sbi_plan* plan, const double* x, int64_t P, int64_t ldx,
                                 double* y, void* stream) {
  if (!plan || plan->kind != 1 || !x || !y || P < 0 || ldx < plan->bary.d_in) {
    sbi::set_error("sbi_bary_eval_f64: invalid argument");
    return SBI_ERR_INVALID;
  }
  if (plan->eval_lds_bytes > (size_t)kMaxLds) {
    sbi::set_error("sbi_bary_eval_f64: prefix basis table exceeds LDS (" +
                   std::to_string(plan->bary.n_prefix_vals) + " rows > 320)");
    return SBI_ERR_CAPACITY;
  }
  if (P == 0) return SBI_OK;
  hipStream_t st = (hipStream_t)stream;
  const int64_t dout = plan->bary.d_out;
  if (plan->prog_lds_bytes) {  // record program (d_out == 1)
    const sbi_bary_desc& d = plan->bary;
    sbi::allow_lds((const void*)bary_eval_prog_kernel);
    hipLaunchKernelGGL(bary_eval_prog_kernel, dim3((unsigned)((P + kTile - 1) / kTile)),
                       dim3(kTile * kWaves), plan->prog_lds_bytes, st, d, x, P, ldx, y);
    SBI_HIP_CHECK(hipGetLastError());
    return SBI_OK;
  }
  if (dout == 1) return launch_eval<1>(plan, x, P, ldx, y, st);  // unit-stride F rows
  if (dout == 2) return launch_eval<2>(plan, x, P, ldx, y, st);
  return launch_eval<4>(plan, x, P, ldx, y, st);
}

extern "C" int sbi_bary_jacobian_f64(const sbi_plan* plan, const double* x, int64_t P, int64_t ldx,
                                     double* J, void* stream) {
  if (!plan || plan->kind != 1 || !x || !J || P < 0 || ldx < plan->bary.d_in ||
      !plan->bary.blk_last_slot || !plan->bary.dim_blk) {
    sbi::set_error("sbi_bary_jacobian_f64: invalid argument");
    return SBI_ERR_INVALID;
  }
  if (P == 0) return SBI_OK;
  const sbi_bary_desc& d = plan->bary;
  // passes of kJO outputs are spread over grid.z until the launch has ~4096 workgroups
  const int64_t tiles = (P + kTile - 1) / kTile;
  const int64_t passes = (d.d_out + kJO - 1) / kJO;
  const int64_t base = tiles * d.d_in;
  const int64_t gz = std::max<int64_t>(1, std::min<int64_t>(passes, (4096 + base - 1) / base));
  if (tiles > 0x7fffffff || d.d_in > 65535) {
    sbi::set_error("sbi_bary_jacobian_f64: too many points or input dims for one launch");
    return SBI_ERR_INVALID;
  }
  dim3 grid((unsigned)tiles, (unsigned)d.d_in, (unsigned)gz);
  const size_t lds = (size_t)(d.max_n + 1) * d.dmax1 * kTile * sizeof(double);
  if (d.max_n > kJMaxN || d.dmax1 > kJDm || lds > (size_t)kMaxLds) {
    sbi::set_error("sbi_bary_jacobian_f64: supports at most 8 active dims, degree 31, and "
                   "(active dims + 1) * (max degree + 1) <= 320");
    return SBI_ERR_UNSUPPORTED;
  }
  sbi::allow_lds((const void*)bary_jacobian_kernel);
  hipLaunchKernelGGL(bary_jacobian_kernel, grid, dim3(64), lds, (hipStream_t)stream, d, x, P, ldx,
                     J);
  SBI_HIP_CHECK(hipGetLastError());
  return SBI_OK;
}

extern "C" int64_t sbi_bary_integral_workspace(const sbi_plan* plan) {
  if (!plan) {
    sbi::set_error("sbi_bary_integral_workspace: null plan");
    return -1;
  }
  return (int64_t)kIntMaxChunks * plan->bary.d_out;
}

template <int kW>
static void integral_launch(const sbi_bary_desc& d, int nc, int64_t chunk, double* partial,
                            hipStream_t st) {
  dim3 g((unsigned)nc, (unsigned)((d.d_out + kW - 1) / kW));
  hipLaunchKernelGGL(bary_integral_partial_kernel<kW>, g, dim3(kIntThreads), 0, st, d.F, d.ent_q,
                     d.k_raw, d.d_out, chunk, partial);
}

// Chunks of the integral's partial-sum pass for a plan: nc chunks of `chunk` entries, output tiles
// of w (the partial sums take nc * d_out doubles).
struct IntSplit {
  int w;
  int64_t nc, chunk, otiles;
};

static IntSplit integral_split(const sbi_bary_desc& d) {
  // output tile width: the smallest power of two >= d_out, at most 64 (4 entry lanes), or 256
  // (one entry lane: a workgroup reads 2 KB runs of each F row) from d_out = 1024 on, where F is
  // large enough to keep ~2048 such workgroups busy (F of 69 185 x 4 096: 75 % against 65 % of
  // HBM; F of 27 003 x 256: 33 against 25 us)
  IntSplit sp;
  const int wmax = d.d_out >= 1024 ? 256 : 64;
  int w = 1;
  while (w < d.d_out && w < wmax) w <<= 1;
  sp.w = w;
  sp.otiles = (d.d_out + w - 1) / w;
  // entry chunks: ~2048 workgroups in all, each chunk at least 4 entries per entry lane
  const int64_t lanes = kIntThreads / w;
  int64_t nc = std::max<int64_t>(1, std::min<int64_t>(kIntMaxChunks, (2048 + sp.otiles - 1) / sp.otiles));
  nc = std::max<int64_t>(1, std::min<int64_t>(nc, (d.k_raw + 4 * lanes - 1) / (4 * lanes)));
  sp.chunk = (d.k_raw + nc - 1) / nc;
  sp.nc = std::max<int64_t>(1, (d.k_raw + sp.chunk - 1) / std::max<int64_t>(sp.chunk, 1));
  return sp;
}

static int integral_run(const sbi_bary_desc& d, const IntSplit& sp, double* Q, double* partial,
                        hipStream_t st) {
  if (sp.otiles > 65535) {
    sbi::set_error("sbi_bary_integral_f64: d_out too large for one launch");
    return SBI_ERR_INVALID;
  }
  const int nc = (int)sp.nc;
  switch (sp.w) {
    case 1: integral_launch<1>(d, nc, sp.chunk, partial, st); break;
    case 2: integral_launch<2>(d, nc, sp.chunk, partial, st); break;
    case 4: integral_launch<4>(d, nc, sp.chunk, partial, st); break;
    case 8: integral_launch<8>(d, nc, sp.chunk, partial, st); break;
    case 16: integral_launch<16>(d, nc, sp.chunk, partial, st); break;
    case 32: integral_launch<32>(d, nc, sp.chunk, partial, st); break;
    case 64: integral_launch<64>(d, nc, sp.chunk, partial, st); break;
    case 128: integral_launch<128>(d, nc, sp.chunk, partial, st); break;
    default: integral_launch<256>(d, nc, sp.chunk, partial, st); break;
  }
  SBI_HIP_CHECK(hipGetLastError());
  dim3 g2((unsigned)((d.d_out + kFinO - 1) / kFinO));
  hipLaunchKernelGGL(bary_integral_final_kernel, g2, dim3(kIntThreads), 0, st, d.d_out, nc,
                     d.offset, partial, Q);
  SBI_HIP_CHECK(hipGetLastError());
  return SBI_OK;
}

extern "C" int sbi_bary_integral_f64(const sbi_plan* plan, double* Q, double* partial,
                                     void* stream) {
  if (!plan || plan->kind != 1 || !Q || !partial || !plan->bary.ent_q) {
    sbi::set_error("sbi_bary_integral_f64: invalid argument (or the plan has no ent_q)");
    return SBI_ERR_INVALID;
  }
  const sbi_bary_desc& d = plan->bary;
  return integral_run(d, integral_split(d), Q, partial, (hipStream_t)stream);
}

// Device address of this device's copy of `scratch`, looked up once per device (with that device
// current) and kept for the process.
static int scratch_address(double** out) {
  static std::mutex mu;
  static double* addr[64] = {};
  int dev = 0;
  SBI_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) {
    sbi::set_error("sbi_bary_integral_f64_noalloc: device index beyond 63");
    return SBI_ERR_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void* ptr = nullptr;
    SBI_HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) {
      sbi::set_error("sbi_bary_integral_f64_noalloc: no device address for the scratch array");
      return SBI_ERR_HIP;
    }
    addr[dev] = (double*)ptr;
  }
  *out = addr[dev];
  return SBI_OK;
}

extern "C" int sbi_bary_integral_f64_noalloc(const sbi_plan* plan, double* Q, void* stream) {
  if (!plan || plan->kind != 1 || !Q || !plan->bary.ent_q) {
    sbi::set_error("sbi_bary_integral_f64_noalloc: invalid argument (or the plan has no ent_q)");
    return SBI_ERR_INVALID;
  }
  const sbi_bary_desc& d = plan->bary;
  const IntSplit sp = integral_split(d);
  if (sp.nc * d.d_out > kIntScratchDoubles) {  // checked before any HIP call
    sbi::set_error("sbi_bary_integral_f64_noalloc: the partial sums (" +
                   std::to_string(sp.nc * d.d_out) + " doubles) exceed the " +
                   std::to_string(kIntScratchDoubles) +
                   "-double device scratch; use sbi_bary_integral_f64 with a workspace");
    return SBI_ERR_CAPACITY;
  }
  double* partial = nullptr;
  const int rc = scratch_address(&partial);
  if (rc != SBI_OK) return rc;
  // every partial sum the final pass reads is written by this call's first pass: nothing to clear
  return integral_run(d, sp, Q, partial, (hipStream_t)stream);
}
