// Coefficient-by-basis contraction on the matrix cores (gfx950 MFMA), for many outputs.
//
//   Y[p, j] = offset[j] + sum_k A[k, p] * B[k, j]        p < P, j < n, k < K
//
// A is the basis matrix written by the basis kernels (K x P, points contiguous: the Newton basis
// H^T or the barycentric entry weights W^T); B is the coefficient matrix (K x n, outputs
// contiguous: Newton surpluses or zeta-scaled node values).  Y is FP64, row stride ldy.
//
// Work split: a workgroup of 4 waves owns a 64-point x (32*FN)-output tile of Y; each wave a
// 32 x 16*FN sub-tile made of 2 x FN MFMA 16x16 fragments.  Every workgroup is independent (no
// inter-workgroup waits, no atomics), so any number of these launches may run concurrently on
// several streams, and each Y row is a function of its own column of A only, summed in a fixed
// order: a point's result does not depend on P, on its position in the batch or on the batch's
// other points (bit for bit).
//
// k order: MFMA 16x16x4 takes k = lane/16 from lane group g; within a 16-wide chunk of k, step s
// of the four MFMAs takes k = k0 + 4g + s, so each lane walks four consecutive rows of A and B.
// A and B are read with one dword (FP32) or qword (FP64) per lane; the 16 lanes of a group read
// 16 consecutive points (outputs) of one row: 64- (128-) byte segments.  Edges are handled by
// clamped loads (rows beyond P or n are computed and discarded) and, in the last chunk, by zeroing
// terms with k >= K.
//
// Accumulation: FP64 operands use the FP64 MFMA (v_mfma_f64_16x16x4), FP32 operands the FP32 MFMA
// (v_mfma_f32_16x16x4); either way the MFMA accumulators hold partial sums of kFlushK terms of k
// that are added into FP64 accumulators (in LDS).  FP32: the accumulation error stays that of a
// 256-term sum and the result's error is dominated by the FP32 rounding of the operands.  FP64:
// the rounding of a K-term sum grows with kFlushK + K / kFlushK instead of K.

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/smolyax_amd.h"
#include "common.h"
#include "gemm.h"

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int kWaves = 4;
constexpr int kBlockM = 64;   // points per workgroup
constexpr int kChunk = 16;    // k per main-loop iteration
constexpr int kFlushK = 256;  // FP32 partial sums are flushed into FP64 every kFlushK terms

template <typename T>
struct Acc;
template <>
struct Acc<float> {
  using type = v4f;
};
template <>
struct Acc<double> {
  using type = v4d;
};

__device__ __forceinline__ v4f mfma(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v4d mfma(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// FP64 head rows k < K64 of A64 (K64 x P) and B64 (K64 x n) into h with the FP64 MFMA (lane group
// g supplies row k0 + g of each 4-row step; lane li point ph[fm] and output jcol[fn]).  Four steps
// per iteration: all their loads are issued before their MFMAs (one memory round trip per 16
// rows instead of per 4), and the MFMAs run in row order, steps past K64 skipped.
template <int WM, int WN>
__device__ __forceinline__ void head_fp64(const double* __restrict__ A64, int64_t lda64,
                                          const double* __restrict__ B64, int64_t ldb64,
                                          int64_t K64, int g, const int64_t (&ph)[WM],
                                          const int64_t (&jcol)[WN], v4d (&h)[WM][WN]) {
  for (int64_t k0 = 0; k0 < K64; k0 += 16) {
    double ha[4][WM], hb[4][WN];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t k = k0 + 4 * u + g;
      const bool in = k < K64;
      const int64_t kc = in ? k : K64 - 1;
#pragma unroll
      for (int fm = 0; fm < WM; ++fm) {
        const double v = A64[kc * lda64 + ph[fm]];
        ha[u][fm] = in ? v : 0.0;
      }
#pragma unroll
      for (int fn = 0; fn < WN; ++fn) {
        const double v = B64[kc * ldb64 + jcol[fn]];
        hb[u][fn] = in ? v : 0.0;
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (k0 + 4 * u >= K64) break;
#pragma unroll
      for (int fm = 0; fm < WM; ++fm)
#pragma unroll
        for (int fn = 0; fn < WN; ++fn) h[fm][fn] = mfma(ha[u][fm], hb[u][fn], h[fm][fn]);
    }
  }
}

// Optional FP64 head (FP32 operands only): rows k < K64 of A64 (K64 x P) and B64 (K64 x n) are
// contracted first with the FP64 MFMA, straight into the FP64 accumulators.  Lane li supplies the
// A row of point 4*(li & 3) + (li >> 2) of its 16-point block, which puts the FP64 MFMA's result
// (row g + 4r in register r of lane group g) into the FP32 layout (row 4g + r).
template <typename T, int FN, bool kK4 = false>
__global__ __launch_bounds__(64 * kWaves) void contract_kernel(
    const T* __restrict__ A, int64_t lda, const T* __restrict__ B, int64_t ldb,
    const double* __restrict__ offset, double* __restrict__ Y, int64_t ldy, int64_t P, int64_t n,
    int64_t K, const double* __restrict__ A64, int64_t lda64, const double* __restrict__ B64,
    int64_t ldb64, int64_t K64) {
  using AccT = typename Acc<T>::type;
  constexpr bool kF32 = sizeof(T) == 4;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int li = lane & 15, g = lane >> 4;

  const int64_t m_base = (int64_t)blockIdx.x * kBlockM + wm * 32;
  const int64_t n_base = (int64_t)blockIdx.y * (32 * FN) + wn * (16 * FN);
  int64_t pcol[2], jcol[FN];
#pragma unroll
  for (int fm = 0; fm < 2; ++fm) pcol[fm] = min(m_base + fm * 16 + li, P - 1);
#pragma unroll
  for (int fn = 0; fn < FN; ++fn) jcol[fn] = min(n_base + fn * 16 + li, n - 1);

  AccT acc[2][FN];
  // FP64 sums of the partial tiles of kFlushK terms (FP32 or FP64 MFMA accumulators): in LDS (one
  // column per thread, no sharing), which keeps the register count low enough for two waves per
  // SIMD.  Two levels keep the FP64 rounding of a long K at that of kFlushK + K / kFlushK terms.
  __shared__ double acc64s[2 * FN * 4][64 * kWaves];
  auto acc64 = [&](int fm, int fn, int r) -> double& {
    return acc64s[(fm * FN + fn) * 4 + r][threadIdx.x];
  };
#pragma unroll
  for (int fm = 0; fm < 2; ++fm)
#pragma unroll
    for (int fn = 0; fn < FN; ++fn) {
      acc[fm][fn] = AccT{0, 0, 0, 0};
#pragma unroll
      for (int r = 0; r < 4; ++r) acc64(fm, fn, r) = 0.0;
    }

  if constexpr (kF32) {
    if (K64 > 0) {
      int64_t ph[2];
#pragma unroll
      for (int fm = 0; fm < 2; ++fm)
        ph[fm] = min(m_base + fm * 16 + 4 * (li & 3) + (li >> 2), P - 1);
      v4d h[2][FN];
#pragma unroll
      for (int fm = 0; fm < 2; ++fm)
#pragma unroll
        for (int fn = 0; fn < FN; ++fn) h[fm][fn] = v4d{0, 0, 0, 0};
      head_fp64<2, FN>(A64, lda64, B64, ldb64, K64, g, ph, jcol, h);
#pragma unroll
      for (int fm = 0; fm < 2; ++fm)
#pragma unroll
        for (int fn = 0; fn < FN; ++fn)
#pragma unroll
          for (int r = 0; r < 4; ++r) acc64(fm, fn, r) = h[fm][fn][r];
    }
  }

  const int64_t K16 = K & ~(int64_t)(kChunk - 1);
  T a[2][4], b[FN][4];
  typedef T v4t __attribute__((ext_vector_type(4)));
  auto load = [&](int64_t k0) {
    if constexpr (kK4) {
      // k-interleaved operands: the four k of a lane's steps are one 16-byte vector
      const int64_t kq = (k0 >> 2) + g;
#pragma unroll
      for (int fm = 0; fm < 2; ++fm) {
        const v4t v = *reinterpret_cast<const v4t*>(A + (kq * lda + pcol[fm]) * 4);
        a[fm][0] = v.x;
        a[fm][1] = v.y;
        a[fm][2] = v.z;
        a[fm][3] = v.w;
      }
#pragma unroll
      for (int fn = 0; fn < FN; ++fn) {
        const v4t v = *reinterpret_cast<const v4t*>(B + (kq * ldb + jcol[fn]) * 4);
        b[fn][0] = v.x;
        b[fn][1] = v.y;
        b[fn][2] = v.z;
        b[fn][3] = v.w;
      }
    } else {
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int64_t k = k0 + 4 * g + s;
        const T* __restrict__ ar = A + k * lda;
        const T* __restrict__ br = B + k * ldb;
#pragma unroll
        for (int fm = 0; fm < 2; ++fm) a[fm][s] = ar[pcol[fm]];
#pragma unroll
        for (int fn = 0; fn < FN; ++fn) b[fn][s] = br[jcol[fn]];
      }
    }
  };
  auto step = [&]() {
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int fm = 0; fm < 2; ++fm)
#pragma unroll
        for (int fn = 0; fn < FN; ++fn) acc[fm][fn] = mfma(a[fm][s], b[fn][s], acc[fm][fn]);
  };
  auto flush = [&]() {
#pragma unroll
    for (int fm = 0; fm < 2; ++fm)
#pragma unroll
      for (int fn = 0; fn < FN; ++fn) {
#pragma unroll
        for (int r = 0; r < 4; ++r) acc64(fm, fn, r) += (double)acc[fm][fn][r];
        acc[fm][fn] = AccT{0, 0, 0, 0};
      }
  };

  auto chunk = [&](int64_t k0) {
    T ca[2][4], cb[FN][4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
#pragma unroll
      for (int fm = 0; fm < 2; ++fm) ca[fm][s] = a[fm][s];
#pragma unroll
      for (int fn = 0; fn < FN; ++fn) cb[fn][s] = b[fn][s];
    }
    if (k0 + kChunk < K16) load(k0 + kChunk);  // next chunk in flight during this one
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int fm = 0; fm < 2; ++fm)
#pragma unroll
        for (int fn = 0; fn < FN; ++fn) acc[fm][fn] = mfma(ca[fm][s], cb[fn][s], acc[fm][fn]);
  };
  if (K16 > 0) {
    load(0);
    if constexpr (kK4 || !kF32) {
      // the chunks of one kFlushK run form an inner loop with the flush after it, so the MFMA
      // accumulators stay in the matrix-core registers across the run (a flush test inside the
      // chunk loop made the compiler copy them to and from ordinary registers at every chunk:
      // 4096 x 4096 outputs, K 10 016: FP64 7.42 -> 5.58 ms, FP32 k4 3.11 -> 2.90 ms).  The
      // plain FP32 layout keeps the flat loop (3.73 ms flat, 4.53 ms nested).  Same sums.
      for (int64_t kc = 0; kc < K16; kc += kFlushK) {
        const int64_t kce = min(K16, kc + (int64_t)kFlushK);
        for (int64_t k0 = kc; k0 < kce; k0 += kChunk) chunk(k0);
        if (kce % kFlushK == 0) flush();  // a run cut short by K16 is flushed after the tail
      }
    } else {
      for (int64_t k0 = 0; k0 < K16; k0 += kChunk) {
        chunk(k0);
        if (((k0 + kChunk) % kFlushK) == 0) flush();
      }
    }
  }
  if (!kK4 && K16 < K) {  // last partial chunk: rows k >= K contribute zero (K4: K % 16 == 0)
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int64_t k = K16 + 4 * g + s;
      const bool in = k < K;
      const int64_t kc = in ? k : K - 1;
#pragma unroll
      for (int fm = 0; fm < 2; ++fm) {
        const T v = A[kc * lda + pcol[fm]];
        a[fm][s] = in ? v : T(0);
      }
#pragma unroll
      for (int fn = 0; fn < FN; ++fn) {
        const T v = B[kc * ldb + jcol[fn]];
        b[fn][s] = in ? v : T(0);
      }
    }
    step();
  }
  flush();

#pragma unroll
  for (int fm = 0; fm < 2; ++fm)
#pragma unroll
    for (int fn = 0; fn < FN; ++fn) {
      const int64_t col = n_base + fn * 16 + li;
      if (col >= n) continue;
      const double off = offset ? offset[col] : 0.0;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        // result layout: FP32 16x16x4 holds row 4g + r in register r of lane group g, FP64
        // 16x16x4 holds row g + 4r (column = lane % 16 in both)
        const int64_t row = m_base + fm * 16 + (kF32 ? 4 * g + r : g + 4 * r);
        const double v = acc64(fm, fn, r);
        if (row < P) Y[row * ldy + col] = off + v;
      }
    }
}

// Outputs per workgroup tile (32 * FN): 128 for wide outputs, 64 when 128-wide tiles would leave
// fewer than two workgroups per CU (8192 points x 256 outputs: 0.337 ms for the mixed step
// against 0.371 ms); a tile's results do not depend on its width (same bits).
inline int64_t tile_n(int64_t P, int64_t n) {
  if (n <= 32) return 32;
  if (n <= 64) return 64;
  const int64_t gm = (P + kBlockM - 1) / kBlockM;
  return gm * ((n + 127) / 128) < 512 ? 64 : 128;
}

// ------------------------------------------------------------------------------------------
// Split-k form of contract_kernel for small batches: the kFlushK-chunks of k are spread over
// blockIdx.z (chunks [z * cps, (z + 1) * cps) per slice) and slice z = nsplit computes the FP64
// head; every chunk's partial tile (the MFMA accumulators of its kFlushK terms, exactly what
// contract_kernel adds into its FP64 sums at that flush) is written to `part`, the head to
// `head`, and contract_reduce_kernel adds them in contract_kernel's order: the same bits.
// ------------------------------------------------------------------------------------------
template <typename T, int FN, bool kK4>
__global__ __launch_bounds__(64 * kWaves) void contract_split_kernel(
    const T* __restrict__ A, int64_t lda, const T* __restrict__ B, int64_t ldb, int64_t P,
    int64_t n, int64_t K, const double* __restrict__ A64, int64_t lda64,
    const double* __restrict__ B64, int64_t ldb64, int64_t K64, int64_t cps, int nsplit,
    T* __restrict__ part, double* __restrict__ head) {
  using AccT = typename Acc<T>::type;  // partial tiles are stored as T (the accumulator's type)
  constexpr bool kF32 = sizeof(T) == 4;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int li = lane & 15, g = lane >> 4;
  const int64_t m_base = (int64_t)blockIdx.x * kBlockM + wm * 32;
  const int64_t n_base = (int64_t)blockIdx.y * (32 * FN) + wn * (16 * FN);
  const int64_t Pn = P * n;
  int64_t pcol[2], jcol[FN];
#pragma unroll
  for (int fm = 0; fm < 2; ++fm) pcol[fm] = min(m_base + fm * 16 + li, P - 1);
#pragma unroll
  for (int fn = 0; fn < FN; ++fn) jcol[fn] = min(n_base + fn * 16 + li, n - 1);
  // element (fm, fn, r) of this lane: row and column of Y (FP32 layout for FP32 operands)
  auto out_at = [&](int fm, int fn, int r, int64_t& row, int64_t& col) {
    row = m_base + fm * 16 + (kF32 ? 4 * g + r : g + 4 * r);
    col = n_base + fn * 16 + li;
  };

  if ((int)blockIdx.z == nsplit) {  // the FP64 head (FP32 operands only)
    if constexpr (kF32) {
      int64_t ph[2];
#pragma unroll
      for (int fm = 0; fm < 2; ++fm)
        ph[fm] = min(m_base + fm * 16 + 4 * (li & 3) + (li >> 2), P - 1);
      v4d h[2][FN];
#pragma unroll
      for (int fm = 0; fm < 2; ++fm)
#pragma unroll
        for (int fn = 0; fn < FN; ++fn) h[fm][fn] = v4d{0, 0, 0, 0};
      head_fp64<2, FN>(A64, lda64, B64, ldb64, K64, g, ph, jcol, h);
#pragma unroll
      for (int fm = 0; fm < 2; ++fm)
#pragma unroll
        for (int fn = 0; fn < FN; ++fn)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            int64_t row, col;
            out_at(fm, fn, r, row, col);
            if (row < P && col < n) head[row * n + col] = h[fm][fn][r];
          }
    }
    return;
  }

  const int64_t nch = (K + kFlushK - 1) / kFlushK;
  const int64_t c0 = (int64_t)blockIdx.z * cps, c1 = min(nch, c0 + cps);
  if (c0 >= c1) return;
  const int64_t K16 = K & ~(int64_t)(kChunk - 1);
  AccT acc[2][FN];
#pragma unroll
  for (int fm = 0; fm < 2; ++fm)
#pragma unroll
    for (int fn = 0; fn < FN; ++fn) acc[fm][fn] = AccT{0, 0, 0, 0};
  int64_t cur = c0;  // chunk whose partial tile is accumulating
  auto emit = [&]() {
#pragma unroll
    for (int fm = 0; fm < 2; ++fm)
#pragma unroll
      for (int fn = 0; fn < FN; ++fn) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int64_t row, col;
          out_at(fm, fn, r, row, col);
          if (row < P && col < n) part[cur * Pn + row * n + col] = acc[fm][fn][r];
        }
        acc[fm][fn] = AccT{0, 0, 0, 0};
      }
    ++cur;
  };

  T a[2][4], b[FN][4];
  typedef T v4t __attribute__((ext_vector_type(4)));
  auto load = [&](int64_t k0) {
    if constexpr (kK4) {
      const int64_t kq = (k0 >> 2) + g;
#pragma unroll
      for (int fm = 0; fm < 2; ++fm) {
        const v4t v = *reinterpret_cast<const v4t*>(A + (kq * lda + pcol[fm]) * 4);
        a[fm][0] = v.x;
        a[fm][1] = v.y;
        a[fm][2] = v.z;
        a[fm][3] = v.w;
      }
#pragma unroll
      for (int fn = 0; fn < FN; ++fn) {
        const v4t v = *reinterpret_cast<const v4t*>(B + (kq * ldb + jcol[fn]) * 4);
        b[fn][0] = v.x;
        b[fn][1] = v.y;
        b[fn][2] = v.z;
        b[fn][3] = v.w;
      }
    } else {
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int64_t k = k0 + 4 * g + s;
        const T* __restrict__ ar = A + k * lda;
        const T* __restrict__ br = B + k * ldb;
#pragma unroll
        for (int fm = 0; fm < 2; ++fm) a[fm][s] = ar[pcol[fm]];
#pragma unroll
        for (int fn = 0; fn < FN; ++fn) b[fn][s] = br[jcol[fn]];
      }
    }
  };

  const int64_t kb = c0 * kFlushK, ke = min(c1 * kFlushK, K16);
  if (kb < ke) {
    load(kb);
    // one inner loop per kFlushK chunk, emitted after it (accumulators stay in place)
    for (int64_t kc = kb; kc < ke; kc += kFlushK) {
      const int64_t kce = min(ke, kc + (int64_t)kFlushK);
      for (int64_t k0 = kc; k0 < kce; k0 += kChunk) {
        T ca[2][4], cb[FN][4];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
#pragma unroll
          for (int fm = 0; fm < 2; ++fm) ca[fm][s] = a[fm][s];
#pragma unroll
          for (int fn = 0; fn < FN; ++fn) cb[fn][s] = b[fn][s];
        }
        if (k0 + kChunk < ke) load(k0 + kChunk);
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
          for (int fm = 0; fm < 2; ++fm)
#pragma unroll
            for (int fn = 0; fn < FN; ++fn)
              acc[fm][fn] = mfma(ca[fm][s], cb[fn][s], acc[fm][fn]);
      }
      if (kce % kFlushK == 0) emit();
    }
  }
  if (!kK4 && c1 == nch && K16 < K) {  // last partial chunk of 16: rows k >= K contribute zero
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int64_t k = K16 + 4 * g + s;
      const bool in = k < K;
      const int64_t kc = in ? k : K - 1;
#pragma unroll
      for (int fm = 0; fm < 2; ++fm) {
        const T v = A[kc * lda + pcol[fm]];
        a[fm][s] = in ? v : T(0);
      }
#pragma unroll
      for (int fn = 0; fn < FN; ++fn) {
        const T v = B[kc * ldb + jcol[fn]];
        b[fn][s] = in ? v : T(0);
      }
    }
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int fm = 0; fm < 2; ++fm)
#pragma unroll
        for (int fn = 0; fn < FN; ++fn) acc[fm][fn] = mfma(a[fm][s], b[fn][s], acc[fm][fn]);
  }
  if (cur < c1) emit();  // the slice's last chunk when K ends inside it
}

// Y[p, j] = offset[j] + (((head + part_0) + part_1) + ...) (+ 0.0 when K is a multiple of
// kFlushK: contract_kernel's final flush then adds an empty partial tile), head = 0 without one.
template <typename E>
__global__ __launch_bounds__(256) void contract_reduce_kernel(const E* __restrict__ part,
                                                              int64_t nch,
                                                              const double* __restrict__ head,
                                                              const double* __restrict__ offset,
                                                              double* __restrict__ Y, int64_t ldy,
                                                              int64_t P, int64_t n, int64_t K) {
  const int64_t Pn = P * n;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < Pn;
       i += (int64_t)gridDim.x * 256) {
    const int64_t row = i / n, col = i - row * n;
    double v = head ? head[i] : 0.0;
    for (int64_t c = 0; c < nch; ++c) v += (double)part[c * Pn + i];
    if (K % kFlushK == 0) v += 0.0;
    const double off = offset ? offset[col] : 0.0;
    Y[row * ldy + col] = off + v;
  }
}

// Split plan for a contraction of P points x n outputs: the slices of k (0: no split) and chunks
// per slice.  A launch of fewer than kSplitMinGroups workgroups is split so that about
// kSplitTarget workgroups run (small batches: the k loop of one workgroup is the latency).
constexpr int64_t kSplitMinGroups = 256;
constexpr int64_t kSplitTarget = 512;
constexpr int64_t kSplitMaxBytes = int64_t(256) << 20;  // workspace of one split call
struct SplitPlan {
  int64_t nsplit = 0, cps = 0, nch = 0, part_bytes = 0, head_bytes = 0;
};
template <typename T>
SplitPlan split_plan(int64_t P, int64_t n, int64_t K, int64_t K64) {
  SplitPlan sp;
  const int64_t gm = (P + kBlockM - 1) / kBlockM;
  const int64_t tn = tile_n(P, n);
  const int64_t groups = gm * ((n + tn - 1) / tn);
  sp.nch = (K + kFlushK - 1) / kFlushK;
  if (P <= 0 || n <= 0 || groups >= kSplitMinGroups || (sp.nch < 2 && K64 == 0)) return sp;
  int64_t want = (kSplitTarget + groups - 1) / groups;
  if (want > sp.nch) want = sp.nch;
  if (want < 1) want = 1;
  sp.cps = sp.nch > 0 ? (sp.nch + want - 1) / want : 0;
  sp.nsplit = sp.cps > 0 ? (sp.nch + sp.cps - 1) / sp.cps : 0;
  if (sp.nsplit + (K64 > 0 ? 1 : 0) < 2) {  // one slice: nothing to gain
    sp = SplitPlan{};
    return sp;
  }
  sp.part_bytes = ((sp.nch * P * n * (int64_t)sizeof(T)) + 255) /
                  256 * 256;
  sp.head_bytes = K64 > 0 ? P * n * 8 : 0;
  if (sp.part_bytes + sp.head_bytes > kSplitMaxBytes || sp.nsplit >= 65535) sp = SplitPlan{};
  return sp;
}

template <typename T, bool kK4>
int contract_split_launch(const SplitPlan& sp, const T* A, int64_t lda, const T* B, int64_t ldb,
                          const double* offset, double* Y, int64_t ldy, int64_t P, int64_t n,
                          int64_t K, const double* A64, int64_t lda64, const double* B64,
                          int64_t ldb64, int64_t K64, void* ws, hipStream_t st) {
  using E = T;
  E* part = (E*)ws;
  double* head = K64 > 0 ? (double*)((char*)ws + sp.part_bytes) : nullptr;
  const dim3 block(64 * kWaves);
  const unsigned gm = (unsigned)((P + kBlockM - 1) / kBlockM);
  const unsigned gz = (unsigned)(sp.nsplit + (K64 > 0 ? 1 : 0));
  const int ns = (int)sp.nsplit;
  const int64_t tn = tile_n(P, n);
  if (tn == 32) {
    hipLaunchKernelGGL((contract_split_kernel<T, 1, kK4>), dim3(gm, (unsigned)((n + 31) / 32), gz),
                       block, 0, st, A, lda, B, ldb, P, n, K, A64, lda64, B64, ldb64, K64, sp.cps,
                       ns, part, head);
  } else if (tn == 64) {
    hipLaunchKernelGGL((contract_split_kernel<T, 2, kK4>), dim3(gm, (unsigned)((n + 63) / 64), gz),
                       block, 0, st, A, lda, B, ldb, P, n, K, A64, lda64, B64, ldb64, K64, sp.cps,
                       ns, part, head);
  } else {
    hipLaunchKernelGGL((contract_split_kernel<T, 4, kK4>),
                       dim3(gm, (unsigned)((n + 127) / 128), gz), block, 0, st, A, lda, B, ldb, P,
                       n, K, A64, lda64, B64, ldb64, K64, sp.cps, ns, part, head);
  }
  SBI_HIP_CHECK(hipGetLastError());
  const int64_t Pn = P * n;
  int64_t rg = (Pn + 255) / 256;
  if (rg > 4096) rg = 4096;
  hipLaunchKernelGGL((contract_reduce_kernel<E>), dim3((unsigned)rg), dim3(256), 0, st, part,
                     sp.nch, head, offset, Y, ldy, P, n, K);
  SBI_HIP_CHECK(hipGetLastError());
  return SBI_OK;
}

template <typename T, bool kK4 = false>
int contract_launch(const T* A, int64_t lda, const T* B, int64_t ldb, const double* offset,
                    double* Y, int64_t ldy, int64_t P, int64_t n, int64_t K, const double* A64,
                    int64_t lda64, const double* B64, int64_t ldb64, int64_t K64, hipStream_t st) {
  const dim3 block(64 * kWaves);
  const unsigned gm = (unsigned)((P + kBlockM - 1) / kBlockM);
  const int64_t tn = tile_n(P, n);
  if (tn == 32) {
    hipLaunchKernelGGL((contract_kernel<T, 1, kK4>), dim3(gm, (unsigned)((n + 31) / 32)), block, 0, st,
                       A, lda, B, ldb, offset, Y, ldy, P, n, K, A64, lda64, B64, ldb64, K64);
  } else if (tn == 64) {
    hipLaunchKernelGGL((contract_kernel<T, 2, kK4>), dim3(gm, (unsigned)((n + 63) / 64)), block, 0, st,
                       A, lda, B, ldb, offset, Y, ldy, P, n, K, A64, lda64, B64, ldb64, K64);
  } else {
    hipLaunchKernelGGL((contract_kernel<T, 4, kK4>), dim3(gm, (unsigned)((n + 127) / 128)), block, 0,
                       st, A, lda, B, ldb, offset, Y, ldy, P, n, K, A64, lda64, B64, ldb64, K64);
  }
  SBI_HIP_CHECK(hipGetLastError());
  return SBI_OK;
}

// ------------------------------------------------------------------------------------------
// BF16 contraction (gfx950 v_mfma_f32_16x16x32_bf16): Y = offset + A64^T B64 + A^T B with A and B
// in BF16, k-interleaved by 8 (element (k, p) of A at A[((k / 8) * lda + p) * 8 + k % 8], the same
// for B; K a multiple of 32, zero rows pad it), and an optional FP64 head of K64 rows in the plain
// layout (FP64 MFMA, as in contract_kernel).  BF16 products are exact in FP32; the FP32 partial
// sums are added into FP64 registers every kBfFlush k.
//
// Work split: a workgroup of 4 waves (2 x 2) owns a tile of TM = 32 * kBfWM points x TN =
// 32 * kBfWN outputs; each wave a sub-tile of kBfWM x kBfWN MFMA 16x16 fragments (4 x 4 for the
// largest grids, 4 x 2 for large ones, 2 x 2 otherwise).  Per 32-k step the workgroup
// copies the step's 16-byte A vectors (4 k-groups x TM points) and B vectors (4 x TN outputs)
// into LDS with direct-to-LDS loads, kBfBuf buffers in rotation, the copies issued two steps
// ahead and one barrier per step (a buffer is rewritten two steps after it was read, and every
// wave has passed the barrier of the step in between); each wave then reads its fragments with
// 16-byte LDS reads.  Lane group g = lane / 16 supplies k = 8g .. 8g + 7 of the step to both
// operands.  The steps of one kBfFlush chunk form an inner loop with the flush after it, so the
// FP32 accumulators stay in the matrix-core registers across the chunk (a flush test inside the
// step loop made the compiler copy them to and from ordinary registers at every step).
// Measured (same bits as each wave loading its own fragments from global memory): 4096 points x
// 4096 outputs, K 10 016: 0.42 ms against 0.66-0.71 ms (4 x 2 fragments; 0.32-0.35 ms with 4 x 4);
// 8192 x 256, K 5 024: 39 us against 66 us.
// Every workgroup is independent and each Y row depends only on its own column of A (fixed order).
// ------------------------------------------------------------------------------------------
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef __attribute__((address_space(3))) v8bf lds_v8bf;
typedef __attribute__((address_space(3))) void lds_void;
constexpr int kBfFlush = 1024;  // k per FP32 partial sum
constexpr int kBfBuf = 4;       // LDS buffers of one 32-k step each

// kBfWM x kBfWN fragments per wave (points, outputs); 2 x 2 waves per workgroup (the body of the
// kernels below; each has its own LDS buffers)
template <int kBfWM, int kBfWN>
__device__ __forceinline__ void contract_bf16_body(
    const uint4* __restrict__ A, int64_t lda, const uint4* __restrict__ B, int64_t ldb,
    int64_t K, const double* __restrict__ A64, int64_t lda64, const double* __restrict__ B64,
    int64_t ldb64, int64_t K64, const double* __restrict__ offset, double* __restrict__ Y,
    int64_t ldy, int64_t P, int64_t n) {
  constexpr int TM = 32 * kBfWM, TN = 32 * kBfWN;
  constexpr int kAB = 4 * TM * 16, kStep = kAB + 4 * TN * 16;  // bytes per step: A, then B
  constexpr int nA = kAB / 1024, nI = kStep / 1024, kPer = nI / 4;  // 1 KB copies
  static_assert(TM % 64 == 0 && TN % 64 == 0 && nI % 4 == 0 && kPer >= 2 && kPer <= 4,
                "tile shape");
  __shared__ __attribute__((aligned(16))) char smem[kBfBuf * kStep];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wm = wave & 1, wn = wave >> 1;
  const int li = lane & 15, g = lane >> 4;
  const int64_t m0 = (int64_t)blockIdx.x * TM, n0 = (int64_t)blockIdx.y * TN;
  const int64_t m_base = m0 + wm * (16 * kBfWM);
  const int64_t n_base = n0 + wn * (16 * kBfWN);
  int64_t jcol[kBfWN];
#pragma unroll
  for (int fn = 0; fn < kBfWN; ++fn) jcol[fn] = min(n_base + fn * 16 + li, n - 1);

  v4d acc64[kBfWM][kBfWN];
#pragma unroll
  for (int fm = 0; fm < kBfWM; ++fm)
#pragma unroll
    for (int fn = 0; fn < kBfWN; ++fn) acc64[fm][fn] = v4d{0, 0, 0, 0};

  if (K64 > 0) {  // FP64 head: lane li supplies point 4 * (li & 3) + (li >> 2) of its block, so
                  // the FP64 MFMA's rows land in the FP32 layout (row 4g + r in register r)
    int64_t ph[kBfWM];
#pragma unroll
    for (int fm = 0; fm < kBfWM; ++fm)
      ph[fm] = min(m_base + fm * 16 + 4 * (li & 3) + (li >> 2), P - 1);
    head_fp64<kBfWM, kBfWN>(A64, lda64, B64, ldb64, K64, g, ph, jcol, acc64);
  }

  v4f acc[kBfWM][kBfWN];
#pragma unroll
  for (int fm = 0; fm < kBfWM; ++fm)
#pragma unroll
    for (int fn = 0; fn < kBfWN; ++fn) acc[fm][fn] = v4f{0, 0, 0, 0};
  auto flush = [&]() {
#pragma unroll
    for (int fm = 0; fm < kBfWM; ++fm)
#pragma unroll
      for (int fn = 0; fn < kBfWN; ++fn) {
#pragma unroll
        for (int r = 0; r < 4; ++r) acc64[fm][fn][r] += (double)acc[fm][fn][r];
        acc[fm][fn] = v4f{0, 0, 0, 0};
      }
  };
  // copy c of a step (this wave: c = wave + 4 i): c < nA -> A vectors of k-group c / (TM / 64),
  // points 64 (c % (TM / 64)) .., else B vectors likewise; it lands at LDS byte c * 1024 +
  // 16 lane, i.e. [k-group][point] then [k-group][output].  Points and outputs past the edge
  // re-read the last one (their results are not stored).
  const uint4* src[kPer];
  int64_t sstep[kPer];
#pragma unroll
  for (int i = 0; i < kPer; ++i) {
    const int c = wave + 4 * i;
    if (c < nA) {
      const int kq = c / (TM / 64), part = c % (TM / 64);
      src[i] = A + (int64_t)kq * lda + min(m0 + part * 64 + lane, P - 1);
      sstep[i] = 4 * lda;
    } else {
      const int kq = (c - nA) / (TN / 64), part = (c - nA) % (TN / 64);
      src[i] = B + (int64_t)kq * ldb + min(n0 + part * 64 + lane, n - 1);
      sstep[i] = 4 * ldb;
    }
  }
  auto issue = [&](int s) {
    char* buf = smem + (s % kBfBuf) * kStep;
#pragma unroll
    for (int i = 0; i < kPer; ++i)
      __builtin_amdgcn_global_load_lds((const void*)(src[i] + (int64_t)s * sstep[i]),
                                       (lds_void*)(buf + (wave + 4 * i) * 1024), 16, 0, 0);
  };
  const int S = (int)(K / 32);
  if (S > 0) {
    issue(0);
    if (S > 1) issue(1);
    for (int sc = 0; sc < S; sc += kBfFlush / 32) {
      const int se = min(S, sc + kBfFlush / 32);
      for (int s = sc; s < se; ++s) {
        // this wave's copies of step s have landed once at most those of s + 1, s + 2 remain
        if (s + 2 < S) {
          issue(s + 2);
          if constexpr (kPer == 4) asm volatile("s_waitcnt vmcnt(8)" ::: "memory");
          else if constexpr (kPer == 3) asm volatile("s_waitcnt vmcnt(6)" ::: "memory");
          else asm volatile("s_waitcnt vmcnt(4)" ::: "memory");
        } else if (s + 1 < S) {
          if constexpr (kPer == 4) asm volatile("s_waitcnt vmcnt(4)" ::: "memory");
          else if constexpr (kPer == 3) asm volatile("s_waitcnt vmcnt(3)" ::: "memory");
          else asm volatile("s_waitcnt vmcnt(2)" ::: "memory");
        } else {
          asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        }
        asm volatile("s_barrier" ::: "memory");  // every wave's copies of step s have landed
        const lds_v8bf* LA = (const lds_v8bf*)(smem + (s % kBfBuf) * kStep);
        const lds_v8bf* LB = (const lds_v8bf*)(smem + (s % kBfBuf) * kStep + kAB);
        v8bf a[kBfWM], b[kBfWN];
#pragma unroll
        for (int fm = 0; fm < kBfWM; ++fm) a[fm] = LA[g * TM + wm * (16 * kBfWM) + fm * 16 + li];
#pragma unroll
        for (int fn = 0; fn < kBfWN; ++fn) b[fn] = LB[g * TN + wn * (16 * kBfWN) + fn * 16 + li];
#pragma unroll
        for (int fm = 0; fm < kBfWM; ++fm)
#pragma unroll
          for (int fn = 0; fn < kBfWN; ++fn)
            acc[fm][fn] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[fm], b[fn], acc[fm][fn], 0, 0,
                                                                  0);
      }
      flush();
    }
  }

#pragma unroll
  for (int fn = 0; fn < kBfWN; ++fn) {
    const int64_t col = n_base + fn * 16 + li;
    if (col >= n) continue;
    const double off = offset ? offset[col] : 0.0;
#pragma unroll
    for (int fm = 0; fm < kBfWM; ++fm)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = m_base + fm * 16 + 4 * g + r;
        if (row < P) Y[row * ldy + col] = off + acc64[fm][fn][r];
      }
  }
}

#define SBI_BF16_ARGS                                                                          \
  const uint4 *__restrict__ A, int64_t lda, const uint4 *__restrict__ B, int64_t ldb, int64_t K, \
      const double *__restrict__ A64, int64_t lda64, const double *__restrict__ B64,             \
      int64_t ldb64, int64_t K64, const double *__restrict__ offset, double *__restrict__ Y,     \
      int64_t ldy, int64_t P, int64_t n
template <int kBfWM, int kBfWN>
__global__ __launch_bounds__(256) void contract_bf16_kernel(SBI_BF16_ARGS) {
  contract_bf16_body<kBfWM, kBfWN>(A, lda, B, ldb, K, A64, lda64, B64, ldb64, K64, offset, Y, ldy,
                                   P, n);
}
// 4 x 4 fragments per wave (workgroup 128 x 128), asked for two waves per SIMD: 243 registers
// and no scratch (one wave per SIMD otherwise).  4096 points x 4096 outputs, K 10 016: 323-350 us
// against 416-485 us for 4 x 2 fragments; same bits.  (The request on the 4 x 2 and 2 x 2 forms
// measured no gain, so it is this kernel's alone.)
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void
contract_bf16_wide_kernel(SBI_BF16_ARGS) {
  contract_bf16_body<4, 4>(A, lda, B, ldb, K, A64, lda64, B64, ldb64, K64, offset, Y, ldy, P, n);
}
#undef SBI_BF16_ARGS

// ------------------------------------------------------------------------------------------
// Jacobian contraction: for each input dim z, J[p, c, z] = sum_{k in [zoff[z], zoff[z+1])}
// A[k, p] B[brow[k * brs], c] (B rows gathered per row of A).  A workgroup owns a
// 64-point x (32 * FN)-output tile for JG consecutive dims, streams their rows
// [zoff[z0], zoff[z0 + JG]) as one pipelined sequence of 16-row chunks and keeps JG accumulator
// tiles in registers: a chunk that spans a dim boundary is applied once per dim it touches, with
// the other rows masked to zero.  Each lane finally writes JG consecutive doubles
// J[p, c, z0 .. z0 + JG) (one 8 * JG-byte run) instead of JG scattered words.
// Workgroup order: the n-tiles of one (point tile, dim group) go to the same XCD (consecutive
// workgroups of one XCD), so its L2 serves their shared A rows; consecutive (point tile, dim
// group) pairs go to different XCDs, which spreads the heavy low dims.
// ------------------------------------------------------------------------------------------
// Tile shapes (dims per workgroup JG, 16-output fragments per wave FN): 8 x 1 for up to 512
// outputs (d_out 256: 1.31 ms against 1.43-1.48 ms for 4 x 2, 2 x 4, 8 x 2), 2 x 4 above (d_out
// 4096: 3.48 ms against 4.10 ms for 8 x 1, 3.72 ms for 4 x 2, 3.58 ms for 2 x 2).

template <int JG, int FN>
__global__ __launch_bounds__(64 * kWaves) void jac_contract_kernel(
    const double* __restrict__ A, int64_t lda, const double* __restrict__ B, int64_t ldb,
    const int32_t* __restrict__ brow, int64_t brs, const int32_t* __restrict__ zoff, int nz,
    double* __restrict__ J, int64_t P, int64_t n, int nt, int ntp, int npairs) {
  const int L = blockIdx.x;
  const int xcd = L & 7, idx = L >> 3;
  const int tn = idx % nt;
  const int pair = (idx / nt) * 8 + xcd;
  if (pair >= npairs) return;
  const int tp = pair % ntp, zg = pair / ntp;

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int li = lane & 15, g = lane >> 4;
  const int64_t m_base = (int64_t)tp * kBlockM + wm * 32;
  const int64_t n_base = (int64_t)tn * (32 * FN) + wn * (16 * FN);
  int64_t pcol[2], jcol[FN];
#pragma unroll
  for (int fm = 0; fm < 2; ++fm) pcol[fm] = min(m_base + fm * 16 + li, P - 1);
#pragma unroll
  for (int fn = 0; fn < FN; ++fn) jcol[fn] = min(n_base + fn * 16 + li, n - 1);

  const int z0 = zg * JG;
  int zb[JG + 1];  // row boundaries of the group's dims (uniform)
#pragma unroll
  for (int jj = 0; jj <= JG; ++jj) zb[jj] = zoff[min(z0 + jj, nz)];
  const int64_t R0 = zb[0], R1 = zb[JG];

  v4d acc[JG][2][FN];
#pragma unroll
  for (int jj = 0; jj < JG; ++jj)
#pragma unroll
    for (int fm = 0; fm < 2; ++fm)
#pragma unroll
      for (int fn = 0; fn < FN; ++fn) acc[jj][fm][fn] = v4d{0, 0, 0, 0};

  // rows k0 + 4g + s of a chunk, clamped into the stream (the extra rows are masked); the B row
  // indices are read one chunk ahead of the B rows themselves
  double a[2][4], b[FN][4];
  int bi[4];
  auto load_idx = [&](int64_t k0) {
#pragma unroll
    for (int s = 0; s < 4; ++s) bi[s] = brow[min(k0 + 4 * g + s, R1 - 1) * brs];
  };
  auto load = [&](int64_t k0) {
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int64_t k = min(k0 + 4 * g + s, R1 - 1);
      const double* __restrict__ ar = A + k * lda;
#pragma unroll
      for (int fm = 0; fm < 2; ++fm) a[fm][s] = ar[pcol[fm]];
      const double* __restrict__ br = B + (int64_t)bi[s] * ldb;
#pragma unroll
      for (int fn = 0; fn < FN; ++fn) b[fn][s] = br[jcol[fn]];
    }
  };
  if (R1 > R0) {
    load_idx(R0);
    load(R0);
    if (R0 + kChunk < R1) load_idx(R0 + kChunk);
  }
  double ca[2][4], cb[FN][4];
  auto advance = [&](int64_t k0) {  // this chunk's operands; the next chunk's loads in flight
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      ca[0][s] = a[0][s];
      ca[1][s] = a[1][s];
#pragma unroll
      for (int fn = 0; fn < FN; ++fn) cb[fn][s] = b[fn][s];
    }
    if (k0 + kChunk < R1) {
      load(k0 + kChunk);
      if (k0 + 2 * kChunk < R1) load_idx(k0 + 2 * kChunk);
    }
  };
  auto apply = [&](int64_t k0) {  // a chunk applied to every dim it touches
    const int64_t klast = min(k0 + kChunk, R1) - 1;
    // dims touched by this chunk: [first row's dim, last row's dim] (uniform)
#pragma unroll
    for (int jj = 0; jj < JG; ++jj) {
      if (zb[jj + 1] <= k0 || zb[jj] > klast || zb[jj] == zb[jj + 1]) continue;
      if (zb[jj] <= k0 && zb[jj + 1] >= k0 + kChunk) {  // all 16 rows belong to this dim
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
          for (int fm = 0; fm < 2; ++fm)
#pragma unroll
            for (int fn = 0; fn < FN; ++fn)
              acc[jj][fm][fn] = mfma(ca[fm][s], cb[fn][s], acc[jj][fm][fn]);
      } else {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int64_t k = k0 + 4 * g + s;
          const bool in = k >= zb[jj] && k < zb[jj + 1];
#pragma unroll
          for (int fn = 0; fn < FN; ++fn) {
            const double bv = in ? cb[fn][s] : 0.0;
#pragma unroll
            for (int fm = 0; fm < 2; ++fm) acc[jj][fm][fn] = mfma(ca[fm][s], bv, acc[jj][fm][fn]);
          }
        }
      }
    }
  };
  // The chunks in stream order: those wholly inside dim jj in a branch-free loop on acc[jj], then
  // the chunk that ends dim jj inside it (or the stream's last, partial chunk) through apply.  At
  // dim jj the next chunk starts at or after the dim's first row, so every chunk is applied once
  // and each dim's sum sees the same MFMAs in the same order as a single chunk loop with a per-dim
  // test (same bits; 320 points x 4096 outputs x 100 dims: 3.14-3.23 ms against 3.45-3.51 ms,
  // 2560 x 256 x 200: 1.15-1.23 ms against 1.20-1.26 ms).
  int64_t k0 = R0;
#pragma unroll
  for (int jj = 0; jj < JG; ++jj) {
    while (k0 < R1 && zb[jj] <= k0 && k0 + kChunk <= zb[jj + 1]) {
      advance(k0);
#pragma unroll
      for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int fm = 0; fm < 2; ++fm)
#pragma unroll
          for (int fn = 0; fn < FN; ++fn)
            acc[jj][fm][fn] = mfma(ca[fm][s], cb[fn][s], acc[jj][fm][fn]);
      k0 += kChunk;
    }
    if (k0 < R1 && k0 < zb[jj + 1]) {
      advance(k0);
      apply(k0);
      k0 += kChunk;
    }
  }

  // FP64 16x16x4 result: row (point) g + 4r in register r, column (output) = lane % 16
#pragma unroll
  for (int fn = 0; fn < FN; ++fn) {
    const int64_t col = n_base + fn * 16 + li;
    if (col >= n) continue;
#pragma unroll
    for (int fm = 0; fm < 2; ++fm)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = m_base + fm * 16 + g + 4 * r;
        if (row >= P) continue;
        double* __restrict__ out = J + (row * n + col) * nz + z0;
#pragma unroll
        for (int jj = 0; jj < JG; ++jj)
          if (z0 + jj < nz) out[jj] = acc[jj][fm][fn][r];
      }
  }
}

}  // namespace

namespace sbi {

template <int JG, int FN>
static int jac_contract_launch(const double* A, int64_t lda, const double* B, int64_t ldb,
                               const int32_t* brow, int64_t brs, const int32_t* zoff, int nz,
                               double* J, int64_t P, int64_t n, hipStream_t st) {
  const int64_t nt = (n + 32 * FN - 1) / (32 * FN), ntp = (P + kBlockM - 1) / kBlockM,
                nzg = (nz + JG - 1) / JG;
  const int64_t npairs = ntp * nzg;
  const int64_t grid = nt * ((npairs + 7) / 8) * 8;
  if (grid > 0x7fffffff || npairs > 0x7fffffff) {
    set_error("contract_jacobian_f64: launch too large");
    return SBI_ERR_INVALID;
  }
  hipLaunchKernelGGL((jac_contract_kernel<JG, FN>), dim3((unsigned)grid), dim3(64 * kWaves), 0,
                     st, A, lda, B, ldb, brow, brs, zoff, nz, J, P, n, (int)nt, (int)ntp,
                     (int)npairs);
  SBI_HIP_CHECK(hipGetLastError());
  return SBI_OK;
}

int contract_jacobian_f64(const double* A, int64_t lda, const double* B, int64_t ldb,
                          const int32_t* brow, int64_t brs, const int32_t* zoff, int nz, double* J,
                          int64_t P, int64_t n, hipStream_t st) {
  if (P <= 0 || n <= 0 || nz <= 0) return SBI_OK;
  if (n > 512) return jac_contract_launch<2, 4>(A, lda, B, ldb, brow, brs, zoff, nz, J, P, n, st);
  return jac_contract_launch<8, 1>(A, lda, B, ldb, brow, brs, zoff, nz, J, P, n, st);
}

}  // namespace sbi

extern "C" int sbi_contract(const void* A, int64_t lda, const void* B, int64_t ldb,
                            const double* offset, double* Y, int64_t ldy, int64_t P, int64_t n,
                            int64_t K, int32_t dtype, void* stream) {
  if (P < 0 || n < 0 || K < 0 || (P > 0 && n > 0 && (!Y || ldy < n)) ||
      (P > 0 && n > 0 && K > 0 && (!A || !B || lda < P || ldb < n)) ||
      P > (int64_t)65535 * kBlockM || n > (int64_t)65535 * 32) {
    sbi::set_error("sbi_contract: invalid argument");
    return SBI_ERR_INVALID;
  }
  if (P == 0 || n == 0) return SBI_OK;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == 0)
    return contract_launch<double>((const double*)A, lda, (const double*)B, ldb, offset, Y, ldy,
                                   P, n, K, nullptr, 0, nullptr, 0, 0, st);
  if (dtype == 1)
    return contract_launch<float>((const float*)A, lda, (const float*)B, ldb, offset, Y, ldy, P,
                                  n, K, nullptr, 0, nullptr, 0, 0, st);
  sbi::set_error("sbi_contract: dtype must be 0 (f64) or 1 (f32)");
  return SBI_ERR_INVALID;
}

extern "C" int sbi_contract_split_k4(const double* A64, int64_t lda64, const double* B64,
                                     int64_t ldb64, int64_t K64, const float* A32, int64_t lda32,
                                     const float* B32, int64_t ldb32, int64_t K32,
                                     const double* offset, double* Y, int64_t ldy, int64_t P,
                                     int64_t n, void* stream) {
  if (P < 0 || n < 0 || K64 < 0 || K32 < 0 || K32 % 16 != 0 ||
      (P > 0 && n > 0 && (!Y || ldy < n)) ||
      (P > 0 && n > 0 && K64 > 0 && (!A64 || !B64 || lda64 < P || ldb64 < n)) ||
      (P > 0 && n > 0 && K32 > 0 &&
       (!A32 || !B32 || lda32 < P || ldb32 < n || ((uintptr_t)A32 & 15) || ((uintptr_t)B32 & 15))) ||
      P > (int64_t)65535 * kBlockM || n > (int64_t)65535 * 32) {
    sbi::set_error("sbi_contract_split_k4: invalid argument (K32 must be a multiple of 16, the "
                   "k-interleaved operands 16-byte aligned)");
    return SBI_ERR_INVALID;
  }
  if (P == 0 || n == 0) return SBI_OK;
  return contract_launch<float, true>(A32, lda32, B32, ldb32, offset, Y, ldy, P, n, K32, A64,
                                      lda64, B64, ldb64, K64, (hipStream_t)stream);
}

extern "C" int64_t sbi_contract_split_k4_workspace(int64_t P, int64_t n, int64_t K64,
                                                   int64_t K32) {
  if (P < 0 || n < 0 || K64 < 0 || K32 < 0 || K32 % 16 != 0) {
    sbi::set_error("sbi_contract_split_k4_workspace: invalid argument");
    return -1;
  }
  const SplitPlan sp = split_plan<float>(P, n, K32, K64);
  return sp.nsplit ? sp.part_bytes + sp.head_bytes : 0;
}

extern "C" int sbi_contract_split_k4_ws(const double* A64, int64_t lda64, const double* B64,
                                        int64_t ldb64, int64_t K64, const float* A32,
                                        int64_t lda32, const float* B32, int64_t ldb32,
                                        int64_t K32, const double* offset, double* Y, int64_t ldy,
                                        int64_t P, int64_t n, void* workspace, int64_t ws_bytes,
                                        void* stream) {
  if (P < 0 || n < 0 || K64 < 0 || K32 < 0 || K32 % 16 != 0 || ws_bytes < 0 ||
      (P > 0 && n > 0 && (!Y || ldy < n)) ||
      (P > 0 && n > 0 && K64 > 0 && (!A64 || !B64 || lda64 < P || ldb64 < n)) ||
      (P > 0 && n > 0 && K32 > 0 &&
       (!A32 || !B32 || lda32 < P || ldb32 < n || ((uintptr_t)A32 & 15) || ((uintptr_t)B32 & 15))) ||
      ((uintptr_t)workspace & 255) || P > (int64_t)65535 * kBlockM || n > (int64_t)65535 * 32) {
    sbi::set_error("sbi_contract_split_k4_ws: invalid argument (K32 must be a multiple of 16, the "
                   "k-interleaved operands 16-byte aligned, the workspace 256-byte aligned)");
    return SBI_ERR_INVALID;
  }
  if (P == 0 || n == 0) return SBI_OK;
  const SplitPlan sp = split_plan<float>(P, n, K32, K64);
  if (sp.nsplit == 0 || !workspace || ws_bytes < sp.part_bytes + sp.head_bytes)
    return contract_launch<float, true>(A32, lda32, B32, ldb32, offset, Y, ldy, P, n, K32, A64,
                                        lda64, B64, ldb64, K64, (hipStream_t)stream);
  return contract_split_launch<float, true>(sp, A32, lda32, B32, ldb32, offset, Y, ldy, P, n,
                                            K32, A64, lda64, B64, ldb64, K64, workspace,
                                            (hipStream_t)stream);
}

extern "C" int sbi_contract_split(const double* A64, int64_t lda64, const double* B64,
                                  int64_t ldb64, int64_t K64, const float* A32, int64_t lda32,
                                  const float* B32, int64_t ldb32, int64_t K32,
                                  const double* offset, double* Y, int64_t ldy, int64_t P,
                                  int64_t n, void* stream) {
  if (P < 0 || n < 0 || K64 < 0 || K32 < 0 || (P > 0 && n > 0 && (!Y || ldy < n)) ||
      (P > 0 && n > 0 && K64 > 0 && (!A64 || !B64 || lda64 < P || ldb64 < n)) ||
      (P > 0 && n > 0 && K32 > 0 && (!A32 || !B32 || lda32 < P || ldb32 < n)) ||
      P > (int64_t)65535 * kBlockM || n > (int64_t)65535 * 32) {
    sbi::set_error("sbi_contract_split: invalid argument");
    return SBI_ERR_INVALID;
  }
  if (P == 0 || n == 0) return SBI_OK;
  return contract_launch<float>(A32, lda32, B32, ldb32, offset, Y, ldy, P, n, K32, A64, lda64,
                                B64, ldb64, K64, (hipStream_t)stream);
}

extern "C" int sbi_contract_bf16_k8(const double* A64, int64_t lda64, const double* B64,
                                    int64_t ldb64, int64_t K64, const void* A16, int64_t lda16,
                                    const void* B16, int64_t ldb16, int64_t K16,
                                    const double* offset, double* Y, int64_t ldy, int64_t P,
                                    int64_t n, void* stream) {
  if (P < 0 || n < 0 || K64 < 0 || K16 < 0 || K16 % 32 != 0 ||
      (P > 0 && n > 0 && (!Y || ldy < n)) ||
      (P > 0 && n > 0 && K64 > 0 && (!A64 || !B64 || lda64 < P || ldb64 < n)) ||
      (P > 0 && n > 0 && K16 > 0 &&
       (!A16 || !B16 || lda16 < P || ldb16 < n || ((uintptr_t)A16 & 15) || ((uintptr_t)B16 & 15))) ||
      (P + 63) / 64 > 0x7fffffff || (n + 63) / 64 > 65535) {
    sbi::set_error("sbi_contract_bf16_k8: invalid argument (K16 must be a multiple of 32, the "
                   "k-interleaved operands 16-byte aligned)");
    return SBI_ERR_INVALID;
  }
  if (P == 0 || n == 0) return SBI_OK;
  // per-wave tiles: 64 x 64 (workgroup 128 x 128, LDS 64 KB) while that gives at least 1024
  // workgroups, else 64 x 32 (128 x 64, 48 KB) likewise, else 32 x 32 (64 x 64, 32 KB): the
  // smaller tiles spread small grids further (2 and 4 waves per SIMD); every element's sum is the
  // same (same bits)
  const int64_t g44 = ((P + 127) / 128) * ((n + 127) / 128);
  const int64_t g42 = ((P + 127) / 128) * ((n + 63) / 64);
  if (g44 >= 1024) {
    const dim3 grid((unsigned)((P + 127) / 128), (unsigned)((n + 127) / 128));
    hipLaunchKernelGGL(contract_bf16_wide_kernel, grid, dim3(256), 0, (hipStream_t)stream,
                       (const uint4*)A16, lda16, (const uint4*)B16, ldb16, K16, A64, lda64, B64,
                       ldb64, K64, offset, Y, ldy, P, n);
  } else if (g42 >= 1024) {
    const dim3 grid((unsigned)((P + 127) / 128), (unsigned)((n + 63) / 64));
    hipLaunchKernelGGL((contract_bf16_kernel<4, 2>), grid, dim3(256), 0, (hipStream_t)stream,
                       (const uint4*)A16, lda16, (const uint4*)B16, ldb16, K16, A64, lda64, B64,
                       ldb64, K64, offset, Y, ldy, P, n);
  } else {
    const dim3 grid((unsigned)((P + 63) / 64), (unsigned)((n + 63) / 64));
    hipLaunchKernelGGL((contract_bf16_kernel<2, 2>), grid, dim3(256), 0, (hipStream_t)stream,
                       (const uint4*)A16, lda16, (const uint4*)B16, ldb16, K16, A64, lda64, B64,
                       ldb64, K64, offset, Y, ldy, P, n);
  }
  SBI_HIP_CHECK(hipGetLastError());
  return SBI_OK;
}
