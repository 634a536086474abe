// Contraction entry points shared by the library's HIP translation units (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sbi {

// FP64 contraction in groups (gemm.hip): for z < nz,
//   Y[z * yzs + p * ldy + c * ycs] = sum_{k in [zoff[z], zoff[z+1])} A[k * lda + p] * B[brow[k * brs] * ldb + c]
// for p < P, c < n (zoff, brow: device arrays).  A group with no rows writes zeros.
int contract_grouped_f64(const double* A, int64_t lda, const double* B, int64_t ldb,
                         const int32_t* brow, int64_t brs, const int32_t* zoff, int nz, double* Y,
                         int64_t ldy, int64_t ycs, int64_t yzs, int64_t P, int64_t n,
                         hipStream_t st);

}  // namespace sbi
