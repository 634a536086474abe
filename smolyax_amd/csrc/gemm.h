// Contraction entry points shared by the library's HIP translation units (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sbi {

// Jacobian contraction (gemm.hip), FP64 matrix cores: for each input dim z < nz,
//   J[(p * n + c) * nz + z] = sum_{k in [zoff[z], zoff[z+1])} A[k * lda + p] * B[brow[k * brs] * ldb + c]
// for p < P, c < n (zoff, brow: device arrays; zoff non-decreasing).  A dim with no rows gets
// zeros.
int contract_jacobian_f64(const double* A, int64_t lda, const double* B, int64_t ldb,
                          const int32_t* brow, int64_t brs, const int32_t* zoff, int nz, double* J,
                          int64_t P, int64_t n, hipStream_t st);

}  // namespace sbi
