// Mixed-precision solver internals (no reference counterpart): low-precision LU with
// equilibration, FP64 iterative refinement / GMRES-IR, FP64 fallback.
#pragma once

#include "common.hpp"

namespace bola {

// precision of the factorization's trailing-update products
enum { LP_F32 = 0, LP_F16 = 1, LP_BF16 = 2, LP_F16X3 = 3 };

// C (fp32, column-major) -= A * B with the products in precision lp; ws holds the 16-bit
// operand copies (lp_gemm_ws_bytes).
size_t lp_gemm_ws_bytes(int lp, size_t m, size_t n, size_t k);
// C = -(A * B), fp32 throughout (C's old contents are not read)
void launch_sgemm_set(hipStream_t st, int m, int n, int k, const float* A, size_t lda, const float* B, size_t ldb,
                      float* C, size_t ldc);
void launch_lp_gemm(hipStream_t st, int lp, int m, int n, int k, const float* A, size_t lda, const float* B,
                    size_t ldb, float* C, size_t ldc, void* ws);

// Recursive LU with tournament-pivoted 32-column leaves on T = (S A C)^T (fp32, see
// mixed_lu.hip).  ipiv: 0-based LAPACK-style interchanges; perm: row origins, updated in
// place (start from the identity).  n <= mixed_lu_max_n().
size_t mixed_lu_ws_bytes(int n, int lp);
// pivot = false: LU without interchanges, for input randomised by butterflies (mixed.hip).
void mixed_lu_factor(bola_hip_handle_s* h, int lp, bool pivot, int n, float* T, size_t ldt, int32_t* ipiv, int* perm,
                     void* ws);
int mixed_lu_max_n();

}  // namespace bola
