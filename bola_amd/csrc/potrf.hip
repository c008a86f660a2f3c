// SYRK and Cholesky on gfx950: syrk.rs:17-249, potrf.rs:101-181, potrf2.rs:25-80.
//
// SYRK keeps the reference's 32-block structure, which fixes its rounding: the 32x32
// diagonal blocks are updated element by element in increasing k -- daxpy-style
// (fused on 4-aligned chunks of each column segment, unfused tail, skipped when
// alpha*a is zero) for trans 'N', dot-product style (4 fused lanes, left-to-right lane
// sum, unfused tail, then c += alpha*dot) for trans 'T' -- and the off-diagonal blocks
// take one 32-deep accumulate-and-fold GEMM step per 32-chunk of k, which is the MFMA
// DGEMM with a 32-deep slab and a block-triangle mask.  For (uplo, trans) = (U, N) and
// (L, T) the reference addresses A with mixed offsets once n or k exceed 32
// (syrk.rs:80, :139); those two combinations are computed here with the BLAS meaning,
// which is what the reference computes whenever n, k <= 32.
//
// POTRF follows the reference's left-looking loop over 32-column blocks: SYRK of the
// diagonal block against all previous columns, POTRF2 of the block, GEMM of the panel
// against the previous columns, TRSM with the factored block.  POTRF2 of a block of at
// most 32 runs as one workgroup in LDS, replaying the reference's recursion (halve;
// factor the first half; TRSM; SYRK; factor the second half; sqrt at the leaves) with
// its exact base-case arithmetic.
#include "common.hpp"

namespace bola {

namespace {

constexpr int SB = 32;  // syrk.rs:9 NB, potrf.rs:8 NB

// ---- diagonal blocks of SYRK, trans 'N': C[i,j] += sum_l (alpha*A[j,l]) * A[i,l]
// Each thread owns one element of one diagonal block and walks l = 0..k-1.
__global__ void syrk_diag_n(int upper, int n, int k, double alpha, const double* __restrict__ A, size_t lda,
                            double* __restrict__ C, size_t ldc, Abort abort) {
    if (abort_stop(abort)) return;
    const int blk = blockIdx.x;
    const int j0 = blk * SB;
    const int nb = min(SB, n - j0);
    const int e = threadIdx.x;  // 0 .. 1023 -> (i, jj) in the 32x32 block
    const int i = e & 31, jj = e >> 5;
    if (i >= nb || jj >= nb) return;
    if (upper ? (i > jj) : (i < jj)) return;
    // daxpy length and fused prefix for column jj of the block
    const int len = upper ? (jj + 1) : (nb - jj);
    const int rel = upper ? i : (i - jj);
    const bool fused = rel < (len & ~3);
    double c = C[(size_t)(j0 + i) + (size_t)(j0 + jj) * ldc];
    const double* ai = A + (j0 + i);
    const double* aj = A + (j0 + jj);
    for (int l = 0; l < k; l++) {
        const double temp = alpha * aj[(size_t)l * lda];
        if (temp != 0.0) {
            const double x = ai[(size_t)l * lda];
            c = fused ? fma(temp, x, c) : c + temp * x;
        }
    }
    C[(size_t)(j0 + i) + (size_t)(j0 + jj) * ldc] = c;
}

// ---- diagonal blocks of SYRK, trans 'T': per 32-chunk p, c += alpha * dot_p(A[:,i], A[:,j])
// with the reference's dot_product order (utilities.rs:22-54).
__device__ __forceinline__ double bola_dot(int len, const double* x, const double* y) {
    const int nc = len / 4;
    double l0 = 0.0, l1 = 0.0, l2 = 0.0, l3 = 0.0;
    for (int t = 0; t < nc; t++) {
        l0 = fma(x[4 * t + 0], y[4 * t + 0], l0);
        l1 = fma(x[4 * t + 1], y[4 * t + 1], l1);
        l2 = fma(x[4 * t + 2], y[4 * t + 2], l2);
        l3 = fma(x[4 * t + 3], y[4 * t + 3], l3);
    }
    double hs = -0.0;
    hs = hs + l0;
    hs = hs + l1;
    hs = hs + l2;
    hs = hs + l3;
    double ts = 0.0;
    ts = ts + hs;
    for (int t = nc * 4; t < len; t++) ts = ts + x[t] * y[t];
    return ts;
}

__global__ void syrk_diag_t(int upper, int n, int k, double alpha, const double* __restrict__ A, size_t lda,
                            double* __restrict__ C, size_t ldc, Abort abort) {
    if (abort_stop(abort)) return;
    const int blk = blockIdx.x;
    const int j0 = blk * SB;
    const int nb = min(SB, n - j0);
    const int e = threadIdx.x;
    const int i = e & 31, jj = e >> 5;
    if (i >= nb || jj >= nb) return;
    if (upper ? (i > jj) : (i < jj)) return;
    double c = C[(size_t)(j0 + i) + (size_t)(j0 + jj) * ldc];
    const double* ci = A + (size_t)(j0 + i) * lda;   // column i of A (k x n)
    const double* cj = A + (size_t)(j0 + jj) * lda;
    for (int p = 0; p < k; p += SB) {
        const int pb = min(SB, k - p);
        const double temp = bola_dot(pb, ci + p, cj + p);
        c = c + alpha * temp;
    }
    C[(size_t)(j0 + i) + (size_t)(j0 + jj) * ldc] = c;
}

// ---- POTRF2 of one block (n <= 32) in LDS
__global__ void __launch_bounds__(256) potrf2_small(int upper, int nb, double* __restrict__ A, size_t lda,
                                                    int32_t* __restrict__ info, int col_off) {
    if (*info != 0) return;
    __shared__ double s[SB][SB + 1];  // s[r][c]
    __shared__ int err;
    const int tid = threadIdx.x;
    for (int e = tid; e < nb * nb; e += blockDim.x) {
        int r = e % nb, c = e / nb;
        s[r][c] = A[r + (size_t)c * lda];
    }
    if (tid == 0) err = 0;
    __syncthreads();
    struct Frame { int off, n, stage; };
    Frame stk[8];
    int sp = 0;
    stk[0] = {0, nb, 0};
    while (sp >= 0) {
        Frame f = stk[sp];
        if (f.n == 0) { sp--; continue; }
        if (f.n == 1) {
            if (tid == 0) {
                double v = s[f.off][f.off];
                if (v > 0.0) s[f.off][f.off] = sqrt(v);
                else err = f.off + 1;
            }
            __syncthreads();
            if (err) break;
            sp--;
            continue;
        }
        const int n1 = f.n / 2, n2 = f.n - n1, o = f.off;
        if (f.stage == 0) {
            stk[sp].stage = 1;
            stk[++sp] = {o, n1, 0};
            continue;
        }
        if (f.stage == 1) {
            if (!upper) {
                // trsm('R','L','T','N', n2, n1, 1, A11, A21): rows of A21 independent (trsm.rs:316-338)
                if (tid < n2) {
                    const int r = o + n1 + tid;
                    for (int j = 0; j < n1; j++) {
                        double bj = s[r][o + j];
                        for (int kk = 0; kk < j; kk++) {
                            const double a = s[o + j][o + kk];
                            if (a != 0.0) bj = bj - a * s[r][o + kk];
                        }
                        bj = bj * (1.0 / s[o + j][o + j]);
                        s[r][o + j] = bj;
                    }
                }
                __syncthreads();
                // syrk('L','N', n2, n1, -1, A21, A22): one block, daxpy order (syrk.rs:213-222)
                for (int e = tid; e < n2 * n2; e += blockDim.x) {
                    const int i = e % n2, jj = e / n2;
                    if (i < jj) continue;
                    const int len = n2 - jj;
                    const bool fused = (i - jj) < (len & ~3);
                    double c = s[o + n1 + i][o + n1 + jj];
                    for (int l = 0; l < n1; l++) {
                        const double temp = -1.0 * s[o + n1 + jj][o + l];
                        if (temp != 0.0) {
                            const double x = s[o + n1 + i][o + l];
                            c = fused ? fma(temp, x, c) : c + temp * x;
                        }
                    }
                    s[o + n1 + i][o + n1 + jj] = c;
                }
            } else {
                // trsm('L','U','T','N', n1, n2, 1, A11, A12): columns of A12 independent (trsm.rs:213-241)
                if (tid < n2) {
                    const int cc = o + n1 + tid;
                    for (int i = 0; i < n1; i++) {
                        double temp = s[o + i][cc];
                        for (int kk = 0; kk < i; kk++) temp = temp - s[o + kk][o + i] * s[o + kk][cc];
                        temp = temp / s[o + i][o + i];
                        s[o + i][cc] = temp;
                    }
                }
                __syncthreads();
                // syrk('U','T', n2, n1, -1, A12, A22): dot order (syrk.rs:226-235)
                for (int e = tid; e < n2 * n2; e += blockDim.x) {
                    const int i = e % n2, jj = e / n2;
                    if (i > jj) continue;
                    const int nc = n1 / 4;
                    double l0 = 0.0, l1 = 0.0, l2 = 0.0, l3 = 0.0;
                    const int ci = o + n1 + i, cj = o + n1 + jj;
                    for (int t = 0; t < nc; t++) {
                        l0 = fma(s[o + 4 * t + 0][ci], s[o + 4 * t + 0][cj], l0);
                        l1 = fma(s[o + 4 * t + 1][ci], s[o + 4 * t + 1][cj], l1);
                        l2 = fma(s[o + 4 * t + 2][ci], s[o + 4 * t + 2][cj], l2);
                        l3 = fma(s[o + 4 * t + 3][ci], s[o + 4 * t + 3][cj], l3);
                    }
                    double hs = -0.0;
                    hs = hs + l0;
                    hs = hs + l1;
                    hs = hs + l2;
                    hs = hs + l3;
                    double ts = 0.0;
                    ts = ts + hs;
                    for (int t = nc * 4; t < n1; t++) ts = ts + s[o + t][ci] * s[o + t][cj];
                    s[o + n1 + i][o + n1 + jj] = s[o + n1 + i][o + n1 + jj] + (-1.0 * ts);
                }
            }
            __syncthreads();
            stk[sp].stage = 2;
            stk[++sp] = {o + n1, n2, 0};
            continue;
        }
        sp--;
    }
    __syncthreads();
    for (int e = tid; e < nb * nb; e += blockDim.x) {
        int r = e % nb, c = e / nb;
        A[r + (size_t)c * lda] = s[r][c];
    }
    if (tid == 0 && err) {
        info[1] = err;
        __hip_atomic_store(info, col_off + err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

__global__ void k_zero2(int32_t* p) {
    if (threadIdx.x < 2) p[threadIdx.x] = 0;
}

}  // namespace

void launch_dsyrk(hipStream_t st, char uplo, char trans, size_t n, size_t k, double alpha, const double* a, size_t lda,
                  double beta, double* c, size_t ldc, Abort abort) {
    if (n == 0 || ((alpha == 0.0 || k == 0) && beta == 1.0)) return;
    const bool upper = (uplo == 'U');
    if (beta != 1.0) launch_scale_tri(st, upper ? 'U' : 'L', c, ldc, n, beta);
    if (alpha == 0.0 || k == 0) return;
    const bool tr = (trans == 'T' || trans == 't' || trans == 'C' || trans == 'c');
    const unsigned nblk = cdiv(n, SB);
    if (!tr)
        hipLaunchKernelGGL(syrk_diag_n, dim3(nblk), dim3(1024), 0, st, upper ? 1 : 0, (int)n, (int)k, alpha, a, lda,
                           c, ldc, abort);
    else
        hipLaunchKernelGGL(syrk_diag_t, dim3(nblk), dim3(1024), 0, st, upper ? 1 : 0, (int)n, (int)k, alpha, a, lda,
                           c, ldc, abort);
    BOLA_LAUNCH_CHECK();
    if (n > (size_t)SB) {
        // off-diagonal blocks: one 32-deep GEMM step per 32-chunk of k
        if (!upper) {
            if (!tr) launch_dgemm_ex(st, 'N', 'T', n, n, k, alpha, a, lda, a, lda, 1.0, c, ldc, abort, SB, 1);
            else launch_dgemm_ex(st, 'T', 'N', n, n, k, alpha, a, lda, a, lda, 1.0, c, ldc, abort, SB, 1);
        } else {
            if (tr) launch_dgemm_ex(st, 'T', 'N', n, n, k, alpha, a, lda, a, lda, 1.0, c, ldc, abort, SB, 2);
            else launch_dgemm_ex(st, 'N', 'T', n, n, k, alpha, a, lda, a, lda, 1.0, c, ldc, abort, SB, 2);
        }
    }
}

namespace {

void potrf2_block(hipStream_t st, char uplo, size_t n, double* a, size_t lda, int32_t* info, int col_off) {
    hipLaunchKernelGGL(potrf2_small, dim3(1), dim3(256), 0, st, uplo == 'U' ? 1 : 0, (int)n, a, lda, info, col_off);
    BOLA_LAUNCH_CHECK();
}

// potrf2 on n > 32: host recursion until blocks of at most 32 (potrf2.rs:50-77)
void potrf2_rec(hipStream_t st, char uplo, size_t n, double* a, size_t lda, int32_t* info, int col_off) {
    if (n == 0) return;
    if (n <= (size_t)SB) {
        potrf2_block(st, uplo, n, a, lda, info, col_off);
        return;
    }
    const size_t n1 = n / 2, n2 = n - n1;
    potrf2_rec(st, uplo, n1, a, lda, info, col_off);
    if (uplo == 'U') {
        double* a12 = a + n1 * lda;
        launch_dtrsm(st, 'L', 'U', 'T', 'N', n1, n2, 1.0, a, lda, a12, lda, info);
        double* a22 = a + n1 + n1 * lda;
        launch_dsyrk(st, 'U', 'T', n2, n1, -1.0, a12, lda, 1.0, a22, lda, info);
        potrf2_rec(st, uplo, n2, a22, lda, info, col_off + (int)n1);
    } else {
        double* a21 = a + n1;
        launch_dtrsm(st, 'R', 'L', 'T', 'N', n2, n1, 1.0, a, lda, a21, lda, info);
        double* a22 = a + n1 + n1 * lda;
        launch_dsyrk(st, 'L', 'N', n2, n1, -1.0, a21, lda, 1.0, a22, lda, info);
        potrf2_rec(st, uplo, n2, a22, lda, info, col_off + (int)n1);
    }
}

}  // namespace

// info[0] = 0 or the global 1-based column of the failing leading minor; info[1] = that
// column within its 32-column block (what the reference's potrf prints; potrf2, and
// potrf on n <= 32, print info[0]).
void chol_factor(bola_hip_handle_s* h, char uplo, size_t n, double* a, size_t lda, int32_t* info, bool recursive_entry) {
    hipStream_t st = h->stream;
    hipLaunchKernelGGL(k_zero2, dim3(1), dim3(64), 0, st, info);
    BOLA_LAUNCH_CHECK();
    if (recursive_entry || n <= (size_t)SB) {
        Phase ph(h, recursive_entry ? "POTRF2" : "POTRF");
        potrf2_rec(st, uplo, n, a, lda, info, 0);
        return;
    }
    Phase ph(h, "POTRF", (double)n * n * n / 3.0);
    for (size_t j = 0; j < n; j += SB) {
        const size_t jb = std::min((size_t)SB, n - j);
        double* ajj = a + j + j * lda;
        if (uplo == 'U') {
            double* a1j = a + j * lda;
            if (j > 0) {
                Phase p(h, "SYRK", (double)jb * jb * j);
                launch_dsyrk(st, 'U', 'T', jb, j, -1.0, a1j, lda, 1.0, ajj, lda, info);
            }
            {
                Phase p(h, "POTRF2");
                potrf2_block(st, 'U', jb, ajj, lda, info, (int)j);
            }
            if (j + jb < n) {
                const size_t n_rem = n - j - jb;
                double* a1jb = a + (j + jb) * lda;
                double* ajjb = a + j + (j + jb) * lda;
                if (j > 0) {
                    Phase p(h, "GEMM", 2.0 * jb * n_rem * j);
                    launch_dgemm(st, 'T', 'N', jb, n_rem, j, -1.0, a1j, lda, a1jb, lda, 1.0, ajjb, lda, info);
                }
                Phase p(h, "TRSM", (double)jb * jb * n_rem);
                launch_dtrsm(st, 'L', 'U', 'T', 'N', jb, n_rem, 1.0, ajj, lda, ajjb, lda, info);
            }
        } else {
            double* aj0 = a + j;
            if (j > 0) {
                Phase p(h, "SYRK", (double)jb * jb * j);
                launch_dsyrk(st, 'L', 'N', jb, j, -1.0, aj0, lda, 1.0, ajj, lda, info);
            }
            {
                Phase p(h, "POTRF2");
                potrf2_block(st, 'L', jb, ajj, lda, info, (int)j);
            }
            if (j + jb < n) {
                const size_t n_rem = n - j - jb;
                double* ajb0 = a + j + jb;
                double* ajbj = a + j + jb + j * lda;
                if (j > 0) {
                    Phase p(h, "GEMM", 2.0 * n_rem * jb * j);
                    launch_dgemm(st, 'N', 'T', n_rem, jb, j, -1.0, ajb0, lda, aj0, lda, 1.0, ajbj, lda, info);
                }
                Phase p(h, "TRSM", (double)jb * jb * n_rem);
                launch_dtrsm(st, 'R', 'L', 'T', 'N', n_rem, jb, 1.0, ajj, lda, ajbj, lda, info);
            }
        }
    }
}

}  // namespace bola
