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
// POTRF gives every element the operations of the reference's left-looking loop over
// 32-column blocks (SYRK of the diagonal block against all previous columns, POTRF2 of the
// block, GEMM of the panel against the previous columns, TRSM with the factored block),
// scheduled right-looking: each finished block column pushes its terms into the later ones
// (see chol_factor).  POTRF2 of a block of at most 32 replays the reference's recursion
// (halve; factor the first half; TRSM; SYRK; factor the second half; sqrt at the leaves)
// with its exact base-case arithmetic.
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

// ---- POTRF2 of a full 32 x 32 block with the recursion unrolled at compile time (no
// run-time stack), the same arithmetic as potrf2_small: halve; factor the first half;
// TRSM of the second half's rows (L: multiply by the reciprocal of the diagonal, which
// the leaf computes once -- the same correctly rounded 1.0 / a as the reference's; U:
// divide); SYRK into the trailing square (L: daxpy order, fused on the 4-aligned prefix
// of each column segment, zero multipliers skipped; U: dot order); factor the second
// half.  Nodes of 8 or fewer columns run in one thread's registers, straight through
// (their parallel steps become independent instruction streams of that thread); the 16-
// and 32-wide levels run their TRSM rows and SYRK elements on separate threads between
// barriers.  A failing leaf records its column and every level returns at once, leaving
// the block as the reference leaves it.

// an 8 x 8 diagonal block in registers: m[r][c] (L: r >= c used; U: r <= c used)
template <bool UP, int N, int O>
__device__ __forceinline__ bool p2_reg(double (&m)[8][8], double (&ri)[8], int& fail) {
    if constexpr (N == 1) {
        const double v = m[O][O];
        if (!(v > 0.0)) {
            fail = O + 1;
            return true;
        }
        const double r = sqrt(v);
        m[O][O] = r;
        if (!UP) ri[O] = 1.0 / r;
        return false;
    } else {
        constexpr int n1 = N / 2, n2 = N - n1;
        if (p2_reg<UP, n1, O>(m, ri, fail)) return true;
        if (!UP) {
#pragma unroll
            for (int t = 0; t < n2; t++) {
                const int r = O + n1 + t;
#pragma unroll
                for (int j = 0; j < n1; j++) {
                    double bj = m[r][O + j];
#pragma unroll
                    for (int kk = 0; kk < j; kk++) {
                        const double a = m[O + j][O + kk];
                        const double y = bj - a * m[r][O + kk];
                        bj = a != 0.0 ? y : bj;
                    }
                    m[r][O + j] = bj * ri[O + j];
                }
            }
#pragma unroll
            for (int jj = 0; jj < n2; jj++)
#pragma unroll
                for (int i = jj; i < n2; i++) {
                    const bool fused = (i - jj) < ((n2 - jj) & ~3);
                    double c = m[O + n1 + i][O + n1 + jj];
#pragma unroll
                    for (int l = 0; l < n1; l++) {
                        const double temp = -1.0 * m[O + n1 + jj][O + l];
                        const double x = m[O + n1 + i][O + l];
                        const double y = fused ? fma(temp, x, c) : c + temp * x;
                        c = temp != 0.0 ? y : c;
                    }
                    m[O + n1 + i][O + n1 + jj] = c;
                }
        } else {
#pragma unroll
            for (int t = 0; t < n2; t++) {
                const int cc = O + n1 + t;
#pragma unroll
                for (int i = 0; i < n1; i++) {
                    double temp = m[O + i][cc];
#pragma unroll
                    for (int kk = 0; kk < i; kk++) temp = temp - m[O + kk][O + i] * m[O + kk][cc];
                    m[O + i][cc] = temp / m[O + i][O + i];
                }
            }
#pragma unroll
            for (int jj = 0; jj < n2; jj++)
#pragma unroll
                for (int i = 0; i <= jj; i++) {
                    constexpr int nc = n1 / 4;
                    double l0 = 0.0, l1 = 0.0, l2 = 0.0, l3 = 0.0;
                    const int ci = O + n1 + i, cj = O + n1 + jj;
#pragma unroll
                    for (int t = 0; t < nc; t++) {
                        l0 = fma(m[O + 4 * t + 0][ci], m[O + 4 * t + 0][cj], l0);
                        l1 = fma(m[O + 4 * t + 1][ci], m[O + 4 * t + 1][cj], l1);
                        l2 = fma(m[O + 4 * t + 2][ci], m[O + 4 * t + 2][cj], l2);
                        l3 = fma(m[O + 4 * t + 3][ci], m[O + 4 * t + 3][cj], l3);
                    }
                    double hs = -0.0;
                    hs = hs + l0;
                    hs = hs + l1;
                    hs = hs + l2;
                    hs = hs + l3;
                    double ts = 0.0;
                    ts = ts + hs;
#pragma unroll
                    for (int t = nc * 4; t < n1; t++) ts = ts + m[O + t][ci] * m[O + t][cj];
                    m[ci][cj] = m[ci][cj] + (-1.0 * ts);
                }
        }
        return p2_reg<UP, n2, O + n1>(m, ri, fail);
    }
}

template <bool UP, int N>
__device__ __forceinline__ bool p2_node(double (*s)[SB + 1], double* rinv, int* err, int o, int tid) {
    if constexpr (N <= 8) {
        if (tid == 0) {
            double m[8][8], ri[8];
#pragma unroll
            for (int r = 0; r < N; r++)
#pragma unroll
                for (int c = 0; c < N; c++)
                    if (UP ? r <= c : r >= c) m[r][c] = s[o + r][o + c];
            int fail = 0;
            p2_reg<UP, N, 0>(m, ri, fail);
#pragma unroll
            for (int r = 0; r < N; r++)
#pragma unroll
                for (int c = 0; c < N; c++)
                    if (UP ? r <= c : r >= c) s[o + r][o + c] = m[r][c];
            if (!UP)
#pragma unroll
                for (int r = 0; r < N; r++) rinv[o + r] = ri[r];
            if (fail) *err = o + fail;
        }
        __syncthreads();
        return *err != 0;
    } else {
        constexpr int n1 = N / 2, n2 = N - n1;
        if (p2_node<UP, n1>(s, rinv, err, o, tid)) return true;
        if (!UP) {
            // trsm('R','L','T','N', n2, n1, 1, A11, A21), rows independent (trsm.rs:316-338)
            if (tid < n2) {
                const int r = o + n1 + tid;
                double xr[n1];
#pragma unroll
                for (int j = 0; j < n1; j++) xr[j] = s[r][o + j];
#pragma unroll
                for (int j = 0; j < n1; j++) {
                    double bj = xr[j];
#pragma unroll
                    for (int kk = 0; kk < j; kk++) {
                        const double a = s[o + j][o + kk];
                        const double y = bj - a * xr[kk];
                        bj = a != 0.0 ? y : bj;
                    }
                    xr[j] = bj * rinv[o + j];
                }
#pragma unroll
                for (int j = 0; j < n1; j++) s[r][o + j] = xr[j];
            }
            __syncthreads();
            // syrk('L','N', n2, n1, -1, A21, A22): one block, daxpy order (syrk.rs:213-222)
            for (int e = tid; e < n2 * n2; e += blockDim.x) {
                const int i = e % n2, jj = e / n2;
                if (i >= jj) {
                    const int len = n2 - jj;
                    const bool fused = (i - jj) < (len & ~3);
                    double c = s[o + n1 + i][o + n1 + jj];
#pragma unroll
                    for (int l = 0; l < n1; l++) {
                        const double temp = -1.0 * s[o + n1 + jj][o + l];
                        const double x = s[o + n1 + i][o + l];
                        const double y = fused ? fma(temp, x, c) : c + temp * x;
                        c = temp != 0.0 ? y : c;
                    }
                    s[o + n1 + i][o + n1 + jj] = c;
                }
            }
        } else {
            // trsm('L','U','T','N', n1, n2, 1, A11, A12), columns independent (trsm.rs:213-241)
            if (tid < n2) {
                const int cc = o + n1 + tid;
                double xc[n1];
#pragma unroll
                for (int i = 0; i < n1; i++) xc[i] = s[o + i][cc];
#pragma unroll
                for (int i = 0; i < n1; i++) {
                    double temp = xc[i];
#pragma unroll
                    for (int kk = 0; kk < i; kk++) temp = temp - s[o + kk][o + i] * xc[kk];
                    xc[i] = temp / s[o + i][o + i];
                }
#pragma unroll
                for (int i = 0; i < n1; i++) s[o + i][cc] = xc[i];
            }
            __syncthreads();
            // syrk('U','T', n2, n1, -1, A12, A22): dot order (syrk.rs:226-235)
            for (int e = tid; e < n2 * n2; e += blockDim.x) {
                const int i = e % n2, jj = e / n2;
                if (i <= jj) {
                    constexpr int nc = n1 / 4;
                    double l0 = 0.0, l1 = 0.0, l2 = 0.0, l3 = 0.0;
                    const int ci = o + n1 + i, cj = o + n1 + jj;
#pragma unroll
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
#pragma unroll
                    for (int t = nc * 4; t < n1; t++) ts = ts + s[o + t][ci] * s[o + t][cj];
                    s[o + n1 + i][o + n1 + jj] = s[o + n1 + i][o + n1 + jj] + (-1.0 * ts);
                }
            }
        }
        __syncthreads();
        return p2_node<UP, n2>(s, rinv, err, o + n1, tid);
    }
}

// ---- one block step of the factorization in one launch: POTRF2 of the 32 x 32 diagonal
// block at (c, c), then the TRSM of the block column below it (L: trsm('R','L','T','N'),
// one thread per row) or the block row right of it (U: trsm('L','U','T','N'), one
// thread per column), each thread running the reference's 32-wide base-case sequence
// (trsm.rs:252-376 / :213-241).  Every workgroup factors the diagonal block itself (the
// same bits everywhere) and then solves its 64 rows (columns); a failing pivot leaves the
// block as left by the failing step, sets the error, and no TRSM runs.  The factored
// block goes to `dout` (32 x 32, ld 32) with `flag` = c + 1, never in place: a workgroup
// that starts after another has finished must still read the unfactored block (on a busy
// GPU the workgroups of one launch start at different times); chol_diag_put copies it into
// place after the launch.
template <bool UP>
__global__ void __launch_bounds__(64) chol_panel(int n, int c, double* __restrict__ A, size_t lda,
                                                  int32_t* __restrict__ info, double* __restrict__ dout,
                                                  int* __restrict__ flag) {
    if (*info != 0) return;
    __shared__ double s[SB][SB + 1];  // s[r][c]
    __shared__ double rinv[SB];
    __shared__ int err;
    const int tid = threadIdx.x;
    double* D = A + (size_t)c + (size_t)c * lda;
    for (int e = tid; e < SB * SB; e += 64) s[e % SB][e / SB] = D[(e % SB) + (size_t)(e / SB) * lda];
    if (tid == 0) err = 0;
    __syncthreads();
    const bool failed = p2_node<UP, SB>(s, rinv, &err, 0, tid);
    if (blockIdx.x == 0) {
        for (int e = tid; e < SB * SB; e += 64) dout[e] = s[e % SB][e / SB];
        if (tid == 0) {
            *flag = c + 1;
            if (failed) {
                info[1] = err;
                __hip_atomic_store(info, c + err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
    if (failed) return;
    const long v = (long)blockIdx.x * 64 + tid;
    if (v >= n - c - SB) return;
    double x[SB];
    if (!UP) {
        double* row = A + (size_t)(c + SB + v) + (size_t)c * lda;
#pragma unroll
        for (int q = 0; q < SB; q++) x[q] = row[(size_t)q * lda];
#pragma unroll
        for (int j = 0; j < SB; j++) {
            double bj = x[j];
#pragma unroll
            for (int k = 0; k < j; k++) {
                const double a = s[j][k];
                const double y = bj - a * x[k];
                bj = a != 0.0 ? y : bj;
            }
            x[j] = bj * rinv[j];
        }
#pragma unroll
        for (int q = 0; q < SB; q++) row[(size_t)q * lda] = x[q];
    } else {
        double* col = A + (size_t)c + (size_t)(c + SB + v) * lda;
#pragma unroll
        for (int q = 0; q < SB; q++) x[q] = col[q];
#pragma unroll
        for (int i = 0; i < SB; i++) {
            double temp = x[i];
#pragma unroll
            for (int k = 0; k < i; k++) temp = temp - s[k][i] * x[k];
            x[i] = temp / s[i][i];
        }
#pragma unroll
        for (int q = 0; q < SB; q++) col[q] = x[q];
    }
}

// the diagonal block chol_panel factored at column c, into place (nothing when that launch
// returned at once behind an earlier failure: flag is then not c + 1)
__global__ void __launch_bounds__(256) chol_diag_put(int c, double* __restrict__ A, size_t lda,
                                                     const double* __restrict__ dout, const int* __restrict__ flag) {
    if (*flag != c + 1) return;
    for (int e = threadIdx.x; e < SB * SB; e += 256) A[(size_t)(c + e % SB) + (size_t)(c + e / SB) * lda] = dout[e];
}

// ---- the blocked factorization, right-looking with the reference's arithmetic
//
// The reference's loop is left-looking: block column j receives, at step j, a SYRK of its
// diagonal block over all earlier columns (one exact chain per element, l = 0 .. j-1),
// and a GEMM of the rows below over all earlier columns (64-deep slabs folded in order,
// the last one 32 deep when j is an odd multiple of 32).  Every one of those terms is
// known as soon as its earlier block column is final, so here each finished block
// column c pushes its terms into the later ones at once: 32 more chain steps into every
// later diagonal block, and the slab folds -- a 32-deep one into block column c + 32
// when c is an even block (it is that column's last, partial slab), a 64-deep one over
// columns c - 32 .. c + 31 into every later block column when c is an odd block.  Each
// element receives the same operations in the same order, so the result is the
// reference's bit for bit; the folds are large masked GEMMs on the matrix cores and the
// critical path is one small POTRF2 and TRSM per block column.
//
// A failing block leaves later block columns updated, which the reference never does:
// the part of the matrix the factorization may write is saved first and restored for
// every block column after the failing one (and below its diagonal block) when a pivot
// fails.

// the block-triangle region of element (r, col): L: r in or below col's diagonal block;
// U: r in or above it
__device__ __forceinline__ bool chol_region(int upper, int r, int col) {
    return upper ? (r / SB <= col / SB) : (r / SB >= col / SB);
}

__global__ void chol_save(int upper, int n, const double* __restrict__ A, size_t lda, double* __restrict__ W) {
    const long total = (long)n * n;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
        const int r = (int)(e % n), col = (int)(e / n);
        if (chol_region(upper, r, col)) W[e] = A[r + (size_t)col * lda];
    }
}

// after a failure at block jf: the region of every later block column (L) / block row
// (U), and of block column / row jf outside its diagonal block, as it was
__global__ void chol_restore(int upper, int n, double* __restrict__ A, size_t lda, const double* __restrict__ W,
                             const int32_t* __restrict__ info) {
    const int v = *info;
    if (v <= 0) return;
    const int bf = (v - 1) / SB;
    const long total = (long)n * n;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
        const int r = (int)(e % n), col = (int)(e / n);
        if (!chol_region(upper, r, col)) continue;
        const int own = upper ? r / SB : col / SB;   // the block step that finalises the element
        if (own < bf || (r / SB == bf && col / SB == bf)) continue;
        A[r + (size_t)col * lda] = W[e];
    }
}

// The chain steps l = l0 .. l0+steps-1 (steps a multiple of 32, taken 32 at a time in
// order) of the diagonal blocks starting at d0, d0 + 32, ..., with the reference's SYRK
// forms: L (syrk.rs, trans N, daxpy order): c += (alpha a[jj,l]) a[i,l], fused on the
// 4-aligned prefix of the column segment, skipped when alpha a[jj,l] == 0; U (trans T,
// dot order per 32-chunk): c = c + alpha * dot(A[l0:l0+32, i], A[l0:l0+32, jj]).
__global__ void __launch_bounds__(1024) chol_diag_steps(int upper, int n, int l0, int steps, int d0,
                                                         double* __restrict__ A, size_t lda,
                                                         const int32_t* __restrict__ abort) {
    if (*abort != 0) return;
    __shared__ double sl[SB][SB + 1];  // sl[r][l]: rows of the block, columns l .. l+31 (L) / transposed (U)
    const int j0 = d0 + (int)blockIdx.x * SB;
    const int nb = min(SB, n - j0);
    const int e = threadIdx.x, i = e & 31, jj = e >> 5;
    const bool act = i < nb && jj < nb && (upper ? (i <= jj) : (i >= jj));
    double* cp = A + (size_t)(j0 + i) + (size_t)(j0 + jj) * lda;
    double v = act ? *cp : 0.0;
    const int len = nb - jj, rel = i - jj;
    const bool fused = rel < (len & ~3);
    for (int lc = l0; lc < l0 + steps; lc += SB) {
        // stage the 32 x 32 operand block: L: A[j0 + r, lc + l]; U: A[lc + l, j0 + r]
        {
            const int r = e >> 5, l = e & 31;
            double w = 0.0;
            if (r < nb)
                w = upper ? A[(size_t)(lc + l) + (size_t)(j0 + r) * lda] : A[(size_t)(j0 + r) + (size_t)(lc + l) * lda];
            sl[r][l] = w;
        }
        __syncthreads();
        if (act) {
            if (!upper) {
#pragma unroll
                for (int l = 0; l < SB; l++) {
                    const double temp = -1.0 * sl[jj][l];
                    const double x = sl[i][l];
                    const double y = fused ? fma(temp, x, v) : v + temp * x;
                    v = temp != 0.0 ? y : v;
                }
            } else {
                double q0 = 0.0, q1 = 0.0, q2 = 0.0, q3 = 0.0;
#pragma unroll
                for (int t = 0; t < SB / 4; t++) {
                    q0 = fma(sl[i][4 * t + 0], sl[jj][4 * t + 0], q0);
                    q1 = fma(sl[i][4 * t + 1], sl[jj][4 * t + 1], q1);
                    q2 = fma(sl[i][4 * t + 2], sl[jj][4 * t + 2], q2);
                    q3 = fma(sl[i][4 * t + 3], sl[jj][4 * t + 3], q3);
                }
                double hs = -0.0;
                hs = hs + q0;
                hs = hs + q1;
                hs = hs + q2;
                hs = hs + q3;
                double ts = 0.0;
                ts = ts + hs;
                v = v + -1.0 * ts;
            }
        }
        __syncthreads();
    }
    if (act) *cp = v;
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
    const int N = (int)n;
    const int up = uplo == 'U' ? 1 : 0;
    hipStream_t s2 = h->aux_stream();
    const unsigned sgrid = (unsigned)std::min<size_t>((n * n + 255) / 256, 4096);
    // the save region, then the factored diagonal block of the current step and its flag
    const size_t wbytes = (n * n * sizeof(double) + 255) & ~(size_t)255;
    double* W = (double*)h->cws.get(wbytes + SB * SB * sizeof(double) + 256);
    double* dout = (double*)((char*)W + wbytes);
    int* dflag = (int*)(dout + SB * SB);
    hipLaunchKernelGGL(k_zero2, dim3(1), dim3(64), 0, st, dflag);
    BOLA_LAUNCH_CHECK();
    hipLaunchKernelGGL(chol_save, dim3(sgrid), dim3(256), 0, st, up, N, a, lda, W);
    BOLA_LAUNCH_CHECK();
    // the reference's GEMM edges (gemm.rs:177-190) for every block step: L -- the last n % 8
    // rows (its m is n - j - 32); U -- the last n % 4 columns (its n is n - j - 32)
    const int edge_r = N - N % 8, edge_c = N - N % 4;
    auto at = [&](int r, int c) { return a + (size_t)r + (size_t)c * lda; };
    // fold over columns (L) / rows (U) [k0, k0 + kd) into the block-triangle region whose
    // diagonal starts at d0, restricted to the nb columns (L) / rows (U) from d0 when nb > 0
    auto fold = [&](hipStream_t sx, int k0, int kd, int slab, int d0, int nb) {
        const int rest = N - d0;
        if (rest <= 0) return;
        const int w = nb > 0 ? std::min(nb, rest) : rest;
        Phase pg(h, "GEMM", (double)kd * w * (2.0 * rest - w), sx);
        if (!up)
            launch_dgemm_ex(sx, 'N', 'T', (size_t)rest, (size_t)w, (size_t)kd, -1.0, at(d0, k0), lda, at(d0, k0), lda,
                            1.0, at(d0, d0), lda, info, slab, 1, edge_r - d0, (long)w);
        else
            launch_dgemm_ex(sx, 'T', 'N', (size_t)w, (size_t)rest, (size_t)kd, -1.0, at(k0, d0), lda, at(k0, d0), lda,
                            1.0, at(d0, d0), lda, info, slab, 2, (long)w, edge_c - d0);
    };
    // Super-blocks of kSuper columns (four 64-deep slabs).  Inside one, each block column
    // pushes its terms into the later block columns of the same super-block only; the
    // super-block's terms for every column after it are applied once it is done -- the
    // chain steps of all later diagonal blocks, then the four slab folds in one product
    // (the next super-block's columns here, the rest on stream 2 while the next
    // super-block is factored).  Every element still receives its terms in the
    // reference's order: slabs by increasing k, chain steps by increasing l.
    constexpr int kSuper = 256;
    hipEvent_t bulk_done = nullptr;
    int ev = 0;
    auto diag_steps = [&](int l0, int steps, int d0, int d1) {
        if (d1 <= d0) return;
        Phase p(h, "SYRK", (double)SB * steps * (d1 - d0), st);
        hipLaunchKernelGGL(chol_diag_steps, dim3((unsigned)((d1 - d0 + SB - 1) / SB)), dim3(1024), 0, st, up, N, l0,
                           steps, d0, a, lda, info);
        BOLA_LAUNCH_CHECK();
    };
    for (int s0 = 0; s0 < N; s0 += kSuper) {
        const int e0 = std::min(N, s0 + kSuper);
        for (int c = s0; c < e0; c += SB) {
            const int jb = std::min(SB, N - c);
            const int n_rem = N - c - jb;
            if (jb == SB) {
                Phase p(h, "POTRF2", (double)jb * jb * n_rem, st);  // POTRF2 and the TRSM below it
                const unsigned g = (unsigned)std::max(1, (n_rem + 63) / 64);
                if (up) hipLaunchKernelGGL(chol_panel<true>, dim3(g), dim3(64), 0, st, N, c, a, lda, info, dout, dflag);
                else hipLaunchKernelGGL(chol_panel<false>, dim3(g), dim3(64), 0, st, N, c, a, lda, info, dout, dflag);
                BOLA_LAUNCH_CHECK();
                hipLaunchKernelGGL(chol_diag_put, dim3(1), dim3(256), 0, st, c, a, lda, dout, dflag);
                BOLA_LAUNCH_CHECK();
            } else {  // the last, partial block: nothing below it
                Phase p(h, "POTRF2", 0.0, st);
                potrf2_block(st, uplo, (size_t)jb, at(c, c), lda, info, c);
            }
            if (n_rem == 0) break;
            diag_steps(c, SB, c + SB, e0);
            if ((c / SB) % 2 == 0) {
                // the partial 32-deep last slab of block column c + 32
                if (c + SB < e0) fold(st, c, SB, SB, c + SB, SB);
            } else if (c + SB < e0) {
                // the 64-deep slab over block columns c - 32 .. c + 31, into the rest of the super-block
                fold(st, c - SB, 2 * SB, 2 * SB, c + SB, e0 - c - SB);
            }
        }
        if (e0 >= N) break;
        diag_steps(s0, e0 - s0, e0, N);
        // the next super-block's columns were the previous bulk update's until it is done
        if (bulk_done) BOLA_CHECK(hipStreamWaitEvent(st, bulk_done, 0));
        fold(st, s0, e0 - s0, 2 * SB, e0, kSuper);
        if (e0 + kSuper < N) {
            hipEvent_t ready = handle_event(h, ev++);
            BOLA_CHECK(hipEventRecord(ready, st));
            BOLA_CHECK(hipStreamWaitEvent(s2, ready, 0));
            fold(s2, s0, e0 - s0, 2 * SB, e0 + kSuper, 0);
            bulk_done = handle_event(h, ev++);
            BOLA_CHECK(hipEventRecord(bulk_done, s2));
        } else {
            bulk_done = nullptr;
        }
    }
    if (bulk_done) BOLA_CHECK(hipStreamWaitEvent(st, bulk_done, 0));
    hipLaunchKernelGGL(chol_restore, dim3(sgrid), dim3(256), 0, st, up, N, a, lda, W, info);
    BOLA_LAUNCH_CHECK();
}

}  // namespace bola
