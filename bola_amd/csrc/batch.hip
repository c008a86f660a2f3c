// Batched solves of many independent small systems (no reference counterpart: BOLA
// solves one system per call).  One persistent launch: every workgroup takes systems from
// a device-wide counter, largest first, and factors and solves each one alone:
//   LU with partial pivoting (LAPACK dgetrf/dgetrs semantics, 1-based ipiv) or Cholesky
//   (lower, dpotrf/dpotrs semantics),
// either in FP64 (A and B overwritten with the factors and the solution, like dgesv /
// dposv) or in FP32 with FP64 iterative refinement (A left unchanged, B overwritten with
// the solution; a system whose refinement does not reach n * 2^-52 is refactored in FP64).
//
// Inside a workgroup the factorization is right-looking with panels of NB columns: the
// panel lives in LDS for the pivot search and the rank-1 steps, U12 is solved column by
// column, and the trailing update runs as 64 x 64 output blocks, 4 x 4 per thread, the
// U12 rows staged in LDS.  NB and the number of workgroups per CU are chosen by the
// autotuner below (bola_hip_batch_tune_dev).
#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <numeric>
#include <vector>

namespace bola {
namespace {

constexpr int BT = 256;     // threads per workgroup
constexpr int BMAX = 512;   // largest order
constexpr int BNR = 16;     // most right-hand sides
constexpr int IR_MAX = 10;  // refinement steps before the FP64 fallback

using Desc = bola_hip_batch_desc;

// per-workgroup scratch, in bytes
constexpr size_t kWsFd = (size_t)BMAX * BMAX * 8;  // FP64 factors (refinement fallback)
constexpr size_t kWsFw = (size_t)BMAX * BMAX * 4;  // FP32 factors
constexpr size_t kWsVec = (size_t)BMAX * BNR * 8;  // one n x nrhs FP64 block
constexpr size_t kWsPerWg = kWsFd + kWsFw + 3 * kWsVec + 2 * BMAX * 4;

struct Lds {
    double red_v[BT / 64];
    int red_i[BT / 64];
    int piv_sh;
    int idx_sh;
    int info_sh;
    int flag_sh;
};

__device__ __forceinline__ void block_max_abs(double v, int i, Lds& s, int& out) {
    // largest |value|, then the smallest index
    const int tid = threadIdx.x;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(v, off);
        const int oi = __shfl_xor(i, off);
        if (ov > v || (ov == v && oi < i)) {
            v = ov;
            i = oi;
        }
    }
    if ((tid & 63) == 0) {
        s.red_v[tid >> 6] = v;
        s.red_i[tid >> 6] = i;
    }
    __syncthreads();
    if (tid == 0) {
        double bv = s.red_v[0];
        int bi = s.red_i[0];
        for (int w = 1; w < BT / 64; w++)
            if (s.red_v[w] > bv || (s.red_v[w] == bv && s.red_i[w] < bi)) {
                bv = s.red_v[w];
                bi = s.red_i[w];
            }
        s.piv_sh = bi;
    }
    __syncthreads();
    out = s.piv_sh;
}

// One 64 x 64 block of a trailing update, F_blk -= L(64 x nb) U(nb x 64), on the matrix
// cores: wavefront w owns the 32 x 32 quadrant (wm, wn), 2 x 2 MFMA tiles of 16 x 16.
// L(r, j) = Lp[j BMAX + lr0 + r] (clamped to the panel), U(j, c) = Up[j ldu + c] (uc
// columns valid).  The block is rows [r0, r0 + 64) and columns [q0, q0 + 64) of the
// trailing matrix at F (ld n), m2 x m2; LOWER: only its lower triangle.  F is read from
// clamped addresses first, so the loads are in flight together.
typedef double bd4 __attribute__((ext_vector_type(4)));
typedef float bf4 __attribute__((ext_vector_type(4)));

template <typename T, int NB, bool LOWER, bool COLS>
__device__ __forceinline__ void block_update(const T* Lp, int lr0, const T* Up, int ldu, int uc, int nb, T* F,
                                             size_t n, int m2, int r0, int q0) {
    using V = typename std::conditional<sizeof(T) == 8, bd4, bf4>::type;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = (w & 1) * 32, wn = (w >> 1) * 32;
    const int lr = lane & 15, lk = lane >> 4;
    V acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = V{0, 0, 0, 0};
#pragma unroll
    for (int k4 = 0; k4 < NB; k4 += 4) {
        const int k = k4 + lk;
        T fa[2], fb[2];
#pragma unroll
        for (int a = 0; a < 2; a++) {
            const T v = Lp[k * BMAX + min(lr0 + wm + 16 * a + lr, BMAX - 1)];
            fa[a] = k < nb ? v : (T)0;
        }
#pragma unroll
        for (int b = 0; b < 2; b++) {
            const T v = Up[k * ldu + min(wn + 16 * b + lr, uc - 1)];
            fb[b] = k < nb ? v : (T)0;
        }
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++) {
                // COLS (FP64): U first, so the tile comes out transposed and a lane's
                // results run down a column of F (whole 128-byte lines) rather than across
                // 16 columns; it costs registers, so the kernel holding the FP64 fallback
                // of the refined variant keeps the plain form
                if constexpr (sizeof(T) == 8 && COLS)
                    acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(fb[b], fa[a], acc[a][b], 0, 0, 0);
                else if constexpr (sizeof(T) == 8)
                    acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[a], fb[b], acc[a][b], 0, 0, 0);
                else
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[a], fb[b], acc[a][b], 0, 0, 0);
            }
    }
    // result element i of tile (a, b): row lr, column lk + 4 i (FP64, COLS); row lk + 4 i,
    // column lr (FP64); row 4 lk + i, column lr (FP32)
    T v[2][2][4];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int rr = r0 + wm + 16 * a + (sizeof(T) == 8 ? (COLS ? lr : lk + 4 * i) : 4 * lk + i);
                const int cc = q0 + wn + 16 * b + (sizeof(T) == 8 && COLS ? lk + 4 * i : lr);
                v[a][b][i] = F[(size_t)min(rr, m2 - 1) + (size_t)min(cc, m2 - 1) * n];
            }
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int rr = r0 + wm + 16 * a + (sizeof(T) == 8 ? (COLS ? lr : lk + 4 * i) : 4 * lk + i);
                const int cc = q0 + wn + 16 * b + (sizeof(T) == 8 && COLS ? lk + 4 * i : lr);
                if (rr < m2 && cc < m2 && (!LOWER || cc <= rr))
                    F[(size_t)rr + (size_t)cc * n] = v[a][b][i] - acc[a][b][i];
            }
}

// LU with partial pivoting of the n x n matrix F (ld n) in place; ipiv 1-based.  Returns
// the LAPACK info (first zero pivot, 1-based, or 0).
template <typename T, int NB, bool COLS>
__device__ __forceinline__ int lu_factor(int n, T* __restrict__ F, int32_t* __restrict__ ipiv, T* P, T* Us, int* spiv, Lds& s) {
    const int tid = threadIdx.x;
    int info = 0;
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int nb = min(NB, n - k0), mp = n - k0;
        for (int c = 0; c < nb; c++)
            for (int r = tid; r < mp; r += BT) P[c * BMAX + r] = F[(size_t)(k0 + r) + (size_t)(k0 + c) * n];
        __syncthreads();
        for (int j = 0; j < nb; j++) {
            double bv = -1.0;
            int bi = mp;
            for (int r = j + tid; r < mp; r += BT) {
                const double v = fabs((double)P[j * BMAX + r]);
                if (v > bv) {
                    bv = v;
                    bi = r;
                }
            }
            int p;
            block_max_abs(bv, bi, s, p);
            if (p >= mp) p = j;  // only NaNs below: keep the diagonal
            if (tid == 0) {
                ipiv[k0 + j] = k0 + p + 1;
                spiv[j] = p;
            }
            const T pv = P[j * BMAX + p];
            if (pv == (T)0 && info == 0) info = k0 + j + 1;
            __syncthreads();
            if (p != j && tid < nb) {
                const T t = P[tid * BMAX + j];
                P[tid * BMAX + j] = P[tid * BMAX + p];
                P[tid * BMAX + p] = t;
            }
            __syncthreads();
            if (pv != (T)0) {
                const T rc = (T)1 / pv;
                for (int r = j + 1 + tid; r < mp; r += BT) {
                    const T l = P[j * BMAX + r] * rc;
                    P[j * BMAX + r] = l;
                    for (int c = j + 1; c < nb; c++) P[c * BMAX + r] -= l * P[c * BMAX + j];
                }
            }
            __syncthreads();
        }
        for (int c = 0; c < nb; c++)
            for (int r = tid; r < mp; r += BT) F[(size_t)(k0 + r) + (size_t)(k0 + c) * n] = P[c * BMAX + r];
        // the panel's interchanges on the other columns
        for (int c = tid; c < n; c += BT) {
            if (c >= k0 && c < k0 + nb) continue;
            T* col = F + (size_t)c * n + k0;
            for (int j = 0; j < nb; j++) {
                const int p = spiv[j];
                if (p != j) {
                    const T t = col[j];
                    col[j] = col[p];
                    col[p] = t;
                }
            }
        }
        __syncthreads();
        const int c2 = k0 + nb, m2 = n - c2;
        if (m2 <= 0) continue;
        // U12 = L11^-1 A12, a column per thread
        for (int c = c2 + tid; c < n; c += BT) {
            T* col = F + (size_t)c * n + k0;
            T x[NB];
#pragma unroll
            for (int j = 0; j < NB; j++) x[j] = j < nb ? col[j] : (T)0;
#pragma unroll
            for (int j = 1; j < NB; j++)
#pragma unroll
                for (int i = 0; i < j; i++) x[j] -= P[i * BMAX + j] * x[i];
#pragma unroll
            for (int j = 0; j < NB; j++)
                if (j < nb) col[j] = x[j];
        }
        __syncthreads();
        // A22 -= L21 U12, by 64 x 64 blocks on the matrix cores
        T* F22 = F + c2 + (size_t)c2 * n;
        for (int cb = 0; cb < m2; cb += 64) {
            for (int e = tid; e < NB * 64; e += BT) {
                const int j = e >> 6, cc = e & 63;
                Us[j * 65 + cc] = (j < nb && cb + cc < m2) ? F[(size_t)(k0 + j) + (size_t)(c2 + cb + cc) * n] : (T)0;
            }
            __syncthreads();
            for (int rb = 0; rb < m2; rb += 64)
                block_update<T, NB, false, COLS>(P, nb + rb, Us, 65, 64, nb, F22, n, m2, rb, cb);
            __syncthreads();
        }
    }
    return info;
}

// Cholesky A = L L^T of the lower triangle of F (ld n) in place; the strict upper triangle
// is not referenced.  Returns the LAPACK info (order of the first non-positive leading
// minor, or 0); the factorization stops there.
template <typename T, int NB, bool COLS>
__device__ __forceinline__ int chol_factor(int n, T* __restrict__ F, T* P, Lds& s) {
    const int tid = threadIdx.x;
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int nb = min(NB, n - k0), mp = n - k0;
        for (int c = 0; c < nb; c++)
            for (int r = tid; r < mp; r += BT)
                P[c * BMAX + r] = r >= c ? F[(size_t)(k0 + r) + (size_t)(k0 + c) * n] : (T)0;
        __syncthreads();
        for (int j = 0; j < nb; j++) {
            const T d = P[j * BMAX + j];
            if (!(d > (T)0)) return k0 + j + 1;  // uniform: every thread read the same d
            const T sd = sqrt(d), rd = (T)1 / sd;
            __syncthreads();
            for (int r = j + tid; r < mp; r += BT) P[j * BMAX + r] = r == j ? sd : P[j * BMAX + r] * rd;
            __syncthreads();
            for (int r = j + 1 + tid; r < mp; r += BT) {
                const T l = P[j * BMAX + r];
                for (int c = j + 1; c < nb && c <= r; c++) P[c * BMAX + r] -= l * P[j * BMAX + c];
            }
            __syncthreads();
        }
        for (int c = 0; c < nb; c++)
            for (int r = c + tid; r < mp; r += BT) F[(size_t)(k0 + r) + (size_t)(k0 + c) * n] = P[c * BMAX + r];
        const int c2 = k0 + nb, m2 = n - c2;
        // A22 -= L21 L21^T, lower triangle only, by 64 x 64 blocks on the matrix cores
        T* F22 = F + c2 + (size_t)c2 * n;
        for (int cb = 0; cb < m2; cb += 64)
            for (int rb = cb; rb < m2; rb += 64)
                block_update<T, NB, true, COLS>(P, nb + rb, P + min(nb + cb, BMAX - 1), BMAX, BMAX - min(nb + cb, BMAX - 1),
                                          nb, F22, n, m2, rb, cb);
        __syncthreads();
    }
    return 0;
}

// X (n x nrhs, ld n, FP64) := L^-1 X with L the lower triangle of F (unit diagonal when
// UNIT), by blocks of NB rows: the diagonal block by one thread per right-hand side, the
// rows below by one thread per row.  D: NB x NB LDS, xs: NB x BNR LDS.
template <typename T, int NB, bool UNIT>
__device__ __forceinline__ void lower_solve(int n, int nrhs, const T* __restrict__ F, double* __restrict__ X, T* D, double* xs) {
    const int tid = threadIdx.x;
    for (int kb = 0; kb < n; kb += NB) {
        const int b = min(NB, n - kb);
        for (int e = tid; e < NB * NB; e += BT) {
            const int i = e / NB, j = e % NB;
            D[i * (NB + 1) + j] = (i < b && j < b) ? F[(size_t)(kb + i) + (size_t)(kb + j) * n] : (T)0;
        }
        __syncthreads();
        if (tid < nrhs) {
            double x[NB];
#pragma unroll
            for (int i = 0; i < NB; i++) x[i] = i < b ? X[(size_t)(kb + i) + (size_t)tid * n] : 0.0;
#pragma unroll
            for (int i = 0; i < NB; i++) {
#pragma unroll
                for (int j = 0; j < i; j++) x[i] -= (double)D[i * (NB + 1) + j] * x[j];
                if (!UNIT && i < b) x[i] /= (double)D[i * (NB + 1) + i];
            }
#pragma unroll
            for (int i = 0; i < NB; i++)
                if (i < b) {
                    X[(size_t)(kb + i) + (size_t)tid * n] = x[i];
                    xs[i * BNR + tid] = x[i];
                }
        }
        __syncthreads();
        for (int r = kb + b + tid; r < n; r += BT) {
            T f[NB];
#pragma unroll
            for (int j = 0; j < NB; j++) f[j] = j < b ? F[(size_t)r + (size_t)(kb + j) * n] : (T)0;
            for (int q = 0; q < nrhs; q++) {
                double sacc = 0.0;
#pragma unroll
                for (int j = 0; j < NB; j++) sacc += (double)f[j] * xs[j * BNR + q];
                X[(size_t)r + (size_t)q * n] -= sacc;
            }
        }
        __syncthreads();
    }
}

// X := U^-1 X with U upper triangular: U = the upper triangle of F (TRANS false) or the
// transpose of its lower triangle (TRANS, the Cholesky L^T); blocks from the bottom.
template <typename T, int NB, bool TRANS>
__device__ __forceinline__ void upper_solve(int n, int nrhs, const T* __restrict__ F, double* __restrict__ X, T* D, double* xs) {
    const int tid = threadIdx.x;
    const int nblk = (n + NB - 1) / NB;
    for (int bk = nblk - 1; bk >= 0; bk--) {
        const int kb = bk * NB, b = min(NB, n - kb);
        for (int e = tid; e < NB * NB; e += BT) {
            const int i = e / NB, j = e % NB;  // U(kb+i, kb+j)
            T v = (T)0;
            if (i < b && j < b) v = TRANS ? F[(size_t)(kb + j) + (size_t)(kb + i) * n] : F[(size_t)(kb + i) + (size_t)(kb + j) * n];
            D[i * (NB + 1) + j] = v;
        }
        __syncthreads();
        if (tid < nrhs) {
            double x[NB];
#pragma unroll
            for (int i = 0; i < NB; i++) x[i] = i < b ? X[(size_t)(kb + i) + (size_t)tid * n] : 0.0;
#pragma unroll
            for (int ii = 0; ii < NB; ii++) {
                const int i = NB - 1 - ii;
#pragma unroll
                for (int j = i + 1; j < NB; j++) x[i] -= (double)D[i * (NB + 1) + j] * x[j];
                if (i < b) x[i] /= (double)D[i * (NB + 1) + i];
            }
#pragma unroll
            for (int i = 0; i < NB; i++)
                if (i < b) {
                    X[(size_t)(kb + i) + (size_t)tid * n] = x[i];
                    xs[i * BNR + tid] = x[i];
                }
        }
        __syncthreads();
        for (int r = tid; r < kb; r += BT) {
            T f[NB];
#pragma unroll
            for (int j = 0; j < NB; j++) {
                T v = (T)0;
                if (j < b) v = TRANS ? F[(size_t)(kb + j) + (size_t)r * n] : F[(size_t)r + (size_t)(kb + j) * n];
                f[j] = v;
            }
            for (int q = 0; q < nrhs; q++) {
                double sacc = 0.0;
#pragma unroll
                for (int j = 0; j < NB; j++) sacc += (double)f[j] * xs[j * BNR + q];
                X[(size_t)r + (size_t)q * n] -= sacc;
            }
        }
        __syncthreads();
    }
}

// row permutation of the LU interchanges: perm[i] = original row of row i
__device__ __forceinline__ void piv_to_perm(int n, const int32_t* __restrict__ ipiv, int* perm) {
    const int tid = threadIdx.x;
    for (int i = tid; i < n; i += BT) perm[i] = i;
    __syncthreads();
    if (tid == 0)
        for (int j = 0; j < n; j++) {
            const int p = ipiv[j] - 1;
            if (p != j) {
                const int t = perm[j];
                perm[j] = perm[p];
                perm[p] = t;
            }
        }
    __syncthreads();
}

// Y = (P) S with P the row permutation perm (LU) or the identity (Cholesky)
__device__ __forceinline__ void gather_rows(int n, int nrhs, const int* perm, const double* __restrict__ S, double* __restrict__ Y) {
    for (int e = threadIdx.x; e < n * nrhs; e += BT) {
        const int i = e % n, q = e / n;
        Y[(size_t)i + (size_t)q * n] = S[(size_t)(perm ? perm[i] : i) + (size_t)q * n];
    }
    __syncthreads();
}

template <typename T, int NB>
__device__ __forceinline__ void solve_with(int n, int nrhs, bool lu, const T* F, const int* perm, const double* S, double* Y, T* D,
                           double* xs) {
    gather_rows(n, nrhs, lu ? perm : nullptr, S, Y);
    if (lu) {
        lower_solve<T, NB, true>(n, nrhs, F, Y, D, xs);
        upper_solve<T, NB, false>(n, nrhs, F, Y, D, xs);
    } else {
        lower_solve<T, NB, false>(n, nrhs, F, Y, D, xs);
        upper_solve<T, NB, true>(n, nrhs, F, Y, D, xs);
    }
}

// R = B - A X (FP64), and the backward error max_q |R_q|inf / (|A|inf |X_q|inf + |B_q|inf);
// anorm: |A|inf, computed when negative on entry.  The column maxima meet in LDS through
// atomic maxima on the bit patterns (non-negative doubles order like their bits).
__device__ __forceinline__ double residual_berr(int n, int nrhs, bool spd, const double* __restrict__ A, const double* __restrict__ B,
                                const double* __restrict__ X, double* __restrict__ R, double& anorm,
                                unsigned long long* mx) {
    const int tid = threadIdx.x;
    // mx: [0, BNR) |R|, [BNR, 2 BNR) |X|, [2 BNR, 3 BNR) |B|, [3 BNR] row sums of |A|
    for (int e = tid; e < 3 * BNR + 1; e += BT) mx[e] = 0ull;
    __syncthreads();
    for (int r = tid; r < n; r += BT) {
        double acc[BNR];
#pragma unroll
        for (int q = 0; q < BNR; q++) acc[q] = q < nrhs ? B[(size_t)r + (size_t)q * n] : 0.0;
        double asum = 0.0;
        for (int c = 0; c < n; c++) {
            // the SPD systems keep only their lower triangle meaningful: mirror it
            const size_t ia = (spd && c > r) ? (size_t)c + (size_t)r * n : (size_t)r + (size_t)c * n;
            const double a = A[ia];
            asum += fabs(a);
#pragma unroll
            for (int q = 0; q < BNR; q++)
                if (q < nrhs) acc[q] -= a * X[(size_t)c + (size_t)q * n];
        }
        atomicMax(&mx[3 * BNR], (unsigned long long)__double_as_longlong(asum));
#pragma unroll
        for (int q = 0; q < BNR; q++)
            if (q < nrhs) {
                R[(size_t)r + (size_t)q * n] = acc[q];
                atomicMax(&mx[q], (unsigned long long)__double_as_longlong(fabs(acc[q])));
                atomicMax(&mx[BNR + q], (unsigned long long)__double_as_longlong(fabs(X[(size_t)r + (size_t)q * n])));
                atomicMax(&mx[2 * BNR + q], (unsigned long long)__double_as_longlong(fabs(B[(size_t)r + (size_t)q * n])));
            }
    }
    __syncthreads();
    if (anorm < 0.0) anorm = __longlong_as_double((long long)mx[3 * BNR]);
    double berr = 0.0;
    for (int q = 0; q < nrhs; q++) {
        const double rq = __longlong_as_double((long long)mx[q]), xq = __longlong_as_double((long long)mx[BNR + q]),
                     bq = __longlong_as_double((long long)mx[2 * BNR + q]);
        const double den = anorm * xq + bq;
        berr = fmax(berr, den > 0.0 ? rq / den : (rq > 0.0 ? INFINITY : 0.0));
    }
    __syncthreads();
    return berr;
}

// One persistent launch over the batch.  order: system indices, largest first; next: the
// shared counter (zero on entry).  VAR 0: FP64; VAR 1: FP32 factors + FP64 refinement.
template <int NB, int VAR>
__global__ void __launch_bounds__(BT) batch_solve(int count, const Desc* __restrict__ desc,
                                                  const int* __restrict__ order, double* __restrict__ arena,
                                                  int32_t* __restrict__ piva, int32_t* __restrict__ info,
                                                  int32_t* __restrict__ iters, int* __restrict__ next,
                                                  char* __restrict__ wsall) {
    // sized for FP64 (the refinement's fallback factors in FP64 too)
    __shared__ double P[NB * BMAX];
    __shared__ double Us[NB * 65];
    __shared__ double Dd[NB * (NB + 1)];
    __shared__ double xs[NB * BNR];
    __shared__ int spiv[NB];
    __shared__ int perm[BMAX];
    __shared__ unsigned long long part[3 * BNR + 1];
    __shared__ Lds s;
    const int tid = threadIdx.x;
    char* ws = wsall + (size_t)blockIdx.x * kWsPerWg;
    double* Fd = (double*)ws;
    float* Fw = (float*)(ws + kWsFd);
    double* Xw = (double*)(ws + kWsFd + kWsFw);
    double* Rw = Xw + BMAX * BNR;
    double* Dw = Rw + BMAX * BNR;
    int32_t* pw = (int32_t*)(Dw + BMAX * BNR);
    for (;;) {
        if (tid == 0) s.idx_sh = atomicAdd(next, 1);
        __syncthreads();
        const int k = s.idx_sh;
        __syncthreads();
        if (k >= count) return;
        const int sys = order[k];
        const Desc d = desc[sys];
        const int n = d.n, nrhs = d.nrhs;
        const bool lu = d.kind == BOLA_HIP_BATCH_LU;
        double* A = arena + d.a;
        double* B = arena + d.b;
        int32_t* ipiv = lu ? piva + d.piv : pw;
        int inf = 0, its = 0;
        if (VAR == 0) {
            inf = lu ? lu_factor<double, NB, true>(n, A, ipiv, P, Us, spiv, s)
                     : chol_factor<double, NB, true>(n, A, P, s);
            if (inf == 0) {
                if (lu) piv_to_perm(n, ipiv, perm);
                solve_with<double, NB>(n, nrhs, lu, A, perm, B, Xw, Dd, xs);
                for (int e = tid; e < n * nrhs; e += BT) B[(size_t)(e % n) + (size_t)(e / n) * n] = Xw[e];
            }
        } else {
            // FP32 factors of A
            for (int e = tid; e < n * n; e += BT) Fw[e] = (float)A[e];
            __syncthreads();
            inf = lu ? lu_factor<float, NB, false>(n, Fw, pw, (float*)P, (float*)Us, spiv, s)
                     : chol_factor<float, NB, false>(n, Fw, (float*)P, s);
            bool done = false;
            const double tol = (double)n * 0x1.0p-52;
            if (inf == 0) {
                if (lu) piv_to_perm(n, pw, perm);
                for (int e = tid; e < n * nrhs; e += BT) Xw[e] = 0.0;
                for (int e = tid; e < n * nrhs; e += BT) Rw[e] = B[(size_t)(e % n) + (size_t)(e / n) * n];
                __syncthreads();
                double anorm = -1.0;
                for (its = 1; its <= IR_MAX; its++) {
                    solve_with<float, NB>(n, nrhs, lu, Fw, perm, Rw, Dw, (float*)Dd, xs);
                    for (int e = tid; e < n * nrhs; e += BT) Xw[e] += Dw[e];
                    __syncthreads();
                    const double be = residual_berr(n, nrhs, !lu, A, B, Xw, Rw, anorm, part);
                    if (be <= tol) {
                        done = true;
                        break;
                    }
                }
            }
            if (!done) {
                // the FP64 path on a copy of A
                for (int e = tid; e < n * n; e += BT) Fd[e] = A[e];
                __syncthreads();
                inf = lu ? lu_factor<double, NB, false>(n, Fd, pw, P, Us, spiv, s)
                         : chol_factor<double, NB, false>(n, Fd, P, s);
                its = -1;
                if (inf == 0) {
                    if (lu) piv_to_perm(n, pw, perm);
                    solve_with<double, NB>(n, nrhs, lu, Fd, perm, B, Xw, Dd, xs);
                }
            }
            __syncthreads();
            if (inf == 0)
                for (int e = tid; e < n * nrhs; e += BT) B[(size_t)(e % n) + (size_t)(e / n) * n] = Xw[e];
            if (lu && tid == 0 && inf == 0)
                for (int j = 0; j < n; j++) piva[d.piv + j] = pw[j];
        }
        if (tid == 0) {
            info[sys] = inf;
            if (iters) iters[sys] = its;
        }
        __syncthreads();
    }
}

__global__ void k_zero1(int* p) {
    if (threadIdx.x == 0) *p = 0;
}

// ---------------------------------------------------------------- host side

double sys_cost(const Desc& d) {
    const double n = d.n;
    return (d.kind == BOLA_HIP_BATCH_LU ? 2.0 * n * n * n / 3.0 : n * n * n / 3.0) + 2.0 * n * n * d.nrhs;
}

using KernelFn = void (*)(int, const Desc*, const int*, double*, int32_t*, int32_t*, int32_t*, int*, char*);

KernelFn pick_kernel(int variant, int nb) {
    if (variant == BOLA_HIP_BATCH_FP64) return nb == 8 ? batch_solve<8, 0> : batch_solve<16, 0>;
    return nb == 8 ? batch_solve<8, 1> : batch_solve<16, 1>;
}

// panels of 8 or 16 columns (the FP64 panel of 16 x 512 fills 64 KiB of LDS); up to four
// workgroups per CU (LDS admits two or three resident, the rest start as those finish)
bool valid_config(const bola_hip_batch_config& c) {
    return (c.nb == 8 || c.nb == 16) && c.wg_per_cu >= 1 && c.wg_per_cu <= 4;
}

// device copies of the descriptors and the largest-first order, plus the counter
struct Staged {
    Desc* desc = nullptr;
    int* order = nullptr;
    int* next = nullptr;
};

int check_descs(bola_hip_handle_s* h, size_t count, const Desc* desc) {
    for (size_t i = 0; i < count; i++) {
        const Desc& d = desc[i];
        if (d.n < 1 || d.n > BMAX || d.nrhs < 1 || d.nrhs > BNR ||
            (d.kind != BOLA_HIP_BATCH_LU && d.kind != BOLA_HIP_BATCH_CHOL) || d.a < 0 || d.b < 0 ||
            (d.kind == BOLA_HIP_BATCH_LU && d.piv < 0)) {
            set_error(h, "batch descriptor " + std::to_string(i) +
                             " is invalid (1 <= n <= 512, 1 <= nrhs <= 16, kind LU or CHOL, offsets >= 0)");
            return -4;
        }
    }
    return 0;
}

bool any_lu(size_t count, const Desc* desc) {
    for (size_t i = 0; i < count; i++)
        if (desc[i].kind == BOLA_HIP_BATCH_LU) return true;
    return false;
}

std::vector<int> largest_first(size_t count, const Desc* desc) {
    std::vector<int> ord(count);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return sys_cost(desc[a]) > sys_cost(desc[b]); });
    return ord;
}

// Uploads into buf (allocated by the caller): [desc | order | counter]
size_t staged_bytes(size_t count) { return ((count * sizeof(Desc) + 255) & ~(size_t)255) + ((count * 4 + 255) & ~(size_t)255) + 256; }

Staged stage_descs(hipStream_t st, char* buf, size_t count, const Desc* desc) {
    Staged s;
    s.desc = (Desc*)buf;
    s.order = (int*)(buf + ((count * sizeof(Desc) + 255) & ~(size_t)255));
    s.next = (int*)((char*)s.order + ((count * 4 + 255) & ~(size_t)255));
    const std::vector<int> ord = largest_first(count, desc);
    BOLA_CHECK(hipMemcpyAsync(s.desc, desc, count * sizeof(Desc), hipMemcpyHostToDevice, st));
    BOLA_CHECK(hipMemcpyAsync(s.order, ord.data(), count * 4, hipMemcpyHostToDevice, st));
    BOLA_CHECK(hipStreamSynchronize(st));  // the host vectors go out of scope
    return s;
}

unsigned grid_for(bola_hip_handle_s* h, const bola_hip_batch_config& c, size_t count) {
    return (unsigned)std::max<size_t>(1, std::min<size_t>(count, (size_t)h->num_cus * c.wg_per_cu));
}

void launch_batch(hipStream_t st, int variant, const bola_hip_batch_config& c, unsigned grid, size_t count,
                  const Staged& s, double* arena, int32_t* piva, int32_t* info, int32_t* iters, char* ws) {
    hipLaunchKernelGGL(k_zero1, dim3(1), dim3(64), 0, st, s.next);  // a kernel, not a memset: graph-safe
    hipLaunchKernelGGL(pick_kernel(variant, c.nb), dim3(grid), dim3(BT), 0, st, (int)count, s.desc, s.order, arena,
                       piva, info, iters, s.next, ws);
    BOLA_LAUNCH_CHECK();
}

long tune_key(int variant, size_t count, const Desc* desc) {
    int nmax = 0;
    for (size_t i = 0; i < count; i++) nmax = std::max(nmax, desc[i].n);
    int nb = 0;  // log2 bucket of the largest order and of the count
    while ((64 << nb) < nmax && nb < 8) nb++;
    int cb = 0;
    while ((size_t(256) << cb) < count && cb < 16) cb++;
    return ((long)variant << 16) | (nb << 8) | cb;
}

std::map<long, bola_hip_batch_config>& tuned_map(bola_hip_handle_s* h) {
    static std::map<bola_hip_handle_s*, std::map<long, bola_hip_batch_config>> all;
    return all[h];
}

}  // namespace
}  // namespace bola

using namespace bola;

struct bola_hip_batch_graph_s {
    bola_hip_handle_s* h = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    char* buf = nullptr;  // staged descriptors
    char* ws = nullptr;   // per-workgroup scratch
    hipStream_t cap = nullptr;  // the capture stream
};

extern "C" {

size_t bola_hip_batch_workspace_bytes(bola_hip_handle_t h, const bola_hip_batch_config* cfg) {
    const int wpc = cfg && valid_config(*cfg) ? cfg->wg_per_cu : 4;
    return (size_t)(h ? h->num_cus : 256) * wpc * kWsPerWg;
}

int bola_hip_batch_solve_dev(bola_hip_handle_t h, int variant, size_t count, const bola_hip_batch_desc* desc,
                             double* arena, int32_t* piv_arena, int32_t* info, int32_t* iters,
                             const bola_hip_batch_config* cfg) {
    if (!h) h = default_handle();
    BOLA_API_BEGIN
    if (variant != BOLA_HIP_BATCH_FP64 && variant != BOLA_HIP_BATCH_FP32_IR) {
        set_error(h, "Argument 2 to batch_solve had an illegal value");
        return -2;
    }
    if (count == 0) return 0;
    if (!desc || !arena || !info) {
        set_error(h, "batch_solve needs descriptors, an arena and an info array");
        return -4;
    }
    if (int r = check_descs(h, count, desc)) return r;
    if (!piv_arena && any_lu(count, desc)) {
        set_error(h, "Argument 6 to batch_solve had an illegal value (LU systems need a pivot arena)");
        return -6;
    }
    bola_hip_batch_config c{16, 2};
    if (cfg) {
        if (!valid_config(*cfg)) {
            set_error(h, "Argument 9 to batch_solve had an illegal value (nb 8 or 16; 1 <= wg_per_cu <= 4)");
            return -9;
        }
        c = *cfg;
    } else {
        auto& m = tuned_map(h);
        auto it = m.find(tune_key(variant, count, desc));
        if (it != m.end()) c = it->second;
    }
    const unsigned grid = grid_for(h, c, count);
    const size_t sb = staged_bytes(count);
    char* base = (char*)h->bws.get(sb + (size_t)grid * kWsPerWg + 256);
    const Staged s = stage_descs(h->stream, base, count, desc);
    {
        Phase ph(h, variant == BOLA_HIP_BATCH_FP64 ? "BATCH_FP64" : "BATCH_FP32_IR");
        launch_batch(h->stream, variant, c, grid, count, s, arena, piv_arena, info, iters, base + sb);
    }
    if (h->profile) collect_profile(h);
    return 0;
    BOLA_API_END(h)
}

int bola_hip_batch_tune_dev(bola_hip_handle_t h, int variant, size_t count, const bola_hip_batch_desc* desc,
                            const double* arena, bola_hip_batch_config* out) {
    if (!h) h = default_handle();
    BOLA_API_BEGIN
    if (variant != BOLA_HIP_BATCH_FP64 && variant != BOLA_HIP_BATCH_FP32_IR) {
        set_error(h, "Argument 2 to batch_tune had an illegal value");
        return -2;
    }
    if (count == 0 || !desc || !arena) {
        set_error(h, "batch_tune needs a non-empty batch");
        return -3;
    }
    if (int r = check_descs(h, count, desc)) return r;
    // a sample of up to 2048 systems (every k-th, so the size mix is kept), copied into a
    // private arena so that the caller's data stays untouched
    const size_t stride = std::max<size_t>(1, count / 2048);
    std::vector<Desc> sd;
    size_t off = 0, poff = 0;
    std::vector<std::pair<size_t, size_t>> copies;  // (src offset, elements)
    for (size_t i = 0; i < count; i += stride) {
        Desc d = desc[i];
        const size_t na = (size_t)d.n * d.n, nbb = (size_t)d.n * d.nrhs;
        copies.push_back({(size_t)d.a, na});
        copies.push_back({(size_t)d.b, nbb});
        d.a = (int64_t)off;
        d.b = (int64_t)(off + na);
        d.piv = (int64_t)poff;
        off += na + nbb;
        poff += d.n;
        sd.push_back(d);
    }
    const size_t sc = sd.size();
    double *pristine = nullptr, *work = nullptr;
    int32_t *piv = nullptr, *inf = nullptr;
    char *buf = nullptr, *ws = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const size_t sb = staged_bytes(sc);
    const size_t wsb = (size_t)h->num_cus * 4 * kWsPerWg;
    BOLA_CHECK(hipMalloc(&pristine, off * 8));
    BOLA_CHECK(hipMalloc(&work, off * 8));
    BOLA_CHECK(hipMalloc(&piv, std::max<size_t>(poff, 1) * 4));
    BOLA_CHECK(hipMalloc(&inf, sc * 4));
    BOLA_CHECK(hipMalloc(&buf, sb));
    BOLA_CHECK(hipMalloc(&ws, wsb));
    BOLA_CHECK(hipEventCreate(&e0));
    BOLA_CHECK(hipEventCreate(&e1));
    size_t dst = 0;
    for (auto& cp : copies) {
        BOLA_CHECK(hipMemcpyAsync(pristine + dst, arena + cp.first, cp.second * 8, hipMemcpyDeviceToDevice, h->stream));
        dst += cp.second;
    }
    const Staged s = stage_descs(h->stream, buf, sc, sd.data());
    bola_hip_batch_config best{16, 2};
    float best_ms = 1e30f;
    const int nbs[2] = {8, 16};
    const int wpcs[3] = {1, 2, 4};
    for (int nb : nbs)
        for (int wpc : wpcs) {
            const bola_hip_batch_config c{nb, wpc};
            const unsigned grid = grid_for(h, c, sc);
            float ms = 0.f;
            for (int rep = 0; rep < 2; rep++) {  // the first run warms up
                BOLA_CHECK(hipMemcpyAsync(work, pristine, off * 8, hipMemcpyDeviceToDevice, h->stream));
                BOLA_CHECK(hipEventRecord(e0, h->stream));
                launch_batch(h->stream, variant, c, grid, sc, s, work, piv, inf, nullptr, ws);
                BOLA_CHECK(hipEventRecord(e1, h->stream));
                BOLA_CHECK(hipEventSynchronize(e1));
                BOLA_CHECK(hipEventElapsedTime(&ms, e0, e1));
            }
            if (ms < best_ms) {
                best_ms = ms;
                best = c;
            }
        }
    BOLA_CHECK(hipStreamSynchronize(h->stream));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(pristine);
    hipFree(work);
    hipFree(piv);
    hipFree(inf);
    hipFree(buf);
    hipFree(ws);
    tuned_map(h)[tune_key(variant, count, desc)] = best;
    if (out) *out = best;
    return 0;
    BOLA_API_END(h)
}

int bola_hip_batch_graph_create(bola_hip_handle_t h, int variant, size_t count, const bola_hip_batch_desc* desc,
                                double* arena, int32_t* piv_arena, int32_t* info, int32_t* iters,
                                const bola_hip_batch_config* cfg, bola_hip_batch_graph_t* out) {
    if (!h) h = default_handle();
    BOLA_API_BEGIN
    if (!out) {
        set_error(h, "batch_graph_create needs an output pointer");
        return -10;
    }
    *out = nullptr;
    if (variant != BOLA_HIP_BATCH_FP64 && variant != BOLA_HIP_BATCH_FP32_IR) {
        set_error(h, "Argument 2 to batch_graph_create had an illegal value");
        return -2;
    }
    if (count == 0 || !desc || !arena || !info) {
        set_error(h, "batch_graph_create needs a non-empty batch");
        return -3;
    }
    if (int r = check_descs(h, count, desc)) return r;
    if (!piv_arena && any_lu(count, desc)) {
        set_error(h, "Argument 6 to batch_graph_create had an illegal value (LU systems need a pivot arena)");
        return -6;
    }
    bola_hip_batch_config c{16, 2};
    if (cfg) {
        if (!valid_config(*cfg)) {
            set_error(h, "Argument 9 to batch_graph_create had an illegal value");
            return -9;
        }
        c = *cfg;
    } else {
        auto& m = tuned_map(h);
        auto it = m.find(tune_key(variant, count, desc));
        if (it != m.end()) c = it->second;
    }
    auto* g = new bola_hip_batch_graph_s;
    g->h = h;
    const unsigned grid = grid_for(h, c, count);
    const size_t sb = staged_bytes(count);
    try {
        BOLA_CHECK(hipMalloc(&g->buf, sb));
        BOLA_CHECK(hipMalloc(&g->ws, (size_t)grid * kWsPerWg));
        const Staged s = stage_descs(h->stream, g->buf, count, desc);
        // captured on a stream of its own (the handle's may be the legacy default stream,
        // which cannot be captured); replays go to the handle's stream
        BOLA_CHECK(hipStreamCreateWithFlags(&g->cap, hipStreamNonBlocking));
        BOLA_CHECK(hipStreamBeginCapture(g->cap, hipStreamCaptureModeThreadLocal));
        try {
            launch_batch(g->cap, variant, c, grid, count, s, arena, piv_arena, info, iters, g->ws);
        } catch (...) {
            hipGraph_t tmp = nullptr;
            hipStreamEndCapture(g->cap, &tmp);
            if (tmp) hipGraphDestroy(tmp);
            throw;
        }
        BOLA_CHECK(hipStreamEndCapture(g->cap, &g->graph));
        BOLA_CHECK(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    } catch (...) {
        if (g->exec) hipGraphExecDestroy(g->exec);
        if (g->graph) hipGraphDestroy(g->graph);
        if (g->cap) hipStreamDestroy(g->cap);
        if (g->buf) hipFree(g->buf);
        if (g->ws) hipFree(g->ws);
        delete g;
        throw;
    }
    *out = g;
    return 0;
    BOLA_API_END(h)
}

int bola_hip_batch_graph_launch(bola_hip_batch_graph_t g) {
    if (!g) return -1;
    bola_hip_handle_s* h = g->h;
    BOLA_API_BEGIN
    BOLA_CHECK(hipGraphLaunch(g->exec, h->stream));
    return 0;
    BOLA_API_END(h)
}

int bola_hip_batch_graph_destroy(bola_hip_batch_graph_t g) {
    if (!g) return 0;
    hipStreamSynchronize(g->h->stream);
    if (g->exec) hipGraphExecDestroy(g->exec);
    if (g->graph) hipGraphDestroy(g->graph);
    if (g->cap) hipStreamDestroy(g->cap);
    if (g->buf) hipFree(g->buf);
    if (g->ws) hipFree(g->ws);
    delete g;
    return 0;
}

}  // extern "C"
