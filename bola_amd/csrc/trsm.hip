// TRSM on gfx950 (trsm.rs:33-376): the reference's recursive structure with its exact
// base cases, the off-diagonal updates on the MFMA DGEMM.
//
// The reference halves the triangle until it is at most 64 wide (m for side 'L', n for
// side 'R') and updates the other half with gemm(alpha = -1, beta = 1).  The base cases
// are small triangular solves with specific roundings: left/no-transpose solves divide
// by the diagonal and update with daxpy (fused on 4-aligned chunks, unfused tail, and
// skipped for a zero right-hand-side entry); left/transpose solves accumulate
// `temp -= a*b` unfused then divide; right solves subtract unfused (skipping zero
// coefficients) then multiply by the reciprocal of the diagonal.  The base-case kernels
// below run those exact sequences, one thread per independent right-hand-side column
// (side 'L') or row (side 'R'), with the triangle and the thread's vector in LDS.
#include "common.hpp"

namespace bola {

namespace {

constexpr int TB = 64;      // threads per base-case block (one wavefront)
constexpr int MAXT = 64;    // base-case triangle size (trsm.rs:8)

// variant ids
enum { LNL = 0, LNU, LTL, LTU, RNL, RNU, RTL, RTU };

template <int V>
__global__ void __launch_bounds__(TB) trsm_base(int unit, int t, int nvec, const double* __restrict__ A, size_t lda,
                                                double* __restrict__ B, size_t ldb, const int32_t* __restrict__ abort) {
    if (abort && *abort != 0) return;
    // t: triangle order (<= 64); nvec: number of independent vectors (columns for 'L',
    // rows for 'R'); vector v of B: 'L' -> B[0..t, v], 'R' -> B[v, 0..t]
    __shared__ double sA[MAXT * (MAXT + 1)];   // sA[r * (MAXT+1) + c] = A[r, c]
    __shared__ double sV[MAXT * TB];           // sV[i * TB + lane] = element i of the lane's vector
    const int lane = threadIdx.x;
    for (int e = lane; e < t * t; e += TB) {
        int r = e % t, c = e / t;
        sA[r * (MAXT + 1) + c] = A[r + (size_t)c * lda];
    }
    const long v = (long)blockIdx.x * TB + lane;
    const bool active = v < nvec;
    constexpr bool LEFT = V <= LTU;
    if (active) {
        for (int i = 0; i < t; i++)
            sV[i * TB + lane] = LEFT ? B[i + (size_t)v * ldb] : B[v + (size_t)i * ldb];
    }
    __syncthreads();
    if (active) {
#define AT(r, c) sA[(r) * (MAXT + 1) + (c)]
#define X(i) sV[(i) * TB + lane]
        if (V == LNL) {
            for (int k = 0; k < t; k++) {
                double bk = X(k);
                if (bk != 0.0) {
                    double u = unit ? bk : bk / AT(k, k);
                    X(k) = u;
                    double al = -u;
                    int len = t - k - 1;
                    if (len > 0 && al != 0.0) {
                        int nf = len & ~3;
                        for (int i = 0; i < len; i++) {
                            int r = k + 1 + i;
                            double y = X(r), x = AT(r, k);
                            X(r) = (i < nf) ? fma(al, x, y) : y + al * x;
                        }
                    }
                }
            }
        } else if (V == LNU) {
            for (int k = t - 1; k >= 0; k--) {
                double bk = X(k);
                if (bk != 0.0) {
                    double u = unit ? bk : bk / AT(k, k);
                    X(k) = u;
                    double al = -u;
                    int len = k;
                    if (len > 0 && al != 0.0) {
                        int nf = len & ~3;
                        for (int i = 0; i < len; i++) {
                            double y = X(i), x = AT(i, k);
                            X(i) = (i < nf) ? fma(al, x, y) : y + al * x;
                        }
                    }
                }
            }
        } else if (V == LTL) {
            for (int i = t - 1; i >= 0; i--) {
                double temp = X(i);
                for (int k = i + 1; k < t; k++) temp = temp - AT(k, i) * X(k);
                if (!unit) temp = temp / AT(i, i);
                X(i) = temp;
            }
        } else if (V == LTU) {
            for (int i = 0; i < t; i++) {
                double temp = X(i);
                for (int k = 0; k < i; k++) temp = temp - AT(k, i) * X(k);
                if (!unit) temp = temp / AT(i, i);
                X(i) = temp;
            }
        } else if (V == RNL) {
            for (int j = t - 1; j >= 0; j--) {
                double bj = X(j);
                for (int k = j + 1; k < t; k++) {
                    double a = AT(k, j);
                    if (a != 0.0) bj = bj - a * X(k);
                }
                if (!unit) bj = bj * (1.0 / AT(j, j));
                X(j) = bj;
            }
        } else if (V == RNU) {
            for (int j = 0; j < t; j++) {
                double bj = X(j);
                for (int k = 0; k < j; k++) {
                    double a = AT(k, j);
                    if (a != 0.0) bj = bj - a * X(k);
                }
                if (!unit) bj = bj * (1.0 / AT(j, j));
                X(j) = bj;
            }
        } else if (V == RTL) {
            for (int j = 0; j < t; j++) {
                double bj = X(j);
                for (int k = 0; k < j; k++) {
                    double a = AT(j, k);
                    if (a != 0.0) bj = bj - a * X(k);
                }
                if (!unit) bj = bj * (1.0 / AT(j, j));
                X(j) = bj;
            }
        } else {  // RTU
            for (int j = t - 1; j >= 0; j--) {
                double bj = X(j);
                for (int k = j + 1; k < t; k++) {
                    double a = AT(j, k);
                    if (a != 0.0) bj = bj - a * X(k);
                }
                if (!unit) bj = bj * (1.0 / AT(j, j));
                X(j) = bj;
            }
        }
        for (int i = 0; i < t; i++) {
            if (LEFT) B[i + (size_t)v * ldb] = X(i);
            else B[v + (size_t)i * ldb] = X(i);
        }
#undef AT
#undef X
    }
}

template <int V>
void base(hipStream_t st, int unit, size_t t, size_t nvec, const double* a, size_t lda, double* b, size_t ldb,
          const int32_t* abort) {
    if (t == 0 || nvec == 0) return;
    unsigned blocks = (unsigned)((nvec + TB - 1) / TB);
    hipLaunchKernelGGL((trsm_base<V>), dim3(blocks), dim3(TB), 0, st, unit, (int)t, (int)nvec, a, lda, b, ldb, abort);
    BOLA_LAUNCH_CHECK();
}

constexpr size_t STOP = 64;  // trsm.rs:8

void lnl(hipStream_t st, int unit, size_t m, size_t n, const double* a, size_t lda, double* b, size_t ldb,
         const int32_t* abort) {
    if (m <= STOP) return base<LNL>(st, unit, m, n, a, lda, b, ldb, abort);
    size_t m1 = m / 2, m2 = m - m1;
    lnl(st, unit, m1, n, a, lda, b, ldb, abort);
    launch_dgemm(st, 'N', 'N', m2, n, m1, -1.0, a + m1, lda, b, ldb, 1.0, b + m1, ldb, abort);
    lnl(st, unit, m2, n, a + m1 + m1 * lda, lda, b + m1, ldb, abort);
}
void lnu(hipStream_t st, int unit, size_t m, size_t n, const double* a, size_t lda, double* b, size_t ldb,
         const int32_t* abort) {
    if (m <= STOP) return base<LNU>(st, unit, m, n, a, lda, b, ldb, abort);
    size_t m1 = m / 2, m2 = m - m1;
    lnu(st, unit, m2, n, a + m1 + m1 * lda, lda, b + m1, ldb, abort);
    launch_dgemm(st, 'N', 'N', m1, n, m2, -1.0, a + m1 * lda, lda, b + m1, ldb, 1.0, b, ldb, abort);
    lnu(st, unit, m1, n, a, lda, b, ldb, abort);
}
void ltl(hipStream_t st, int unit, size_t m, size_t n, const double* a, size_t lda, double* b, size_t ldb,
         const int32_t* abort) {
    if (m <= STOP) return base<LTL>(st, unit, m, n, a, lda, b, ldb, abort);
    size_t m1 = m / 2, m2 = m - m1;
    ltl(st, unit, m2, n, a + m1 + m1 * lda, lda, b + m1, ldb, abort);
    launch_dgemm(st, 'T', 'N', m1, n, m2, -1.0, a + m1, lda, b + m1, ldb, 1.0, b, ldb, abort);
    ltl(st, unit, m1, n, a, lda, b, ldb, abort);
}
void ltu(hipStream_t st, int unit, size_t m, size_t n, const double* a, size_t lda, double* b, size_t ldb,
         const int32_t* abort) {
    if (m <= STOP) return base<LTU>(st, unit, m, n, a, lda, b, ldb, abort);
    size_t m1 = m / 2, m2 = m - m1;
    ltu(st, unit, m1, n, a, lda, b, ldb, abort);
    launch_dgemm(st, 'T', 'N', m2, n, m1, -1.0, a + m1 * lda, lda, b, ldb, 1.0, b + m1, ldb, abort);
    ltu(st, unit, m2, n, a + m1 + m1 * lda, lda, b + m1, ldb, abort);
}
void rnl(hipStream_t st, int unit, size_t m, size_t n, const double* a, size_t lda, double* b, size_t ldb,
         const int32_t* abort) {
    if (n <= STOP) return base<RNL>(st, unit, n, m, a, lda, b, ldb, abort);
    size_t n1 = n / 2, n2 = n - n1;
    rnl(st, unit, m, n2, a + n1 + n1 * lda, lda, b + n1 * ldb, ldb, abort);
    launch_dgemm(st, 'N', 'N', m, n1, n2, -1.0, b + n1 * ldb, ldb, a + n1, lda, 1.0, b, ldb, abort);
    rnl(st, unit, m, n1, a, lda, b, ldb, abort);
}
void rnu(hipStream_t st, int unit, size_t m, size_t n, const double* a, size_t lda, double* b, size_t ldb,
         const int32_t* abort) {
    if (n <= STOP) return base<RNU>(st, unit, n, m, a, lda, b, ldb, abort);
    size_t n1 = n / 2, n2 = n - n1;
    rnu(st, unit, m, n1, a, lda, b, ldb, abort);
    launch_dgemm(st, 'N', 'N', m, n2, n1, -1.0, b, ldb, a + n1 * lda, lda, 1.0, b + n1 * ldb, ldb, abort);
    rnu(st, unit, m, n2, a + n1 + n1 * lda, lda, b + n1 * ldb, ldb, abort);
}
void rtl(hipStream_t st, int unit, size_t m, size_t n, const double* a, size_t lda, double* b, size_t ldb,
         const int32_t* abort) {
    if (n <= STOP) return base<RTL>(st, unit, n, m, a, lda, b, ldb, abort);
    size_t n1 = n / 2, n2 = n - n1;
    rtl(st, unit, m, n1, a, lda, b, ldb, abort);
    launch_dgemm(st, 'N', 'T', m, n2, n1, -1.0, b, ldb, a + n1, lda, 1.0, b + n1 * ldb, ldb, abort);
    rtl(st, unit, m, n2, a + n1 + n1 * lda, lda, b + n1 * ldb, ldb, abort);
}
void rtu(hipStream_t st, int unit, size_t m, size_t n, const double* a, size_t lda, double* b, size_t ldb,
         const int32_t* abort) {
    if (n <= STOP) return base<RTU>(st, unit, n, m, a, lda, b, ldb, abort);
    size_t n1 = n / 2, n2 = n - n1;
    rtu(st, unit, m, n2, a + n1 + n1 * lda, lda, b + n1 * ldb, ldb, abort);
    launch_dgemm(st, 'N', 'T', m, n1, n2, -1.0, b + n1 * ldb, ldb, a + n1 * lda, lda, 1.0, b, ldb, abort);
    rtu(st, unit, m, n1, a, lda, b, ldb, abort);
}

bool eqc(char c, char w) { return c == w || c == (char)(w + 32); }

}  // namespace

void launch_dtrsm(hipStream_t st, char side, char uplo, char transa, char diag, size_t m, size_t n, double alpha,
                  const double* a, size_t lda, double* b, size_t ldb, const int32_t* abort) {
    if (m == 0 || n == 0) return;
    if (alpha != 1.0) launch_scale(st, b, ldb, m, n, alpha);
    bool lside = eqc(side, 'L'), upper = eqc(uplo, 'U'), notrans = eqc(transa, 'N');
    int unit = eqc(diag, 'U') ? 1 : 0;
    if (lside) {
        if (notrans) {
            if (upper) lnu(st, unit, m, n, a, lda, b, ldb, abort);
            else lnl(st, unit, m, n, a, lda, b, ldb, abort);
        } else if (upper) ltu(st, unit, m, n, a, lda, b, ldb, abort);
        else ltl(st, unit, m, n, a, lda, b, ldb, abort);
    } else if (notrans) {
        if (upper) rnu(st, unit, m, n, a, lda, b, ldb, abort);
        else rnl(st, unit, m, n, a, lda, b, ldb, abort);
    } else if (upper) rtu(st, unit, m, n, a, lda, b, ldb, abort);
    else rtl(st, unit, m, n, a, lda, b, ldb, abort);
}

}  // namespace bola
