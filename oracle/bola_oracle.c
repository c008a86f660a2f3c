/*
 * bola_oracle.c -- CPU restatement of BOLA's x86_64 (AVX2+FMA) dense-solver path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle that the GPU library is checked
 * against, and the CPU baseline that the benchmark times beside the GPU.  The product
 * (bola_amd/, libbola_hip.so) never links, loads or calls it.  Only tests/, the smoke
 * check and the benchmark's cpu_baseline leg may use it.
 *
 * It restates, function by function, the algorithm of the Rust crate bola 0.1.2
 * (the src/ directory, one .rs file per routine): the same block constants, loop orders, thresholds, and the same placement
 * of fused (explicit _mm256_fmadd_pd) versus unfused (Rust never contracts a*b+c)
 * arithmetic.  Build with -ffp-contract=off so C's `a += t*x` stays unfused like Rust's.
 * The three rayon sites (gemm over 16-column blocks, laswp over 32-column chunks,
 * left-trsm base cases over right-hand-side columns) become OpenMP loops; every task
 * writes disjoint columns, so results do not depend on the thread count.
 *
 * Error returns: 0 = Ok; otherwise a nonzero code and the exact text of BOLA's
 * Err(String) in a thread-local buffer (bola_oracle_last_error()).
 */
#include <immintrin.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stddef.h>
#include <string.h>
#include <ctype.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define EXPORT __attribute__((visibility("default")))

static __thread char g_err[256];

EXPORT const char* bola_oracle_last_error(void) { return g_err; }

static size_t min_sz(size_t a, size_t b) { return a < b ? a : b; }
static int eq_ic(char c, char want) { return toupper((unsigned char)c) == want; }

/* ---------------------------------------------------------------- utilities.rs */

/* dot_product, utilities.rs:22-54 (x86_64). 4 FMA lanes, horizontal sum, unfused tail. */
EXPORT double bola_oracle_dot_product(size_t n, const double* x, size_t incx, const double* y, size_t incy) {
    if (n == 0) return 0.0;
    if (incx == 1 && incy == 1) {
        size_t n_chunks = n / 4;
        __m256d sum_vec = _mm256_setzero_pd();
        for (size_t i = 0; i < n_chunks; i++) {
            __m256d xv = _mm256_loadu_pd(x + i * 4);
            __m256d yv = _mm256_loadu_pd(y + i * 4);
            sum_vec = _mm256_fmadd_pd(xv, yv, sum_vec);
        }
        double lanes[4];
        _mm256_storeu_pd(lanes, sum_vec);
        /* temp_sum = 0.0; temp_sum += lanes.iter().sum() -- the fold order is l0,l1,l2,l3. */
        double hs = -0.0;
        hs = hs + lanes[0];
        hs = hs + lanes[1];
        hs = hs + lanes[2];
        hs = hs + lanes[3];
        double temp_sum = 0.0;
        temp_sum += hs;
        for (size_t i = n_chunks * 4; i < n; i++) temp_sum += x[i] * y[i];
        return temp_sum;
    } else {
        double temp_sum = 0.0;
        for (size_t i = 0; i < n; i++) temp_sum += x[i * incx] * y[i * incy];
        return temp_sum;
    }
}

/* daxpy, utilities.rs:107-132. y = fma(alpha, x, y) on 4-chunks, y += alpha*x tail. */
EXPORT void bola_oracle_daxpy(size_t n, double alpha, const double* x, size_t incx, double* y, size_t incy) {
    if (n == 0 || alpha == 0.0) return;
    if (incx == 1 && incy == 1) {
        size_t n_chunks = n / 4;
        __m256d av = _mm256_set1_pd(alpha);
        for (size_t i = 0; i < n_chunks; i++) {
            const double* xp = x + i * 4;
            double* yp = y + i * 4;
            _mm256_storeu_pd(yp, _mm256_fmadd_pd(av, _mm256_loadu_pd(xp), _mm256_loadu_pd(yp)));
        }
        for (size_t i = n_chunks * 4; i < n; i++) y[i] += alpha * x[i];
    } else {
        for (size_t i = 0; i < n; i++) y[i * incy] += alpha * x[i * incx];
    }
}

/* ---------------------------------------------------------------- scal.rs */

/* scal_kernel, scal.rs:10-22: multiply 4 at a time; returns the count processed. */
EXPORT size_t bola_oracle_scal_kernel(size_t m, double inv_diag, double* col) {
    __m256d iv = _mm256_set1_pd(inv_diag);
    size_t m_chunks = m / 4;
    double* p = col;
    for (size_t i = 0; i < m_chunks; i++) {
        _mm256_storeu_pd(p, _mm256_mul_pd(_mm256_loadu_pd(p), iv));
        p += 4;
    }
    return m_chunks * 4;
}

/* scal, scal.rs:45-56 */
EXPORT void bola_oracle_scal(size_t n, double alpha, double* x, size_t incx) {
    if (incx == 1) {
        size_t processed = bola_oracle_scal_kernel(n, alpha, x);
        for (size_t i = processed; i < n; i++) x[i] *= alpha;
    } else {
        for (size_t i = 0; i < n; i++) x[i * incx] *= alpha;
    }
}

/* ---------------------------------------------------------------- idamax.rs */

/* idamax, idamax.rs:23-133 (x86_64 variant): 16 (value, index) accumulators, one per
 * residue i mod 16, strict ordered '>' (NaN never wins), merged vector 0<-1<-2<-3
 * lane-wise, then lanes 0<-1<-2<-3, then the scalar tail. */
EXPORT size_t bola_oracle_idamax(size_t n, const double* x) {
    if (n < 16) {
        double max_val = -1.0;
        size_t max_idx = 0;
        for (size_t i = 0; i < n; i++) {
            double v = fabs(x[i]);
            if (v > max_val) { max_val = v; max_idx = i; }
        }
        return max_idx;
    }
    double mv[4][4], mi[4][4]; /* [vector g][lane l] */
    for (int g = 0; g < 4; g++)
        for (int l = 0; l < 4; l++) { mv[g][l] = 0.0; mi[g][l] = 0.0; }
    size_t n_chunks = n / 16;
    for (size_t c = 0; c < n_chunks; c++) {
        for (int g = 0; g < 4; g++) {
            for (int l = 0; l < 4; l++) {
                size_t i = c * 16 + (size_t)g * 4 + (size_t)l;
                double v = fabs(x[i]);
                if (v > mv[g][l]) { mv[g][l] = v; mi[g][l] = (double)i; }
            }
        }
    }
    for (int g = 1; g < 4; g++)
        for (int l = 0; l < 4; l++)
            if (mv[g][l] > mv[0][l]) { mv[0][l] = mv[g][l]; mi[0][l] = mi[g][l]; }
    double max_val = mv[0][0];
    size_t max_idx = (size_t)mi[0][0];
    for (int l = 1; l < 4; l++)
        if (mv[0][l] > max_val) { max_val = mv[0][l]; max_idx = (size_t)mi[0][l]; }
    for (size_t i = n_chunks * 16; i < n; i++) {
        double v = fabs(x[i]);
        if (v > max_val) { max_val = v; max_idx = i; }
    }
    return max_idx;
}

/* ---------------------------------------------------------------- laswp.rs */

#define LASWP_PARALLEL_THRESHOLD 256

static void laswp_cols(double* a, size_t lda, size_t c0, size_t c1, size_t ix0, int32_t i1, int32_t i2,
                       int32_t i_inc, int32_t incx, const int32_t* ipiv) {
    for (size_t k = c0; k < c1; k++) {
        double* col = a + k * lda;
        size_t ix = ix0;
        int32_t i = i1;
        while (i_inc > 0 ? i <= i2 : i >= i2) {
            size_t ip = (size_t)ipiv[ix];
            if (ip != (size_t)i) {
                double t = col[i - 1];
                col[i - 1] = col[ip - 1];
                col[ip - 1] = t;
            }
            ix = (size_t)((int64_t)ix + incx);
            i += i_inc;
        }
    }
}

/* laswp, laswp.rs:22-132.  Row interchanges i <-> ipiv[ix]-1 (1-based pivots) applied in
 * order for i = k1..k2 (reverse when incx < 0).  Pure moves; the column partition
 * (4-column unroll, 32-column rayon chunks, row-outer tail) does not change results. */
EXPORT void bola_oracle_laswp(size_t n, double* a, size_t lda, int32_t k1, int32_t k2, const int32_t* ipiv,
                              int32_t incx) {
    if (n == 0 || incx == 0 || k1 > k2) return;
    size_t ix0;
    int32_t i1, i2, i_inc;
    if (incx > 0) {
        ix0 = (size_t)(k1 - 1); i1 = k1; i2 = k2; i_inc = 1;
    } else {
        size_t start = (size_t)(k1 - 1);
        size_t num = (size_t)(k2 - k1);
        ix0 = start + num * (size_t)(incx < 0 ? -(int64_t)incx : incx);
        i1 = k2; i2 = k1; i_inc = -1;
    }
    if (n < LASWP_PARALLEL_THRESHOLD) {
        laswp_cols(a, lda, 0, n, ix0, i1, i2, i_inc, incx, ipiv);
        return;
    }
    size_t n32 = (n / 32) * 32;
    if (n32 != 0) {
        long nchunks = (long)(n32 / 32);
#pragma omp parallel for schedule(static)
        for (long c = 0; c < nchunks; c++)
            laswp_cols(a, lda, (size_t)c * 32, (size_t)c * 32 + 32, ix0, i1, i2, i_inc, incx, ipiv);
    }
    if (n32 != n) laswp_cols(a, lda, n32, n, ix0, i1, i2, i_inc, incx, ipiv);
}

/* ---------------------------------------------------------------- ger.rs */

/* ger, ger.rs:28-60 with ger_kernel :67-90.  Column pairs; FMA on rows < 4*(m/4),
 * unfused tail rows; odd last column unfused; a pair with both alpha*y == 0 is skipped. */
EXPORT void bola_oracle_ger(size_t m, size_t n, double alpha, const double* x, const double* y, size_t ldy, double* a,
                            size_t lda) {
    size_t n_unrolled = n - (n % 2);
    for (size_t j = 0; j < n_unrolled; j += 2) {
        double t0 = alpha * y[j * ldy];
        double t1 = alpha * y[(j + 1) * ldy];
        if (t0 == 0.0 && t1 == 0.0) continue;
        double* a0 = a + j * lda;
        double* a1 = a + (j + 1) * lda;
        __m256d tv0 = _mm256_set1_pd(t0), tv1 = _mm256_set1_pd(t1);
        size_t m_chunks = m / 4;
        for (size_t c = 0; c < m_chunks; c++) {
            __m256d xv = _mm256_loadu_pd(x + c * 4);
            _mm256_storeu_pd(a0 + c * 4, _mm256_fmadd_pd(xv, tv0, _mm256_loadu_pd(a0 + c * 4)));
            _mm256_storeu_pd(a1 + c * 4, _mm256_fmadd_pd(xv, tv1, _mm256_loadu_pd(a1 + c * 4)));
        }
        for (size_t i = m_chunks * 4; i < m; i++) {
            a0[i] += t0 * x[i];
            a1[i] += t1 * x[i];
        }
    }
    if (n % 2 != 0) {
        size_t j = n - 1;
        double t = alpha * y[j * ldy];
        if (t != 0.0)
            for (size_t i = 0; i < m; i++) a[i + j * lda] += t * x[i];
    }
}

/* ---------------------------------------------------------------- gemm.rs */

#define KC 64
#define MC 32
#define NC 16
#define MR 8
#define NR 4
#define GEMM_PARALLEL_THRESHOLD 128

/* pack, gemm.rs:49-65.  Note the case-sensitive test trans == 'N'. */
static void pack(char trans, size_t rows, size_t cols, const double* src, size_t ld_src, double* dst) {
    if (trans == 'N') {
        for (size_t j = 0; j < cols; j++)
            for (size_t i = 0; i < rows; i++) dst[i + j * rows] = src[i + j * ld_src];
    } else {
        for (size_t j = 0; j < cols; j++)
            for (size_t i = 0; i < rows; i++) dst[i + j * rows] = src[j + i * ld_src];
    }
}

/* add_8x4_kernel, gemm.rs:232-270: acc = sum_l fma(a_il, b_lj, acc) from zero over the
 * slab, then C = fma(alpha, acc, C). */
static void add_8x4_kernel(size_t k, double alpha, const double* a, size_t lda, const double* b, size_t ldb, double* c,
                           size_t ldc) {
    __m256d c0[4], c1[4];
    for (int j = 0; j < 4; j++) { c0[j] = _mm256_setzero_pd(); c1[j] = _mm256_setzero_pd(); }
    for (size_t l = 0; l < k; l++) {
        __m256d a0 = _mm256_loadu_pd(a + l * lda);
        __m256d a1 = _mm256_loadu_pd(a + l * lda + 4);
        for (int j = 0; j < 4; j++) {
            __m256d bj = _mm256_set1_pd(b[l + (size_t)j * ldb]);
            c0[j] = _mm256_fmadd_pd(a0, bj, c0[j]);
            c1[j] = _mm256_fmadd_pd(a1, bj, c1[j]);
        }
    }
    __m256d av = _mm256_set1_pd(alpha);
    for (int j = 0; j < NR; j++) {
        double* cp = c + (size_t)j * ldc;
        _mm256_storeu_pd(cp, _mm256_fmadd_pd(av, c0[j], _mm256_loadu_pd(cp)));
        _mm256_storeu_pd(cp + 4, _mm256_fmadd_pd(av, c1[j], _mm256_loadu_pd(cp + 4)));
    }
}

/* add_scalar_kernel, gemm.rs:357-377: C += (alpha*b_lj)*a_il, unfused, l outer. */
static void add_scalar_kernel(size_t m, size_t n, size_t k, double alpha, const double* a, size_t lda, const double* b,
                              size_t ldb, double* c, size_t ldc) {
    for (size_t j = 0; j < n; j++)
        for (size_t l = 0; l < k; l++) {
            double temp = alpha * b[l + j * ldb];
            for (size_t i = 0; i < m; i++) c[i + j * ldc] += temp * a[i + l * lda];
        }
}

static void gemm_core(size_t jc, int notransa, int notransb, char transa, char transb, size_t m, size_t n, size_t k,
                      double alpha, const double* a, size_t lda, const double* b, size_t ldb, double* c, size_t ldc) {
    double a_packed[MC * KC];
    double b_packed[KC * NC];
    size_t nb = min_sz(n - jc, NC);
    for (size_t pc = 0; pc < k; pc += KC) {
        size_t pb = min_sz(k - pc, KC);
        const double* bp = notransb ? b + pc + jc * ldb : b + jc + pc * ldb;
        pack(transb, pb, nb, bp, ldb, b_packed);
        for (size_t ic = 0; ic < m; ic += MC) {
            size_t mb = min_sz(m - ic, MC);
            const double* ap = notransa ? a + ic + pc * lda : a + pc + ic * lda;
            pack(transa, mb, pb, ap, lda, a_packed);
            for (size_t j = 0; j < nb; j += NR)
                for (size_t i = 0; i < mb; i += MR) {
                    double* cb = c + ic + i + (jc + j) * ldc;
                    const double* bpk = b_packed + j * pb;
                    if (i + MR > mb || j + NR > nb)
                        add_scalar_kernel(min_sz(mb - i, MR), min_sz(nb - j, NR), pb, alpha, a_packed + i, mb, bpk, pb,
                                          cb, ldc);
                    else
                        add_8x4_kernel(pb, alpha, a_packed + i, mb, bpk, pb, cb, ldc);
                }
        }
    }
}

/* gemm, gemm.rs:88-211 */
EXPORT void bola_oracle_gemm(char transa, char transb, size_t m, size_t n, size_t k, double alpha, const double* a,
                             size_t lda, const double* b, size_t ldb, double beta, double* c, size_t ldc) {
    if (m == 0 || n == 0) return;
    if (beta == 0.0) {
        for (size_t j = 0; j < n; j++)
            for (size_t i = 0; i < m; i++) c[i + j * ldc] = 0.0;
    } else if (beta != 1.0) {
        for (size_t j = 0; j < n; j++)
            for (size_t i = 0; i < m; i++) c[i + j * ldc] *= beta;
    }
    if (alpha == 0.0) return;
    int notransa = eq_ic(transa, 'N');
    int notransb = eq_ic(transb, 'N');
    if (n > GEMM_PARALLEL_THRESHOLD) {
        long nblk = (long)((n + NC - 1) / NC);
#pragma omp parallel for schedule(dynamic, 1)
        for (long bi = 0; bi < nblk; bi++)
            gemm_core((size_t)bi * NC, notransa, notransb, transa, transb, m, n, k, alpha, a, lda, b, ldb, c, ldc);
    } else {
        for (size_t jc = 0; jc < n; jc += NC)
            gemm_core(jc, notransa, notransb, transa, transb, m, n, k, alpha, a, lda, b, ldb, c, ldc);
    }
}

/* ---------------------------------------------------------------- trsm.rs */

#define TRSM_RECURSION_STOP 64
#define TRSM_PARALLEL_THRESHOLD 32

/* trsm_lnl, trsm.rs:95-131 */
static void trsm_lnl(int unit, size_t m, size_t n, const double* a, size_t lda, double* b, size_t ldb) {
    if (m <= TRSM_RECURSION_STOP) {
#pragma omp parallel for schedule(static) if (n > TRSM_PARALLEL_THRESHOLD)
        for (long jj = 0; jj < (long)n; jj++) {
            double* bc = b + (size_t)jj * ldb;
            for (size_t kk = 0; kk < m; kk++) {
                double bkj = bc[kk];
                if (bkj != 0.0) {
                    double u = !unit ? bkj / a[kk + kk * lda] : bkj;
                    bc[kk] = u;
                    if (m > kk + 1) bola_oracle_daxpy(m - kk - 1, -u, a + kk + 1 + kk * lda, 1, bc + kk + 1, 1);
                }
            }
        }
        return;
    }
    size_t m1 = m / 2, m2 = m - m1;
    trsm_lnl(unit, m1, n, a, lda, b, ldb);
    bola_oracle_gemm('N', 'N', m2, n, m1, -1.0, a + m1, lda, b, ldb, 1.0, b + m1, ldb);
    trsm_lnl(unit, m2, n, a + m1 + m1 * lda, lda, b + m1, ldb);
}

/* trsm_lnu, trsm.rs:135-170 */
static void trsm_lnu(int unit, size_t m, size_t n, const double* a, size_t lda, double* b, size_t ldb) {
    if (m <= TRSM_RECURSION_STOP) {
#pragma omp parallel for schedule(static) if (n > TRSM_PARALLEL_THRESHOLD)
        for (long jj = 0; jj < (long)n; jj++) {
            double* bc = b + (size_t)jj * ldb;
            for (size_t kk = m; kk-- > 0;) {
                double bkj = bc[kk];
                if (bkj != 0.0) {
                    double u = !unit ? bkj / a[kk + kk * lda] : bkj;
                    bc[kk] = u;
                    if (kk > 0) bola_oracle_daxpy(kk, -u, a + kk * lda, 1, bc, 1);
                }
            }
        }
        return;
    }
    size_t m1 = m / 2, m2 = m - m1;
    trsm_lnu(unit, m2, n, a + m1 + m1 * lda, lda, b + m1, ldb);
    bola_oracle_gemm('N', 'N', m1, n, m2, -1.0, a + m1 * lda, lda, b + m1, ldb, 1.0, b, ldb);
    trsm_lnu(unit, m1, n, a, lda, b, ldb);
}

/* trsm_ltl, trsm.rs:174-209 */
static void trsm_ltl(int unit, size_t m, size_t n, const double* a, size_t lda, double* b, size_t ldb) {
    if (m <= TRSM_RECURSION_STOP) {
#pragma omp parallel for schedule(static) if (n > TRSM_PARALLEL_THRESHOLD)
        for (long jj = 0; jj < (long)n; jj++) {
            double* bc = b + (size_t)jj * ldb;
            for (size_t i = m; i-- > 0;) {
                double temp = bc[i];
                for (size_t kk = i + 1; kk < m; kk++) temp -= a[kk + i * lda] * bc[kk];
                if (!unit) temp /= a[i + i * lda];
                bc[i] = temp;
            }
        }
        return;
    }
    size_t m1 = m / 2, m2 = m - m1;
    trsm_ltl(unit, m2, n, a + m1 + m1 * lda, lda, b + m1, ldb);
    bola_oracle_gemm('T', 'N', m1, n, m2, -1.0, a + m1, lda, b + m1, ldb, 1.0, b, ldb);
    trsm_ltl(unit, m1, n, a, lda, b, ldb);
}

/* trsm_ltu, trsm.rs:213-248 */
static void trsm_ltu(int unit, size_t m, size_t n, const double* a, size_t lda, double* b, size_t ldb) {
    if (m <= TRSM_RECURSION_STOP) {
#pragma omp parallel for schedule(static) if (n > TRSM_PARALLEL_THRESHOLD)
        for (long jj = 0; jj < (long)n; jj++) {
            double* bc = b + (size_t)jj * ldb;
            for (size_t i = 0; i < m; i++) {
                double temp = bc[i];
                for (size_t kk = 0; kk < i; kk++) temp -= a[kk + i * lda] * bc[kk];
                if (!unit) temp /= a[i + i * lda];
                bc[i] = temp;
            }
        }
        return;
    }
    size_t m1 = m / 2, m2 = m - m1;
    trsm_ltu(unit, m1, n, a, lda, b, ldb);
    bola_oracle_gemm('T', 'N', m2, n, m1, -1.0, a + m1 * lda, lda, b, ldb, 1.0, b + m1, ldb);
    trsm_ltu(unit, m2, n, a + m1 + m1 * lda, lda, b + m1, ldb);
}

/* trsm_rnl, trsm.rs:252-280 */
static void trsm_rnl(int unit, size_t m, size_t n, const double* a, size_t lda, double* b, size_t ldb) {
    if (n <= TRSM_RECURSION_STOP) {
        for (size_t j = n; j-- > 0;) {
            double* bj = b + j * ldb;
            for (size_t kk = j + 1; kk < n; kk++) {
                double akj = a[kk + j * lda];
                if (akj != 0.0) {
                    const double* bk = b + kk * ldb;
                    for (size_t i = 0; i < m; i++) bj[i] -= akj * bk[i];
                }
            }
            if (!unit) {
                double inv = 1.0 / a[j + j * lda];
                for (size_t i = 0; i < m; i++) bj[i] *= inv;
            }
        }
        return;
    }
    size_t n1 = n / 2, n2 = n - n1;
    trsm_rnl(unit, m, n2, a + n1 + n1 * lda, lda, b + n1 * ldb, ldb);
    bola_oracle_gemm('N', 'N', m, n1, n2, -1.0, b + n1 * ldb, ldb, a + n1, lda, 1.0, b, ldb);
    trsm_rnl(unit, m, n1, a, lda, b, ldb);
}

/* trsm_rnu, trsm.rs:284-312 */
static void trsm_rnu(int unit, size_t m, size_t n, const double* a, size_t lda, double* b, size_t ldb) {
    if (n <= TRSM_RECURSION_STOP) {
        for (size_t j = 0; j < n; j++) {
            double* bj = b + j * ldb;
            for (size_t kk = 0; kk < j; kk++) {
                double akj = a[kk + j * lda];
                if (akj != 0.0) {
                    const double* bk = b + kk * ldb;
                    for (size_t i = 0; i < m; i++) bj[i] -= akj * bk[i];
                }
            }
            if (!unit) {
                double inv = 1.0 / a[j + j * lda];
                for (size_t i = 0; i < m; i++) bj[i] *= inv;
            }
        }
        return;
    }
    size_t n1 = n / 2, n2 = n - n1;
    trsm_rnu(unit, m, n1, a, lda, b, ldb);
    bola_oracle_gemm('N', 'N', m, n2, n1, -1.0, b, ldb, a + n1 * lda, lda, 1.0, b + n1 * ldb, ldb);
    trsm_rnu(unit, m, n2, a + n1 + n1 * lda, lda, b + n1 * ldb, ldb);
}

/* trsm_rtl, trsm.rs:316-344 */
static void trsm_rtl(int unit, size_t m, size_t n, const double* a, size_t lda, double* b, size_t ldb) {
    if (n <= TRSM_RECURSION_STOP) {
        for (size_t j = 0; j < n; j++) {
            double* bj = b + j * ldb;
            for (size_t kk = 0; kk < j; kk++) {
                double ajk = a[j + kk * lda];
                if (ajk != 0.0) {
                    const double* bk = b + kk * ldb;
                    for (size_t i = 0; i < m; i++) bj[i] -= ajk * bk[i];
                }
            }
            if (!unit) {
                double inv = 1.0 / a[j + j * lda];
                for (size_t i = 0; i < m; i++) bj[i] *= inv;
            }
        }
        return;
    }
    size_t n1 = n / 2, n2 = n - n1;
    trsm_rtl(unit, m, n1, a, lda, b, ldb);
    bola_oracle_gemm('N', 'T', m, n2, n1, -1.0, b, ldb, a + n1, lda, 1.0, b + n1 * ldb, ldb);
    trsm_rtl(unit, m, n2, a + n1 + n1 * lda, lda, b + n1 * ldb, ldb);
}

/* trsm_rtu, trsm.rs:348-376 */
static void trsm_rtu(int unit, size_t m, size_t n, const double* a, size_t lda, double* b, size_t ldb) {
    if (n <= TRSM_RECURSION_STOP) {
        for (size_t j = n; j-- > 0;) {
            double* bj = b + j * ldb;
            for (size_t kk = j + 1; kk < n; kk++) {
                double ajk = a[j + kk * lda];
                if (ajk != 0.0) {
                    const double* bk = b + kk * ldb;
                    for (size_t i = 0; i < m; i++) bj[i] -= ajk * bk[i];
                }
            }
            if (!unit) {
                double inv = 1.0 / a[j + j * lda];
                for (size_t i = 0; i < m; i++) bj[i] *= inv;
            }
        }
        return;
    }
    size_t n1 = n / 2, n2 = n - n1;
    trsm_rtu(unit, m, n2, a + n1 + n1 * lda, lda, b + n1 * ldb, ldb);
    bola_oracle_gemm('N', 'T', m, n1, n2, -1.0, b + n1 * ldb, ldb, a + n1 * lda, lda, 1.0, b, ldb);
    trsm_rtu(unit, m, n1, a, lda, b, ldb);
}

/* trsm, trsm.rs:33-91 */
EXPORT void bola_oracle_trsm(char side, char uplo, char transa, char diag, size_t m, size_t n, double alpha,
                             const double* a, size_t lda, double* b, size_t ldb) {
    if (m == 0 || n == 0) return;
    if (alpha != 1.0)
        for (size_t j = 0; j < n; j++)
            for (size_t i = 0; i < m; i++) b[i + j * ldb] *= alpha;
    int lside = eq_ic(side, 'L'), upper = eq_ic(uplo, 'U'), notrans = eq_ic(transa, 'N'), unit = eq_ic(diag, 'U');
    if (lside) {
        if (notrans) {
            if (upper) trsm_lnu(unit, m, n, a, lda, b, ldb);
            else trsm_lnl(unit, m, n, a, lda, b, ldb);
        } else if (upper) trsm_ltu(unit, m, n, a, lda, b, ldb);
        else trsm_ltl(unit, m, n, a, lda, b, ldb);
    } else if (notrans) {
        if (upper) trsm_rnu(unit, m, n, a, lda, b, ldb);
        else trsm_rnl(unit, m, n, a, lda, b, ldb);
    } else if (upper) trsm_rtu(unit, m, n, a, lda, b, ldb);
    else trsm_rtl(unit, m, n, a, lda, b, ldb);
}

/* ---------------------------------------------------------------- syrk.rs */

#define SYRK_NB 32

/* syrk_kernel, syrk.rs:190-249 (case-sensitive tests as in the reference) */
static void syrk_kernel(char uplo, char trans, size_t n, size_t k, double alpha, const double* a, size_t lda,
                        double* c, size_t ldc, char original_trans) {
    if (trans == 'N') {
        if (uplo == 'U') {
            for (size_t j = 0; j < n; j++)
                for (size_t l = 0; l < k; l++) {
                    double temp = alpha * a[j + l * lda];
                    if (temp != 0.0) bola_oracle_daxpy(j + 1, temp, a + l * lda, 1, c + j * ldc, 1);
                }
        } else {
            for (size_t j = 0; j < n; j++)
                for (size_t l = 0; l < k; l++) {
                    double temp = alpha * a[j + l * lda];
                    if (temp != 0.0) bola_oracle_daxpy(n - j, temp, a + j + l * lda, 1, c + j + j * ldc, 1);
                }
        }
    } else {
        int ot = original_trans == 'T';
        size_t lda_a = ot ? 1 : lda;
        if (uplo == 'U') {
            for (size_t j = 0; j < n; j++)
                for (size_t i = 0; i <= j; i++) {
                    const double* ri = ot ? a + i * lda : a + i;
                    const double* rj = ot ? a + j * lda : a + j;
                    double temp = bola_oracle_dot_product(k, ri, lda_a, rj, lda_a);
                    c[i + j * ldc] += alpha * temp;
                }
        } else {
            for (size_t j = 0; j < n; j++)
                for (size_t i = j; i < n; i++) {
                    const double* ri = ot ? a + i * lda : a + i;
                    const double* rj = ot ? a + j * lda : a + j;
                    double temp = bola_oracle_dot_product(k, ri, lda_a, rj, lda_a);
                    c[i + j * ldc] += alpha * temp;
                }
        }
    }
}

/* syrk, syrk.rs:17-186.  Restated as written, including its operand offsets for the
 * (U,N) and (L,T) combinations. */
EXPORT void bola_oracle_syrk(char uplo, char trans, size_t n, size_t k, double alpha, const double* a, size_t lda,
                             double beta, double* c, size_t ldc) {
    if (n == 0 || ((alpha == 0.0 || k == 0) && beta == 1.0)) return;
    if (beta == 0.0) {
        for (size_t j = 0; j < n; j++) {
            if (uplo == 'U') for (size_t i = 0; i <= j; i++) c[i + j * ldc] = 0.0;
            else for (size_t i = j; i < n; i++) c[i + j * ldc] = 0.0;
        }
    } else if (beta != 1.0) {
        for (size_t j = 0; j < n; j++) {
            if (uplo == 'U') for (size_t i = 0; i <= j; i++) c[i + j * ldc] *= beta;
            else for (size_t i = j; i < n; i++) c[i + j * ldc] *= beta;
        }
    }
    if (alpha == 0.0) return;
    if (uplo == 'U') {
        for (size_t j = 0; j < n; j += SYRK_NB) {
            size_t jb = min_sz(n - j, SYRK_NB);
            for (size_t p = 0; p < k; p += SYRK_NB) {
                size_t pb = min_sz(k - p, SYRK_NB);
                syrk_kernel('U', trans, jb, pb, alpha, a + j * lda + p, lda, c + j + j * ldc, ldc, trans);
                for (size_t i = 0; i < j; i += SYRK_NB) {
                    size_t ib = min_sz(j - i, SYRK_NB);
                    if (trans == 'T')
                        bola_oracle_gemm('T', 'N', ib, jb, pb, alpha, a + p + i * lda, lda, a + p + j * lda, lda, 1.0,
                                         c + i + j * ldc, ldc);
                    else
                        bola_oracle_gemm('N', 'T', ib, jb, pb, alpha, a + i + p * lda, lda, a + j + p * lda, lda, 1.0,
                                         c + i + j * ldc, ldc);
                }
            }
        }
    } else {
        for (size_t j = 0; j < n; j += SYRK_NB) {
            size_t jb = min_sz(n - j, SYRK_NB);
            for (size_t p = 0; p < k; p += SYRK_NB) {
                size_t pb = min_sz(k - p, SYRK_NB);
                syrk_kernel('L', trans, jb, pb, alpha, a + j + p * lda, lda, c + j + j * ldc, ldc, trans);
                for (size_t i = j + jb; i < n; i += SYRK_NB) {
                    size_t ib = min_sz(n - i, SYRK_NB);
                    if (trans == 'N')
                        bola_oracle_gemm('N', 'T', ib, jb, pb, alpha, a + i + p * lda, lda, a + j + p * lda, lda, 1.0,
                                         c + i + j * ldc, ldc);
                    else
                        bola_oracle_gemm('T', 'N', ib, jb, pb, alpha, a + p + i * lda, lda, a + p + j * lda, lda, 1.0,
                                         c + i + j * ldc, ldc);
                }
            }
        }
    }
}

/* ---------------------------------------------------------------- getrf2.rs / getrf.rs */

#define GETRF2_RECURSION_STOP 32
#define GETRF_NB 32

/* iterative_getrf2, getrf2.rs:131-185.  Returns 0, or the 1-based panel-local column of a
 * zero pivot (the reference's message text uses that local number). */
static int iterative_getrf2(size_t m, size_t n, double* a, size_t lda, int32_t* ipiv) {
    size_t min_mn = min_sz(m, n);
    for (size_t j = 0; j < min_mn; j++) {
        size_t jp = j + bola_oracle_idamax(m - j, a + j + j * lda);
        ipiv[j] = (int32_t)(jp + 1);
        if (a[jp + j * lda] != 0.0) {
            if (jp != j)
                for (size_t k = 0; k < n; k++) {
                    double t = a[j + k * lda];
                    a[j + k * lda] = a[jp + k * lda];
                    a[jp + k * lda] = t;
                }
            if (j < m - 1) {
                double ajj = a[j + j * lda];
                double inv = 1.0 / ajj;
                size_t m_rem = m - (j + 1);
                double* col = a + j + 1 + j * lda;
                size_t processed = bola_oracle_scal_kernel(m_rem, inv, col);
                for (size_t i = processed; i < m_rem; i++) col[i] *= inv;
            }
        } else {
            snprintf(g_err, sizeof g_err, "Matrix is singular. The pivot in column %zu is zero.", j + 1);
            return (int)(j + 1);
        }
        if (j < min_mn - 1)
            bola_oracle_ger(m - j - 1, n - j - 1, -1.0, a + j + 1 + j * lda, a + j + (j + 1) * lda, lda,
                            a + j + 1 + (j + 1) * lda, lda);
    }
    return 0;
}

/* getrf2, getrf2.rs:39-97 (recursive split at min(m,n)/2 above 32). */
EXPORT int bola_oracle_getrf2(size_t m, size_t n, double* a, size_t lda, int32_t* ipiv) {
    size_t mx = m > 1 ? m : 1;
    if (lda < mx) {
        snprintf(g_err, sizeof g_err, "Argument 4 to getrf2 had an illegal value of %zu", lda);
        return -4;
    }
    if (m == 0 || n == 0) return 0;
    if (m <= GETRF2_RECURSION_STOP || n <= GETRF2_RECURSION_STOP) return iterative_getrf2(m, n, a, lda, ipiv);
    size_t n1 = min_sz(m, n) / 2;
    size_t n2 = n - n1;
    int r = bola_oracle_getrf2(m, n1, a, lda, ipiv);
    if (r) return r;
    bola_oracle_laswp(n2, a + n1 * lda, lda, 1, (int32_t)n1, ipiv, 1);
    bola_oracle_trsm('L', 'L', 'N', 'U', n1, n2, 1.0, a, lda, a + n1 * lda, lda);
    if (n1 < m)
        bola_oracle_gemm('N', 'N', m - n1, n2, n1, -1.0, a + n1, lda, a + n1 * lda, lda, 1.0, a + n1 + n1 * lda, lda);
    if (n1 < m) {
        r = bola_oracle_getrf2(m - n1, n2, a + n1 + n1 * lda, lda, ipiv + n1);
        if (r) return r;
    }
    size_t mn = min_sz(m, n);
    for (size_t i = n1; i < mn; i++) ipiv[i] += (int32_t)n1;
    bola_oracle_laswp(n1, a, lda, (int32_t)(n1 + 1), (int32_t)mn, ipiv, 1);
    return 0;
}

/* getrf, getrf.rs:37-101.  Blocked right-looking LU, NB = 32. */
EXPORT int bola_oracle_getrf(size_t m, size_t n, double* a, size_t lda, int32_t* ipiv) {
    size_t mx = m > 1 ? m : 1;
    if (lda < mx) {
        snprintf(g_err, sizeof g_err, "Argument 4 to getrf had an illegal value of %zu", lda);
        return -4;
    }
    if (m == 0 || n == 0) return 0;
    size_t min_mn = min_sz(m, n);
    if (GETRF_NB <= 1 || GETRF_NB >= min_mn) return bola_oracle_getrf2(m, n, a, lda, ipiv);
    for (size_t j = 0; j < min_mn; j += GETRF_NB) {
        size_t jb = min_sz(GETRF_NB, min_mn - j);
        int r = bola_oracle_getrf2(m - j, jb, a + j + j * lda, lda, ipiv + j);
        if (r) return r;
        for (size_t i = j; i < j + jb; i++) ipiv[i] += (int32_t)j;
        int32_t k1 = (int32_t)(j + 1), k2 = (int32_t)(j + jb);
        if (j > 0) bola_oracle_laswp(j, a, lda, k1, k2, ipiv, 1);
        if (j + jb < n) {
            bola_oracle_laswp(n - j - jb, a + (j + jb) * lda, lda, k1, k2, ipiv, 1);
            double* l11 = a + j + j * lda;
            double* u12 = a + j + (j + jb) * lda;
            bola_oracle_trsm('L', 'L', 'N', 'U', jb, n - j - jb, 1.0, l11, lda, u12, lda);
            if (j + jb < m)
                bola_oracle_gemm('N', 'N', m - j - jb, n - j - jb, jb, -1.0, a + (j + jb) + j * lda, lda, u12, lda,
                                 1.0, a + (j + jb) + (j + jb) * lda, lda);
        }
    }
    return 0;
}

/* ---------------------------------------------------------------- potrf2.rs / potrf.rs */

/* potrf2_recursive, potrf2.rs:36-80: returns 0 or the 1-based failing column of the block. */
static int potrf2_recursive(char uplo, size_t n, double* a, size_t lda) {
    if (n == 0) return 0;
    if (n == 1) {
        double a11 = a[0];
        if (a11 > 0.0) { a[0] = sqrt(a11); return 0; }
        return 1;
    }
    size_t n1 = n / 2, n2 = n - n1;
    int r = potrf2_recursive(uplo, n1, a, lda);
    if (r) return r;
    if (uplo == 'U') {
        double* a12 = a + n1 * lda;
        bola_oracle_trsm('L', 'U', 'T', 'N', n1, n2, 1.0, a, lda, a12, lda);
        double* a22 = a + n1 + n1 * lda;
        bola_oracle_syrk('U', 'T', n2, n1, -1.0, a12, lda, 1.0, a22, lda);
        r = potrf2_recursive(uplo, n2, a22, lda);
        if (r) return r + (int)n1;
    } else {
        double* a21 = a + n1;
        bola_oracle_trsm('R', 'L', 'T', 'N', n2, n1, 1.0, a, lda, a21, lda);
        double* a22 = a + n1 + n1 * lda;
        bola_oracle_syrk('L', 'N', n2, n1, -1.0, a21, lda, 1.0, a22, lda);
        r = potrf2_recursive(uplo, n2, a22, lda);
        if (r) return r + (int)n1;
    }
    return 0;
}

/* potrf2, potrf2.rs:25-33 */
EXPORT int bola_oracle_potrf2(char uplo, size_t n, double* a, size_t lda) {
    int r = potrf2_recursive(uplo, n, a, lda);
    if (r) snprintf(g_err, sizeof g_err, "Matrix is not positive-definite. Failure at column %d.", r);
    return r;
}

#define POTRF_NB 32

/* potrf, potrf.rs:101-181.  Blocked left-looking Cholesky, NB = 32. */
EXPORT int bola_oracle_potrf(char uplo, size_t n, double* a, size_t lda) {
    if (uplo != 'U' && uplo != 'L') {
        snprintf(g_err, sizeof g_err, "Argument 1 to potrf had an illegal value");
        return -1;
    }
    size_t mx = n > 1 ? n : 1;
    if (lda < mx) {
        snprintf(g_err, sizeof g_err, "Argument 4 to potrf had an illegal value");
        return -4;
    }
    if (POTRF_NB <= 1 || POTRF_NB >= n) return bola_oracle_potrf2(uplo, n, a, lda);
    if (uplo == 'U') {
        for (size_t j = 0; j < n; j += POTRF_NB) {
            size_t jb = min_sz(POTRF_NB, n - j);
            double* ajj = a + j + j * lda;
            double* a1j = a + j * lda;
            bola_oracle_syrk('U', 'T', jb, j, -1.0, a1j, lda, 1.0, ajj, lda);
            int r = bola_oracle_potrf2('U', jb, ajj, lda);
            if (r) return r;
            if (j + jb < n) {
                size_t n_rem = n - j - jb;
                double* a1jb = a + (j + jb) * lda;
                double* ajjb = a + j + (j + jb) * lda;
                bola_oracle_gemm('T', 'N', jb, n_rem, j, -1.0, a1j, lda, a1jb, lda, 1.0, ajjb, lda);
                bola_oracle_trsm('L', 'U', 'T', 'N', jb, n_rem, 1.0, ajj, lda, ajjb, lda);
            }
        }
    } else {
        for (size_t j = 0; j < n; j += POTRF_NB) {
            size_t jb = min_sz(POTRF_NB, n - j);
            double* ajj = a + j + j * lda;
            double* aj0 = a + j;
            bola_oracle_syrk('L', 'N', jb, j, -1.0, aj0, lda, 1.0, ajj, lda);
            int r = bola_oracle_potrf2('L', jb, ajj, lda);
            if (r) return r;
            if (j + jb < n) {
                size_t n_rem = n - j - jb;
                double* ajb0 = a + j + jb;
                double* ajbj = a + j + jb + j * lda;
                bola_oracle_gemm('N', 'T', n_rem, jb, j, -1.0, ajb0, lda, aj0, lda, 1.0, ajbj, lda);
                bola_oracle_trsm('R', 'L', 'T', 'N', n_rem, jb, 1.0, ajj, lda, ajbj, lda);
            }
        }
    }
    return 0;
}

/* ---------------------------------------------------------------- misc */

EXPORT int bola_oracle_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

EXPORT void bola_oracle_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
