/*
 * bola_hip.h -- C ABI of libbola_hip.so, the gfx950 (MI355X) dense-solver path.
 *
 * Every entry point below replaces one public function of the Rust crate bola 0.1.2
 * (cited as file:line of that crate) or one composition the crate's callers build from
 * those functions.  Conventions follow the crate exactly:
 *   - column-major f64 matrices, element (i,j) at a[i + j*lda];
 *   - pivots are int32 and 1-based (ipiv[i] = row swapped with row i+1);
 *   - the character arguments are ASCII ('N','T','L','U','R'); a Rust `char` is narrowed
 *     to one byte by the binding.
 *
 * Two flavours per routine:
 *   bola_hip_X(...)            host pointers, synchronous: copies the operands to HBM,
 *                              runs on the calling thread's default handle, copies the
 *                              results back.  Same in-place semantics as the crate.
 *   bola_hip_X_dev(h, ...)     device pointers on handle h's stream; never copies.
 *
 * Return codes (int):
 *   0                    Ok
 *   -i  (i = 1..)        argument i had an illegal value (the crate's "Argument i to ..."
 *                        errors; getrf lda -> -4, potrf uplo -> -1, potrf lda -> -4)
 *   +k                   getrf: pivot k (1-based, global column) is exactly zero;
 *                        potrf: leading minor k is not positive definite
 *   BOLA_HIP_ERR_DEVICE  a HIP runtime call failed (no counterpart in the crate)
 * For every nonzero code, bola_hip_last_error() (host flavour) or
 * bola_hip_error_string(h) (device flavour) holds the text the crate's Err(String)
 * would carry, e.g. "Matrix is singular. The pivot in column 9 is zero." -- the crate
 * numbers that column within its 32-column panel, and so does the text.
 */
#ifndef BOLA_HIP_H
#define BOLA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BOLA_HIP_ERR_DEVICE (-1000)
#define BOLA_HIP_ERR_CANCELLED (-1001) /* a served request dropped when its server closed */

typedef struct bola_hip_handle_s* bola_hip_handle_t;

/* ------------------------------------------------------------------ handle */

/* Creates a handle on the current HIP device with its own stream.  No crate
 * counterpart (the crate has no state); it owns the stream and the device workspace. */
int bola_hip_create(bola_hip_handle_t* out);
int bola_hip_destroy(bola_hip_handle_t h);
/* Runs subsequent _dev calls on `stream` (a hipStream_t; NULL = the null stream). */
int bola_hip_set_stream(bola_hip_handle_t h, void* stream);
void* bola_hip_get_stream(bola_hip_handle_t h);
/* Waits for all work enqueued on the handle's stream. */
int bola_hip_synchronize(bola_hip_handle_t h);
/* No crate counterpart (a device-placement setting): exclusive != 0 (the default) lets the
 * handle assume its work has the device to itself -- an LU of up to 12288 rows and columns
 * then runs its critical path on streams restricted to a quarter of the CUs, and GETRS may
 * run each triangular solve as one launch whose workgroups wait for one another.  Set 0 for
 * a handle whose work shares the GPU with other heavy kernels (the server's slots do). */
int bola_hip_set_exclusive(bola_hip_handle_t h, int exclusive);
/* Text of the last failing call on h / on this thread's host-flavour calls. */
const char* bola_hip_error_string(bola_hip_handle_t h);
const char* bola_hip_last_error(void);

/* ------------------------------------------------------------------ profiling
 * Replaces the crate's `profiling` feature (profiling.rs:10-140): per-phase call counts
 * and GPU time under the same phase names (GEMM, GETRF, GETRF2, LASWP, TRSM, SYRK,
 * POTRF, POTRF2, IDAMAX, ...), measured with HIP events on the handle's stream. */
int bola_hip_profile_enable(bola_hip_handle_t h, int on);
/* reset_counters, profiling.rs:60-64 */
int bola_hip_profile_reset(bola_hip_handle_t h);
/* One phase: total GPU ms, number of calls, algorithmic flops attributed to it. */
int bola_hip_profile_get(bola_hip_handle_t h, const char* name, double* ms, long* calls, double* flops);
/* print_profile, profiling.rs:70-115: the same table, written into buf. */
int bola_hip_profile_report(bola_hip_handle_t h, double iterations, char* buf, size_t len);

/* ------------------------------------------------------------------ BLAS-1/2 */

/* idamax, idamax.rs:23 -- 0-based index of the largest |x_i| under the crate's x86_64
 * tie and NaN rule.  Host flavour returns it; device flavour writes *result_dev. */
size_t bola_hip_idamax(size_t n, const double* x);
int bola_hip_idamax_dev(bola_hip_handle_t h, size_t n, const double* x, int64_t* result_dev);

/* laswp, laswp.rs:22 -- row interchanges k1..k2 (1-based) of ipiv, applied to n columns. */
int bola_hip_dlaswp(size_t n, double* a, size_t lda, int32_t k1, int32_t k2, const int32_t* ipiv, int32_t incx);
int bola_hip_dlaswp_dev(bola_hip_handle_t h, size_t n, double* a, size_t lda, int32_t k1, int32_t k2,
                        const int32_t* ipiv, int32_t incx);

/* ger, ger.rs:28 -- A += alpha * x * y^T, y with stride ldy. */
int bola_hip_dger(size_t m, size_t n, double alpha, const double* x, const double* y, size_t ldy, double* a,
                  size_t lda);
int bola_hip_dger_dev(bola_hip_handle_t h, size_t m, size_t n, double alpha, const double* x, const double* y,
                      size_t ldy, double* a, size_t lda);

/* scal, scal.rs:45 -- x *= alpha. */
int bola_hip_dscal(size_t n, double alpha, double* x, size_t incx);
int bola_hip_dscal_dev(bola_hip_handle_t h, size_t n, double alpha, double* x, size_t incx);

/* daxpy, utilities.rs:107 -- y += alpha * x. */
int bola_hip_daxpy(size_t n, double alpha, const double* x, size_t incx, double* y, size_t incy);
int bola_hip_daxpy_dev(bola_hip_handle_t h, size_t n, double alpha, const double* x, size_t incx, double* y,
                       size_t incy);

/* dot_product, utilities.rs:22 -- sum x_i * y_i. */
double bola_hip_ddot(size_t n, const double* x, size_t incx, const double* y, size_t incy);
int bola_hip_ddot_dev(bola_hip_handle_t h, size_t n, const double* x, size_t incx, const double* y, size_t incy,
                      double* result_dev);

/* ------------------------------------------------------------------ BLAS-3 */

/* gemm, gemm.rs:88-102 -- C = alpha*op(A)*op(B) + beta*C. */
int bola_hip_dgemm(char transa, char transb, size_t m, size_t n, size_t k, double alpha, const double* a,
                   size_t lda, const double* b, size_t ldb, double beta, double* c, size_t ldc);
int bola_hip_dgemm_dev(bola_hip_handle_t h, char transa, char transb, size_t m, size_t n, size_t k, double alpha,
                       const double* a, size_t lda, const double* b, size_t ldb, double beta, double* c,
                       size_t ldc);

/* trsm, trsm.rs:33-45 -- op(A) X = alpha B (side 'L') or X op(A) = alpha B ('R'). */
int bola_hip_dtrsm(char side, char uplo, char transa, char diag, size_t m, size_t n, double alpha, const double* a,
                   size_t lda, double* b, size_t ldb);
int bola_hip_dtrsm_dev(bola_hip_handle_t h, char side, char uplo, char transa, char diag, size_t m, size_t n,
                       double alpha, const double* a, size_t lda, double* b, size_t ldb);

/* syrk, syrk.rs:17-28 -- C = alpha*A*A^T + beta*C ('N') or alpha*A^T*A + beta*C ('T'),
 * one triangle. */
int bola_hip_dsyrk(char uplo, char trans, size_t n, size_t k, double alpha, const double* a, size_t lda,
                   double beta, double* c, size_t ldc);
int bola_hip_dsyrk_dev(bola_hip_handle_t h, char uplo, char trans, size_t n, size_t k, double alpha,
                       const double* a, size_t lda, double beta, double* c, size_t ldc);

/* ------------------------------------------------------------------ LAPACK */

/* getrf, getrf.rs:37 -- P*A = L*U with partial pivoting. */
int bola_hip_dgetrf(size_t m, size_t n, double* a, size_t lda, int32_t* ipiv);
/* Device flavour: if info_dev is NULL the call waits and returns the status; otherwise
 * it returns at once and the status (0 / +k) is written to *info_dev. */
int bola_hip_dgetrf_dev(bola_hip_handle_t h, size_t m, size_t n, double* a, size_t lda, int32_t* ipiv,
                        int32_t* info_dev);
/* No counterpart in the reference: bola_hip_dgetrf_dev for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, kernels and results (bit for bit); the factorization's
 * scratch (panel exchange slots, flags, row maps) is a static device array of 4 MiB that
 * each GPU holds until the process ends, enough for m up to 32768 at any n.  Nothing is
 * allocated or freed.  Calls that overlap on one GPU race on that array.  -2: the
 * factorization needs more scratch than the array holds.  info_dev NULL: synchronous,
 * as bola_hip_dgetrf_dev. */
int bola_hip_dgetrf_noalloc_dev(bola_hip_handle_t h, size_t m, size_t n, double* a, size_t lda, int32_t* ipiv,
                                int32_t* info_dev);

/* getrf2, getrf2.rs:39 -- the crate's unblocked/recursive panel LU, callable directly. */
int bola_hip_dgetrf2(size_t m, size_t n, double* a, size_t lda, int32_t* ipiv);
int bola_hip_dgetrf2_dev(bola_hip_handle_t h, size_t m, size_t n, double* a, size_t lda, int32_t* ipiv,
                         int32_t* info_dev);

/* potrf, potrf.rs:101 -- A = L*L^T ('L') or U^T*U ('U'). */
int bola_hip_dpotrf(char uplo, size_t n, double* a, size_t lda);
int bola_hip_dpotrf_dev(bola_hip_handle_t h, char uplo, size_t n, double* a, size_t lda, int32_t* info_dev);

/* potrf2, potrf2.rs:25 -- the crate's recursive Cholesky, callable directly. */
int bola_hip_dpotrf2(char uplo, size_t n, double* a, size_t lda);
int bola_hip_dpotrf2_dev(bola_hip_handle_t h, char uplo, size_t n, double* a, size_t lda, int32_t* info_dev);

/* LU solve.  The crate has no getrs; its callers compose laswp(nrhs, B, 1, n, ipiv, 1),
 * trsm('L','L','N','U') and trsm('L','U','N','N') (laswp.rs:22, trsm.rs:33).  Same
 * composition, same results. */
int bola_hip_dgetrs(size_t n, size_t nrhs, const double* lu, size_t lda, const int32_t* ipiv, double* b,
                    size_t ldb);
int bola_hip_dgetrs_dev(bola_hip_handle_t h, size_t n, size_t nrhs, const double* lu, size_t lda,
                        const int32_t* ipiv, double* b, size_t ldb);

/* Cholesky solve: 'L' -> trsm('L','L','N','N'), trsm('L','L','T','N');
 *                 'U' -> trsm('L','U','T','N'), trsm('L','U','N','N'). */
int bola_hip_dpotrs(char uplo, size_t n, size_t nrhs, const double* f, size_t lda, double* b, size_t ldb);
int bola_hip_dpotrs_dev(bola_hip_handle_t h, char uplo, size_t n, size_t nrhs, const double* f, size_t lda,
                        double* b, size_t ldb);

/* ------------------------------------------------------------------ mixed precision
 * No crate counterpart: the crate solves in FP64 only.  A X = B (n x n, nrhs <= 16
 * right-hand sides, column-major f64) to FP64 accuracy from a low-precision LU:
 * power-of-two row/column equilibration, LU of the scaled matrix in fp32 storage with
 * tournament pivoting (or, with BOLA_HIP_MIX_RBT, butterfly randomisation and no
 * interchanges) and trailing products in fp32, fp16, bf16 or split fp16 on the MFMA units,
 * iterative refinement in FP64 (r = b - A x on the original A), GMRES-IR when plain
 * refinement stagnates, and the FP64 GETRF + GETRS of this library as the last resort.
 * Stops when every column's backward error |r|_inf / (|A|_inf |x|_inf + |b|_inf) is at most
 * n * 2^-52.  A and B are not modified.  n <= 16384.
 * Returns 0, -i for an illegal argument i, or +k when even the FP64 fallback meets an
 * exactly zero pivot in column k. */
#define BOLA_HIP_PREC_FP32 0
#define BOLA_HIP_PREC_FP16 1
#define BOLA_HIP_PREC_BF16 2
/* Split fp16: every fp32 operand x as the pair h = fp16(x), l = fp16((x - h) 2^11); the
 * product Ah Bh + 2^-11 (Ah Bl + Al Bh) takes three fp16 MFMA steps with fp32
 * accumulation and is about as accurate as an fp32 product. */
#define BOLA_HIP_PREC_FP16X3 3
/* Or-ed into `precision`: randomise A with two-level random butterflies on both sides
 * and factor without interchanges (instead of tournament pivoting); for n divisible by 4,
 * otherwise ignored.  The refinement and the fallback are unchanged. */
#define BOLA_HIP_MIX_RBT 0x100
typedef struct bola_hip_mixed_stats {
    int ir_iterations;    /* refinement steps with the low-precision factors */
    int gmres_cycles;     /* GMRES-IR restarts */
    int gmres_iterations; /* GMRES steps in total */
    int fallback;         /* 1: the FP64 GETRF + GETRS produced the result */
    int converged;        /* 1: final backward error <= tolerance */
    int pad;
    double first_berr;    /* backward error after the first correction */
    double final_berr;    /* largest over the columns, at return */
    double tolerance;     /* n * 2^-52 */
    double factor_ms;     /* equilibration + conversion + factorization (GPU time) */
    double refine_ms;     /* refinement and GMRES-IR */
    double fallback_ms;   /* FP64 fallback, 0 when not needed */
} bola_hip_mixed_stats;
int bola_hip_dgesv_mixed(int precision, size_t n, size_t nrhs, const double* a, size_t lda, const double* b,
                         size_t ldb, double* x, size_t ldx, bola_hip_mixed_stats* stats);
int bola_hip_dgesv_mixed_dev(bola_hip_handle_t h, int precision, size_t n, size_t nrhs, const double* a, size_t lda,
                             const double* b, size_t ldb, double* x, size_t ldx, bola_hip_mixed_stats* stats);
/* ------------------------------------------------------------------ factorization objects
 * No crate counterpart (the crate's callers keep an (a, ipiv) pair and call getrs-style
 * compositions).  A factorization kept in device memory for repeated right-hand sides, in
 * one format for the plain FP64 path, the mixed-precision solver and the solve server:
 *   BOLA_HIP_FACTOR_LU      getrf (bit-identical to bola_hip_dgetrf) + getrs
 *   BOLA_HIP_FACTOR_CHOL_L  potrf('L') + potrs('L'); _CHOL_U likewise
 *   BOLA_HIP_FACTOR_MIXED   the mixed solver's factorization (precision as in
 *                           bola_hip_dgesv_mixed): scalings, low-precision factors, pivots or
 *                           butterflies, and a copy of A for the FP64 residuals; each solve
 *                           refines to backward error <= n * 2^-52, and the FP64 LU of A is
 *                           made the first time refinement fails and kept for later solves.
 * A is read from host or device memory and not modified.  create returns 0, -i for an
 * illegal argument i, or +k (the reference's Err: zero pivot / not positive definite; no
 * object is made).  solve: b and x on the device, any nrhs (the mixed solver works 16
 * columns at a time), x may equal b for the FP64 kinds only. */
#define BOLA_HIP_FACTOR_LU 0
#define BOLA_HIP_FACTOR_CHOL_L 1
#define BOLA_HIP_FACTOR_CHOL_U 2
#define BOLA_HIP_FACTOR_MIXED 3
typedef struct bola_hip_factor_s* bola_hip_factor_t;
typedef struct bola_hip_factor_desc {
    int32_t kind, precision;
    size_t n;
    size_t bytes;              /* device memory held */
    const void* factor;        /* FP64 LU / L / U (ld n), or the fp32 factors of the scaled A^T (ld) */
    size_t ld;
    const int32_t* ipiv;       /* 1-based pivots (LU, mixed with pivoting) or NULL */
    const double* row_scale;   /* mixed: r (powers of two), A_s = diag(r) A diag(c) */
    const double* col_scale;   /* mixed: c */
    const double* matrix;      /* mixed: the FP64 copy of A the residuals use */
    int32_t butterfly;         /* mixed: 1 when randomised by butterflies (no pivots) */
    int32_t pad;
} bola_hip_factor_desc;
int bola_hip_factor_create_dev(bola_hip_handle_t h, int kind, int precision, size_t n, const double* a, size_t lda,
                               bola_hip_factor_t* out);
int bola_hip_factor_solve_dev(bola_hip_handle_t h, bola_hip_factor_t f, size_t nrhs, const double* b, size_t ldb,
                              double* x, size_t ldx, bola_hip_mixed_stats* stats);
int bola_hip_factor_describe(bola_hip_factor_t f, bola_hip_factor_desc* d);
size_t bola_hip_factor_bytes(int kind, int precision, size_t n);
int bola_hip_factor_destroy(bola_hip_factor_t f);

/* The trailing-update product of that factorization, exposed for tests and measurement:
 * C -= A B with A (m x k), B (k x n), C (m x n) column-major fp32 device arrays, the
 * products in `precision` (BOLA_HIP_PREC_*), accumulated in fp32. */
int bola_hip_sgemm_lp_dev(bola_hip_handle_t h, int precision, size_t m, size_t n, size_t k, const float* a,
                          size_t lda, const float* b, size_t ldb, float* c, size_t ldc);

/* ---- Batched solves (no reference counterpart: BOLA solves one system per call) ------
 * Many independent small systems in one persistent launch.  Every system lives in one
 * device arena of doubles: A (n x n, column-major, lda = n) at element offset `a`, B
 * (n x nrhs, ldb = n) at offset `b`; LU systems also own ipiv[n] (1-based, LAPACK
 * meaning) at offset `piv` of an int32 arena.  Descriptors are host memory. */
#define BOLA_HIP_BATCH_LU 0   /* general: LU with partial pivoting (dgesv) */
#define BOLA_HIP_BATCH_CHOL 1 /* symmetric positive definite, lower triangle (dposv) */
#define BOLA_HIP_BATCH_FP64 0     /* FP64 factors: A := its factors, B := X */
#define BOLA_HIP_BATCH_FP32_IR 1  /* FP32 factors, FP64 refinement to n * 2^-52 (FP64 refactor if
                                     it stalls): A unchanged, B := X */
typedef struct bola_hip_batch_desc {
    int32_t n;    /* 1 .. 512 */
    int32_t nrhs; /* 1 .. 16 */
    int32_t kind; /* BOLA_HIP_BATCH_LU or BOLA_HIP_BATCH_CHOL */
    int32_t pad;
    int64_t a, b, piv;
} bola_hip_batch_desc;
/* Launch shape: panel width nb (8 or 16) and workgroups per CU (1 .. 4). */
typedef struct bola_hip_batch_config {
    int32_t nb;
    int32_t wg_per_cu;
} bola_hip_batch_config;
/* info[i] (device, count entries): 0, or the LAPACK info of system i (first zero pivot /
 * first non-positive leading minor, 1-based).  iters (device, optional): refinement steps
 * of system i, -1 when it was refactored in FP64, 0 for the FP64 variant.  cfg NULL: the
 * configuration bola_hip_batch_tune_dev chose for a batch like this one, else a default. */
int bola_hip_batch_solve_dev(bola_hip_handle_t h, int variant, size_t count, const bola_hip_batch_desc* desc,
                             double* arena, int32_t* piv_arena, int32_t* info, int32_t* iters,
                             const bola_hip_batch_config* cfg);
/* Autotuner: times every configuration on a sample of the batch (copied, so the arena is
 * not modified), keeps the fastest for batches of this variant, largest order and count,
 * and returns it in *out. */
int bola_hip_batch_tune_dev(bola_hip_handle_t h, int variant, size_t count, const bola_hip_batch_desc* desc,
                            const double* arena, bola_hip_batch_config* out);
/* Device scratch a batched solve takes from the handle for a configuration (NULL: the
 * largest). */
size_t bola_hip_batch_workspace_bytes(bola_hip_handle_t h, const bola_hip_batch_config* cfg);
/* hipGraph capture of one batched solve with fixed descriptors and buffers: replay it with
 * new contents in the same arena.  Launches go to the handle's stream. */
typedef struct bola_hip_batch_graph_s* bola_hip_batch_graph_t;
int bola_hip_batch_graph_create(bola_hip_handle_t h, int variant, size_t count, const bola_hip_batch_desc* desc,
                                double* arena, int32_t* piv_arena, int32_t* info, int32_t* iters,
                                const bola_hip_batch_config* cfg, bola_hip_batch_graph_t* out);
int bola_hip_batch_graph_launch(bola_hip_batch_graph_t g);
int bola_hip_batch_graph_destroy(bola_hip_batch_graph_t g);

/* ---- Serving (no reference counterpart: the crate is a library of one-shot calls) ----
 * A server owns `slots` worker threads, each with its own handle and HIP stream.  Solves
 * submitted from any thread are admitted in arrival order as soon as a slot is free and
 * their device buffers fit in the HBM budget; a request that does not fit yet may be
 * passed by later ones a bounded number of times.  The slots' work overlaps on the
 * device.  Factorizations submitted with a nonzero factor_key stay resident (counted
 * against the budget, least recently used idle ones evicted first) and serve later
 * requests with the same key, kind and n, which then only upload B and solve.  Results
 * are those of bola_hip_dgetrf_dev + dgetrs_dev (LU) or dpotrf_dev + dpotrs_dev ('L'),
 * bit for bit, whatever else is running. */
#define BOLA_HIP_SERVE_LU 0
#define BOLA_HIP_SERVE_CHOL 1
/* request.precision: BOLA_HIP_SERVE_FP64 -- the reference's FP64 factorization (results
 * bit-identical to the one-shot entry points); BOLA_HIP_SERVE_MIXED(p) with p a
 * BOLA_HIP_PREC_* value, optionally | BOLA_HIP_MIX_RBT -- LU requests only: the
 * mixed-precision factorization object, each solve refined to FP64 accuracy
 * (backward error <= n * 2^-52). */
#define BOLA_HIP_SERVE_FP64 0
#define BOLA_HIP_SERVE_MIXED(p) (1 + (p))
typedef struct bola_hip_request {
    int32_t kind; /* BOLA_HIP_SERVE_LU or BOLA_HIP_SERVE_CHOL (lower triangle used) */
    int32_t precision;
    size_t n, nrhs;
    const double* a;     /* n x n column-major, lda = n, host or device memory; may be NULL on a
                            cache hit.  A request with a key must present the same matrix. */
    const double* b;     /* n x nrhs, ldb = n */
    double* x;           /* n x nrhs, ldx = n: receives the solution */
    uint64_t factor_key; /* 0: not cached */
} bola_hip_request;
typedef struct bola_hip_result {
    int32_t status;    /* 0, or BOLA_HIP_ERR_DEVICE / BOLA_HIP_ERR_CANCELLED */
    int32_t info;      /* the factorization's info (x is not written when nonzero) */
    int32_t cache_hit; /* 1: served from a resident factorization */
    int32_t slot;
    double queue_us;   /* submission -> admission */
    double latency_us; /* submission -> solution in x */
} bola_hip_result;
typedef struct bola_hip_server_stats {
    uint64_t submitted, completed, rejected, cache_hits, evictions;
    double p50_us, p99_us, mean_us; /* submission -> result */
    double solves_per_s;            /* completed / (last completion - first submission) */
    size_t hbm_in_use, hbm_peak, hbm_budget, hbm_cached;
} bola_hip_server_stats;
typedef struct bola_hip_server_s* bola_hip_server_t;
/* The server runs on the current device.  hbm_budget: bytes of request buffers and
 * resident factorizations (each slot's handle keeps its own small workspace besides);
 * 1 <= slots <= 16. */
int bola_hip_server_create(size_t hbm_budget, int slots, bola_hip_server_t* out);
/* Queues a request; its buffers must stay valid until the ticket completes.  -3: the
 * request alone exceeds the budget (or names no matrix and no resident factorization). */
int bola_hip_server_submit(bola_hip_server_t s, const bola_hip_request* r, uint64_t* ticket);
/* Blocks until the ticket completes and releases it. */
int bola_hip_server_wait(bola_hip_server_t s, uint64_t ticket, bola_hip_result* res);
/* 1: still pending; 0: complete (*res filled, the ticket released). */
int bola_hip_server_poll(bola_hip_server_t s, uint64_t ticket, bola_hip_result* res);
/* Blocks until nothing is queued or running. */
int bola_hip_server_drain(bola_hip_server_t s);
int bola_hip_server_get_stats(bola_hip_server_t s, bola_hip_server_stats* st);
int bola_hip_server_reset_stats(bola_hip_server_t s);
/* Cancels what is still queued, finishes what is running, frees everything. */
int bola_hip_server_destroy(bola_hip_server_t s);

#ifdef __cplusplus
}
#endif

#endif /* BOLA_HIP_H */
