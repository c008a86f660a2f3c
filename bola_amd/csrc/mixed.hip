// Mixed-precision LU solve: A X = B to FP64 accuracy from a low-precision factorization.
// No reference counterpart (the reference solves in FP64 only); the FP64 fallback is the
// reference-exact GETRF + GETRS of this library.
//
//   1. equilibration: row scales s (powers of two bringing every row's max |a| into
//      [1/2, 1)), then column scales t on S A likewise; A_s = S A C is exact.
//   2. T = A_s^T in fp32 and the LU of A_s (mixed_lu.hip), trailing products in fp32,
//      fp16 or bf16 on the MFMA units.
//   3. refinement in FP64: r = B - A X on the original FP64 A, d = M^-1 r with
//      M^-1 = C (L U)^-1 P S applied in FP64 arithmetic to the fp32 factors, X += d, until
//      the backward error |r|_inf / (|A|_inf |x|_inf + |b|_inf) <= n * eps64 in every
//      column.  Stagnation switches to GMRES-IR: restarted GMRES in FP64 on the
//      left-preconditioned system M^-1 A d = M^-1 r, one Krylov space per right-hand
//      side, advanced in lockstep so the matrix is read once per step for all of them.
//   4. If GMRES-IR stalls too, the FP64 reference-exact GETRF + GETRS solves the system.
//
// Triangular solves: one persistent kernel per triangle, blocks of 128 rows, one
// workgroup per block; a block polls (bounded) the epoch-tagged solved values of the
// blocks it depends on with device-coherent loads.  Workgroups are dispatched in
// dependency order, so a waiting workgroup never blocks one it waits for.
#include "mixed.hpp"

#include <algorithm>
#include <cmath>
#include <functional>
#include <string>
#include <vector>

namespace bola {


namespace {

constexpr int TB = 128;       // triangular-solve block (rows per workgroup of trsv_blk)
constexpr int NRMAX = 16;     // right-hand sides per solve call
constexpr int RCH = 32;       // column chunks of the residual kernel
constexpr long kTrsvSpin = 1L << 22;

__device__ __forceinline__ void ast(double* p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double ald(const double* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t tld(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void tst(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double pow2_scale(double m) {
    if (!(m > 0.0) || !(m < INFINITY)) return 1.0;
    int e;
    frexp(m, &e);  // m = f 2^e, f in [1/2, 1)
    return ldexp(1.0, -e);
}
__device__ __forceinline__ double rdl(double v, int l) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)u, l);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(u >> 32), l);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// ---------------------------------------------------------------- equilibration
__global__ void eq_row_part(int n, const double* __restrict__ A, size_t lda, double* __restrict__ pmax,
                            double* __restrict__ psum, int cw) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int c0 = blockIdx.y * cw, c1 = min(n, c0 + cw);
    if (r >= n) return;
    double mx = 0.0, sm = 0.0;
    for (int c = c0; c < c1; c++) {
        const double v = fabs(A[(size_t)r + (size_t)c * lda]);
        mx = fmax(mx, v);
        sm += v;
    }
    pmax[(size_t)blockIdx.y * n + r] = mx;
    psum[(size_t)blockIdx.y * n + r] = sm;
}
__global__ void eq_row_fin(int n, int nch, const double* __restrict__ pmax, const double* __restrict__ psum,
                           double* __restrict__ s, double* __restrict__ rsum) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double mx = 0.0, sm = 0.0;
    for (int ch = 0; ch < nch; ch++) {
        mx = fmax(mx, pmax[(size_t)ch * n + r]);
        sm += psum[(size_t)ch * n + r];
    }
    s[r] = pow2_scale(mx);
    rsum[r] = sm;
}
__global__ void __launch_bounds__(256) eq_col(int n, const double* __restrict__ A, size_t lda,
                                              const double* __restrict__ s, double* __restrict__ t) {
    __shared__ double red[4];
    const int c = blockIdx.x, tid = threadIdx.x;
    double mx = 0.0;
    for (int r = tid; r < n; r += 256) mx = fmax(mx, fabs(s[r] * A[(size_t)r + (size_t)c * lda]));
    for (int off = 32; off >= 1; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
    if ((tid & 63) == 0) red[tid >> 6] = mx;
    __syncthreads();
    if (tid == 0) t[c] = pow2_scale(fmax(fmax(red[0], red[1]), fmax(red[2], red[3])));
}
// T[c + r*ldt] = s_r A(r,c) t_c in fp32 (64 x 64 tiles through LDS)
__global__ void __launch_bounds__(256) to_T(int n, const double* __restrict__ A, size_t lda,
                                            const double* __restrict__ s, const double* __restrict__ t,
                                            float* __restrict__ T, size_t ldt) {
    __shared__ float tile[64][65];
    const int r0 = blockIdx.x * 64, c0 = blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int cc = ty; cc < 64; cc += 4) {
        const int r = r0 + tx, c = c0 + cc;
        tile[cc][tx] = (r < n && c < n) ? (float)(s[r] * A[(size_t)r + (size_t)c * lda] * t[c]) : 0.f;
    }
    __syncthreads();
    for (int rr = ty; rr < 64; rr += 4) {
        const int r = r0 + rr, c = c0 + tx;
        if (r < n && c < n) T[(size_t)c + (size_t)r * ldt] = tile[tx][rr];
    }
}
__global__ void k_iota(int n, int* p) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

// ---------------------------------------------------------------- butterflies
// One level of a recursive butterfly transformation (Parker): blocks of 2h rows (or
// columns); inside a block the pair (a, b) = (x_i, x_{i+h}) becomes
//   transposed:  (R_i (a + b), R_{i+h} (a - b)) / sqrt 2
//   plain:       (R_i a + R_{i+h} b, R_i a - R_{i+h} b) / sqrt 2
// with R an n-vector of random factors near one.  Randomising A this way makes LU
// without interchanges safe with high probability.
constexpr double kInvSqrt2 = 0.70710678118654752440;

// On T (fp32).  LEFT: pairs of rows of A (columns of T, contiguous); RIGHT: pairs of
// columns of A (rows of T).  Both use the transposed pair rule: the left factor is W_L^T,
// and the right factor's plain form on the columns of A is the transposed rule on the
// rows of T.  Threads run along T's contiguous index in both cases.
template <bool LEFT>
__global__ void rbt_T(int n, float* __restrict__ T, size_t ldt, int h, const double* __restrict__ R) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    for (int o = blockIdx.y; o < (LEFT ? n / 2 : n); o += gridDim.y) {
        float *pa, *pb;
        int i;
        if (LEFT) {  // o: pair, t: element along the column of T
            if (t >= n) return;
            i = (o / h) * 2 * h + o % h;
            pa = T + (size_t)t + (size_t)i * ldt;
            pb = pa + (size_t)h * ldt;
        } else {  // o: column of T, t: pair
            if (t >= n / 2) return;
            i = (t / h) * 2 * h + t % h;
            pa = T + (size_t)i + (size_t)o * ldt;
            pb = pa + h;
        }
        const float r0 = (float)(R[i] * kInvSqrt2), r1 = (float)(R[i + h] * kInvSqrt2);
        const float a = *pa, b = *pb;
        *pa = r0 * (a + b);
        *pb = r1 * (a - b);
    }
}
// on n x nrhs FP64 columns: TRANS: the transposed form (W^T z); else the plain form (W y)
template <bool TRANS>
__global__ void rbt_vec(int n, int nrhs, double* __restrict__ Z, size_t ldz, int h, const double* __restrict__ R) {
    const int pr = blockIdx.x * blockDim.x + threadIdx.x, q = blockIdx.y;
    if (pr >= n / 2) return;
    const int blk = pr / h, i = blk * 2 * h + pr % h;
    double* z = Z + (size_t)q * ldz;
    const double a = z[i], b = z[i + h], r0 = R[i] * kInvSqrt2, r1 = R[i + h] * kInvSqrt2;
    if (TRANS) {
        z[i] = r0 * (a + b);
        z[i + h] = r1 * (a - b);
    } else {
        z[i] = r0 * a + r1 * b;
        z[i + h] = r0 * a - r1 * b;
    }
}

// ---------------------------------------------------------------- M^-1 pieces
// Z(i, q) = s[perm[i]] R(perm[i], q)
__global__ void gather_scale(int n, int nrhs, const double* __restrict__ R, size_t ldr, const int* __restrict__ perm,
                             const double* __restrict__ s, double* __restrict__ Z, size_t ldz) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int p = perm[i];
    const double sc = s[p];
    for (int q = 0; q < nrhs; q++) Z[(size_t)i + (size_t)q * ldz] = sc * R[(size_t)p + (size_t)q * ldr];
}
// D(c, q) = t[c] W(c, q)
__global__ void col_scale(int n, int nrhs, const double* __restrict__ W, size_t ldw, const double* __restrict__ t,
                          double* __restrict__ D, size_t ldd) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double sc = t[i];
    for (int q = 0; q < nrhs; q++) D[(size_t)i + (size_t)q * ldd] = sc * W[(size_t)i + (size_t)q * ldw];
}

// Inverses of the TB x TB diagonal blocks of L (unit lower, blockIdx.y == 0) and U
// (blockIdx.y == 1) in T, in fp32, row-major per block: D[((y nb + bi) TB + r) TB + c].
// Blocks past n are completed with the identity.  Lane b substitutes column b of the
// inverse, two accumulators per row.
__global__ void __launch_bounds__(TB) diag_inv(int n, const float* __restrict__ T, size_t ldt, float* __restrict__ D,
                                               int nb) {
    __shared__ float f[TB][TB + 1];
    __shared__ float x[TB][TB + 1];
    const int bi = blockIdx.x, up = blockIdx.y, b = threadIdx.x;
    const int r0 = bi * TB, rows = min(TB, n - r0);
    for (int r = 0; r < TB; r++)  // f[r][c] = factor(r0 + r, r0 + c)
        f[r][b] = (r < rows && b < rows) ? T[(size_t)(r0 + b) + (size_t)(r0 + r) * ldt] : (r == b ? 1.f : 0.f);
    __syncthreads();
#pragma unroll 1
    for (int i = 0; i < TB; i++) {
        const int a = up ? TB - 1 - i : i;
        const int lo = up ? a + 1 : 0, hi = up ? TB : a;
        float v0 = a == b ? 1.f : 0.f, v1 = 0.f;
        int c = lo;
#pragma unroll 4
        for (; c + 1 < hi; c += 2) {
            v0 = fmaf(-f[a][c], x[c][b], v0);
            v1 = fmaf(-f[a][c + 1], x[c + 1][b], v1);
        }
        if (c < hi) v0 = fmaf(-f[a][c], x[c][b], v0);
        v0 += v1;
        x[a][b] = up ? v0 / f[a][a] : v0;
    }
    __syncthreads();
    float* o = D + ((size_t)up * nb + bi) * TB * TB;
    for (int r = 0; r < TB; r++) o[(size_t)r * TB + b] = x[r][b];
}

// Z := L^-1 Z (LOWER, L unit lower) or U^-1 Z (upper), factors in T (fp32), FP64
// arithmetic; Z is n x nrhs (ld ldz).  Block bi = rows [TB bi, TB bi + TB), one workgroup
// each, dispatched in the order the blocks are solved.  A block subtracts the products
// with the blocks it depends on (the oldest first), then applies the inverse of its
// diagonal block (diag_inv).  A solved block publishes its rows twice: plainly in Z, for
// the kernels that follow, and in zt (same indexing, two 64-bit words per double), each
// word carrying the call's epoch in its upper half beside 32 bits of the value.  A block
// that depends on it polls those words until they carry the epoch: no store waits for
// another and there is no flag.  zt starts zeroed and epochs are unique per solver.
// Thread (r, h) holds row r's entries in columns [64 h, 64 h + 64) of a block; the two
// halves meet in LDS.
template <bool LOWER>
__global__ void __launch_bounds__(256) trsv_blk(int n, int nrhs, const float* __restrict__ T, size_t ldt,
                                                const float* __restrict__ Dinv, double* __restrict__ Z, size_t ldz,
                                                uint64_t* __restrict__ zt, int epoch, int nb, long spin_limit,
                                                int* __restrict__ err) {
    constexpr int HW = TB / 2;
    __shared__ double ys[TB][NRMAX + 1];       // the incoming block of the solution
    __shared__ double acc[TB][NRMAX + 1];      // this block's right-hand side, updated
    __shared__ double part[2][TB][NRMAX + 1];  // the halves' partial products
    __shared__ int ok_sh;
    const int bi = LOWER ? (int)blockIdx.x : nb - 1 - (int)blockIdx.x;
    const int r0 = bi * TB, rows = min(TB, n - r0);
    const int tid = threadIdx.x, r = tid % TB, h = tid / TB;
    for (int e = tid; e < TB * nrhs; e += 256) {
        const int rr = e % TB, q = e / TB;
        acc[rr][q] = rr < rows ? Z[(size_t)(r0 + rr) + (size_t)q * ldz] : 0.0;
    }
    // this row's half of the diagonal block's inverse, wanted last: loaded first
    float dr[HW];
    {
        const float4* drow = reinterpret_cast<const float4*>(Dinv + (((size_t)(LOWER ? 0 : 1) * nb + bi) * TB + r) * TB +
                                                             h * HW);
#pragma unroll
        for (int c4 = 0; c4 < HW / 4; c4++) {
            const float4 v = drow[c4];
            dr[4 * c4] = v.x;
            dr[4 * c4 + 1] = v.y;
            dr[4 * c4 + 2] = v.z;
            dr[4 * c4 + 3] = v.w;
        }
    }
    const int ndep = LOWER ? bi : nb - 1 - bi;
    for (int jj = 0; jj < ndep; jj++) {
        // dependencies in the order they are solved: the oldest first
        const int j = LOWER ? jj : nb - 1 - jj;
        const int c0 = j * TB, cols = min(TB, n - c0);
        // this row's factor entries in block j do not depend on the flag: load them first
        // (clamped addresses, so that all the loads are in flight at once; masked after)
        float lr[HW];
        const float* lrow = T + (size_t)c0 + (size_t)(r0 + min(r, rows - 1)) * ldt;
        if (cols == TB) {  // a full block: the row's 64 entries as 16-byte loads (rows of T are 256-byte aligned)
            const float4* p4 = reinterpret_cast<const float4*>(lrow + h * HW);
#pragma unroll
            for (int c4 = 0; c4 < HW / 4; c4++) {
                const float4 v = p4[c4];
                lr[4 * c4] = v.x;
                lr[4 * c4 + 1] = v.y;
                lr[4 * c4 + 2] = v.z;
                lr[4 * c4 + 3] = v.w;
            }
        } else {
#pragma unroll
            for (int c = 0; c < HW; c++) lr[c] = lrow[min(h * HW + c, cols - 1)];
#pragma unroll
            for (int c = 0; c < HW; c++) lr[c] = keep_or_zero(lr[c], h * HW + c < cols);
        }
        // block j's solution, word by word until every word carries the epoch; the block
        // solved next polls eagerly, the others (usually long solved) sparingly
        const bool next = j == (LOWER ? bi - 1 : bi + 1);
        int ok = 1;
        for (int e = tid; e < TB * nrhs; e += 256) {
            const int c = e % TB, q = e / TB;
            double v = 0.0;
            if (c < cols) {
                const uint64_t* w = zt + 2 * ((size_t)(c0 + c) + (size_t)q * ldz);
                uint64_t lo, hi;
                long sp = 0;
                for (;;) {
                    lo = tld(w);
                    hi = tld(w + 1);
                    if ((uint32_t)(lo >> 32) == (uint32_t)epoch && (uint32_t)(hi >> 32) == (uint32_t)epoch) break;
                    if (++sp > spin_limit ||
                        __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                        ok = 0;
                        break;
                    }
                    if (!next) __builtin_amdgcn_s_sleep(32);
                }
                v = __longlong_as_double((long long)((lo & 0xffffffffull) | (hi << 32)));
            }
            ys[c][q] = v;
        }
        if (!__syncthreads_and(ok)) {
            if (tid == 0) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return;
        }
        for (int q = 0; q < nrhs; q++) {
            double s0 = 0.0, s1 = 0.0;
#pragma unroll
            for (int c = 0; c < HW; c += 2) {
                s0 = fma((double)lr[c], ys[h * HW + c][q], s0);
                s1 = fma((double)lr[c + 1], ys[h * HW + c + 1][q], s1);
            }
            part[h][r][q] = s0 + s1;
        }
        __syncthreads();
        for (int e = tid; e < TB * nrhs; e += 256) {
            const int rr = e % TB, q = e / TB;
            acc[rr][q] -= part[0][rr][q] + part[1][rr][q];
        }
    }
    __syncthreads();
    // the diagonal block: y = D^-1 acc
    for (int q = 0; q < nrhs; q++) {
        double s0 = 0.0, s1 = 0.0;
#pragma unroll
        for (int c = 0; c < HW; c += 2) {
            s0 = fma((double)dr[c], acc[h * HW + c][q], s0);
            s1 = fma((double)dr[c + 1], acc[h * HW + c + 1][q], s1);
        }
        part[h][r][q] = s0 + s1;
    }
    __syncthreads();
    const uint64_t tag = (uint64_t)(uint32_t)epoch << 32;
    for (int e = tid; e < TB * nrhs; e += 256) {
        const int rr = e % TB, q = e / TB;
        if (rr < rows) {
            const double y = part[0][rr][q] + part[1][rr][q];
            const size_t at = (size_t)(r0 + rr) + (size_t)q * ldz;
            const uint64_t b = (uint64_t)__double_as_longlong(y);
            Z[at] = y;
            tst(zt + 2 * at, tag | (b & 0xffffffffull));
            tst(zt + 2 * at + 1, tag | (b >> 32));
        }
    }
}

// ---------------------------------------------------------------- residual / products
// P[(ch * NRMAX + q) * n + r] = sum over the chunk's columns c of A(r, c) X(c, q)
__global__ void __launch_bounds__(256) prod_part(int n, int nrhs, const double* __restrict__ A, size_t lda,
                                                 const double* __restrict__ X, size_t ldx, double* __restrict__ P,
                                                 int cw) {
    __shared__ double xs[64][NRMAX];
    const int r = blockIdx.x * 256 + threadIdx.x;
    const int c0 = blockIdx.y * cw, c1 = min(n, c0 + cw);
    double acc[NRMAX];
#pragma unroll
    for (int q = 0; q < NRMAX; q++) acc[q] = 0.0;
    for (int cb = c0; cb < c1; cb += 64) {
        const int cn = min(64, c1 - cb);
        __syncthreads();
        for (int e = threadIdx.x; e < 64 * NRMAX; e += 256) {
            const int c = e & 63, q = e >> 6;
            xs[c][q] = (c < cn && q < nrhs) ? X[(size_t)(cb + c) + (size_t)q * ldx] : 0.0;
        }
        __syncthreads();
        if (r < n) {
            if (nrhs == 1) {
                for (int c = 0; c < cn; c++) acc[0] = fma(A[(size_t)r + (size_t)(cb + c) * lda], xs[c][0], acc[0]);
            } else {
                for (int c = 0; c < cn; c++) {
                    const double a = A[(size_t)r + (size_t)(cb + c) * lda];
#pragma unroll
                    for (int q = 0; q < NRMAX; q++) acc[q] = fma(a, xs[c][q], acc[q]);
                }
            }
        }
    }
    if (r < n)
        for (int q = 0; q < nrhs; q++) P[((size_t)blockIdx.y * NRMAX + q) * n + r] = acc[q];
}
// R(r, q) = (B ? B(r, q) : 0) - sum_ch P  (fixed order)
__global__ void prod_fin(int n, int nrhs, int nch, const double* __restrict__ P, const double* __restrict__ B,
                         size_t ldb, double* __restrict__ R, size_t ldr) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    for (int q = 0; q < nrhs; q++) {
        double s = 0.0;
        for (int ch = 0; ch < nch; ch++) s += P[((size_t)ch * NRMAX + q) * n + r];
        R[(size_t)r + (size_t)q * ldr] = (B ? B[(size_t)r + (size_t)q * ldb] : 0.0) - s;
    }
}

// ---------------------------------------------------------------- vector kernels
// out[q] = max_i |V(i, q)|
__global__ void __launch_bounds__(256) col_amax(int n, const double* __restrict__ V, size_t ldv, double* __restrict__ out) {
    __shared__ double red[4];
    const int q = blockIdx.x, tid = threadIdx.x;
    double m = 0.0;
    for (int i = tid; i < n; i += 256) {
        const double v = fabs(V[(size_t)i + (size_t)q * ldv]);
        m = (v > m || v != v) ? v : m;  // NaN propagates
    }
    for (int off = 32; off >= 1; off >>= 1) {
        const double o = __shfl_xor(m, off, 64);
        m = (o > m || o != o) ? o : m;
    }
    if ((tid & 63) == 0) red[tid >> 6] = m;
    __syncthreads();
    if (tid == 0) {
        double g = red[0];
        for (int w = 1; w < 4; w++) g = (red[w] > g || red[w] != red[w]) ? red[w] : g;
        out[q] = g;
    }
}
// out[j * nrhs + q] = sum_i V_j(i, q) W(i, q), j < nv, V_j at V + j*jstride (column q at + q*qstride)
__global__ void __launch_bounds__(256) multi_dot(int n, int nrhs, int nv, const double* __restrict__ V, size_t jstride,
                                                 size_t qstride, const double* __restrict__ W, size_t ldw,
                                                 double* __restrict__ out) {
    __shared__ double red[4];
    const int j = blockIdx.x, q = blockIdx.y, tid = threadIdx.x;
    const double* v = V + (size_t)j * jstride + (size_t)q * qstride;
    const double* w = W + (size_t)q * ldw;
    double s = 0.0;
    for (int i = tid; i < n; i += 256) s = fma(v[i], w[i], s);
    for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((tid & 63) == 0) red[tid >> 6] = s;
    __syncthreads();
    if (tid == 0) out[(size_t)j * nrhs + q] = (red[0] + red[1]) + (red[2] + red[3]);
    (void)nv;
}
// W(i, q) -= sum_{j < nv} H[j * nrhs + q] V_j(i, q)
__global__ void multi_axpy(int n, int nrhs, int nv, const double* __restrict__ V, size_t jstride, size_t qstride,
                           const double* __restrict__ H, double* __restrict__ W, size_t ldw) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, q = blockIdx.y;
    if (i >= n) return;
    double s = W[(size_t)i + (size_t)q * ldw];
    for (int j = 0; j < nv; j++) s = fma(-H[(size_t)j * nrhs + q], V[(size_t)j * jstride + (size_t)q * qstride + i], s);
    W[(size_t)i + (size_t)q * ldw] = s;
}
// Y(i, q) = a[q] X(i, q)   (a per column, from device memory)
__global__ void col_mul(int n, int nrhs, const double* __restrict__ a, const double* __restrict__ X, size_t ldx,
                        double* __restrict__ Y, size_t ldy) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, q = blockIdx.y;
    if (i >= n) return;
    Y[(size_t)i + (size_t)q * ldy] = a[q] * X[(size_t)i + (size_t)q * ldx];
}
// X(i, q) += D(i, q)
__global__ void mat_add(int n, int nrhs, const double* __restrict__ D, size_t ldd, double* __restrict__ X, size_t ldx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, q = blockIdx.y;
    if (i >= n) return;
    X[(size_t)i + (size_t)q * ldx] += D[(size_t)i + (size_t)q * ldd];
}
__global__ void mat_copy(int n, int nrhs, const double* __restrict__ S, size_t lds, double* __restrict__ D, size_t ldd) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, q = blockIdx.y;
    if (i >= n) return;
    D[(size_t)i + (size_t)q * ldd] = S[(size_t)i + (size_t)q * lds];
}
__global__ void mat_zero(int n, int nrhs, double* __restrict__ D, size_t ldd) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, q = blockIdx.y;
    if (i >= n) return;
    D[(size_t)i + (size_t)q * ldd] = 0.0;
}
__global__ void k_zero_int(int n, int* p) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

// ---------------------------------------------------------------- host side
struct Solver {
    bola_hip_handle_s* h;
    hipStream_t st;
    int n, nrhs, lp;
    const double* A;
    size_t lda;
    // device buffers
    float* T;
    float* Dinv;  // inverses of the diagonal blocks of L and U (diag_inv)
    size_t ldt;
    int32_t* ipiv;
    int* perm;
    double *s, *t, *rsum, *pmax, *psum;
    double *R, *Z, *D, *W, *P, *nrm, *dots;
    double* Vb;  // GMRES basis: [q][j][n]
    uint64_t* zt;  // tagged copy of Z for the blocked triangular solves
    int* err;
    void* luws;
    int epoch = 0;
    int mgm;  // GMRES restart length
    bool rbt = false;       // butterfly-randomised, factored without interchanges
    double* rl[2] = {};     // left butterflies, levels 1 (h = n/2) and 2 (h = n/4)
    double* rr[2] = {};     // right butterflies
};

dim3 g1(int n) { return dim3(cdiv((size_t)n, 256)); }
dim3 g2(int n, int nrhs) { return dim3(cdiv((size_t)n, 256), (unsigned)nrhs); }

// Y = M^-1 Xr (all n x nrhs, ld n); uses Z as scratch
void apply_minv(Solver& S, const double* Xr, size_t ldx, double* Y) {
    hipLaunchKernelGGL(gather_scale, g1(S.n), dim3(256), 0, S.st, S.n, S.nrhs, Xr, ldx, S.perm, S.s, S.Z, (size_t)S.n);
    BOLA_LAUNCH_CHECK();
    const dim3 gp(cdiv((size_t)S.n / 2, 256), (unsigned)S.nrhs);
    if (S.rbt) {  // W_L^T z = B1^T (B2^T z)
        hipLaunchKernelGGL(rbt_vec<true>, gp, dim3(256), 0, S.st, S.n, S.nrhs, S.Z, (size_t)S.n, S.n / 4, S.rl[1]);
        hipLaunchKernelGGL(rbt_vec<true>, gp, dim3(256), 0, S.st, S.n, S.nrhs, S.Z, (size_t)S.n, S.n / 2, S.rl[0]);
        BOLA_LAUNCH_CHECK();
    }
    const int nb = (S.n + TB - 1) / TB;
    S.epoch++;
    hipLaunchKernelGGL(trsv_blk<true>, dim3(nb), dim3(256), 0, S.st, S.n, S.nrhs, S.T, S.ldt, S.Dinv, S.Z,
                       (size_t)S.n, S.zt, S.epoch, nb, kTrsvSpin, S.err);
    BOLA_LAUNCH_CHECK();
    S.epoch++;
    hipLaunchKernelGGL(trsv_blk<false>, dim3(nb), dim3(256), 0, S.st, S.n, S.nrhs, S.T, S.ldt, S.Dinv, S.Z,
                       (size_t)S.n, S.zt, S.epoch, nb, kTrsvSpin, S.err);
    BOLA_LAUNCH_CHECK();
    if (S.rbt) {  // W_R y = B2 (B1 y)
        hipLaunchKernelGGL(rbt_vec<false>, gp, dim3(256), 0, S.st, S.n, S.nrhs, S.Z, (size_t)S.n, S.n / 2, S.rr[0]);
        hipLaunchKernelGGL(rbt_vec<false>, gp, dim3(256), 0, S.st, S.n, S.nrhs, S.Z, (size_t)S.n, S.n / 4, S.rr[1]);
        BOLA_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(col_scale, g1(S.n), dim3(256), 0, S.st, S.n, S.nrhs, S.Z, (size_t)S.n, S.t, Y, (size_t)S.n);
    BOLA_LAUNCH_CHECK();
}

// R = (B ? B : 0) - A X
void residual(Solver& S, const double* B, size_t ldb, const double* X, size_t ldx, double* R) {
    const int cw = (S.n + RCH - 1) / RCH;
    const int nch = (S.n + cw - 1) / cw;
    hipLaunchKernelGGL(prod_part, dim3(cdiv((size_t)S.n, 256), (unsigned)nch), dim3(256), 0, S.st, S.n, S.nrhs, S.A,
                       S.lda, X, ldx, S.P, cw);
    BOLA_LAUNCH_CHECK();
    hipLaunchKernelGGL(prod_fin, g1(S.n), dim3(256), 0, S.st, S.n, S.nrhs, nch, S.P, B, ldb, R, (size_t)S.n);
    BOLA_LAUNCH_CHECK();
}

void amax(Solver& S, const double* V, size_t ldv, double* out_dev) {
    hipLaunchKernelGGL(col_amax, dim3((unsigned)S.nrhs), dim3(256), 0, S.st, S.n, V, ldv, out_dev);
    BOLA_LAUNCH_CHECK();
}

void put(Solver& S, double* dev, const std::vector<double>& v) {
    BOLA_CHECK(hipMemcpyAsync(dev, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice, S.st));
    BOLA_CHECK(hipStreamSynchronize(S.st));  // v may go out of scope right after
}

std::vector<double> fetch(Solver& S, const double* dev, int count) {
    std::vector<double> v(count);
    BOLA_CHECK(hipMemcpyAsync(v.data(), dev, sizeof(double) * count, hipMemcpyDeviceToHost, S.st));
    BOLA_CHECK(hipStreamSynchronize(S.st));
    return v;
}

// backward error per column of the current X (R must hold B - A X)
std::vector<double> berr(Solver& S, const double* X, size_t ldx, const std::vector<double>& bn, double anorm) {
    amax(S, X, ldx, S.nrm);
    amax(S, S.R, (size_t)S.n, S.nrm + NRMAX);
    std::vector<double> v = fetch(S, S.nrm, 2 * NRMAX);
    std::vector<double> e(S.nrhs);
    for (int q = 0; q < S.nrhs; q++) {
        const double den = anorm * v[q] + bn[q];
        e[q] = den > 0 ? v[NRMAX + q] / den : (v[NRMAX + q] == 0 ? 0.0 : INFINITY);
        if (!(e[q] == e[q])) e[q] = INFINITY;
    }
    return e;
}

double maxv(const std::vector<double>& v) {
    double m = 0;
    for (double x : v) m = (x > m || x != x) ? x : m;
    return m;
}

// one restarted GMRES cycle on M^-1 A d = M^-1 r for every column; returns steps taken
int gmres_cycle(Solver& S, double* Dout, double tol_rel) {
    const int n = S.n, nr = S.nrhs, M = S.mgm;
    const size_t js = (size_t)n, qs = (size_t)(M + 1) * n;
    // v0 = M^-1 r / beta
    apply_minv(S, S.R, (size_t)n, S.W);
    std::vector<std::vector<double>> H(nr, std::vector<double>((M + 1) * M, 0.0));
    std::vector<std::vector<double>> cs(nr, std::vector<double>(M)), sn(nr, std::vector<double>(M));
    std::vector<std::vector<double>> g(nr, std::vector<double>(M + 1, 0.0));
    std::vector<int> kdone(nr, -1);
    // beta
    hipLaunchKernelGGL(multi_dot, dim3(1, (unsigned)nr), dim3(256), 0, S.st, n, nr, 1, S.W, (size_t)0, (size_t)n, S.W,
                       (size_t)n, S.dots);
    BOLA_LAUNCH_CHECK();
    std::vector<double> b2 = fetch(S, S.dots, nr);
    std::vector<double> beta(nr), inv(nr);
    for (int q = 0; q < nr; q++) {
        beta[q] = std::sqrt(std::max(0.0, b2[q]));
        inv[q] = beta[q] > 0 ? 1.0 / beta[q] : 0.0;
        g[q][0] = beta[q];
        if (beta[q] == 0) kdone[q] = 0;
    }
    put(S, S.nrm, inv);
    hipLaunchKernelGGL(col_mul, g2(n, nr), dim3(256), 0, S.st, n, nr, S.nrm, S.W, (size_t)n, S.Vb, qs);
    BOLA_LAUNCH_CHECK();
    int k = 0;
    for (; k < M; k++) {
        bool all = true;
        for (int q = 0; q < nr; q++) all = all && kdone[q] >= 0;
        if (all) break;
        // W = M^-1 A v_k
        residual(S, nullptr, 0, S.Vb + (size_t)k * js, qs, S.D);  // D = -A v_k
        apply_minv(S, S.D, (size_t)n, S.W);                        // W = -M^-1 A v_k
        // classical Gram-Schmidt, twice; coefficients accumulate in H (sign folded in)
        std::vector<double> hk(nr * (k + 2), 0.0);
        for (int pass = 0; pass < 2; pass++) {
            hipLaunchKernelGGL(multi_dot, dim3((unsigned)(k + 1), (unsigned)nr), dim3(256), 0, S.st, n, nr, k + 1,
                               S.Vb, js, qs, S.W, (size_t)n, S.dots);
            BOLA_LAUNCH_CHECK();
            hipLaunchKernelGGL(multi_axpy, g2(n, nr), dim3(256), 0, S.st, n, nr, k + 1, S.Vb, js, qs, S.dots, S.W,
                               (size_t)n);
            BOLA_LAUNCH_CHECK();
            std::vector<double> d = fetch(S, S.dots, (k + 1) * nr);
            for (int j = 0; j <= k; j++)
                for (int q = 0; q < nr; q++) hk[j * nr + q] += d[j * nr + q];
        }
        hipLaunchKernelGGL(multi_dot, dim3(1, (unsigned)nr), dim3(256), 0, S.st, n, nr, 1, S.W, (size_t)0, (size_t)n,
                           S.W, (size_t)n, S.dots);
        BOLA_LAUNCH_CHECK();
        std::vector<double> w2 = fetch(S, S.dots, nr);
        std::vector<double> sc(nr);
        for (int q = 0; q < nr; q++) {
            const double hn = std::sqrt(std::max(0.0, w2[q]));
            // W holds -(M^-1 A v_k - sum h v): the Hessenberg column is the negation
            double* Hq = H[q].data();
            for (int j = 0; j <= k; j++) Hq[j * M + k] = -hk[j * nr + q];
            Hq[(k + 1) * M + k] = hn;
            sc[q] = hn > 0 ? -1.0 / hn : 0.0;  // v_{k+1} = -W / hn
            if (kdone[q] >= 0) continue;
            // previous rotations, then a new one
            for (int j = 0; j < k; j++) {
                const double a = Hq[j * M + k], b = Hq[(j + 1) * M + k];
                Hq[j * M + k] = cs[q][j] * a + sn[q][j] * b;
                Hq[(j + 1) * M + k] = -sn[q][j] * a + cs[q][j] * b;
            }
            const double a = Hq[k * M + k], b = Hq[(k + 1) * M + k];
            const double rr = std::hypot(a, b);
            cs[q][k] = rr > 0 ? a / rr : 1.0;
            sn[q][k] = rr > 0 ? b / rr : 0.0;
            Hq[k * M + k] = rr;
            Hq[(k + 1) * M + k] = 0.0;
            g[q][k + 1] = -sn[q][k] * g[q][k];
            g[q][k] = cs[q][k] * g[q][k];
            if (std::fabs(g[q][k + 1]) <= tol_rel * beta[q] || hn == 0.0) kdone[q] = k + 1;
        }
        put(S, S.nrm, sc);
        hipLaunchKernelGGL(col_mul, g2(n, nr), dim3(256), 0, S.st, n, nr, S.nrm, S.W, (size_t)n,
                           S.Vb + (size_t)(k + 1) * js, qs);
        BOLA_LAUNCH_CHECK();
    }
    // y = R^-1 g per column; d = V y
    std::vector<double> Y((size_t)M * nr, 0.0);
    for (int q = 0; q < nr; q++) {
        const int kq = kdone[q] >= 0 ? kdone[q] : k;
        const double* Hq = H[q].data();
        for (int i = kq - 1; i >= 0; i--) {
            double v = g[q][i];
            for (int j = i + 1; j < kq; j++) v -= Hq[i * M + j] * Y[(size_t)j * nr + q];
            Y[(size_t)i * nr + q] = Hq[i * M + i] != 0 ? v / Hq[i * M + i] : 0.0;
        }
    }
    hipLaunchKernelGGL(mat_zero, g2(n, nr), dim3(256), 0, S.st, n, nr, Dout, (size_t)n);
    BOLA_LAUNCH_CHECK();
    // Dout -= sum_j (-y_j) v_j
    std::vector<double> negY(Y.size());
    for (size_t e = 0; e < Y.size(); e++) negY[e] = -Y[e];
    put(S, S.dots, negY);
    hipLaunchKernelGGL(multi_axpy, g2(n, nr), dim3(256), 0, S.st, n, nr, std::max(k, 1), S.Vb, js, qs, S.dots, Dout,
                       (size_t)n);
    BOLA_LAUNCH_CHECK();
    return k;
}

}  // namespace

}  // namespace bola

namespace bola {

// ---- the mixed-precision factorization as an object: factor once, solve many times
//
// Owned by a MixedFactor: the scaled, transposed low-precision factors T (ldt x n), the
// inverses of their diagonal blocks, pivots (tournament) or butterflies (RBT), the
// power-of-two scalings, |A| row sums (for |A|_inf); the FP64 matrix the residuals use
// (the caller's, or a copy); and, once refinement has failed for it, its FP64 LU.
struct MixedFactor {
    int n = 0, lp = 0;
    bool rbt = false;
    size_t ldt = 0;
    float* T = nullptr;
    float* Dinv = nullptr;
    int32_t* ipiv = nullptr;
    int* perm = nullptr;
    double *s = nullptr, *t = nullptr, *rsum = nullptr;
    double* rl[2] = {};
    double* rr[2] = {};
    const double* A = nullptr;
    size_t lda = 0;
    double anorm = 0.0;
    float factor_ms = 0.f;
    double* F = nullptr;     // FP64 LU of A (ld n) and its pivots, after a failed refinement
    int32_t* fpiv = nullptr;
    int finfo = 0;
    bool fready = false;
};

namespace {

struct Carve {
    size_t off = 0;
    size_t take(size_t b) {
        size_t o = off;
        off += (b + 255) & ~(size_t)255;
        return o;
    }
};

// the factor's own buffers, in one block
size_t factor_layout(MixedFactor& f, char* base) {
    const int n = f.n;
    Carve c;
    const size_t oT = c.take(sizeof(float) * f.ldt * n);
    const size_t odi = c.take(sizeof(float) * 2 * (size_t)((n + TB - 1) / TB) * TB * TB);
    const size_t oipiv = c.take(sizeof(int32_t) * n), operm = c.take(sizeof(int) * n);
    const size_t os = c.take(8 * (size_t)n), ot = c.take(8 * (size_t)n), orsum = c.take(8 * (size_t)n);
    const size_t orb = c.take(8 * 4 * (size_t)n);
    if (base) {
        f.T = (float*)(base + oT);
        f.Dinv = (float*)(base + odi);
        f.ipiv = (int32_t*)(base + oipiv);
        f.perm = (int*)(base + operm);
        f.s = (double*)(base + os);
        f.t = (double*)(base + ot);
        f.rsum = (double*)(base + orsum);
        for (int l = 0; l < 2; l++) {
            f.rl[l] = (double*)(base + orb) + (size_t)l * n;
            f.rr[l] = (double*)(base + orb) + (size_t)(2 + l) * n;
        }
    }
    return c.off;
}

// scratch of the factorization step
struct BuildScratch {
    double *pmax, *psum;
    void* luws;
    int* err;
};
size_t build_layout(int n, int lp, char* base, BuildScratch* b) {
    Carve c;
    const size_t opmax = c.take(8 * (size_t)n * RCH), opsum = c.take(8 * (size_t)n * RCH);
    const size_t olu = c.take(mixed_lu_ws_bytes(n, lp));
    const size_t oerr = c.take(sizeof(int) * 4);
    const size_t onrm = c.take(8 * 4);
    if (base) {
        b->pmax = (double*)(base + opmax);
        b->psum = (double*)(base + opsum);
        b->luws = base + olu;
        b->err = (int*)(base + oerr);
        (void)onrm;
    }
    return c.off;
}

// scratch of one solve (refinement and GMRES-IR vectors)
size_t solve_layout(Solver& S, char* base) {
    const int n = S.n;
    Carve c;
    const size_t nv = (size_t)n * NRMAX;
    const size_t oR = c.take(8 * nv), oZ = c.take(8 * nv), oD = c.take(8 * nv), oW = c.take(8 * nv);
    const size_t oP = c.take(8 * (size_t)n * NRMAX * RCH);
    const size_t onrm = c.take(8 * 4 * NRMAX), odots = c.take(8 * (size_t)(S.mgm + 2) * NRMAX * 2);
    const size_t oVb = c.take(8 * (size_t)S.nrhs * (S.mgm + 1) * n);
    const size_t oerr = c.take(sizeof(int) * 4);
    const size_t ozt = c.take(2 * 8 * nv);
    if (base) {
        S.R = (double*)(base + oR);
        S.Z = (double*)(base + oZ);
        S.D = (double*)(base + oD);
        S.W = (double*)(base + oW);
        S.P = (double*)(base + oP);
        S.nrm = (double*)(base + onrm);
        S.dots = (double*)(base + odots);
        S.Vb = (double*)(base + oVb);
        S.err = (int*)(base + oerr);
        S.zt = (uint64_t*)(base + ozt);
    }
    return c.off;
}

// Equilibration, conversion, (butterflies,) the low-precision LU and the diagonal-block
// inverses of A into f (whose buffers are carved); f.A / f.lda must be set.
void mixed_build(bola_hip_handle_s* h, MixedFactor& f, char* scratch) {
    hipStream_t st = h->stream;
    const int n = f.n;
    BuildScratch b;
    build_layout(n, f.lp, scratch, &b);
    if (f.rbt) {
        // random factors exp(u / 10), u uniform in [-1/2, 1/2), from a fixed-seed splitmix64
        std::vector<double> rv(4 * (size_t)n);
        uint64_t x = 0x5EEDB077E8F1ull;
        for (auto& v : rv) {
            uint64_t z = (x += 0x9E3779B97F4A7C15ull);
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            z ^= z >> 31;
            v = std::exp(((double)(z >> 11) * 0x1.0p-53 - 0.5) / 10.0);
        }
        BOLA_CHECK(hipMemcpyAsync(f.rl[0], rv.data(), sizeof(double) * rv.size(), hipMemcpyHostToDevice, st));
        BOLA_CHECK(hipStreamSynchronize(st));
    }
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, st, 4, b.err);
    BOLA_LAUNCH_CHECK();
    hipEvent_t ev[2];
    for (auto& e : ev) BOLA_CHECK(hipEventCreate(&e));
    BOLA_CHECK(hipEventRecord(ev[0], st));
    {
        Phase ph(h, "MIXED_FACTOR", 2.0 * (double)n * n * n / 3.0);
        // 1. equilibration and the fp32 transpose
        const int cw = (n + RCH - 1) / RCH;
        hipLaunchKernelGGL(eq_row_part, dim3(cdiv((size_t)n, 256), (unsigned)RCH), dim3(256), 0, st, n, f.A, f.lda,
                           b.pmax, b.psum, cw);
        hipLaunchKernelGGL(eq_row_fin, g1(n), dim3(256), 0, st, n, RCH, b.pmax, b.psum, f.s, f.rsum);
        hipLaunchKernelGGL(eq_col, dim3((unsigned)n), dim3(256), 0, st, n, f.A, f.lda, f.s, f.t);
        hipLaunchKernelGGL(to_T, dim3(cdiv((size_t)n, 64), cdiv((size_t)n, 64)), dim3(256), 0, st, n, f.A, f.lda, f.s,
                           f.t, f.T, f.ldt);
        hipLaunchKernelGGL(k_iota, g1(n), dim3(256), 0, st, n, f.perm);
        BOLA_LAUNCH_CHECK();
        if (f.rbt) {  // T <- (W_L^T A_s W_R)^T: left B2^T then B1^T, right B2 then B1
            const dim3 gl(cdiv((size_t)n, 256), (unsigned)std::min(n / 2, 16384));
            const dim3 gr(cdiv((size_t)n / 2, 256), (unsigned)std::min(n, 16384));
            hipLaunchKernelGGL(rbt_T<true>, gl, dim3(256), 0, st, n, f.T, f.ldt, n / 4, f.rl[1]);
            hipLaunchKernelGGL(rbt_T<true>, gl, dim3(256), 0, st, n, f.T, f.ldt, n / 2, f.rl[0]);
            hipLaunchKernelGGL(rbt_T<false>, gr, dim3(256), 0, st, n, f.T, f.ldt, n / 4, f.rr[1]);
            hipLaunchKernelGGL(rbt_T<false>, gr, dim3(256), 0, st, n, f.T, f.ldt, n / 2, f.rr[0]);
            BOLA_LAUNCH_CHECK();
        }
        // 2. the low-precision LU
        mixed_lu_factor(h, f.lp, !f.rbt, n, f.T, f.ldt, f.ipiv, f.perm, b.luws);
        const int nbk = (n + TB - 1) / TB;
        hipLaunchKernelGGL(diag_inv, dim3((unsigned)nbk, 2), dim3(TB), 0, st, n, f.T, f.ldt, f.Dinv, nbk);
        BOLA_LAUNCH_CHECK();
    }
    BOLA_CHECK(hipEventRecord(ev[1], st));
    // |A|_inf from the row sums
    double* nrm = (double*)((char*)b.err + 256);
    hipLaunchKernelGGL(col_amax, dim3(1), dim3(256), 0, st, n, f.rsum, (size_t)n, nrm);
    BOLA_LAUNCH_CHECK();
    BOLA_CHECK(hipMemcpyAsync(&f.anorm, nrm, sizeof(double), hipMemcpyDeviceToHost, st));
    BOLA_CHECK(hipStreamSynchronize(st));
    float ms = 0.f;
    BOLA_CHECK(hipEventElapsedTime(&ms, ev[0], ev[1]));
    for (auto& e : ev) (void)hipEventDestroy(e);
    f.factor_ms = ms;
}

// FP64 LU of f.A into f.F / f.fpiv (which the caller provides), once
void mixed_fp64_factor(bola_hip_handle_s* h, MixedFactor& f) {
    hipStream_t st = h->stream;
    const int n = f.n;
    Phase ph(h, "MIXED_FALLBACK");
    BOLA_CHECK(hipMemcpy2DAsync(f.F, sizeof(double) * n, f.A, sizeof(double) * f.lda, sizeof(double) * n, n,
                                hipMemcpyDeviceToDevice, st));
    lu_factor(h, n, n, f.F, n, f.fpiv, h->d_info, false);
    BOLA_CHECK(hipMemcpyAsync(&f.finfo, h->d_info, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    BOLA_CHECK(hipStreamSynchronize(st));
    f.fready = true;
}

// A X = B from the factorization f: refinement, GMRES-IR, and the FP64 LU as the last
// resort (made on first need by fp64_mem, which points f.F / f.fpiv at memory).
int mixed_solve(bola_hip_handle_s* h, MixedFactor& f, int nrhs, const double* B, size_t ldb, double* X, size_t ldx,
                bola_hip_mixed_stats& so, char* scratch, const std::function<void()>& fp64_mem) {
    const int n = f.n;
    hipStream_t st = h->stream;
    Solver S;
    S.h = h;
    S.st = st;
    S.n = n;
    S.nrhs = nrhs;
    S.lp = f.lp;
    S.A = f.A;
    S.lda = f.lda;
    S.mgm = 30;
    S.ldt = f.ldt;
    S.rbt = f.rbt;
    S.T = f.T;
    S.Dinv = f.Dinv;
    S.ipiv = f.ipiv;
    S.perm = f.perm;
    S.s = f.s;
    S.t = f.t;
    S.rsum = f.rsum;
    for (int l = 0; l < 2; l++) {
        S.rl[l] = f.rl[l];
        S.rr[l] = f.rr[l];
    }
    solve_layout(S, scratch);
    so.factor_ms = f.factor_ms;
    hipLaunchKernelGGL(k_zero_int, dim3(256), dim3(256), 0, st, (int)(4 * (size_t)n * NRMAX), (int*)S.zt);
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, st, 4, S.err);
    BOLA_LAUNCH_CHECK();
    hipEvent_t ev[3];
    for (auto& e : ev) BOLA_CHECK(hipEventCreate(&e));
    BOLA_CHECK(hipEventRecord(ev[0], st));
    hipLaunchKernelGGL(col_amax, dim3((unsigned)nrhs), dim3(256), 0, st, n, B, ldb, S.nrm + 3 * NRMAX);
    BOLA_LAUNCH_CHECK();
    const std::vector<double> bn = fetch(S, S.nrm + 3 * NRMAX, nrhs);
    const double anorm = f.anorm;
    const double tol = (double)n * 2.220446049250313e-16;
    so.tolerance = tol;
    std::vector<double> e(nrhs, INFINITY);
    bool done = false;
    if (!f.fready) {
        // 3. refinement: X = 0, R = B
        hipLaunchKernelGGL(mat_zero, g2(n, nrhs), dim3(256), 0, st, n, nrhs, X, ldx);
        hipLaunchKernelGGL(mat_copy, g2(n, nrhs), dim3(256), 0, st, n, nrhs, B, ldb, S.R, (size_t)n);
        BOLA_LAUNCH_CHECK();
        double prev = INFINITY;
        int stall = 0;
        Phase ph(h, "MIXED_REFINE");
        for (int it = 0; it < 30 && !done; it++) {
            apply_minv(S, S.R, (size_t)n, S.D);
            hipLaunchKernelGGL(mat_add, g2(n, nrhs), dim3(256), 0, st, n, nrhs, S.D, (size_t)n, X, ldx);
            BOLA_LAUNCH_CHECK();
            residual(S, B, ldb, X, ldx, S.R);
            e = berr(S, X, ldx, bn, anorm);
            so.ir_iterations = it + 1;
            const double m = maxv(e);
            if (it == 0) so.first_berr = m;
            if (m <= tol) { done = true; break; }
            if (!(m < 0.5 * prev)) stall++;
            else stall = 0;
            prev = std::min(prev, m);
            if (stall >= 2 || !(m < 1.0)) break;
        }
        // GMRES-IR
        if (!done) {
            for (int outer = 0; outer < 10 && !done; outer++) {
                so.gmres_iterations += gmres_cycle(S, S.D, 1e-8);
                hipLaunchKernelGGL(mat_add, g2(n, nrhs), dim3(256), 0, st, n, nrhs, S.D, (size_t)n, X, ldx);
                BOLA_LAUNCH_CHECK();
                residual(S, B, ldb, X, ldx, S.R);
                e = berr(S, X, ldx, bn, anorm);
                so.gmres_cycles = outer + 1;
                const double m = maxv(e);
                if (m <= tol) { done = true; break; }
                if (!(m < 0.5 * prev) && outer > 0) break;
                prev = std::min(prev, m);
            }
        }
        int herr = 0;
        BOLA_CHECK(hipMemcpyAsync(&herr, S.err, sizeof(int), hipMemcpyDeviceToHost, st));
        BOLA_CHECK(hipStreamSynchronize(st));
        if (herr) done = false;  // a triangular solve gave up waiting: trust nothing it produced
    }
    BOLA_CHECK(hipEventRecord(ev[1], st));
    // 4. fallback: the FP64 reference-exact path
    if (!done) {
        so.fallback = 1;
        if (!f.fready) {
            fp64_mem();
            mixed_fp64_factor(h, f);
        }
        if (f.finfo != 0) {
            for (auto& x : ev) (void)hipEventDestroy(x);
            set_error(h, "Matrix is singular to working precision (FP64 fallback factorization failed).");
            return f.finfo > 0 ? f.finfo : BOLA_HIP_ERR_DEVICE;
        }
        hipLaunchKernelGGL(mat_copy, g2(n, nrhs), dim3(256), 0, st, n, nrhs, B, ldb, X, ldx);
        BOLA_LAUNCH_CHECK();
        int r = bola_hip_dgetrs_dev(h, n, nrhs, f.F, n, f.fpiv, X, ldx);
        if (r) {
            for (auto& x : ev) (void)hipEventDestroy(x);
            return r;
        }
        residual(S, B, ldb, X, ldx, S.R);
        e = berr(S, X, ldx, bn, anorm);
    }
    BOLA_CHECK(hipEventRecord(ev[2], st));
    BOLA_CHECK(hipEventSynchronize(ev[2]));
    float t01 = 0, t12 = 0;
    BOLA_CHECK(hipEventElapsedTime(&t01, ev[0], ev[1]));
    BOLA_CHECK(hipEventElapsedTime(&t12, ev[1], ev[2]));
    for (auto& x : ev) (void)hipEventDestroy(x);
    so.refine_ms = t01;
    so.fallback_ms = t12;
    so.final_berr = maxv(e);
    so.converged = so.final_berr <= tol ? 1 : 0;
    return 0;
}

int mixed_check_args(bola_hip_handle_s* h, int precision, size_t n, size_t lda, const char* who) {
    precision &= 0xFF;
    if (precision > LP_F16X3) {
        set_error(h, std::string("Argument 2 to ") + who + " had an illegal value");
        return -2;
    }
    if (lda < std::max<size_t>(1, n)) {
        set_error(h, std::string("Argument 6 to ") + who + " had an illegal value");
        return -6;
    }
    if (n > (size_t)mixed_lu_max_n()) {
        set_error(h, std::string("Argument 3 to ") + who + " had an illegal value (n above the supported maximum)");
        return -3;
    }
    return 0;
}

void mixed_init(MixedFactor& f, int precision, int n) {
    const bool want_rbt = (precision & BOLA_HIP_MIX_RBT) != 0;
    f.n = n;
    f.lp = precision & 0xFF;
    f.ldt = ((size_t)n + 63) & ~(size_t)63;
    f.rbt = want_rbt && n % 4 == 0 && n >= 8;  // two butterfly levels need n divisible by 4
}

}  // namespace
}  // namespace bola

using namespace bola;

extern "C" int bola_hip_dgesv_mixed_dev(bola_hip_handle_t h, int precision, size_t n_, size_t nrhs_, const double* A,
                                        size_t lda, const double* B, size_t ldb, double* X, size_t ldx,
                                        bola_hip_mixed_stats* stats) {
    BOLA_API_BEGIN
    if (!h) return -1;
    if (int r = mixed_check_args(h, precision, n_, lda, "dgesv_mixed")) return r;
    if (nrhs_ > (size_t)NRMAX) {
        set_error(h, "Argument 4 to dgesv_mixed had an illegal value (at most 16 right-hand sides)");
        return -4;
    }
    bola_hip_mixed_stats loc{};
    bola_hip_mixed_stats& so = stats ? *stats : loc;
    so = bola_hip_mixed_stats{};
    if (n_ == 0 || nrhs_ == 0) return 0;
    const int n = (int)n_, nrhs = (int)nrhs_;
    MixedFactor f;
    mixed_init(f, precision, n);
    f.A = A;
    f.lda = lda;
    // one workspace: the factor's buffers, then the build scratch, then the solve scratch
    Solver probe;
    probe.n = n;
    probe.nrhs = nrhs;
    probe.mgm = 30;
    const size_t fb = factor_layout(f, nullptr);
    const size_t bb = build_layout(n, f.lp, nullptr, nullptr);
    const size_t sb = solve_layout(probe, nullptr);
    char* base = (char*)h->mws.get(fb + bb + sb);
    factor_layout(f, base);
    mixed_build(h, f, base + fb);
    return mixed_solve(h, f, nrhs, B, ldb, X, ldx, so, base + fb + bb, [&] {
        f.F = (double*)h->mws2.get(sizeof(double) * (size_t)n * n + sizeof(int32_t) * n + 512);
        f.fpiv = (int32_t*)(f.F + (size_t)n * n);
    });
    BOLA_API_END(h)
}


// ------------------------------------------------------------------ factorization objects

struct bola_hip_factor_s {
    int kind = 0, precision = 0;
    size_t n = 0;
    int device = 0;
    bool stream_ordered = false;  // memory from hipMallocAsync (the server's), else hipMalloc
    void* mem = nullptr;      // every device buffer the object owns except the fallback
    size_t bytes = 0;
    double* a64 = nullptr;    // FP64 LU / Cholesky factor, or (mixed) the copy of A (ld n)
    int32_t* ipiv = nullptr;  // LU pivots (1-based, global)
    bola::MixedFactor mf;
    void* fb_mem = nullptr;   // mixed: the FP64 LU made on first need
};

namespace bola {

// Builds a factorization object of A (host or device memory); its device memory comes from
// the stream-ordered allocator on the handle's stream when `stream_ordered` (serving), else
// from hipMalloc.  Returns null with *info set when the reference's factorization reports
// an error (zero pivot, not positive definite; the text is in the handle's error string).
bola_hip_factor_s* factor_build(bola_hip_handle_s* h, int kind, int precision, size_t n, const double* a, size_t lda,
                                int* info, bool stream_ordered) {
    hipStream_t st = h->stream;
    *info = 0;
    auto* f = new bola_hip_factor_s;
    f->kind = kind;
    f->precision = kind == BOLA_HIP_FACTOR_MIXED ? precision : 0;
    f->n = n;
    f->device = h->device;
    f->stream_ordered = stream_ordered;
    try {
        const size_t mat = (sizeof(double) * n * n + 255) & ~(size_t)255;
        size_t extra = 0;
        if (kind == BOLA_HIP_FACTOR_LU) extra = (sizeof(int32_t) * n + 255) & ~(size_t)255;
        if (kind == BOLA_HIP_FACTOR_MIXED) {
            mixed_init(f->mf, precision, (int)n);
            extra = factor_layout(f->mf, nullptr);
        }
        f->bytes = mat + extra;
        if (stream_ordered) BOLA_CHECK(hipMallocAsync(&f->mem, f->bytes, st));
        else BOLA_CHECK(hipMalloc(&f->mem, f->bytes));
        f->a64 = (double*)f->mem;
        BOLA_CHECK(hipMemcpy2DAsync(f->a64, sizeof(double) * n, a, sizeof(double) * lda, sizeof(double) * n, n,
                                    hipMemcpyDefault, st));
        int32_t rc = 0;
        if (kind == BOLA_HIP_FACTOR_LU) {
            f->ipiv = (int32_t*)((char*)f->mem + mat);
            rc = bola_hip_dgetrf_dev(h, n, n, f->a64, n, f->ipiv, nullptr);
        } else if (kind == BOLA_HIP_FACTOR_CHOL_L || kind == BOLA_HIP_FACTOR_CHOL_U) {
            rc = bola_hip_dpotrf_dev(h, kind == BOLA_HIP_FACTOR_CHOL_L ? 'L' : 'U', n, f->a64, n, nullptr);
        } else {
            factor_layout(f->mf, (char*)f->mem + mat);
            f->mf.A = f->a64;
            f->mf.lda = n;
            char* scratch = (char*)h->mws.get(build_layout((int)n, f->mf.lp, nullptr, nullptr));
            mixed_build(h, f->mf, scratch);
        }
        if (rc < 0) throw HipFailure{hipErrorUnknown, "factorization", __LINE__};
        BOLA_CHECK(hipStreamSynchronize(st));
        if (rc > 0) {
            *info = rc;
            factor_release(f, st);
            return nullptr;
        }
        return f;
    } catch (...) {
        factor_release(f, st);
        throw;
    }
}

// Frees the object's device memory (stream-ordered on st when it was so allocated).
void factor_release(bola_hip_factor_s* f, hipStream_t st) {
    if (!f) return;
    if (f->stream_ordered) {
        if (f->mem) (void)hipFreeAsync(f->mem, st);
        if (f->fb_mem) (void)hipFreeAsync(f->fb_mem, st);
    } else {
        if (f->mem) (void)hipFree(f->mem);
        if (f->fb_mem) (void)hipFree(f->fb_mem);
    }
    delete f;
}

}  // namespace bola

extern "C" int bola_hip_factor_create_dev(bola_hip_handle_t h, int kind, int precision, size_t n, const double* a,
                                          size_t lda, bola_hip_factor_t* out) {
    if (!h) h = default_handle();
    if (!out) return -7;
    *out = nullptr;
    BOLA_API_BEGIN
    if (kind < BOLA_HIP_FACTOR_LU || kind > BOLA_HIP_FACTOR_MIXED) {
        set_error(h, "Argument 2 to factor_create had an illegal value");
        return -2;
    }
    if (kind == BOLA_HIP_FACTOR_MIXED) {
        if (int r = mixed_check_args(h, precision, n, lda, "factor_create")) return r;
    } else if (lda < std::max<size_t>(1, n)) {
        set_error(h, "Argument 6 to factor_create had an illegal value");
        return -6;
    }
    if (n == 0 || !a) {
        set_error(h, "factor_create needs n > 0 and a matrix");
        return -4;
    }
    int info = 0;
    bola_hip_factor_s* f = factor_build(h, kind, precision, n, a, lda, &info, false);
    if (!f) return info;
    *out = f;
    return 0;
    BOLA_API_END(h)
}

extern "C" int bola_hip_factor_solve_dev(bola_hip_handle_t h, bola_hip_factor_t f, size_t nrhs, const double* b,
                                         size_t ldb, double* x, size_t ldx, bola_hip_mixed_stats* stats) {
    if (!h) h = default_handle();
    BOLA_API_BEGIN
    if (!f) {
        set_error(h, "Argument 2 to factor_solve had an illegal value");
        return -2;
    }
    const size_t n = f->n;
    if (ldb < std::max<size_t>(1, n) || ldx < std::max<size_t>(1, n)) {
        set_error(h, "factor_solve: leading dimensions must be at least n");
        return -5;
    }
    bola_hip_mixed_stats loc{};
    bola_hip_mixed_stats& so = stats ? *stats : loc;
    so = bola_hip_mixed_stats{};
    if (nrhs == 0) return 0;
    hipStream_t st = h->stream;
    if (f->kind != BOLA_HIP_FACTOR_MIXED) {
        if (x != b)
            BOLA_CHECK(hipMemcpy2DAsync(x, sizeof(double) * ldx, b, sizeof(double) * ldb, sizeof(double) * n, nrhs,
                                        hipMemcpyDeviceToDevice, st));
        int r = f->kind == BOLA_HIP_FACTOR_LU
                    ? bola_hip_dgetrs_dev(h, n, nrhs, f->a64, n, f->ipiv, x, ldx)
                    : bola_hip_dpotrs_dev(h, f->kind == BOLA_HIP_FACTOR_CHOL_L ? 'L' : 'U', n, nrhs, f->a64, n, x, ldx);
        if (r) return r;
        so.converged = 1;
        return 0;
    }
    // mixed: NRMAX right-hand sides at a time
    const int ni = (int)n;
    Solver probe;
    probe.n = ni;
    probe.nrhs = NRMAX;
    probe.mgm = 30;
    char* scratch = (char*)h->mws.get(solve_layout(probe, nullptr));
    auto fp64_mem = [&] {
        const size_t fbb = sizeof(double) * n * n + sizeof(int32_t) * n + 512;
        if (!f->fb_mem) {
            if (f->stream_ordered) BOLA_CHECK(hipMallocAsync(&f->fb_mem, fbb, st));
            else BOLA_CHECK(hipMalloc(&f->fb_mem, fbb));
        }
        f->mf.F = (double*)f->fb_mem;
        f->mf.fpiv = (int32_t*)(f->mf.F + n * n);
    };
    double worst = 0.0;
    int conv = 1;
    for (size_t c0 = 0; c0 < nrhs; c0 += NRMAX) {
        const int nc = (int)std::min<size_t>(NRMAX, nrhs - c0);
        bola_hip_mixed_stats part{};
        int r = mixed_solve(h, f->mf, nc, b + c0 * ldb, ldb, x + c0 * ldx, ldx, part, scratch, fp64_mem);
        if (r) return r;
        so.ir_iterations += part.ir_iterations;
        so.gmres_cycles += part.gmres_cycles;
        so.gmres_iterations += part.gmres_iterations;
        so.fallback |= part.fallback;
        so.refine_ms += part.refine_ms;
        so.fallback_ms += part.fallback_ms;
        so.tolerance = part.tolerance;
        if (c0 == 0) so.first_berr = part.first_berr;
        worst = std::max(worst, part.final_berr);
        conv = conv && part.converged;
    }
    so.factor_ms = f->mf.factor_ms;
    so.final_berr = worst;
    so.converged = conv;
    return 0;
    BOLA_API_END(h)
}

extern "C" int bola_hip_factor_describe(bola_hip_factor_t f, bola_hip_factor_desc* d) {
    if (!f || !d) return -1;
    *d = bola_hip_factor_desc{};
    d->kind = f->kind;
    d->precision = f->precision;
    d->n = f->n;
    d->bytes = f->bytes + (f->fb_mem ? sizeof(double) * f->n * f->n + sizeof(int32_t) * f->n + 512 : 0);
    if (f->kind == BOLA_HIP_FACTOR_MIXED) {
        d->factor = f->mf.T;
        d->ld = f->mf.ldt;
        d->ipiv = f->mf.rbt ? nullptr : f->mf.ipiv;
        d->row_scale = f->mf.s;
        d->col_scale = f->mf.t;
        d->matrix = f->a64;
        d->butterfly = f->mf.rbt ? 1 : 0;
    } else {
        d->factor = f->a64;
        d->ld = f->n;
        d->ipiv = f->ipiv;
    }
    return 0;
}

extern "C" size_t bola_hip_factor_bytes(int kind, int precision, size_t n) {
    const size_t mat = (sizeof(double) * n * n + 255) & ~(size_t)255;
    if (kind == BOLA_HIP_FACTOR_LU) return mat + ((sizeof(int32_t) * n + 255) & ~(size_t)255);
    if (kind == BOLA_HIP_FACTOR_MIXED) {
        MixedFactor mf;
        mixed_init(mf, precision, (int)n);
        return mat + factor_layout(mf, nullptr);
    }
    return mat;
}

extern "C" int bola_hip_factor_destroy(bola_hip_factor_t f) {
    if (!f) return 0;
    (void)hipSetDevice(f->device);
    (void)hipDeviceSynchronize();  // no solve may still read it (callers' streams included)
    factor_release(f, nullptr);
    return 0;
}
