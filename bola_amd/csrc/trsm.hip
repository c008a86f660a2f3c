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
// below run those exact sequences: the daxpy-form solves with one lane per row, the
// dot-form solves with one thread per independent vector held in registers.
#include "common.hpp"

namespace bola {

namespace {

constexpr int MAXT = 64;    // base-case triangle size (trsm.rs:8)

// variant ids
enum { LNL = 0, LNU, LTL, LTU, RNL, RNU, RTL, RTU };

// ---- left, no-transpose (daxpy form): one wavefront per right-hand-side column (two
// columns per wavefront when t <= 32), one lane per row.  Each lane loads its row of the
// triangle once; step k reads the solved entry of lane k with a lane read (no LDS), and
// every lane below (LNL) / above (LNU) applies its own fused or unfused update, the same
// operation the reference's daxpy applies to that element.
__device__ __forceinline__ double rdlane(double v, int l) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)u, l);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(u >> 32), l);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Unit diagonal: each step is straight-line code, the reference's skips (a zero
// right-hand-side entry) becoming selects, so a wavefront alone on its SIMD runs no branch
// per step.  Non-unit diagonal: the division sits on the step chain and the skips stay
// branches (measured faster for that case).  Every element sees exactly the reference's
// operation or none.
template <bool LOWER, int TW, bool UNIT>
__global__ void __launch_bounds__(256) trsm_ln_lanes(int t, int nvec, const double* __restrict__ A, size_t lda,
                                                     double* __restrict__ B, size_t ldb,
                                                     Abort abort) {
    if (abort_stop(abort)) return;
    constexpr bool TWO = TW == 32;
    const int lane = threadIdx.x & 63;
    const int sub = TWO ? (lane >> 5) : 0;
    const int i = TWO ? (lane & 31) : lane;
    const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long v = TWO ? wave * 2 + sub : wave;
    const bool act = v < nvec && i < t;
    double arow[TW];  // row i of the triangle
#pragma unroll
    for (int k = 0; k < TW; k++) {
        const bool need = LOWER ? k <= i : k >= i;
        const double a = A[(size_t)min(i, t - 1) + (size_t)min(k, t - 1) * lda];  // clamped: loads in flight together
        arow[k] = keep_or_zero(a, i < t && k < t && need);
    }
    const double xb = B[min(i, t - 1) + (size_t)min(v, nvec - 1) * ldb];
    double x = keep_or_zero(xb, act);
    if constexpr (UNIT) {
        // steps k >= t need no test: lane k is then inactive, its x is zero, and a zero
        // entry changes nothing
#pragma unroll
        for (int kk = 0; kk < TW; kk++) {
            const int k = LOWER ? kk : (TW - 1 - kk);
            double bk = rdlane(x, k);
            if (TWO) {
                const double b1 = rdlane(x, 32 + k);
                bk = sub ? b1 : bk;
            }
            // reference: nothing happens for a zero entry; else x_k = b_k and the rows past
            // k take -x_k times their coefficient, fused on the 4-aligned part
            const bool nz = bk != 0.0;
            const double al = -bk;
            const double a = arow[k];
            const int nf = LOWER ? ((t - k - 1) & ~3) : (k & ~3);
            const bool fused = LOWER ? (i - k - 1 < nf) : (i < nf);
            const double y = fused ? fma(al, a, x) : x + al * a;
            const bool upd = nz && act && (LOWER ? i > k : i < k);
            x = upd ? y : x;
        }
    } else {
#pragma unroll
        for (int kk = 0; kk < TW; kk++) {
            const int k = LOWER ? kk : (TW - 1 - kk);
            if (k < t) {
                double bk = rdlane(x, k);
                if (TWO) {
                    const double b1 = rdlane(x, 32 + k);
                    bk = sub ? b1 : bk;
                }
                if (bk != 0.0) {
                    const double u = bk / rdlane(arow[k], k);
                    if (i == k) x = u;
                    const double al = -u;
                    if (al != 0.0) {
                        if (LOWER) {
                            const int nf = (t - k - 1) & ~3;
                            if (act && i > k) x = (i - k - 1 < nf) ? fma(al, arow[k], x) : x + al * arow[k];
                        } else {
                            const int nf = k & ~3;
                            if (act && i < k) x = (i < nf) ? fma(al, arow[k], x) : x + al * arow[k];
                        }
                    }
                }
            }
        }
    }
    if (act) B[i + (size_t)v * ldb] = x;
}

// ---- dot-form variants (LTL, LTU and the four right-side ones): one thread per
// independent vector, the vector in registers (compile-time indices), the triangle in
// LDS (broadcast reads).  Sequences as in trsm.rs:174-241 (left) and :252-376 (right).
// TT > 0: the triangle is exactly TT wide (the recursion's usual 64, and the Cholesky
// block's 32), so every step is straight-line code with no size test; TT == 0: any
// width up to 64, tested at run time.  The right-side forms multiply by the reciprocal of
// the diagonal entry; it is computed once per workgroup (the same correctly rounded
// 1.0 / a as the reference's per-use one).
template <int V, int TT>
__global__ void __launch_bounds__(64) trsm_dot_regs(int unit, int t_rt, int nvec, const double* __restrict__ A,
                                                    size_t lda, double* __restrict__ B, size_t ldb,
                                                    Abort abort) {
    if (abort_stop(abort)) return;
    constexpr int W = TT > 0 ? TT : MAXT;
    const int t = TT > 0 ? TT : t_rt;
    __shared__ double sA[W * (W + 1)];  // sA[r * (W+1) + c] = A[r, c]
    __shared__ double rinv[W];
    const int lane = threadIdx.x;
    for (int e = lane; e < t * t; e += 64) {
        int r = e % t, c = e / t;
        sA[r * (W + 1) + c] = A[r + (size_t)c * lda];
    }
    constexpr bool LEFT = V <= LTU;
    if (!LEFT && !unit)
        for (int q = lane; q < t; q += 64) rinv[q] = 1.0 / A[q + (size_t)q * lda];
    __syncthreads();
    const long v = (long)blockIdx.x * 64 + lane;
    if (v >= nvec) return;
#define AT(r, c) sA[(r) * (W + 1) + (c)]
    double x[W];
#pragma unroll
    for (int q = 0; q < W; q++)
        x[q] = LEFT ? B[min(q, t - 1) + (size_t)v * ldb] : B[v + (size_t)min(q, t - 1) * ldb];  // clamped
    if (TT == 0) {
#pragma unroll
        for (int q = 0; q < W; q++) x[q] = keep_or_zero(x[q], q < t);
    }
    if (V == LTL) {
#pragma unroll
        for (int ii = 0; ii < W; ii++) {
            const int i = W - 1 - ii;
            if (TT > 0 || i < t) {
                double temp = x[i];
#pragma unroll
                for (int k = i + 1; k < W; k++)
                    if (TT > 0 || k < t) temp = temp - AT(k, i) * x[k];
                if (!unit) temp = temp / AT(i, i);
                x[i] = temp;
            }
        }
    } else if (V == LTU) {
#pragma unroll
        for (int i = 0; i < W; i++) {
            if (TT > 0 || i < t) {
                double temp = x[i];
#pragma unroll
                for (int k = 0; k < i; k++) temp = temp - AT(k, i) * x[k];
                if (!unit) temp = temp / AT(i, i);
                x[i] = temp;
            }
        }
    } else if (V == RNL || V == RTU) {
#pragma unroll
        for (int jj = 0; jj < W; jj++) {
            const int j = W - 1 - jj;
            if (TT > 0 || j < t) {
                double bj = x[j];
#pragma unroll
                for (int k = j + 1; k < W; k++) {
                    if (TT > 0 || k < t) {
                        const double a = (V == RNL) ? AT(k, j) : AT(j, k);
                        const double y = bj - a * x[k];
                        bj = a != 0.0 ? y : bj;
                    }
                }
                if (!unit) bj = bj * rinv[j];
                x[j] = bj;
            }
        }
    } else {  // RNU, RTL
#pragma unroll
        for (int j = 0; j < W; j++) {
            if (TT > 0 || j < t) {
                double bj = x[j];
#pragma unroll
                for (int k = 0; k < j; k++) {
                    const double a = (V == RNU) ? AT(k, j) : AT(j, k);
                    const double y = bj - a * x[k];
                    bj = a != 0.0 ? y : bj;
                }
                if (!unit) bj = bj * rinv[j];
                x[j] = bj;
            }
        }
    }
#undef AT
#pragma unroll
    for (int q = 0; q < W; q++)
        if (TT > 0 || q < t) {
            if (LEFT) B[q + (size_t)v * ldb] = x[q];
            else B[v + (size_t)q * ldb] = x[q];
        }
}

template <int V>
void base(hipStream_t st, int unit, size_t t, size_t nvec, const double* a, size_t lda, double* b, size_t ldb,
          Abort abort) {
    if (t == 0 || nvec == 0) return;
    if (V == LNL || V == LNU) {
        const size_t waves = t <= 32 ? (nvec + 1) / 2 : nvec;
        const unsigned blocks = (unsigned)((waves + 3) / 4);
        const int tt = (int)t, nv = (int)nvec;
        if (t <= 32) {
            if (unit)
                hipLaunchKernelGGL((trsm_ln_lanes<V == LNL, 32, true>), dim3(blocks), dim3(256), 0, st, tt, nv, a, lda,
                                   b, ldb, abort);
            else
                hipLaunchKernelGGL((trsm_ln_lanes<V == LNL, 32, false>), dim3(blocks), dim3(256), 0, st, tt, nv, a,
                                   lda, b, ldb, abort);
        } else {
            if (unit)
                hipLaunchKernelGGL((trsm_ln_lanes<V == LNL, 64, true>), dim3(blocks), dim3(256), 0, st, tt, nv, a, lda,
                                   b, ldb, abort);
            else
                hipLaunchKernelGGL((trsm_ln_lanes<V == LNL, 64, false>), dim3(blocks), dim3(256), 0, st, tt, nv, a,
                                   lda, b, ldb, abort);
        }
    } else {
        const unsigned blocks = (unsigned)((nvec + 63) / 64);
        if (t == 64)
            hipLaunchKernelGGL((trsm_dot_regs<V, 64>), dim3(blocks), dim3(64), 0, st, unit, (int)t, (int)nvec, a, lda,
                               b, ldb, abort);
        else if (t == 32)
            hipLaunchKernelGGL((trsm_dot_regs<V, 32>), dim3(blocks), dim3(64), 0, st, unit, (int)t, (int)nvec, a, lda,
                               b, ldb, abort);
        else
            hipLaunchKernelGGL((trsm_dot_regs<V, 0>), dim3(blocks), dim3(64), 0, st, unit, (int)t, (int)nvec, a, lda,
                               b, ldb, abort);
    }
    BOLA_LAUNCH_CHECK();
}

// ---- persistent left/no-transpose solve (one launch for the whole recursion)
//
// For B of m x n with m = 64 * 2^k (2 to 256 leaves) and n % 16 == 0, the reference's
// recursion (trsm.rs:95-170) reduces, element by element, to a fixed sequence: every gemm
// of the recursion has k a multiple of 64 starting on a leaf boundary, m a multiple of 8
// and n a multiple of 4, so a row of leaf b receives one fold per 64-row slab it depends
// on -- acc = sum of the slab's 64 products by an FMA chain from zero, C = fma(-1, acc, C)
// (gemm.rs:232-270) -- and then leaf b's base-case solve.  The lower solve folds slabs
// 0 .. b-1 in increasing order; the upper solve folds the recursion's right siblings from
// the outermost in, each sibling's slabs in increasing order.  One workgroup per (leaf,
// 16-column group) walks that list, waiting for each slab's leaf to be published, folds it
// with v_mfma_f64_16x16x4f64 (which rounds like four chained FMAs in k order), solves its
// leaf and publishes it.  A workgroup waits only for leaves that come earlier in the
// solve, and workgroups are numbered in that order, so the one that is waited for has
// always been dispatched.
constexpr int PLN_COLS = 16;  // right-hand-side columns per workgroup
typedef double pln_d4 __attribute__((ext_vector_type(4)));

template <int K>
__device__ __forceinline__ double row_bcast(double x) {  // lane K of each row of 16, to the row
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), 0x150 + K, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), 0x150 + K, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// Leaf of 64 rows for the 4 columns of one wavefront: lane (r = lane & 15, cq = lane >> 4)
// holds rows 16 q + r of its column in x[q].  The leaf is solved in 16-row blocks: block p's
// 16 steps (the wavefront's rows of block p exchange the solved entry with a row broadcast),
// then the rows of the other blocks that the steps update take them, in step order, from
// the multipliers each lane kept -- the same operation on each element, in the same order,
// as the reference's step-by-step daxpy form (trsm.rs:95-170, utilities.rs:107-132).

// step S of block P's own 16 steps (step k = S for the lower solve, 15 - S for the upper);
// ad: this lane's row of the block's diagonal part, dg: the diagonal, both in registers so
// the step chain waits on nothing but the row broadcast (and the division)
template <bool LOWER, bool UNIT, int P, int S>
__device__ __forceinline__ void pln_diag(double& xp, double (&al)[16], bool (&act)[16], const double (&ad)[16],
                                         const double (&dg)[16], int r) {
    if constexpr (S < 16) {
        constexpr int k = LOWER ? S : 15 - S;
        constexpr int kk = 16 * P + k;
        const int i = 16 * P + r;
        const double bk = row_bcast<k>(xp);
        const double u = UNIT ? bk : bk / dg[k];
        const bool nz = bk != 0.0;
        if constexpr (!UNIT) xp = (nz && r == k) ? u : xp;  // with a unit diagonal u is lane k's own x
        al[k] = -u;
        act[k] = nz && al[k] != 0.0;
        const bool fused = LOWER ? (i - kk - 1 < ((63 - kk) & ~3)) : (i < (kk & ~3));
        const double y = fused ? fma(al[k], ad[k], xp) : xp + al[k] * ad[k];
        xp = (act[k] && (LOWER ? r > k : r < k)) ? y : xp;
        pln_diag<LOWER, UNIT, P, S + 1>(xp, al, act, ad, dg, r);
    }
}

template <bool LOWER, bool UNIT, int P>
__device__ __forceinline__ void pln_block(double (&x)[4], const double (*la)[65], int r) {
    double al[16], ad[16], dg[16];
    bool act[16];
#pragma unroll
    for (int k = 0; k < 16; k++) {
        ad[k] = la[16 * P + r][16 * P + k];
        dg[k] = UNIT ? 1.0 : la[16 * P + k][16 * P + k];
    }
    pln_diag<LOWER, UNIT, P, 0>(x[P], al, act, ad, dg, r);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        if (LOWER ? q > P : q < P) {
            const int i = 16 * q + r;
            double aq[16];
#pragma unroll
            for (int k = 0; k < 16; k++) aq[k] = la[i][16 * P + k];
            double xq = x[q];
#pragma unroll
            for (int s = 0; s < 16; s++) {
                const int k = LOWER ? s : 15 - s;
                const int kk = 16 * P + k;
                const bool fused = LOWER ? (i - kk - 1 < ((63 - kk) & ~3)) : (i < (kk & ~3));
                const double y = fused ? fma(al[k], aq[k], xq) : xq + al[k] * aq[k];
                xq = act[k] ? y : xq;
            }
            x[q] = xq;
        }
    }
}

// Upper solve: the products leaf b keeps at most at once (its first sibling group's size
// less one; 0 for the last leaf), and the number kept by the leaves before it.  Leaves with
// j leading one bits (of log2 nb) form the range [nb - nb/2^j, nb - nb/2^(j+1)) and keep
// nb/2^(j+1) - 1 each.
__host__ __device__ inline int pln_lead_ones(int nb, int b) {
    int j = 0;
    for (int h = nb >> 1; h >= 1 && (b & h); h >>= 1) j++;
    return j;
}
__host__ __device__ inline int pln_kept_slots(int nb, int b) {
    const int g = nb >> (pln_lead_ones(nb, b) + 1);
    return g > 0 ? g - 1 : 0;
}
__host__ __device__ inline long pln_kept_offset(int nb, int b) {
    const int j = pln_lead_ones(nb, b);
    long off = 0;
    for (int q = 0; q < j; q++) {
        const long g = nb >> (q + 1);
        off += g * (g - 1);  // range q: g leaves keeping g - 1 each
    }
    const long start = nb - (nb >> j);
    return off + (b - start) * (long)pln_kept_slots(nb, b);
}

// The slabs leaf b folds, in the recursion's order: the ranges of the siblings on its path
// from the root (left siblings for the lower solve, right ones for the upper), outermost first.
struct PlnWalk {
    int lo, hi;  // the node whose sibling ranges are still to come
    int s, e;    // the current range
};
template <bool LOWER>
__device__ __forceinline__ bool pln_next(PlnWalk& wk, int b, int& out) {
    while (wk.s >= wk.e) {
        if (wk.hi - wk.lo <= 1) return false;
        const int mid = (wk.lo + wk.hi) >> 1;
        if (LOWER) {
            if (b >= mid) {
                wk.s = wk.lo;
                wk.e = mid;
                wk.lo = mid;
            } else {
                wk.hi = mid;
            }
        } else {
            if (b < mid) {
                wk.s = mid;
                wk.e = wk.hi;
                wk.hi = mid;
            } else {
                wk.lo = mid;
            }
        }
    }
    out = wk.s++;
    return true;
}

template <bool LOWER, bool UNIT>
__global__ void __launch_bounds__(256) trsm_ln_persistent(int nb, int ncg, const double* __restrict__ A, size_t lda,
                                                          double* __restrict__ B, size_t ldb, int* __restrict__ flags,
                                                          double* __restrict__ kept, long spin_limit) {
    __shared__ double la[64][65];             // the leaf's diagonal block of the triangle, [row][col]
    __shared__ double tb[PLN_COLS][65];       // the folded block, [col][row]
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int pos = blockIdx.x / ncg, cg = blockIdx.x - pos * ncg;
    const int b = LOWER ? pos : nb - 1 - pos;
    const size_t r0 = (size_t)b * 64, c0 = (size_t)cg * PLN_COLS;
    for (int e = tid; e < 64 * 64; e += 256) {
        const int rr = e & 63, cc = e >> 6;
        const bool need = LOWER ? cc <= rr : cc >= rr;
        la[rr][cc] = need ? A[r0 + rr + (r0 + cc) * lda] : 0.0;
    }
    // this wavefront's 16 x 16 tile of the block: rows 16 w + lk + 4 i, column lr
    const int lr = lane & 15, lk = lane >> 4;
    double crun[4];
#pragma unroll
    for (int i = 0; i < 4; i++) crun[i] = B[r0 + 16 * w + lk + 4 * i + (c0 + lr) * ldb];
    int bad = 0;
    // operands of slab s: this tile's rows of the triangle (no wait) and the slab's solved rows
    auto load_a = [&](int s, double (&fa)[16]) {
        const double* ap = A + r0 + 16 * w + lr + ((size_t)s * 64 + lk) * lda;
#pragma unroll
        for (int t = 0; t < 16; t++) fa[t] = ap[(size_t)4 * t * lda];
    };
    auto ready = [&](int s) {
        return __hip_atomic_load(flags + (size_t)s * ncg + cg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
    };
    auto wait = [&](int s) {
        long sp = 0;
        while (!ready(s)) {
            if (++sp > spin_limit) { bad = 1; break; }
            __builtin_amdgcn_s_sleep(1);
        }
    };
    auto load_b = [&](int s, double (&fb)[16]) {
        const double* bp = B + (size_t)s * 64 + lk + (c0 + lr) * ldb;
#pragma unroll
        for (int t = 0; t < 16; t++) fb[t] = __hip_atomic_load(bp + 4 * t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    };
    auto product = [&](int s) {
        double fa[16], fb[16];
        load_a(s, fa);
        wait(s);
        load_b(s, fb);
        pln_d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int t = 0; t < 16; t++) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[t], fb[t], acc, 0, 0, 0);
        return acc;
    };
    if constexpr (LOWER) {
        // the slabs arrive in the order they are folded: each is folded as it comes, the next
        // slab's operands loaded while this one's products run
        PlnWalk wk{0, nb, 0, 0};
        int s = 0;
        bool have = pln_next<LOWER>(wk, b, s);
        double fa[16], fb[16];
        if (have) {
            load_a(s, fa);
            wait(s);
            load_b(s, fb);
        }
        while (have) {
            int sn = 0;
            const bool hn = pln_next<LOWER>(wk, b, sn);
            double na[16], nv[16];
            bool got = false;
            if (hn) {
                load_a(sn, na);
                if (ready(sn)) {
                    load_b(sn, nv);
                    got = true;
                }
            }
            pln_d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int t = 0; t < 16; t++) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[t], fb[t], acc, 0, 0, 0);
#pragma unroll
            for (int i = 0; i < 4; i++) crun[i] = fma(-1.0, acc[i], crun[i]);
            if (hn) {
                if (!got) {
                    wait(sn);
                    load_b(sn, nv);
                }
#pragma unroll
                for (int t = 0; t < 16; t++) {
                    fa[t] = na[t];
                    fb[t] = nv[t];
                }
            }
            have = hn;
        }
    } else {
        // a right sibling's slabs arrive last-to-first, but are folded first-to-last: the
        // products of all but its first slab are formed as the slabs arrive and kept (this
        // lane's 4 entries of each, in global scratch), and once the first slab is there
        // they are folded in order, each with one fma
        double* kp = kept + ((size_t)pln_kept_offset(nb, b) * ncg + (size_t)cg * pln_kept_slots(nb, b)) * 1024 +
                     (size_t)tid * 4;
        int lo = 0, hi = nb;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (b < mid) {
                const int glo = mid, ghi = hi;
                for (int s = ghi - 1; s > glo; s--) {
                    const pln_d4 acc = product(s);
                    *reinterpret_cast<pln_d4*>(kp + (size_t)(s - glo - 1) * 1024) = acc;
                }
                const pln_d4 acc0 = product(glo);
#pragma unroll
                for (int i = 0; i < 4; i++) crun[i] = fma(-1.0, acc0[i], crun[i]);
                for (int s0 = glo + 1; s0 < ghi; s0 += 8) {
                    pln_d4 kv[8];
#pragma unroll
                    for (int u = 0; u < 8; u++)
                        if (s0 + u < ghi) kv[u] = *reinterpret_cast<const pln_d4*>(kp + (size_t)(s0 + u - glo - 1) * 1024);
#pragma unroll
                    for (int u = 0; u < 8; u++)
                        if (s0 + u < ghi)
#pragma unroll
                            for (int i = 0; i < 4; i++) crun[i] = fma(-1.0, kv[u][i], crun[i]);
                }
                hi = mid;
            } else {
                lo = mid;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 4; i++) tb[lr][16 * w + lk + 4 * i] = crun[i];
    __syncthreads();
    // the leaf, in the layout of pln_block
    const int r = lane & 15, c = 4 * w + (lane >> 4);
    double x[4];
#pragma unroll
    for (int q = 0; q < 4; q++) x[q] = tb[c][16 * q + r];
    if (LOWER) {
        pln_block<true, UNIT, 0>(x, la, r);
        pln_block<true, UNIT, 1>(x, la, r);
        pln_block<true, UNIT, 2>(x, la, r);
        pln_block<true, UNIT, 3>(x, la, r);
    } else {
        pln_block<false, UNIT, 3>(x, la, r);
        pln_block<false, UNIT, 2>(x, la, r);
        pln_block<false, UNIT, 1>(x, la, r);
        pln_block<false, UNIT, 0>(x, la, r);
    }
    // a wait that gave up leaves NaN rather than a wrong solution
    if (__any(bad)) bad = 1;
    bad = __syncthreads_or(bad);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const double v = bad ? __builtin_nan("") : x[q];
        __hip_atomic_store(B + r0 + 16 * q + r + (c0 + c) * ldb, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();  // every wavefront's stores have completed
    if (tid == 0) __hip_atomic_store(flags + (size_t)b * ncg + cg, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void pln_zero(int* __restrict__ p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

constexpr size_t STOP = 64;  // trsm.rs:8

void lnl(hipStream_t st, int unit, size_t m, size_t n, const double* a, size_t lda, double* b, size_t ldb,
         Abort abort) {
    if (m <= STOP) return base<LNL>(st, unit, m, n, a, lda, b, ldb, abort);
    size_t m1 = m / 2, m2 = m - m1;
    lnl(st, unit, m1, n, a, lda, b, ldb, abort);
    launch_dgemm(st, 'N', 'N', m2, n, m1, -1.0, a + m1, lda, b, ldb, 1.0, b + m1, ldb, abort);
    lnl(st, unit, m2, n, a + m1 + m1 * lda, lda, b + m1, ldb, abort);
}
void lnu(hipStream_t st, int unit, size_t m, size_t n, const double* a, size_t lda, double* b, size_t ldb,
         Abort abort) {
    if (m <= STOP) return base<LNU>(st, unit, m, n, a, lda, b, ldb, abort);
    size_t m1 = m / 2, m2 = m - m1;
    lnu(st, unit, m2, n, a + m1 + m1 * lda, lda, b + m1, ldb, abort);
    launch_dgemm(st, 'N', 'N', m1, n, m2, -1.0, a + m1 * lda, lda, b + m1, ldb, 1.0, b, ldb, abort);
    lnu(st, unit, m1, n, a, lda, b, ldb, abort);
}
void ltl(hipStream_t st, int unit, size_t m, size_t n, const double* a, size_t lda, double* b, size_t ldb,
         Abort abort) {
    if (m <= STOP) return base<LTL>(st, unit, m, n, a, lda, b, ldb, abort);
    size_t m1 = m / 2, m2 = m - m1;
    ltl(st, unit, m2, n, a + m1 + m1 * lda, lda, b + m1, ldb, abort);
    launch_dgemm(st, 'T', 'N', m1, n, m2, -1.0, a + m1, lda, b + m1, ldb, 1.0, b, ldb, abort);
    ltl(st, unit, m1, n, a, lda, b, ldb, abort);
}
void ltu(hipStream_t st, int unit, size_t m, size_t n, const double* a, size_t lda, double* b, size_t ldb,
         Abort abort) {
    if (m <= STOP) return base<LTU>(st, unit, m, n, a, lda, b, ldb, abort);
    size_t m1 = m / 2, m2 = m - m1;
    ltu(st, unit, m1, n, a, lda, b, ldb, abort);
    launch_dgemm(st, 'T', 'N', m2, n, m1, -1.0, a + m1 * lda, lda, b, ldb, 1.0, b + m1, ldb, abort);
    ltu(st, unit, m2, n, a + m1 + m1 * lda, lda, b + m1, ldb, abort);
}
void rnl(hipStream_t st, int unit, size_t m, size_t n, const double* a, size_t lda, double* b, size_t ldb,
         Abort abort) {
    if (n <= STOP) return base<RNL>(st, unit, n, m, a, lda, b, ldb, abort);
    size_t n1 = n / 2, n2 = n - n1;
    rnl(st, unit, m, n2, a + n1 + n1 * lda, lda, b + n1 * ldb, ldb, abort);
    launch_dgemm(st, 'N', 'N', m, n1, n2, -1.0, b + n1 * ldb, ldb, a + n1, lda, 1.0, b, ldb, abort);
    rnl(st, unit, m, n1, a, lda, b, ldb, abort);
}
void rnu(hipStream_t st, int unit, size_t m, size_t n, const double* a, size_t lda, double* b, size_t ldb,
         Abort abort) {
    if (n <= STOP) return base<RNU>(st, unit, n, m, a, lda, b, ldb, abort);
    size_t n1 = n / 2, n2 = n - n1;
    rnu(st, unit, m, n1, a, lda, b, ldb, abort);
    launch_dgemm(st, 'N', 'N', m, n2, n1, -1.0, b, ldb, a + n1 * lda, lda, 1.0, b + n1 * ldb, ldb, abort);
    rnu(st, unit, m, n2, a + n1 + n1 * lda, lda, b + n1 * ldb, ldb, abort);
}
void rtl(hipStream_t st, int unit, size_t m, size_t n, const double* a, size_t lda, double* b, size_t ldb,
         Abort abort) {
    if (n <= STOP) return base<RTL>(st, unit, n, m, a, lda, b, ldb, abort);
    size_t n1 = n / 2, n2 = n - n1;
    rtl(st, unit, m, n1, a, lda, b, ldb, abort);
    launch_dgemm(st, 'N', 'T', m, n2, n1, -1.0, b, ldb, a + n1, lda, 1.0, b + n1 * ldb, ldb, abort);
    rtl(st, unit, m, n2, a + n1 + n1 * lda, lda, b + n1 * ldb, ldb, abort);
}
void rtu(hipStream_t st, int unit, size_t m, size_t n, const double* a, size_t lda, double* b, size_t ldb,
         Abort abort) {
    if (n <= STOP) return base<RTU>(st, unit, n, m, a, lda, b, ldb, abort);
    size_t n1 = n / 2, n2 = n - n1;
    rtu(st, unit, m, n2, a + n1 + n1 * lda, lda, b + n1 * ldb, ldb, abort);
    launch_dgemm(st, 'N', 'T', m, n1, n2, -1.0, b + n1 * ldb, ldb, a + n1 * lda, lda, 1.0, b, ldb, abort);
    rtu(st, unit, m, n1, a, lda, b, ldb, abort);
}

bool eqc(char c, char w) { return c == w || c == (char)(w + 32); }

}  // namespace

// scratch: the flags, then (upper solve) the kept products, 8 KB per product and group
static size_t pln_flag_bytes(size_t m, size_t n) {
    return (sizeof(int) * (m / 64) * (n / PLN_COLS) + 255) & ~(size_t)255;
}
static size_t pln_kept_bytes(size_t m, size_t n) {
    const int nb = (int)(m / 64);
    return (size_t)pln_kept_offset(nb, nb - 1) * (n / PLN_COLS) * 1024 * sizeof(double);
}

bool ln_persistent_fits(size_t m, size_t n) {
    const size_t nb = m / 64;
    if (!(n > 0 && n % PLN_COLS == 0 && m % 64 == 0 && nb >= 2 && nb <= 256 && (nb & (nb - 1)) == 0)) return false;
    return pln_flag_bytes(m, n) + pln_kept_bytes(m, n) <= ((size_t)1 << 30);
}

size_t ln_persistent_ws_bytes(size_t m, size_t n) { return pln_flag_bytes(m, n) + pln_kept_bytes(m, n); }

void launch_dtrsm_ln_persistent(hipStream_t st, bool upper, int unit, size_t m, size_t n, const double* a, size_t lda,
                                double* b, size_t ldb, void* ws) {
    const int nb = (int)(m / 64), ncg = (int)(n / PLN_COLS), nf = nb * ncg;
    int* flags = (int*)ws;
    double* kept = (double*)((char*)ws + pln_flag_bytes(m, n));
    hipLaunchKernelGGL(pln_zero, dim3((nf + 255) / 256), dim3(256), 0, st, flags, nf);
    BOLA_LAUNCH_CHECK();
    const long spin = 1L << 22;
    if (upper) {
        if (unit)
            hipLaunchKernelGGL((trsm_ln_persistent<false, true>), dim3(nf), dim3(256), 0, st, nb, ncg, a, lda, b, ldb,
                               flags, kept, spin);
        else
            hipLaunchKernelGGL((trsm_ln_persistent<false, false>), dim3(nf), dim3(256), 0, st, nb, ncg, a, lda, b, ldb,
                               flags, kept, spin);
    } else {
        if (unit)
            hipLaunchKernelGGL((trsm_ln_persistent<true, true>), dim3(nf), dim3(256), 0, st, nb, ncg, a, lda, b, ldb,
                               flags, kept, spin);
        else
            hipLaunchKernelGGL((trsm_ln_persistent<true, false>), dim3(nf), dim3(256), 0, st, nb, ncg, a, lda, b, ldb,
                               flags, kept, spin);
    }
    BOLA_LAUNCH_CHECK();
}

void launch_dtrsm(hipStream_t st, char side, char uplo, char transa, char diag, size_t m, size_t n, double alpha,
                  const double* a, size_t lda, double* b, size_t ldb, Abort abort) {
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
