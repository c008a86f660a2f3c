// DGEMM on gfx950 MFMA (v_mfma_f64_16x16x4_f64), gemm.rs:88-211.
//
// C = alpha*op(A)*op(B) + beta*C, column-major.  The kernel keeps the reference's
// numerical structure so that every element sees the same sequence of roundings:
//   * beta: C*beta (one multiply) or 0 when beta == 0, applied as C is loaded
//     (gemm.rs:110-122);
//   * k is consumed in slabs of 64 (KC, gemm.rs:15): a slab's products accumulate from
//     zero in increasing k (the MFMA chain), then C = fma(alpha, acc, C)
//     (gemm.rs:245-268); the slab partials are never summed among themselves;
//   * v_mfma_f64_16x16x4_f64 rounds exactly as four chained correctly rounded FMAs in
//     increasing k (measured on gfx950), so the slab accumulator equals the reference's
//     _mm256_fmadd_pd chain bit for bit;
//   * the last m % 8 rows and the last n % 4 columns are the reference's partial 8x4
//     micro-tiles (gemm.rs:177-190), which it computes unfused as C += (alpha*b)*a in
//     increasing k; a separate kernel does exactly that for those elements.
//
// Tiling: 256 threads = 4 wavefronts in a 2x2 grid (or, for the 128 x 128 tile of
// short-k updates, 512 threads = 8 wavefronts in a 2x4 grid); each wavefront owns a
// (BM/2)x(BN/(NT/128)) block of 16x16 MFMA tiles.  K advances BK at a time through a
// double-buffered LDS stage: 16 for the large tiles, 64 for 32x32 tiles whose slab is a
// multiple of 64 (the latency-bound narrow products of the triangular solves, where each
// step is one global-load round trip).  LDS layouts put each MFMA operand read on 32 distinct
// banks: A as [k][BM] (or [m][BK+1] when op(A) = A^T), B as [n][BK+1] (or [k][BN] when
// op(B) = B^T).  Block ids are remapped so each XCD's L2 serves a compact C region.
// Plain products made of whole 128 x 64 (or, short-k, 128 x 128) tiles skip the staging
// registers: gfx950's global_load_lds_dwordx4 writes each operand stage straight into one of
// three LDS buffers, so loads run two stages ahead with no registers held for them (the
// 128 x 64 tile at 8192^3 goes from 60 % to 82 % of the FP64 matrix peak).
#include "common.hpp"

namespace bola {

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int KC = 64;  // reference slab depth

template <int BM, int BN, bool TA, bool TB, int BK, int NT = 256>
struct GemmCfg {
    // NT / 64 wavefronts: two along m, the rest along n
    static constexpr int WGN = NT / 128;
    static constexpr int WM = BM / 2, WN = BN / WGN;  // wavefront tile
    static constexpr int TM = WM / 16, TN = WN / 16;  // MFMA tiles per wavefront
    // LDS strides (doubles)
    static constexpr int A_STRIDE = TA ? (BK + 1) : BM;   // TA: [m][BK+1]; else [k][BM]
    static constexpr int A_SIZE = TA ? BM * (BK + 1) : BK * BM;
    static constexpr int B_STRIDE = TB ? BN : (BK + 1);   // TB: [k][BN]; else [n][BK+1]
    static constexpr int B_SIZE = TB ? BK * BN : BN * (BK + 1);
    static constexpr int A_LOADS = BM * BK / NT;
    static constexpr int B_LOADS = BN * BK / NT;
};

// Element (i, kk) of the BMxBK tile of op(A) handled by thread t for load slot s.
template <int BM, bool TA, int BK, int NT>
__device__ __forceinline__ void a_coord(int t, int s, int& i, int& kk) {
    int e = t + NT * s;
    if (!TA) { i = e % BM; kk = e / BM; }       // column-major A: m fastest
    else { kk = e % BK; i = e / BK; }           // A^T: k fastest
}
template <int BN, bool TB, int BK, int NT>
__device__ __forceinline__ void b_coord(int t, int s, int& kk, int& j) {
    int e = t + NT * s;
    if (!TB) { kk = e % BK; j = e / BK; }       // column-major B: k fastest
    else { j = e % BN; kk = e / BN; }           // B^T: n fastest
}

// OCC: workgroups per CU the register budget is fitted to.  2 for stand-alone products;
// 1 (more registers, lighter on the machine) for the updates that run beside a
// factorization's latency-bound panel kernel.  BK divides the caller's slab.
// DL: operands go straight from global memory into LDS (global_load_lds_dwordx4, a
// wavefront writes 1 KB of LDS per instruction) through three LDS stages, with no staging
// registers: A as [k][BM] (a 128-row column of A per instruction), B as [n][BK] with the
// 16-byte units of each column XOR-swizzled by the column index (unit u of column j at
// position u ^ (j % 8)), so the operand reads of sixteen consecutive columns fall on
// distinct banks.  Only whole tiles of op(A) = A, op(B) = B (the launcher checks).
template <int BM, int BN, bool TA, bool TB, int OCC, int BK, int NT = 256, bool DL = false>
__global__ void __launch_bounds__(NT, OCC) dgemm_mfma(int m, int n, int k, double alpha, const double* __restrict__ A,
                                                  size_t lda, const double* __restrict__ B, size_t ldb, double beta,
                                                  double* __restrict__ C, size_t ldc, int tiles_m, int tiles_n,
                                                  int edge_m, int edge_n, Abort abort, int slab, int mask) {
    using Cfg = GemmCfg<BM, BN, TA, TB, BK, NT>;
    if (k > 0) {  // behind a failed pivot: only the k columns of panels before it
        k = abort_count(abort, k, 1);
        if (k == 0) return;
    }
    static_assert(!DL || (!TA && !TB && BM == 128 && BK == 16 && ((BN == 64 && NT == 256) || (BN == 128 && NT == 512))),
                  "direct-load shape");
    __shared__ __attribute__((aligned(16))) double sA[DL ? 1 : 2][Cfg::A_SIZE];
    __shared__ __attribute__((aligned(16))) double sB[DL ? 1 : 2][Cfg::B_SIZE];

    // XCD-aware block remap: consecutive tile indices land on the same XCD.
    const int ntiles = tiles_m * tiles_n;
    int bid = blockIdx.x;
    if ((ntiles & 7) == 0) bid = (bid & 7) * (ntiles >> 3) + (bid >> 3);
    // grouped ordering: 8 tile-rows at a time, column-major inside a group
    const int GROUP = 8;
    int group = bid / (GROUP * tiles_n);
    int first_m = group * GROUP;
    int gsz = min(tiles_m - first_m, GROUP);
    int tm = first_m + (bid % (GROUP * tiles_n)) % gsz;
    int tn = (bid % (GROUP * tiles_n)) / gsz;
    const int row0 = tm * BM, col0 = tn * BN;
    // triangle mask (SYRK off-diagonal blocks): 1 = rows in a later 32-block than the
    // column, 2 = rows in an earlier 32-block; skip tiles with no such element
    if (mask == 1 && (row0 + BM - 1) / 32 <= col0 / 32) return;
    if (mask == 2 && row0 / 32 >= (col0 + BN - 1) / 32) return;

    const int t = threadIdx.x;
    const int lane = t & 63, w = t >> 6;
    const int wm = (w & 1) * Cfg::WM, wn = (w >> 1) * Cfg::WN;
    const int lr = lane & 15, lk = lane >> 4;

    // running C (reference: C after beta, updated once per 64-slab) and slab accumulator
    d4 crun[Cfg::TM][Cfg::TN];
    d4 acc[Cfg::TM][Cfg::TN];
    // C from clamped addresses, every load in flight at once (C is not read when beta == 0)
#pragma unroll
    for (int a = 0; a < Cfg::TM; a++)
#pragma unroll
        for (int b = 0; b < Cfg::TN; b++) {
            const int gc = col0 + wn + 16 * b + lr;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int gr = row0 + wm + 16 * a + lk + 4 * r;
                double v = 0.0;
                if (beta != 0.0) v = keep_or_zero(C[min(gr, m - 1) + (size_t)min(gc, n - 1) * ldc], gr < m && gc < n);
                crun[a][b][r] = v;
                acc[a][b][r] = 0.0;
            }
        }
    if (beta != 0.0 && beta != 1.0)
#pragma unroll
        for (int a = 0; a < Cfg::TM; a++)
#pragma unroll
            for (int b = 0; b < Cfg::TN; b++)
#pragma unroll
                for (int r = 0; r < 4; r++) crun[a][b][r] = crun[a][b][r] * beta;

    // Two register stages: the global loads for step t+2 are issued while step t computes,
    // so each load has two steps of MFMA work to arrive in.  The loads are raw (clamped
    // addresses); out-of-range elements are zeroed only when the stage is stored to LDS,
    // after the compute step, so that no instruction waits on a load before then.
    double ra0[Cfg::A_LOADS], rb0[Cfg::B_LOADS], ra1[Cfg::A_LOADS], rb1[Cfg::B_LOADS];
    auto gload = [&](int k0, double (&ra)[Cfg::A_LOADS], double (&rb)[Cfg::B_LOADS]) {
#pragma unroll
        for (int s = 0; s < Cfg::A_LOADS; s++) {
            int i, kk;
            a_coord<BM, TA, BK, NT>(t, s, i, kk);
            const int ci = min(row0 + i, m - 1), ck = min(k0 + kk, k - 1);  // clamped: loads in flight together
            ra[s] = TA ? A[ck + (size_t)ci * lda] : A[ci + (size_t)ck * lda];
        }
#pragma unroll
        for (int s = 0; s < Cfg::B_LOADS; s++) {
            int kk, j;
            b_coord<BN, TB, BK, NT>(t, s, kk, j);
            const int cj = min(col0 + j, n - 1), ck = min(k0 + kk, k - 1);
            rb[s] = TB ? B[cj + (size_t)ck * ldb] : B[ck + (size_t)cj * ldb];
        }
    };
    auto lstore = [&](int buf, int k0, const double (&ra)[Cfg::A_LOADS], const double (&rb)[Cfg::B_LOADS]) {
        // the whole stage inside the matrices (the common case): no masks
        const bool full = row0 + BM <= m && col0 + BN <= n && k0 + BK <= k;
#pragma unroll
        for (int s = 0; s < Cfg::A_LOADS; s++) {
            int i, kk;
            a_coord<BM, TA, BK, NT>(t, s, i, kk);
            const double v = full ? ra[s] : keep_or_zero(ra[s], row0 + i < m && k0 + kk < k);
            if (!TA) sA[buf][kk * BM + i] = v;
            else sA[buf][i * (BK + 1) + kk] = v;
        }
#pragma unroll
        for (int s = 0; s < Cfg::B_LOADS; s++) {
            int kk, j;
            b_coord<BN, TB, BK, NT>(t, s, kk, j);
            const double v = full ? rb[s] : keep_or_zero(rb[s], col0 + j < n && k0 + kk < k);
            if (!TB) sB[buf][j * (BK + 1) + kk] = v;
            else sB[buf][kk * BN + j] = v;
        }
    };
    const int nsteps = (k + BK - 1) / BK;
    // one BK-deep step from LDS buffer buf, then the slab fold when a slab (or k) ends
    auto compute = [&](const double* pa, const double* pb, int step) {
#pragma unroll
        for (int k4 = 0; k4 < BK; k4 += 4) {
            double fa[Cfg::TM], fb[Cfg::TN];
#pragma unroll
            for (int a = 0; a < Cfg::TM; a++) {
                int i = wm + 16 * a + lr, kk = k4 + lk;
                fa[a] = TA ? pa[i * (BK + 1) + kk] : pa[kk * BM + i];
            }
#pragma unroll
            for (int b = 0; b < Cfg::TN; b++) {
                int j = wn + 16 * b + lr, kk = k4 + lk;
                if (DL) fb[b] = pb[j * BK + 2 * ((kk >> 1) ^ (j & 7)) + (kk & 1)];
                else fb[b] = TB ? pb[kk * BN + j] : pb[j * (BK + 1) + kk];
            }
#pragma unroll
            for (int a = 0; a < Cfg::TM; a++)
#pragma unroll
                for (int b = 0; b < Cfg::TN; b++)
                    acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[a], fb[b], acc[a][b], 0, 0, 0);
        }
        const int kdone = (step + 1) * BK;
        if ((kdone % slab) == 0 || step + 1 == nsteps) {
#pragma unroll
            for (int a = 0; a < Cfg::TM; a++)
#pragma unroll
                for (int b = 0; b < Cfg::TN; b++)
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        crun[a][b][r] = fma(alpha, acc[a][b][r], crun[a][b][r]);
                        acc[a][b][r] = 0.0;
                    }
        }
    };

    if constexpr (DL) {
        // three stages in separate LDS objects, so that the compiler's bookkeeping of the
        // direct loads in flight lets a stage be read while the next ones are still loading
        __shared__ __attribute__((aligned(16))) double dA0[BK * BM], dA1[BK * BM], dA2[BK * BM];
        __shared__ __attribute__((aligned(16))) double dB0[BN * BK], dB1[BN * BK], dB2[BN * BK];
        typedef __attribute__((address_space(3))) void* lds_ptr;
        // a wavefront's share of a stage, in 1 KB direct loads: AQ columns of A (BM = 128
        // doubles each), BQ groups of 8 columns of B (16 doubles each)
        constexpr int NW = NT / 64, AQ = BK / NW, BQ = BN / 8 / NW;
        static_assert(AQ + BQ <= 15, "vmcnt field");
        // stage <- k0 .. k0+BK: A columns AQ w .. AQ w + AQ-1; B columns 8(BQ w + q) + lane/8,
        // unit lane % 8 of each holding rows 2u, 2u+1 with u = (lane % 8) ^ (column % 8)
        auto dissue = [&](double* la, double* lb, int k0) {
#pragma unroll
            for (int q = 0; q < AQ; q++) {
                const int kk = AQ * w + q;
                __builtin_amdgcn_global_load_lds(A + (size_t)row0 + 2 * lane + (size_t)(k0 + kk) * lda,
                                                 (lds_ptr)(la + kk * BM), 16, 0, 0);
            }
#pragma unroll
            for (int q = 0; q < BQ; q++) {
                const int jb = 8 * (BQ * w + q), j = jb + (lane >> 3), u = (lane & 7) ^ (j & 7);
                __builtin_amdgcn_global_load_lds(B + (size_t)(k0 + 2 * u) + (size_t)(col0 + j) * ldb,
                                                 (lds_ptr)(lb + jb * BK), 16, 0, 0);
            }
        };
        // one step: its stage (ca, cb) has landed when at most the next stage's AQ + BQ loads
        // per lane are pending; after the barrier every wavefront's part is in and the stage
        // read in the previous step (ia, ib) can take step + 2
        auto body = [&](int st, const double* ca, const double* cb, double* ia, double* ib) {
            if (st + 1 < nsteps) __builtin_amdgcn_s_waitcnt(0x0F70 | (AQ + BQ));  // vmcnt(AQ + BQ)
            else __builtin_amdgcn_s_waitcnt(0x0F70);                               // vmcnt(0)
            __builtin_amdgcn_s_barrier();
            if (st + 2 < nsteps) dissue(ia, ib, (st + 2) * BK);
            compute(ca, cb, st);
        };
        if (nsteps > 0) dissue(dA0, dB0, 0);
        if (nsteps > 1) dissue(dA1, dB1, BK);
        for (int step = 0; step < nsteps; step += 3) {
            body(step, dA0, dB0, dA2, dB2);
            if (step + 1 >= nsteps) break;
            body(step + 1, dA1, dB1, dA0, dB0);
            if (step + 2 >= nsteps) break;
            body(step + 2, dA2, dB2, dA1, dB1);
        }
    } else {
    if (nsteps > 0) {  // k == 0: C = beta C, and A, B are not read
        gload(0, ra0, rb0);
        lstore(0, 0, ra0, rb0);
        if (nsteps > 1) gload(BK, ra1, rb1);
    }
    __syncthreads();
    for (int step = 0; step < nsteps; step += 2) {
        // even step: LDS 0 holds step, registers 1 hold step + 1
        if (step + 2 < nsteps) gload((step + 2) * BK, ra0, rb0);
        compute(sA[0], sB[0], step);
        if (step + 1 >= nsteps) break;
        lstore(1, (step + 1) * BK, ra1, rb1);
        __syncthreads();
        // odd step: LDS 1 holds step + 1, registers 0 hold step + 2
        if (step + 3 < nsteps) gload((step + 3) * BK, ra1, rb1);
        compute(sA[1], sB[1], step + 1);
        if (step + 2 < nsteps) {
            lstore(0, (step + 2) * BK, ra0, rb0);
            __syncthreads();
        }
    }
    }

#pragma unroll
    for (int a = 0; a < Cfg::TM; a++)
#pragma unroll
        for (int b = 0; b < Cfg::TN; b++) {
            int gc = col0 + wn + 16 * b + lr;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                int gr = row0 + wm + 16 * a + lk + 4 * r;
                bool keep = gr < m && gc < n && gr < edge_m && gc < edge_n;
                if (mask == 1) keep = keep && (gr / 32 > gc / 32);
                if (mask == 2) keep = keep && (gr / 32 < gc / 32);
                if (keep) C[gr + (size_t)gc * ldc] = crun[a][b][r];
            }
        }
}

// The reference's partial micro-tiles: rows >= edge_m or columns >= edge_n.
// C = C*beta (or 0), then C += (alpha*b_lj)*a_il unfused for l = 0..k-1.
// The chain is sequential in l; its operands are independent of it, so they are loaded
// EK steps at a time into one of two register buffers while the other buffer's steps run
// (an element's loads for steps l+EK .. l+2EK-1 are in flight during steps l .. l+EK-1).
constexpr int EK = 32;
template <bool TA, bool TB>
__global__ void __launch_bounds__(64) dgemm_edge(int m, int n, int k, double alpha, const double* __restrict__ A,
                                                  size_t lda, const double* __restrict__ B, size_t ldb, double beta,
                                                  double* __restrict__ C, size_t ldc, int edge_m, int edge_n,
                                                  Abort abort, int mask) {
    if (k > 0) {
        k = abort_count(abort, k, 1);
        if (k == 0) return;
    }
    // enumerate edge elements: first the edge rows for all columns, then the edge columns
    // for the interior rows
    const int er = m - edge_m, ec = n - edge_n;
    const long n1 = (long)er * n;
    const long total = n1 + (long)edge_m * ec;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
        int i, j;
        if (e < n1) { i = edge_m + (int)(e % er); j = (int)(e / er); }
        else { long f = e - n1; i = (int)(f % edge_m); j = edge_n + (int)(f / edge_m); }
        if (mask == 1 && !(i / 32 > j / 32)) continue;
        if (mask == 2 && !(i / 32 < j / 32)) continue;
        double c = 0.0;
        if (beta != 0.0) {
            c = C[i + (size_t)j * ldc];
            if (beta != 1.0) c = c * beta;
        }
        if (alpha != 0.0 && k > 0) {
            double a0[EK], b0[EK], a1[EK], b1[EK];
            auto load = [&](int l0, double (&av)[EK], double (&bv)[EK]) {
#pragma unroll
                for (int s = 0; s < EK; s++) {
                    const int l = min(l0 + s, k - 1);  // clamped: every load in flight at once
                    bv[s] = TB ? B[j + (size_t)l * ldb] : B[l + (size_t)j * ldb];
                    av[s] = TA ? A[l + (size_t)i * lda] : A[i + (size_t)l * lda];
                }
            };
            auto steps = [&](int l0, const double (&av)[EK], const double (&bv)[EK]) {
#pragma unroll
                for (int s = 0; s < EK; s++) {
                    if (l0 + s < k) {
                        const double temp = alpha * bv[s];
                        c = c + temp * av[s];
                    }
                }
            };
            load(0, a0, b0);
            for (int l0 = 0; l0 < k; l0 += 2 * EK) {
                if (l0 + EK < k) load(l0 + EK, a1, b1);
                steps(l0, a0, b0);
                if (l0 + EK >= k) break;
                if (l0 + 2 * EK < k) load(l0 + 2 * EK, a0, b0);
                steps(l0 + EK, a1, b1);
            }
        }
        C[i + (size_t)j * ldc] = c;
    }
}

template <int BM, int BN, bool TA, bool TB, int BK = 16, int NT = 256, bool DL = false>
void launch_tile(hipStream_t st, int m, int n, int k, double alpha, const double* A, size_t lda, const double* B,
                 size_t ldb, double beta, double* C, size_t ldc, int edge_m, int edge_n, Abort abort,
                 int slab, int mask, bool fast) {
    int tm = (m + BM - 1) / BM, tn = (n + BN - 1) / BN;
    if constexpr (NT == 512)  // one workgroup of 8 wavefronts per CU (two per SIMD)
        hipLaunchKernelGGL((dgemm_mfma<BM, BN, TA, TB, 1, BK, NT, DL>), dim3(tm * tn), dim3(NT), 0, st, m, n, k, alpha,
                           A, lda, B, ldb, beta, C, ldc, tm, tn, edge_m, edge_n, abort, slab, mask);
    else if (fast)
        hipLaunchKernelGGL((dgemm_mfma<BM, BN, TA, TB, (BM == 128 ? 2 : BK == 64 ? 2 : BK == 32 ? 4 : 3), BK, 256, DL>),
                           dim3(tm * tn), dim3(256), 0, st, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, tm, tn,
                           edge_m, edge_n, abort, slab, mask);
    else
        hipLaunchKernelGGL((dgemm_mfma<BM, BN, TA, TB, 1, BK>), dim3(tm * tn), dim3(256), 0, st, m, n, k, alpha, A,
                           lda, B, ldb, beta, C, ldc, tm, tn, edge_m, edge_n, abort, slab, mask);
    BOLA_LAUNCH_CHECK();
}

// the reference's partial micro-tiles of a product (rows >= edge_m, columns >= edge_n)
template <bool TA, bool TB>
void launch_edge(hipStream_t st, int m, int n, int k, double alpha, const double* A, size_t lda, const double* B,
                 size_t ldb, double beta, double* C, size_t ldc, Abort abort, int mask, int edge_m, int edge_n) {
    long nedge = (long)(m - edge_m) * n + (long)edge_m * (n - edge_n);
    if (nedge > 0) {
        // one wavefront per workgroup, so that a few thousand elements spread over the CUs
        unsigned blocks = (unsigned)std::min<long>((nedge + 63) / 64, 16384);
        hipLaunchKernelGGL((dgemm_edge<TA, TB>), dim3(blocks), dim3(64), 0, st, m, n, k, alpha, A, lda, B, ldb, beta,
                           C, ldc, edge_m, edge_n, abort, mask);
        BOLA_LAUNCH_CHECK();
    }
}

template <bool TA, bool TB>
void launch_t(hipStream_t st, int m, int n, int k, double alpha, const double* A, size_t lda, const double* B,
              size_t ldb, double beta, double* C, size_t ldc, int num_cus, Abort abort, int slab,
              int mask, int edge_m, int edge_n, bool fast) {
    if (edge_m > 0 && edge_n > 0) {
        const long t128 = (long)((m + 127) / 128) * ((n + 63) / 64);
        const long t64 = (long)((m + 63) / 64) * ((n + 63) / 64);
        // 64-deep steps need fold points on multiples of 64 and at least one full step
        const bool deep = slab % 64 == 0 && k >= 64;
        // short-k updates (the LU's and the Cholesky's trailing updates, k = 256) with at
        // least four 128 x 128 tiles per CU: 8-wavefront tiles, a third less operand traffic
        // per flop (long-k products run faster on the 128 x 64 tile at two workgroups per CU)
        const long t128x128 = (long)((m + 127) / 128) * ((n + 127) / 128);
        // whole tiles of A B (m % 128, n % BN, k % 16 zero) on 16-byte aligned operands take
        // the direct-to-LDS loads
        const bool dl_ok = m % 128 == 0 && k % 16 == 0 && slab % 16 == 0 && lda % 2 == 0 && ldb % 2 == 0 &&
                           ((uintptr_t)A & 15) == 0 && ((uintptr_t)B & 15) == 0;
        if (k <= 1024 && t128x128 >= 4L * num_cus) {
            bool done = false;
            if constexpr (!TA && !TB) {
                if (dl_ok && n % 128 == 0) {
                    launch_tile<128, 128, TA, TB, 16, 512, true>(st, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc,
                                                                 edge_m, edge_n, abort, slab, mask, fast);
                    done = true;
                }
            }
            if (!done)
                launch_tile<128, 128, TA, TB, 16, 512>(st, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, edge_m,
                                                       edge_n, abort, slab, mask, fast);
        } else if (t128 >= 2L * num_cus) {
            const bool dl = fast && dl_ok && n % 64 == 0;
            bool done = false;
            if constexpr (!TA && !TB) {
                if (dl) {
                    launch_tile<128, 64, TA, TB, 16, 256, true>(st, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc,
                                                                edge_m, edge_n, abort, slab, mask, fast);
                    done = true;
                }
            }
            if (!done)
                launch_tile<128, 64, TA, TB>(st, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, edge_m, edge_n, abort,
                                             slab, mask, fast);
        }
        else if (deep && fast && t64 < 2L * num_cus)  // stand-alone: below 2 64x64 tiles per CU
            launch_tile<32, 32, TA, TB, 32>(st, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, edge_m, edge_n, abort,
                                            slab, mask, fast);
        else if (deep && t64 < num_cus)
            launch_tile<32, 32, TA, TB, 64>(st, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, edge_m, edge_n, abort,
                                            slab, mask, fast);
        else if (t64 >= num_cus || m * (long)n <= 64L * 64)
            launch_tile<64, 64, TA, TB>(st, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, edge_m, edge_n, abort, slab,
                                        mask, fast);
        else
            launch_tile<32, 32, TA, TB>(st, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, edge_m, edge_n, abort, slab,
                                        mask, fast);
    }
    launch_edge<TA, TB>(st, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, abort, mask, edge_m, edge_n);
}

int g_num_cus = 256;

}  // namespace

void set_num_cus(int n) { g_num_cus = n > 0 ? n : 256; }

// Requires m, n, k < 2^31.  alpha == 0 must be handled by the caller (beta-only pass).
void launch_dgemm(hipStream_t st, char transa, char transb, size_t m, size_t n, size_t k, double alpha,
                  const double* a, size_t lda, const double* b, size_t ldb, double beta, double* c, size_t ldc,
                  Abort abort, bool fast) {
    launch_dgemm_ex(st, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, abort, KC, 0, -1, -1, fast);
}

void launch_dgemm_ex(hipStream_t st, char transa, char transb, size_t m, size_t n, size_t k, double alpha,
                     const double* a, size_t lda, const double* b, size_t ldb, double beta, double* c, size_t ldc,
                     Abort abort, int slab, int mask, long edge_m, long edge_n, bool fast, int ncu) {
    const int cus = ncu > 0 ? ncu : g_num_cus;
    bool ta = !(transa == 'N' || transa == 'n');
    bool tb = !(transb == 'N' || transb == 'n');
    int M = (int)m, N = (int)n, K = (int)k;
    // elements in rows >= em or columns >= en take the reference's unfused edge path;
    // by default the last m % 8 rows and n % 4 columns of this call (gemm.rs:177-190)
    const int em = edge_m < 0 ? M - (M % 8) : (int)std::min<long>(std::max<long>(edge_m, 0), M);
    const int en = edge_n < 0 ? N - (N % 4) : (int)std::min<long>(std::max<long>(edge_n, 0), N);
    if (!ta && !tb) launch_t<false, false>(st, M, N, K, alpha, a, lda, b, ldb, beta, c, ldc, cus, abort, slab, mask, em, en, fast);
    else if (!ta && tb) launch_t<false, true>(st, M, N, K, alpha, a, lda, b, ldb, beta, c, ldc, cus, abort, slab, mask, em, en, fast);
    else if (ta && !tb) launch_t<true, false>(st, M, N, K, alpha, a, lda, b, ldb, beta, c, ldc, cus, abort, slab, mask, em, en, fast);
    else launch_t<true, true>(st, M, N, K, alpha, a, lda, b, ldb, beta, c, ldc, cus, abort, slab, mask, em, en, fast);
}

}  // namespace bola
