// Low-precision trailing-update GEMM for the mixed-precision LU (no reference
// counterpart: the reference computes in FP64 only).
//
//   C (fp32, column-major) -= A * B
//
// with the products formed on the MFMA units in one of three precisions:
//   LP_F32  : A, B fp32, v_mfma_f32_16x16x4_f32 (FP32 matrix peak 157 TF/s)
//   LP_F16  : A, B rounded to fp16, v_mfma_f32_16x16x32_f16, fp32 accumulation
//   LP_BF16 : A, B rounded to bf16, v_mfma_f32_16x16x32_bf16, fp32 accumulation
//   LP_F16X3: A, B split into fp16 pairs, x = h + 2^-11 l with h = fp16(x) and
//             l = fp16((x - h) 2^11); C -= Ah Bh + 2^-11 (Ah Bl + Al Bh), three
//             v_mfma_f32_16x16x32_f16 per step into two fp32 accumulators.  The dropped
//             2^-22 Al Bl term and the rounding of l are both near 2^-22 relative, so the
//             products are about as accurate as fp32 ones at three sixteenths of the fp32
//             MFMA cost (fp16 matrix peak 2.5 PF/s).
//
// Tiles of 128 x 128 per 256-thread workgroup, 2 x 2 wavefronts of 64 x 64 (4 x 4 MFMA
// blocks each, 16 accumulators); K steps staged through double-buffered LDS.  LDS rows
// are k-contiguous with a stride of 36 words, which makes every fragment read free of
// bank conflicts.  The 16-bit variants take A already transposed (k-contiguous rows,
// produced by the conversion kernel below), so both operands load as 16-byte chunks.
#include "mixed.hpp"

namespace bola {

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));

constexpr int GT = 256;
constexpr int GBM = 128, GBN = 128;
constexpr int RW = 36;  // LDS row stride in 32-bit words

__device__ __forceinline__ int xcd_remap(int bid, int nblk) {
    // consecutive tiles to one XCD (8 XCDs, round-robin dispatch)
    const int per = nblk / 8;
    if (per == 0 || bid >= per * 8) return bid;
    return (bid % 8) * per + bid / 8;
}

// A lane's four accumulator entries of an MFMA 16 x 16 tile are rows ib .. ib+3 of one column
// j of C: 16-byte accesses when they lie inside the matrix and C's columns are 16-byte
// aligned (c16), element by element (rows clamped on load, masked on store) otherwise.
__device__ __forceinline__ f4 ld_c4(const float* __restrict__ C, size_t ldc, int ib, int j, int m, bool c16) {
    if (c16 && ib + 3 < m) return *(const f4*)(C + (size_t)ib + (size_t)j * ldc);
    f4 v;
#pragma unroll
    for (int r = 0; r < 4; r++) v[r] = C[(size_t)min(ib + r, m - 1) + (size_t)j * ldc];
    return v;
}
__device__ __forceinline__ void st_c4(float* __restrict__ C, size_t ldc, int ib, int j, int m, bool c16, const f4& v) {
    float* cp = C + (size_t)ib + (size_t)j * ldc;
    if (c16 && ib + 3 < m) {
        *(f4*)cp = v;
        return;
    }
#pragma unroll
    for (int r = 0; r < 4; r++)
        if (ib + r < m) cp[r] = v[r];
}

// ---- fp32 operands.  A column-major (lda), B column-major (ldb).  BK = 32 floats.
__global__ void __launch_bounds__(GT) sgemm_f32(int m, int n, int k, const float* __restrict__ A, size_t lda,
                                                const float* __restrict__ B, size_t ldb, float* __restrict__ C,
                                                size_t ldc, int tiles_m, int tiles_n, int overwrite) {
    constexpr int BK = 32;
    __shared__ float As[2][GBM * RW];
    __shared__ float Bs[2][GBN * RW];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int bid = xcd_remap(blockIdx.x, tiles_m * tiles_n);
    const int tm = bid % tiles_m, tn = bid / tiles_m;
    const int i0 = tm * GBM, j0 = tn * GBN;
    const int wm = (wv & 1) * 64, wn = (wv >> 1) * 64;
    // The accumulators start from C (clamped addresses; zero with overwrite) and A enters
    // negated, so they end as C - A B.  D layout: reg r of lane l -> (row 4*(l/16)+r, col l%16).
    const bool c16 = (ldc & 3) == 0 && ((uintptr_t)C & 15) == 0;
    f4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const int j = min(j0 + wn + 16 * b + (lane & 15), n - 1);
            const int ib = i0 + wm + 16 * a + 4 * (lane >> 4);
            acc[a][b] = overwrite ? f4{0.f, 0.f, 0.f, 0.f} : ld_c4(C, ldc, ib, j, m, c16);
        }

    // loaders: A tile 128 x 32 as 4 rows x 1 k per thread-chunk; B tile 32 x 128 as 4 k x 1 col.
    // Tiles wholly inside the operands with 16-byte aligned columns load 4 floats at once.
    const bool vec = i0 + GBM <= m && j0 + GBN <= n && (k % 32) == 0 && (lda % 4) == 0 && (ldb % 4) == 0 &&
                     (((uintptr_t)A | (uintptr_t)B) & 15) == 0;
    float ra[4][4], rb[4][4];
    auto gload = [&](int k0) {
        if (vec) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int e = tid + q * GT;
                const int ii = (e & 31) * 4, kk = e >> 5;
                const f4 va = *(const f4*)(A + (size_t)(i0 + ii) + (size_t)(k0 + kk) * lda);
                const int kb = (e & 7) * 4, jj = e >> 3;
                const f4 vb = *(const f4*)(B + (size_t)(k0 + kb) + (size_t)(j0 + jj) * ldb);
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    ra[q][r] = va[r];
                    rb[q][r] = vb[r];
                }
            }
            return;
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int e = tid + q * GT;       // 0..1023
            const int ii = (e & 31) * 4;      // row chunk within tile
            const int kk = e >> 5;            // k within tile
            const int gi = i0 + ii, gk = k0 + kk;
#pragma unroll
            for (int r = 0; r < 4; r++)
                ra[q][r] = A[(size_t)min(gi + r, m - 1) + (size_t)min(gk, k - 1) * lda];  // clamped
#pragma unroll
            for (int r = 0; r < 4; r++) ra[q][r] = keep_or_zero(ra[q][r], gi + r < m && gk < k);
            const int kb = (e & 7) * 4;       // k chunk
            const int jj = e >> 3;            // column within tile
            const int gj = j0 + jj, gkb = k0 + kb;
#pragma unroll
            for (int r = 0; r < 4; r++)
                rb[q][r] = B[(size_t)min(gkb + r, k - 1) + (size_t)min(gj, n - 1) * ldb];  // clamped
#pragma unroll
            for (int r = 0; r < 4; r++) rb[q][r] = keep_or_zero(rb[q][r], gj < n && gkb + r < k);
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int e = tid + q * GT;
            const int ii = (e & 31) * 4, kk = e >> 5;
#pragma unroll
            for (int r = 0; r < 4; r++) As[buf][(ii + r) * RW + kk] = -ra[q][r];
            const int kb = (e & 7) * 4, jj = e >> 3;
            *(f4*)&Bs[buf][jj * RW + kb] = f4{rb[q][0], rb[q][1], rb[q][2], rb[q][3]};
        }
    };
    const int nk = (k + BK - 1) / BK;
    gload(0);
    sstore(0);
    __syncthreads();
    for (int t = 0; t < nk; t++) {
        const int cur = t & 1;
        if (t + 1 < nk) gload((t + 1) * BK);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            float fa[4], fb[4];
#pragma unroll
            for (int a = 0; a < 4; a++) fa[a] = As[cur][(wm + 16 * a + (lane & 15)) * RW + kk + (lane >> 4)];
#pragma unroll
            for (int b = 0; b < 4; b++) fb[b] = Bs[cur][(wn + 16 * b + (lane & 15)) * RW + kk + (lane >> 4)];
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int b = 0; b < 4; b++) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[a], fb[b], acc[a][b], 0, 0, 0);
        }
        if (t + 1 < nk) sstore(cur ^ 1);
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const int j = j0 + wn + 16 * b + (lane & 15);
            const int ib = i0 + wm + 16 * a + 4 * (lane >> 4);
            if (j < n) st_c4(C, ldc, ib, j, m, c16, acc[a][b]);
        }
}

// ---- 16-bit operands.  At: A transposed, row i at At + i*lda (k-contiguous); B column
// j at B + j*ldb (k-contiguous).  BK = 64 elements.
template <typename T, typename V8>
__global__ void __launch_bounds__(GT) sgemm_h(int m, int n, int k, const T* __restrict__ At, size_t lda,
                                              const T* __restrict__ B, size_t ldb, float* __restrict__ C, size_t ldc,
                                              int tiles_m, int tiles_n) {
    constexpr int BK = 64;
    __shared__ uint32_t As[2][GBM * RW];
    __shared__ uint32_t Bs[2][GBN * RW];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int bid = xcd_remap(blockIdx.x, tiles_m * tiles_n);
    const int tm = bid % tiles_m, tn = bid / tiles_m;
    const int i0 = tm * GBM, j0 = tn * GBN;
    const int wm = (wv & 1) * 64, wn = (wv >> 1) * 64;
    f4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = f4{0.f, 0.f, 0.f, 0.f};
    // a tile is 128 rows x 64 elements = 128 x 8 chunks of 16 bytes; 4 chunks per thread
    uint4 ra[4], rb[4];
    auto gload = [&](int k0) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int e = tid + q * GT;
            const int row = e >> 3, ch = (e & 7) * 8;
            const int gk = k0 + ch;
            const int gi = i0 + row, gj = j0 + row;
            if (gi < m && gk + 8 <= k) {
                ra[q] = *(const uint4*)(At + (size_t)gi * lda + gk);
            } else {
                uint16_t tmp[8];
#pragma unroll
                for (int r = 0; r < 8; r++)
                    tmp[r] = (gi < m && gk + r < k) ? ((const uint16_t*)At)[(size_t)gi * lda + gk + r] : (uint16_t)0;
                ra[q] = uint4{tmp[0] | ((uint32_t)tmp[1] << 16), tmp[2] | ((uint32_t)tmp[3] << 16),
                              tmp[4] | ((uint32_t)tmp[5] << 16), tmp[6] | ((uint32_t)tmp[7] << 16)};
            }
            if (gj < n && gk + 8 <= k) {
                rb[q] = *(const uint4*)(B + (size_t)gj * ldb + gk);
            } else {
                uint16_t tmp[8];
#pragma unroll
                for (int r = 0; r < 8; r++)
                    tmp[r] = (gj < n && gk + r < k) ? ((const uint16_t*)B)[(size_t)gj * ldb + gk + r] : (uint16_t)0;
                rb[q] = uint4{tmp[0] | ((uint32_t)tmp[1] << 16), tmp[2] | ((uint32_t)tmp[3] << 16),
                              tmp[4] | ((uint32_t)tmp[5] << 16), tmp[6] | ((uint32_t)tmp[7] << 16)};
            }
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int e = tid + q * GT;
            const int row = e >> 3, w4 = (e & 7) * 4;
            *(uint4*)&As[buf][row * RW + w4] = ra[q];
            *(uint4*)&Bs[buf][row * RW + w4] = rb[q];
        }
    };
    const int nk = (k + BK - 1) / BK;
    gload(0);
    sstore(0);
    __syncthreads();
    for (int t = 0; t < nk; t++) {
        const int cur = t & 1;
        if (t + 1 < nk) gload((t + 1) * BK);
#pragma unroll
        for (int kk = 0; kk < BK / 2; kk += 16) {  // in words: 32 elements per MFMA
            V8 fa[4], fb[4];
#pragma unroll
            for (int a = 0; a < 4; a++) {
                const uint4 u = *(const uint4*)&As[cur][(wm + 16 * a + (lane & 15)) * RW + kk + 4 * (lane >> 4)];
                fa[a] = __builtin_bit_cast(V8, u);
            }
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const uint4 u = *(const uint4*)&Bs[cur][(wn + 16 * b + (lane & 15)) * RW + kk + 4 * (lane >> 4)];
                fb[b] = __builtin_bit_cast(V8, u);
            }
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    if constexpr (sizeof(T) == 2 && __is_same(T, _Float16))
                        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fa[a], fb[b], acc[a][b], 0, 0, 0);
                    else
                        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[a], fb[b], acc[a][b], 0, 0, 0);
                }
        }
        if (t + 1 < nk) sstore(cur ^ 1);
        __syncthreads();
    }
    // C -= acc, C read from clamped addresses first (every load in flight at once)
    const bool c16 = (ldc & 3) == 0 && ((uintptr_t)C & 15) == 0;
    f4 cv[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const int j = min(j0 + wn + 16 * b + (lane & 15), n - 1);
            const int ib = i0 + wm + 16 * a + 4 * (lane >> 4);
            cv[a][b] = ld_c4(C, ldc, ib, j, m, c16);
        }
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const int j = j0 + wn + 16 * b + (lane & 15);
            const int ib = i0 + wm + 16 * a + 4 * (lane >> 4);
            if (j < n) st_c4(C, ldc, ib, j, m, c16, cv[a][b] - acc[a][b]);
        }
}

// fp32 -> 16-bit, transposed: dst row i (length k, stride ldd) = src row i (src column-major).
template <typename T>
__global__ void __launch_bounds__(256) cvt_t16(int m, int k, const float* __restrict__ src, size_t lds,
                                               T* __restrict__ dst, size_t ldd) {
    __shared__ float tile[64][65];
    const int bi = blockIdx.x * 64, bk = blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int r = ty; r < 64; r += 4) {
        const int i = bi + tx, kk = bk + r;
        tile[r][tx] = (i < m && kk < k) ? src[(size_t)i + (size_t)kk * lds] : 0.f;
    }
    __syncthreads();
    for (int r = ty; r < 64; r += 4) {
        const int i = bi + r, kk = bk + tx;
        if (i < m && kk < k) dst[(size_t)i * ldd + kk] = (T)tile[tx][r];
    }
}

// fp32 -> 16-bit, same layout (column-major, columns k-contiguous)
template <typename T>
__global__ void cvt_n16(int k, int n, const float* __restrict__ src, size_t lds, T* __restrict__ dst, size_t ldd) {
    const size_t total = (size_t)k * n;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const size_t j = e / k, kk = e % k;
        dst[kk + j * ldd] = (T)src[kk + j * lds];
    }
}

// ---- split fp16 operands (LP_F16X3).  Ath/Atl: A transposed (row i k-contiguous), Bh/Bl:
// B columns k-contiguous.  256 x 128 tiles, 8 wavefronts (two per SIMD, so one hides the
// other's waits) of 64 x 64; BK = 32 elements (one MFMA step) a stage.
//   DL = false: the four operand tiles double-buffered in LDS (120 KB) through registers,
//               rows padded to 20 words;
//   DL = true:  (k % 32 == 0) operands k-chunked in memory ([k / 32][rows][32], so a stage of
//               16 rows is 1 KB of contiguous memory), three LDS stages of 48 KB filled by
//               global_load_lds_dwordx4 straight from memory, two stages ahead; a row's four
//               16-byte chunks are stored at chunk c ^ ((row / 4) % 4), so the sixteen rows an
//               MFMA operand read touches fall on distinct banks without padding.
constexpr int H3M = 256, H3N = 128, H3T = 512, H3RW = 20;  // LDS row stride in words: 16 data + 4 pad
template <bool DL>
__global__ void __launch_bounds__(H3T) sgemm_h3(int m, int n, int k, const _Float16* __restrict__ Ath,
                                                const _Float16* __restrict__ Atl, size_t lda,
                                                const _Float16* __restrict__ Bh, const _Float16* __restrict__ Bl,
                                                size_t ldb, float* __restrict__ C, size_t ldc, int tiles_m,
                                                int tiles_n) {
    constexpr int BK = 32;
    __shared__ uint32_t SA[DL ? 1 : 2][2][DL ? 4 : H3M * H3RW];
    __shared__ uint32_t SB[DL ? 1 : 2][2][DL ? 4 : H3N * H3RW];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int bid = xcd_remap(blockIdx.x, tiles_m * tiles_n);
    const int tm = bid % tiles_m, tn = bid / tiles_m;
    const int i0 = tm * H3M, j0 = tn * H3N;
    const int wm = (wv & 3) * 64, wn = (wv >> 2) * 64;
    // am = Ah Bh and ac = Ah Bl + Al Bh from zero; C enters once, at the end (accumulating
    // into C instead would round every step at C's magnitude, which in a trailing update can
    // exceed the product's by far)
    f4 am[4][4], ac[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) {
            am[a][b] = f4{0.f, 0.f, 0.f, 0.f};
            ac[a][b] = f4{0.f, 0.f, 0.f, 0.f};
        }
    // 16-byte chunks: an A tile is 256 rows x 4 chunks (2 a thread), a B tile 128 x 4 (1)
    uint4 ra[2][2], rb[2];
    auto ld8 = [&](const _Float16* src, int gr, int lim, size_t ld, int gk) -> uint4 {
        if (gr < lim && gk + 8 <= k) return *(const uint4*)(src + (size_t)gr * ld + gk);
        uint16_t tmp[8];
#pragma unroll
        for (int r = 0; r < 8; r++)
            tmp[r] = (gr < lim && gk + r < k) ? ((const uint16_t*)src)[(size_t)gr * ld + gk + r] : (uint16_t)0;
        return uint4{tmp[0] | ((uint32_t)tmp[1] << 16), tmp[2] | ((uint32_t)tmp[3] << 16),
                     tmp[4] | ((uint32_t)tmp[5] << 16), tmp[6] | ((uint32_t)tmp[7] << 16)};
    };
    auto gload = [&](int k0) {
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const int e = tid + q * H3T;
            const int row = e >> 2, gk = k0 + (e & 3) * 8;
            ra[0][q] = ld8(Ath, i0 + row, m, lda, gk);
            ra[1][q] = ld8(Atl, i0 + row, m, lda, gk);
        }
        const int row = tid >> 2, gk = k0 + (tid & 3) * 8;
        rb[0] = ld8(Bh, j0 + row, n, ldb, gk);
        rb[1] = ld8(Bl, j0 + row, n, ldb, gk);
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const int e = tid + q * H3T;
            const int o = (e >> 2) * H3RW + (e & 3) * 4;
            *(uint4*)&SA[buf][0][o] = ra[0][q];
            *(uint4*)&SA[buf][1][o] = ra[1][q];
        }
        const int o = (tid >> 2) * H3RW + (tid & 3) * 4;
        *(uint4*)&SB[buf][0][o] = rb[0];
        *(uint4*)&SB[buf][1][o] = rb[1];
    };
    const int nk = (k + BK - 1) / BK;
    // one BK step from operand tiles (ah, al: 256 rows; bh, bl: 128 rows) with row stride rw
    // words; swz: chunks stored at c ^ ((row / 4) % 4)
    auto step = [&](const uint32_t* pah, const uint32_t* pal, const uint32_t* pbh, const uint32_t* pbl, int rw,
                    bool swz) {
        h8 fbh[4], fbl[4];
        const int c = lane >> 4;
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const int r = wn + 16 * b + (lane & 15);
            const int o = r * rw + 4 * (swz ? (c ^ ((r >> 2) & 3)) : c);
            fbh[b] = __builtin_bit_cast(h8, *(const uint4*)&pbh[o]);
            fbl[b] = __builtin_bit_cast(h8, *(const uint4*)&pbl[o]);
        }
#pragma unroll
        for (int a = 0; a < 4; a++) {
            const int r = wm + 16 * a + (lane & 15);
            const int o = r * rw + 4 * (swz ? (c ^ ((r >> 2) & 3)) : c);
            const h8 fah = __builtin_bit_cast(h8, *(const uint4*)&pah[o]);
            const h8 fal = __builtin_bit_cast(h8, *(const uint4*)&pal[o]);
#pragma unroll
            for (int b = 0; b < 4; b++) {
                am[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fah, fbh[b], am[a][b], 0, 0, 0);
                ac[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fah, fbl[b], ac[a][b], 0, 0, 0);
                ac[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fal, fbh[b], ac[a][b], 0, 0, 0);
            }
        }
    };
    if constexpr (DL) {
        // three stages, each A hi | A lo | B hi | B lo; the barrier is a bare s_barrier (LDS
        // needs no cache fence), so no wait on the loads in flight is implied (offsets in words)
        constexpr int OAH = 0, OAL = H3M * 16, OBH = 2 * H3M * 16, OBL = 2 * H3M * 16 + H3N * 16;
        constexpr int STW = 2 * (H3M + H3N) * 16;
        __shared__ __attribute__((aligned(16))) uint32_t stg[3][STW];
        typedef __attribute__((address_space(3))) void* lds_ptr;
        // a direct load moves 16 rows (64 lanes x one 16-byte chunk): lane L takes row L / 4,
        // physical chunk L % 4, i.e. logical chunk (L % 4) ^ ((L / 16) % 4)
        const int lr = lane >> 2, lc = (lane & 3) ^ ((lane >> 4) & 3);
        // operands k-chunked ([k / 32][rows][32]): a direct load reads 1 KB of contiguous memory
        auto dl16 = [&](const _Float16* src, int row, int lim, int k0, uint32_t* dst) {
            __builtin_amdgcn_global_load_lds(src + ((((size_t)(k0 >> 5)) * lim + min(row, lim - 1)) << 5) + 8 * lc,
                                             (lds_ptr)dst, 16, 0, 0);
        };
        auto dissue = [&](uint32_t* sg, int k0) {
            uint32_t *ah = sg + OAH, *al = sg + OAL, *bh = sg + OBH, *bl = sg + OBL;
#pragma unroll
            for (int q = 0; q < 2; q++) {  // A: rows 32 wv + 16 q ..
                const int rb = 32 * wv + 16 * q;
                dl16(Ath, i0 + rb + lr, m, k0, ah + rb * 16);
                dl16(Atl, i0 + rb + lr, m, k0, al + rb * 16);
            }
            const int rb = 16 * wv;  // B: rows 16 wv ..
            dl16(Bh, j0 + rb + lr, n, k0, bh + rb * 16);
            dl16(Bl, j0 + rb + lr, n, k0, bl + rb * 16);
        };
        auto body = [&](int t, const uint32_t* sg, uint32_t* nsg) {
            if (t + 1 < nk) __builtin_amdgcn_s_waitcnt(0x0F76);  // vmcnt(6): the next stage may still load
            else __builtin_amdgcn_s_waitcnt(0x0F70);
            __builtin_amdgcn_s_barrier();
            if (t + 2 < nk) dissue(nsg, (t + 2) * BK);
            step(sg + OAH, sg + OAL, sg + OBH, sg + OBL, 16, true);
        };
        if (nk > 0) dissue(stg[0], 0);
        if (nk > 1) dissue(stg[1], BK);
        for (int t = 0; t < nk; t++) body(t, stg[t % 3], stg[(t + 2) % 3]);
    } else {
        gload(0);
        sstore(0);
        __syncthreads();
        for (int t = 0; t < nk; t++) {
            const int cur = t & 1;
            if (t + 1 < nk) gload((t + 1) * BK);
            step(SA[cur][0], SA[cur][1], SB[cur][0], SB[cur][1], H3RW, false);
            if (t + 1 < nk) sstore(cur ^ 1);
            __syncthreads();
        }
    }
    // C -= am + 2^-11 ac, a row block at a time, C read first
    const bool c16 = (ldc & 3) == 0 && ((uintptr_t)C & 15) == 0;
#pragma unroll
    for (int a = 0; a < 4; a++) {
        const int ib = i0 + wm + 16 * a + 4 * (lane >> 4);
        f4 cv[4];
#pragma unroll
        for (int b = 0; b < 4; b++) cv[b] = ld_c4(C, ldc, ib, min(j0 + wn + 16 * b + (lane & 15), n - 1), m, c16);
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const int j = j0 + wn + 16 * b + (lane & 15);
            if (j < n) st_c4(C, ldc, ib, j, m, c16, cv[b] - (am[a][b] + ac[a][b] * 0x1p-11f));
        }
    }
}

__device__ __forceinline__ void split16(float x, _Float16& h, _Float16& l) {
    h = (_Float16)x;
    l = (_Float16)((x - (float)h) * 2048.0f);
}

// fp32 -> split fp16 pairs, transposed (as cvt_t16)
// chunked: destination [k / 32][m][32] instead of rows of ldd
__global__ void __launch_bounds__(256) cvt_t16s(int m, int k, const float* __restrict__ src, size_t lds,
                                                _Float16* __restrict__ hi, _Float16* __restrict__ lo, size_t ldd,
                                                bool chunked) {
    __shared__ float tile[64][65];
    const int bi = blockIdx.x * 64, bk = blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int r = ty; r < 64; r += 4) {
        const int i = bi + tx, kk = bk + r;
        tile[r][tx] = (i < m && kk < k) ? src[(size_t)i + (size_t)kk * lds] : 0.f;
    }
    __syncthreads();
    for (int r = ty; r < 64; r += 4) {
        const int i = bi + r, kk = bk + tx;
        if (i < m && kk < k) {
            _Float16 h, l;
            split16(tile[tx][r], h, l);
            const size_t o = chunked ? ((((size_t)(kk >> 5)) * m + i) << 5) + (kk & 31) : (size_t)i * ldd + kk;
            hi[o] = h;
            lo[o] = l;
        }
    }
}

// fp32 -> split fp16 pairs, same layout (as cvt_n16)
__global__ void cvt_n16s(int k, int n, const float* __restrict__ src, size_t lds, _Float16* __restrict__ hi,
                         _Float16* __restrict__ lo, size_t ldd, bool chunked) {
    const size_t total = (size_t)k * n;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const size_t j = e / k, kk = e % k;
        _Float16 h, l;
        split16(src[kk + j * lds], h, l);
        const size_t o = chunked ? (((kk >> 5) * n + j) << 5) + (kk & 31) : kk + j * ldd;
        hi[o] = h;
        lo[o] = l;
    }
}

}  // namespace

void launch_sgemm_set(hipStream_t st, int m, int n, int k, const float* A, size_t lda, const float* B, size_t ldb,
                      float* C, size_t ldc) {
    if (m <= 0 || n <= 0 || k <= 0) return;
    const int tiles_m = (m + GBM - 1) / GBM, tiles_n = (n + GBN - 1) / GBN;
    hipLaunchKernelGGL(sgemm_f32, dim3((unsigned)(tiles_m * tiles_n)), dim3(GT), 0, st, m, n, k, A, lda, B, ldb, C, ldc,
                       tiles_m, tiles_n, 1);
    BOLA_LAUNCH_CHECK();
}

size_t lp_gemm_ws_bytes(int lp, size_t m, size_t n, size_t k) {
    if (lp == LP_F32) return 0;
    const size_t kp = (k + 7) & ~(size_t)7;
    return (lp == LP_F16X3 ? 4 : 2) * (m * kp + n * kp) + 256;
}

void launch_lp_gemm(hipStream_t st, int lp, int m, int n, int k, const float* A, size_t lda, const float* B,
                    size_t ldb, float* C, size_t ldc, void* ws) {
    if (m <= 0 || n <= 0 || k <= 0) return;
    const int tiles_m = (m + GBM - 1) / GBM, tiles_n = (n + GBN - 1) / GBN;
    const unsigned grid = (unsigned)(tiles_m * tiles_n);
    if (lp == LP_F32) {
        hipLaunchKernelGGL(sgemm_f32, dim3(grid), dim3(GT), 0, st, m, n, k, A, lda, B, ldb, C, ldc, tiles_m, tiles_n,
                           0);
        BOLA_LAUNCH_CHECK();
        return;
    }
    // 16-bit copies: At (m rows of kp), Bh (n columns of kp); kp keeps rows 16-byte aligned
    const size_t kp = ((size_t)k + 7) & ~(size_t)7;
    char* base = (char*)ws;
    void* At = base;
    void* Bh = base + 2 * (size_t)m * kp;
    const dim3 tg((unsigned)((m + 63) / 64), (unsigned)((k + 63) / 64));
    const unsigned cg = (unsigned)std::min<size_t>(((size_t)k * n + 255) / 256, 4096);
    if (lp == LP_F16X3) {
        // hi and lo copies of each operand: At_hi, At_lo (m rows of kp), B_hi, B_lo (n columns of kp)
        _Float16* ath = (_Float16*)base;
        _Float16* atl = ath + (size_t)m * kp;
        _Float16* bh = atl + (size_t)m * kp;
        _Float16* bl = bh + (size_t)n * kp;
        // whole 32-deep stages (the workspace is 256-byte aligned): k-chunked copies and
        // direct-to-LDS operand loads
        const bool dl = k % 32 == 0 && ((uintptr_t)ws & 15) == 0;
        hipLaunchKernelGGL(cvt_t16s, tg, dim3(256), 0, st, m, k, A, lda, ath, atl, kp, dl);
        hipLaunchKernelGGL(cvt_n16s, dim3(cg), dim3(256), 0, st, k, n, B, ldb, bh, bl, kp, dl);
        const int t3m = (m + H3M - 1) / H3M, t3n = (n + H3N - 1) / H3N;
        if (dl)
            hipLaunchKernelGGL(sgemm_h3<true>, dim3((unsigned)(t3m * t3n)), dim3(H3T), 0, st, m, n, k, ath, atl, kp,
                               bh, bl, kp, C, ldc, t3m, t3n);
        else
            hipLaunchKernelGGL(sgemm_h3<false>, dim3((unsigned)(t3m * t3n)), dim3(H3T), 0, st, m, n, k, ath, atl, kp,
                               bh, bl, kp, C, ldc, t3m, t3n);
    } else if (lp == LP_F16) {
        hipLaunchKernelGGL(cvt_t16<_Float16>, tg, dim3(256), 0, st, m, k, A, lda, (_Float16*)At, kp);
        hipLaunchKernelGGL(cvt_n16<_Float16>, dim3(cg), dim3(256), 0, st, k, n, B, ldb, (_Float16*)Bh, kp);
        hipLaunchKernelGGL((sgemm_h<_Float16, h8>), dim3(grid), dim3(GT), 0, st, m, n, k, (const _Float16*)At, kp,
                           (const _Float16*)Bh, kp, C, ldc, tiles_m, tiles_n);
    } else {
        hipLaunchKernelGGL(cvt_t16<__bf16>, tg, dim3(256), 0, st, m, k, A, lda, (__bf16*)At, kp);
        hipLaunchKernelGGL(cvt_n16<__bf16>, dim3(cg), dim3(256), 0, st, k, n, B, ldb, (__bf16*)Bh, kp);
        hipLaunchKernelGGL((sgemm_h<__bf16, b8>), dim3(grid), dim3(GT), 0, st, m, n, k, (const __bf16*)At, kp,
                           (const __bf16*)Bh, kp, C, ldc, tiles_m, tiles_n);
    }
    BOLA_LAUNCH_CHECK();
}

}  // namespace bola

using namespace bola;

extern "C" int bola_hip_sgemm_lp_dev(bola_hip_handle_t h, int precision, size_t m, size_t n, size_t k, const float* a,
                                     size_t lda, const float* b, size_t ldb, float* c, size_t ldc) {
    BOLA_API_BEGIN
    if (!h) return -1;
    if (precision < 0 || precision > LP_F16X3) return -2;
    void* ws = lp_gemm_ws_bytes(precision, m, n, k) ? h->mws.get(lp_gemm_ws_bytes(precision, m, n, k)) : nullptr;
    launch_lp_gemm(h->stream, precision, (int)m, (int)n, (int)k, a, lda, b, ldb, c, ldc, ws);
    return 0;
    BOLA_API_END(h)
}
