// Low-precision LU with partial (tournament) pivoting for the mixed-precision solver.
// No reference counterpart: the reference factors in FP64 only.
//
// Storage.  The working matrix is T = (S A C)^T in fp32, column-major with leading
// dimension ldt: T[c + r*ldt] = s_r * A(r, c) * t_c, with S = diag(s), C = diag(t) powers of
// two (equilibration, exact).  Row r of A is then a contiguous column of T, so the row
// interchanges of partial pivoting are coalesced column swaps, a panel's rows are
// contiguous, and every trailing update is a column-major GEMM on blocks of T:
//   A22 -= A21 A12   <=>   T(c1.., r1..) -= T(c1.., r0..) * T(c0.., r1..).
//
// Factorization: recursive LU (column halves, leaves of 32 columns).  A leaf panel is
// factored with tournament pivoting (CALU): every workgroup runs partial pivoting on its
// block of rows and proposes its pivot rows; one workgroup runs partial pivoting on the
// proposals (original values) and picks the panel's pivots; the interchanges are applied
// to whole rows, and the panel is factored without further pivoting (LU of the 32 x 32
// top block, L21 = A21 U11^-1).  There is no per-column synchronisation between
// workgroups.  TRSM (right, upper, unit on T) and GEMM in the chosen product precision.
#include "mixed.hpp"

#include <algorithm>
#include <cmath>

namespace bola {

namespace {

constexpr int LPW = 32;        // leaf panel width
constexpr int kMixNB = 256;    // block width of the right-looking factorization

// 64-bit max over the wavefront: DPP within rows of 16 lanes, lane reads across rows;
// the result is uniform.
template <int CTRL>
__device__ __forceinline__ uint64_t dpp64(uint64_t v) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)v, CTRL, 0xF, 0xF, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(v >> 32), CTRL, 0xF, 0xF, false);
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t rl64(uint64_t v, int l) {
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, l), hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), l);
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t umax64(uint64_t a, uint64_t b) { return a > b ? a : b; }
__device__ __forceinline__ uint64_t wave_max64(uint64_t k) {
    k = umax64(k, dpp64<0xB1>(k));   // quad_perm [1,0,3,2]
    k = umax64(k, dpp64<0x4E>(k));   // quad_perm [2,3,0,1]
    k = umax64(k, dpp64<0x124>(k));  // row_ror 4
    k = umax64(k, dpp64<0x128>(k));  // row_ror 8: every lane holds its row's maximum
    return umax64(umax64(rl64(k, 0), rl64(k, 16)), umax64(rl64(k, 32), rl64(k, 48)));
}
// pivot key: larger |x| first, then the smaller set index; NaN above nothing, dead rows 0
__device__ __forceinline__ uint64_t piv_key(float x, int idx, bool live) {
    const float v = fabsf(x);
    const uint32_t kv = !live ? 0u : (v == v ? __builtin_bit_cast(uint32_t, v) + 2u : 1u);
    return ((uint64_t)kv << 32) | (uint32_t)(0xFFFFFFFFu - (uint32_t)idx);
}

// LAPACK-style interchanges for picks (absolute rows, in order) at rows r0..: lane s
// follows the current position of pick s through the sequence of swaps.
__device__ void picks_to_ipiv(const int* picks, int k, int r0, int32_t* __restrict__ ipiv) {
    const int lane = threadIdx.x & 63;
    int cur = lane < k ? picks[lane] : -1;
    for (int t = 0; t < k; t++) {
        const int pos = __builtin_amdgcn_readlane(cur, t);
        const int dst = r0 + t;
        if (lane == 0) ipiv[dst] = pos;
        cur = cur == dst ? pos : (cur == pos ? dst : cur);
    }
}

// ---- the workgroup tournament: 4 wavefronts, RPT rows per thread in registers, two
// barriers per step (argmax across wavefronts, then the pivot row).  Fully unrolled over
// the 32 steps, so columns already eliminated leave the registers.
template <int J, int RPT>
__device__ __forceinline__ void gstep(float (&own)[RPT][LPW], bool (&live)[RPT], int steps, int (&pick)[LPW],
                                      uint64_t (*red)[4], float (*prow)[LPW]) {
    if constexpr (J < LPW) {
        const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
        if (J < steps) {
            uint64_t k = 0;
#pragma unroll
            for (int q = 0; q < RPT; q++) k = umax64(k, piv_key(own[q][J], tid + q * 256, live[q]));
            k = wave_max64(k);
            if (lane == 0) red[J & 1][wv] = k;
            __syncthreads();
            k = umax64(umax64(red[J & 1][0], red[J & 1][1]), umax64(red[J & 1][2], red[J & 1][3]));
            const int p = (int)(0xFFFFFFFFu - (uint32_t)k);
            pick[J] = p;
            const int qp = p >> 8, tp = p & 255;
            if (tid == tp) {
                // the owner publishes the pivot row, its slot chosen with bit masks (a branch
                // per slot would put the rows in memory)
#pragma unroll
                for (int c = J; c < LPW; c++) {
                    uint32_t v = 0;
#pragma unroll
                    for (int q = 0; q < RPT; q++) v |= __builtin_bit_cast(uint32_t, own[q][c]) & (q == qp ? ~0u : 0u);
                    prow[J & 1][c] = __builtin_bit_cast(float, v);
                }
            }
#pragma unroll
            for (int q = 0; q < RPT; q++) live[q] = live[q] && !(q == qp && tid == tp);
            __syncthreads();
            float pv[LPW];
#pragma unroll
            for (int c = J; c < LPW; c++) pv[c] = prow[J & 1][c];
            const float piv = pv[J];
            if (piv != 0.f) {
                const float inv = 1.f / piv;
#pragma unroll
                for (int q = 0; q < RPT; q++) {
                    const float lm = live[q] ? -(own[q][J] * inv) : 0.f;
#pragma unroll
                    for (int c = J + 1; c < LPW; c++) own[q][c] = fmaf(lm, pv[c], own[q][c]);
                }
            }
        }
        gstep<J + 1, RPT>(own, live, steps, pick, red, prow);
    }
}

// Partial pivoting of the workgroup on rows set(i), i < rows (absolute A-rows; set == null:
// rows base + i); the picks (absolute rows) go to out[].  Returns the number of picks.
template <int RPT>
__device__ __forceinline__ int wg_tournament(const int* set, int base, int rows, int w, int c0,
                                             const float* __restrict__ T, size_t ldt, int* out, uint64_t (*red)[4], float (*prow)[LPW], int* pk_sh) {
    const int tid = threadIdx.x;
    float own[RPT][LPW];
    bool live[RPT];
#pragma unroll
    for (int q = 0; q < RPT; q++) {
        const int i = tid + q * 256;
        live[q] = i < rows;
        const int rid = live[q] ? (set ? set[i] : base + i) : c0;  // c0 = r0: a valid row for idle threads
        const float* src = T + (size_t)c0 + (size_t)rid * ldt;
#pragma unroll
        for (int c = 0; c < LPW; c++) own[q][c] = src[min(c, w - 1)];  // clamped: loads in flight together
#pragma unroll
        for (int c = 0; c < LPW; c++) own[q][c] = keep_or_zero(own[q][c], live[q] && c < w);
    }
    const int steps = min(w, rows);
    int pk[LPW];
    gstep<0, RPT>(own, live, steps, pk, red, prow);
    if (tid == 0)
#pragma unroll
        for (int j = 0; j < LPW; j++) pk_sh[j] = pk[j];
    __syncthreads();
    if (tid < steps) out[tid] = set ? set[pk_sh[tid]] : base + pk_sh[tid];
    __syncthreads();
    return steps;
}

constexpr int GROWS = 1024;  // rows per workgroup in the tournament's first stage (4 per thread)
constexpr int GFAN = 16;     // tree fan-in: 16 proposals of 32 rows, 2 per thread

// lists: [level][node][LPW] absolute rows and counts; cnt: [level][node] arrival counters
// (zero on entry; the root clears them for the next panel).  G workgroups of 256 threads:
// the tournament's first stage on GROWS rows each, then a tree with fan-in GFAN where the last workgroup of a
// group to arrive merges the group's proposals.  Nobody waits for anybody.  The root
// writes the panel's interchanges ipiv[r0 .. r0+k).
__global__ void __launch_bounds__(256) calu_wg(int n, int r0, int c0, int w, const float* __restrict__ T, size_t ldt,
                                               int* __restrict__ lists, int* __restrict__ lcount, int* __restrict__ cnt,
                                               int lvl_stride, int32_t* __restrict__ ipiv) {
    __shared__ uint64_t red[2][4];
    __shared__ float prow[2][LPW];
    __shared__ int pk_sh[LPW];
    __shared__ int mylist[LPW];
    __shared__ int merged[GFAN * LPW];
    __shared__ int go;
    const int tid = threadIdx.x;
    const int G = gridDim.x;
    const int rb = r0 + blockIdx.x * GROWS;
    const int rows = max(0, min(GROWS, n - rb));
    int k = wg_tournament<GROWS / 256>(nullptr, rb, rows, w, c0, T, ldt, mylist, red, prow, pk_sh);
    if (G == 1) {
        if (tid < 64) picks_to_ipiv(mylist, k, r0, ipiv);
        return;
    }
    int node = blockIdx.x, nodes = G;
    for (int lvl = 0;; lvl++) {
        int* L = lists + ((size_t)lvl * lvl_stride + node) * LPW;
        if (tid < k) L[tid] = mylist[tid];
        if (tid == 0) lcount[lvl * lvl_stride + node] = k;
        const int parent = node / GFAN;
        const int first = parent * GFAN, nchild = min(GFAN, nodes - first);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (tid == 0) {
            const int old = __hip_atomic_fetch_add(cnt + lvl * lvl_stride + parent, 1, __ATOMIC_ACQ_REL,
                                                   __HIP_MEMORY_SCOPE_AGENT);
            go = old == nchild - 1;
        }
        __syncthreads();
        if (!go) return;
        // the last to arrive merges the group's proposals
        int tot = 0;
        for (int ch = 0; ch < nchild; ch++) {
            const int cc = __hip_atomic_load(lcount + lvl * lvl_stride + first + ch, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
            const int* CL = lists + ((size_t)lvl * lvl_stride + first + ch) * LPW;
            if (tid < cc) merged[tot + tid] = __hip_atomic_load(CL + tid, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            tot += cc;
        }
        __syncthreads();
        const int pnodes = (nodes + GFAN - 1) / GFAN;
        const bool root = pnodes == 1;
        k = wg_tournament<GFAN * LPW / 256>(merged, 0, tot, w, c0, T, ldt, mylist, red, prow, pk_sh);
        node = parent;
        nodes = pnodes;
        if (root) {
            if (tid < 64) picks_to_ipiv(mylist, k, r0, ipiv);
            for (int l2 = 0; l2 <= lvl; l2++)
                for (int e = tid; e < lvl_stride; e += 256) cnt[l2 * lvl_stride + e] = 0;
            return;
        }
    }
}


__global__ void k_iota_piv(int n, int32_t* p) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

__global__ void k_zero_cnt(int* p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

// interchanges ipiv[k1 .. k2) applied to the columns [cl, ch) of A (rows of T) and, with
// perm != null, to the row-origin vector; one thread per column of A.
__global__ void laswp_t(int cl, int ch, float* __restrict__ T, size_t ldt, int k1, int k2,
                        const int32_t* __restrict__ ipiv, int* __restrict__ perm) {
    const int c = cl + blockIdx.x * blockDim.x + threadIdx.x;
    if (c > ch || (c == ch && !perm)) return;
    for (int i = k1; i < k2; i++) {
        const int p = ipiv[i];
        if (p == i) continue;
        if (c < ch) {
            const float t = T[(size_t)c + (size_t)i * ldt];
            T[(size_t)c + (size_t)i * ldt] = T[(size_t)c + (size_t)p * ldt];
            T[(size_t)c + (size_t)p * ldt] = t;
        } else {
            const int t = perm[i];
            perm[i] = perm[p];
            perm[p] = t;
        }
    }
}

// Ui = U^-1 for the w x w upper triangle U (unit: ones on the diagonal, not read),
// both in LDS; one wavefront, lane b computes column b by back substitution.  A zero
// diagonal entry is taken as one (the factorization is then useless and the refinement
// will say so).
__device__ void upper_inverse_lds(const float (*U)[LPW + 1], float (*Ui)[LPW + 1], int w, bool unit) {
    const int b = threadIdx.x;
    if (b >= 64) return;
    float x[LPW];
#pragma unroll
    for (int a = LPW - 1; a >= 0; a--) {
        float v = (a == b) ? 1.f : 0.f;
#pragma unroll
        for (int c = a + 1; c < LPW; c++) v = fmaf(-U[a][c], x[c], v);
        asm volatile("" ::: "memory");
        const float d = unit ? 1.f : U[a][a];
        x[a] = (a < w && b < w && a <= b) ? (d != 0.f ? v / d : v) : 0.f;
    }
    if (b < LPW)
#pragma unroll
        for (int a = 0; a < LPW; a++) Ui[a][b] = x[a];
}

// the panel without pivoting, after the interchanges: L21 = A21 U11^-1 for the rows
// below the top block (one thread per row, a product with the inverse).  Every workgroup
// factors the top block itself (LU without pivoting: its rows are the chosen pivots, in
// order); with top_only, one workgroup stores it instead, in a launch after the others
// have read the unfactored block.
template <bool FULL>  // FULL: w == 32, loads and stores without tests
__global__ void __launch_bounds__(256) panel_nopiv(int n, int r0, int c0, int w, float* __restrict__ T, size_t ldt,
                                                   int top_only) {
    __shared__ float U[LPW][LPW + 1];  // U[row][col] of the top block, A orientation
    __shared__ float Ui[LPW][LPW + 1];
    const int tid = threadIdx.x;
    if (tid < 64) {
        // LU without pivoting of the top block (its rows are the chosen pivots, in order):
        // lane c holds column c; multipliers come from lane j by lane reads
        const int c = tid;
        float col[LPW];
#pragma unroll
        for (int r = 0; r < LPW; r++)
            col[r] = T[(size_t)(c0 + min(c, w - 1)) + (size_t)(r0 + min(r, w - 1)) * ldt];  // clamped
#pragma unroll
        for (int r = 0; r < LPW; r++) col[r] = (c < w && r < w) ? col[r] : (r == c ? 1.f : 0.f);
#pragma unroll
        for (int j = 0; j < LPW; j++) {
            const float piv = __builtin_amdgcn_readlane(__builtin_bit_cast(int, col[j]), j) == 0
                                  ? 0.f
                                  : __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, col[j]), j));
            const float inv = piv != 0.f ? 1.f / piv : 0.f;
#pragma unroll
            for (int r = j + 1; r < LPW; r++) {
                const float l = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, col[r]), j)) * inv;
                col[r] = c > j ? fmaf(-l, col[j], col[r]) : (c == j ? l : col[r]);
            }
        }
        if (c < LPW)
#pragma unroll
            for (int r = 0; r < LPW; r++) U[r][c] = col[r];
    }
    __syncthreads();
    if (top_only) {
        for (int e = tid; e < w * w; e += 256) {
            const int r = e / w, c = e % w;
            T[(size_t)(c0 + c) + (size_t)(r0 + r) * ldt] = U[r][c];
        }
        return;
    }
    upper_inverse_lds(U, Ui, w, false);
    __syncthreads();
    const int r = r0 + w + blockIdx.x * 256 + tid;
    if (r >= n) return;
    float x[LPW], l[LPW];
    float* col = T + (size_t)c0 + (size_t)r * ldt;
#pragma unroll
    for (int b = 0; b < LPW; b++) {
        x[b] = col[FULL ? b : min(b, w - 1)];  // clamped: loads in flight together
        x[b] = FULL ? x[b] : keep_or_zero(x[b], b < w);
        l[b] = 0.f;
    }
#pragma unroll
    for (int a = 0; a < LPW; a++) {
#pragma unroll
        for (int b = a; b < LPW; b++) l[b] = fmaf(x[a], Ui[a][b], l[b]);
        asm volatile("" ::: "memory");  // keeps the Ui reads row by row (registers)
    }
#pragma unroll
    for (int b = 0; b < LPW; b++)
        if (FULL || b < w) col[b] = l[b];
}

// X U = B for unit upper U (w <= 32): X, B the nr x w column-major block at X (ld), U the
// w x w block at U (ld) (its strictly upper part).  X = B U^-1 with the inverse formed in
// LDS; one thread per row.
template <bool FULL>  // FULL: w == 32, loads and stores without tests
__global__ void __launch_bounds__(256) strsm_rupu_base(int nr, int w, const float* __restrict__ Ub, float* __restrict__ X,
                                                       size_t ld) {
    __shared__ float U[LPW][LPW + 1];
    __shared__ float Ui[LPW][LPW + 1];
    const int tid = threadIdx.x;
    for (int e = tid; e < LPW * LPW; e += 256) {
        const int a = e % LPW, b = e / LPW;
        U[a][b] = (a < w && b < w) ? Ub[(size_t)a + (size_t)b * ld] : 0.f;
    }
    __syncthreads();
    upper_inverse_lds(U, Ui, w, true);
    __syncthreads();
    const int row = blockIdx.x * 256 + tid;
    if (row >= nr) return;
    float* xr = X + row;  // element b at xr[b * ld]: a per-lane base and a uniform offset
    float x[LPW], y[LPW];
#pragma unroll
    for (int b = 0; b < LPW; b++) {
        x[b] = (FULL || b < w) ? xr[(size_t)b * ld] : 0.f;
        y[b] = 0.f;
    }
#pragma unroll
    for (int a = 0; a < LPW; a++) {
#pragma unroll
        for (int b = a; b < LPW; b++) y[b] = fmaf(x[a], Ui[a][b], y[b]);
        asm volatile("" ::: "memory");  // keeps the Ui reads row by row (registers)
    }
#pragma unroll
    for (int b = 0; b < LPW; b++)
        if (FULL || b < w) xr[(size_t)b * ld] = y[b];
}

// ---- the butterfly mode: LU without interchanges of a panel -----------------------------
//
// A panel A[r0:n, r0:r0+w) (w <= kMixNB) is factored in two launches:
//   rbt_diag   one workgroup factors the diagonal block A11 = L11 U11 in registers and
//              publishes its tiles in the forms the second launch reads;
//   rbt_below  L21 = A21 U11^-1 (blocked, a wavefront per 16 rows) and, in four more
//              workgroups, -L11^-1 for the U12 products of the trailing updates.
// No workgroup waits for another.
//
// 16 x 16 tiles, "row layout" in registers: lane (g, r) = (lane / 16, lane % 16) holds row
// r, columns 4g .. 4g+3 of a tile.  v_mfma_f32_16x16x4_f32 forms D = A B + C with lane l
// supplying A[l%16][l/16] and B[l/16][l%16] and holding D[4(l/16)+v][l%16], so a tile in
// row layout is the accumulator layout of its transpose.  mm16 forms C + X Y in row
// layout as (C + X Y)^T = C^T + Y^T X^T in four MFMAs; in step kk, k-slot g stands for the
// inner index 4g + kk, so the operands are
//   X: lane (g, r), step kk: X[r][4g + kk]   -- X's row layout, value kk
//   Y: lane (g, r), step kk: Y[4g + kk][r]   -- "column layout" of Y, here called colv
constexpr int DT = 16;            // tile order
constexpr int DNT = kMixNB / DT;  // tiles along a panel side
constexpr int DLS = DT + 4;       // LDS row stride (floats) of a tile: 16-byte aligned rows
constexpr size_t kPackTile = DT * DT;  // floats of a packed tile (4 per lane, lane-major)
typedef float f4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f4v mm16(const f4v& y, const f4v& x, f4v c) {
#pragma unroll
    for (int kk = 0; kk < 4; kk++) c = __builtin_amdgcn_mfma_f32_16x16x4f32(y[kk], x[kk], c, 0, 0, 0);
    return c;
}
__device__ __forceinline__ f4v ld4(const float* p) { return *reinterpret_cast<const f4v*>(p); }
__device__ __forceinline__ void st4(float* p, const f4v& v) { *reinterpret_cast<f4v*>(p) = v; }
__device__ __forceinline__ float rlf(float v, int l) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}

// Packed tiles, 256 floats each, lane-major (lane l: floats 4l .. 4l+3), DNT x DNT slots:
//   upk[t][c], c > t: colv(U_tc)      upk[t][t]: colv(U_tt^-1)
//   lpk[t][k], k < t: row(L_tk)       lpk[t][t]: row(L_tt^-1)
__device__ __forceinline__ float* pk_slot(float* base, int i, int j) {
    return base + ((size_t)i * DNT + j) * kPackTile;
}

// The diagonal block, one workgroup of 8 wavefronts.  Wavefront (a, b) = (wv / 2, wv % 2)
// holds tiles (a + 4p, b + 2q), p < 4, q < 8, in row layout (128 registers a lane); tiles
// past the block are the identity.  Step t (a tile column):
//   A1  the owner of tile (t, t) factors it (LU without interchanges, lanes 0..15 a row
//       each, the pivot row by readlane);
//   A2  wavefront 0 forms U_tt^-1, wavefront 1 L_tt^-1 (a column per lane);
//   B   L_it = A_it U_tt^-1 (i > t) and U_tc = L_tt^-1 A_tc (c > t) by their owners, into
//       LDS for everybody and packed for rbt_below;
//   C   A_ic -= L_it U_tc for i, c > t.
// The LU goes back to A11 at the end.
constexpr int DWA = 4, DWB = 2;                     // tile-row and tile-column residues
constexpr int DNP = DNT / DWA, DNQ = DNT / DWB;     // tiles a wavefront holds per row/column
__global__ void __launch_bounds__(64 * DWA * DWB) rbt_diag(int r0, int w, float* __restrict__ T, size_t ldt,
                                                           float* __restrict__ upk, float* __restrict__ lpk) {
    __shared__ float dg[DT][DT + 1];
    __shared__ __attribute__((aligned(16))) float uit[DT][DLS];  // uit[r][k] = U_tt^-1[k][r]
    __shared__ __attribute__((aligned(16))) float li[DT][DLS];   // li[r][k] = L_tt^-1[r][k]
    __shared__ __attribute__((aligned(16))) float ls[DNT][DT][DLS];  // ls[i][r][k] = L_it[r][k]
    __shared__ __attribute__((aligned(16))) float us[DNT][DT][DLS];  // us[c][r][k] = U_tc[k][r]
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63, g = lane >> 4, r = lane & 15;
    const int a = wv / DWB, b = wv % DWB;
    const int nt = (w + DT - 1) / DT;
    float* const P = T + (size_t)r0 + (size_t)r0 * ldt;  // A11(0, 0)
    const f4v zero = {0.f, 0.f, 0.f, 0.f};
    f4v x[DNP][DNQ];
#pragma unroll
    for (int p = 0; p < DNP; p++)
#pragma unroll
        for (int q = 0; q < DNQ; q++) {
            const int row = DT * (a + DWA * p) + r, col = DT * (b + DWB * q) + 4 * g;
            if (row < w && col + 3 < w) {
                x[p][q] = ld4(P + col + (size_t)row * ldt);
            } else {
#pragma unroll
                for (int v = 0; v < 4; v++)
                    x[p][q][v] = (row < w && col + v < w) ? P[col + v + (size_t)row * ldt] : (row == col + v ? 1.f : 0.f);
            }
        }
    for (int t = 0; t < nt; t++) {
        const int tp = t / DWA, tq = t / DWB, ta = t % DWA, tb = t % DWB;
        // A1
        if (a == ta && b == tb) {
#pragma unroll
            for (int p = 0; p < DNP; p++)
#pragma unroll
                for (int q = 0; q < DNQ; q++)
                    if (p == tp && q == tq)
#pragma unroll
                        for (int v = 0; v < 4; v++) dg[r][4 * g + v] = x[p][q][v];
            if (lane < DT) {
                float rv[DT];
#pragma unroll
                for (int c = 0; c < DT; c++) rv[c] = dg[lane][c];
#pragma unroll
                for (int k = 0; k < DT; k++) {
                    // rows at or above k take a zero multiplier: their entries stay as they are
                    const float piv = rlf(rv[k], k);
                    const float ip = piv != 0.f ? __builtin_amdgcn_rcpf(piv) : 0.f;
                    const float l = lane > k ? rv[k] * ip : 0.f;
                    if (lane > k) rv[k] = l;
#pragma unroll
                    for (int c = k + 1; c < DT; c++) rv[c] = fmaf(-l, rlf(rv[c], k), rv[c]);
                    __builtin_amdgcn_sched_barrier(0);
                }
#pragma unroll
                for (int c = 0; c < DT; c++) dg[lane][c] = rv[c];
            }
#pragma unroll
            for (int p = 0; p < DNP; p++)
#pragma unroll
                for (int q = 0; q < DNQ; q++)
                    if (p == tp && q == tq)
#pragma unroll
                        for (int v = 0; v < 4; v++) x[p][q][v] = dg[r][4 * g + v];
        }
        __syncthreads();
        // A2: column j = lane of U_tt^-1 (back substitution) and of L_tt^-1 (forward); the
        // factored rows in registers (lane k: row k), their entries broadcast by readlane
        if (wv < 2 && lane < DT) {
            const int j = lane;
            float rv[DT], xv[DT];
#pragma unroll
            for (int c = 0; c < DT; c++) rv[c] = dg[lane][c];
            if (wv == 0) {
#pragma unroll
                for (int k = DT - 1; k >= 0; k--) {
                    float s0 = k == j ? 1.f : 0.f, s1 = 0.f;
#pragma unroll
                    for (int c = k + 1; c < DT; c += 2) {
                        s0 = fmaf(-rlf(rv[c], k), xv[c], s0);
                        if (c + 1 < DT) s1 = fmaf(-rlf(rv[c + 1], k), xv[c + 1], s1);
                    }
                    const float d = rlf(rv[k], k), s = s0 + s1;
                    xv[k] = d != 0.f ? s * __builtin_amdgcn_rcpf(d) : s;
                    __builtin_amdgcn_sched_barrier(0);  // the broadcasts of a step stay in it
                }
#pragma unroll
                for (int k = 0; k < DT; k++) uit[j][k] = xv[k];
            } else {
#pragma unroll
                for (int k = 0; k < DT; k++) {
                    float s0 = k == j ? 1.f : 0.f, s1 = 0.f;
#pragma unroll
                    for (int c = 0; c < k; c += 2) {
                        s0 = fmaf(-rlf(rv[c], k), xv[c], s0);
                        if (c + 1 < k) s1 = fmaf(-rlf(rv[c + 1], k), xv[c + 1], s1);
                    }
                    xv[k] = s0 + s1;
                    __builtin_amdgcn_sched_barrier(0);
                }
#pragma unroll
                for (int k = 0; k < DT; k++) li[k][j] = xv[k];
            }
        }
        __syncthreads();
        // B
        if (b == tb) {
            const f4v ui = ld4(&uit[r][4 * g]);
#pragma unroll
            for (int p = 0; p < DNP; p++) {
                const int i = a + DWA * p;
                if (i > t && i < nt)
#pragma unroll
                    for (int q = 0; q < DNQ; q++)
                        if (q == tq) {
                            const f4v l = mm16(ui, x[p][q], zero);
                            x[p][q] = l;
                            st4(&ls[i][r][4 * g], l);
                            st4(pk_slot(lpk, i, t) + 4 * lane, l);
                        }
            }
        }
        if (a == ta) {
            const f4v lr = ld4(&li[r][4 * g]);
#pragma unroll
            for (int q = 0; q < DNQ; q++) {
                const int c = b + DWB * q;
                if (c > t && c < nt)
#pragma unroll
                    for (int p = 0; p < DNP; p++)
                        if (p == tp) {
                            // colv(A_tc) through LDS, then U_tc = L_tt^-1 A_tc
#pragma unroll
                            for (int v = 0; v < 4; v++) us[c][4 * g + v][r] = x[p][q][v];
                            const f4v ya = ld4(&us[c][r][4 * g]);
                            const f4v u = mm16(ya, lr, zero);
                            x[p][q] = u;
#pragma unroll
                            for (int v = 0; v < 4; v++) us[c][4 * g + v][r] = u[v];
                        }
            }
        }
        if (wv == DWA * DWB - 1) {
            st4(pk_slot(upk, t, t) + 4 * lane, ld4(&uit[r][4 * g]));
            st4(pk_slot(lpk, t, t) + 4 * lane, ld4(&li[r][4 * g]));
        }
        __syncthreads();
        // C
        if (a == ta)
#pragma unroll
            for (int q = 0; q < DNQ; q++) {
                const int c = b + DWB * q;
                if (c > t && c < nt) st4(pk_slot(upk, t, c) + 4 * lane, ld4(&us[c][r][4 * g]));
            }
#pragma unroll
        for (int p = 0; p < DNP; p++) {
            const int i = a + DWA * p;
            if (i > t && i < nt) {
                const f4v xl = ld4(&ls[i][r][4 * g]);
#pragma unroll
                for (int q = 0; q < DNQ; q++) {
                    const int c = b + DWB * q;
                    if (c > t && c < nt) x[p][q] = mm16(-ld4(&us[c][r][4 * g]), xl, x[p][q]);
                }
            }
        }
    }
#pragma unroll
    for (int p = 0; p < DNP; p++)
#pragma unroll
        for (int q = 0; q < DNQ; q++) {
            const int row = DT * (a + DWA * p) + r, col = DT * (b + DWB * q) + 4 * g;
            if (row < w && col + 3 < w) {
                st4(P + col + (size_t)row * ldt, x[p][q]);
            } else if (row < w) {
#pragma unroll
                for (int v = 0; v < 4; v++)
                    if (col + v < w) P[col + v + (size_t)row * ldt] = x[p][q][v];
            }
        }
}

// After rbt_diag.  Workgroups 0 .. nbw-1: the rows below A11, a wavefront per 16 rows,
// L21 = A21 U11^-1 by tile columns (L_t = A_t U_tt^-1, then A_c -= L_t U_tc for c > t).
// Workgroups nbw .. nbw+3: wavefront s writes block column s of -(L11^-1)^T into nlt
// (nlt[c + r w] = -L11^-1(r, c)): with Z = L11^-T (upper triangular),
//   Z_ss = (L_ss^-1)^T,  Z_st = -(sum_{s <= k < t} Z_sk L_tk^T) (L_tt^-1)^T  (t > s);
// row(L_tk) is colv(L_tk^T), so the packed L tiles serve as the right factors.
__global__ void __launch_bounds__(256) rbt_below(int n, int r0, int w, float* __restrict__ T, size_t ldt,
                                                 const float* __restrict__ upk, const float* __restrict__ lpk,
                                                 float* __restrict__ nlt, int nbw) {
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
    const int nt = (w + DT - 1) / DT;
    const f4v zero = {0.f, 0.f, 0.f, 0.f};
    if ((int)blockIdx.x >= nbw) {
        const int s = ((int)blockIdx.x - nbw) * 4 + wv;
        if (s >= nt) return;
        f4v z[DNT];  // z[d] = Z_{s, s+d}
        f4v id;
#pragma unroll
        for (int v = 0; v < 4; v++) id[v] = r == 4 * g + v ? 1.f : 0.f;
        z[0] = mm16(ld4(pk_slot(const_cast<float*>(lpk), s, s) + 4 * lane), id, zero);
#pragma unroll
        for (int d = 1; d < DNT; d++) {
            if (s + d < nt) {
                const int t = s + d;
                f4v acc = zero;
#pragma unroll
                for (int e = 0; e < d; e++) acc = mm16(ld4(pk_slot(const_cast<float*>(lpk), t, s + e) + 4 * lane), z[e], acc);
                z[d] = -mm16(ld4(pk_slot(const_cast<float*>(lpk), t, t) + 4 * lane), acc, zero);
            }
        }
        // nlt = -Z: tile (s, t) of Z holds Z(16s + r, 16t + 4g + v) at nlt[(16s + r) + (16t + 4g + v) w]
        const int cz = DT * s + r;
        if (cz >= w) return;
        for (int t = 0; t < nt; t++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int rz = DT * t + 4 * g + v;
                if (rz >= w) continue;
                float val = 0.f;
#pragma unroll
                for (int d = 0; d < DNT; d++)
                    if (t == s + d) val = -z[d][v];
                nlt[(size_t)cz + (size_t)rz * w] = val;
            }
        return;
    }
    const long row0 = (long)r0 + w + DT * ((long)blockIdx.x * 4 + wv);
    if (row0 >= n) return;
    const long row = row0 + r;
    const bool live = row < n;
    float* const src = T + (size_t)r0 + (size_t)(live ? row : row0) * ldt;
    f4v x[DNT];
#pragma unroll
    for (int c = 0; c < DNT; c++) {
        const int col = DT * c + 4 * g;
        if (live && col + 3 < w) {
            x[c] = ld4(src + col);
        } else {
#pragma unroll
            for (int v = 0; v < 4; v++) x[c][v] = (live && col + v < w) ? src[col + v] : 0.f;
        }
    }
    // the U tiles of step t (cur) were loaded during step t - 1
    f4v cur[DNT];
#pragma unroll
    for (int c = 0; c < DNT; c++) cur[c] = c < nt ? ld4(pk_slot(const_cast<float*>(upk), 0, c) + 4 * lane) : zero;
#pragma unroll
    for (int t = 0; t < DNT; t++) {
        if (t < nt) {
            f4v nxt[DNT];
#pragma unroll
            for (int c = t + 1; c < DNT; c++)
                nxt[c] = (t + 1 < nt && c < nt) ? ld4(pk_slot(const_cast<float*>(upk), t + 1, c) + 4 * lane) : zero;
            x[t] = mm16(cur[t], x[t], zero);
#pragma unroll
            for (int c = t + 1; c < DNT; c++)
                if (c < nt) x[c] = mm16(-cur[c], x[t], x[c]);
#pragma unroll
            for (int c = t + 1; c < DNT; c++) cur[c] = nxt[c];
        }
    }
    if (!live) return;
#pragma unroll
    for (int c = 0; c < DNT; c++) {
        const int col = DT * c + 4 * g;
        if (col + 3 < w) {
            st4(src + col, x[c]);
        } else {
#pragma unroll
            for (int v = 0; v < 4; v++)
                if (col + v < w) src[col + v] = x[c][v];
        }
    }
}

// dst (rows x cols, column-major, ld ldd) = src (ld lds)
__global__ void copy2d(int rows, int cols, const float* __restrict__ src, size_t lds, float* __restrict__ dst,
                       size_t ldd) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows) return;
    for (int j = blockIdx.y; j < cols; j += gridDim.y) dst[(size_t)i + (size_t)j * ldd] = src[(size_t)i + (size_t)j * lds];
}

}  // namespace


namespace {

void launch_laswp_t(hipStream_t st, int cl, int ch, float* T, size_t ldt, int k1, int k2, const int32_t* ipiv,
                    int* perm) {
    if (ch <= cl && !perm) return;
    hipLaunchKernelGGL(laswp_t, dim3(cdiv((size_t)(ch - cl) + 1, 256)), dim3(256), 0, st, cl, ch, T, ldt, k1, k2, ipiv,
                       perm);
    BOLA_LAUNCH_CHECK();
}

struct LuCtx {
    hipStream_t st;
    bool pivot;            // tournament pivoting; false: no interchanges (butterfly-randomised input)
    int swap_lo, swap_hi;  // columns of A that take a leaf's interchanges at once
    int n, lp;
    float* T;
    size_t ldt;
    int32_t* ipiv;
    int* perm;
    int* lists;
    int* lcount;
    int* cnt;
    int lvl_stride;
    void* gws;
    float* upk;  // the butterfly mode: the current panel's packed U11 tiles (rbt_diag)
    float* lpk;  // the butterfly mode: the current panel's packed L11 tiles (rbt_diag)
    float* nlt;  // the butterfly mode: -(L11^-1)^T of the current panel (kMixNB x kMixNB)
    float* ubuf; // the butterfly mode: U12^T blocks, row c for A-column c (ld n)
};

void calu_panel(const LuCtx& c, int r0, int w) {
    const int rows = c.n - r0;
    if (c.pivot) {
        const int G = (rows + GROWS - 1) / GROWS;
        hipLaunchKernelGGL(calu_wg, dim3(G), dim3(256), 0, c.st, c.n, r0, r0, w, c.T, c.ldt, c.lists, c.lcount, c.cnt,
                           c.lvl_stride, c.ipiv);
        BOLA_LAUNCH_CHECK();
        const int k = std::min(w, rows);
        launch_laswp_t(c.st, c.swap_lo, c.swap_hi, c.T, c.ldt, r0, r0 + k, c.ipiv, c.perm);
    }
    const int below = std::max(0, rows - w);
    if (below > 0) {
        if (w == LPW)
            hipLaunchKernelGGL(panel_nopiv<true>, dim3(cdiv((size_t)below, 256)), dim3(256), 0, c.st, c.n, r0, r0, w,
                               c.T, c.ldt, 0);
        else
            hipLaunchKernelGGL(panel_nopiv<false>, dim3(cdiv((size_t)below, 256)), dim3(256), 0, c.st, c.n, r0, r0, w,
                               c.T, c.ldt, 0);
        BOLA_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(panel_nopiv<false>, dim3(1), dim3(256), 0, c.st, c.n, r0, r0, w, c.T, c.ldt, 1);
    BOLA_LAUNCH_CHECK();
}

float* Tp(const LuCtx& c, int col, int row) { return c.T + (size_t)col + (size_t)row * c.ldt; }

// The butterfly mode's panel [k, k+nb): the diagonal block, then the rows below and
// -L11^-1 for the U12 products.
void launch_panel_rbt(const LuCtx& c, int k, int nb) {
    hipLaunchKernelGGL(rbt_diag, dim3(1), dim3(64 * DWA * DWB), 0, c.st, k, nb, c.T, c.ldt, c.upk, c.lpk);
    BOLA_LAUNCH_CHECK();
    const int below = c.n - k - nb;
    const int nbw = (int)cdiv((size_t)std::max(below, 0), 4 * DT);
    hipLaunchKernelGGL(rbt_below, dim3((unsigned)(nbw + 4)), dim3(256), 0, c.st, c.n, k, nb, c.T, c.ldt, c.upk, c.lpk,
                       c.nlt, nbw);
    BOLA_LAUNCH_CHECK();
}

// X U = B on T blocks: X/B at (xc, r0), nr x w; U at (r0c, r0), w x w unit upper
void strsm_rupu(const LuCtx& c, int nr, int w, int uc, int ur, int xc) {
    if (nr <= 0 || w <= 0) return;
    if (w <= LPW) {
        if (w == LPW)
            hipLaunchKernelGGL(strsm_rupu_base<true>, dim3(cdiv((size_t)nr, 256)), dim3(256), 0, c.st, nr, w,
                               Tp(c, uc, ur), Tp(c, xc, ur), c.ldt);
        else
            hipLaunchKernelGGL(strsm_rupu_base<false>, dim3(cdiv((size_t)nr, 256)), dim3(256), 0, c.st, nr, w,
                               Tp(c, uc, ur), Tp(c, xc, ur), c.ldt);
        BOLA_LAUNCH_CHECK();
        return;
    }
    const int h1 = ((w / 2) + LPW - 1) / LPW * LPW, h2 = w - h1;
    strsm_rupu(c, nr, h1, uc, ur, xc);
    // B2 -= X1 * U12 : (nr x h2) -= (nr x h1)(h1 x h2), in fp32 (small)
    launch_lp_gemm(c.st, LP_F32, nr, h2, h1, Tp(c, xc, ur), c.ldt, Tp(c, uc, ur + h1), c.ldt, Tp(c, xc, ur + h1),
                   c.ldt, c.gws);
    strsm_rupu(c, nr, h2, uc + h1, ur + h1, xc);
}

void lu_rec(const LuCtx& c, int r0, int w) {
    if (w <= LPW) {
        calu_panel(c, r0, w);
        return;
    }
    const int w1 = ((w / 2) + LPW - 1) / LPW * LPW, nr = w - w1;
    lu_rec(c, r0, w1);
    // U12 = L11^-1 A12  <=>  X (nr x w1 at T(r0+w1, r0)) * L11^T = B
    strsm_rupu(c, nr, w1, r0, r0, r0 + w1);
    // A22 -= A21 A12
    const int below = c.n - r0 - w1;
    if (below > 0)  // inside a panel: narrow, kept in fp32
        launch_lp_gemm(c.st, LP_F32, nr, below, w1, Tp(c, r0 + w1, r0), c.ldt, Tp(c, r0, r0 + w1), c.ldt,
                       Tp(c, r0 + w1, r0 + w1), c.ldt, c.gws);
    lu_rec(c, r0 + w1, nr);
}

// U12 and the trailing update of panel [k, k+nb) on the columns [c_lo, c_hi) of A
void update_cols(const LuCtx& c, int k, int nb, int c_lo, int c_hi) {
    const int ncols = c_hi - c_lo;
    if (ncols <= 0) return;
    const int below = c.n - k - nb;
    if (c.pivot) {
        strsm_rupu(c, ncols, nb, k, k, c_lo);
        if (below > 0)
            launch_lp_gemm(c.st, c.lp, ncols, below, nb, Tp(c, c_lo, k), c.ldt, Tp(c, k, k + nb), c.ldt,
                           Tp(c, c_lo, k + nb), c.ldt, c.gws);
        return;
    }
    // U12^T = A12^T L11^-T as one product into ubuf (copied into place afterwards), and
    // the trailing update reads it from there
    float* ub = c.ubuf + c_lo;
    launch_sgemm_set(c.st, ncols, nb, nb, Tp(c, c_lo, k), c.ldt, c.nlt, (size_t)nb, ub, (size_t)c.n);
    if (below > 0)
        launch_lp_gemm(c.st, c.lp, ncols, below, nb, ub, (size_t)c.n, Tp(c, k, k + nb), c.ldt, Tp(c, c_lo, k + nb),
                       c.ldt, c.gws);
    hipLaunchKernelGGL(copy2d, dim3(cdiv((size_t)ncols, 256), (unsigned)std::min(nb, 256)), dim3(256), 0, c.st, ncols,
                       nb, ub, (size_t)c.n, Tp(c, c_lo, k), c.ldt);
    BOLA_LAUNCH_CHECK();
}

// The butterfly mode's two consecutive panels [k, k+nb) and [k+nb, k+2nb) (nb = kMixNB) on
// the columns [c_lo, c_hi): the first panel's U rows, its update of the second panel's
// rows, the second panel's U rows, then the rows below both in one 2nb-deep product --
// half the passes over the trailing matrix of two single-panel updates.
void update_pair(const LuCtx& c, int k, const float* nlt_a, const float* nlt_b, int c_lo, int c_hi) {
    const int ncols = c_hi - c_lo, nb = kMixNB;
    if (ncols <= 0) return;
    float* ub = c.ubuf + c_lo;  // columns 0 .. 2nb-1 of U12^T, ld n
    launch_sgemm_set(c.st, ncols, nb, nb, Tp(c, c_lo, k), c.ldt, nlt_a, (size_t)nb, ub, (size_t)c.n);
    launch_lp_gemm(c.st, c.lp, ncols, nb, nb, ub, (size_t)c.n, Tp(c, k, k + nb), c.ldt, Tp(c, c_lo, k + nb), c.ldt,
                   c.gws);
    launch_sgemm_set(c.st, ncols, nb, nb, Tp(c, c_lo, k + nb), c.ldt, nlt_b, (size_t)nb, ub + (size_t)nb * c.n,
                     (size_t)c.n);
    const int below = c.n - k - 2 * nb;
    if (below > 0)
        launch_lp_gemm(c.st, c.lp, ncols, below, 2 * nb, ub, (size_t)c.n, Tp(c, k, k + 2 * nb), c.ldt,
                       Tp(c, c_lo, k + 2 * nb), c.ldt, c.gws);
    hipLaunchKernelGGL(copy2d, dim3(cdiv((size_t)ncols, 256), (unsigned)std::min(2 * nb, 256)), dim3(256), 0, c.st,
                       ncols, 2 * nb, ub, (size_t)c.n, Tp(c, c_lo, k), c.ldt);
    BOLA_LAUNCH_CHECK();
}

}  // namespace

size_t mixed_lu_ws_bytes(int n, int lp) {
    const size_t G = (size_t)(n + GROWS - 1) / GROWS;
    size_t b = 8 * G * (LPW + 2) * sizeof(int) + LPW * LPW * sizeof(float) + 4096;
    b += 2 * (lp_gemm_ws_bytes(lp, (size_t)n, (size_t)n, (size_t)2 * kMixNB) + 256);
    b += 2 * DNT * DNT * kPackTile * sizeof(float) + 1024;
    b += (4 * (size_t)kMixNB * kMixNB + (size_t)n * 2 * kMixNB) * sizeof(float) + 1024;
    return b;
}

// Right-looking blocked LU, panels of kMixNB columns factored recursively (tournament
// leaves), with look-ahead: stream 1 factors panel k, applies its interchanges and
// updates the next panel's columns; stream 2 meanwhile applies panel k to all columns
// further right.  The streams never work on the same columns at the same time.
void mixed_lu_factor(bola_hip_handle_s* h, int lp, bool pivot, int n, float* T, size_t ldt, int32_t* ipiv, int* perm,
                     void* ws) {
    LuCtx c;
    c.st = h->stream;
    c.pivot = pivot;
    c.n = n;
    c.lp = lp;
    c.T = T;
    c.ldt = ldt;
    c.ipiv = ipiv;
    c.perm = perm;
    const int G = (n + GROWS - 1) / GROWS;
    const int levels = 8;
    c.lvl_stride = G;
    char* b = (char*)ws;
    size_t o = 0;
    c.lists = (int*)(b + o);
    o += ((size_t)levels * G * LPW * sizeof(int) + 255) & ~(size_t)255;
    c.lcount = (int*)(b + o);
    o += ((size_t)levels * G * sizeof(int) + 255) & ~(size_t)255;
    c.cnt = (int*)(b + o);
    o += ((size_t)levels * G * sizeof(int) + 255) & ~(size_t)255;
    const size_t gsz = (lp_gemm_ws_bytes(lp, (size_t)n, (size_t)n, (size_t)2 * kMixNB) + 511) & ~(size_t)255;
    c.gws = b + o;
    c.upk = (float*)(b + o + 2 * gsz);
    c.lpk = c.upk + DNT * DNT * kPackTile;
    // four panels' -(L11^-1)^T (panel index mod 4): stream 2 may still read a pair's while
    // stream 1 factors the next pair
    float* const nltb = (float*)((char*)(c.lpk + DNT * DNT * kPackTile) + 1024);
    c.ubuf = nltb + 4 * (size_t)kMixNB * kMixNB;
    LuCtx c2 = c;  // stream 2: its own 16-bit operand buffers
    c2.st = h->aux_stream();
    c2.gws = b + o + gsz;
    hipLaunchKernelGGL(k_zero_cnt, dim3(1), dim3(256), 0, c.st, c.cnt, levels * G);
    BOLA_LAUNCH_CHECK();
    if (!pivot) {  // identity interchanges: the solve path stays the same
        hipLaunchKernelGGL(k_iota_piv, dim3(cdiv((size_t)n, 256)), dim3(256), 0, c.st, n, ipiv);
        BOLA_LAUNCH_CHECK();
    }
    int ev = 0;
    hipEvent_t bulk_done = nullptr;
    if (!pivot) {
        // Butterfly mode, panels in pairs.  Stream 1: panel a, its update of panel b's
        // columns, panel b, then both panels on the next pair's columns; stream 2 meanwhile
        // applies the previous pair to every column after that.
        auto nltp = [&](int k) { return nltb + (size_t)((k / kMixNB) % 4) * kMixNB * kMixNB; };
        auto panel = [&](int k, int nb) {
            c.nlt = nltp(k);
            launch_panel_rbt(c, k, nb);
        };
        c2.ubuf = c.ubuf;
        for (int k = 0; k < n; k += 2 * kMixNB) {
            const int nba = std::min(kMixNB, n - k);
            const int kb = k + nba, nbb = std::min(kMixNB, n - kb);
            panel(k, nba);
            if (nbb > 0) {
                update_cols(c, k, nba, kb, kb + nbb);  // c.nlt: panel a's
                panel(kb, nbb);
            }
            const int lo = kb + std::max(nbb, 0);  // n unless both panels are full
            if (lo >= n) break;
            // the next pair's columns were the previous bulk update's until it is done
            if (bulk_done) BOLA_CHECK(hipStreamWaitEvent(c.st, bulk_done, 0));
            const int la_hi = std::min(n, lo + 2 * kMixNB);
            update_pair(c, k, nltp(k), nltp(kb), lo, la_hi);
            if (la_hi < n) {
                hipEvent_t la = handle_event(h, ev++);
                BOLA_CHECK(hipEventRecord(la, c.st));
                BOLA_CHECK(hipStreamWaitEvent(c2.st, la, 0));
                update_pair(c2, k, nltp(k), nltp(kb), la_hi, n);
                bulk_done = handle_event(h, ev++);
                BOLA_CHECK(hipEventRecord(bulk_done, c2.st));
            } else {
                bulk_done = nullptr;
            }
        }
        if (bulk_done) BOLA_CHECK(hipStreamWaitEvent(c.st, bulk_done, 0));
        return;
    }
    for (int k = 0; k < n; k += kMixNB) {
        const int nb = std::min(kMixNB, n - k);
        // the panel: its own columns take the interchanges at once
        c.swap_lo = k;
        c.swap_hi = k + nb;
        lu_rec(c, k, nb);
        // the columns right of the panel are stream 2's until its previous bulk update is done
        if (bulk_done) BOLA_CHECK(hipStreamWaitEvent(c.st, bulk_done, 0));
        if (c.pivot) {
            launch_laswp_t(c.st, 0, k, T, ldt, k, k + nb, ipiv, nullptr);
            launch_laswp_t(c.st, k + nb, n, T, ldt, k, k + nb, ipiv, nullptr);
        }
        // look-ahead: the next panel's columns
        const int c_mid = std::min(n, k + nb + kMixNB);
        update_cols(c, k, nb, k + nb, c_mid);
        if (c_mid < n) {
            hipEvent_t la = handle_event(h, ev++);
            BOLA_CHECK(hipEventRecord(la, c.st));
            BOLA_CHECK(hipStreamWaitEvent(c2.st, la, 0));
            update_cols(c2, k, nb, c_mid, n);
            bulk_done = handle_event(h, ev++);
            BOLA_CHECK(hipEventRecord(bulk_done, c2.st));
        } else {
            bulk_done = nullptr;
        }
    }
    if (bulk_done) BOLA_CHECK(hipStreamWaitEvent(c.st, bulk_done, 0));
}

int mixed_lu_max_n() { return 1 << 17; }

}  // namespace bola
