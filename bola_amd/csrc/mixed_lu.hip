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

// device-scope relaxed float accesses: coherent across XCDs without cache-wide fences
__device__ __forceinline__ void xstf(float* p, float v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float xldf(const float* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---- the butterfly mode: LU without interchanges, one launch per panel ----------------

constexpr int NPLD = kMixNB + 4;      // LDS row stride (floats) of panel rows
constexpr long kPanelSpin = 1L << 22;  // bound on the polls of one flag
constexpr size_t kInvPanel = 2 * (size_t)(kMixNB / LPW) * LPW * LPW;  // floats of one panel's inverses

// column b (a lane) of the inverse of d's unit lower triangle (LOWER) or of its upper
// triangle, substituted in o's column b.  d is the identity outside its leading block, so
// the inverse is too.  Rolled loops over LDS: unrolled, the compiler hoists the d reads
// and runs out of registers.
template <bool LOWER>
__device__ void tri_inv_col(const float (*d)[LPW + 1], float (*o)[LPW + 1], int b) {
    if (b >= LPW) return;
#pragma unroll 1
    for (int i = 0; i < LPW; i++) {
        const int a = LOWER ? i : LPW - 1 - i;
        const int lo = LOWER ? 0 : a + 1, hi = LOWER ? a : LPW;
        float v0 = a == b ? 1.f : 0.f, v1 = 0.f;
        int c = lo;
#pragma unroll 4
        for (; c + 1 < hi; c += 2) {
            v0 = fmaf(-d[a][c], o[c][b], v0);
            v1 = fmaf(-d[a][c + 1], o[c + 1][b], v1);
        }
        if (c < hi) v0 = fmaf(-d[a][c], o[c][b], v0);
        v0 += v1;
        if (LOWER) {
            o[a][b] = v0;
        } else {
            const float dd = d[a][a];
            o[a][b] = dd != 0.f ? v0 / dd : v0;
        }
    }
}

// one leaf (columns cj .. cj+32, zero-padded past the panel) applied to the 32 rows in a:
// L = A[:, leaf] U11^-1, then A[:, right of the leaf] -= L U12 with U12 in ub
__device__ __forceinline__ void rbt_apply_leaf(float (*a)[NPLD], const float (*ub)[NPLD], const float (*ui)[LPW + 1],
                                               int cj, int lw, int w) {
    const int tid = threadIdx.x;
    {
        const int r = tid >> 3, q = (tid & 7) * 4;
        float l[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 8
        for (int s = 0; s < LPW; s++) {
            const float x = a[r][cj + s];
#pragma unroll
            for (int i = 0; i < 4; i++) l[i] = fmaf(x, ui[s][q + i], l[i]);
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 4; i++)
            if (q + i < lw) a[r][cj + q + i] = l[i];
    }
    __syncthreads();
    const int c0 = cj + lw;
    if (c0 < w) {
        const int rq = (tid >> 6) * 8, cl = tid & 63;
        float acc[8][4];
#pragma unroll
        for (int rr = 0; rr < 8; rr++)
#pragma unroll
            for (int i = 0; i < 4; i++) acc[rr][i] = 0.f;
#pragma unroll 4
        for (int s = 0; s < LPW; s++) {
            float u[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int c = c0 + cl + 64 * i;
                u[i] = c < w ? ub[s][c] : 0.f;
            }
#pragma unroll
            for (int rr = 0; rr < 8; rr++) {
                const float x = a[rq + rr][cj + s];
#pragma unroll
                for (int i = 0; i < 4; i++) acc[rr][i] = fmaf(x, u[i], acc[rr][i]);
            }
        }
#pragma unroll
        for (int rr = 0; rr < 8; rr++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int c = c0 + cl + 64 * i;
                if (c < w) a[rq + rr][c] -= acc[rr][i];
            }
    }
    __syncthreads();
}

// LU without interchanges of the 32 x 32 block a[0..32)[cj..cj+32) (identity past lw) into
// d (L\U), by lanes 0..31 of one wavefront, lane r holding row r; the pivot row goes
// through LDS (one wavefront: its LDS accesses stay in program order).  Kept out of line:
// inlined beside the inverses it costs the kernel hundreds of registers.
__device__ __forceinline__ void diag_lu_rows(const float (*a)[NPLD], float (*d)[LPW + 1], int cj, int lw) {
    const int r = threadIdx.x;
    float rv[LPW];
#pragma unroll
    for (int c = 0; c < LPW; c++) rv[c] = (r < lw && c < lw) ? a[r][cj + c] : (r == c ? 1.f : 0.f);
#pragma unroll
    for (int j = 0; j < LPW; j++) {
        if (r == j)
#pragma unroll
            for (int c = j; c < LPW; c++) d[j][c] = rv[c];
        asm volatile("" ::: "memory");
        const float piv = d[j][j];
        const float iv = piv != 0.f ? 1.f / piv : 0.f;
        const float l = rv[j] * iv;
        if (r > j) {
            rv[j] = l;
#pragma unroll
            for (int c = j + 1; c < LPW; c++) rv[c] = fmaf(-l, d[j][c], rv[c]);
        }
    }
#pragma unroll
    for (int c = 0; c < LPW; c++)
        if (c < r) d[r][c] = rv[c];
}

// One launch factors the panel A[r0:n, r0:r0+w) (w <= kMixNB) without interchanges.  Each
// workgroup holds 32 rows in LDS.  Workgroup j < leaves owns leaf j (panel columns
// 32j ..): it applies leaves 0..j-1 to its rows, factors its diagonal block, forms
// U12 = L11^-1 A12 for the panel columns right of the leaf, stores its rows and the block
// inverses (inv[j] = U11^-1 then L11^-1) and raises flags[j] to epoch.  Every other
// workgroup applies the leaves as their flags rise.  Workgroups wait only for
// lower-numbered ones, which are dispatched first; every wait is bounded and a timeout
// raises err (the factors are then not to be trusted).
__global__ void __launch_bounds__(256) panel_rbt(int n, int r0, int w, float* __restrict__ T, size_t ldt,
                                                 float* __restrict__ inv, int* __restrict__ flags, int epoch,
                                                 int* __restrict__ err) {
    __shared__ float a[LPW][NPLD];   // this workgroup's rows, panel columns 0 .. w
    __shared__ float ub[LPW][NPLD];  // the current leaf's rows
    __shared__ float ui[LPW][LPW + 1], li[LPW][LPW + 1], d[LPW][LPW + 1];
    __shared__ int bail;
    const int tid = threadIdx.x, wb = blockIdx.x;
    const int row0 = r0 + wb * LPW;
    const int nr = min(LPW, n - row0);
    const int leaves = (w + LPW - 1) / LPW;
    const int col = tid;  // one panel column per thread in loads and stores
    float* const P = T + (size_t)r0;  // column r0 of A
    {
        // loads from clamped (valid) addresses, all in flight at once, then the selects
        const int cc = min(col, w - 1);
        float v[LPW];
#pragma unroll
        for (int r = 0; r < LPW; r++) v[r] = P[cc + (size_t)(row0 + min(r, nr - 1)) * ldt];
#pragma unroll
        for (int r = 0; r < LPW; r++) a[r][col] = (col < w && r < nr) ? v[r] : 0.f;
    }
    __syncthreads();
    const int napply = min(wb, leaves);
    for (int j = 0; j < napply; j++) {
        const int cj = j * LPW, lw = min(LPW, w - cj);
        if (tid == 0) {
            long sp = 0;
            int ok = 1;
            // the next owner polls eagerly; the others sparingly, so that hundreds of
            // pollers do not crowd the flag's cache line
            const bool next = wb == j + 1;
            while (__hip_atomic_load(flags + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < epoch) {
                if (++sp > kPanelSpin || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                    ok = 0;
                    break;
                }
                if (next) __builtin_amdgcn_s_sleep(1);
                else __builtin_amdgcn_s_sleep(32);
            }
            if (!ok) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            bail = !ok;
        }
        __syncthreads();
        if (bail) return;
        {
            // the owner's rows and inverses were stored coherently before its flag rose:
            // read them coherently too (no cache-wide acquire)
            const int cc = min(max(col, cj), w - 1);
            float v[LPW];
#pragma unroll
            for (int r = 0; r < LPW; r++) v[r] = xldf(&P[cc + (size_t)(r0 + cj + min(r, lw - 1)) * ldt]);
#pragma unroll
            for (int r = 0; r < LPW; r++) ub[r][col] = (r < lw && col >= cj && col < w) ? v[r] : 0.f;
        }
        for (int e = tid; e < LPW * LPW; e += 256) ui[e / LPW][e % LPW] = xldf(&inv[(size_t)j * 2 * LPW * LPW + e]);
        __syncthreads();
        rbt_apply_leaf(a, ub, ui, cj, lw, w);
    }
    if (wb < leaves) {
        // the owner of leaf wb: its rows are the leaf's diagonal block (a partial leaf only
        // occurs in a square last panel, where it is also the last rows)
        const int cj = wb * LPW, lw = min(LPW, w - cj);
        if (tid < LPW) diag_lu_rows(a, d, cj, lw);
        __syncthreads();
        // side by side: U11^-1 (wavefront 0), L11^-1 (wavefront 1) and U12 = L11^-1 A12 by
        // forward substitution, a column per thread (wavefronts 2 and 3)
        if (tid < 64) {
            tri_inv_col<false>(d, ui, tid);
        } else if (tid < 128) {
            tri_inv_col<true>(d, li, tid - 64);
        } else {
            for (int c = cj + lw + tid - 128; c < w; c += 128) {
#pragma unroll 1
                for (int r = 1; r < LPW; r++) {
                    float v0 = a[r][c], v1 = 0.f;
                    int s2 = 0;
#pragma unroll 4
                    for (; s2 + 1 < r; s2 += 2) {
                        v0 = fmaf(-d[r][s2], a[s2][c], v0);
                        v1 = fmaf(-d[r][s2 + 1], a[s2 + 1][c], v1);
                    }
                    if (s2 < r) v0 = fmaf(-d[r][s2], a[s2][c], v0);
                    a[r][c] = v0 + v1;
                }
            }
        }
        __syncthreads();
        if (tid < lw)
            for (int r = 0; r < lw; r++) a[r][cj + tid] = d[r][tid];
        __syncthreads();
        // the leaf's rows and inverses, stored coherently (device scope, past the XCD's L2)
        // and complete before the flag rises: the readers need no cache-wide fences
        for (int r = 0; r < nr; r++)
            if (col < w) xstf(&P[col + (size_t)(row0 + r) * ldt], a[r][col]);
        float* io = inv + (size_t)wb * 2 * LPW * LPW;
        for (int e = tid; e < LPW * LPW; e += 256) {
            xstf(&io[e], ui[e / LPW][e % LPW]);
            xstf(&io[LPW * LPW + e], li[e / LPW][e % LPW]);
        }
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (tid == 0) __hip_atomic_store(flags + wb, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    for (int r = 0; r < nr; r++)
        if (col < w) P[col + (size_t)(row0 + r) * ldt] = a[r][col];
}

// -L11^-1 of a factored panel (w x w, unit lower), stored transposed: nlt[c + r w] =
// -L11^-1(r, c), for the U12 products.  First its diagonal 32 x 32 blocks (the leaves'
// inverses, inv[j] + 32*32) and zeros above them; then, doubling the block size h = 32,
// 64, 128, .., the off-diagonal block of every pair of h-blocks:
//   [A 0; B C]^-1 = [A^-1 0; -C^-1 B A^-1  C^-1].
// L(r, c) is the panel's entry in T; one workgroup per 32 x 32 tile of -C^-1 B A^-1.
__global__ void __launch_bounds__(256) linv_diag(int w, const float* __restrict__ inv, float* __restrict__ nlt) {
    const int r = blockIdx.x;  // row of L11^-1
    for (int c = threadIdx.x; c < w; c += 256) {
        const int j = r / LPW;
        float v = 0.f;
        if (c / LPW == j) v = -inv[(size_t)j * 2 * LPW * LPW + LPW * LPW + (r % LPW) * LPW + (c % LPW)];
        if (c / LPW < j) continue;  // filled by the doubling steps
        nlt[(size_t)c + (size_t)r * w] = v;
    }
}

__global__ void __launch_bounds__(256) linv_level(int w, int h, const float* __restrict__ T, size_t ldt, int r0,
                                                  float* __restrict__ nlt) {
    const int tiles = (h + LPW - 1) / LPW;
    const int pair = blockIdx.x / (tiles * tiles), tt = blockIdx.x % (tiles * tiles);
    const int tr = tt / tiles, tc = tt % tiles;  // tile of Y: rows 32 tr.., columns 32 tc..
    const int i0 = 2 * h * pair;
    const int hc = min(h, w - i0 - h);           // rows of the C block
    if (hc <= 0 || 32 * tr >= hc) return;
    const int tid = threadIdx.x, lane_c = tid & 31, grp = tid >> 5;
    const int rmax = min(32 * tr + 32, hc);      // X rows needed: t < rmax
    constexpr int HM = kMixNB / 2;               // largest h
    // LDS kept under what a CU has left beside one trailing-update workgroup, so that this
    // launch is not starved by the update running on the other stream
    __shared__ float Bs[HM][LPW + 1];            // B(t, s) = L(i0 + h + t, i0 + s), t < rmax, one 32-chunk of s
    __shared__ float Ai[HM][LPW + 1];            // A^-1(s, c) for the tile's columns
    __shared__ float Xs[HM][LPW + 1];            // X = B A^-1, the tile's columns
    __shared__ float Ci[LPW][HM + 1];            // C^-1 rows of the tile
    const float* P = T + r0;
    auto Lv = [&](int r, int c) { return P[c + (size_t)(r0 + r) * ldt]; };           // L(r, c), r > c
    auto Li = [&](int r, int c) { return -nlt[(size_t)c + (size_t)r * w]; };       // L11^-1(r, c)
    // loads in batches of 8 per thread, all in flight together (clamped addresses)
    constexpr int NB8 = 8;
    for (int e0 = tid; e0 < h * LPW; e0 += 256 * NB8) {
        float v[NB8];
#pragma unroll
        for (int q = 0; q < NB8; q++) {
            const int e = min(e0 + 256 * q, h * LPW - 1), s2 = e / LPW, gc = min(32 * tc + e % LPW, h - 1);
            v[q] = Li(i0 + s2, i0 + gc);
        }
#pragma unroll
        for (int q = 0; q < NB8; q++) {
            const int e = e0 + 256 * q;
            if (e >= h * LPW) continue;
            const int s2 = e / LPW, c = e % LPW, gc = 32 * tc + c;
            Ai[s2][c] = (gc < h && s2 >= gc) ? v[q] : 0.f;
        }
    }
    for (int e0 = tid; e0 < LPW * hc; e0 += 256 * NB8) {
        float v[NB8];
#pragma unroll
        for (int q = 0; q < NB8; q++) {
            const int e = min(e0 + 256 * q, LPW * hc - 1), gr = min(32 * tr + e / hc, hc - 1), t = e % hc;
            v[q] = Li(i0 + h + gr, i0 + h + t);
        }
#pragma unroll
        for (int q = 0; q < NB8; q++) {
            const int e = e0 + 256 * q;
            if (e >= LPW * hc) continue;
            const int r = e / hc, t = e % hc, gr = 32 * tr + r;
            Ci[r][t] = (gr < hc && t <= gr) ? v[q] : 0.f;
        }
    }
    // X = B A^-1 over 32-column chunks of B; each thread keeps its rows' sums in registers
    constexpr int RPT = HM / 8;  // rows t per thread: grp, grp + 8, ..
    float xs[RPT];
#pragma unroll
    for (int q = 0; q < RPT; q++) xs[q] = 0.f;
    for (int s0 = 0; s0 < h; s0 += LPW) {
        __syncthreads();  // the previous chunk has been read
        for (int e0 = tid; e0 < rmax * LPW; e0 += 256 * NB8) {
            float v[NB8];
#pragma unroll
            for (int q = 0; q < NB8; q++) {
                const int e = min(e0 + 256 * q, rmax * LPW - 1);
                v[q] = Lv(i0 + h + e / LPW, i0 + s0 + e % LPW);
            }
#pragma unroll
            for (int q = 0; q < NB8; q++) {
                const int e = e0 + 256 * q;
                if (e < rmax * LPW) Bs[e / LPW][e % LPW] = v[q];
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < RPT; q++) {
            const int t = grp + 8 * q;
            if (t < rmax) {
                float x0 = 0.f, x1 = 0.f, x2 = 0.f, x3 = 0.f;
#pragma unroll
                for (int s2 = 0; s2 < LPW; s2 += 4) {
                    x0 = fmaf(Bs[t][s2], Ai[s0 + s2][lane_c], x0);
                    x1 = fmaf(Bs[t][s2 + 1], Ai[s0 + s2 + 1][lane_c], x1);
                    x2 = fmaf(Bs[t][s2 + 2], Ai[s0 + s2 + 2][lane_c], x2);
                    x3 = fmaf(Bs[t][s2 + 3], Ai[s0 + s2 + 3][lane_c], x3);
                }
                xs[q] += (x0 + x1) + (x2 + x3);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < RPT; q++) {
        const int t = grp + 8 * q;
        if (t < rmax) Xs[t][lane_c] = xs[q];
    }
    __syncthreads();
    for (int r = grp; r < LPW; r += 8) {
        const int gr = 32 * tr + r, gc = 32 * tc + lane_c;
        if (gr >= hc || gc >= h) continue;
        float y0 = 0.f, y1 = 0.f, y2 = 0.f, y3 = 0.f;
        int t = 0;
#pragma unroll 4
        for (; t + 3 <= gr; t += 4) {
            y0 = fmaf(Ci[r][t], Xs[t][lane_c], y0);
            y1 = fmaf(Ci[r][t + 1], Xs[t + 1][lane_c], y1);
            y2 = fmaf(Ci[r][t + 2], Xs[t + 2][lane_c], y2);
            y3 = fmaf(Ci[r][t + 3], Xs[t + 3][lane_c], y3);
        }
        for (; t <= gr; t++) y0 = fmaf(Ci[r][t], Xs[t][lane_c], y0);
        const float y = (y0 + y1) + (y2 + y3);
        // L11^-1(i0 + h + gr, i0 + gc) = -y; nlt holds the negative
        nlt[(size_t)(i0 + gc) + (size_t)(i0 + h + gr) * w] = y;
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
    float* inv;  // the butterfly mode: the current panel's diagonal-block inverses
    float* nlt;  // the butterfly mode: -(L11^-1)^T of the current panel (kMixNB x kMixNB)
    float* ubuf; // the butterfly mode: U12^T blocks, row c for A-column c (ld n)
    int* flags;  // the butterfly mode: per-leaf flags of panel_rbt
    int* err;
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

// The butterfly mode's panel [k, k+nb): one launch, the leaf owners and every row below.
void launch_panel_rbt(const LuCtx& c, int k, int nb) {
    hipLaunchKernelGGL(panel_rbt, dim3(cdiv((size_t)(c.n - k), LPW)), dim3(256), 0, c.st, c.n, k, nb, c.T, c.ldt,
                       c.inv, c.flags, k / kMixNB + 1, c.err);
    BOLA_LAUNCH_CHECK();
    // -L11^-1 for the U12 products
    hipLaunchKernelGGL(linv_diag, dim3((unsigned)nb), dim3(256), 0, c.st, nb, c.inv, c.nlt);
    BOLA_LAUNCH_CHECK();
    for (int h = LPW; h < nb; h *= 2) {
        const int tiles = (h + LPW - 1) / LPW;
        const int pairs = (nb + 2 * h - 1) / (2 * h);
        hipLaunchKernelGGL(linv_level, dim3((unsigned)(pairs * tiles * tiles)), dim3(256), 0, c.st, nb, h, c.T, c.ldt,
                           k, c.nlt);
        BOLA_LAUNCH_CHECK();
    }
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
    b += 2 * kInvPanel * sizeof(float) + 1024;
    b += (4 * (size_t)kMixNB * kMixNB + (size_t)n * 2 * kMixNB) * sizeof(float) + 1024;
    return b;
}

// Right-looking blocked LU, panels of kMixNB columns factored recursively (tournament
// leaves), with look-ahead: stream 1 factors panel k, applies its interchanges and
// updates the next panel's columns; stream 2 meanwhile applies panel k to all columns
// further right.  The streams never work on the same columns at the same time.
void mixed_lu_factor(bola_hip_handle_s* h, int lp, bool pivot, int n, float* T, size_t ldt, int32_t* ipiv, int* perm,
                     void* ws, int* err) {
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
    float* const invb = (float*)(b + o + 2 * gsz);  // two panels' inverses, by panel parity
    c.flags = (int*)(invb + 2 * kInvPanel);
    // four panels' -(L11^-1)^T (panel index mod 4): stream 2 may still read a pair's while
    // stream 1 factors the next pair
    float* const nltb = (float*)((char*)c.flags + 1024);
    c.ubuf = nltb + 4 * (size_t)kMixNB * kMixNB;
    c.err = err;
    LuCtx c2 = c;  // stream 2: its own 16-bit operand buffers
    c2.st = h->aux_stream();
    c2.gws = b + o + gsz;
    hipLaunchKernelGGL(k_zero_cnt, dim3(1), dim3(256), 0, c.st, c.cnt, levels * G);
    hipLaunchKernelGGL(k_zero_cnt, dim3(1), dim3(64), 0, c.st, c.flags, 2 * kMixNB / LPW);
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
            c.inv = c2.inv = invb + (size_t)((k / kMixNB) & 1) * kInvPanel;
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
        if (pivot) {
            lu_rec(c, k, nb);
        } else {
            // stream 2 may still read the previous panel's inverses: alternate two buffers
            const int par = (k / kMixNB) & 1;
            c.inv = c2.inv = invb + (size_t)par * kInvPanel;
            c.nlt = c2.nlt = nltb + (size_t)par * kMixNB * kMixNB;
            c2.ubuf = c.ubuf;
            launch_panel_rbt(c, k, nb);
        }
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
