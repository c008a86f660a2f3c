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
constexpr int CRB = 512;       // rows per workgroup in the local tournament round
constexpr int CMAX = 1024;     // proposals one merge workgroup holds
constexpr int CRS = LPW + 1;   // LDS row stride (floats)

struct Best {
    float v;
    int r;
};
__device__ __forceinline__ bool beats(float v, int r, float bv, int br) {
    return v > bv || (v == bv && r < br);
}
__device__ __forceinline__ Best wave_best(Best b) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const float ov = __shfl_xor(b.v, off, 64);
        const int orr = __shfl_xor(b.r, off, 64);
        if (beats(ov, orr, b.v, b.r)) { b.v = ov; b.r = orr; }
    }
    return b;
}

// Partial pivoting on the rows held in registers (own[q][c], row tid + q*NT of the set,
// `rows` rows, w <= 32 columns).  The step loop is unrolled, so column j of a row is a
// register; the pivot row goes through LDS once per step.  At the end every row holds
// its L\U row as far as it was factored; pick[s] = set index of the row chosen at step s,
// and stepof[q] = the step at which this thread's row q was chosen (-1 if never).
template <int J, int NT, int RPT>
__device__ __forceinline__ void pp_step(float (&own)[RPT][LPW], bool (&live)[RPT], int steps, int* pick,
                                        float (*prow)[LPW], Best (*red)[2], int (&stepof)[RPT]) {
    if constexpr (J < LPW) {
        constexpr int NWV = NT / 64;
        const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
        if (J < steps) {  // uniform
            Best b{-1.f, 0x7fffffff};
#pragma unroll
            for (int q = 0; q < RPT; q++) {
                if (live[q]) {
                    float v = fabsf(own[q][J]);
                    if (!(v == v)) v = -0.5f;  // NaN never wins over a number
                    if (beats(v, tid + q * NT, b.v, b.r)) { b.v = v; b.r = tid + q * NT; }
                }
            }
            b = wave_best(b);
            if (lane == 0) red[wv][J & 1] = b;
            __syncthreads();
            Best g = red[0][J & 1];
#pragma unroll
            for (int q = 1; q < NWV; q++)
                if (beats(red[q][J & 1].v, red[q][J & 1].r, g.v, g.r)) g = red[q][J & 1];
            const int p = g.r;
            if (tid == 0) pick[J] = p;
            // the owner publishes the pivot row (selected with bit masks, not a branch per row)
            uint32_t msk[RPT];
            uint32_t any = 0;
#pragma unroll
            for (int q = 0; q < RPT; q++) {
                msk[q] = (tid + q * NT == p) ? ~0u : 0u;
                any |= msk[q];
                live[q] = live[q] && !msk[q];
                stepof[q] = msk[q] ? J : stepof[q];
            }
            if (any) {
#pragma unroll
                for (int c = 0; c < LPW; c++) {
                    uint32_t v = 0;
#pragma unroll
                    for (int q = 0; q < RPT; q++) v |= __builtin_bit_cast(uint32_t, own[q][c]) & msk[q];
                    prow[J & 1][c] = __builtin_bit_cast(float, v);
                }
            }
            __syncthreads();
            const float piv = prow[J & 1][J];
            if (piv != 0.f) {
                const float inv = 1.f / piv;
                float pr[LPW];
#pragma unroll
                for (int c = J + 1; c < LPW; c++) pr[c] = prow[J & 1][c];
#pragma unroll
                for (int q = 0; q < RPT; q++) {
                    const float l = own[q][J] * inv;
                    own[q][J] = live[q] ? l : own[q][J];
                    const float lm = live[q] ? -l : 0.f;
#pragma unroll
                    for (int c = J + 1; c < LPW; c++) {
                        const float u = fmaf(lm, pr[c], own[q][c]);
                        own[q][c] = live[q] ? u : own[q][c];
                    }
                }
            }
        }
        pp_step<J + 1, NT, RPT>(own, live, steps, pick, prow, red, stepof);
    }
}

// Partial pivoting on the rows held in registers (own[q][c], row tid + q*NT of the set,
// `rows` rows, w <= 32 columns).  The steps are expanded at compile time, so column j of a
// row is a register; the pivot row goes through LDS once per step.  At the end every row
// holds its L\U row as far as it was factored; pick[s] = set index of the row chosen at
// step s, stepof[q] = the step at which this thread's row q was chosen (-1 if never).
template <int NT, int RPT>
__device__ __forceinline__ void reg_partial_pivot(float (&own)[RPT][LPW], int rows, int w, int* pick,
                                                  float (*prow)[LPW], Best (*red)[2], int (&stepof)[RPT]) {
    bool live[RPT];
#pragma unroll
    for (int q = 0; q < RPT; q++) {
        live[q] = (int)threadIdx.x + q * NT < rows;
        stepof[q] = -1;
    }
    pp_step<0, NT, RPT>(own, live, min(w, rows), pick, prow, red, stepof);
}

// round 1: every workgroup proposes min(w, its rows) pivot rows (absolute A-row indices)
__global__ void __launch_bounds__(256) calu_local(int n, int r0, int c0, int w, const float* __restrict__ T,
                                                  size_t ldt, int* __restrict__ prop, int* __restrict__ nprop) {
    constexpr int RPT = CRB / 256;
    __shared__ float prow[2][LPW];
    __shared__ Best red[4][2];
    __shared__ int pick[LPW];
    const int g = blockIdx.x, tid = threadIdx.x;
    const int rb = r0 + g * CRB;
    const int rows = min(CRB, n - rb);
    float own[RPT][LPW];
#pragma unroll
    for (int q = 0; q < RPT; q++) {
        const int row = tid + q * 256;
        const float* src = T + (size_t)c0 + (size_t)(rb + min(row, rows - 1)) * ldt;
#pragma unroll
        for (int c = 0; c < LPW; c++) own[q][c] = (row < rows && c < w) ? src[c] : 0.f;
    }
    int stepof[RPT];
    reg_partial_pivot<256, RPT>(own, rows, w, pick, prow, red, stepof);
    __syncthreads();
    const int k = min(w, rows);
    if (tid < k) prop[g * LPW + tid] = rb + pick[tid];
    if (tid == 0) nprop[g] = k;
}

// round 2: one workgroup runs partial pivoting on the proposals' original rows; its picks
// become LAPACK-style interchanges ipiv[r0 .. r0+w) (absolute, 0-based), and the picked
// rows' L\U rows, in pick order, are the factored top block (top[s * LPW + c]).
__global__ void __launch_bounds__(1024) calu_merge(int n, int r0, int c0, int w, const float* __restrict__ T,
                                                   size_t ldt, const int* __restrict__ prop,
                                                   const int* __restrict__ nprop, int nblk,
                                                   int32_t* __restrict__ ipiv, float* __restrict__ top) {
    __shared__ int rowid[CMAX];
    __shared__ float prow[2][LPW];
    __shared__ Best red[16][2];
    __shared__ int pick[LPW];
    __shared__ int total;
    __shared__ int at_row[2 * LPW], at_src[2 * LPW];  // moved positions -> original rows
    const int tid = threadIdx.x;
    if (tid == 0) {
        int t = 0;
        for (int g = 0; g < nblk; g++) {
            for (int q = 0; q < nprop[g]; q++) rowid[t + q] = prop[g * LPW + q];
            t += nprop[g];
        }
        total = t;
    }
    __syncthreads();
    const int rows = total;
    float own[1][LPW];
    {
        const float* src = T + (size_t)c0 + (size_t)rowid[min(tid, rows - 1)] * ldt;
#pragma unroll
        for (int c = 0; c < LPW; c++) own[0][c] = (tid < rows && c < w) ? src[c] : 0.f;
    }
    int stepof[1];
    reg_partial_pivot<1024, 1>(own, rows, w, pick, prow, red, stepof);
    if (stepof[0] >= 0) {
#pragma unroll
        for (int c = 0; c < LPW; c++) top[stepof[0] * LPW + c] = own[0][c];
    }
    __syncthreads();
    if (tid == 0) {
        // picks -> sequential interchanges: track where each original row currently is
        const int k = min(w, rows);
        int nm = 0;
        auto content = [&](int pos) {
            for (int q = 0; q < nm; q++)
                if (at_row[q] == pos) return at_src[q];
            return pos;
        };
        auto set = [&](int pos, int src) {
            for (int q = 0; q < nm; q++)
                if (at_row[q] == pos) { at_src[q] = src; return; }
            at_row[nm] = pos;
            at_src[nm] = src;
            nm++;
        };
        for (int s = 0; s < k; s++) {
            const int want = rowid[pick[s]];
            int pos = want;  // where the original row `want` is now
            for (int q = 0; q < nm; q++)
                if (at_src[q] == want) { pos = at_row[q]; break; }
            const int dst = r0 + s;
            ipiv[dst] = pos;
            if (pos != dst) {
                const int a = content(dst);
                set(dst, want);
                set(pos, a);
            }
        }
    }
}

// interchanges ipiv[k1 .. k2) applied to whole rows (all columns of A = all rows of T)
// and to the row-origin vector perm; one thread per column of A.
__global__ void laswp_t(int n, float* __restrict__ T, size_t ldt, int k1, int k2, const int32_t* __restrict__ ipiv,
                        int* __restrict__ perm) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c > n) return;
    for (int i = k1; i < k2; i++) {
        const int p = ipiv[i];
        if (p == i) continue;
        if (c < n) {
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

// the panel without pivoting, after the interchanges: L21 = A21 U11^-1 for the rows
// below the top block (one thread per row), with L11\U11 from the merge round; workgroup
// 0 also stores the factored top block.
__global__ void __launch_bounds__(256) panel_nopiv(int n, int r0, int c0, int w, float* __restrict__ T, size_t ldt,
                                                   const float* __restrict__ top) {
    __shared__ float U[LPW][LPW + 1];  // U[row][col] of the top block, A orientation
    const int tid = threadIdx.x;
    for (int e = tid; e < w * w; e += 256) {
        const int r = e / w, c = e % w;
        U[r][c] = top[r * LPW + c];
    }
    __syncthreads();
    if (blockIdx.x == 0) {
        for (int e = tid; e < w * w; e += 256) {
            const int r = e / w, c = e % w;
            T[(size_t)(c0 + c) + (size_t)(r0 + r) * ldt] = U[r][c];
        }
    }
    const int r = r0 + w + blockIdx.x * 256 + tid;
    if (r >= n) return;
    float x[LPW];
    float* col = T + (size_t)c0 + (size_t)r * ldt;
#pragma unroll
    for (int b = 0; b < LPW; b++) x[b] = b < w ? col[b] : 0.f;
#pragma unroll
    for (int b = 0; b < LPW; b++) {
        if (b < w) {
            float v = x[b];
#pragma unroll
            for (int a = 0; a < b; a++) v = fmaf(-x[a], U[a][b], v);
            x[b] = U[b][b] != 0.f ? v / U[b][b] : v;
        }
    }
#pragma unroll
    for (int b = 0; b < LPW; b++)
        if (b < w) col[b] = x[b];
}

// X U = B for unit upper U (w <= 32): X, B the nr x w column-major block at X (ld), U the
// w x w block at U (ld) (its strictly upper part).  One thread per row.
__global__ void __launch_bounds__(256) strsm_rupu_base(int nr, int w, const float* __restrict__ Ub, float* __restrict__ X,
                                                       size_t ld) {
    __shared__ float U[LPW][LPW + 1];
    const int tid = threadIdx.x;
    for (int e = tid; e < w * w; e += 256) {
        const int a = e % w, b = e / w;
        U[a][b] = Ub[(size_t)a + (size_t)b * ld];
    }
    __syncthreads();
    const int row = blockIdx.x * 256 + tid;
    if (row >= nr) return;
    float x[LPW];
#pragma unroll
    for (int b = 0; b < LPW; b++) x[b] = b < w ? X[(size_t)row + (size_t)b * ld] : 0.f;
#pragma unroll
    for (int b = 1; b < LPW; b++) {
        if (b < w) {
            float v = x[b];
#pragma unroll
            for (int a = 0; a < b; a++) v = fmaf(-x[a], U[a][b], v);
            x[b] = v;
        }
    }
#pragma unroll
    for (int b = 0; b < LPW; b++)
        if (b < w) X[(size_t)row + (size_t)b * ld] = x[b];
}

}  // namespace

size_t mixed_lu_ws_bytes(int n, int lp) {
    const size_t nblk = (size_t)(n + CRB - 1) / CRB;
    size_t b = nblk * LPW * sizeof(int) + nblk * sizeof(int) + LPW * LPW * sizeof(float) + 2048;
    b += lp_gemm_ws_bytes(lp, (size_t)n, (size_t)n, (size_t)n / 2 + 64);
    return b;
}

namespace {

struct LuCtx {
    hipStream_t st;
    int n, lp;
    float* T;
    size_t ldt;
    int32_t* ipiv;
    int* perm;
    int* prop;
    int* nprop;
    float* top;
    void* gws;
};

void calu_panel(const LuCtx& c, int r0, int w) {
    const int rows = c.n - r0;
    const int nblk = (rows + CRB - 1) / CRB;
    hipLaunchKernelGGL(calu_local, dim3(nblk), dim3(256), 0, c.st, c.n, r0, r0, w, c.T, c.ldt, c.prop, c.nprop);
    BOLA_LAUNCH_CHECK();
    hipLaunchKernelGGL(calu_merge, dim3(1), dim3(1024), 0, c.st, c.n, r0, r0, w, c.T, c.ldt, c.prop, c.nprop, nblk,
                       c.ipiv, c.top);
    BOLA_LAUNCH_CHECK();
    const int k = std::min(w, rows);
    hipLaunchKernelGGL(laswp_t, dim3(cdiv((size_t)c.n + 1, 256)), dim3(256), 0, c.st, c.n, c.T, c.ldt, r0, r0 + k,
                       c.ipiv, c.perm);
    BOLA_LAUNCH_CHECK();
    const int below = std::max(0, rows - w);
    hipLaunchKernelGGL(panel_nopiv, dim3(std::max(1u, cdiv((size_t)below, 256))), dim3(256), 0, c.st, c.n, r0, r0, w,
                       c.T, c.ldt, c.top);
    BOLA_LAUNCH_CHECK();
}

float* Tp(const LuCtx& c, int col, int row) { return c.T + (size_t)col + (size_t)row * c.ldt; }

// X U = B on T blocks: X/B at (xc, r0), nr x w; U at (r0c, r0), w x w unit upper
void strsm_rupu(const LuCtx& c, int nr, int w, int uc, int ur, int xc) {
    if (nr <= 0 || w <= 0) return;
    if (w <= LPW) {
        hipLaunchKernelGGL(strsm_rupu_base, dim3(cdiv((size_t)nr, 256)), dim3(256), 0, c.st, nr, w, Tp(c, uc, ur),
                           Tp(c, xc, ur), c.ldt);
        BOLA_LAUNCH_CHECK();
        return;
    }
    const int h1 = ((w / 2) + LPW - 1) / LPW * LPW, h2 = w - h1;
    strsm_rupu(c, nr, h1, uc, ur, xc);
    // B2 -= X1 * U12 : (nr x h2) -= (nr x h1)(h1 x h2)
    launch_lp_gemm(c.st, c.lp, nr, h2, h1, Tp(c, xc, ur), c.ldt, Tp(c, uc, ur + h1), c.ldt, Tp(c, xc, ur + h1),
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
    if (below > 0)
        launch_lp_gemm(c.st, c.lp, nr, below, w1, Tp(c, r0 + w1, r0), c.ldt, Tp(c, r0, r0 + w1), c.ldt,
                       Tp(c, r0 + w1, r0 + w1), c.ldt, c.gws);
    lu_rec(c, r0 + w1, nr);
}

}  // namespace

void mixed_lu_factor(hipStream_t st, int lp, int n, float* T, size_t ldt, int32_t* ipiv, int* perm, void* ws) {
    LuCtx c;
    c.st = st;
    c.n = n;
    c.lp = lp;
    c.T = T;
    c.ldt = ldt;
    c.ipiv = ipiv;
    c.perm = perm;
    const size_t nblk = (size_t)(n + CRB - 1) / CRB;
    char* b = (char*)ws;
    c.prop = (int*)b;
    c.nprop = (int*)(b + nblk * LPW * sizeof(int));
    const size_t o_top = (nblk * LPW * sizeof(int) + nblk * sizeof(int) + 255) & ~(size_t)255;
    c.top = (float*)(b + o_top);
    c.gws = b + o_top + ((LPW * LPW * sizeof(float) + 255) & ~(size_t)255);
    lu_rec(c, 0, n);
}

int mixed_lu_max_n() { return CMAX / LPW * CRB; }

}  // namespace bola
