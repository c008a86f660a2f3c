// LU with partial pivoting on gfx950: getrf.rs:37-101, getrf2.rs:39-185.
//
// The factorization keeps the reference's structure -- 32-column panels factored by
// the unblocked iterative_getrf2, the panel's interchanges applied to the columns on
// either side, TRSM(L,L,N,U) of the U12 block row and a rank-32 GEMM update -- because
// that structure fixes the rounding of every element.  Each step runs on the GPU:
//
//   * getf2_panel: the unblocked panel as one persistent multi-workgroup kernel.  Rows
//     of the m x w panel (w <= 32) live in registers, 256*R rows per workgroup.  Rows are
//     never moved during the panel: each thread tracks the logical position of its rows,
//     which is what the reference's physical swaps change.  Per column: every workgroup
//     reduces its candidates under the reference's IDAMAX rule (idamax.rs:23-133, keys
//     relative to the column's diagonal row), publishes its best vector-phase and tail
//     candidates with their row data, and raises its own flag; wavefront 0 of every
//     workgroup waits for all flags and picks the same pivot.  Then scale by the
//     reciprocal (getrf2.rs:156-164) and the GER with its pair/fused/tail structure
//     (ger.rs:28-90), all in registers.  At the end rows are stored at their final
//     positions and the composed permutation is emitted as a row map.
//   * the row map is applied to the left and right column ranges (laswp.rs semantics,
//     pure moves), the U12 TRSM uses the reference's base-case arithmetic, and the
//     trailing update is the slab-exact MFMA DGEMM.
//   * the m <= 32 < n case of iterative_getrf2 (one short, wide panel) runs as three
//     small kernels per column.
//
// Errors: a zero pivot stops the panel; every later kernel of the factorization sees
// the status word and does nothing, so the matrix is left as the reference leaves it
// (partially factored, panel pivots panel-local).
#include "common.hpp"

#include <mutex>

namespace bola {

namespace {

constexpr int PW = 32;     // panel width (getrf.rs:11 NB, getrf2.rs:11)
constexpr int PT = 256;    // threads per panel workgroup
constexpr uint64_t NOKEY = ~0ull;
constexpr uint64_t KEYMASK = (1ull << 60) - 1;

struct Cand {
    double v;
    uint64_t k;
};

// A published pivot candidate: its key, and its row as it stood before the previous
// step's deferred rank-1 update (x, fr: that row's multiplier and fused-row flag).
struct PanelSlot {
    double v;
    uint64_t key;
    double x;
    uint64_t flags;  // bit 0: fused row, bit 1: row updated
    double row[PW];
};
// The same for the wavefront panel, with the row window relative to the current column
// (win[c] = column j + c), so a candidate row is staged without per-column tests.
struct WaveSlot {
    double v;
    uint64_t key;
    double x;
    uint64_t flags;
    double win[PW];
};
static_assert(sizeof(WaveSlot) == sizeof(PanelSlot), "slot layouts share the workspace");
// A wavefront slot in global memory: each 32-bit half of a WaveSlot word travels in a
// 64-bit word whose upper half is the step's epoch, so readers poll the data itself.
constexpr int TW = 2 * (int)(sizeof(WaveSlot) / 8);

__device__ __forceinline__ bool better(double v, uint64_t k, double cv, uint64_t ck) {
    return v > cv || (v == cv && k < ck);
}
__device__ __forceinline__ uint64_t vec_key(uint64_t i) {
    return ((i & 3ull) << 62) | (((i >> 2) & 3ull) << 60) | i;
}

__device__ __forceinline__ double getc(const double (&row)[PW], int j) {
    double v = 0.0;
#pragma unroll
    for (int c = 0; c < PW; c++)
        if (c == j) v = row[c];
    return v;
}
__device__ __forceinline__ void setc(double (&row)[PW], int j, double v) {
#pragma unroll
    for (int c = 0; c < PW; c++)
        if (c == j) row[c] = v;
}

// reduce (v, k) pairs over the workgroup; result broadcast through LDS to every thread
__device__ void wg_reduce(Cand& c, Cand* sh, Cand* out) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int off = 32; off >= 1; off >>= 1) {
        double ov = __shfl_xor(c.v, off, 64);
        uint64_t ok = (uint64_t)__shfl_xor((long long)c.k, off, 64);
        if (better(ov, ok, c.v, c.k)) { c.v = ov; c.k = ok; }
    }
    if (lane == 0) sh[w] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        Cand b = sh[0];
        for (int q = 1; q < PT / 64; q++)
            if (better(sh[q].v, sh[q].k, b.v, b.k)) b = sh[q];
        *out = b;
    }
    __syncthreads();
    c = *out;
}

// reduce two candidate streams at once; every thread gets both results
__device__ __forceinline__ void wg_reduce2(Cand& a, Cand& b, Cand (*sh)[2]) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int off = 32; off >= 1; off >>= 1) {
        double ov = __shfl_xor(a.v, off, 64);
        uint64_t ok = (uint64_t)__shfl_xor((long long)a.k, off, 64);
        if (better(ov, ok, a.v, a.k)) { a.v = ov; a.k = ok; }
        ov = __shfl_xor(b.v, off, 64);
        ok = (uint64_t)__shfl_xor((long long)b.k, off, 64);
        if (better(ov, ok, b.v, b.k)) { b.v = ov; b.k = ok; }
    }
    if (lane == 0) { sh[w][0] = a; sh[w][1] = b; }
    __syncthreads();
    a = sh[0][0];
    b = sh[0][1];
#pragma unroll
    for (int q = 1; q < PT / 64; q++) {
        if (better(sh[q][0].v, sh[q][0].k, a.v, a.k)) a = sh[q][0];
        if (better(sh[q][1].v, sh[q][1].k, b.v, b.k)) b = sh[q][1];
    }
}

// the reference's final IDAMAX rule (idamax.rs:27-38, 91-130) on the best vector-phase
// and tail candidates; returns the winning offset from the diagonal row
__device__ __forceinline__ long idamax_rule(int len, const Cand& gv, const Cand& gt, bool& from_tail) {
    from_tail = false;
    if (len < 16) {
        from_tail = gt.v >= 0.0;
        return from_tail ? (long)gt.k : 0;
    }
    double V = 0.0;
    long vi = 0;
    if (gv.v > 0.0) { V = gv.v; vi = (long)(gv.k & KEYMASK); }
    if (gt.v > V) { from_tail = true; return (long)gt.k; }
    return vi;
}

constexpr int MAXWG_FAST = 32;  // up to this many workgroups, one read round per column

// One element of the reference's GER (ger.rs:28-90) for step j: column offset rc from
// j+1 (pair = rc/2), tc = -pivot_row[c], tp = -pivot_row[partner].  `np`, `nu`: columns
// right of j and their even part; `frow`: row inside the 4-aligned fused prefix.
__device__ __forceinline__ double ger_elem(double a, double x, double tc, double tp, int rc, int np, int nu,
                                           bool frow, bool upd) {
    const bool valid = upd && rc < np;
    const bool paired = rc < nu;
    const bool skip = !valid || (paired ? (tc == 0.0 && tp == 0.0) : (tc == 0.0));
    const double f = fma(x, tc, a);
    const double u = a + tc * x;
    return skip ? a : ((paired && frow) ? f : u);
}

// Deferred part of step j's GER on the row windows of one thread (slot 0 = column j+1):
// columns j+2.. (window slots 1..PW-1, rc = slot).  pr: step j's pivot row (absolute
// columns), zmask: bit i set when pr[i] == 0.  Which columns are touched (the pair skip
// of ger.rs) and which take the fused form are the same for every row, so they become
// two scalar bit masks; a row inside the fused prefix then costs one FMA per touched
// column, a row outside it the unfused multiply-add, a row that is not updated nothing.
template <int R>
__device__ __forceinline__ void ger_rest(double (&a)[R][PW], const double (&x)[R], const bool (&fr)[R],
                                         const bool (&up)[R], const double* pr, uint64_t zmask, int j, int np,
                                         int nu) {
    constexpr uint64_t EVEN = 0x5555555555555555ull;
    const uint64_t z = zmask >> (j + 1);                       // bit c: pr[j+1+c] == 0
    const uint64_t zpair = z & (((z & EVEN) << 1) | ((z >> 1) & EVEN));  // c and its partner c^1
    const uint64_t paired = (1ull << nu) - 1;
    const uint64_t skip = (zpair & paired) | (z & ~paired);
    const uint64_t touch = ((1ull << np) - 1) & ~skip & ~1ull;
    double tcv[PW];
#pragma unroll
    for (int c = 1; c < PW; c++) tcv[c] = -1.0 * pr[j + 1 + c];
#pragma unroll
    for (int s = 0; s < R; s++) {
        if (up[s]) {
            if (fr[s]) {
#pragma unroll
                for (int c = 1; c < PW; c++) {
                    if ((touch >> c) & 1) {
                        if ((paired >> c) & 1) {
                            a[s][c] = fma(x[s], tcv[c], a[s][c]);
                        } else {
                            // at most one such column; the empty asm keeps this a branch
                            // rather than a select between both forms in every column
                            asm volatile("");
                            a[s][c] = a[s][c] + tcv[c] * x[s];
                        }
                    }
                }
            } else {
#pragma unroll
                for (int c = 1; c < PW; c++)
                    if ((touch >> c) & 1) a[s][c] = a[s][c] + tcv[c] * x[s];
            }
        }
    }
}

// Exchange between workgroups uses relaxed device-scope atomic loads and stores, which
// are coherent across XCDs without cache flushes; a writer waits for its stores to
// complete before raising its flag, a reader loads the data after seeing the flag.
__device__ __forceinline__ void xst(double* p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double xld(const double* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void xstu(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t xldu(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Registers hold the not-yet-retired columns of each row as a window that shifts left
// by one column per step, so the current column is always slot 0 and every register
// index is a compile-time constant; a retired column goes to LDS (sret).  Step j's GER
// is split: column j+1 is updated at once (it is the next pivot column); columns j+2..
// are updated during step j+1 while the workgroups exchange candidates.  Candidate rows
// are published before that deferred update together with their multiplier, and the
// receiver applies it to the one winning row, so every element still sees the
// reference's sequence of operations.  MULTI: more than one workgroup (exchange through
// global memory); otherwise everything stays in LDS.
template <int R, bool MULTI>
__global__ void __launch_bounds__(PT) getf2_panel(int m, int w, double* __restrict__ A, size_t lda,
                                                  int32_t* __restrict__ ipiv, int32_t* __restrict__ info, int col_off,
                                                  long row_off, int* __restrict__ flags, PanelSlot* __restrict__ slots,
                                                  PanelSlot* __restrict__ cur, RowMap* __restrict__ map, int nwg,
                                                  long spin_limit) {
    __shared__ Cand red[PT / 64][2];
    __shared__ double prow[2][2 * PW + 2];  // pivot rows of consecutive steps (parity)
    __shared__ uint64_t pzero[2];           // zero masks of those rows
    __shared__ int ctl[2];
    __shared__ double sret[PW * PT * R];  // sret[c * PT*R + s*PT + tid]
    __shared__ PanelSlot lslot[MULTI ? 2 * MAXWG_FAST + 1 : 1];  // all workgroups' slots (exchange)
    __shared__ PanelSlot stg[3];  // this workgroup's candidates and the row at position j
    __shared__ int curmark;       // step whose row at position j this workgroup holds
    if (__hip_atomic_load(info, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;

    const int g = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    double a[R][PW];
    int pos[R];
    long rowid[R];
    bool valid[R];
#pragma unroll
    for (int s = 0; s < R; s++) {
        long r = (long)g * PT * R + (long)s * PT + tid;
        rowid[s] = r;
        valid[s] = r < m;
        pos[s] = (int)r;
#pragma unroll
        for (int c = 0; c < PW; c++) {  // clamped addresses: every load in flight at once
            const double v = A[min(r, (long)m - 1) + (size_t)min(c, w - 1) * lda];
            a[s][c] = keep_or_zero(v, valid[s] && c < w);
        }
    }
    for (int e = tid; e < 2 * (2 * PW + 2); e += PT) (&prow[0][0])[e] = 0.0;
    if (tid < 2) pzero[tid] = ~0ull;
    if (tid == 0) curmark = -1;
    __syncthreads();

    const int min_mn = min(m, w);
    const bool fast = nwg <= MAXWG_FAST;
    const int CUR = 2 * MAXWG_FAST;
    double xprev[R];
    bool fprev[R], uprev[R];
#pragma unroll
    for (int s = 0; s < R; s++) { xprev[s] = 0.0; fprev[s] = false; uprev[s] = false; }
    int dnp = 0, dnu = 0, dj = -1;  // previous step index and its column counts
    int nret = 0;
    for (int j = 0; j < min_mn; j++) {
        const int len = m - j;
        const int T = len < 16 ? 0 : (len & ~15);
        const int par = j & 1;
        Cand vb{-1.0, NOKEY}, tb{-1.0, NOKEY};
#pragma unroll
        for (int s = 0; s < R; s++) {
            if (valid[s] && pos[s] >= j) {
                double v = fabs(a[s][0]);
                if (v == v) {
                    uint64_t rel = (uint64_t)(pos[s] - j);
                    if ((long)rel < T) {
                        uint64_t k = vec_key(rel);
                        if (better(v, k, vb.v, vb.k)) { vb.v = v; vb.k = k; }
                    } else if (better(v, rel, tb.v, tb.k)) {
                        tb.v = v;
                        tb.k = rel;
                    }
                }
            }
        }
        wg_reduce2(vb, tb, red);
        // stage this workgroup's candidate rows (pre-update form) and the row at position j
        // in LDS: stg[0] vector-phase best, stg[1] tail best, stg[2] position j
        {
            int sv = -1, st = -1, sc = -1;
#pragma unroll
            for (int s = 0; s < R; s++) {
                if (valid[s] && pos[s] >= j) {
                    const uint64_t rel = (uint64_t)(pos[s] - j);
                    if ((long)rel < T) { if (vb.v >= 0.0 && vec_key(rel) == vb.k) sv = s; }
                    else if (tb.v >= 0.0 && rel == tb.k) st = s;
                    if (rel == 0) sc = s;
                }
            }
#pragma unroll
            for (int q = 0; q < 3; q++) {
                const int sx = q == 0 ? sv : (q == 1 ? st : sc);
                if (sx >= 0) {
                    PanelSlot* d = stg + q;
                    double xv = xprev[0];
                    bool f0 = fprev[0], u0 = uprev[0];
#pragma unroll
                    for (int s = 1; s < R; s++)
                        if (sx == s) { xv = xprev[s]; f0 = fprev[s]; u0 = uprev[s]; }
                    d->x = xv;
                    d->flags = (f0 ? 1ull : 0ull) | (u0 ? 2ull : 0ull);
#pragma unroll
                    for (int c = 0; c < PW; c++) {
                        if (c < w - j) {
                            double v = a[0][c];
#pragma unroll
                            for (int s = 1; s < R; s++)
                                if (sx == s) v = a[s][c];
                            d->row[j + c] = v;
                        }
                    }
                    if (q == 2) curmark = j;
                }
            }
        }
        if (tid == 0) {
            stg[0].v = vb.v; stg[0].key = vb.k;
            stg[1].v = tb.v; stg[1].key = tb.k;
            if (!MULTI) {
                bool ft;
                const long rel = idamax_rule(len, vb, tb, ft);
                ctl[0] = (int)rel;
                ctl[1] = rel == 0 ? 2 : (ft ? 1 : 0);
            }
        }
        __syncthreads();
        if (MULTI && wv == 0) {
            // wavefront 0 copies the staged slots out and raises this workgroup's flag once
            // its stores have completed
            PanelSlot* mine = slots + ((size_t)par * nwg + g) * 2;
            const bool hc = curmark == j;
            constexpr int NW = sizeof(PanelSlot) / 8;
            for (int e = lane; e < 3 * NW; e += 64) {
                const int q = e / NW, f = e % NW;
                if (q < 2 || hc) {
                    uint64_t* dst = (uint64_t*)(q < 2 ? mine + q : cur + par);
                    xstu(dst + f, ((const uint64_t*)(stg + q))[f]);
                }
            }
            __builtin_amdgcn_s_waitcnt(0);
            if (lane == 0) __hip_atomic_store(flags + g, j + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        // the deferred columns of the previous step, while the other workgroups catch up
        if (dj >= 0) ger_rest<R>(a, xprev, fprev, uprev, prow[dj & 1], pzero[dj & 1], dj, dnp, dnu);
        if (MULTI) {
            int ok = 1;
            for (int q = tid; q < nwg; q += PT) {
                long sp = 0;
                // a zero pivot (info > 0) is no reason to stop: the flags workgroup 0 saw
                // reach every workgroup, which then stops on the same pivot
                while (__hip_atomic_load(flags + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < j + 1) {
                    if (++sp > spin_limit) {  // timed out: report it unless a status is set
                        int zero = 0;
                        __hip_atomic_compare_exchange_strong(info, &zero, BOLA_HIP_ERR_DEVICE, __ATOMIC_RELAXED,
                                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        ok = 0;
                        break;
                    }
                    if (__hip_atomic_load(info, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < 0) {
                        ok = 0;
                        break;
                    }
                }
            }
            ok = __syncthreads_and(ok);
            if (!ok) return;
            const PanelSlot* base = slots + (size_t)par * nwg * 2;
            if (fast) {
                // one read round: every candidate (key, multiplier, row) and the row at position j;
                // all loads of a thread are issued before the first is used
                constexpr int NW = sizeof(PanelSlot) / 8;
                constexpr int NLD = ((2 * MAXWG_FAST + 1) * NW + PT - 1) / PT;
                const int nslot = 2 * nwg;
                const int total = (nslot + 1) * NW;
                uint64_t tmp[NLD];
#pragma unroll
                for (int k = 0; k < NLD; k++) {
                    const int e = tid + k * PT;
                    if (e < total) {
                        const int q = e / NW, f = e % NW;
                        const uint64_t* srcp =
                            (q < nslot) ? (const uint64_t*)(base + q) : (const uint64_t*)(cur + par);
                        tmp[k] = xldu(srcp + f);
                    }
                }
#pragma unroll
                for (int k = 0; k < NLD; k++) {
                    const int e = tid + k * PT;
                    if (e < total) {
                        const int q = e / NW, f = e % NW;
                        ((uint64_t*)(lslot + (q < nslot ? q : CUR)))[f] = tmp[k];
                    }
                }
                __syncthreads();
            }
            if (wv == 0) {
                Cand gv{-1.0, NOKEY}, gt{-1.0, NOKEY};
                int gvs = -1, gts = -1;
                for (int q = lane; q < nwg; q += 64) {
                    double v0, v1;
                    uint64_t k0, k1;
                    if (fast) {
                        v0 = lslot[2 * q].v; v1 = lslot[2 * q + 1].v;
                        k0 = lslot[2 * q].key; k1 = lslot[2 * q + 1].key;
                    } else {
                        v0 = xld(&base[2 * q].v); v1 = xld(&base[2 * q + 1].v);
                        k0 = xldu(&base[2 * q].key); k1 = xldu(&base[2 * q + 1].key);
                    }
                    if (better(v0, k0, gv.v, gv.k)) { gv.v = v0; gv.k = k0; gvs = 2 * q; }
                    if (better(v1, k1, gt.v, gt.k)) { gt.v = v1; gt.k = k1; gts = 2 * q + 1; }
                }
                for (int off = 32; off >= 1; off >>= 1) {
                    double ov = __shfl_xor(gv.v, off, 64);
                    uint64_t okk = (uint64_t)__shfl_xor((long long)gv.k, off, 64);
                    int os = __shfl_xor(gvs, off, 64);
                    if (better(ov, okk, gv.v, gv.k)) { gv.v = ov; gv.k = okk; gvs = os; }
                    ov = __shfl_xor(gt.v, off, 64);
                    okk = (uint64_t)__shfl_xor((long long)gt.k, off, 64);
                    os = __shfl_xor(gts, off, 64);
                    if (better(ov, okk, gt.v, gt.k)) { gt.v = ov; gt.k = okk; gts = os; }
                }
                if (lane == 0) {
                    bool ft;
                    const long r2 = idamax_rule(len, gv, gt, ft);
                    ctl[0] = (int)r2;
                    ctl[1] = r2 == 0 ? CUR : (ft ? gts : gvs);
                }
            }
        }
        __syncthreads();
        // the winning row: apply the previous step's deferred GER to it, then it is the pivot row
        if (wv == 0) {
            const int sidx = ctl[1];
            const bool in_lds = !MULTI || fast;
            const int c = lane;  // absolute column
            double v = 0.0;
            if (c >= j && c < w) {
                if (in_lds) {
                    const PanelSlot* P = MULTI ? lslot + sidx : stg + sidx;
                    v = P->row[c];
                    if (dj >= 0 && c > j) {
                        const int rc = c - (dj + 1);
                        const double* pr = prow[dj & 1];
                        const double tc = -1.0 * pr[c];
                        const double tp = -1.0 * ((rc & 1) ? pr[c - 1] : pr[c + 1]);
                        v = ger_elem(v, P->x, tc, tp, rc, dnp, dnu, (P->flags & 1) != 0, (P->flags & 2) != 0);
                    }
                } else {
                    const PanelSlot* P = (sidx == CUR) ? cur + par : slots + (size_t)par * nwg * 2 + sidx;
                    v = xld(&P->row[c]);
                    if (dj >= 0 && c > j) {
                        const double px = xld(&P->x);
                        const uint64_t pf = xldu(&P->flags);
                        const int rc = c - (dj + 1);
                        const double* pr = prow[dj & 1];
                        const double tc = -1.0 * pr[c];
                        const double tp = -1.0 * ((rc & 1) ? pr[c - 1] : pr[c + 1]);
                        v = ger_elem(v, px, tc, tp, rc, dnp, dnu, (pf & 1) != 0, (pf & 2) != 0);
                    }
                }
            }
            if (c < PW) prow[par][c] = v;
            const uint64_t zm = __ballot(v == 0.0);  // lanes >= w carry 0.0
            if (lane == 0) pzero[par] = zm;
        }
        __syncthreads();
        const int jp = j + ctl[0];
        const double piv = prow[par][j];
        if (g == 0 && tid == 0) ipiv[j] = jp + 1;
        if (!(piv != 0.0)) {
            if (g == 0 && tid == 0) {
                info[1] = j + 1;
                __hip_atomic_store(info, col_off + j + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            break;
        }
        if (jp != j) {
#pragma unroll
            for (int s = 0; s < R; s++) {
                if (pos[s] == j) pos[s] = jp;
                else if (pos[s] == jp) pos[s] = j;
            }
        }
        if (j < m - 1) {
            const double inv = 1.0 / piv;
#pragma unroll
            for (int s = 0; s < R; s++)
                if (valid[s] && pos[s] > j) a[s][0] = a[s][0] * inv;
        }
        // step j's GER: column j+1 now, the rest deferred to the next step
        const int np = w - j - 1;
        const int nu = np - (np & 1);
        const int mf = (m - j - 1) & ~3;
        const bool do_ger = j < min_mn - 1;
        {
            const double* pr = prow[par];
            const double tc = -1.0 * pr[j + 1];
            const double tp = -1.0 * pr[j + 2];
#pragma unroll
            for (int s = 0; s < R; s++) {
                const bool upd = do_ger && valid[s] && pos[s] > j;
                const bool frow = (pos[s] - (j + 1)) < mf;
                a[s][1] = ger_elem(a[s][1], a[s][0], tc, tp, 0, np, nu, frow, upd);
                xprev[s] = a[s][0];
                fprev[s] = frow;
                uprev[s] = upd;
            }
        }
        dj = do_ger ? j : -1;
        dnp = np;
        dnu = nu;
        // retire column j and shift the window
#pragma unroll
        for (int s = 0; s < R; s++) {
            sret[(size_t)j * (PT * R) + s * PT + tid] = a[s][0];
#pragma unroll
            for (int c = 0; c < PW - 1; c++) a[s][c] = a[s][c + 1];
            a[s][PW - 1] = 0.0;
        }
        nret = j + 1;
    }
    // (no deferred update is pending here: the last step has no GER, and an error stops
    // the loop after the previous step's deferred columns were applied)
    // rows to their final positions; moved rows form the panel's row map
#pragma unroll
    for (int s = 0; s < R; s++) {
        if (valid[s]) {
            for (int c = 0; c < nret; c++)
                A[(size_t)pos[s] + (size_t)c * lda] = sret[(size_t)c * (PT * R) + s * PT + tid];
#pragma unroll
            for (int c = 0; c < PW; c++)
                if (nret + c < w) A[(size_t)pos[s] + (size_t)(nret + c) * lda] = a[s][c];
            if (pos[s] != rowid[s]) {
                int idx = atomicAdd(&map->count, 1);
                if (idx < ROWMAP_MAX) {
                    map->row[idx] = row_off + pos[s];
                    map->src[idx] = row_off + rowid[s];
                }
            }
        }
    }
}

// ---- the panel with one wavefront per workgroup and no workgroup barriers
//
// Every wavefront holds 64*R rows and takes part in the pivot search on its own: it
// reduces its candidates with DPP row operations, stages the candidate rows in its LDS
// and copies them to its global slots, every word tagged with the step's epoch; after
// the deferred columns of the previous step it polls every wavefront's keys until they
// carry the epoch, picks the same pivot as all the others and reads the winning row
// (again until its words carry the epoch).  No store waits for another: a word is
// valid when its tag is, so the exchange costs the stores' visibility, not their
// acknowledgement.  With one wavefront there is no exchange.
//
// Register layout of a row: a window of columns j .. w-2 that shifts left by one column
// per step (slot 0 is the current column, every index is a compile-time constant) and
// the panel's last column w-1 in a register of its own.  The last column is the only one
// a GER can leave unpaired (ger.rs:52-59), so every window column takes the same form
// and the deferred update is straight-line code.
constexpr int WV = 64;
constexpr int WW = PW - 1;  // window slots

// Best (v, k) under `better` over the wavefront, every lane receiving it, with the lane
// that held it (idx: that lane's value of idx).  The largest v comes from shuffles; when
// one lane holds it (the usual case) its key is read from that lane, when several do
// (equal magnitudes) the smallest of their keys decides.  Candidates are never NaN and
// an empty lane holds (-1, NOKEY), so an empty wavefront yields (-1, NOKEY) again.
__device__ __forceinline__ double rdlane_d(double v, int l) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ uint64_t lane_u64(uint64_t x, int l) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), l);
    return ((uint64_t)hi << 32) | lo;
}
// x from another lane by a DPP row operation (control C, rows in RM; lanes outside RM
// keep their own x)
template <int C, int RM = 0xF>
__device__ __forceinline__ double dpp_f64(double x) {
    const int lo = __builtin_amdgcn_update_dpp(__double2loint(x), __double2loint(x), C, RM, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(x), __double2hiint(x), C, RM, 0xF, false);
    return __hiloint2double(hi, lo);
}
// the maxima of a and b over the wavefront, in every lane: quad swaps and row rotations
// leave each row of 16 with its maximum, two row broadcasts carry it to lane 63
__device__ __forceinline__ void wave_max2(double& a, double& b) {
    a = fmax(a, dpp_f64<0xB1>(a)); b = fmax(b, dpp_f64<0xB1>(b));                // quad_perm 1,0,3,2
    a = fmax(a, dpp_f64<0x4E>(a)); b = fmax(b, dpp_f64<0x4E>(b));                // quad_perm 2,3,0,1
    a = fmax(a, dpp_f64<0x124>(a)); b = fmax(b, dpp_f64<0x124>(b));              // row_ror 4
    a = fmax(a, dpp_f64<0x128>(a)); b = fmax(b, dpp_f64<0x128>(b));              // row_ror 8
    a = fmax(a, dpp_f64<0x142, 0xA>(a)); b = fmax(b, dpp_f64<0x142, 0xA>(b));    // row_bcast 15
    a = fmax(a, dpp_f64<0x143, 0xC>(a)); b = fmax(b, dpp_f64<0x143, 0xC>(b));    // row_bcast 31
    a = __longlong_as_double((long long)lane_u64((uint64_t)__double_as_longlong(a), 63));
    b = __longlong_as_double((long long)lane_u64((uint64_t)__double_as_longlong(b), 63));
}
// the lane holding the best candidate once the maximum m is known
__device__ __forceinline__ int best_lane(const Cand& a, double m) {
    uint64_t h = __ballot(a.v == m);
    if (__popcll(h) != 1) {  // equal magnitudes in several lanes: the smallest key
        uint64_t k = a.v == m ? a.k : NOKEY;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const uint64_t o = (uint64_t)__shfl_xor((long long)k, off, 64);
            k = o < k ? o : k;
        }
        h = __ballot(a.v == m && a.k == k);
    }
    return __ffsll((unsigned long long)h) - 1;
}
// best (v, k) of two candidate streams over the wavefront, every lane receiving both;
// la, lb: the lanes that held them
__device__ __forceinline__ void wave_best2(Cand& a, Cand& b, int& la, int& lb) {
    double ma = a.v, mb = b.v;
    wave_max2(ma, mb);
    la = best_lane(a, ma);
    lb = best_lane(b, mb);
    a.k = lane_u64(a.k, la);
    b.k = lane_u64(b.k, lb);
    a.v = ma;
    b.v = mb;
}

__device__ __forceinline__ void wave_reduce2(Cand& a, Cand& b) {
    int la, lb;
    wave_best2(a, b, la, lb);
}

// Deferred part of step dj's GER (ger.rs:28-90): window slots 1.. (column dj+1+c, rc = c)
// and the last column (rc = np-1, when np >= 2).  pr: step dj's pivot row (absolute
// columns), zmask: bit i set when pr[i] == 0.  Slots past the window's valid part meet
// coefficients of columns >= w-1 and turn into values that are never read.
template <int R>
__device__ __forceinline__ void wave_ger_rest(double (&a)[R][WW], double (&al)[R], const double (&x)[R],
                                              const bool (&fr)[R], const bool (&up)[R], const double* pr,
                                              uint64_t zmask, int dj, int np, int nu) {
    constexpr uint64_t EVEN = 0x5555555555555555ull;
    const uint64_t z = zmask >> (dj + 1);  // bit rc: pr[dj+1+rc] == 0
    const uint64_t zpair = z & (((z & EVEN) << 1) | ((z >> 1) & EVEN));
    const uint64_t paired = (1ull << nu) - 1;
    const uint64_t skip = (zpair & paired) | (z & ~paired);
    const uint64_t touch = ((1ull << np) - 1) & ~skip & ~1ull;
    double tcv[WW];
#pragma unroll
    for (int c = 1; c < WW; c++) tcv[c] = -1.0 * pr[dj + 1 + c];
    const uint64_t want = paired & ~1ull;
    if ((touch & want) == want) {
#pragma unroll
        for (int s = 0; s < R; s++) {
            if (up[s]) {
                if (fr[s]) {
#pragma unroll
                    for (int c = 1; c < WW; c++) a[s][c] = fma(x[s], tcv[c], a[s][c]);
                } else {
#pragma unroll
                    for (int c = 1; c < WW; c++) a[s][c] = a[s][c] + tcv[c] * x[s];
                }
            }
        }
    } else {
#pragma unroll
        for (int s = 0; s < R; s++) {
            if (up[s]) {
                if (fr[s]) {
#pragma unroll
                    for (int c = 1; c < WW; c++)
                        if ((touch >> c) & 1) a[s][c] = fma(x[s], tcv[c], a[s][c]);
                } else {
#pragma unroll
                    for (int c = 1; c < WW; c++)
                        if ((touch >> c) & 1) a[s][c] = a[s][c] + tcv[c] * x[s];
                }
            }
        }
    }
    if (np >= 2 && ((touch >> (np - 1)) & 1)) {
        const double tl = -1.0 * pr[dj + np];  // column w-1
        const bool lp = !(np & 1);             // the last column is paired
#pragma unroll
        for (int s = 0; s < R; s++)
            if (up[s]) al[s] = (lp && fr[s]) ? fma(x[s], tl, al[s]) : al[s] + tl * x[s];
    }
}

__device__ __forceinline__ uint64_t join32(uint64_t lo, uint64_t hi) { return (lo & 0xffffffffull) | (hi << 32); }

// A wavefront stops waiting: when its own wait timed out (some lane's `timed`) it reports
// a device error, unless a status is already set -- a zero pivot's column is never
// overwritten.  A wavefront that stopped because another one gave up reports nothing.
__device__ __forceinline__ void give_up(int32_t* info, int timed, int lane) {
    if (__any(timed) && lane == 0) {
        int zero = 0;
        __hip_atomic_compare_exchange_strong(info, &zero, BOLA_HIP_ERR_DEVICE, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
    }
}

// slots: [2 parities][nwv][2 candidates][TW] tagged words, cur: [2 parities][TW]; a step's
// epoch is col_off + j + 1, unique within a factorization whose slots start zeroed.
template <int R, bool MULTI>
__global__ void __launch_bounds__(WV) getf2_wave(int m, int w, double* __restrict__ A, size_t lda,
                                                 int32_t* __restrict__ ipiv, int32_t* __restrict__ info, int col_off,
                                                 long row_off, uint64_t* __restrict__ slots,
                                                 uint64_t* __restrict__ cur, RowMap* __restrict__ map, int nwv,
                                                 long spin_limit) {
    __shared__ double sret[PW * WV * R];  // retired columns: sret[c * WV*R + s*WV + lane]
    // every lane's rows as they stand before the deferred update: [0] multiplier x, [1] form
    // flags, [2 + c] window slot c (column j + c); 35 doubles a row, so that the lanes' stores
    // of one slot meet at most two to a bank and a row reads without conflicts
    constexpr int RS = 35;
    __shared__ double rst[R][WV][RS];
    __shared__ double prow[2][2 * PW + 2];
    if (__hip_atomic_load(info, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;

    const int g = blockIdx.x, lane = threadIdx.x;
    const int wl = w - 1;  // the last column
    double a[R][WW], al[R];
    int pos[R];
    long rowid[R];
    bool valid[R];
#pragma unroll
    for (int s = 0; s < R; s++) {
        long r = (long)g * WV * R + (long)s * WV + lane;
        rowid[s] = r;
        valid[s] = r < m;
        pos[s] = (int)r;
        const long rc = min(r, (long)m - 1);  // clamped addresses: every load in flight at once
#pragma unroll
        for (int c = 0; c < WW; c++) {
            const double v = A[rc + (size_t)min(c, max(wl - 1, 0)) * lda];
            a[s][c] = keep_or_zero(v, valid[s] && c < wl);
        }
        const double vl = A[rc + (size_t)wl * lda];
        al[s] = keep_or_zero(vl, valid[s]);
    }
    for (int e = lane; e < 2 * (2 * PW + 2); e += WV) (&prow[0][0])[e] = 0.0;
    uint64_t pz0 = ~0ull, pz1 = ~0ull;  // zero masks of prow[0], prow[1]
    __syncthreads();

    const int min_mn = min(m, w);
    double xprev[R];
    bool fprev[R], uprev[R];
#pragma unroll
    for (int s = 0; s < R; s++) { xprev[s] = 0.0; fprev[s] = false; uprev[s] = false; }
    int dnp = 0, dnu = 0, dj = -1;
    int nret = 0;
    for (int j = 0; j < min_mn; j++) {
        const int len = m - j;
        const int T = len < 16 ? 0 : (len & ~15);
        const int par = j & 1;
        const bool lastcol = j == wl;  // the current column is the last one
        Cand vb{-1.0, NOKEY}, tb{-1.0, NOKEY};
#pragma unroll
        for (int s = 0; s < R; s++) {
            if (valid[s] && pos[s] >= j) {
                double v = fabs(lastcol ? al[s] : a[s][0]);
                if (v == v) {
                    uint64_t rel = (uint64_t)(pos[s] - j);
                    if ((long)rel < T) {
                        uint64_t k = vec_key(rel);
                        if (better(v, k, vb.v, vb.k)) { vb.v = v; vb.k = k; }
                    } else if (better(v, rel, tb.v, tb.k)) {
                        tb.v = v;
                        tb.k = rel;
                    }
                }
            }
        }
        int lv, lt;  // the lanes holding the wavefront's best candidates
        wave_best2(vb, tb, lv, lt);
        // every lane stages its rows (pre-update form); the candidates' and the row at
        // position j are read from there by lane
        int sv = -1, st = -1, sc = -1;
#pragma unroll
        for (int s = 0; s < R; s++) {
            if (valid[s] && pos[s] >= j) {
                const uint64_t rel = (uint64_t)(pos[s] - j);
                if ((long)rel < T) { if (vb.v >= 0.0 && vec_key(rel) == vb.k) sv = s; }
                else if (tb.v >= 0.0 && rel == tb.k) st = s;
                if (rel == 0) sc = s;
            }
        }
        const uint64_t hc = __ballot(sc >= 0);
        const bool hascur = hc != 0;
        const int lc = hascur ? __ffsll((unsigned long long)hc) - 1 : 0;
        // (lane, row slot) of each staged row: the vector-phase candidate, the tail one, the row at j
        int ql[3], qs[3];
        ql[0] = lv;
        ql[1] = lt;
        ql[2] = lc;
        qs[0] = R == 1 ? 0 : max(__builtin_amdgcn_readlane(sv, lv), 0);
        qs[1] = R == 1 ? 0 : max(__builtin_amdgcn_readlane(st, lt), 0);
        qs[2] = R == 1 ? 0 : max(__builtin_amdgcn_readlane(sc, lc), 0);
#pragma unroll
        for (int s = 0; s < R; s++) {
            double* d = rst[s][lane];
            d[0] = xprev[s];
            d[1] = __longlong_as_double((long long)((fprev[s] ? 1ull : 0ull) | (uprev[s] ? 2ull : 0ull)));
#pragma unroll
            for (int c = 0; c < WW; c++) d[2 + c] = a[s][c];  // slots past the window are never read
            d[2 + wl - j] = al[s];
        }
        __syncthreads();  // one wavefront: orders the staging writes before the reads below
        const uint32_t ep = (uint32_t)(col_off + j + 1);
        if (MULTI) {
            // publish without waiting for the stores: the keys always, a row only for a
            // candidate that exists (an empty one cannot win), the row at j when held here
            const uint64_t tag = (uint64_t)ep << 32;
            uint64_t* mine = slots + ((size_t)par * nwv + g) * 2 * TW;
            constexpr int NE = (3 * TW + WV - 1) / WV;
            uint32_t hv[NE];
#pragma unroll
            for (int i = 0; i < NE; i++) {  // every LDS read in flight before the first store
                const int e = min(lane + i * WV, 3 * TW - 1);
                const int q = e / TW, f2 = e % TW, f = f2 >> 1;
                // words 0, 1: the candidate's magnitude and key; from 2 on: the staged row
                const int qi = q == 0 ? 0 : (q == 1 ? 1 : 2);
                const uint64_t head = f == 0 ? (uint64_t)__double_as_longlong(qi == 0 ? vb.v : tb.v)
                                             : (qi == 0 ? vb.k : tb.k);
                const double* rrow = rst[qi == 0 ? qs[0] : (qi == 1 ? qs[1] : qs[2])]
                                        [qi == 0 ? ql[0] : (qi == 1 ? ql[1] : ql[2])];
                const uint32_t body = ((const uint32_t*)rrow)[max(f2 - 4, 0)];
                hv[i] = f >= 2 ? body : (uint32_t)(head >> (32 * (f2 & 1)));
            }
#pragma unroll
            for (int i = 0; i < NE; i++) {
                const int e = lane + i * WV;
                const int q = e / TW, f2 = e % TW;
                const bool need = e < 3 * TW && (q < 2 ? (f2 < 4 || (q == 0 ? vb.v : tb.v) >= 0.0) : hascur);
                if (need) {
                    uint64_t* dst = q < 2 ? mine + q * TW : cur + (size_t)par * TW;
                    xstu(dst + f2, tag | hv[i]);
                }
            }
        }
        // the deferred columns of the previous step, while the other wavefronts catch up
        if (dj >= 0) wave_ger_rest<R>(a, al, xprev, fprev, uprev, prow[dj & 1], (dj & 1) ? pz1 : pz0, dj, dnp, dnu);
        // the pivot: every wavefront reaches the same decision
        const int c = lane;  // absolute column held by this lane in the pivot row
        double v = 0.0, px = 0.0;
        uint64_t pf = 0;
        long rel;
        if (MULTI) {
            int ok = 1, timed = 0;
            const uint64_t* base = slots + (size_t)par * nwv * 2 * TW;
            Cand gv{-1.0, NOKEY}, gt{-1.0, NOKEY};
            int gvs = -1, gts = -1;
            for (int q0 = 0; q0 < nwv && ok; q0 += 2 * WV) {  // two wavefronts' keys per lane
                uint64_t t[2][8];
                long sp = 0;
                const bool two = q0 + WV < nwv;  // a second wavefront for some lane
                for (;;) {  // until every word carries this step's epoch
#pragma unroll
                    for (int h = 0; h < 2; h++) {
                        const uint64_t* s0 = base + (size_t)2 * min(q0 + h * WV + lane, nwv - 1) * TW;
                        if (h == 0 || two) {
#pragma unroll
                            for (int i = 0; i < 4; i++) {
                                t[h][i] = xldu(s0 + i);
                                t[h][4 + i] = xldu(s0 + TW + i);
                            }
                        } else {
#pragma unroll
                            for (int i = 0; i < 8; i++) t[h][i] = t[0][i];
                        }
                    }
                    bool good = true;
#pragma unroll
                    for (int h = 0; h < 2; h++)
#pragma unroll
                        for (int i = 0; i < 8; i++) good = good && (uint32_t)(t[h][i] >> 32) == ep;
                    if (good) break;
                    // a zero pivot found by wavefront 0 (info > 0) is no reason to stop: the
                    // step's words it saw reach every wavefront, which then stops on it too
                    if (++sp > spin_limit) { ok = 0; timed = 1; break; }
                    if (__hip_atomic_load(info, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < 0) { ok = 0; break; }
                }
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const int q = q0 + h * WV + lane;
                    if (q < nwv) {
                        const double v0 = __longlong_as_double((long long)join32(t[h][0], t[h][1]));
                        const double v1 = __longlong_as_double((long long)join32(t[h][4], t[h][5]));
                        const uint64_t k0 = join32(t[h][2], t[h][3]), k1 = join32(t[h][6], t[h][7]);
                        if (better(v0, k0, gv.v, gv.k)) { gv.v = v0; gv.k = k0; gvs = 2 * q; }
                        if (better(v1, k1, gt.v, gt.k)) { gt.v = v1; gt.k = k1; gts = 2 * q + 1; }
                    }
                }
            }
            if (!__all(ok)) {
                give_up(info, timed, lane);
                return;
            }
            int lv, lt;
            wave_best2(gv, gt, lv, lt);
            gvs = __builtin_amdgcn_readlane(gvs, lv);
            gts = __builtin_amdgcn_readlane(gts, lt);
            bool ft;
            rel = idamax_rule(len, gv, gt, ft);
            const uint64_t* P = rel == 0 ? cur + (size_t)par * TW : base + (size_t)(ft ? gts : gvs) * TW;
            const bool inw = c >= j && c < w;
            const int wi = 2 * (4 + (inw ? c - j : 0));  // win[c - j]; x and flags are words 2 and 3
            uint64_t r0, r1, x0, x1, f0, f1;
            long sp = 0;
            for (;;) {
                r0 = xldu(P + wi);
                r1 = xldu(P + wi + 1);
                x0 = xldu(P + 4);
                x1 = xldu(P + 5);
                f0 = xldu(P + 6);
                f1 = xldu(P + 7);
                const bool good = (!inw || ((uint32_t)(r0 >> 32) == ep && (uint32_t)(r1 >> 32) == ep)) &&
                                  (uint32_t)(x0 >> 32) == ep && (uint32_t)(x1 >> 32) == ep &&
                                  (uint32_t)(f0 >> 32) == ep && (uint32_t)(f1 >> 32) == ep;
                if (good) break;
                if (++sp > spin_limit) { ok = 0; timed = 1; break; }
                if (__hip_atomic_load(info, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < 0) { ok = 0; break; }
            }
            if (!__all(ok)) {
                give_up(info, timed, lane);
                return;
            }
            if (inw) v = __longlong_as_double((long long)join32(r0, r1));
            px = __longlong_as_double((long long)join32(x0, x1));
            pf = join32(f0, f1);
        } else {
            bool ft;
            rel = idamax_rule(len, vb, tb, ft);
            const int qi = rel == 0 ? 2 : (ft ? 1 : 0);
            const double* P = rst[qs[qi]][ql[qi]];
            if (c >= j && c < w) v = P[2 + c - j];
            px = P[0];
            pf = (uint64_t)__double_as_longlong(P[1]);
        }
        // the winning row receives the previous step's deferred update and becomes the pivot row
        if (dj >= 0 && c > j && c < w) {
            const int rc = c - (dj + 1);
            const double* pr = prow[dj & 1];
            const double tc = -1.0 * pr[c];
            const double tp = -1.0 * ((rc & 1) ? pr[c - 1] : pr[c + 1]);
            v = ger_elem(v, px, tc, tp, rc, dnp, dnu, (pf & 1) != 0, (pf & 2) != 0);
        }
        if (c < PW) prow[par][c] = v;
        const uint64_t zm = __ballot(v == 0.0);  // lanes outside [j, w) carry 0.0
        if (par) pz1 = zm; else pz0 = zm;
        __syncthreads();
        const int jp = j + (int)rel;
        const double piv = prow[par][j];
        if (g == 0 && lane == 0) ipiv[j] = jp + 1;
        if (!(piv != 0.0)) {
            if (g == 0 && lane == 0) {
                info[1] = j + 1;
                __hip_atomic_store(info, col_off + j + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            break;
        }
        if (jp != j) {
#pragma unroll
            for (int s = 0; s < R; s++) {
                if (pos[s] == j) pos[s] = jp;
                else if (pos[s] == jp) pos[s] = j;
            }
        }
        // scale (getrf2.rs:156-164) and step j's GER: column j+1 now, the rest deferred
        const int np = w - j - 1;
        const int nu = np - (np & 1);
        const int mf = (m - j - 1) & ~3;
        const bool do_ger = j < min_mn - 1;
        const double inv = 1.0 / piv;
        const double* pr = prow[par];
        const double tc = -1.0 * pr[j + 1];
        const double tp = -1.0 * pr[j + 2];
#pragma unroll
        for (int s = 0; s < R; s++) {
            double xj = lastcol ? al[s] : a[s][0];
            if (j < m - 1 && valid[s] && pos[s] > j) xj = xj * inv;
            const bool upd = do_ger && valid[s] && pos[s] > j;
            const bool frow = (pos[s] - (j + 1)) < mf;
            if (j + 1 == wl) al[s] = ger_elem(al[s], xj, tc, tp, 0, np, nu, frow, upd);
            else a[s][1] = ger_elem(a[s][1], xj, tc, tp, 0, np, nu, frow, upd);
            xprev[s] = xj;
            fprev[s] = frow;
            uprev[s] = upd;
            // retire column j and shift the window
            sret[(size_t)j * (WV * R) + s * WV + lane] = xj;
#pragma unroll
            for (int c2 = 0; c2 < WW - 1; c2++) a[s][c2] = a[s][c2 + 1];
            a[s][WW - 1] = 0.0;
        }
        dj = do_ger ? j : -1;
        dnp = np;
        dnu = nu;
        nret = j + 1;
    }
    // rows to their final positions; moved rows form the panel's row map
#pragma unroll
    for (int s = 0; s < R; s++) {
        if (valid[s]) {
            for (int c = 0; c < nret; c++)
                A[(size_t)pos[s] + (size_t)c * lda] = sret[(size_t)c * (WV * R) + s * WV + lane];
#pragma unroll
            for (int c = 0; c < WW; c++)
                if (nret + c < wl) A[(size_t)pos[s] + (size_t)(nret + c) * lda] = a[s][c];
            if (nret <= wl) A[(size_t)pos[s] + (size_t)wl * lda] = al[s];
            if (pos[s] != rowid[s]) {
                int idx = atomicAdd(&map->count, 1);
                if (idx < ROWMAP_MAX) {
                    map->row[idx] = row_off + pos[s];
                    map->src[idx] = row_off + rowid[s];
                }
            }
        }
    }
}

// ---- the panel with two wavefronts per workgroup
//
// The exchange of getf2_wave with the work on a block of 64*R rows split between two
// wavefronts of one workgroup, which run on two SIMDs of one CU, so that the pivot search
// no longer waits for the rank-1 updates (a lone wavefront issues one instruction every
// few cycles, and the updates, the staging and the window shift were most of them):
//   * the pivot wavefront (wave 0) holds column j of the block: it reduces the candidates
//     of step j, publishes its best one, polls every workgroup's, reads the winning row,
//     scales column j, and brings column j+1 -- handed over by the update wavefront --
//     through step j;
//   * the update wavefront (wave 1) holds columns j+1 .. w-1: while the pivot wavefront
//     searches step j it applies step j-1's rank-1 update to them, hands column j+1 over,
//     stages the rows in LDS and, once the step's candidate is known, publishes the
//     candidate row and the row at position j, already through step j-1, so that a
//     receiver applies nothing to them.
// Two workgroup barriers a step: X_j (the candidate is known, column j+1 and the rows are
// staged) and Y_j (pivot, multipliers and pivot row of step j are in LDS).  Each element
// receives the reference's operations in the reference's order and form (getrf2.rs:131-185,
// ger.rs:28-90); only the wavefront that applies them changes.
//
// Pivot candidates form one stream keyed so that the reference's IDAMAX rule (idamax.rs:
// 23-133) reduces to "largest magnitude, then smallest key": a vector-phase row (offset
// rel < T from the diagonal) has key (rel mod 4, (rel div 4) mod 4, rel) in bits 34-35,
// 32-33 and 0-31, a tail row 2^36 + rel, so vector-phase rows win ties against tail rows;
// an empty candidate is (0, NOKEY).  The winner (v, k) maps to the reference's index as
// idamax_rule does: nothing -> 0; len < 16 -> rel; else rel when v > 0, 0 otherwise.
//
// Global words (zeroed per factorization, each 32-bit half tagged with the step's epoch):
// keys [2 parities][4][nwv] (candidate x lo, hi, key lo, hi; word-major so a poll reads
// consecutive words), rows [2][nwv][2][PW] (candidate row at absolute columns: the low
// halves, then the high halves), cur [2][2 + 2 PW] (x, then the row at position j laid out
// as a candidate row).
constexpr int DUO_SR = 33;      // staged row stride (doubles)
constexpr int DUO_CW = 2 + 2 * PW;
constexpr int kDuoMax = 256;    // workgroups: all must be resident at once
constexpr int kDuoPollPause = 6;  // s_sleep units (64 clocks) between publishing and the first poll

__device__ __forceinline__ uint64_t duo_key(int rel, int T) {
    const uint64_t r = (uint64_t)(uint32_t)rel;
    return rel < T ? (((r & 3ull) << 34) | (((r >> 2) & 3ull) << 32) | r) : ((1ull << 36) | r);
}
__device__ __forceinline__ double wave_max1(double a) {
    a = fmax(a, dpp_f64<0xB1>(a));            // quad_perm 1,0,3,2
    a = fmax(a, dpp_f64<0x4E>(a));            // quad_perm 2,3,0,1
    a = fmax(a, dpp_f64<0x124>(a));           // row_ror 4
    a = fmax(a, dpp_f64<0x128>(a));           // row_ror 8
    a = fmax(a, dpp_f64<0x142, 0xA>(a));      // row_bcast 15
    a = fmax(a, dpp_f64<0x143, 0xC>(a));      // row_bcast 31
    return __longlong_as_double((long long)lane_u64((uint64_t)__double_as_longlong(a), 63));
}
// best (v, k) over the wavefront into every lane; returns the lane that held it
__device__ __forceinline__ int wave_best1(double& v, uint64_t& k) {
    const double mx = wave_max1(v);
    const int l = best_lane(Cand{v, k}, mx);
    k = lane_u64(k, l);
    v = mx;
    return l;
}

// Step dj's rank-1 update of the update wavefront's window (slot c = column dj+2+c, rc =
// c+1) and of the last column; the forms of wave_ger_rest.
template <int R>
__device__ __forceinline__ void duo_ger_rest(double (&a)[R][WW], double (&al)[R], const double (&x)[R],
                                             const bool (&fr)[R], const bool (&up)[R], const double* pr,
                                             uint64_t zmask, int dj, int np, int nu) {
    constexpr uint64_t EVEN = 0x5555555555555555ull;
    const uint64_t z = zmask >> (dj + 1);  // bit rc: pr[dj+1+rc] == 0
    const uint64_t zpair = z & (((z & EVEN) << 1) | ((z >> 1) & EVEN));
    const uint64_t paired = (1ull << nu) - 1;
    const uint64_t skip = (zpair & paired) | (z & ~paired);
    const uint64_t touch = ((1ull << np) - 1) & ~skip & ~1ull;
    double tcv[WW - 1];
#pragma unroll
    for (int c = 0; c < WW - 1; c++) tcv[c] = -1.0 * pr[dj + 2 + c];
    const uint64_t want = paired & ~1ull;
    if ((touch & want) == want) {
#pragma unroll
        for (int s = 0; s < R; s++) {
            if (up[s]) {
                if (fr[s]) {
#pragma unroll
                    for (int c = 0; c < WW - 1; c++) a[s][c] = fma(x[s], tcv[c], a[s][c]);
                } else {
#pragma unroll
                    for (int c = 0; c < WW - 1; c++) a[s][c] = a[s][c] + tcv[c] * x[s];
                }
            }
        }
    } else {
#pragma unroll
        for (int s = 0; s < R; s++) {
            if (up[s]) {
                if (fr[s]) {
#pragma unroll
                    for (int c = 0; c < WW - 1; c++)
                        if ((touch >> (c + 1)) & 1) a[s][c] = fma(x[s], tcv[c], a[s][c]);
                } else {
#pragma unroll
                    for (int c = 0; c < WW - 1; c++)
                        if ((touch >> (c + 1)) & 1) a[s][c] = a[s][c] + tcv[c] * x[s];
                }
            }
        }
    }
    if (np >= 2 && ((touch >> (np - 1)) & 1)) {
        const double tl = -1.0 * pr[dj + np];  // column w-1
        const bool lp = !(np & 1);             // the last column is paired
#pragma unroll
        for (int s = 0; s < R; s++)
            if (up[s]) al[s] = (lp && fr[s]) ? fma(x[s], tl, al[s]) : al[s] + tl * x[s];
    }
}

// m > 64 R >= w: at least two workgroups, min(m, w) = w.
template <int R>
__global__ void __launch_bounds__(2 * WV) getf2_duo(int m, int w, double* __restrict__ A, size_t lda,
                                                    int32_t* __restrict__ ipiv, int32_t* __restrict__ info,
                                                    int col_off, long row_off, uint64_t* __restrict__ slots,
                                                    uint64_t* __restrict__ cur, RowMap* __restrict__ map, int nwv,
                                                    long spin_limit) {
    __shared__ double sret[PW][R * WV];    // retired columns (pivot wavefront)
    __shared__ double stg[R][WV][DUO_SR];  // columns j+1 .. w-1 of every row (update wavefront)
    __shared__ double colx[R][WV];         // column j+1 through step j-1: update -> pivot wavefront
    __shared__ double lbx[R][WV];          // step j's multipliers: pivot -> update wavefront
    __shared__ int lbf[R][WV];             // their form flags: bit 0 fused row, bit 1 updated row
    __shared__ double ubuf[2 * PW + 2];    // step j's pivot row at absolute columns, 0 outside [j, w)
    __shared__ uint64_t uzero;             // its zero mask: bit c set when column c is 0
    __shared__ int ctl[3];                 // stop; candidate (lane * R + s, or -1); row at j (or -1)

    const int g = blockIdx.x, lane = threadIdx.x & 63;
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wl = w - 1;
    uint64_t* keys = slots;                    // [2][4][nwv]
    uint64_t* rows = slots + (size_t)8 * nwv;  // [2][nwv][2 PW]
    long rowid[R];
    bool valid[R];
#pragma unroll
    for (int s = 0; s < R; s++) {
        rowid[s] = (long)g * WV * R + (long)s * WV + lane;
        valid[s] = rowid[s] < m;
    }

    if (wid == 0) {
        // ---------------- pivot wavefront
        double xc[R];  // column j of the rows, through step j-1
        int pos[R];
#pragma unroll
        for (int s = 0; s < R; s++) {
            pos[s] = (int)rowid[s];
            xc[s] = keep_or_zero(A[min(rowid[s], (long)m - 1)], valid[s]);
        }
        if (lane == 0) ctl[0] = __hip_atomic_load(info, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ? 2 : 0;
        __syncthreads();  // B0
        if (ctl[0]) return;
        int jstop = w, stop = 0;
        for (int j = 0; j < w; j++) {
            const int len = m - j;
            const int T = len < 16 ? 0 : (len & ~15);
            const int par = j & 1;
            const uint32_t ep = (uint32_t)(col_off + j + 1);
            const uint64_t tag = (uint64_t)ep << 32;
            // this block's best candidate and the row at position j
            double bv = 0.0, bx = 0.0, xcur = 0.0;
            uint64_t bk = NOKEY;
            int bs = 0, scur = -1;
#pragma unroll
            for (int s = 0; s < R; s++) {
                if (valid[s] && pos[s] >= j) {
                    const int rel = pos[s] - j;
                    if (rel == 0) { scur = s; xcur = xc[s]; }
                    const double v = fabs(xc[s]);
                    if (v == v) {
                        const uint64_t k = duo_key(rel, T);
                        if (better(v, k, bv, bk)) { bv = v; bk = k; bx = xc[s]; bs = s; }
                    }
                }
            }
            const int lw = wave_best1(bv, bk);
            bx = rdlane_d(bx, lw);
            bs = __builtin_amdgcn_readlane(bs, lw);
            const uint64_t hc = __ballot(scur >= 0);
            const int lc = hc ? __ffsll((unsigned long long)hc) - 1 : -1;
            const double xcu = lc >= 0 ? rdlane_d(xcur, lc) : 0.0;
            const int scu = lc >= 0 ? __builtin_amdgcn_readlane(scur, lc) : 0;
            if (lane < 4) {
                const uint64_t src = lane < 2 ? (uint64_t)__double_as_longlong(bx) : bk;
                xstu(keys + (size_t)(par * 4 + lane) * nwv + g, tag | (uint32_t)(src >> (32 * (lane & 1))));
            } else if (lane < 6 && lc >= 0) {
                xstu(cur + (size_t)par * DUO_CW + (lane - 4),
                     tag | (uint32_t)((uint64_t)__double_as_longlong(xcu) >> (32 * (lane - 4))));
            }
            if (lane == 0) {
                ctl[1] = bk != NOKEY ? lw * R + bs : -1;
                ctl[2] = lc >= 0 ? lc * R + scu : -1;
            }
            __syncthreads();  // X_j
            double xn[R];
#pragma unroll
            for (int s = 0; s < R; s++) xn[s] = colx[s][lane];
            // a pause first: polls issued while the other workgroups' stores are still in
            // flight slow those stores down (n = 8192 LU: 102.4 -> 97.5 us per panel launch)
            __builtin_amdgcn_s_sleep(kDuoPollPause);
            // every workgroup's candidate, until its words carry the epoch
            int ok = 1, timed = 0;
            double gv = 0.0, gx = 0.0;
            uint64_t gk = NOKEY;
            int gsl = 0;
            uint64_t cx0 = 0, cx1 = 0;
            const uint64_t* kb = keys + (size_t)par * 4 * nwv;
            const uint64_t* cxw = cur + (size_t)par * DUO_CW;
            for (int q0 = 0; q0 < nwv && ok; q0 += 2 * WV) {
                uint64_t t[2][4];
                const bool two = q0 + WV < nwv;
                const bool withx = q0 == 0 && lane == 0;
                const int qa = min(q0 + lane, nwv - 1), qb = min(q0 + WV + lane, nwv - 1);
                long sp = 0;
                for (;;) {
                    // one round trip an attempt: the keys, the x of the row at j and the
                    // status word (another workgroup gave up) are loaded together
#pragma unroll
                    for (int i = 0; i < 4; i++) t[0][i] = xldu(kb + (size_t)i * nwv + qa);
                    if (two) {
#pragma unroll
                        for (int i = 0; i < 4; i++) t[1][i] = xldu(kb + (size_t)i * nwv + qb);
                    } else {
#pragma unroll
                        for (int i = 0; i < 4; i++) t[1][i] = t[0][i];
                    }
                    if (withx) {
                        cx0 = xldu(cxw);
                        cx1 = xldu(cxw + 1);
                    }
                    const int stw = __hip_atomic_load(info, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    bool good = !withx || ((uint32_t)(cx0 >> 32) == ep && (uint32_t)(cx1 >> 32) == ep);
#pragma unroll
                    for (int h = 0; h < 2; h++)
#pragma unroll
                        for (int i = 0; i < 4; i++) good = good && (uint32_t)(t[h][i] >> 32) == ep;
                    if (good) break;
                    if (++sp > spin_limit) { ok = 0; timed = 1; break; }
                    if (stw < 0) { ok = 0; break; }
                }
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const int q = q0 + h * WV + lane;
                    if (q < nwv) {
                        const double x = __longlong_as_double((long long)join32(t[h][0], t[h][1]));
                        const uint64_t k = join32(t[h][2], t[h][3]);
                        const double v = fabs(x);
                        if (better(v, k, gv, gk)) { gv = v; gk = k; gx = x; gsl = q; }
                    }
                }
            }
            long rel = 0;
            double piv = 0.0, u = 0.0;
            double l[R];
            bool fr[R], up[R];
#pragma unroll
            for (int s = 0; s < R; s++) { l[s] = 0.0; fr[s] = false; up[s] = false; }
            if (!__all(ok)) {
                give_up(info, timed, lane);
                stop = 2;
            } else {
                const int lg = wave_best1(gv, gk);
                gx = rdlane_d(gx, lg);
                gsl = __builtin_amdgcn_readlane(gsl, lg);
                if (gk == NOKEY) rel = 0;
                else if (len < 16) rel = (long)(gk & 0xffffffffull);
                else rel = gv > 0.0 ? (long)(gk & 0xffffffffull) : 0;
                piv = rel == 0 ? __longlong_as_double((long long)lane_u64(join32(cx0, cx1), 0)) : gx;
                // the pivot row, columns j+1 .. w-1 by lane: the first loads go out now, the
                // scaling runs while they are in flight
                const uint64_t* prw =
                    rel == 0 ? cur + (size_t)par * DUO_CW + 2 : rows + ((size_t)par * nwv + gsl) * (2 * PW);
                const bool need = lane > j && lane < w;
                const uint64_t* src = prw + (need ? lane : 0);
                uint64_t r0 = 0, r1 = 0;
                if (need) {
                    r0 = xldu(src);
                    r1 = xldu(src + PW);
                }
                if (g == 0 && lane == 0) ipiv[j] = j + (int)rel + 1;
                if (!(piv != 0.0)) {
                    if (g == 0 && lane == 0) {
                        info[1] = j + 1;
                        __hip_atomic_store(info, col_off + j + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                    stop = 1;
                } else {
                    const int jp = j + (int)rel;
                    if (jp != j) {
#pragma unroll
                        for (int s = 0; s < R; s++) {
                            if (pos[s] == j) pos[s] = jp;
                            else if (pos[s] == jp) pos[s] = j;
                        }
                    }
                    // scale (getrf2.rs:156-164); multipliers and forms of step j's GER
                    const int mf = (m - j - 1) & ~3;
                    const bool do_ger = j < w - 1;
                    const double inv = 1.0 / piv;
#pragma unroll
                    for (int s = 0; s < R; s++) {
                        double xj = xc[s];
                        if (j < m - 1 && valid[s] && pos[s] > j) xj = xj * inv;
                        up[s] = do_ger && valid[s] && pos[s] > j;
                        fr[s] = (pos[s] - (j + 1)) < mf;
                        l[s] = xj;
                        sret[j][s * WV + lane] = xj;
                        lbx[s][lane] = xj;
                        lbf[s][lane] = (fr[s] ? 1 : 0) | (up[s] ? 2 : 0);
                    }
                    if (need) {
                        long sp = 0;
                        int stw = 0;
                        while ((uint32_t)(r0 >> 32) != ep || (uint32_t)(r1 >> 32) != ep) {
                            if (++sp > spin_limit) { ok = 0; timed = 1; break; }
                            if (stw < 0) { ok = 0; break; }
                            r0 = xldu(src);
                            r1 = xldu(src + PW);
                            stw = __hip_atomic_load(info, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        }
                    }
                    if (!__all(ok)) {
                        give_up(info, timed, lane);
                        stop = 2;
                    } else {
                        u = need ? __longlong_as_double((long long)join32(r0, r1)) : (lane == j ? piv : 0.0);
                        ubuf[lane] = u;
                        const uint64_t zm = __ballot(u == 0.0);  // lanes outside [j, w) carry 0.0
                        if (lane == 0) uzero = zm;
                    }
                }
            }
            if (lane == 0) ctl[0] = stop;
            __syncthreads();  // Y_j
            if (stop) {
                jstop = j;
                break;
            }
            if (j + 1 < w) {
                // column j+1 through step j (ger.rs:28-90; its partner is column j+2)
                const int np = w - j - 1, nu = np - (np & 1);
                const double tc = -1.0 * rdlane_d(u, j + 1);
                const double tp = -1.0 * rdlane_d(u, j + 2);
#pragma unroll
                for (int s = 0; s < R; s++) xc[s] = ger_elem(xn[s], l[s], tc, tp, 0, np, nu, fr[s], up[s]);
            }
        }
        if (stop == 2) return;
        // rows to their final positions; moved rows form the panel's row map.  After a zero
        // pivot at step jstop: columns before it retired, column jstop as it stood, the
        // columns after it as staged for that step (through step jstop-1).
#pragma unroll
        for (int s = 0; s < R; s++) {
            if (valid[s]) {
                for (int c = 0; c < jstop; c++) A[(size_t)pos[s] + (size_t)c * lda] = sret[c][s * WV + lane];
                if (jstop < w) {
                    A[(size_t)pos[s] + (size_t)jstop * lda] = xc[s];
                    for (int c = jstop + 1; c < w; c++)
                        A[(size_t)pos[s] + (size_t)c * lda] = stg[s][lane][c - jstop - 1];
                }
                if (pos[s] != rowid[s]) {
                    int idx = atomicAdd(&map->count, 1);
                    if (idx < ROWMAP_MAX) {
                        map->row[idx] = row_off + pos[s];
                        map->src[idx] = row_off + rowid[s];
                    }
                }
            }
        }
    } else {
        // ---------------- update wavefront: window slot c = column j+1+c, the last column apart
        double a[R][WW], al[R];
#pragma unroll
        for (int s = 0; s < R; s++) {
            const long rc = min(rowid[s], (long)m - 1);  // clamped addresses: every load in flight at once
#pragma unroll
            for (int c = 0; c < WW; c++) {
                const double v = A[rc + (size_t)min(1 + c, max(wl - 1, 0)) * lda];
                a[s][c] = keep_or_zero(v, valid[s] && 1 + c < wl);
            }
            al[s] = keep_or_zero(A[rc + (size_t)wl * lda], valid[s] && wl >= 1);
        }
        if (lane < 2) ubuf[2 * PW + lane] = 0.0;  // read past the panel by the window loop, never written
        __syncthreads();  // B0
        if (ctl[0]) return;
        for (int j = 0; j < w; j++) {
            const int par = j & 1;
            const uint64_t tag = (uint64_t)(uint32_t)(col_off + j + 1) << 32;
            if (j >= 1) {
                // step j-1's GER on columns j+1 .. w-1
                double x[R];
                bool fr[R], up[R];
#pragma unroll
                for (int s = 0; s < R; s++) {
                    x[s] = lbx[s][lane];
                    const int f = lbf[s][lane];
                    fr[s] = (f & 1) != 0;
                    up[s] = (f & 2) != 0;
                }
                const int np = w - j;  // columns right of column j-1
                duo_ger_rest<R>(a, al, x, fr, up, ubuf, uzero, j - 1, np, np - (np & 1));
            }
            if (j + 1 < w) {
#pragma unroll
                for (int s = 0; s < R; s++) {
                    colx[s][lane] = (j + 1 == wl) ? al[s] : a[s][0];
                    double* d = stg[s][lane];
#pragma unroll
                    for (int c = 0; c < WW; c++) d[c] = a[s][c];  // slots past the panel are never read
                    d[wl - j - 1] = al[s];
                }
            }
            __syncthreads();  // X_j
            {
                // the candidate row (lanes 0-31) and the row at position j (lanes 32-63), by
                // column: the low halves of the row's words, then its high halves, so each store
                // instruction fills whole lines
                const int h = lane >> 5, c = lane & 31;
                const int q = h == 0 ? ctl[1] : ctl[2];
                if (q >= 0 && c > j && c < w) {
                    const uint64_t b = (uint64_t)__double_as_longlong(stg[q % R][q / R][c - j - 1]);
                    uint64_t* dst = h == 0 ? rows + ((size_t)par * nwv + g) * (2 * PW) + c
                                           : cur + (size_t)par * DUO_CW + 2 + c;
                    xstu(dst, tag | (uint32_t)b);
                    xstu(dst + PW, tag | (uint32_t)(b >> 32));
                }
            }
#pragma unroll
            for (int s = 0; s < R; s++) {
#pragma unroll
                for (int c = 0; c < WW - 1; c++) a[s][c] = a[s][c + 1];
                a[s][WW - 1] = 0.0;
            }
            __syncthreads();  // Y_j
            if (ctl[0]) return;
        }
    }
}

// ---- the U rows of a factored super-panel for one block of 32 columns, in one launch
//
// For panels q = 0 .. G-1 of the super-panel (rows j0 = 32 q): first the GEMM of those
// rows against the U rows of the earlier panels, then the unit-lower TRSM with the
// panel's L11 -- the operations the separate launches perform, in their order and with
// their arithmetic:
//   * the GEMM as dgemm_mfma<32, 32> computes it with a 32-deep slab (getrf.rs:75-86):
//     each wavefront owns a 16 x 16 tile, every 32-deep slab is an 8-step
//     mfma_f64_16x16x4f64 chain over k = 0, 4, 8, .. from a zero accumulator, folded as
//     C = fma(-1, acc, C);
//   * the TRSM as trsm_ln_lanes<true, 32> computes it (trsm.rs:95-171, unit diagonal):
//     a lane per row, columns eliminated in order, a zero multiplier skipped, the first
//     (31 - k) & ~3 rows below the pivot fused, the rest unfused.
// A workgroup takes CB columns; their 32G x CB block lives in LDS between the steps.  L: the super-panel's first row and
// column (a + s0 + s0 lda); Bm: its first row in the first of the columns.
constexpr int UR_CB = 32;  // columns per column block of the panel update below
constexpr int UR_W = 8;    // columns per workgroup of urows_fused: one per half-wavefront in the TRSM
typedef double ur_d4 __attribute__((ext_vector_type(4)));
template <int CB>
__global__ void __launch_bounds__(256) urows_fused(int G, const double* __restrict__ L, size_t lda,
                                                   double* __restrict__ Bm, int ncols, Abort abort) {
    // behind a zero pivot in panel GA of the super-panel: panels before it are applied in
    // full, the rows of the later ones only receive those panels' GEMM updates
    const int GA = abort_count(abort, G, PW);
    if (GA == 0) return;
    __shared__ double sb[8 * PW][CB + 1];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int cb0 = blockIdx.x * CB;
    const int rows = G * PW;
    for (int e = tid; e < rows * CB; e += 256) {
        const int r = e % rows, c = e / rows;
        sb[r][c] = cb0 + c < ncols ? Bm[(size_t)r + (size_t)(cb0 + c) * lda] : 0.0;
    }
    __syncthreads();
    const int lr = lane & 15, lk = lane >> 4;
    const int wm = (w & 1) * 16, wn = (w >> 1) * 16;
    // a 16-column MFMA tile: with CB < 32 the wavefronts past the block sit the GEMM out and
    // the lanes past it read a clamped column and store nothing
    const bool gemm_wave = wn < CB;
    const int gc = min(wn + lr, CB - 1);
    const bool gcol = wn + lr < CB;
    const int i = lane & 31, sub = lane >> 5;
    for (int q = 0; q < G; q++) {
        const int j0 = q * PW;
        const int qs = min(q, GA);  // earlier panels whose GEMM these rows receive
        ur_d4 crun = {0.0, 0.0, 0.0, 0.0};
        if (qs > 0 && gemm_wave) {
            ur_d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int r = 0; r < 4; r++) crun[r] = sb[j0 + wm + lk + 4 * r][gc];
            const double* Lr = L + (size_t)(j0 + wm + lr);
            // every slab's L entries in flight at once (at most 7 earlier panels)
            double fa[7][PW / 4];
#pragma unroll
            for (int sl = 0; sl < 7; sl++)
#pragma unroll
                for (int t = 0; t < PW / 4; t++)
                    fa[sl][t] = Lr[(size_t)(min(sl, qs - 1) * PW + 4 * t + lk) * lda];  // clamped; slabs >= qs unused
#pragma unroll
            for (int sl = 0; sl < 7; sl++) {  // one 32-deep slab per earlier panel
                if (sl < qs) {
#pragma unroll
                    for (int t = 0; t < PW / 4; t++)
                        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[sl][t], sb[sl * PW + 4 * t + lk][gc], acc, 0,
                                                                   0, 0);
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        crun[r] = fma(-1.0, acc[r], crun[r]);
                        acc[r] = 0.0;
                    }
                }
            }
        }
        if (qs > 0) {
            __syncthreads();  // every wavefront has read the rows it multiplies
            if (gemm_wave && gcol)
#pragma unroll
                for (int r = 0; r < 4; r++) sb[j0 + wm + lk + 4 * r][gc] = crun[r];
            __syncthreads();
        }
        if (q >= GA) continue;  // (uniform over the workgroup)
        // TRSM of rows j0 .. j0+31 with the panel's unit-lower L11: lane i holds row i of
        // L11; each wavefront solves its CB / 4 columns together, CB / 8 per half, whose
        // eliminations are independent and interleave
        double arow[PW];
#pragma unroll
        for (int k = 0; k < PW; k++) {
            const double a = L[(size_t)(j0 + i) + (size_t)(j0 + min(k, i)) * lda];  // clamped: loads in flight together
            arow[k] = keep_or_zero(a, k <= i);
        }
        constexpr int NC = CB / 8;  // columns per lane
        const int cbase = w * (CB / 4) + sub * NC;
        double x[NC];
#pragma unroll
        for (int m2 = 0; m2 < NC; m2++) x[m2] = sb[j0 + i][cbase + m2];
#pragma unroll
        for (int k = 0; k < PW; k++) {
            const int nf = (PW - k - 1) & ~3;
#pragma unroll
            for (int m2 = 0; m2 < NC; m2++) {
                double bk = rdlane_d(x[m2], k);
                const double b1 = rdlane_d(x[m2], 32 + k);
                bk = sub ? b1 : bk;
                const double al = -bk;
                // a zero multiplier (bk == 0, so al == 0) leaves the column as it is; a select,
                // not a branch, since the two half-wavefronts' multipliers differ
                const double y = (i - k - 1 < nf) ? fma(al, arow[k], x[m2]) : x[m2] + al * arow[k];
                x[m2] = (bk != 0.0 && i > k) ? y : x[m2];
            }
        }
#pragma unroll
        for (int m2 = 0; m2 < NC; m2++) sb[j0 + i][cbase + m2] = x[m2];
        __syncthreads();
    }
    for (int e = tid; e < rows * CB; e += 256) {
        const int r = e % rows, c = e / rows;
        if (cb0 + c < ncols) Bm[(size_t)r + (size_t)(cb0 + c) * lda] = sb[r][c];
    }
}

// ---- after a panel inside its super-panel: the TRSM of its U rows and the GEMM of the
// rows below, in one launch
//
// U := L11^-1 U (the panel's 32 rows in the super-panel's nr columns right of it) as
// trsm_ln_lanes<true, 32> computes it, then C -= L21 U for the mb rows below as
// dgemm_mfma computes a 32-deep single-slab product (a zero accumulator, the 8-step
// mfma_f64_16x16x4f64 chain over k = 0, 4, .., then C = fma(-1, acc, C)).  The first
// PSU_S workgroups of column block ct each solve 8 of its 32 columns (a wavefront per two,
// one per half, as the TRSM launch does), store them with device-coherent stores and raise
// their flag; the block's other workgroups each update 64 rows of its columns once all
// PSU_S flags are up.  A block's solvers precede its updaters in dispatch order, so a
// waiting workgroup never holds back the ones it waits for.
//
// The panel's row interchanges (its row map, absolute rows) are applied here too, in place
// of a separate launch before this one: rows below the panel's 32 only ever receive rows
// that started among those 32, and the rows of the 32 receive rows from anywhere.  So each
// solver column gathers its 32 rows through the map and first moves the displaced rows of
// the 32 to their rows below (all read before its solved rows overwrite the 32, and stored
// before its flag rises); the updaters read their rows ahead of the flag and read again,
// after it, the few that received a displaced row.  The first nlwg workgroups (they wait for
// nothing) apply the map to the `left` columns at Aleft, a wavefront a column, and the next
// one turns the panel's pivots (fix, 1-based within the panel) into global ones.
constexpr int PSU_S = UR_CB / 8;  // solver workgroups per column block
__global__ void __launch_bounds__(256) panel_solve_update(int mb, int nr, int nrt, const double* __restrict__ L11,
                                                          double* __restrict__ U, const double* __restrict__ L21,
                                                          double* __restrict__ C, size_t lda,
                                                          int* __restrict__ flags, long spin_limit,
                                                          int32_t* __restrict__ info, int col,
                                                          const RowMap* __restrict__ map, double* __restrict__ Aleft,
                                                          int left, int32_t* __restrict__ fix, int nlwg) {
    if (abort_stop(Abort::panel(info, col))) return;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int cnt = min(map->count, ROWMAP_MAX);
    // loaded beside the count, used only below it
    const int64_t mrow = map->row[tid & (ROWMAP_MAX - 1)], msrc = map->src[tid & (ROWMAP_MAX - 1)];
    if ((int)blockIdx.x < nlwg) {
        const int cc = (int)blockIdx.x * 4 + w;
        if (cc >= left) return;
        double* c = Aleft + (size_t)cc * lda;
        double v0 = 0.0, v1 = 0.0;
        int64_t d0 = -1, d1 = -1;
        if (lane < cnt) { d0 = map->row[lane]; v0 = c[map->src[lane]]; }
        if (lane + 64 < cnt) { d1 = map->row[lane + 64]; v1 = c[map->src[lane + 64]]; }
        if (d0 >= 0) c[d0] = v0;
        if (d1 >= 0) c[d1] = v1;
        return;
    }
    const int bid = (int)blockIdx.x - nlwg;
    if (bid == 0 && tid < PW) fix[tid] += col;
    const int ct = bid / (nrt + PSU_S), rt = bid % (nrt + PSU_S);
    const int c0 = ct * UR_CB;
    const int64_t top = col;  // U's first row (absolute)
    if (rt < PSU_S) {
        __shared__ int64_t tsrc[PW], brow[PW], bsrc[PW];
        __shared__ int nbelow;
        if (tid < PW) tsrc[tid] = top + tid;
        if (tid == 0) nbelow = 0;
        __syncthreads();
        if (tid < cnt) {
            const int64_t r = mrow, sr = msrc;
            if (r < top + PW) {
                tsrc[r - top] = sr;
            } else {
                const int k = atomicAdd(&nbelow, 1);
                if (k < PW) { brow[k] = r; bsrc[k] = sr; }
            }
        }
        __syncthreads();
        const int i = lane & 31, sub = lane >> 5;
        double arow[PW];
#pragma unroll
        for (int k = 0; k < PW; k++) {
            const double a = L11[(size_t)i + (size_t)min(k, i) * lda];  // clamped: loads in flight together
            arow[k] = keep_or_zero(a, k <= i);
        }
        constexpr int NC = 1;
        const int cbase = rt * 8 + w * 2 + sub;  // this lane's column within the block
        const bool cval = c0 + cbase < nr;
        const size_t cof = (size_t)min(c0 + cbase, nr - 1) * lda;
        double x[NC];
        x[0] = keep_or_zero(U[(size_t)(tsrc[i] - top) + cof], cval);
        if (cval && i < min(nbelow, PW)) {
            const double v = U[(size_t)(bsrc[i] - top) + cof];
            __builtin_amdgcn_s_waitcnt(0);  // this column's gathered rows are in before any store
            xst(U + (size_t)(brow[i] - top) + cof, v);
        }
#pragma unroll
        for (int k = 0; k < PW; k++) {
            const int nf = (PW - k - 1) & ~3;
#pragma unroll
            for (int m2 = 0; m2 < NC; m2++) {
                double bk = rdlane_d(x[m2], k);
                const double b1 = rdlane_d(x[m2], 32 + k);
                bk = sub ? b1 : bk;
                const double al = -bk;
                // a zero multiplier leaves the column as it is (a select, as above)
                const double y = (i - k - 1 < nf) ? fma(al, arow[k], x[m2]) : x[m2] + al * arow[k];
                x[m2] = (bk != 0.0 && i > k) ? y : x[m2];
            }
        }
#pragma unroll
        for (int m2 = 0; m2 < NC; m2++)
            if (c0 + cbase + m2 < nr) xst(U + (size_t)i + (size_t)(c0 + cbase + m2) * lda, x[m2]);
        __builtin_amdgcn_s_waitcnt(0);  // the solved rows are stored before the flag rises
        __syncthreads();
        if (tid == 0) __hip_atomic_store(flags + ct * PSU_S + rt, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    // 64 rows of the update: wavefront w owns rows [16 w, 16 w + 16) as two 16 x 16 tiles
    const int r0 = (rt - PSU_S) * 64 + 16 * w;
    const int lr = lane & 15, lk = lane >> 4;
    const int row = min(r0 + lr, mb - 1);
    double fa[PW / 4];
#pragma unroll
    for (int t = 0; t < PW / 4; t++) fa[t] = L21[(size_t)row + (size_t)(4 * t + lk) * lda];
    ur_d4 crun[2];
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
        for (int r = 0; r < 4; r++)
            crun[b][r] = C[(size_t)min(r0 + lk + 4 * r, mb - 1) + (size_t)min(c0 + 16 * b + lr, nr - 1) * lda];
    __shared__ int ok_sh;
    __shared__ unsigned long long rmask;  // which of the workgroup's 64 rows receive a displaced row
    if (tid == 0) {
        ok_sh = 1;
        rmask = 0;
    }
    __syncthreads();
    const int64_t tb = top + PW + (int64_t)(rt - PSU_S) * 64;  // the workgroup's first row (absolute)
    if (tid < cnt) {
        const int64_t r = mrow;
        if (r >= tb && r < tb + 64) atomicOr(&rmask, 1ull << (r - tb));
    }
    if (tid < PSU_S) {  // one thread per solver flag
        long sp = 0;
        while (__hip_atomic_load(flags + ct * PSU_S + tid, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
            if (++sp > spin_limit) {  // this wait timed out: report it unless a status is set
                int zero = 0;
                __hip_atomic_compare_exchange_strong(info, &zero, BOLA_HIP_ERR_DEVICE, __ATOMIC_RELAXED,
                                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok_sh = 0;
                break;
            }
            if (__hip_atomic_load(info, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < 0) {  // another gave up
                ok_sh = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
    }
    __syncthreads();
    if (!ok_sh) return;
    if (const unsigned long long rm = rmask) {
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int r = 0; r < 4; r++)
                if ((rm >> (16 * w + lk + 4 * r)) & 1)
                    crun[b][r] = xld(C + (size_t)min(r0 + lk + 4 * r, mb - 1) +
                                     (size_t)min(c0 + 16 * b + lr, nr - 1) * lda);
    }
#pragma unroll
    for (int b = 0; b < 2; b++) {
        const int col = min(c0 + 16 * b + lr, nr - 1);
        ur_d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int t = 0; t < PW / 4; t++)
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[t], xld(U + (size_t)(4 * t + lk) + (size_t)col * lda), acc,
                                                       0, 0, 0);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int rr = r0 + lk + 4 * r, cc = c0 + 16 * b + lr;
            if (rr < mb && cc < nr) C[(size_t)rr + (size_t)cc * lda] = fma(-1.0, acc[r], crun[b][r]);
        }
    }
}

__global__ void k_zero_i32(int32_t* p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0;
}

// ---- the m <= 32 < n panel (one thread per column; three launches per pivot column)
__global__ void wide_pivot(int m, double* __restrict__ A, size_t lda, int j, int32_t* __restrict__ ipiv,
                           int32_t* __restrict__ info, int col_off, int32_t* __restrict__ jp_out) {
    if (*info != 0) return;
    const int lane = threadIdx.x;  // 64 threads, m - j <= 32
    const int len = m - j;
    const int T = len < 16 ? 0 : (len & ~15);
    Cand vb{-1.0, NOKEY}, tb{-1.0, NOKEY};
    if (lane < len) {
        double v = fabs(A[(size_t)(j + lane) + (size_t)j * lda]);
        if (v == v) {
            if (lane < T) { vb.v = v; vb.k = vec_key((uint64_t)lane); }
            else { tb.v = v; tb.k = (uint64_t)lane; }
        }
    }
    for (int off = 32; off >= 1; off >>= 1) {
        double ov = __shfl_xor(vb.v, off, 64);
        uint64_t ok = (uint64_t)__shfl_xor((long long)vb.k, off, 64);
        if (better(ov, ok, vb.v, vb.k)) { vb.v = ov; vb.k = ok; }
        ov = __shfl_xor(tb.v, off, 64);
        ok = (uint64_t)__shfl_xor((long long)tb.k, off, 64);
        if (better(ov, ok, tb.v, tb.k)) { tb.v = ov; tb.k = ok; }
    }
    if (lane == 0) {
        long rel;
        if (len < 16) rel = tb.v >= 0.0 ? (long)tb.k : 0;
        else {
            double V = 0.0;
            long vi = 0;
            if (vb.v > 0.0) { V = vb.v; vi = (long)(vb.k & KEYMASK); }
            rel = (tb.v > V) ? (long)tb.k : vi;
        }
        const int jp = j + (int)rel;
        ipiv[j] = jp + 1;
        *jp_out = jp;
        if (!(A[(size_t)jp + (size_t)j * lda] != 0.0)) {
            info[1] = j + 1;
            info[0] = col_off + j + 1;
        }
    }
}

__global__ void wide_swap_scale(int m, int n, double* __restrict__ A, size_t lda, int j,
                                const int32_t* __restrict__ jp_in, const int32_t* __restrict__ info) {
    if (*info != 0) return;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int jp = *jp_in;
    double* col = A + (size_t)c * lda;
    if (jp != j) {
        double t = col[j];
        col[j] = col[jp];
        col[jp] = t;
    }
    if (c == j && j < m - 1) {
        const double inv = 1.0 / col[j];
        for (int i = j + 1; i < m; i++) col[i] *= inv;
    }
}

__global__ void wide_ger(int m, int n, double* __restrict__ A, size_t lda, int j, const int32_t* __restrict__ info) {
    if (*info != 0) return;
    const int c = j + 1 + blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int np = n - j - 1, nu = np - (np & 1), rc = c - (j + 1);
    const int mf = (m - j - 1) & ~3;
    const double tc = -1.0 * A[(size_t)j + (size_t)c * lda];
    const bool paired = rc < nu;
    bool skip;
    if (paired) {
        const int pc = (rc & 1) ? c - 1 : c + 1;
        const double tp = -1.0 * A[(size_t)j + (size_t)pc * lda];
        skip = tc == 0.0 && tp == 0.0;
    } else {
        skip = tc == 0.0;
    }
    if (skip) return;
    const double* x = A + (size_t)j * lda;
    double* col = A + (size_t)c * lda;
    for (int i = j + 1; i < m; i++) {
        const bool fr = (i - (j + 1)) < mf;
        col[i] = (paired && fr) ? fma(x[i], tc, col[i]) : col[i] + tc * x[i];
    }
}

__global__ void k_ipiv_add(int32_t* __restrict__ ipiv, int cnt, int off, const int32_t* __restrict__ info) {
    if (*info != 0) return;
    for (int i = threadIdx.x; i < cnt; i += blockDim.x) ipiv[i] += off;
}

}  // namespace

// ---------------------------------------------------------------- host drivers

namespace {

struct LuWs {
    int* flags;         // [npanels][nwg_max]
    int* uflags;        // [npanels][8 * 4]: the fused panel update's solver flags
    PanelSlot* slots;   // [2][nwg_max][2]
    PanelSlot* cur;     // [2]
    RowMap* maps;       // [npanels]
    int32_t* jp;        // wide-panel pivot scratch
    int nwg_max;
    int npanels;
};

int panel_rows_per_thread(int m) { return m <= PT ? 1 : 2; }
int panel_nwg(int m) {
    int r = panel_rows_per_thread(m);
    return (m + PT * r - 1) / (PT * r);
}
// the wavefront panel: rows per lane and wavefronts; taller panels use the workgroup kernel
constexpr int kWaveMax = 512;
int wave_rows_per_lane(int m) { return m <= 8192 ? 1 : 2; }
int wave_count(int m) {
    int r = wave_rows_per_lane(m);
    return (m + WV * r - 1) / (WV * r);
}
int panel_slots(int m) { return std::max(panel_nwg(m), wave_count(m)); }

// The factorization's scratch: bytes, and the carve-up into `base` (which must hold them);
// everything carved here is zeroed on the handle's stream by a kernel.
size_t lu_ws_layout(int m, int npanels, LuWs* w, char* base) {
    LuWs l;
    l.nwg_max = std::max(1, panel_slots(m));
    l.npanels = std::max(1, npanels);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    };
    // the slots start zeroed: the wavefront panel's epochs are unique only within one
    // factorization
    const size_t slot_bytes = std::max(sizeof(PanelSlot), sizeof(uint64_t) * TW);
    size_t o_flags = take(sizeof(int) * (size_t)l.npanels * l.nwg_max);
    size_t o_uflags = take(sizeof(int) * (size_t)l.npanels * 32);
    size_t o_maps = take(sizeof(RowMap) * (size_t)l.npanels);
    size_t o_jp = take(sizeof(int32_t) * 4);
    size_t o_slots = take(slot_bytes * 4 * (size_t)l.nwg_max);
    size_t o_cur = take(slot_bytes * 2);
    if (w) {
        l.flags = (int*)(base + o_flags);
        l.uflags = (int*)(base + o_uflags);
        l.maps = (RowMap*)(base + o_maps);
        l.jp = (int32_t*)(base + o_jp);
        l.slots = (PanelSlot*)(base + o_slots);
        l.cur = (PanelSlot*)(base + o_cur);
        *w = l;
    }
    return off;
}

// fixed: caller-provided scratch (the no-allocation entry point), else the handle's
LuWs lu_workspace(bola_hip_handle_s* h, int m, int npanels, const FixedScratch* fixed) {
    LuWs w;
    const size_t bytes = lu_ws_layout(m, npanels, nullptr, nullptr);
    char* base = fixed ? (char*)fixed->ptr : (char*)h->ws.get(bytes);
    if (fixed && bytes > fixed->bytes) throw HipFailure{hipErrorInvalidValue, "LU scratch too small", __LINE__};
    lu_ws_layout(m, npanels, &w, base);
    size_t nzero = bytes / 4;
    hipLaunchKernelGGL(k_zero_i32, dim3(std::min<size_t>((nzero + 255) / 256, 1024)), dim3(256), 0, h->stream,
                       (int32_t*)base, nzero);
    BOLA_LAUNCH_CHECK();
    return w;
}

const long kSpinLimit = 1L << 22;

// iterative_getrf2 (getrf2.rs:131-185) on an m x n block, n <= 32 or m <= 32.
void getf2(hipStream_t st, int m, int n, double* a, size_t lda, int32_t* ipiv, int32_t* info, int col_off,
           long row_off, const LuWs& w, int panel_idx) {
    if (n <= PW && wave_count(m) <= kWaveMax) {
        const int r = wave_rows_per_lane(m);
        const int nwv = wave_count(m);
        uint64_t* wslots = reinterpret_cast<uint64_t*>(w.slots);
        uint64_t* wcur = reinterpret_cast<uint64_t*>(w.cur);
        RowMap* map = w.maps + panel_idx;
        if (nwv == 1) {
            if (r == 1)
                hipLaunchKernelGGL((getf2_wave<1, false>), dim3(1), dim3(WV), 0, st, m, n, a, lda, ipiv, info, col_off,
                                   row_off, wslots, wcur, map, nwv, kSpinLimit);
            else
                hipLaunchKernelGGL((getf2_wave<2, false>), dim3(1), dim3(WV), 0, st, m, n, a, lda, ipiv, info, col_off,
                                   row_off, wslots, wcur, map, nwv, kSpinLimit);
        } else if (nwv <= kDuoMax) {
            if (r == 1)
                hipLaunchKernelGGL((getf2_duo<1>), dim3(nwv), dim3(2 * WV), 0, st, m, n, a, lda, ipiv, info, col_off,
                                   row_off, wslots, wcur, map, nwv, kSpinLimit);
            else
                hipLaunchKernelGGL((getf2_duo<2>), dim3(nwv), dim3(2 * WV), 0, st, m, n, a, lda, ipiv, info, col_off,
                                   row_off, wslots, wcur, map, nwv, kSpinLimit);
        } else {
            if (r == 1)
                hipLaunchKernelGGL((getf2_wave<1, true>), dim3(nwv), dim3(WV), 0, st, m, n, a, lda, ipiv, info,
                                   col_off, row_off, wslots, wcur, map, nwv, kSpinLimit);
            else
                hipLaunchKernelGGL((getf2_wave<2, true>), dim3(nwv), dim3(WV), 0, st, m, n, a, lda, ipiv, info,
                                   col_off, row_off, wslots, wcur, map, nwv, kSpinLimit);
        }
        BOLA_LAUNCH_CHECK();
        return;
    }
    if (n <= PW) {
        const int r = panel_rows_per_thread(m);
        const int nwg = panel_nwg(m);
        int* flags = w.flags + (size_t)panel_idx * w.nwg_max;
        RowMap* map = w.maps + panel_idx;
        if (nwg == 1) {
            if (r == 1)
                hipLaunchKernelGGL((getf2_panel<1, false>), dim3(1), dim3(PT), 0, st, m, n, a, lda, ipiv, info, col_off,
                                   row_off, flags, w.slots, w.cur, map, nwg, kSpinLimit);
            else
                hipLaunchKernelGGL((getf2_panel<2, false>), dim3(1), dim3(PT), 0, st, m, n, a, lda, ipiv, info, col_off,
                                   row_off, flags, w.slots, w.cur, map, nwg, kSpinLimit);
        } else {
            if (r == 1)
                hipLaunchKernelGGL((getf2_panel<1, true>), dim3(nwg), dim3(PT), 0, st, m, n, a, lda, ipiv, info,
                                   col_off, row_off, flags, w.slots, w.cur, map, nwg, kSpinLimit);
            else
                hipLaunchKernelGGL((getf2_panel<2, true>), dim3(nwg), dim3(PT), 0, st, m, n, a, lda, ipiv, info,
                                   col_off, row_off, flags, w.slots, w.cur, map, nwg, kSpinLimit);
        }
        BOLA_LAUNCH_CHECK();
        return;
    }
    // m <= 32 < n
    const int min_mn = std::min(m, n);
    for (int j = 0; j < min_mn; j++) {
        hipLaunchKernelGGL(wide_pivot, dim3(1), dim3(64), 0, st, m, a, lda, j, ipiv, info, col_off, w.jp);
        BOLA_LAUNCH_CHECK();
        hipLaunchKernelGGL(wide_swap_scale, dim3(cdiv(n, 256)), dim3(256), 0, st, m, n, a, lda, j, w.jp, info);
        BOLA_LAUNCH_CHECK();
        if (j < min_mn - 1 && j + 1 < n) {
            hipLaunchKernelGGL(wide_ger, dim3(cdiv(n - j - 1, 256)), dim3(256), 0, st, m, n, a, lda, j, info);
            BOLA_LAUNCH_CHECK();
        }
    }
}

// getrf2, getrf2.rs:39-97 (recursive above 32 in both dimensions)
void getrf2_rec(bola_hip_handle_s* h, int m, int n, double* a, size_t lda, int32_t* ipiv, int32_t* info, int col_off,
                const LuWs& w, void* laswp_ws) {
    if (m == 0 || n == 0) return;
    if (m <= PW || n <= PW) {
        // fresh flags / row map for this leaf panel
        hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(256), 0, h->stream, (int32_t*)w.flags, (size_t)w.nwg_max);
        BOLA_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, h->stream, (int32_t*)w.maps, (size_t)1);
        BOLA_LAUNCH_CHECK();
        getf2(h->stream, m, n, a, lda, ipiv, info, col_off, 0, w, 0);
        // the panel kernel's own row map is not needed here: the recursion applies
        // interchanges through laswp with the pivots, as the reference does
        return;
    }
    hipStream_t st = h->stream;
    const int n1 = std::min(m, n) / 2, n2 = n - n1;
    getrf2_rec(h, m, n1, a, lda, ipiv, info, col_off, w, laswp_ws);
    launch_laswp(st, (size_t)n2, a + (size_t)n1 * lda, lda, 1, n1, ipiv, 1, laswp_ws, info);
    launch_dtrsm(st, 'L', 'L', 'N', 'U', (size_t)n1, (size_t)n2, 1.0, a, lda, a + (size_t)n1 * lda, lda, info);
    if (n1 < m)
        launch_dgemm(st, 'N', 'N', (size_t)(m - n1), (size_t)n2, (size_t)n1, -1.0, a + n1, lda, a + (size_t)n1 * lda,
                     lda, 1.0, a + n1 + (size_t)n1 * lda, lda, info);
    if (n1 < m) getrf2_rec(h, m - n1, n2, a + n1 + (size_t)n1 * lda, lda, ipiv + n1, info, col_off + n1, w, laswp_ws);
    const int mn = std::min(m, n);
    hipLaunchKernelGGL(k_ipiv_add, dim3(1), dim3(256), 0, st, ipiv + n1, mn - n1, n1, info);
    BOLA_LAUNCH_CHECK();
    launch_laswp(st, (size_t)n1, a, lda, n1 + 1, mn, ipiv, 1, laswp_ws, info);
}

}  // namespace

namespace {

// Row/column where the reference's trailing GEMM of the panel at column j (width jb)
// switches to its unfused edge path: rows >= m - (m-j-jb)%8, columns >= n - (n-j-jb)%4.
struct Edge {
    long row, col;
};
inline Edge panel_edge(int m, int n, int j, int jb) {
    return Edge{(long)m - (long)((m - j - jb) % 8), (long)n - (long)((n - j - jb) % 4)};
}

// Update of the columns [c0, c1) right of the super-panel [s0, s0+sw) whose panels are
// factored: the panels' interchanges, then for each panel its U rows (GEMM from the
// earlier panels of the super-panel, TRSM with its L11), then one GEMM for all rows
// below the super-panel with one 32-deep fold per panel.  Every element receives the
// reference's operations in the reference's order (getrf.rs:53-98).
// ncu, fast: the CUs the stream may use (0: all) and whether its GEMMs have them to
// themselves (two workgroups per CU)
void update_cols(bola_hip_handle_s* h, hipStream_t st, int m, int n, double* a, size_t lda, int32_t* info,
                 const LuWs& w, int s0, int sw, int c0, int c1, int ncu = 0, bool fast = false) {
    if (c1 <= c0) return;
    const int ncols = c1 - c0;
    const int p0 = s0 / PW, G = (sw + PW - 1) / PW;
    double* B = a + (size_t)c0 * lda;
    {
        Phase pl(h, "LASWP", 0.0, st);
        launch_rowmap_apply(st, B, lda, (size_t)ncols, 0, 0, w.maps + p0, G, Abort::span(info, s0), nullptr, 0, 0);
    }
    const bool uniform = (sw % PW) == 0;
    if (uniform) {
        const Edge e = panel_edge(m, n, s0, PW);  // the same for every full-width panel
        if (e.row >= s0 + sw && e.col >= c1 && G <= 8) {
            // no edge elements among these rows and columns: every panel's GEMM and TRSM in
            // one launch, a workgroup per 8 columns
            Phase pt(h, "TRSM", (double)sw * sw * ncols, st);
            hipLaunchKernelGGL(urows_fused<UR_W>, dim3((unsigned)((ncols + UR_W - 1) / UR_W)), dim3(256), 0, st, G,
                               a + s0 + (size_t)s0 * lda, lda, a + s0 + (size_t)c0 * lda, ncols, Abort::span(info, s0));
            BOLA_LAUNCH_CHECK();
        } else
        for (int q = 0; q < G; q++) {
            const int j = s0 + q * PW;
            if (q > 0) {
                Phase pg(h, "GEMM", 2.0 * PW * ncols * (q * PW), st);
                launch_dgemm_ex(st, 'N', 'N', PW, ncols, (size_t)q * PW, -1.0, a + j + (size_t)s0 * lda, lda,
                                a + s0 + (size_t)c0 * lda, lda, 1.0, a + j + (size_t)c0 * lda, lda, Abort::span(info, s0), PW, 0,
                                e.row - j, e.col - c0, fast, ncu);
            }
            Phase pt(h, "TRSM", (double)PW * PW * ncols, st);
            launch_dtrsm(st, 'L', 'L', 'N', 'U', PW, ncols, 1.0, a + j + (size_t)j * lda, lda, a + j + (size_t)c0 * lda,
                         lda, Abort::panel(info, j));
        }
        if (s0 + sw < m) {
            Phase pg(h, "GEMM", 2.0 * (double)(m - s0 - sw) * ncols * sw, st);
            launch_dgemm_ex(st, 'N', 'N', (size_t)(m - s0 - sw), ncols, (size_t)sw, -1.0, a + (s0 + sw) + (size_t)s0 * lda,
                            lda, a + s0 + (size_t)c0 * lda, lda, 1.0, a + (s0 + sw) + (size_t)c0 * lda, lda, Abort::span(info, s0), PW, 0,
                            e.row - (s0 + sw), e.col - c0, fast, ncu);
        }
    } else {
        for (int q = 0; q < G; q++) {
            const int j = s0 + q * PW;
            const int jb = std::min(PW, s0 + sw - j);
            {
                Phase pt(h, "TRSM", (double)jb * jb * ncols, st);
                launch_dtrsm(st, 'L', 'L', 'N', 'U', jb, ncols, 1.0, a + j + (size_t)j * lda, lda,
                             a + j + (size_t)c0 * lda, lda, Abort::panel(info, j));
            }
            if (j + jb < m) {
                const Edge e = panel_edge(m, n, j, jb);
                Phase pg(h, "GEMM", 2.0 * (double)(m - j - jb) * ncols * jb, st);
                launch_dgemm_ex(st, 'N', 'N', (size_t)(m - j - jb), ncols, (size_t)jb, -1.0,
                                a + (j + jb) + (size_t)j * lda, lda, a + j + (size_t)c0 * lda, lda, 1.0,
                                a + (j + jb) + (size_t)c0 * lda, lda, Abort::panel(info, j), PW, 0, e.row - (j + jb), e.col - c0,
                                fast, ncu);
            }
        }
    }
}

// The columns [s0, s0+sw) themselves, panel by panel: the reference's loop restricted
// to these columns (their left/right interchanges, TRSM and GEMM stay inside them).
// after_first (may be null): an event the columns right of the first panel wait for; the
// first panel itself runs before it.
void factor_super(bola_hip_handle_s* h, hipStream_t st, int m, int n, double* a, size_t lda, int32_t* ipiv,
                  int32_t* info, const LuWs& w, int s0, int sw, hipEvent_t after_first) {
    const int G = (sw + PW - 1) / PW;
    for (int q = 0; q < G; q++) {
        const int j = s0 + q * PW;
        const int jb = std::min(PW, s0 + sw - j);
        const int p = j / PW;
        {
            Phase pp(h, "GETRF2_IT", 0.0, st);
            getf2(st, m - j, jb, a + j + (size_t)j * lda, lda, ipiv + j, info, j, j, w, p);
        }
        if (q == 0 && after_first) BOLA_CHECK(hipStreamWaitEvent(st, after_first, 0));
        const int nr = s0 + sw - j - jb;
        const Edge ef = panel_edge(m, n, j, jb);
        const bool fused = nr > 0 && jb == PW && j + jb < m && nr <= 8 * UR_CB && ef.row >= m && ef.col >= j + jb + nr;
        if (!fused) {
            Phase pl(h, "LASWP", 0.0, st);
            launch_rowmap_apply(st, a + (size_t)s0 * lda, lda, (size_t)(j - s0), (size_t)(j + jb - s0),
                                (size_t)(s0 + sw - j - jb), w.maps + p, 1, Abort::panel(info, j), ipiv + j, jb, j);
        }
        if (fused) {
            // no edge element among the rows below and these columns: the interchanges, the
            // TRSM and the GEMM in one launch
            const int mb = m - j - jb, nrt = (mb + 63) / 64, ncb = (nr + UR_CB - 1) / UR_CB;
            const int left = j - s0, nlwg = (left + 3) / 4;
            Phase pt(h, "TRSM", (double)jb * jb * nr + 2.0 * (double)mb * nr * jb, st);
            hipLaunchKernelGGL(panel_solve_update, dim3((unsigned)(ncb * (nrt + PSU_S) + nlwg)), dim3(256), 0, st, mb, nr,
                               nrt, a + j + (size_t)j * lda, a + j + (size_t)(j + jb) * lda,
                               a + (j + jb) + (size_t)j * lda, a + (j + jb) + (size_t)(j + jb) * lda, lda,
                               w.uflags + (size_t)p * 32, kSpinLimit, info, j, w.maps + p, a + (size_t)s0 * lda, left,
                               ipiv + j, nlwg);
            BOLA_LAUNCH_CHECK();
        } else if (nr > 0) {
            {
                Phase pt(h, "TRSM", (double)jb * jb * nr, st);
                launch_dtrsm(st, 'L', 'L', 'N', 'U', jb, nr, 1.0, a + j + (size_t)j * lda, lda,
                             a + j + (size_t)(j + jb) * lda, lda, Abort::panel(info, j));
            }
            if (j + jb < m) {
                const Edge e = panel_edge(m, n, j, jb);
                Phase pg(h, "GEMM", 2.0 * (double)(m - j - jb) * nr * jb, st);
                launch_dgemm_ex(st, 'N', 'N', (size_t)(m - j - jb), (size_t)nr, (size_t)jb, -1.0,
                                a + (j + jb) + (size_t)j * lda, lda, a + j + (size_t)(j + jb) * lda, lda, 1.0,
                                a + (j + jb) + (size_t)(j + jb) * lda, lda, Abort::panel(info, j), PW, 0, e.row - (j + jb),
                                e.col - (j + jb));
            }
        }
    }
}

hipEvent_t lu_event(bola_hip_handle_s* h, int i) {
    while ((int)h->sync_events.size() <= i) {
        hipEvent_t e;
        BOLA_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        h->sync_events.push_back(e);
    }
    return h->sync_events[i];
}

}  // namespace

// Status words: info[0] = 0 or global 1-based failing column (or BOLA_HIP_ERR_DEVICE);
// info[1] = the column number the reference prints (local to its panel).
//
// Schedule (m a multiple of 8): super-panels of W columns.  Deferring a super-panel's
// interchanges on the columns to its right past the updates of its earlier panels is
// exact only while no row takes the reference's unfused GEMM edge path, i.e. m % 8 == 0;
// otherwise the reference's panel-by-panel order is kept.  Stream 1 factors a
// super-panel and then brings
// the next super-panel's columns up to date (look-ahead); stream 2 meanwhile applies the
// super-panel to all columns further right (the bulk GEMM), so the next factorization
// overlaps the bulk update.  Work on the two streams never touches the same columns at
// the same time; events order the hand-offs.
size_t lu_scratch_bytes(size_t m_, size_t n_, bool recursive_entry) {
    const int m = (int)m_, n = (int)n_;
    const int min_mn = std::min(m, n);
    if (m == 0 || n == 0) return 0;
    if (recursive_entry || PW >= min_mn)
        return lu_ws_layout(m, 1, nullptr, nullptr) + ((laswp_ws_bytes((size_t)min_mn) + 255) & ~(size_t)255);
    return lu_ws_layout(m, (min_mn + PW - 1) / PW, nullptr, nullptr);
}

// A stream on the CUs i with ((i / 8) % k == 0) == crit.  Selecting CUs in runs of 8 with
// k <= 4 gives every XCD some of each set whether CU numbers run XCD by XCD (32 each) or
// alternate between the 8 XCDs; a set that left an XCD without CUs could strand the
// workgroups dispatched to it.
hipStream_t lu_masked_stream(int ncu, int k, bool crit) {
    uint32_t mask[64] = {0};
    const int n = std::min(ncu, 64 * 32);
    for (int i = 0; i < n; i++)
        if ((((i / 8) % k) == 0) == crit) mask[i / 32] |= 1u << (i % 32);
    hipStream_t s = nullptr;
    if (hipExtStreamCreateWithCUMask(&s, (uint32_t)((n + 31) / 32), mask) != hipSuccess) return nullptr;
    return s;
}

// The LU's critical path on a quarter of the CUs, its bulk update and left interchanges
// on the other three quarters: sharing CUs with the bulk GEMM slowed the latency-bound
// critical kernels by 30-270 % (n = 8192: 39.2 -> 37.6 ms per DGETRF + DGETRS).  From
// about 12K on, the bulk GEMM on three quarters of the CUs becomes the longer path (n =
// 16384: 136 -> 144 ms), so larger matrices keep undivided streams.  Only for a handle
// that has the device to itself.
constexpr int kLuCuSplit = 4;
constexpr int kLuCuSplitMax = 12288;  // largest m and n that take the split

bool lu_split_streams(bola_hip_handle_s* h) {
    if (!h->exclusive_device) return false;
    if (h->lu_split < 0) {
        h->lu_split = kLuCuSplit;
        h->lu_crit = lu_masked_stream(h->num_cus, h->lu_split, true);
        h->lu_bulk = lu_masked_stream(h->num_cus, h->lu_split, false);
        h->lu_left = lu_masked_stream(h->num_cus, h->lu_split, false);
        if (!h->lu_crit || !h->lu_bulk || !h->lu_left) h->lu_split = 0;
    }
    return h->lu_split > 0;
}

void lu_factor(bola_hip_handle_s* h, size_t m_, size_t n_, double* a, size_t lda, int32_t* ipiv, int32_t* info,
               bool recursive_entry, const FixedScratch* fixed, LuHooks* hooks) {
    hipStream_t st = h->stream;
    const int m = (int)m_, n = (int)n_;
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, st, info, (size_t)2);
    BOLA_LAUNCH_CHECK();
    if (m == 0 || n == 0) return;
    const int min_mn = std::min(m, n);
    if (recursive_entry || PW >= min_mn) {
        Phase ph(h, recursive_entry ? "GETRF2" : "GETRF");
        const size_t lbytes = (laswp_ws_bytes((size_t)min_mn) + 255) & ~(size_t)255;
        void* lws = fixed ? fixed->ptr : h->ws_aux.get(lbytes);
        FixedScratch rest;  // the LU scratch after the LASWP group maps
        if (fixed) {
            if (lbytes > fixed->bytes) throw HipFailure{hipErrorInvalidValue, "LU scratch too small", __LINE__};
            rest.ptr = (char*)fixed->ptr + lbytes;
            rest.bytes = fixed->bytes - lbytes;
        }
        LuWs w = lu_workspace(h, m, 1, fixed ? &rest : nullptr);
        if (hooks) hooks->need_cols(st, n);
        getrf2_rec(h, m, n, a, lda, ipiv, info, 0, w, lws);
        if (hooks) hooks->rows_done(0, m, &st, 1);
        return;
    }
    Phase ph(h, "GETRF", 2.0 * (double)min_mn * min_mn * (3.0 * std::max(m, n) - min_mn) / 6.0);
    const int npanels = (min_mn + PW - 1) / PW;
    LuWs w = lu_workspace(h, m, npanels, fixed);
    if (m % 8 != 0) {
        if (hooks) hooks->need_cols(st, n);
        // The reference's GEMM computes the last (m-j-32) % 8 rows of each trailing update
        // unfused; later interchanges move such rows, so interchanges cannot be deferred
        // past an update.  Keep the reference's own order: panel, interchanges, TRSM, GEMM.
        for (int p = 0; p < npanels; p++) {
            const int j = p * PW;
            const int jb = std::min(PW, min_mn - j);
            {
                Phase pp(h, "GETRF2_IT", 0.0, st);
                getf2(st, m - j, jb, a + j + (size_t)j * lda, lda, ipiv + j, info, j, j, w, p);
            }
            {
                Phase pl(h, "LASWP", 0.0, st);
                const size_t right = (j + jb < n) ? (size_t)(n - j - jb) : 0;
                launch_rowmap_apply(st, a, lda, (size_t)j, (size_t)(j + jb), right, w.maps + p, 1, Abort::panel(info, j), ipiv + j, jb,
                                    j);
            }
            if (j + jb < n) {
                {
                    Phase pt(h, "TRSM", (double)jb * jb * (n - j - jb), st);
                    launch_dtrsm(st, 'L', 'L', 'N', 'U', (size_t)jb, (size_t)(n - j - jb), 1.0,
                                 a + j + (size_t)j * lda, lda, a + j + (size_t)(j + jb) * lda, lda, Abort::panel(info, j));
                }
                if (j + jb < m) {
                    Phase pg(h, "GEMM", 2.0 * (double)(m - j - jb) * (n - j - jb) * jb, st);
                    launch_dgemm(st, 'N', 'N', (size_t)(m - j - jb), (size_t)(n - j - jb), (size_t)jb, -1.0,
                                 a + (j + jb) + (size_t)j * lda, lda, a + j + (size_t)(j + jb) * lda, lda, 1.0,
                                 a + (j + jb) + (size_t)(j + jb) * lda, lda, Abort::panel(info, j));
                }
            }
        }
        if (hooks) hooks->rows_done(0, m, &st, 1);
        return;
    }
    // Three streams.  Stream 1 (critical): the super-panel's panels, then the next
    // super-panel's columns.  Stream 2 (bulk): every column after the next super-panel.
    // Stream 3: the super-panel's interchanges on the columns to its left, which nothing
    // after it reads until the end.  Each column still receives the same updates in the
    // same order; the events only order the streams that touch the same columns.
    const int W = min_mn >= 2048 ? 256 : 128;
    hipStream_t s2 = h->aux_stream();
    hipStream_t s3 = h->aux_stream3();
    int ev = 0;
    hipStream_t user = st;
    const bool split = m <= kLuCuSplitMax && n <= kLuCuSplitMax && lu_split_streams(h);
    // CUs of the critical stream's partition (runs of 8, one in kLuCuSplit)
    const int crit_cus = split ? 8 * ((h->num_cus / 8 + kLuCuSplit - 1) / kLuCuSplit) : 0;
    if (split) {
        // the critical path on its own CUs, the bulk update and left interchanges on the rest
        hipEvent_t e_in = lu_event(h, ev++);
        BOLA_CHECK(hipEventRecord(e_in, user));
        st = h->lu_crit;
        s2 = h->lu_bulk;
        s3 = h->lu_left;
        BOLA_CHECK(hipStreamWaitEvent(st, e_in, 0));
    }
    hipEvent_t bulk_done = nullptr;  // the previous super-panel's bulk update (stream 2)
    hipEvent_t left_done = nullptr;  // the last left interchanges (stream 3)
    for (int s0 = 0; s0 < min_mn; s0 += W) {
        const int sw = std::min(W, min_mn - s0);
        if (hooks) hooks->need_cols(st, s0 + sw);
        factor_super(h, st, m, n, a, lda, ipiv, info, w, s0, sw, nullptr);
        hipEvent_t fact_done = lu_event(h, ev++);
        BOLA_CHECK(hipEventRecord(fact_done, st));
        if (s0 > 0) {
            // the rows it permutes include the previous super-panel, which the previous
            // bulk update reads
            BOLA_CHECK(hipStreamWaitEvent(s3, fact_done, 0));
            if (bulk_done) BOLA_CHECK(hipStreamWaitEvent(s3, bulk_done, 0));
            Phase pl(h, "LASWP", 0.0, s3);
            launch_rowmap_apply(s3, a, lda, (size_t)s0, 0, 0, w.maps + s0 / PW, (sw + PW - 1) / PW, Abort::span(info, s0), nullptr, 0, 0);
            left_done = lu_event(h, ev++);
            BOLA_CHECK(hipEventRecord(left_done, s3));
        }
        // from here on stream 1 touches columns the previous bulk update wrote
        if (bulk_done) BOLA_CHECK(hipStreamWaitEvent(st, bulk_done, 0));
        const int c_lo = s0 + sw;
        const int c_mid = std::min(n, c_lo + W);
        if (hooks) hooks->need_cols(st, c_mid);
        update_cols(h, st, m, n, a, lda, info, w, s0, sw, c_lo, c_mid, crit_cus, split);
        if (c_mid < n) {
            BOLA_CHECK(hipStreamWaitEvent(s2, fact_done, 0));
            if (hooks) hooks->need_cols(s2, n);
            update_cols(h, s2, m, n, a, lda, info, w, s0, sw, c_mid, n, split ? h->num_cus - crit_cus : 0, split);
            bulk_done = lu_event(h, ev++);
            BOLA_CHECK(hipEventRecord(bulk_done, s2));
        } else {
            bulk_done = nullptr;
        }
        if (hooks) {
            // rows [s0, s0 + sw) are final once this super-panel's work on every stream is
            // done: no later interchange or update touches them (the last strip: to m)
            hipStream_t after[3] = {st, s2, s3};
            const bool last = s0 + W >= min_mn;
            hooks->rows_done(s0, last ? m : s0 + sw, after, 3);
        }
    }
    if (bulk_done) BOLA_CHECK(hipStreamWaitEvent(st, bulk_done, 0));
    if (left_done) BOLA_CHECK(hipStreamWaitEvent(st, left_done, 0));
    if (st != user) {
        hipEvent_t e_out = lu_event(h, ev++);
        BOLA_CHECK(hipEventRecord(e_out, st));
        BOLA_CHECK(hipStreamWaitEvent(user, e_out, 0));
    }
}

// ---- scratch of the no-allocation LU entry point
//
// A file-scope device array: every GPU the library runs on holds its own copy for the
// life of the process, and every call on that GPU shares it (calls that overlap on one
// GPU race on it).  In host code `scratch` names a host-side stand-in, not device memory:
// the device address comes from hipGetSymbolAddress, once per device, with that device
// current, and only that address is handed to kernels.
constexpr size_t kNoAllocScratchBytes = (size_t)4 << 20;
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};
}  // namespace

void* noalloc_scratch(bola_hip_handle_s* h, size_t* bytes) {
    int dev = -1;
    BOLA_CHECK(hipGetDevice(&dev));
    if (dev != h->device) throw HipFailure{hipErrorInvalidDevice, "the handle's device is not current", __LINE__};
    if (dev < 0 || dev >= kMaxDevices) throw HipFailure{hipErrorInvalidDevice, "device index", __LINE__};
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        BOLA_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_dev[dev] = p;
    }
    *bytes = kNoAllocScratchBytes;
    return g_scratch_dev[dev];
}

}  // namespace bola
