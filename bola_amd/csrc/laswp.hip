// LASWP on gfx950 (laswp.rs:22-132): row interchanges applied to n columns.
//
// The pivot sequence is cut into groups of up to 64 interchanges.  A prologue kernel
// (one wavefront per group) composes each group's transpositions into a compact row map
// {row r_s receives the old value of row src_s}, at most 128 entries.  The apply kernel
// then walks the groups in order for its own columns: lanes span the map entries, so the
// rows k1..k2 (contiguous in a column) are read with coalesced loads, every load of a
// group completes before its stores, and a column is only ever touched by one
// wavefront.  Pure moves: the result is bit-identical to the sequential swaps.
#include "common.hpp"

namespace bola {

namespace {

constexpr int GROUP = 64;  // interchanges per group; a group touches at most 2*GROUP rows
static_assert(2 * GROUP <= ROWMAP_MAX, "row map too small");
using GroupMap = RowMap;

// One wavefront per group: slots live in registers, two per lane (slot = lane, lane+64).
__global__ void __launch_bounds__(64) laswp_compose(int32_t k1, int32_t k2, const int32_t* __restrict__ ipiv,
                                                    int32_t incx, int ngroups, GroupMap* __restrict__ maps) {
    int g = blockIdx.x;
    if (g >= ngroups) return;
    int lane = threadIdx.x;
    const int npiv = k2 - k1 + 1;
    const int first = g * GROUP;
    const int last = min(npiv, first + GROUP);
    // (position, source) for slots lane and lane+64; -1 = empty
    int64_t r0 = -1, s0 = -1, r1 = -1, s1 = -1;
    int count = 0;
    const int64_t absinc = incx < 0 ? -(int64_t)incx : (int64_t)incx;
    for (int t = first; t < last; t++) {
        // the t-th interchange in application order
        int64_t i, ix;
        if (incx > 0) {
            i = (int64_t)k1 + t;
            ix = (int64_t)(k1 - 1) + (int64_t)t * absinc;
        } else {
            i = (int64_t)k2 - t;
            ix = (int64_t)(k1 - 1) + (int64_t)(npiv - 1 - t) * absinc;
        }
        int64_t p = (int64_t)ipiv[ix];
        if (p == i) continue;
        int64_t ri = i - 1, rp = p - 1;
        // find or append slots for ri and rp
        int64_t want[2] = {ri, rp};
        int slot[2];
        for (int q = 0; q < 2; q++) {
            unsigned long long b0 = __ballot(r0 == want[q]);
            unsigned long long b1 = __ballot(r1 == want[q]);
            if (b0) slot[q] = __ffsll(b0) - 1;
            else if (b1) slot[q] = 64 + __ffsll(b1) - 1;
            else {
                int s = count++;
                if (s < 64) { if (lane == s) { r0 = want[q]; s0 = want[q]; } }
                else { if (lane == s - 64) { r1 = want[q]; s1 = want[q]; } }
                slot[q] = s;
            }
        }
        // exchange the sources held in the two slots
        int a = slot[0], b = slot[1];
        int64_t sa = (a < 64) ? __shfl(s0, a, 64) : __shfl(s1, a - 64, 64);
        int64_t sb = (b < 64) ? __shfl(s0, b, 64) : __shfl(s1, b - 64, 64);
        if (a < 64) { if (lane == a) s0 = sb; } else { if (lane == a - 64) s1 = sb; }
        if (b < 64) { if (lane == b) s0 = sa; } else { if (lane == b - 64) s1 = sa; }
    }
    GroupMap* M = maps + g;
    if (lane == 0) M->count = count;
    if (lane < count) { M->row[lane] = r0; M->src[lane] = s0; }
    if (lane + 64 < count) { M->row[lane + 64] = r1; M->src[lane + 64] = s1; }
}

// Each wavefront owns whole columns; per group: load all mapped sources, then store.
__global__ void __launch_bounds__(256) laswp_apply(size_t left, size_t skip_to, size_t right, double* __restrict__ a,
                                                   size_t lda, int ngroups, const GroupMap* __restrict__ maps,
                                                   Abort abort, int32_t* __restrict__ fix_ipiv,
                                                   int fix_cnt, int fix_off) {
    // behind a failed pivot: only the maps of panels before it (one map per panel)
    ngroups = abort_count(abort, ngroups, 32);
    if (ngroups == 0) return;
    if (fix_ipiv && blockIdx.x == 0 && (int)threadIdx.x < fix_cnt) fix_ipiv[threadIdx.x] += fix_off;
    const int lane = threadIdx.x & 63;
    const size_t wave = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const size_t nwaves = ((size_t)gridDim.x * blockDim.x) >> 6;
    const size_t n = left + right;
    for (size_t cc = wave; cc < n; cc += nwaves) {
        const size_t col = cc < left ? cc : skip_to + (cc - left);
        double* c = a + col * lda;
        for (int g = 0; g < ngroups; g++) {
            const GroupMap* M = maps + g;
            int cnt = M->count;
            double v0 = 0.0, v1 = 0.0;
            int64_t d0 = -1, d1 = -1;
            if (lane < cnt) { d0 = M->row[lane]; v0 = c[M->src[lane]]; }
            if (lane + 64 < cnt) { d1 = M->row[lane + 64]; v1 = c[M->src[lane + 64]]; }
            if (d0 >= 0) c[d0] = v0;
            if (d1 >= 0) c[d1] = v1;
            // the next group's loads (other lanes, same wavefront) must see these stores
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
        }
    }
}

}  // namespace

size_t laswp_ws_bytes(size_t npiv) {
    size_t ng = (npiv + GROUP - 1) / GROUP;
    return sizeof(GroupMap) * (ng ? ng : 1);
}

void launch_rowmap_apply(hipStream_t st, double* a, size_t lda, size_t left, size_t skip_to, size_t right,
                         const RowMap* maps, int nmaps, Abort abort, int32_t* fix_ipiv, int fix_cnt,
                         int fix_off) {
    size_t waves = left + right;
    unsigned blocks = (unsigned)std::max<size_t>(1, std::min<size_t>((waves + 3) / 4, 2048));
    hipLaunchKernelGGL(laswp_apply, dim3(blocks), dim3(256), 0, st, left, skip_to, right, a, lda, nmaps, maps, abort,
                       fix_ipiv, fix_cnt, fix_off);
    BOLA_LAUNCH_CHECK();
}

void launch_laswp(hipStream_t st, size_t n, double* a, size_t lda, int32_t k1, int32_t k2, const int32_t* ipiv,
                  int32_t incx, void* ws, Abort abort) {
    if (n == 0 || incx == 0 || k1 > k2) return;
    size_t npiv = (size_t)(k2 - k1) + 1;
    int ngroups = (int)((npiv + GROUP - 1) / GROUP);
    GroupMap* maps = (GroupMap*)ws;
    hipLaunchKernelGGL(laswp_compose, dim3(ngroups), dim3(64), 0, st, k1, k2, ipiv, incx, ngroups, maps);
    BOLA_LAUNCH_CHECK();
    launch_rowmap_apply(st, a, lda, n, 0, 0, maps, ngroups, abort, nullptr, 0, 0);
}

}  // namespace bola
