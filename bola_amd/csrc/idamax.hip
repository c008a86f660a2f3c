// IDAMAX on gfx950: the reference's x86_64 selection rule, reproduced exactly
// (idamax.rs:23-133), as a two-stage wavefront-shuffle argmax.
//
// The reference keeps 16 (value, index) accumulators, one per residue i mod 16, updated
// on strict ordered '>' and merged lane-wise, then across lanes, then a scalar tail.
// As a selection rule that is: over the first T = 16*floor(n/16) positions take the
// largest non-NaN |x_i|, breaking ties by the smallest key (i mod 4, (i mod 16) div 4, i),
// and report index 0 if that maximum is 0; a tail position i >= T wins only if it is
// strictly larger, first occurrence first.  Below 16 elements: the first strict maximum
// starting from -1 (all-NaN -> 0).  Every candidate here carries that key, so the
// reduction order on the GPU cannot change the answer.
#include "common.hpp"

namespace bola {

namespace {

constexpr int IDAMAX_THREADS = 256;
constexpr uint64_t NOKEY = ~0ull;

struct Cand {
    double v;    // |x|, or -1 for "no candidate"
    uint64_t k;  // tie-break key, smaller wins
};

__device__ __forceinline__ bool better(double v, uint64_t k, const Cand& c) {
    return v > c.v || (v == c.v && k < c.k);
}

__device__ __forceinline__ uint64_t vec_key(uint64_t i) {
    return ((i & 3ull) << 62) | (((i >> 2) & 3ull) << 60) | i;
}

__device__ __forceinline__ void wave_reduce(Cand& c) {
    for (int off = 32; off >= 1; off >>= 1) {
        double ov = __shfl_xor(c.v, off, 64);
        uint64_t ok = (uint64_t)__shfl_xor((long long)c.k, off, 64);
        if (better(ov, ok, c)) { c.v = ov; c.k = ok; }
    }
}

__device__ void block_reduce(Cand& c, Cand* sh) {
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    wave_reduce(c);
    if (lane == 0) sh[w] = c;
    __syncthreads();
    if (w == 0) {
        int nw = blockDim.x >> 6;
        c = lane < nw ? sh[lane] : Cand{-1.0, NOKEY};
        wave_reduce(c);
    }
    __syncthreads();
}

// Stage 1: each block reduces a grid-stride slice into (vector-phase best, tail best).
__global__ void __launch_bounds__(IDAMAX_THREADS) idamax_partial(size_t n, const double* __restrict__ x,
                                                                  Cand* __restrict__ part) {
    __shared__ Cand sh[IDAMAX_THREADS / 64];
    const size_t T = (n < 16) ? 0 : (n & ~(size_t)15);
    Cand vb{-1.0, NOKEY}, tb{-1.0, NOKEY};
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        double v = fabs(x[i]);
        if (v != v) continue;  // NaN never wins
        if (i < T) {
            uint64_t k = vec_key(i);
            if (better(v, k, vb)) { vb.v = v; vb.k = k; }
        } else {
            if (better(v, i, tb)) { tb.v = v; tb.k = i; }
        }
    }
    block_reduce(vb, sh);
    block_reduce(tb, sh);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = vb;
        part[2 * blockIdx.x + 1] = tb;
    }
}

// Stage 2: one block combines the partials and applies the final rule.
__global__ void __launch_bounds__(IDAMAX_THREADS) idamax_final(size_t n, int nparts, const Cand* __restrict__ part,
                                                                int64_t* __restrict__ out) {
    __shared__ Cand sh[IDAMAX_THREADS / 64];
    Cand vb{-1.0, NOKEY}, tb{-1.0, NOKEY};
    for (int p = threadIdx.x; p < nparts; p += blockDim.x) {
        Cand a = part[2 * p], b = part[2 * p + 1];
        if (better(a.v, a.k, vb)) vb = a;
        if (better(b.v, b.k, tb)) tb = b;
    }
    block_reduce(vb, sh);
    block_reduce(tb, sh);
    if (threadIdx.x == 0) {
        int64_t r;
        if (n < 16) {
            r = tb.v >= 0.0 ? (int64_t)tb.k : 0;
        } else {
            double V = 0.0;
            int64_t vi = 0;
            if (vb.v > 0.0) { V = vb.v; vi = (int64_t)(vb.k & ((1ull << 60) - 1)); }
            r = (tb.v > V) ? (int64_t)tb.k : vi;
        }
        *out = r;
    }
}

int idamax_blocks(size_t n) {
    size_t b = (n + IDAMAX_THREADS * 8 - 1) / (IDAMAX_THREADS * 8);
    if (b < 1) b = 1;
    if (b > 1024) b = 1024;
    return (int)b;
}

}  // namespace

size_t idamax_ws_bytes(size_t n) { return sizeof(Cand) * 2 * (size_t)idamax_blocks(n); }

void launch_idamax(hipStream_t st, size_t n, const double* x, int64_t* out, void* ws) {
    int nb = idamax_blocks(n);
    Cand* part = (Cand*)ws;
    hipLaunchKernelGGL(idamax_partial, dim3(nb), dim3(IDAMAX_THREADS), 0, st, n, x, part);
    BOLA_LAUNCH_CHECK();
    hipLaunchKernelGGL(idamax_final, dim3(1), dim3(IDAMAX_THREADS), 0, st, n, nb, part, out);
    BOLA_LAUNCH_CHECK();
}

}  // namespace bola
