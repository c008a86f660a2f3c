// BLAS-1/2 kernels: SCAL, AXPY, DOT, GER and the beta pre-scaling helpers.
//
// Elementwise kernels reproduce the reference's rounding exactly: the library is built
// with -ffp-contract=off, and a fused multiply-add appears only where the reference
// calls _mm256_fmadd_pd (4-wide chunks), with the unfused `y += a*x` form in its
// scalar tails.  DOT is a parallel tree reduction (tolerance parity, not bitwise).
#include "common.hpp"

namespace bola {

namespace {

constexpr int TPB = 256;

unsigned grid_for(size_t n, int per_thread = 1) {
    size_t b = (n + (size_t)TPB * per_thread - 1) / ((size_t)TPB * per_thread);
    if (b < 1) b = 1;
    if (b > 65535u * 8) b = 65535u * 8;
    return (unsigned)b;
}

// scal, scal.rs:45-56: x *= alpha (a plain multiply on every element).
__global__ void k_scal(size_t n, double alpha, double* __restrict__ x, size_t incx) {
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) x[i * incx] *= alpha;
}

// daxpy, utilities.rs:107-132: fused on the first 4*floor(n/4) elements when both
// strides are 1, unfused elsewhere.
__global__ void k_daxpy(size_t n, double alpha, const double* __restrict__ x, size_t incx, double* __restrict__ y,
                        size_t incy) {
    size_t stride = (size_t)gridDim.x * blockDim.x;
    const bool unit = incx == 1 && incy == 1;
    const size_t nf = unit ? (n & ~(size_t)3) : 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        double xv = x[i * incx];
        double yv = y[i * incy];
        y[i * incy] = (i < nf) ? fma(alpha, xv, yv) : yv + alpha * xv;
    }
}

__global__ void k_ddot_partial(size_t n, const double* __restrict__ x, size_t incx, const double* __restrict__ y,
                               size_t incy, double* __restrict__ part) {
    __shared__ double sh[TPB / 64];
    double s = 0.0;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        s = fma(x[i * incx], y[i * incy], s);
    for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < TPB / 64; w++) t += sh[w];
        part[blockIdx.x] = t;
    }
}

__global__ void k_ddot_final(int nparts, const double* __restrict__ part, double* __restrict__ out) {
    __shared__ double sh[TPB / 64];
    double s = 0.0;
    for (int p = threadIdx.x; p < nparts; p += blockDim.x) s += part[p];
    for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < TPB / 64; w++) t += sh[w];
        *out = t;
    }
}

// ger, ger.rs:28-60: column pairs; a pair whose two alpha*y values are both zero is
// skipped; fused on rows below 4*floor(m/4) of a pair, unfused on the tail rows and on
// an odd last column (which is skipped when its alpha*y is zero).
__global__ void k_dger(size_t m, size_t n, double alpha, const double* __restrict__ x, const double* __restrict__ y,
                       size_t ldy, double* __restrict__ a, size_t lda) {
    size_t j = blockIdx.y;
    const size_t n_unrolled = n - (n % 2);
    const size_t mf = m & ~(size_t)3;
    double t = alpha * y[j * ldy];
    bool paired = j < n_unrolled;
    if (paired) {
        size_t j0 = j & ~(size_t)1;
        double t0 = alpha * y[j0 * ldy], t1 = alpha * y[(j0 + 1) * ldy];
        if (t0 == 0.0 && t1 == 0.0) return;
    } else if (t == 0.0) {
        return;
    }
    double* col = a + j * lda;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
        double av = col[i];
        col[i] = (paired && i < mf) ? fma(x[i], t, av) : av + t * x[i];
    }
}

__global__ void k_fill(double* __restrict__ a, size_t lda, size_t m, size_t n, double v) {
    size_t j = blockIdx.y;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (size_t)gridDim.x * blockDim.x)
        a[i + j * lda] = v;
}

__global__ void k_scale(double* __restrict__ a, size_t lda, size_t m, size_t n, double beta) {
    size_t j = blockIdx.y;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (size_t)gridDim.x * blockDim.x)
        a[i + j * lda] *= beta;
}

// beta pre-scaling of one triangle (syrk.rs:37-61): beta == 0 writes zeros.
__global__ void k_scale_tri(int upper, double* __restrict__ a, size_t lda, size_t n, double beta) {
    size_t j = blockIdx.y;
    size_t i0 = upper ? 0 : j, i1 = upper ? j + 1 : n;
    for (size_t i = i0 + (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < i1; i += (size_t)gridDim.x * blockDim.x)
        a[i + j * lda] = (beta == 0.0) ? 0.0 : a[i + j * lda] * beta;
}

}  // namespace

void launch_dscal(hipStream_t st, size_t n, double alpha, double* x, size_t incx) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_scal, dim3(grid_for(n, 4)), dim3(TPB), 0, st, n, alpha, x, incx);
    BOLA_LAUNCH_CHECK();
}

void launch_daxpy(hipStream_t st, size_t n, double alpha, const double* x, size_t incx, double* y, size_t incy) {
    if (n == 0 || alpha == 0.0) return;
    hipLaunchKernelGGL(k_daxpy, dim3(grid_for(n, 4)), dim3(TPB), 0, st, n, alpha, x, incx, y, incy);
    BOLA_LAUNCH_CHECK();
}

size_t ddot_ws_bytes(size_t n) { return sizeof(double) * (size_t)std::min<unsigned>(grid_for(n, 8), 1024u); }

void launch_ddot(hipStream_t st, size_t n, const double* x, size_t incx, const double* y, size_t incy, double* out,
                 void* ws) {
    unsigned nb = std::min<unsigned>(grid_for(n, 8), 1024u);
    double* part = (double*)ws;
    hipLaunchKernelGGL(k_ddot_partial, dim3(nb), dim3(TPB), 0, st, n, x, incx, y, incy, part);
    BOLA_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_ddot_final, dim3(1), dim3(TPB), 0, st, (int)nb, part, out);
    BOLA_LAUNCH_CHECK();
}

void launch_dger(hipStream_t st, size_t m, size_t n, double alpha, const double* x, const double* y, size_t ldy,
                 double* a, size_t lda) {
    if (m == 0 || n == 0) return;
    unsigned gx = std::min<unsigned>(grid_for(m, 4), 64u);
    hipLaunchKernelGGL(k_dger, dim3(gx, (unsigned)n), dim3(TPB), 0, st, m, n, alpha, x, y, ldy, a, lda);
    BOLA_LAUNCH_CHECK();
}

void launch_fill(hipStream_t st, double* a, size_t lda, size_t m, size_t n, double v) {
    if (m == 0 || n == 0) return;
    unsigned gx = std::min<unsigned>(grid_for(m, 4), 64u);
    hipLaunchKernelGGL(k_fill, dim3(gx, (unsigned)n), dim3(TPB), 0, st, a, lda, m, n, v);
    BOLA_LAUNCH_CHECK();
}

void launch_scale(hipStream_t st, double* a, size_t lda, size_t m, size_t n, double beta) {
    if (m == 0 || n == 0) return;
    unsigned gx = std::min<unsigned>(grid_for(m, 4), 64u);
    hipLaunchKernelGGL(k_scale, dim3(gx, (unsigned)n), dim3(TPB), 0, st, a, lda, m, n, beta);
    BOLA_LAUNCH_CHECK();
}

void launch_scale_tri(hipStream_t st, char uplo, double* a, size_t lda, size_t n, double beta) {
    if (n == 0) return;
    unsigned gx = std::min<unsigned>(grid_for(n, 4), 64u);
    hipLaunchKernelGGL(k_scale_tri, dim3(gx, (unsigned)n), dim3(TPB), 0, st, uplo == 'U' ? 1 : 0, a, lda, n, beta);
    BOLA_LAUNCH_CHECK();
}

}  // namespace bola
