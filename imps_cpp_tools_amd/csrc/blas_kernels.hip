// FP64 vector kernels for the Krylov drivers (GMRES, Lanczos): inner products against a block of
// basis vectors, their linear combinations, scaling.  HBM-bound BLAS-1/2 work.
//
// Determinism: every reduction is two-stage with a fixed tree.  Stage one: each workgroup reduces a
// fixed contiguous range of rows (grid-stride order fixed by the launch shape) into a partial per
// basis vector; stage two: one workgroup sums the partials in index order.  Results are bitwise
// reproducible for a given problem size.
#include <hip/hip_runtime.h>

#include "blas.hpp"

namespace imps {

constexpr int kBlasThreads = 256;
constexpr int kRedBlocks = 256;  // stage-one workgroups of a reduction (partials per basis vector)

// sum over the workgroup of one double per thread (fixed tree); result valid in thread 0
__device__ __forceinline__ double block_sum(double v, double* sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int s = kBlasThreads / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] += sh[t + s];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// partial[b*nb + i] = sum over rows of block b of V[i*ldv + r] * w[r], i < nb
__global__ __launch_bounds__(kBlasThreads) void multidot_partial(const double* __restrict__ V, long ldv, int nb,
                                                                 const double* __restrict__ w, long n, double* __restrict__ partial) {
    __shared__ double sh[kBlasThreads];
    const long per = (n + gridDim.x - 1) / gridDim.x;
    const long r0 = blockIdx.x * per, r1 = min(n, r0 + per);
    for (int i = 0; i < nb; ++i) {
        double s = 0;
        for (long r = r0 + threadIdx.x; r < r1; r += kBlasThreads) s += V[i * ldv + r] * w[r];
        s = block_sum(s, sh);
        if (threadIdx.x == 0) partial[(long)blockIdx.x * nb + i] = s;
    }
}

// out[i] = sum_b partial[b*nb + i] (b ascending), one thread per i
__global__ void multidot_final(const double* __restrict__ partial, int nblocks, int nb, double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    double s = 0;
    for (int b = 0; b < nblocks; ++b) s += partial[(long)b * nb + i];
    out[i] = s;
}

// w[r] -= sum_i V[i*ldv + r] * h[i]
__global__ __launch_bounds__(kBlasThreads) void gemv_sub(const double* __restrict__ V, long ldv, int nb, const double* __restrict__ h,
                                                         double* __restrict__ w, long n) {
    for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < n; r += (long)gridDim.x * blockDim.x) {
        double s = 0;
        for (int i = 0; i < nb; ++i) s += V[i * ldv + r] * h[i];
        w[r] -= s;
    }
}

// y[r] = a*x[r] + b*y[r]  (a, b read from device scalars when given, else the immediates)
__global__ void axpby_kernel(double a, const double* __restrict__ da, const double* __restrict__ x, double b, double* __restrict__ y,
                             long n, int inv_a) {
    double aa = da ? *da : a;
    if (inv_a) aa = 1.0 / aa;
    // BLAS convention: a zero coefficient does not read its operand (fresh buffers may hold NaN)
    for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < n; r += (long)gridDim.x * blockDim.x)
        y[r] = (aa != 0.0 ? aa * x[r] : 0.0) + (b != 0.0 ? b * y[r] : 0.0);
}

static unsigned grid_for(long n) {
    long g = (n + kBlasThreads - 1) / kBlasThreads;
    return (unsigned)(g < 1 ? 1 : (g > 2048 ? 2048 : g));
}

hipError_t blas_multidot(const double* V, long ldv, int nb, const double* w, long n, double* out, double* scratch, hipStream_t s) {
    if (nb <= 0) return hipSuccess;
    const int nblk = (int)std::min<long>(kRedBlocks, std::max<long>(1, (n + kBlasThreads - 1) / kBlasThreads));
    hipLaunchKernelGGL(multidot_partial, dim3(nblk), dim3(kBlasThreads), 0, s, V, ldv, nb, w, n, scratch);
    hipLaunchKernelGGL(multidot_final, dim3((nb + 63) / 64), dim3(64), 0, s, scratch, nblk, nb, out);
    return hipGetLastError();
}

hipError_t blas_gemv_sub(const double* V, long ldv, int nb, const double* h, double* w, long n, hipStream_t s) {
    if (nb <= 0 || n <= 0) return hipSuccess;
    hipLaunchKernelGGL(gemv_sub, dim3(grid_for(n)), dim3(kBlasThreads), 0, s, V, ldv, nb, h, w, n);
    return hipGetLastError();
}

hipError_t blas_axpby(double a, const double* da, bool inv_a, const double* x, double b, double* y, long n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(axpby_kernel, dim3(grid_for(n)), dim3(kBlasThreads), 0, s, a, da, x, b, y, n, inv_a ? 1 : 0);
    return hipGetLastError();
}

size_t blas_scratch_doubles(int nb) { return (size_t)kRedBlocks * (size_t)(nb < 1 ? 1 : nb); }

}  // namespace imps
