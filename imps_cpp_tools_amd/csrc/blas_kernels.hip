// FP64 vector kernels for the Krylov drivers (GMRES, Lanczos): inner products against a block of
// basis vectors, their linear combinations, scaling.  HBM-bound BLAS-1/2 work.
//
// Determinism: every reduction is two-stage with a fixed order.  Stage one: workgroup (c, i) reduces
// row chunk c of basis vector i (thread strides, then a fixed wave64 shuffle tree and the four wave
// sums in order) into a partial; stage two sums the chunk partials of each vector, either in chunk
// order (multidot_final) or, inside the kernel that consumes them, by a fixed lane-stride and
// shuffle-tree order (sum_partials_lds).  Results are bitwise reproducible for a given problem size.
#include <hip/hip_runtime.h>

#include "blas.hpp"

namespace imps {

constexpr int kBlasThreads = 256;
constexpr int kRedChunks = 64;     // at most this many row chunks per reduction
constexpr long kChunkRows = 2048;  // at least this many rows per chunk (fewer chunks for short vectors)
constexpr long kMaxGrid = 2048;    // workgroups of the row-parallel kernels (grid-stride beyond)

// sum over the workgroup of one double per thread: wave64 shuffle tree, then the wave sums in order;
// result valid in thread 0
__device__ __forceinline__ double block_sum(double v, double* sh) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0)
        for (int k = 0; k < kBlasThreads / 64; ++k) r += sh[k];
    return r;
}

// grid (nchunks, nb): partial[c*nb + i] = sum over the rows of chunk c of V[i*ldv + r] * w[r]
__global__ __launch_bounds__(kBlasThreads) void multidot_partial(const double* __restrict__ V, long ldv, int nb,
                                                                 const double* __restrict__ w, long n, double* __restrict__ partial) {
    __shared__ double sh[kBlasThreads / 64];
    const int c = blockIdx.x, i = blockIdx.y;
    const long per = (n + gridDim.x - 1) / gridDim.x;
    const long r0 = c * per, r1 = min(n, r0 + per);
    const double* __restrict__ v = V + i * ldv;
    double s = 0;
#pragma unroll 4
    for (long r = r0 + threadIdx.x; r < r1; r += kBlasThreads) s += v[r] * w[r];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) partial[(long)c * nb + i] = s;
}

// out[i] = sum_c partial[c*nb + i] (c ascending), one thread per i
__global__ void multidot_final(const double* __restrict__ partial, int nchunks, int nb, double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    double s = 0;
    for (int c = 0; c < nchunks; ++c) s += partial[(long)c * nb + i];
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

// hs[i] = sum over c < nch of p[c*nb + i] for i < nb, in a fixed order (lane l sums c = l, l + 64,
// ... in turn, then the wave64 shuffle tree; wave w takes i = w, w + 4, ...), identical in every
// workgroup; workgroup 0 also stores the sums at hout (when given).  Ends with a barrier.
__device__ __forceinline__ void sum_partials_lds(const double* __restrict__ p, int nch, int nb, double* hs, double* __restrict__ hout) {
    const int lane = threadIdx.x & 63;
    for (int i = threadIdx.x >> 6; i < nb; i += kBlasThreads / 64) {
        double t = 0;
        for (int c = lane; c < nch; c += 64) t += p[(long)c * nb + i];
        for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
        if (lane == 0) {
            hs[i] = t;
            if (hout && blockIdx.x == 0) hout[i] = t;
        }
    }
    __syncthreads();
}

// w[r] -= sum_i V[i*ldv + r] * h[i], h[i] summed from the multidot chunk partials p by every
// workgroup into LDS (workgroup 0 also stores h at hout); with qpart, qpart[workgroup] = that
// workgroup's part of ||w||^2 after the update.  Dynamic LDS: nb + 4 doubles.
__global__ __launch_bounds__(kBlasThreads) void gemv_sub_partials(const double* __restrict__ V, long ldv, int nb, const double* __restrict__ p,
                                                                  int nch, double* __restrict__ hout, double* __restrict__ w, long n,
                                                                  double* __restrict__ qpart) {
    extern __shared__ double hs[];
    sum_partials_lds(p, nch, nb, hs, hout);
    double q = 0;
    for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < n; r += (long)gridDim.x * blockDim.x) {
        double s = 0;
        int i = 0;
        for (; i + 4 <= nb; i += 4) {  // four independent loads in flight
            const double a0 = V[i * ldv + r], a1 = V[(i + 1) * ldv + r], a2 = V[(i + 2) * ldv + r], a3 = V[(i + 3) * ldv + r];
            s += a0 * hs[i];
            s += a1 * hs[i + 1];
            s += a2 * hs[i + 2];
            s += a3 * hs[i + 3];
        }
        for (; i < nb; ++i) s += V[i * ldv + r] * hs[i];
        const double x = w[r] - s;
        w[r] = x;
        q += x * x;
    }
    if (!qpart) return;
    q = block_sum(q, hs + nb);
    if (threadIdx.x == 0) qpart[blockIdx.x] = q;
}

// y[r] /= sqrt(Q) with Q = sum of the nq partials (fixed order, in every workgroup; y = 0 when
// Q <= 0); workgroup 0 stores Q at qout.
__global__ __launch_bounds__(kBlasThreads) void scale_inv_norm(const double* __restrict__ qpart, int nq, double* __restrict__ qout,
                                                               double* __restrict__ y, long n) {
    __shared__ double sh[kBlasThreads / 64 + 1];
    double t = 0;
    for (int b = threadIdx.x; b < nq; b += kBlasThreads) t += qpart[b];
    t = block_sum(t, sh);
    if (threadIdx.x == 0) {
        sh[kBlasThreads / 64] = t;
        if (blockIdx.x == 0) *qout = t;
    }
    __syncthreads();
    const double Q = sh[kBlasThreads / 64];
    const double a = Q > 0.0 ? 1.0 / sqrt(Q) : 0.0;
    for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < n; r += (long)gridDim.x * blockDim.x) y[r] = a != 0.0 ? a * y[r] : 0.0;
}

// out[r] = x[r] - y[r] + t * z[r], t = the summed nch partials of one inner product (every
// workgroup sums them in the same order)
__global__ __launch_bounds__(kBlasThreads) void diff_plus_scaled(const double* __restrict__ x, const double* __restrict__ y,
                                                                 const double* __restrict__ z, const double* __restrict__ p, int nch,
                                                                 double* __restrict__ out, long n) {
    __shared__ double t1;
    sum_partials_lds(p, nch, 1, &t1, nullptr);
    const double t = t1;
    for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < n; r += (long)gridDim.x * blockDim.x) out[r] = (x[r] - y[r]) + t * z[r];
}

// y[r] = a*x[r] + b*y[r]  (a read from the device scalar da when given, as `mode` says)
__global__ void axpby_kernel(double a, const double* __restrict__ da, const double* __restrict__ x, double b, double* __restrict__ y,
                             long n, int mode) {
    double aa = a;
    if (da) {
        aa = *da;
        if (mode == AXPBY_INV_DA) aa = 1.0 / aa;
        else if (mode == AXPBY_INV_SQRT_DA) aa = aa > 0.0 ? 1.0 / sqrt(aa) : 0.0;
    }
    // BLAS convention: a zero coefficient does not read its operand (fresh buffers may hold NaN)
    for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < n; r += (long)gridDim.x * blockDim.x)
        y[r] = (aa != 0.0 ? aa * x[r] : 0.0) + (b != 0.0 ? b * y[r] : 0.0);
}

static unsigned grid_for(long n) {
    long g = (n + kBlasThreads - 1) / kBlasThreads;
    return (unsigned)(g < 1 ? 1 : (g > kMaxGrid ? kMaxGrid : g));
}

hipError_t blas_multidot(const double* V, long ldv, int nb, const double* w, long n, double* out, double* scratch, hipStream_t s) {
    if (nb <= 0) return hipSuccess;
    if (nb > 65535) return hipErrorInvalidValue;
    const int nch = (int)std::min<long>(kRedChunks, std::max<long>(1, n / kChunkRows));
    // one chunk: the partials are the results
    hipLaunchKernelGGL(multidot_partial, dim3(nch, nb), dim3(kBlasThreads), 0, s, V, ldv, nb, w, n, nch == 1 ? out : scratch);
    if (nch > 1) hipLaunchKernelGGL(multidot_final, dim3((nb + 63) / 64), dim3(64), 0, s, scratch, nch, nb, out);
    return hipGetLastError();
}

int blas_chunks(long n) { return (int)std::min<long>(kRedChunks, std::max<long>(1, n / kChunkRows)); }
int blas_gemv_grid(long n) { return (int)grid_for(n); }

hipError_t blas_multidot_partials(const double* V, long ldv, int nb, const double* w, long n, double* partials, hipStream_t s) {
    if (nb <= 0) return hipSuccess;
    if (nb > 65535) return hipErrorInvalidValue;
    hipLaunchKernelGGL(multidot_partial, dim3(blas_chunks(n), nb), dim3(kBlasThreads), 0, s, V, ldv, nb, w, n, partials);
    return hipGetLastError();
}

hipError_t blas_gemv_sub_partials(const double* V, long ldv, int nb, const double* partials, int nch, double* hout, double* w, long n,
                                  double* qpart, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    nb = std::max(nb, 0);
    if (nb > 8000) return hipErrorInvalidValue;  // h in LDS
    hipLaunchKernelGGL(gemv_sub_partials, dim3(grid_for(n)), dim3(kBlasThreads), sizeof(double) * (nb + kBlasThreads / 64), s, V, ldv, nb,
                       partials, nch, hout, w, n, qpart);
    return hipGetLastError();
}

hipError_t blas_scale_inv_norm(const double* qpart, int nq, double* qout, double* y, long n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(scale_inv_norm, dim3(grid_for(n)), dim3(kBlasThreads), 0, s, qpart, nq, qout, y, n);
    return hipGetLastError();
}

hipError_t blas_diff_plus_scaled(const double* x, const double* y, const double* z, const double* partials, int nch, double* out, long n,
                                 hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(diff_plus_scaled, dim3(grid_for(n)), dim3(kBlasThreads), 0, s, x, y, z, partials, nch, out, n);
    return hipGetLastError();
}

hipError_t blas_gemv_sub(const double* V, long ldv, int nb, const double* h, double* w, long n, hipStream_t s) {
    if (nb <= 0 || n <= 0) return hipSuccess;
    hipLaunchKernelGGL(gemv_sub, dim3(grid_for(n)), dim3(kBlasThreads), 0, s, V, ldv, nb, h, w, n);
    return hipGetLastError();
}

hipError_t blas_axpby(double a, const double* da, int mode, const double* x, double b, double* y, long n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(axpby_kernel, dim3(grid_for(n)), dim3(kBlasThreads), 0, s, a, da, x, b, y, n, mode);
    return hipGetLastError();
}

size_t blas_scratch_doubles(int nb) { return (size_t)kRedChunks * (size_t)(nb < 1 ? 1 : nb) + kMaxGrid; }

}  // namespace imps
