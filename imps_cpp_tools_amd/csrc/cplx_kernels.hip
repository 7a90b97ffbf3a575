// Batched complex vector kernels of the excitation engine (cplx.hpp): the Krylov steps of the
// momentum-batched GMRES and Lanczos solvers and the block gather-scale-sum of the ApplyHeff tape.
// HBM-bound BLAS-1/2 work; every reduction is two-stage with a fixed order (bitwise reproducible).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "cplx.hpp"

namespace imps {

namespace {

constexpr int kT = 256;             // threads per workgroup
constexpr int kMaxChunks = 64;      // row chunks of an inner product
constexpr long kChunkRows = 2048;   // at least this many rows per chunk
constexpr long kMaxRowGroups = 4096;  // row-parallel workgroups summed over all momenta

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// sums of two doubles per thread over the workgroup; valid in thread 0.  sh: 2 * kT / 64 doubles.
__device__ __forceinline__ void block_sum2(double& a, double& b, double* sh) {
    a = wave_sum(a);
    b = wave_sum(b);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sh[2 * w] = a;
        sh[2 * w + 1] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a = 0;
        b = 0;
        for (int k = 0; k < kT / 64; ++k) {
            a += sh[2 * k];
            b += sh[2 * k + 1];
        }
    }
}

// grid (nch, nb, nmom)
__global__ __launch_bounds__(kT) void cdot_partial_kernel(const double* __restrict__ V, long ldv, int nb, const double* __restrict__ w, long n,
                                                          double* __restrict__ part) {
    __shared__ double sh[2 * kT / 64];
    const int c = blockIdx.x, i = blockIdx.y, t = blockIdx.z, nch = gridDim.x;
    const long per = (n + nch - 1) / nch;
    const long r0 = c * per, r1 = min(n, r0 + per);
    const double* __restrict__ vr = V + i * ldv + 2 * t * n;
    const double* __restrict__ vi = vr + n;
    const double* __restrict__ wr = w + 2 * t * n;
    const double* __restrict__ wi = wr + n;
    double sr = 0, si = 0;
    for (long r = r0 + threadIdx.x; r < r1; r += kT) {
        const double a = vr[r], b = vi[r], x = wr[r], y = wi[r];
        sr += a * x + b * y;  // conj(v) w
        si += a * y - b * x;
    }
    block_sum2(sr, si, sh);
    if (threadIdx.x == 0) {
        double* p = part + (((long)t * nb + i) * nch + c) * 2;
        p[0] = sr;
        p[1] = si;
    }
}

// h_i of momentum t summed into LDS (hs[2 i], hs[2 i + 1]) in a fixed order; workgroup 0 of the
// momentum stores them at hout.  Ends with a barrier.
__device__ __forceinline__ void sum_cpartials(const double* __restrict__ part, int nch, int nb, int t, double* hs, double* __restrict__ hout,
                                              long hstride) {
    const int lane = threadIdx.x & 63;
    for (int i = threadIdx.x >> 6; i < nb; i += kT / 64) {
        const double* p = part + ((long)t * nb + i) * nch * 2;
        double a = 0, b = 0;
        for (int c = lane; c < nch; c += 64) {
            a += p[2 * c];
            b += p[2 * c + 1];
        }
        a = wave_sum(a);
        b = wave_sum(b);
        if (lane == 0) {
            hs[2 * i] = a;
            hs[2 * i + 1] = b;
            if (hout && blockIdx.x == 0) {
                hout[t * hstride + 2 * i] = a;
                hout[t * hstride + 2 * i + 1] = b;
            }
        }
    }
    __syncthreads();
}

// grid (gch, nmom); dynamic LDS 2 nb + 2 kT / 64 doubles
__global__ __launch_bounds__(kT) void cgemv_sub_kernel(const double* __restrict__ V, long ldv, int nb, const double* __restrict__ part, int nch,
                                                       double* __restrict__ hout, long hstride, double* __restrict__ w, long n,
                                                       double* __restrict__ qpart) {
    extern __shared__ double hs[];
    const int t = blockIdx.y;
    sum_cpartials(part, nch, nb, t, hs, hout, hstride);
    double* __restrict__ wr = w + 2 * t * n;
    double* __restrict__ wi = wr + n;
    double q = 0, dummy = 0;
    for (long r = blockIdx.x * (long)kT + threadIdx.x; r < n; r += (long)gridDim.x * kT) {
        double xr = wr[r], xi = wi[r];
        for (int i = 0; i < nb; ++i) {
            const double* v = V + i * ldv + 2 * t * n;
            const double a = v[r], b = v[n + r];
            const double hr = hs[2 * i], hi = hs[2 * i + 1];
            xr -= hr * a - hi * b;
            xi -= hr * b + hi * a;
        }
        wr[r] = xr;
        wi[r] = xi;
        q += xr * xr + xi * xi;
    }
    if (!qpart) return;
    block_sum2(q, dummy, hs + 2 * nb);
    if (threadIdx.x == 0) qpart[(long)t * gridDim.x + blockIdx.x] = q;
}

// grid (gch, nmom)
__global__ __launch_bounds__(kT) void cscale_kernel(const double* __restrict__ qpart, int gch, double* __restrict__ qout, long qstride,
                                                    double* __restrict__ w, long n) {
    __shared__ double sh[2 * kT / 64 + 1];
    const int t = blockIdx.y;
    double a = 0, dummy = 0;
    for (int b = threadIdx.x; b < gch; b += kT) a += qpart[(long)t * gch + b];
    block_sum2(a, dummy, sh);
    if (threadIdx.x == 0) {
        sh[2 * kT / 64] = a;
        if (blockIdx.x == 0 && qout) qout[t * qstride] = a;
    }
    __syncthreads();
    const double Q = sh[2 * kT / 64];
    const double s = Q > 0.0 ? 1.0 / sqrt(Q) : 0.0;
    double* __restrict__ wr = w + 2 * t * n;
    for (long r = blockIdx.x * (long)kT + threadIdx.x; r < 2 * n; r += (long)gridDim.x * kT) wr[r] = s != 0.0 ? s * wr[r] : 0.0;
}

// grid (gch, nmom)
__global__ __launch_bounds__(kT) void cgemv_add_kernel(const double* __restrict__ V, long ldv, int nb, const double* __restrict__ coef,
                                                       long cstride, double* __restrict__ x, long n) {
    const int t = blockIdx.y;
    const double* c = coef + t * cstride;
    double* __restrict__ xr = x + 2 * t * n;
    double* __restrict__ xi = xr + n;
    for (long r = blockIdx.x * (long)kT + threadIdx.x; r < n; r += (long)gridDim.x * kT) {
        double ar = xr[r], ai = xi[r];
        for (int i = 0; i < nb; ++i) {
            const double* v = V + i * ldv + 2 * t * n;
            const double a = v[r], b = v[n + r], cr = c[2 * i], ci = c[2 * i + 1];
            ar += cr * a - ci * b;
            ai += cr * b + ci * a;
        }
        xr[r] = ar;
        xi[r] = ai;
    }
}

// grid (gch, nmom, nk)
__global__ __launch_bounds__(kT) void cgemm_basis_kernel(const double* __restrict__ V, long ldv, int nb, const double* __restrict__ S, int nk,
                                                         double* __restrict__ Y, long ldy, long n) {
    const int t = blockIdx.y, i = blockIdx.z;
    const double* s = S + ((long)t * nk + i) * nb * 2;
    double* __restrict__ yr = Y + i * ldy + 2 * t * n;
    double* __restrict__ yi = yr + n;
    for (long r = blockIdx.x * (long)kT + threadIdx.x; r < n; r += (long)gridDim.x * kT) {
        double ar = 0, ai = 0;
        for (int j = 0; j < nb; ++j) {
            const double* v = V + j * ldv + 2 * t * n;
            const double a = v[r], b = v[n + r], cr = s[2 * j], ci = s[2 * j + 1];
            ar += cr * a - ci * b;
            ai += cr * b + ci * a;
        }
        yr[r] = ar;
        yi[r] = ai;
    }
}

// grid (nch, nmom)
__global__ __launch_bounds__(kT) void rdot_partial_kernel(const double* __restrict__ wv, const double* __restrict__ z, long n,
                                                          double* __restrict__ part) {
    __shared__ double sh[2 * kT / 64];
    const int c = blockIdx.x, t = blockIdx.y, nch = gridDim.x;
    const long per = (n + nch - 1) / nch;
    const long r0 = c * per, r1 = min(n, r0 + per);
    const double* __restrict__ zr = z + 2 * t * n;
    const double* __restrict__ zi = zr + n;
    double sr = 0, si = 0;
    for (long r = r0 + threadIdx.x; r < r1; r += kT) {
        sr += wv[r] * zr[r];
        si += wv[r] * zi[r];
    }
    block_sum2(sr, si, sh);
    if (threadIdx.x == 0) {
        part[((long)t * nch + c) * 2] = sr;
        part[((long)t * nch + c) * 2 + 1] = si;
    }
}

// grid (gch, nmom)
__global__ __launch_bounds__(kT) void geo_combine_kernel(double* __restrict__ out, double a, const double* __restrict__ z,
                                                         const double* __restrict__ Tz, const double* __restrict__ fac,
                                                         const double* __restrict__ gam, const double* __restrict__ tpart, int nch,
                                                         const double* __restrict__ rout, long n) {
    __shared__ double tr[2];
    const int t = blockIdx.y;
    if (tpart) {
        if (threadIdx.x < 64) {
            double x = 0, y = 0;
            for (int c = threadIdx.x; c < nch; c += 64) {
                x += tpart[((long)t * nch + c) * 2];
                y += tpart[((long)t * nch + c) * 2 + 1];
            }
            x = wave_sum(x);
            y = wave_sum(y);
            if (threadIdx.x == 0) {
                tr[0] = gam[t] * x;
                tr[1] = gam[t] * y;
            }
        }
        __syncthreads();
    }
    const double fr = Tz ? fac[2 * t] : 0.0, fi = Tz ? fac[2 * t + 1] : 0.0;
    const double gr = tpart ? tr[0] : 0.0, gi = tpart ? tr[1] : 0.0;
    const double* __restrict__ zr = z + 2 * t * n;
    const double* __restrict__ zi = zr + n;
    double* __restrict__ orr = out + 2 * t * n;
    double* __restrict__ oi = orr + n;
    const double* __restrict__ tr_ = Tz ? Tz + 2 * t * n : nullptr;
    for (long r = blockIdx.x * (long)kT + threadIdx.x; r < n; r += (long)gridDim.x * kT) {
        double xr = a * zr[r], xi = a * zi[r];
        if (tr_) {
            const double ur = tr_[r], ui = tr_[n + r];
            xr -= fr * ur - fi * ui;
            xi -= fr * ui + fi * ur;
        }
        if (tpart) {
            xr += gr * rout[r];
            xi += gi * rout[r];
        }
        orr[r] = xr;
        oi[r] = xi;
    }
}

// grid (blocks, nmom)
__global__ __launch_bounds__(kT) void cinterleave_kernel(const double* __restrict__ planar, double* __restrict__ inter, long n) {
    const int t = blockIdx.y;
    const double* pr = planar + 2 * t * n;
    double2* o = reinterpret_cast<double2*>(inter) + t * n;
    for (long r = blockIdx.x * (long)kT + threadIdx.x; r < n; r += (long)gridDim.x * kT) o[r] = make_double2(pr[r], pr[n + r]);
}

__global__ __launch_bounds__(kT) void cdeinterleave_kernel(const double* __restrict__ inter, double* __restrict__ planar, long n) {
    const int t = blockIdx.y;
    double* pr = planar + 2 * t * n;
    const double2* in = reinterpret_cast<const double2*>(inter) + t * n;
    for (long r = blockIdx.x * (long)kT + threadIdx.x; r < n; r += (long)gridDim.x * kT) {
        const double2 v = in[r];
        pr[r] = v.x;
        pr[n + r] = v.y;
    }
}

// grid (ntiles, nmom or 1)
__global__ __launch_bounds__(kT) void comb_kernel(const CombTile* __restrict__ tiles, const CombTerm* __restrict__ terms,
                                                  const double* __restrict__ coef, double* __restrict__ arena, long out_ps, int out_cplx, int nmom) {
    const CombTile tl = tiles[blockIdx.x];
    const int t = blockIdx.y;
    for (int e = threadIdx.x; e < tl.len; e += kT) {
        double xr = 0, xi = 0;
        for (int k = 0; k < tl.nt; ++k) {
            const CombTerm tm = terms[tl.t0 + k];
            const double cr = coef[((long)tm.coef * nmom + t) * 2], ci = coef[((long)tm.coef * nmom + t) * 2 + 1];
            if (tm.src_ps == 0) {
                const double x = arena[tm.src_off + e];
                xr += cr * x;
                xi += ci * x;
            } else {
                const double a = arena[tm.src_off + 2 * t * tm.src_ps + e], b = arena[tm.src_off + (2 * t + 1) * tm.src_ps + e];
                xr += cr * a - ci * b;
                xi += cr * b + ci * a;
            }
        }
        if (out_cplx) {
            arena[tl.out_off + 2 * t * out_ps + e] = xr;
            arena[tl.out_off + (2 * t + 1) * out_ps + e] = xi;
        } else {
            arena[tl.out_off + e] = xr;
        }
    }
}

unsigned row_groups(long n, int nmom) {
    const long per = std::max<long>(8, kMaxRowGroups / std::max(1, nmom));
    long g = (n + kT - 1) / kT;
    return (unsigned)std::max<long>(1, std::min<long>(g, per));
}

}  // namespace

int cblas_chunks(long n) { return (int)std::min<long>(kMaxChunks, std::max<long>(1, n / kChunkRows)); }
int cblas_gemv_grid(long n, int nmom) { return (int)row_groups(n, nmom); }

hipError_t cdot_partials(const double* V, long ldv, int nb, const double* w, long n, int nmom, double* part, hipStream_t s) {
    if (nb <= 0 || nmom <= 0) return hipSuccess;
    if (nb > 65535 || nmom > 65535) return hipErrorInvalidValue;
    hipLaunchKernelGGL(cdot_partial_kernel, dim3(cblas_chunks(n), nb, nmom), dim3(kT), 0, s, V, ldv, nb, w, n, part);
    return hipGetLastError();
}

// the gemv kernels use row_groups(n, nmom) workgroups per momentum; its partial count is what
// cscale_inv_norm must be given
static unsigned gch_of(long n, int nmom) { return row_groups(n, nmom); }

hipError_t cgemv_sub_partials(const double* V, long ldv, int nb, const double* part, int nch, double* hout, long hstride, double* w, long n,
                              int nmom, double* qpart, hipStream_t s) {
    if (n <= 0 || nmom <= 0) return hipSuccess;
    nb = std::max(nb, 0);
    if (nb > 4000 || nmom > 65535) return hipErrorInvalidValue;
    hipLaunchKernelGGL(cgemv_sub_kernel, dim3(gch_of(n, nmom), nmom), dim3(kT), sizeof(double) * (2 * nb + 2 * kT / 64), s, V, ldv, nb, part,
                       nch, hout, hstride, w, n, qpart);
    return hipGetLastError();
}

hipError_t cscale_inv_norm(const double* qpart, int gch, double* qout, long qstride, double* w, long n, int nmom, hipStream_t s) {
    if (n <= 0 || nmom <= 0) return hipSuccess;
    if ((unsigned)gch != gch_of(n, nmom)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(cscale_kernel, dim3(gch_of(n, nmom), nmom), dim3(kT), 0, s, qpart, gch, qout, qstride, w, n);
    return hipGetLastError();
}

hipError_t cgemv_add(const double* V, long ldv, int nb, const double* coef, long cstride, double* x, long n, int nmom, hipStream_t s) {
    if (n <= 0 || nb <= 0 || nmom <= 0) return hipSuccess;
    hipLaunchKernelGGL(cgemv_add_kernel, dim3(gch_of(n, nmom), nmom), dim3(kT), 0, s, V, ldv, nb, coef, cstride, x, n);
    return hipGetLastError();
}

hipError_t cgemm_basis(const double* V, long ldv, int nb, const double* S, int nk, double* Y, long ldy, long n, int nmom, hipStream_t s) {
    if (n <= 0 || nk <= 0 || nmom <= 0) return hipSuccess;
    if (nk > 65535) return hipErrorInvalidValue;
    hipLaunchKernelGGL(cgemm_basis_kernel, dim3(gch_of(n, nmom), nmom, nk), dim3(kT), 0, s, V, ldv, nb, S, nk, Y, ldy, n);
    return hipGetLastError();
}

hipError_t rdot_partials(const double* wv, const double* z, long n, int nmom, double* part, hipStream_t s) {
    if (nmom <= 0) return hipSuccess;
    hipLaunchKernelGGL(rdot_partial_kernel, dim3(cblas_chunks(n), nmom), dim3(kT), 0, s, wv, z, n, part);
    return hipGetLastError();
}

hipError_t geo_combine(double* out, double a, const double* z, const double* Tz, const double* fac, const double* gam, const double* tpart,
                       int nch, const double* rout, long n, int nmom, hipStream_t s) {
    if (n <= 0 || nmom <= 0) return hipSuccess;
    hipLaunchKernelGGL(geo_combine_kernel, dim3(gch_of(n, nmom), nmom), dim3(kT), 0, s, out, a, z, Tz, fac, gam, tpart, nch, rout, n);
    return hipGetLastError();
}

hipError_t cinterleave(const double* planar, double* inter, long n, int nmom, hipStream_t s) {
    if (n <= 0 || nmom <= 0) return hipSuccess;
    hipLaunchKernelGGL(cinterleave_kernel, dim3(gch_of(n, nmom), nmom), dim3(kT), 0, s, planar, inter, n);
    return hipGetLastError();
}

hipError_t cdeinterleave(const double* inter, double* planar, long n, int nmom, hipStream_t s) {
    if (n <= 0 || nmom <= 0) return hipSuccess;
    hipLaunchKernelGGL(cdeinterleave_kernel, dim3(gch_of(n, nmom), nmom), dim3(kT), 0, s, inter, planar, n);
    return hipGetLastError();
}

hipError_t comb_launch(const CombTile* tiles, int ntiles, const CombTerm* terms, const double* coef, double* arena, long out_ps, int out_cplx,
                       int nmom, hipStream_t s) {
    if (ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(comb_kernel, dim3(ntiles, out_cplx ? nmom : 1), dim3(kT), 0, s, tiles, terms, coef, arena, out_ps, out_cplx, nmom);
    return hipGetLastError();
}

}  // namespace imps
