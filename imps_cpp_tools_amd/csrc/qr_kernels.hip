// Batched per-sector QR with non-negative R diagonal (qr_pos, helpers.hpp:181-195) of the dense
// regrouped sector matrices of an MPS tensor (GetDenseMat, MPSBlockMat.hpp:468-526), FP64.
//
// One 16-wave workgroup per charge sector.  The sector matrix (rows = sum of the parts' row counts,
// cols = the sector size) is gathered from the MPS buffer into LDS, or, when it does not fit, into
// a device work area of its own (the workgroup barriers order its global accesses as they do the
// LDS ones), and factorised by Householder reflections (LAPACK dlarfg convention, v(0) = 1).  Per
// reflector, each wave updates four trailing columns at a time (their dot products interleaved,
// each a fixed-order shuffle reduction within the wave, so the result is deterministic); one
// workgroup barrier per reflector.  R is written out, then Q (economy, rows x cols) is formed in
// place by applying the reflectors backwards to the leading columns of the identity (the order of
// LAPACK dorg2r), with the same column-parallel updates, and written back in block layout.  The
// "right" form factorises the transposed sector matrix (dirtype r), with a transposed gather and
// scatter.
#include <hip/hip_runtime.h>

#include "qr.hpp"

namespace imps {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Apply H = I - tau v v^T (v(j) = 1, v(r) = D(r, j) * vs for r > j) from the left to the columns
// c in [c0, n) of rows j .. m-1; each wave takes four columns at a time, and the row loops are
// unrolled so that a lane has several rows' loads in flight.  When `nrm` is given, the wave that
// updates column j + 1 also writes the squared norm of its rows j + 2 .. m-1 there (the next
// reflector's input).
template <int WAVES>
__device__ __forceinline__ void apply_reflector(double* D, int m, int n, int j, int c0, double tau, double vs, int wave, int lane,
                                                double* nrm) {
    const double* vj = D + (long)j * m;
    for (int cb = c0 + 4 * wave; cb < n; cb += 4 * WAVES) {
        const int nc = min(4, n - cb);
        double* col[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) col[q] = D + (long)(cb + min(q, nc - 1)) * m;
        double p[4] = {0, 0, 0, 0};
#pragma unroll 4
        for (int r = j + 1 + lane; r < m; r += 64) {
            const double v = vj[r];
#pragma unroll
            for (int q = 0; q < 4; ++q) p[q] += v * col[q][r];
        }
        double f[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) f[q] = tau * (vs * wave_sum(p[q]) + col[q][j]);
        const bool next = nrm && cb == j + 1;  // column j + 1 is this group's first
        double s2 = 0;
#pragma unroll 4
        for (int r = j + 1 + lane; r < m; r += 64) {
            const double v = vs * vj[r];
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (q < nc) {
                    const double x = col[q][r] - f[q] * v;
                    col[q][r] = x;
                    if (q == 0 && next && r > j + 1) s2 += x * x;
                }
        }
        if (next) {
            s2 = wave_sum(s2);
            if (lane == 0) *nrm = s2;
        }
        if (lane == 0) {
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (q < nc) col[q][j] -= f[q];
        }
    }
}

template <bool GLOBAL>
__global__ __launch_bounds__(64 * kQrWaves) void sector_qr_kernel(const QrSector* __restrict__ secs, const QrPart* __restrict__ parts,
                                                                 const double* __restrict__ A, double* __restrict__ Q,
                                                                 double* __restrict__ R, int right, double* __restrict__ gwork) {
    extern __shared__ double lds[];
    __shared__ double nrm2[2];  // squared norms below the diagonal of the next reflector's column (double-buffered)
    const QrSector sc = secs[blockIdx.x];
    const int m = sc.m, n = sc.n, k = min(m, n);
    double* D = GLOBAL ? gwork + sc.woff : lds;  // m x n, column-major
    double* taus = D + (long)m * n;              // k
    double* scals = taus + k;                    // k: reflector scale (the stored v is unscaled)
    double* betas = scals + k;                   // k: R's diagonal before the sign fix
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    // gather:  left form D(r, c) = A[src + (r - off) + c*ld];  right form D(r, c) = A[src + c + (r - off)*ld]
    for (int p = sc.part0; p < sc.part0 + sc.nparts; ++p) {
        const QrPart pt = parts[p];
        for (long e = t; e < (long)pt.len * n; e += 64 * kQrWaves) {
            const int rr = (int)(e % pt.len), c = (int)(e / pt.len);
            const long src = right ? pt.src + c + (long)rr * pt.ld : pt.src + rr + (long)c * pt.ld;
            D[(pt.off + rr) + (long)c * m] = A[src];
        }
    }
    __syncthreads();
    if (wave == 0) {
        double s = 0;
        for (int r = 1 + lane; r < m; r += 64) s += D[r] * D[r];
        s = wave_sum(s);
        if (lane == 0) nrm2[0] = s;
    }
    __syncthreads();
    // Householder factorisation, one barrier per reflector: D(j, j) and the rows below of column j
    // are final at the start of step j, and its squared norm below the diagonal is in nrm2[j & 1]
    for (int j = 0; j < k; ++j) {
        const double sig = nrm2[j & 1];
        const double x0 = D[j + (long)j * m];
        double tau = 0, beta = x0, scal = 0;
        if (sig > 0) {
            const double nr = sqrt(x0 * x0 + sig);
            beta = x0 >= 0 ? -nr : nr;
            tau = (beta - x0) / beta;
            scal = 1.0 / (x0 - beta);
        }
        double* nxt = &nrm2[(j + 1) & 1];
        if (tau != 0) {
            apply_reflector<kQrWaves>(D, m, n, j, j + 1, tau, scal, wave, lane, nxt);
        } else if (j + 1 < n && wave == 0) {  // no update: column j + 1's norm as it stands
            double s2 = 0;
            for (int r = j + 2 + lane; r < m; r += 64) s2 += D[r + (long)(j + 1) * m] * D[r + (long)(j + 1) * m];
            s2 = wave_sum(s2);
            if (lane == 0) *nxt = s2;
        }
        if (t == 0) { taus[j] = tau; scals[j] = scal; betas[j] = beta; }
        __syncthreads();
    }
    // R with rows scaled to a non-negative diagonal; left form stores the k x n block,
    // right form its transpose (n x k)
    for (long e = t; e < (long)k * n; e += 64 * kQrWaves) {
        const int r = (int)(e % k), c = (int)(e / k);
        const double sg = betas[r] < 0 ? -1.0 : 1.0;
        const double val = r < c ? sg * D[r + (long)c * m] : r == c ? sg * betas[r] : 0.0;
        if (right) R[sc.rdst + c + (long)r * n] = val;
        else R[sc.rdst + r + (long)c * k] = val;
    }
    __syncthreads();
    // Q = H_0 ... H_{k-1} [I; 0] in place (dorg2r): backwards over the reflectors, H_j applied to
    // the columns j+1 .. k-1 (already Q columns), then column j becomes H_j e_j
    for (int j = k - 1; j >= 0; --j) {
        const double tau = taus[j], scal = scals[j];
        if (j + 1 < k && tau != 0) apply_reflector<kQrWaves>(D, m, k, j, j + 1, tau, scal, wave, lane, nullptr);
        __syncthreads();
        if (wave == 0) {
            const double a = -tau * scal;
            for (int r = j + 1 + lane; r < m; r += 64) D[r + (long)j * m] *= a;
            for (int r = lane; r < j; r += 64) D[r + (long)j * m] = 0.0;
            if (lane == 0) D[j + (long)j * m] = 1.0 - tau;
        }
        __syncthreads();
    }
    // scatter Q's columns, each times the sign of its R row
    for (int p = sc.part0; p < sc.part0 + sc.nparts; ++p) {
        const QrPart pt = parts[p];
        for (long e = t; e < (long)pt.len * k; e += 64 * kQrWaves) {
            const int rr = (int)(e % pt.len), c = (int)(e / pt.len);
            const double val = (betas[c] < 0 ? -1.0 : 1.0) * D[(pt.off + rr) + (long)c * m];
            if (right) Q[pt.qdst + c + (long)rr * k] = val;   // part block k x len
            else Q[pt.qdst + rr + (long)c * pt.len] = val;    // part block len x k
        }
    }
}

hipError_t launch_sector_qr(const QrSector* secs, int nlds, int nglobal, const QrPart* parts, const double* A, double* Q, double* R,
                            bool right, size_t lds_bytes, double* gwork, hipStream_t s) {
    if (nlds > 0) {
        if (lds_bytes > 64 * 1024) {  // beyond the default dynamic-LDS limit of a launch
            const hipError_t e = hipFuncSetAttribute((const void*)sector_qr_kernel<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                     (int)lds_bytes);
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(sector_qr_kernel<false>, dim3(nlds), dim3(64 * kQrWaves), lds_bytes, s, secs, parts, A, Q, R, right ? 1 : 0, gwork);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (nglobal > 0) {
        if (!gwork) return hipErrorInvalidValue;
        hipLaunchKernelGGL(sector_qr_kernel<true>, dim3(nglobal), dim3(64 * kQrWaves), 0, s, secs + nlds, parts, A, Q, R, right ? 1 : 0,
                           gwork);
    }
    return hipGetLastError();
}

}  // namespace imps
