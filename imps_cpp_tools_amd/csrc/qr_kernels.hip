// Batched per-sector QR with non-negative R diagonal (qr_pos, helpers.hpp:181-195) of the dense
// regrouped sector matrices of an MPS tensor (GetDenseMat, MPSBlockMat.hpp:468-526), FP64.
//
// One workgroup per charge sector.  The sector matrix (rows = sum of the parts' row counts, cols =
// the sector size) is gathered from the MPS buffer into LDS (4 waves), or, when it does not fit,
// into a device work area of its own (16 waves; the workgroup barriers order its global accesses
// as they do the LDS ones), and factorised by Householder reflections (LAPACK dlarfg convention,
// v(0) = 1).  Per reflector, the waves split the
// trailing columns between them; each column's dot product is a fixed-order shuffle reduction
// within one wave, so the result is deterministic and needs one workgroup barrier per reflector.
// Q (economy, rows x min(rows, cols)) is then formed column by column (columns split across waves)
// and written back in block layout together with R.  The "right" form factorises the transposed
// sector matrix (dirtype r), with a transposed gather and scatter.
#include <hip/hip_runtime.h>

#include "qr.hpp"

namespace imps {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <int kQrWaves, bool GLOBAL>
__global__ __launch_bounds__(64 * kQrWaves) void sector_qr_kernel(const QrSector* __restrict__ secs, const QrPart* __restrict__ parts,
                                                                 const double* __restrict__ A, double* __restrict__ Q,
                                                                 double* __restrict__ R, int right, double* __restrict__ gwork) {
    extern __shared__ double lds[];
    const QrSector sc = secs[blockIdx.x];
    const int m = sc.m, n = sc.n, k = min(m, n);
    double* D = GLOBAL ? gwork + sc.woff : lds;  // m x n, column-major
    double* taus = D + (long)m * n;             // k
    double* vs = taus + k;                      // kQrWaves x m column scratch
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
    for (int j = 0; j < k; ++j) {
        // reflector for column j (computed redundantly by every wave: identical results)
        double s = 0;
        for (int r = j + 1 + lane; r < m; r += 64) s += D[r + (long)j * m] * D[r + (long)j * m];
        const double sig = wave_sum(s);
        const double x0 = D[j + (long)j * m];
        double tau = 0, beta = x0, scal = 0;
        if (sig > 0) {
            const double nrm = sqrt(x0 * x0 + sig);
            beta = x0 >= 0 ? -nrm : nrm;
            tau = (beta - x0) / beta;
            scal = 1.0 / (x0 - beta);
        }
        // trailing columns, split across waves; v(r) = D(r, j) * scal for r > j (not yet scaled in LDS)
        if (tau != 0) {
            for (int c = j + 1 + wave; c < n; c += kQrWaves) {
                double p = 0;
                for (int r = j + 1 + lane; r < m; r += 64) p += D[r + (long)j * m] * D[r + (long)c * m];
                const double f = tau * (scal * wave_sum(p) + D[j + (long)c * m]);
                for (int r = j + 1 + lane; r < m; r += 64) D[r + (long)c * m] -= f * scal * D[r + (long)j * m];
                if (lane == 0) D[j + (long)c * m] -= f;
            }
        }
        __syncthreads();
        if (wave == 0) {
            for (int r = j + 1 + lane; r < m; r += 64) D[r + (long)j * m] *= scal;
            if (lane == 0) { D[j + (long)j * m] = beta; taus[j] = tau; }
        }
        __syncthreads();
    }
    // R with rows scaled to a non-negative diagonal; left form stores the k x n block,
    // right form its transpose (n x k)
    for (long e = t; e < (long)k * n; e += 64 * kQrWaves) {
        const int r = (int)(e % k), c = (int)(e / k);
        const double sgn = D[r + (long)r * m] < 0 ? -1.0 : 1.0;
        const double val = r <= c ? sgn * D[r + (long)c * m] : 0.0;
        if (right) R[sc.rdst + c + (long)r * n] = val;
        else R[sc.rdst + r + (long)c * k] = val;
    }
    // Q columns: Q e_c = H_0 ... H_{k-1} e_c, columns split across waves
    double* v = vs + (long)wave * m;
    for (int c = wave; c < k; c += kQrWaves) {
        for (int r = lane; r < m; r += 64) v[r] = (r == c) ? 1.0 : 0.0;
        for (int j = min(c, k - 1); j >= 0; --j) {
            const double tau = taus[j];
            if (tau == 0) continue;
            double p = 0;
            for (int r = j + 1 + lane; r < m; r += 64) p += D[r + (long)j * m] * v[r];
            const double f = tau * (wave_sum(p) + v[j]);
            for (int r = j + 1 + lane; r < m; r += 64) v[r] -= f * D[r + (long)j * m];
            if (lane == 0) v[j] -= f;
        }
        const double sgn = D[c + (long)c * m] < 0 ? -1.0 : 1.0;
        for (int p = sc.part0; p < sc.part0 + sc.nparts; ++p) {
            const QrPart pt = parts[p];
            for (int rr = lane; rr < pt.len; rr += 64) {
                const double val = sgn * v[pt.off + rr];
                if (right) Q[pt.qdst + c + (long)rr * k] = val;   // part block k x len
                else Q[pt.qdst + rr + (long)c * pt.len] = val;    // part block len x k
            }
        }
    }
}

hipError_t launch_sector_qr(const QrSector* secs, int nlds, int nglobal, const QrPart* parts, const double* A, double* Q, double* R,
                            bool right, size_t lds_bytes, double* gwork, hipStream_t s) {
    if (nlds > 0) {
        if (lds_bytes > 64 * 1024) {  // beyond the default dynamic-LDS limit of a launch
            const hipError_t e = hipFuncSetAttribute((const void*)sector_qr_kernel<kQrWavesLds, false>,
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL((sector_qr_kernel<kQrWavesLds, false>), dim3(nlds), dim3(64 * kQrWavesLds), lds_bytes, s, secs, parts, A, Q, R,
                           right ? 1 : 0, gwork);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (nglobal > 0) {
        if (!gwork) return hipErrorInvalidValue;
        hipLaunchKernelGGL((sector_qr_kernel<kQrWavesGlobal, true>), dim3(nglobal), dim3(64 * kQrWavesGlobal), 0, s, secs + nlds, parts, A,
                           Q, R, right ? 1 : 0, gwork);
    }
    return hipGetLastError();
}

}  // namespace imps
