// Grouped block-GEMM kernels for gfx950 (CDNA4).
//
// One workgroup (4 waves, 256 lanes) owns one 64x64 output tile of one output block and sums all
// of that tile's segments C += alpha * op(A) * op(B) in order.  Operand panels are staged through
// LDS in K-chunks ([k][row] and [k][col] layouts, so the MFMA fragment reads are consecutive
// across the 16 lanes of a row group), zero-padded at ragged block edges.  Each wave computes a
// 32x32 quadrant as a 2x2 grid of 16x16 MFMA tiles:
//   FP64   : v_mfma_f64_16x16x4f64,  K-chunk 16
//   FP32   : v_mfma_f32_16x16x4f32,  K-chunk 16
//   BF16X3 : v_mfma_f32_16x16x32_bf16, K-chunk 32; each FP32 operand is split as hi = bf16(a),
//            lo = bf16(a - hi) and the product is hi*hi + hi*lo + lo*hi (FP32 accumulation).
// Fragment layouts (16x16 output tiles, measured on gfx950 with one-hot operands):
//   A(i,k) from lane i + 16*k, B(k,j) from lane j + 16*k (bf16 form: lane l holds the 8 consecutive
//   k = 8*(l/16) + e of row/column l%16);
//   result register r of lane l: FP32 and bf16 forms D(4*(l/16) + r, l%16),
//                                FP64 form D(4*r + l/16, l%16).
// Quadrants or 16x16 sub-tiles that lie wholly outside the block are skipped (wave-uniform).

#include <hip/hip_runtime.h>

#include "engine.hpp"

namespace imps {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));

constexpr int kThreads = 256;
constexpr int kPad = 4;  // LDS row padding (elements)

template <int PREC> struct Traits;
template <> struct Traits<IMPS_FP64> { using T = double; using Acc = d4_t; static constexpr int KC = 16; };
template <> struct Traits<IMPS_FP32> { using T = float; using Acc = f4_t; static constexpr int KC = 16; };
template <> struct Traits<IMPS_BF16X3> { using T = float; using Acc = f4_t; static constexpr int KC = 32; };

template <int PREC, typename TOut>
__global__ __launch_bounds__(kThreads) void grouped_gemm_kernel(const Task* __restrict__ tasks, const Seg* __restrict__ segs,
                                                                Slots slots, TOut* __restrict__ out, long out_stride) {
    using T = typename Traits<PREC>::T;
    using Acc = typename Traits<PREC>::Acc;
    constexpr int KC = Traits<PREC>::KC;
    __shared__ T sA[KC][kTile + kPad];
    __shared__ T sB[KC][kTile + kPad];

    const Task t = tasks[blockIdx.x];
    const int v = blockIdx.y;
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    const int mrem = t.M - t.m0, nrem = t.N - t.n0;  // valid extent of this tile
    const bool wave_active = wr < mrem && wc < nrem;

    Acc acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = Acc{0, 0, 0, 0};

    for (int si = 0; si < t.nseg; ++si) {
        const Seg sg = segs[t.seg0 + si];
        // The input-vector slot is always FP64 and is converted to the compute type while staging.
        const bool a_x = sg.a_slot == SLOT_X, b_x = sg.b_slot == SLOT_X;
        const T* Ap = static_cast<const T*>(slots.base[sg.a_slot]) + slots.stride[sg.a_slot] * v + sg.a_off;
        const T* Bp = static_cast<const T*>(slots.base[sg.b_slot]) + slots.stride[sg.b_slot] * v + sg.b_off;
        const double* Axp = static_cast<const double*>(slots.base[SLOT_X]) + slots.stride[SLOT_X] * v + sg.a_off;
        const double* Bxp = static_cast<const double*>(slots.base[SLOT_X]) + slots.stride[SLOT_X] * v + sg.b_off;
        const T alpha = static_cast<T>(sg.alpha);
        const int K = sg.K;
        for (int k0 = 0; k0 < K; k0 += KC) {
            // stage op(A)[m0 : m0+64, k0 : k0+KC] -> sA[k][i]
            for (int idx = tid; idx < kTile * KC; idx += kThreads) {
                int i, k;
                if (!sg.a_tr) { i = idx % kTile; k = idx / kTile; } else { k = idx % KC; i = idx / KC; }
                const int gi = t.m0 + i, gk = k0 + k;
                T val = T(0);
                if (gi < t.M && gk < K) {
                    const long e = sg.a_tr ? gk + (long)gi * sg.a_ld : gi + (long)gk * sg.a_ld;
                    val = a_x ? static_cast<T>(Axp[e]) : Ap[e];
                }
                sA[k][i] = val * alpha;
            }
            // stage op(B)[k0 : k0+KC, n0 : n0+64] -> sB[k][j]
            for (int idx = tid; idx < kTile * KC; idx += kThreads) {
                int j, k;
                if (!sg.b_tr) { k = idx % KC; j = idx / KC; } else { j = idx % kTile; k = idx / kTile; }
                const int gj = t.n0 + j, gk = k0 + k;
                T val = T(0);
                if (gj < t.N && gk < K) {
                    const long e = sg.b_tr ? gj + (long)gk * sg.b_ld : gk + (long)gj * sg.b_ld;
                    val = b_x ? static_cast<T>(Bxp[e]) : Bp[e];
                }
                sB[k][j] = val;
            }
            __syncthreads();
            if (wave_active) {
                const int kmax = min(KC, K - k0);
                if constexpr (PREC == IMPS_BF16X3) {
                    (void)kmax;
                    bf8_t ahi[2], alo[2], bhi[2], blo[2];
                    const int kb = 8 * (lane >> 4);
#pragma unroll
                    for (int a = 0; a < 2; ++a) {
#pragma unroll
                        for (int e = 0; e < 8; ++e) {
                            const float f = sA[kb + e][wr + a * 16 + (lane & 15)];
                            const __bf16 h = (__bf16)f;
                            ahi[a][e] = h;
                            alo[a][e] = (__bf16)(f - (float)h);
                        }
                    }
#pragma unroll
                    for (int b = 0; b < 2; ++b) {
#pragma unroll
                        for (int e = 0; e < 8; ++e) {
                            const float f = sB[kb + e][wc + b * 16 + (lane & 15)];
                            const __bf16 h = (__bf16)f;
                            bhi[b][e] = h;
                            blo[b][e] = (__bf16)(f - (float)h);
                        }
                    }
#pragma unroll
                    for (int a = 0; a < 2; ++a)
#pragma unroll
                        for (int b = 0; b < 2; ++b) {
                            if (wr + a * 16 >= mrem || wc + b * 16 >= nrem) continue;
                            acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo[a], bhi[b], acc[a][b], 0, 0, 0);
                            acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[a], blo[b], acc[a][b], 0, 0, 0);
                            acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[a], bhi[b], acc[a][b], 0, 0, 0);
                        }
                } else {
                    for (int kk = 0; kk < kmax; kk += 4) {
                        const int kr = kk + (lane >> 4);
                        T af[2], bf[2];
#pragma unroll
                        for (int a = 0; a < 2; ++a) af[a] = sA[kr][wr + a * 16 + (lane & 15)];
#pragma unroll
                        for (int b = 0; b < 2; ++b) bf[b] = sB[kr][wc + b * 16 + (lane & 15)];
#pragma unroll
                        for (int a = 0; a < 2; ++a)
#pragma unroll
                            for (int b = 0; b < 2; ++b) {
                                if (wr + a * 16 >= mrem || wc + b * 16 >= nrem) continue;
                                if constexpr (PREC == IMPS_FP64)
                                    acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[a], bf[b], acc[a][b], 0, 0, 0);
                                else
                                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[a], bf[b], acc[a][b], 0, 0, 0);
                            }
                    }
                }
            }
            __syncthreads();
        }
    }
    if (!wave_active) return;
    TOut* C = out + out_stride * v + t.c_off;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int col = t.n0 + wc + b * 16 + (lane & 15);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int rr = (PREC == IMPS_FP64) ? 4 * r + (lane >> 4) : 4 * (lane >> 4) + r;
                const int row = t.m0 + wr + a * 16 + rr;
                if (row < t.M && col < t.N) C[row + (long)col * t.c_ld] = static_cast<TOut>(acc[a][b][r]);
            }
        }
}

__global__ void d2f_kernel(const double* __restrict__ in, float* __restrict__ out, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) out[i] = (float)in[i];
}

hipError_t launch_d2f(const double* in, float* out, long n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const int threads = 256;
    long blocks = (n + threads - 1) / threads;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(d2f_kernel, dim3((unsigned)blocks), dim3(threads), 0, stream, in, out, n);
    return hipGetLastError();
}

template <int PREC>
static hipError_t launch_prec(const DevStage& st, const Slots& slots, void* out, long out_stride, bool out_double, int nvec, hipStream_t stream) {
    using T = typename Traits<PREC>::T;
    dim3 grid((unsigned)st.ntasks, (unsigned)nvec);
    if (out_double)
        hipLaunchKernelGGL((grouped_gemm_kernel<PREC, double>), grid, dim3(kThreads), 0, stream, st.tasks, st.segs, slots, (double*)out, out_stride);
    else
        hipLaunchKernelGGL((grouped_gemm_kernel<PREC, T>), grid, dim3(kThreads), 0, stream, st.tasks, st.segs, slots, (T*)out, out_stride);
    return hipGetLastError();
}

hipError_t launch_stage(int prec, const DevStage& st, const Slots& slots, void* out, long out_stride, bool out_double, int nvec,
                        hipStream_t stream) {
    if (st.ntasks == 0 || nvec == 0) return hipSuccess;
    switch (prec) {
        case IMPS_FP64: return launch_prec<IMPS_FP64>(st, slots, out, out_stride, out_double, nvec, stream);
        case IMPS_FP32: return launch_prec<IMPS_FP32>(st, slots, out, out_stride, out_double, nvec, stream);
        case IMPS_BF16X3: return launch_prec<IMPS_BF16X3>(st, slots, out, out_stride, out_double, nvec, stream);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace imps
