// Grouped block-GEMM kernels for gfx950 (CDNA4).
//
// Work decomposition.  Output blocks are cut into tiles of at most 32x32.  One workgroup of four
// waves owns one tile for one vector.  The tile's work, the sum over its segments
// C += alpha * op(A) * op(B), is cut into K-chunks that are dealt round-robin to the four waves;
// each wave accumulates its share in registers and the four partial tiles are summed through LDS
// in fixed wave order (deterministic, no atomics).  Splitting the chunk sequence shortens each
// wave's chain of dependent loads and multiplies the loads in flight.
//
// Each wave holds the tile as a 2x2 grid of 16x16 MFMA accumulators.  Operand fragments are loaded
// straight from global memory into registers in the MFMA lane layout (no LDS staging, no
// barriers); lanes outside a ragged block edge load zeros.  Row/column groups of 16 wholly outside
// the block are skipped, loads included (wave-uniform branches).
//   FP64   : v_mfma_f64_16x16x4f64      (K step 4)
//   FP32   : v_mfma_f32_16x16x4f32      (K step 4)
//   BF16X3 : v_mfma_f32_16x16x32_bf16   (K step 32); each FP32 operand a = hi + lo with
//            hi = bf16(a), lo = bf16(a - hi); a*b ~ lo*hi + hi*lo + hi*hi, FP32 accumulation.
// Fragment layouts (16x16 output tiles, measured on gfx950 with one-hot operands):
//   A(i,k) from lane i + 16*k, B(k,j) from lane j + 16*k (bf16 form: lane l holds the 8 consecutive
//   k = 8*(l/16) + e of row/column l%16);
//   result register r of lane l: FP32 and bf16 forms D(4*(l/16) + r, l%16),
//                                FP64 form D(4*r + l/16, l%16).
// The input-vector slot is FP64 in memory and is converted to the compute type on load.

#include <hip/hip_runtime.h>

#include "engine.hpp"

namespace imps {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));

constexpr int kWaves = 4;  // waves per workgroup; they split one tile's K-chunks

template <int PREC> struct Traits;
template <> struct Traits<IMPS_FP64> { using T = double; using Acc = d4_t; };
template <> struct Traits<IMPS_FP32> { using T = float; using Acc = f4_t; };
template <> struct Traits<IMPS_BF16X3> { using T = float; using Acc = f4_t; };

// Load of element e of an operand slot: FP64 input-vector slot (converted) or compute-type slot.
template <typename T, bool X>
__device__ __forceinline__ T slot_load(const void* base, long e) {
    if constexpr (X) return static_cast<T>(static_cast<const double*>(base)[e]);
    else return static_cast<const T*>(base)[e];
}

// Element (i, k) of op(S), S column-major with leading dimension ld (tr: op(S) = S^T), read
// unconditionally from a clamped in-block address and zeroed by a select when out of range, so
// that all loads of a chunk issue back to back.
template <typename T, bool X>
__device__ __forceinline__ T load_op(const void* base, long off, int i, int k, int rows, int K, int ld, int tr) {
    const int ic = min(i, rows - 1), kc = min(k, K - 1);
    const long e = off + (tr ? (long)kc + (long)ic * ld : (long)ic + (long)kc * ld);
    const T v = slot_load<T, X>(base, e);
    return (i < rows && k < K) ? v : T(0);
}

// N consecutive elements e, e+1, ... of a slot (k-contiguous operand runs).  Runs that end inside
// the block are read with wide loads (4-byte alignment suffices on gfx950); the last run of a
// ragged block is read element-wise from clamped addresses.  Entries past `n_ok` are zero.
typedef float f4u_t __attribute__((ext_vector_type(4), aligned(4)));
typedef double d2u_t __attribute__((ext_vector_type(2), aligned(8)));
template <int N, typename T, bool X>
__device__ __forceinline__ void load_run(T (&dst)[N], const void* base, long e, int n_ok) {
    if (n_ok >= N) {
        if constexpr (X) {
            const double* p = static_cast<const double*>(base) + e;
#pragma unroll
            for (int h = 0; h < N / 2; ++h) {
                const d2u_t w = *reinterpret_cast<const d2u_t*>(p + 2 * h);
                dst[2 * h] = static_cast<T>(w[0]);
                dst[2 * h + 1] = static_cast<T>(w[1]);
            }
        } else if constexpr (sizeof(T) == 4) {
            const float* p = static_cast<const float*>(base) + e;
#pragma unroll
            for (int h = 0; h < N / 4; ++h) {
                const f4u_t w = *reinterpret_cast<const f4u_t*>(p + 4 * h);
#pragma unroll
                for (int c = 0; c < 4; ++c) dst[4 * h + c] = w[c];
            }
        } else {
            const double* p = static_cast<const double*>(base) + e;
#pragma unroll
            for (int h = 0; h < N / 2; ++h) {
                const d2u_t w = *reinterpret_cast<const d2u_t*>(p + 2 * h);
                dst[2 * h] = w[0];
                dst[2 * h + 1] = w[1];
            }
        }
    } else {
        const int last = n_ok > 0 ? n_ok - 1 : 0;
#pragma unroll
        for (int c = 0; c < N; ++c) {
            const T x = slot_load<T, X>(base, e + min(c, last));
            dst[c] = c < n_ok ? x : T(0);
        }
    }
}

// Fragment of op(S): rows/cols `i` (one per lane), N consecutive k starting at k (one run per
// lane).  kcontig: op(S) is stored with k fastest (S^T of a column-major block), so the run is
// contiguous; otherwise its elements are ld apart.
template <int N, typename T, bool X>
__device__ __forceinline__ void load_frag(T (&dst)[N], const void* base, long off, int i, int k, int rows, int K, int ld, bool kcontig) {
    if (kcontig) {
        const int ic = min(i, rows - 1);
        const int n_ok = i < rows ? min(N, K - k) : 0;
        load_run<N, T, X>(dst, base, off + (long)ic * ld + min(k, K - 1), n_ok);
    } else {
#pragma unroll
        for (int c = 0; c < N; ++c) dst[c] = load_op<T, X>(base, off, i, k + c, rows, K, ld, 0);
    }
}

template <int PREC> using Acc_t = typename Traits<PREC>::Acc;

// Contribution of one segment to the wave's 2x2 grid of 16x16 accumulators.  The segment's K
// range is cut into chunks; chunk number `g` (counted over all segments of the tile) belongs to
// wave g % kWaves, so the waves of a workgroup split the tile's work evenly.  Loads are issued
// only for the 16-row / 16-column groups inside the block (na, nb wave-uniform).  Each lane reads
// runs of consecutive k: 8 for the bf16 form; 4 for the 16x16x4 forms, whose q-th MFMA takes
// k = k0 + 4*(lane/16) + q (a permutation of the K sum, identical for A and B).  AX / BX: operand
// A / B is the FP64 input vector.
template <int PREC, bool AX, bool BX>
__device__ __forceinline__ void segment(Acc_t<PREC> (&acc)[2][2], int na, int nb, int& g, int wave, const Task& t, const Seg& sg,
                                        const void* Ab, long aoff, const void* Bb, long boff, int li, int lk) {
    using T = typename Traits<PREC>::T;
    const T alpha = static_cast<T>(sg.alpha);
    const int K = sg.K;
    // op(A)(i,k) is k-contiguous when A is used transposed; op(B)(k,j), read as op(B)^T(j,k), when B is not
    const bool a_kc = sg.a_tr != 0, b_kc = sg.b_tr == 0;
    constexpr int KC = PREC == IMPS_BF16X3 ? 32 : 16;
    constexpr int R = PREC == IMPS_BF16X3 ? 8 : 4;  // run length per lane
    for (int k0 = 0; k0 < K; k0 += KC, ++g) {
        if (g % kWaves != wave) continue;
        const int kb = k0 + R * lk;
        T fa[2][R], fb[2][R];
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            if (a >= na) break;
            load_frag<R, T, AX>(fa[a], Ab, aoff, t.m0 + 16 * a + li, kb, t.M, K, sg.a_ld, a_kc);
        }
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            if (b >= nb) break;
            load_frag<R, T, BX>(fb[b], Bb, boff, t.n0 + 16 * b + li, kb, t.N, K, sg.b_ld, b_kc);
        }
        if constexpr (PREC == IMPS_BF16X3) {
            bf8_t ahi[2], alo[2], bhi[2], blo[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                if (a >= na) break;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const float f = alpha * fa[a][e];
                    const __bf16 h = (__bf16)f;
                    ahi[a][e] = h;
                    alo[a][e] = (__bf16)(f - (float)h);
                }
            }
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                if (b >= nb) break;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const float f = fb[b][e];
                    const __bf16 h = (__bf16)f;
                    bhi[b][e] = h;
                    blo[b][e] = (__bf16)(f - (float)h);
                }
            }
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                if (a >= na) break;
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    if (b >= nb) break;
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo[a], bhi[b], acc[a][b], 0, 0, 0);
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[a], blo[b], acc[a][b], 0, 0, 0);
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[a], bhi[b], acc[a][b], 0, 0, 0);
                }
            }
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (k0 + 4 * q >= K && k0 + q >= K) break;
#pragma unroll
                for (int a = 0; a < 2; ++a) {
                    if (a >= na) break;
#pragma unroll
                    for (int b = 0; b < 2; ++b) {
                        if (b >= nb) break;
                        if constexpr (PREC == IMPS_FP64)
                            acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(alpha * fa[a][q], fb[b][q], acc[a][b], 0, 0, 0);
                        else
                            acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(alpha * fa[a][q], fb[b][q], acc[a][b], 0, 0, 0);
                    }
                }
            }
        }
    }
}

template <int PREC, typename TOut>
__global__ __launch_bounds__(64 * kWaves) void grouped_gemm_kernel(const Task* __restrict__ tasks, const Seg* __restrict__ segs,
                                                                   Slots slots, TOut* __restrict__ out, long out_stride) {
    using T = typename Traits<PREC>::T;
    using Acc = typename Traits<PREC>::Acc;
    __shared__ Acc part[kWaves][4][64];  // [wave][sub-tile][lane]
    const Task t = tasks[blockIdx.x];
    const int v = blockIdx.y;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & 15, lk = lane >> 4;
    const int na = min(2, (t.M - t.m0 + 15) >> 4), nb = min(2, (t.N - t.n0 + 15) >> 4);

    Acc acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = Acc{0, 0, 0, 0};

    int g = 0;
    for (int si = 0; si < t.nseg; ++si) {
        const Seg sg = segs[t.seg0 + si];
        const void* Ab = slots.base[sg.a_slot];
        const void* Bb = slots.base[sg.b_slot];
        const long aoff = sg.a_off + slots.stride[sg.a_slot] * v;
        const long boff = sg.b_off + slots.stride[sg.b_slot] * v;
        const bool ax = sg.a_slot == SLOT_X, bx = sg.b_slot == SLOT_X;
        if (ax && bx) segment<PREC, true, true>(acc, na, nb, g, wave, t, sg, Ab, aoff, Bb, boff, li, lk);
        else if (ax) segment<PREC, true, false>(acc, na, nb, g, wave, t, sg, Ab, aoff, Bb, boff, li, lk);
        else if (bx) segment<PREC, false, true>(acc, na, nb, g, wave, t, sg, Ab, aoff, Bb, boff, li, lk);
        else segment<PREC, false, false>(acc, na, nb, g, wave, t, sg, Ab, aoff, Bb, boff, li, lk);
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) part[wave][2 * a + b][lane] = acc[a][b];
    __syncthreads();
    // wave w reduces sub-tile w over the waves in fixed order and writes it
    const int a = wave >> 1, b = wave & 1;
    if (a >= na || b >= nb) return;
    Acc sum = part[0][wave][lane];
#pragma unroll
    for (int w = 1; w < kWaves; ++w) sum += part[w][wave][lane];
    TOut* C = out + out_stride * v + t.c_off;
    const int col = t.n0 + 16 * b + li;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int rr = (PREC == IMPS_FP64) ? 4 * r + lk : 4 * lk + r;
        const int row = t.m0 + 16 * a + rr;
        if (row < t.M && col < t.N) C[row + (long)col * t.c_ld] = static_cast<TOut>(sum[r]);
    }
    (void)sizeof(T);
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
static hipError_t launch_prec(const DevStage& st, const Slots& slots, void* out, long out_stride, bool out_double, int nvec,
                              hipStream_t stream) {
    using T = typename Traits<PREC>::T;
    dim3 grid((unsigned)st.ntasks, (unsigned)nvec);
    dim3 block(64 * kWaves);
    if (out_double)
        hipLaunchKernelGGL((grouped_gemm_kernel<PREC, double>), grid, block, 0, stream, st.tasks, st.segs, slots, (double*)out, out_stride);
    else
        hipLaunchKernelGGL((grouped_gemm_kernel<PREC, T>), grid, block, 0, stream, st.tasks, st.segs, slots, (T*)out, out_stride);
    return hipGetLastError();
}

hipError_t launch_stage(int prec, const DevStage& st, const Slots& slots, void* out, long out_stride, bool out_double, int nvec,
                        hipStream_t stream) {
    if (st.ntasks == 0 || nvec == 0) return hipSuccess;
    if (nvec > 65535) return hipErrorInvalidValue;
    switch (prec) {
        case IMPS_FP64: return launch_prec<IMPS_FP64>(st, slots, out, out_stride, out_double, nvec, stream);
        case IMPS_FP32: return launch_prec<IMPS_FP32>(st, slots, out, out_stride, out_double, nvec, stream);
        case IMPS_BF16X3: return launch_prec<IMPS_BF16X3>(st, slots, out, out_stride, out_double, nvec, stream);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace imps
