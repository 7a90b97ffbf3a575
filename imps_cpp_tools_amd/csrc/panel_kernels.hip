// Panel grouped-GEMM kernels for gfx950 (CDNA4): operator blocks pre-packed in MFMA register
// layout (panels), the vector/workspace factor read raw in FP64 through buffer loads.
//
// Work decomposition.  A workgroup has four waves.  With K-split S (1, 2 or 4) it serves 4/S
// vectors of one output tile (<= 32 x 32); the S waves of one vector take the tile's chunks
// round-robin and their partial tiles are summed through LDS in fixed wave order (deterministic,
// no atomics).  Each wave holds the tile as 2 x 2 MFMA sub-tiles of 16 x 16 and computes the
// transposed tile C^T = op(B)^T op(A)^T, whose MFMA result layout gives each lane consecutive
// output rows, so the column-major output is written with coalesced stores.
//
// Per chunk (one segment, KC consecutive k) a wave issues, unconditionally and in a fixed count,
// two panel-fragment loads (16-byte coalesced per lane; 32 bytes for FP64) and 2 x RUN raw FP64
// loads.  All are buffer loads: an element outside the block (or an absent second panel row
// group) gets an offset beyond the buffer, which returns zero without a memory access, so ragged
// and small blocks move only the cache lines they occupy.  Raw factors stored k-contiguous are
// read with consecutive lanes on consecutive k and re-laid through LDS.  The loads of the next
// chunk are issued before the MFMAs of the current one (double buffer).
//
// MFMA forms (fragment layouts measured on gfx950 with one-hot operands):
//   FP64   v_mfma_f64_16x16x4f64     KC 16, RUN 4; lane k-run permuted (k = 4 (l/16) + q)
//   FP32   v_mfma_f32_16x16x4f32     KC 16, RUN 4; same permutation, identical for both factors
//   BF16X3 v_mfma_f32_16x16x32_bf16  KC 32, RUN 8; a*b ~ hi*lo + lo*hi + hi*hi, FP32 accumulation
//   result register r of lane l: FP32/BF16 D(4 (l/16) + r, l % 16); FP64 D(4 r + l/16, l % 16).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "engine.hpp"
#include "panel.hpp"

namespace imps {

namespace {

typedef double pd4_t __attribute__((ext_vector_type(4)));
typedef float pf4_t __attribute__((ext_vector_type(4)));
typedef __bf16 pbf8_t __attribute__((ext_vector_type(8)));
typedef int pi2_t __attribute__((ext_vector_type(2)));
typedef double pd2_t __attribute__((ext_vector_type(2)));
typedef int pi4_t __attribute__((ext_vector_type(4)));

constexpr int kPWaves = 4;
constexpr int kBufDword3 = 0x00020000;  // raw buffer resource word 3 for gfx9-family (32-bit data format)

template <int PREC> struct PTraits;
template <> struct PTraits<IMPS_FP64> {
    static constexpr int KC = 16, RUN = 4;
    using Acc = pd4_t;
    struct Pan { pd4_t v; };
};
template <> struct PTraits<IMPS_FP32> {
    static constexpr int KC = 16, RUN = 4;
    using Acc = pf4_t;
    struct Pan { pf4_t v; };
};
template <> struct PTraits<IMPS_BF16X3> {
    static constexpr int KC = 32, RUN = 8;
    using Acc = pf4_t;
    struct Pan { pbf8_t hi, lo; };
};

__device__ __forceinline__ void split_bf16(float f, __bf16& hi, __bf16& lo) {
    hi = (__bf16)f;
    lo = (__bf16)(f - (float)hi);
}

// ---------------------------------------------------------------------------------------------
// Panel packing: one workgroup per panel; thread e handles (fragment e / 64, lane e % 64).
template <int PREC>
__global__ __launch_bounds__(256) void panelize_kernel(const PanelDesc* __restrict__ panels, const double* __restrict__ op,
                                                       void* __restrict__ store) {
    using TR = PTraits<PREC>;
    const PanelDesc pd = panels[blockIdx.x];
    const int nf = pd.nrg * pd.nkc;
    for (int e = threadIdx.x; e < nf * 64; e += blockDim.x) {
        const int f = e >> 6, lane = e & 63;
        const int rg = f / pd.nkc, kc = f - rg * pd.nkc;
        const int r = 16 * rg + (lane & 15);
        const int kb = TR::KC * kc + TR::RUN * (lane >> 4);
        double v[TR::RUN];
#pragma unroll
        for (int q = 0; q < TR::RUN; ++q) {
            const int k = kb + q;
            double x = 0.0;
            if (r < pd.R && k < pd.K) x = pd.alpha * op[pd.src + (pd.tr ? (long)k + (long)r * pd.ld : (long)r + (long)k * pd.ld)];
            v[q] = x;
        }
        const long frag = pd.frag0 + f;
        if constexpr (PREC == IMPS_FP64) {
            pd4_t w = {v[0], v[1], v[2], v[3]};
            reinterpret_cast<pd4_t*>(store)[frag * 64 + lane] = w;
        } else if constexpr (PREC == IMPS_FP32) {
            pf4_t w = {(float)v[0], (float)v[1], (float)v[2], (float)v[3]};
            reinterpret_cast<pf4_t*>(store)[frag * 64 + lane] = w;
        } else {
            pbf8_t hi, lo;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                __bf16 h, l;
                split_bf16((float)v[q], h, l);
                hi[q] = h;
                lo[q] = l;
            }
            pbf8_t* base = reinterpret_cast<pbf8_t*>(store) + frag * 128;
            base[lane] = hi;
            base[64 + lane] = lo;
        }
    }
}

// ---------------------------------------------------------------------------------------------
template <int PREC>
struct Frag {
    typename PTraits<PREC>::Pan p[2];
    double r[2][PTraits<PREC>::RUN];
};

__device__ __forceinline__ double buf_load_f64(__amdgpu_buffer_rsrc_t rs, unsigned voff) {
    const pi2_t w = __builtin_amdgcn_raw_buffer_load_b64(rs, voff, 0, 0);
    return __builtin_bit_cast(double, w);
}

__device__ __forceinline__ pi4_t buf_load_b128(__amdgpu_buffer_rsrc_t rs, unsigned voff) {
    return __builtin_amdgcn_raw_buffer_load_b128(rs, voff, 0, 0);
}

constexpr unsigned kOOB = 0x80000000u;  // beyond every buffer: the load returns zero, no memory access

// Issue every load of one chunk (fixed count, no branches).  Raw element t of a lane:
//   direct (fragment layout): idx = 16 (t / RUN) + lane % 16, k = RUN (lane / 16) + t % RUN
//   staged (coalesced):       idx = (64 / KC) t + lane / KC,   k = lane % KC   (k-contiguous raw)
// Elements with idx >= nidx (outside the block) or k >= krem read as zero through kOOB, as does
// the second panel row group when the panel has none.
template <int PREC>
__device__ __forceinline__ void load_chunk(Frag<PREC>& F, const PChunk& c, __amdgpu_buffer_rsrc_t rp, __amdgpu_buffer_rsrc_t rx,
                                           __amdgpu_buffer_rsrc_t rw, int nidx, int li, int lk, int lane) {
    using TR = PTraits<PREC>;
    constexpr int KC = TR::KC, RUN = TR::RUN;
    constexpr unsigned FB = PREC == IMPS_FP32 ? 64 * 16 : 64 * 32;
    const unsigned p0 = (unsigned)c.pan * FB;
    const unsigned p1 = c.pan_sub >= 0 ? (unsigned)(c.pan + c.pan_sub) * FB : kOOB;
    if constexpr (PREC == IMPS_BF16X3) {
        F.p[0].hi = __builtin_bit_cast(pbf8_t, buf_load_b128(rp, p0 + 16u * lane));
        F.p[0].lo = __builtin_bit_cast(pbf8_t, buf_load_b128(rp, p0 + 1024u + 16u * lane));
        F.p[1].hi = __builtin_bit_cast(pbf8_t, buf_load_b128(rp, p1 + 16u * lane));
        F.p[1].lo = __builtin_bit_cast(pbf8_t, buf_load_b128(rp, p1 + 1024u + 16u * lane));
    } else if constexpr (PREC == IMPS_FP32) {
        F.p[0].v = __builtin_bit_cast(pf4_t, buf_load_b128(rp, p0 + 16u * lane));
        F.p[1].v = __builtin_bit_cast(pf4_t, buf_load_b128(rp, p1 + 16u * lane));
    } else {
        const pd2_t a0 = __builtin_bit_cast(pd2_t, buf_load_b128(rp, p0 + 32u * lane));
        const pd2_t a1 = __builtin_bit_cast(pd2_t, buf_load_b128(rp, p0 + 32u * lane + 16u));
        const pd2_t b0 = __builtin_bit_cast(pd2_t, buf_load_b128(rp, p1 + 32u * lane));
        const pd2_t b1 = __builtin_bit_cast(pd2_t, buf_load_b128(rp, p1 + 32u * lane + 16u));
        F.p[0].v = pd4_t{a0[0], a0[1], a1[0], a1[1]};
        F.p[1].v = pd4_t{b0[0], b0[1], b1[0], b1[1]};
    }
    const __amdgpu_buffer_rsrc_t rs = (c.flags & PCF_RAW_W) ? rw : rx;
    const bool stage = (c.flags & PCF_STAGE) != 0;
    const long lbase = stage ? (long)(lane % KC) + (long)(lane / KC) * c.s_idx : (long)li * c.s_idx + (long)(RUN * lk) * c.s_k;
    const unsigned v0 = (unsigned)(c.raw_off + lbase) * 8u;
    // per-lane limits: direct layout k = RUN lk + q, idx = 16 s + li; staged k = lane % KC, idx = (64/KC) t + lane / KC
    const int kq = stage ? (lane % KC < c.krem ? RUN : 0) : c.krem - RUN * lk;  // q < kq valid
    const int i0 = stage ? lane / KC : li;
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int q = 0; q < RUN; ++q) {
            const int t = s * RUN + q;
            const int idx = i0 + (stage ? (64 / KC) * t : 16 * s);
            const bool ok = idx < nidx && (stage ? kq > 0 : q < kq);
            const unsigned inc = stage ? (unsigned)((64 / KC) * t * c.s_idx) * 8u : (unsigned)(16 * s * c.s_idx + q * c.s_k) * 8u;
            F.r[s][q] = buf_load_f64(rs, ok ? v0 + inc : kOOB);
        }
}

// Staged raw values (see load_chunk) pass through the wave's LDS region (32 rows x KC + 2 pad
// doubles) and come back in the fragment layout; the same wave writes and reads (LDS operations
// of one wave complete in order), so no barrier is needed.
template <int PREC>
__device__ __forceinline__ void restage(double (&r)[2][PTraits<PREC>::RUN], double* __restrict__ lds, int li, int lk, int lane) {
    constexpr int KC = PTraits<PREC>::KC, RUN = PTraits<PREC>::RUN, ROW = KC + 2;
    double* w = lds + (lane / KC) * ROW + (lane % KC);
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int q = 0; q < RUN; ++q) w[(64 / KC) * (s * RUN + q) * ROW] = r[s][q];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const double* rd = lds + (16 * s + li) * ROW + RUN * lk;
#pragma unroll
        for (int q = 0; q < RUN; q += 2) {
            const pd2_t v = *reinterpret_cast<const pd2_t*>(rd + q);
            r[s][q] = v[0];
            r[s][q + 1] = v[1];
        }
    }
}

// acc[a][b] += (col-side fragment b)(row-side fragment a), the transposed sub-tile (a, b).
template <int PREC>
__device__ __forceinline__ void compute_chunk(typename PTraits<PREC>::Acc (&acc)[2][2], const Frag<PREC>& F, bool panel_is_a, int na, int nb) {
    using TR = PTraits<PREC>;
    if constexpr (PREC == IMPS_BF16X3) {
        pbf8_t rh[2], rl[2];
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                __bf16 h, l;
                split_bf16((float)F.r[s][e], h, l);
                rh[s][e] = h;
                rl[s][e] = l;
            }
        if (panel_is_a) {  // rows: panel (a), columns: raw (b)
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                if (a >= na) break;
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    if (b >= nb) break;
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(rh[b], F.p[a].lo, acc[a][b], 0, 0, 0);
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(rl[b], F.p[a].hi, acc[a][b], 0, 0, 0);
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(rh[b], F.p[a].hi, acc[a][b], 0, 0, 0);
                }
            }
        } else {           // rows: raw (a), columns: panel (b)
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                if (a >= na) break;
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    if (b >= nb) break;
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(F.p[b].hi, rl[a], acc[a][b], 0, 0, 0);
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(F.p[b].lo, rh[a], acc[a][b], 0, 0, 0);
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(F.p[b].hi, rh[a], acc[a][b], 0, 0, 0);
                }
            }
        }
    } else {
        using T = typename std::conditional<PREC == IMPS_FP64, double, float>::type;
        T rv[2][4];
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int q = 0; q < 4; ++q) rv[s][q] = (T)F.r[s][q];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                if (a >= na) break;
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    if (b >= nb) break;
                    const T col = panel_is_a ? rv[b][q] : (T)F.p[b].v[q];
                    const T row = panel_is_a ? (T)F.p[a].v[q] : rv[a][q];
                    if constexpr (PREC == IMPS_FP64)
                        acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(col, row, acc[a][b], 0, 0, 0);
                    else
                        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(col, row, acc[a][b], 0, 0, 0);
                }
            }
        }
    }
    (void)sizeof(TR);
}

// Store the transposed accumulator sub-tile (a, b) into the column-major output block.
template <int PREC>
__device__ __forceinline__ void store_sub(const typename PTraits<PREC>::Acc& v, double* __restrict__ C, const PTask& t, int a, int b,
                                          int li, int lk) {
    const int i = t.m0 + 16 * a + li;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int j = t.n0 + 16 * b + (PREC == IMPS_FP64 ? 4 * r + lk : 4 * lk + r);
        if (i < t.M && j < t.N) C[i + (long)j * t.c_ld] = (double)v[r];
    }
}

template <int PREC, int S>
__global__ __launch_bounds__(64 * kPWaves) void panel_gemm_kernel(const PTask* __restrict__ tasks, const PChunk* __restrict__ chunks,
                                                                  const char* __restrict__ pstore, long pstore_bytes, const double* X, long x_stride,
                                                                  long x_len, const double* W, long w_stride, long w_len,
                                                                  double* __restrict__ out, long out_stride, int nvec, int nvg) {
    using Acc = typename PTraits<PREC>::Acc;
    constexpr int V = kPWaves / S;  // vectors per workgroup
    constexpr int KC = PTraits<PREC>::KC;
    constexpr int kStage = 32 * (KC + 2);                     // doubles of one wave's staging region
    constexpr int kPart = S > 1 ? 4 * 64 * (int)sizeof(Acc) / 8 : 0;  // doubles of one wave's partial tile
    __shared__ double lds[kPWaves * (kStage > kPart ? kStage : kPart)];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int li = lane & 15, lk = lane >> 4;
    const int ti = blockIdx.x / nvg, vg = blockIdx.x - ti * nvg;
    const int vl = wave / S, part = wave - vl * S;
    const int v = vg * V + vl;
    const bool active = v < nvec;
    const PTask t = tasks[ti];
    const int na = min(2, (t.M - t.m0 + 15) >> 4), nb = min(2, (t.N - t.n0 + 15) >> 4);

    Acc acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = Acc{0, 0, 0, 0};

    if (active && part < t.nch) {
        const int vv = v;
        const __amdgpu_buffer_rsrc_t rx = __builtin_amdgcn_make_buffer_rsrc((void*)(X + x_stride * vv), 0, (int)(x_len * 8), kBufDword3);
        const __amdgpu_buffer_rsrc_t rw = __builtin_amdgcn_make_buffer_rsrc((void*)(W + w_stride * vv), 0, (int)(w_len * 8), kBufDword3);
        const __amdgpu_buffer_rsrc_t rp = __builtin_amdgcn_make_buffer_rsrc((void*)pstore, 0, (int)pstore_bytes, kBufDword3);
        const int nrow = t.M - t.m0, ncol = t.N - t.n0;  // raw-side extent: columns if the panel is A, rows otherwise
        const int cend = t.ch0 + t.nch;
        int c = t.ch0 + part;
        double* my = lds + wave * kStage;
        Frag<PREC> F0, F1;
        PChunk d0 = chunks[c];
        load_chunk<PREC>(F0, d0, rp, rx, rw, (d0.flags & PCF_PANEL_IS_A) ? ncol : nrow, li, lk, lane);
        while (true) {
            // prefetch the next chunk (the current one again at the end: same count of loads)
            const int c1 = c + S < cend ? c + S : c;
            const PChunk d1 = chunks[c1];
            load_chunk<PREC>(F1, d1, rp, rx, rw, (d1.flags & PCF_PANEL_IS_A) ? ncol : nrow, li, lk, lane);
            if (d0.flags & PCF_STAGE) restage<PREC>(F0.r, my, li, lk, lane);
            compute_chunk<PREC>(acc, F0, (d0.flags & PCF_PANEL_IS_A) != 0, na, nb);
            if (c + S >= cend) break;
            c = c1;
            const int c2 = c + S < cend ? c + S : c;
            d0 = chunks[c2];
            load_chunk<PREC>(F0, d0, rp, rx, rw, (d0.flags & PCF_PANEL_IS_A) ? ncol : nrow, li, lk, lane);
            if (d1.flags & PCF_STAGE) restage<PREC>(F1.r, my, li, lk, lane);
            compute_chunk<PREC>(acc, F1, (d1.flags & PCF_PANEL_IS_A) != 0, na, nb);
            if (c + S >= cend) break;
            c = c2;
        }
    }

    double* C = out + out_stride * v + t.c_off;
    if constexpr (S == 1) {
        if (!active) return;
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
                if (a < na && b < nb) store_sub<PREC>(acc[a][b], C, t, a, b, li, lk);
    } else {
        // the staging regions are reused for the partial tiles once every wave has left its loop
        Acc (*partial)[4][64] = reinterpret_cast<Acc (*)[4][64]>(lds);
        __syncthreads();
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) partial[wave][2 * a + b][lane] = acc[a][b];
        __syncthreads();
        if (!active) return;
        // the S waves of vector vl reduce 4/S sub-tiles each, summing partials in wave order
#pragma unroll
        for (int u = 0; u < 4 / S; ++u) {
            const int sub = part * (4 / S) + u;
            const int a = sub >> 1, b = sub & 1;
            if (a >= na || b >= nb) continue;
            Acc sum = partial[vl * S][sub][lane];
#pragma unroll
            for (int p = 1; p < S; ++p) sum += partial[vl * S + p][sub][lane];
            store_sub<PREC>(sum, C, t, a, b, li, lk);
        }
    }
}

template <int PREC, int S>
hipError_t launch_ps(const DevPStage& st, const void* pstore, long pstore_bytes, const double* X, long x_stride, long x_len, const double* W, long w_stride,
                     long w_len, double* out, long out_stride, int nvec, hipStream_t stream) {
    constexpr int V = kPWaves / S;
    const int nvg = (nvec + V - 1) / V;
    const long nblk = (long)st.ntasks * nvg;
    if (nblk > 0x7fffffffL) return hipErrorInvalidValue;
    hipLaunchKernelGGL((panel_gemm_kernel<PREC, S>), dim3((unsigned)nblk), dim3(64 * kPWaves), 0, stream, st.tasks, st.chunks,
                       (const char*)pstore, pstore_bytes, X, x_stride, x_len, W, w_stride, w_len, out, out_stride, nvec, nvg);
    return hipGetLastError();
}

template <int PREC>
hipError_t launch_p(const DevPStage& st, int ksplit, const void* pstore, long pstore_bytes, const double* X, long x_stride, long x_len, const double* W,
                    long w_stride, long w_len, double* out, long out_stride, int nvec, hipStream_t stream) {
    switch (ksplit) {
        case 1: return launch_ps<PREC, 1>(st, pstore, pstore_bytes, X, x_stride, x_len, W, w_stride, w_len, out, out_stride, nvec, stream);
        case 2: return launch_ps<PREC, 2>(st, pstore, pstore_bytes, X, x_stride, x_len, W, w_stride, w_len, out, out_stride, nvec, stream);
        case 4: return launch_ps<PREC, 4>(st, pstore, pstore_bytes, X, x_stride, x_len, W, w_stride, w_len, out, out_stride, nvec, stream);
        default: return hipErrorInvalidValue;
    }
}

__global__ void zero_kernel(double* __restrict__ p, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = 0.0;
}

}  // namespace

hipError_t launch_zero(double* p, long n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    long blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(zero_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, p, n);
    return hipGetLastError();
}

hipError_t launch_panelize(int prec, const PanelDesc* d_panels, int npanels, const double* op, void* store, hipStream_t stream) {
    if (npanels == 0) return hipSuccess;
    switch (prec) {
        case IMPS_FP64: hipLaunchKernelGGL(panelize_kernel<IMPS_FP64>, dim3(npanels), dim3(256), 0, stream, d_panels, op, store); break;
        case IMPS_FP32: hipLaunchKernelGGL(panelize_kernel<IMPS_FP32>, dim3(npanels), dim3(256), 0, stream, d_panels, op, store); break;
        case IMPS_BF16X3: hipLaunchKernelGGL(panelize_kernel<IMPS_BF16X3>, dim3(npanels), dim3(256), 0, stream, d_panels, op, store); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_panel_stage(int prec, const DevPStage& st, int ksplit, const void* pstore, long pstore_bytes, const double* X, long x_stride, long x_len,
                              const double* W, long w_stride, long w_len, double* out, long out_stride, int nvec, hipStream_t stream) {
    if (st.ntasks == 0 || nvec == 0) return hipSuccess;
    switch (prec) {
        case IMPS_FP64: return launch_p<IMPS_FP64>(st, ksplit, pstore, pstore_bytes, X, x_stride, x_len, W, w_stride, w_len, out, out_stride, nvec, stream);
        case IMPS_FP32: return launch_p<IMPS_FP32>(st, ksplit, pstore, pstore_bytes, X, x_stride, x_len, W, w_stride, w_len, out, out_stride, nvec, stream);
        case IMPS_BF16X3:
            return launch_p<IMPS_BF16X3>(st, ksplit, pstore, pstore_bytes, X, x_stride, x_len, W, w_stride, w_len, out, out_stride, nvec, stream);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace imps
