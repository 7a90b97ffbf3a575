// Panel grouped-GEMM kernels for gfx950 (CDNA4): operator blocks pre-packed in MFMA register
// layout (panels), the vector/workspace factor read from its split copy (compute format, every
// block stored as is and transposed, so the contracted index is always contiguous).
//
// Work decomposition.  A workgroup has four waves.  With K-split S (1, 2 or 4) it serves 4/S
// vectors of one output tile (<= 32 x 32); the S waves of one vector take the tile's chunks
// round-robin and their partial tiles are summed through LDS in fixed wave order (deterministic,
// no atomics).  Each wave holds the tile as 2 x 2 MFMA sub-tiles of 16 x 16 and computes the
// transposed tile C^T = op(B)^T op(A)^T, whose MFMA result layout gives each lane consecutive
// output rows, so the column-major output is written with coalesced stores.
//
// Per chunk (one segment, KC consecutive k) a wave issues, in a fixed count and without
// branches, 16-byte buffer loads: per 16 x KC fragment one (FP32) or two (FP64; BF16X3 hi + lo)
// for the panel and as many for the raw factor.  A sub-tile outside the block, a raw row past
// the block edge or a k-run wholly past K gets an offset beyond the buffer, which returns zero
// without a memory access; k past K inside a run reads finite neighbouring data that meets zero
// panel entries.  The loads of the next chunk are issued before the MFMAs of the current one.
//
// MFMA forms (fragment layouts measured on gfx950 with one-hot operands):
//   FP64   v_mfma_f64_16x16x4f64     KC 16, RUN 4; lane k-run permuted (k = 4 (l/16) + q)
//   FP32   v_mfma_f32_16x16x4f32     KC 16, RUN 4; same permutation, identical for both factors
//   BF16X3 v_mfma_f32_16x16x32_bf16  KC 32, RUN 8; a*b ~ hi*lo + lo*hi + hi*hi, FP32 accumulation
//   result register r of lane l: FP32/BF16 D(4 (l/16) + r, l % 16); FP64 D(4 r + l/16, l % 16).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdlib>
#include <type_traits>

#include "engine.hpp"
#include "panel.hpp"

namespace imps {

namespace {

typedef double pd4_t __attribute__((ext_vector_type(4)));
typedef double pd2_t __attribute__((ext_vector_type(2)));
typedef float pf4_t __attribute__((ext_vector_type(4)));
typedef __bf16 pbf8_t __attribute__((ext_vector_type(8)));
typedef int pi4_t __attribute__((ext_vector_type(4)));

constexpr int kPWaves = 4;
constexpr int kBufDword3 = 0x00020000;  // raw buffer resource word 3 for gfx9-family (32-bit data format)
constexpr unsigned kOOB = 0x80000000u;  // beyond every buffer: the load returns zero, no memory access

template <int PREC> struct PTraits;
template <> struct PTraits<IMPS_FP64> {
    static constexpr int KC = 16, RUN = 4;
    using Acc = pd4_t;
    struct Frg { pd4_t v; };  // one 16 x KC fragment in registers
};
template <> struct PTraits<IMPS_FP32> {
    static constexpr int KC = 16, RUN = 4;
    using Acc = pf4_t;
    struct Frg { pf4_t v; };
};
template <> struct PTraits<IMPS_BF16X3> {
    static constexpr int KC = 32, RUN = 8;
    using Acc = pf4_t;
    struct Frg { pbf8_t hi, lo; };
};

__device__ __forceinline__ void split_bf16(float f, __bf16& hi, __bf16& lo) {
    hi = (__bf16)f;
    lo = (__bf16)(f - (float)hi);
}

__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void* p, long bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, (int)bytes, kBufDword3);
}

__device__ __forceinline__ pi4_t ld16(__amdgpu_buffer_rsrc_t rs, unsigned voff) {
    return __builtin_amdgcn_raw_buffer_load_b128(rs, voff, 0, 0);
}

// One 16 x KC fragment (lane l: row l % 16, RUN consecutive k from RUN (l / 16)) whose lane data
// start at byte `off` (hi) and `off + lo` (lo plane, BF16X3); kOOB-based offsets read zeros.
template <int PREC>
__device__ __forceinline__ typename PTraits<PREC>::Frg load_frg(__amdgpu_buffer_rsrc_t rs, unsigned off, unsigned lo) {
    typename PTraits<PREC>::Frg f;
    if constexpr (PREC == IMPS_BF16X3) {
        f.hi = __builtin_bit_cast(pbf8_t, ld16(rs, off));
        f.lo = __builtin_bit_cast(pbf8_t, ld16(rs, off + lo));
    } else if constexpr (PREC == IMPS_FP32) {
        f.v = __builtin_bit_cast(pf4_t, ld16(rs, off));
        (void)lo;
    } else {
        const pd2_t a = __builtin_bit_cast(pd2_t, ld16(rs, off));
        const pd2_t b = __builtin_bit_cast(pd2_t, ld16(rs, off + 16u));
        f.v = pd4_t{a[0], a[1], b[0], b[1]};
        (void)lo;
    }
    return f;
}

// ---------------------------------------------------------------------------------------------
// Panel packing: one workgroup per panel; thread e handles (fragment e / 64, lane e % 64).
template <int PREC>
__global__ __launch_bounds__(256) void panelize_kernel(const PanelDesc* __restrict__ panels, const double* __restrict__ op,
                                                       void* __restrict__ store) {
    using TR = PTraits<PREC>;
    // by reference: a local copy indexed by lane group would live in scratch memory
    const PanelDesc& pd = panels[blockIdx.x];
    const int nf = pd.nrg * pd.nkc;
    for (int e = threadIdx.x; e < nf * 64; e += blockDim.x) {
        const int f = e >> 6, lane = e & 63;
        const int rg = f / pd.nkc, kc = f - rg * pd.nkc;
        const int r = 16 * rg + (lane & 15) + (pd.npk > 0 ? pd.r0 : 0);
        const int kb = TR::KC * kc + TR::RUN * (lane >> 4);
        double v[TR::RUN];
        if (pd.npk > 0) {
            // packed chunk: lane group g holds block g's k = 0 .. pK[g] - 1 of rows r0 ..
            const int g = lane >> 4;
#pragma unroll
            for (int q = 0; q < TR::RUN; ++q) {
                double x = 0.0;
                if (g < pd.npk && r < pd.R && q < pd.pK[g])
                    x = pd.palpha[g] * op[pd.psrc[g] + (pd.ptr[g] ? (long)q + (long)r * pd.pld[g] : (long)r + (long)q * pd.pld[g])];
                v[q] = x;
            }
        } else {
#pragma unroll
            for (int q = 0; q < TR::RUN; ++q) {
                const int k = kb + q;
                double x = 0.0;
                if (r < pd.R && k < pd.K) x = pd.alpha * op[pd.src + (pd.tr ? (long)k + (long)r * pd.ld : (long)r + (long)k * pd.ld)];
                v[q] = x;
            }
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
// Split copy: the value v of a block element goes to plane N at n_at and to plane T at t_at, in
// the compute format (BF16X3: hi planes, then lo planes).
template <int PREC>
__device__ __forceinline__ void put_split(void* __restrict__ sp, long pl, long n_at, long t_at, double v) {
    if constexpr (PREC == IMPS_FP64) {
        double* p = static_cast<double*>(sp);
        p[n_at] = v;
        p[pl + t_at] = v;
    } else if constexpr (PREC == IMPS_FP32) {
        float* p = static_cast<float*>(sp);
        p[n_at] = (float)v;
        p[pl + t_at] = (float)v;
    } else {
        __bf16* p = static_cast<__bf16*>(sp);
        __bf16 h, l;
        split_bf16((float)v, h, l);
        p[n_at] = h;
        p[pl + t_at] = h;
        p[2 * pl + n_at] = l;
        p[3 * pl + t_at] = l;
    }
}

// The two halves of put_split: plane N (as stored) and plane T (transposed).
template <int PREC>
__device__ __forceinline__ void put_split_n(void* __restrict__ sp, long pl, long n_at, double v) {
    if constexpr (PREC == IMPS_FP64) {
        static_cast<double*>(sp)[n_at] = v;
    } else if constexpr (PREC == IMPS_FP32) {
        static_cast<float*>(sp)[n_at] = (float)v;
    } else {
        __bf16* p = static_cast<__bf16*>(sp);
        __bf16 h, l;
        split_bf16((float)v, h, l);
        p[n_at] = h;
        p[2 * pl + n_at] = l;
    }
}

template <int PREC>
__device__ __forceinline__ void put_split_t(void* __restrict__ sp, long pl, long t_at, double v) {
    if constexpr (PREC == IMPS_FP64) {
        static_cast<double*>(sp)[pl + t_at] = v;
    } else if constexpr (PREC == IMPS_FP32) {
        static_cast<float*>(sp)[pl + t_at] = (float)v;
    } else {
        __bf16* p = static_cast<__bf16*>(sp);
        __bf16 h, l;
        split_bf16((float)v, h, l);
        p[pl + t_at] = h;
        p[3 * pl + t_at] = l;
    }
}

// grid (ntiles, nvec): one workgroup per (tile of at most 64 x 64 block elements, vector).  Element
// (r, c) of a block goes to plane N at off + r + c*rows (as stored) and to plane T at
// off + c + r*cols.  The tile is read column by column (lanes along r) and written to plane N in
// the same order; plane T is written row by row (lanes along c) from an LDS copy, so that both
// planes are stored in contiguous runs.
template <int PREC>
__global__ __launch_bounds__(256) void split_kernel(const SplitUnit* __restrict__ units, const SplitSpan* __restrict__ spans,
                                                    const double* __restrict__ x, long x_stride, char* __restrict__ split,
                                                    long split_stride, long pl, int write_n) {
    using T = typename std::conditional<PREC == IMPS_FP64, double, float>::type;
    __shared__ long soff[SPLIT_UNIT_SPANS];
    __shared__ int srows[SPLIT_UNIT_SPANS], scols[SPLIT_UNIT_SPANS];
    const SplitUnit u = units[blockIdx.x];
    const long v = blockIdx.y;
    const double* xv = x + x_stride * v;
    void* sp = split + split_stride * v;
    if ((int)threadIdx.x < u.nb) {
        const SplitSpan b = spans[u.b0 + threadIdx.x];
        soff[threadIdx.x] = b.off;
        srows[threadIdx.x] = b.rows;
        scols[threadIdx.x] = b.cols;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < u.n; e += 256) {
        const long o = u.o0 + e;
        int lo = 0, hi = u.nb - 1;  // last span starting at or before o
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (soff[mid] <= o) lo = mid;
            else hi = mid - 1;
        }
        const unsigned rows = (unsigned)srows[lo], cols = (unsigned)scols[lo];
        const unsigned t = (unsigned)(o - soff[lo]);
        if (t >= rows * cols) continue;  // between two spans
        const unsigned r = t / cols, c = t - r * cols;
        put_split_t<PREC>(sp, pl, o, (double)(T)xv[soff[lo] + r + (long)c * rows]);
        if (write_n) put_split_n<PREC>(sp, pl, o, (double)(T)xv[o]);
    }
}

// ---------------------------------------------------------------------------------------------
template <int PREC>
struct Frag {
    typename PTraits<PREC>::Frg p[2], r[2];
};

// The raw block a lane group reads in one chunk: raw(idx0 + idx, k) at ro + idx * si + k for
// k < kr (k-major planes), or, PCF_RAW_XT read from the input, at xo + idx + k * xk.  A packed
// chunk takes them from its RawPack entry (one product per lane group), any other chunk from its
// descriptor, shifted to the group's k-run.
struct RawRef {
    long ro;
    int si, kr, xo, xk;
};
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ int pick4(const int* v, int lk) {  // v[lk] from four uniform values
    const int a = uni(v[0]), b = uni(v[1]), c = uni(v[2]), d = uni(v[3]);
    return lk < 2 ? (lk == 0 ? a : b) : (lk == 2 ? c : d);
}
template <int RUN>
__device__ __forceinline__ RawRef raw_ref(const PChunk& c, const RawPack* __restrict__ packs, int lk) {
    RawRef r;
    if (c.flags & PCF_PACKED) {
        // the whole entry by uniform (scalar) loads, then the lane group's fields by selects: a
        // per-lane load would put a vector-memory round trip between the descriptor and the data
        const RawPack* p = packs + c.raw_off;
        const int* ro = reinterpret_cast<const int*>(p->raw_off);
        int lo[4], hi[4];
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            lo[g] = ro[2 * g];
            hi[g] = ro[2 * g + 1];
        }
        r.ro = (long)(((unsigned long)(unsigned)pick4(hi, lk) << 32) | (unsigned)pick4(lo, lk));
        r.si = pick4(p->s_idx, lk);
        r.kr = pick4(p->krem, lk);
        r.xo = pick4(p->xoff, lk);
        r.xk = pick4(p->xks, lk);
    } else {
        r.ro = c.raw_off + RUN * lk;
        r.si = c.s_idx;
        r.kr = c.krem - RUN * lk;
        r.xo = c.xoff + RUN * lk * c.xks;
        r.xk = c.xks;
    }
    return r;
}

// Raw fragments of an input-vector chunk read from the FP64 input itself (rxn): lane (li, lk) takes
// raw(idx0 + li (+ 16), k0 + RUN lk + q) for q < RUN.  PCF_RAW_XT: eight-byte loads along the stored
// columns, each zero (no access) outside the block; PCF_RAW_XN: the lane's k-run is contiguous (16-
// byte loads; k past K inside a run reads finite neighbouring data that meets zero panel entries, a
// run wholly past K reads zeros).  FP32 and BF16X3 convert as the split copy does.
template <int PREC>
__device__ __forceinline__ void load_raw_x(Frag<PREC>& F, const PChunk& c, const RawRef& rr, __amdgpu_buffer_rsrc_t rxn, int nidx, int li) {
    constexpr int RUN = PTraits<PREC>::RUN;
    const bool xt = (c.flags & PCF_RAW_XT) != 0;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const int idx = li + 16 * s;
        double d[RUN];
        if (xt) {
#pragma unroll
            for (int q = 0; q < RUN; ++q) {
                const unsigned off = (idx < nidx && q < rr.kr) ? (unsigned)(rr.xo + idx + q * rr.xk) * 8u : kOOB;
                d[q] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rxn, off, 0, 0));
            }
        } else {
            const bool ok = idx < nidx && rr.kr > 0;
            const unsigned off = ok ? (unsigned)(rr.ro + (long)idx * rr.si) * 8u : kOOB;
#pragma unroll
            for (int q = 0; q < RUN; q += 2) {
                const pd2_t v = __builtin_bit_cast(pd2_t, ld16(rxn, ok ? off + 8u * q : kOOB));
                d[q] = v[0];
                d[q + 1] = v[1];
            }
        }
        if constexpr (PREC == IMPS_FP64) {
            F.r[s].v = pd4_t{d[0], d[1], d[2], d[3]};
        } else if constexpr (PREC == IMPS_FP32) {
            F.r[s].v = pf4_t{(float)d[0], (float)d[1], (float)d[2], (float)d[3]};
        } else {
#pragma unroll
            for (int q = 0; q < RUN; ++q) {
                __bf16 h, l;
                split_bf16((float)d[q], h, l);
                F.r[s].hi[q] = h;
                F.r[s].lo[q] = l;
            }
        }
    }
}

// The raw fragments of one chunk: output-side indices idx0 + [0, 16) and idx0 + [16, 32) (nidx of
// them valid).  XD: the apply reads every input-vector chunk from the input itself (instantiated
// apart, so that the split-copy kernels carry none of its registers); otherwise from the split copy
// (rx), the workspace's (rw) or, FP64 plane-N chunks, the input (rxn).  Resources are passed and
// selected by value: a selected reference to one (a lambda capture) makes the compiler keep the
// candidates in scratch memory and load the chosen one.
template <int PREC, bool XD>
__device__ __forceinline__ void load_raw_frags(Frag<PREC>& F, const PChunk& c, const RawPack* __restrict__ packs, __amdgpu_buffer_rsrc_t rx,
                                               __amdgpu_buffer_rsrc_t rxn, __amdgpu_buffer_rsrc_t rw, unsigned xlo, unsigned wlo, int nidx,
                                               int li, int lk) {
    constexpr int RUN = PTraits<PREC>::RUN;
    constexpr unsigned EB = PREC == IMPS_FP64 ? 8 : PREC == IMPS_FP32 ? 4 : 2;
    const RawRef rr = raw_ref<RUN>(c, packs, lk);
    if constexpr (XD) {
        if (c.flags & (PCF_RAW_XT | (PREC == IMPS_FP64 ? 0 : PCF_RAW_XN))) {
            load_raw_x(F, c, rr, rxn, nidx, li);
            return;
        }
    }
    const bool w = (c.flags & PCF_RAW_W) != 0;
    const __amdgpu_buffer_rsrc_t rs = w ? rw : ((PREC == IMPS_FP64 && (c.flags & PCF_RAW_XN)) ? rxn : rx);
    const unsigned lo = w ? wlo : xlo;
    const bool kok = rr.kr > 0;
    const unsigned e0 = (unsigned)(rr.ro + (long)li * rr.si) * EB;
    const unsigned es = (unsigned)(16 * rr.si) * EB;
    F.r[0] = load_frg<PREC>(rs, (kok && li < nidx) ? e0 : kOOB, lo);
    F.r[1] = load_frg<PREC>(rs, (kok && 16 + li < nidx) ? e0 + es : kOOB, lo);
}

// Issue every load of one chunk: panel fragments (row groups rg0, rg0 + 1) and the raw fragments.
template <int PREC, bool XD>
__device__ __forceinline__ void load_chunk(Frag<PREC>& F, const PChunk& c, const RawPack* __restrict__ packs, __amdgpu_buffer_rsrc_t rp,
                                           __amdgpu_buffer_rsrc_t rx, __amdgpu_buffer_rsrc_t rxn, __amdgpu_buffer_rsrc_t rw, unsigned xlo,
                                           unsigned wlo, int nidx, int li, int lk, int lane) {
    constexpr unsigned FB = PREC == IMPS_FP32 ? 64 * 16 : 64 * 32;  // bytes of one panel fragment
    constexpr unsigned LB = PREC == IMPS_FP64 ? 32 : 16;            // bytes of one lane's run (hi plane)
    const unsigned p0 = (unsigned)c.pan * FB + LB * lane;
    const unsigned p1 = c.pan_sub >= 0 ? (unsigned)(c.pan + c.pan_sub) * FB + LB * lane : kOOB;
    F.p[0] = load_frg<PREC>(rp, p0, 1024u);
    F.p[1] = load_frg<PREC>(rp, p1, 1024u);
    load_raw_frags<PREC, XD>(F, c, packs, rx, rxn, rw, xlo, wlo, nidx, li, lk);
}

// acc[a][b] += (col-side fragment b)(row-side fragment a), the transposed sub-tile (a, b), for
// a < NA, b < NB as template arguments (shared-panel kernel, one loop per tile shape): the MFMA
// sequence of a chunk is straight-line code.  With runtime counts in that kernel's loop, the
// compiler placed an accumulator read one instruction after a branch from the last MFMA writing
// it, without the MFMA result wait states (wrong values, found by the full-size parity test).
template <int PREC, int NA, int NB>
__device__ __forceinline__ void mma_tiles(typename PTraits<PREC>::Acc (&acc)[2][2], const typename PTraits<PREC>::Frg (&row)[2],
                                          const typename PTraits<PREC>::Frg (&col)[2]) {
    if constexpr (PREC == IMPS_BF16X3) {
#pragma unroll
        for (int a = 0; a < NA; ++a) {
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(col[b].hi, row[a].lo, acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(col[b].lo, row[a].hi, acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(col[b].hi, row[a].hi, acc[a][b], 0, 0, 0);
            }
        }
    } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
#pragma unroll
            for (int a = 0; a < NA; ++a) {
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    if constexpr (PREC == IMPS_FP64)
                        acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(col[b].v[q], row[a].v[q], acc[a][b], 0, 0, 0);
                    else
                        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(col[b].v[q], row[a].v[q], acc[a][b], 0, 0, 0);
                }
            }
        }
    }
}

// The same with runtime tile counts na, nb <= 2 (per-wave kernel; its generated code is checked
// for MFMA result reads across branches by the test suite).
template <int PREC>
__device__ __forceinline__ void mma_tiles_n(typename PTraits<PREC>::Acc (&acc)[2][2], const typename PTraits<PREC>::Frg (&row)[2],
                                            const typename PTraits<PREC>::Frg (&col)[2], int na, int nb) {
    if constexpr (PREC == IMPS_BF16X3) {
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            if (a >= na) break;
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                if (b >= nb) break;
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(col[b].hi, row[a].lo, acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(col[b].lo, row[a].hi, acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(col[b].hi, row[a].hi, acc[a][b], 0, 0, 0);
            }
        }
    } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                if (a >= na) break;
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    if (b >= nb) break;
                    if constexpr (PREC == IMPS_FP64)
                        acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(col[b].v[q], row[a].v[q], acc[a][b], 0, 0, 0);
                    else
                        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(col[b].v[q], row[a].v[q], acc[a][b], 0, 0, 0);
                }
            }
        }
    }
}

// Rows of C come from op(A), columns from op(B): the panel is A or B as the chunk says.  The
// operands are chosen by value (selects), so that the MFMA sequence exists once.
template <int PREC>
__device__ __forceinline__ void select_operands(typename PTraits<PREC>::Frg (&row)[2], typename PTraits<PREC>::Frg (&col)[2], const Frag<PREC>& F,
                                                bool panel_is_a) {
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        if constexpr (PREC == IMPS_BF16X3) {
            row[s].hi = panel_is_a ? F.p[s].hi : F.r[s].hi;
            row[s].lo = panel_is_a ? F.p[s].lo : F.r[s].lo;
            col[s].hi = panel_is_a ? F.r[s].hi : F.p[s].hi;
            col[s].lo = panel_is_a ? F.r[s].lo : F.p[s].lo;
        } else {
            row[s].v = panel_is_a ? F.p[s].v : F.r[s].v;
            col[s].v = panel_is_a ? F.r[s].v : F.p[s].v;
        }
    }
}

template <int PREC>
__device__ __forceinline__ void compute_chunk_n(typename PTraits<PREC>::Acc (&acc)[2][2], const Frag<PREC>& F, bool panel_is_a, int na, int nb) {
    typename PTraits<PREC>::Frg row[2], col[2];
    select_operands<PREC>(row, col, F, panel_is_a);
    mma_tiles_n<PREC>(acc, row, col, na, nb);
}

template <int PREC, int NA, int NB>
__device__ __forceinline__ void compute_chunk(typename PTraits<PREC>::Acc (&acc)[2][2], const Frag<PREC>& F, bool panel_is_a) {
    typename PTraits<PREC>::Frg row[2], col[2];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        if constexpr (PREC == IMPS_BF16X3) {
            row[s].hi = panel_is_a ? F.p[s].hi : F.r[s].hi;
            row[s].lo = panel_is_a ? F.p[s].lo : F.r[s].lo;
            col[s].hi = panel_is_a ? F.r[s].hi : F.p[s].hi;
            col[s].lo = panel_is_a ? F.r[s].lo : F.p[s].lo;
        } else {
            row[s].v = panel_is_a ? F.p[s].v : F.r[s].v;
            col[s].v = panel_is_a ? F.r[s].v : F.p[s].v;
        }
    }
    mma_tiles<PREC, NA, NB>(acc, row, col);
}

// Calls f(integral_constant<NA>, integral_constant<NB>) for the runtime tile counts na, nb in {1, 2}.
template <typename F>
__device__ __forceinline__ void with_tile_shape(int na, int nb, F&& f) {
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    if (na == 2) {
        if (nb == 2) f(I2{}, I2{});
        else f(I2{}, I1{});
    } else {
        if (nb == 2) f(I1{}, I2{});
        else f(I1{}, I1{});
    }
}

// Store the transposed accumulator sub-tile (a, b): to the FP64 output (column-major block), or,
// for a workspace stage, to the workspace's split copy.
template <int PREC>
__device__ __forceinline__ void store_sub(const typename PTraits<PREC>::Acc& v, double* __restrict__ C, void* __restrict__ wsplit,
                                          long plw, const PTask& t, int a, int b, int li, int lk) {
    const int i = t.m0 + 16 * a + li;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int j = t.n0 + 16 * b + (PREC == IMPS_FP64 ? 4 * r + lk : 4 * lk + r);
        if (i < t.M && j < t.N) {
            if (wsplit) put_split<PREC>(wsplit, plw, t.c_off + i + (long)j * t.M, t.c_off + j + (long)i * t.N, (double)v[r]);
            else C[i + (long)j * t.c_ld] = (double)v[r];
        }
    }
}

struct PArgs {
    const PTask* tasks;
    const PChunk* chunks;
    const RawPack* packs;  // the stage's packed-chunk table
    const char* pstore;
    long pstore_bytes;
    const char* xs;  // split copy of the input vectors
    long xs_stride;  // bytes per vector
    const char* xd;  // FP64 input vectors read directly by plane-N chunks, or nullptr
    long xd_stride;  // bytes per vector
    int xt;          // with xd: plane-T chunks (PCF_RAW_XT) read the input directly too (no split copy)
    const char* ws;  // split copy of the workspace
    long ws_stride;
    long plx, plw;   // plane lengths (elements)
    double* out;     // FP64 output (final stage) or nullptr
    long out_stride;
    char* wout;      // split workspace output (workspace stage) or nullptr
    int nvec, nvg;
};

template <int PREC, int S, bool XD>
__global__ __launch_bounds__(64 * kPWaves) void panel_gemm_kernel(const PArgs A) {
    using Acc = typename PTraits<PREC>::Acc;
    constexpr int V = kPWaves / S;  // vectors per workgroup
    constexpr unsigned EB = PREC == IMPS_FP64 ? 8 : PREC == IMPS_FP32 ? 4 : 2;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int li = lane & 15, lk = lane >> 4;
    const int ti = blockIdx.x / A.nvg, vg = blockIdx.x - ti * A.nvg;
    const int vl = wave / S, part = wave - vl * S;
    const int v = vg * V + vl;
    const bool active = v < A.nvec;
    const PTask t = A.tasks[ti];
    const int na = min(2, (t.M - t.m0 + 15) >> 4), nb = min(2, (t.N - t.n0 + 15) >> 4);

    Acc acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = Acc{0, 0, 0, 0};

    if (active && part < t.nch) {
        const __amdgpu_buffer_rsrc_t rx = rsrc(A.xs + A.xs_stride * v, A.xs_stride);
        const __amdgpu_buffer_rsrc_t rxn = A.xd ? rsrc(A.xd + A.xd_stride * v, A.xd_stride) : rx;
        const __amdgpu_buffer_rsrc_t rw = rsrc(A.ws + A.ws_stride * v, A.ws_stride);
        const __amdgpu_buffer_rsrc_t rp = rsrc(A.pstore, A.pstore_bytes);
        const unsigned xlo = (unsigned)(2 * A.plx) * EB, wlo = (unsigned)(2 * A.plw) * EB;  // lo planes (BF16X3)
        const int nrow = t.M - t.m0, ncol = t.N - t.n0;  // raw-side extent: columns if the panel is A, rows otherwise
        const int cend = t.ch0 + t.nch;
        {
            int c = t.ch0 + part;
            Frag<PREC> F0, F1;
            PChunk d0 = A.chunks[c];
            load_chunk<PREC, XD>(F0, d0, A.packs, rp, rx, rxn, rw, xlo, wlo, (d0.flags & PCF_PANEL_IS_A) ? ncol : nrow, li, lk, lane);
            while (true) {
                // prefetch the next chunk (the current one again at the end: same count of loads)
                const int c1 = c + S < cend ? c + S : c;
                const PChunk d1 = A.chunks[c1];
                load_chunk<PREC, XD>(F1, d1, A.packs, rp, rx, rxn, rw, xlo, wlo, (d1.flags & PCF_PANEL_IS_A) ? ncol : nrow, li, lk, lane);
                compute_chunk_n<PREC>(acc, F0, (d0.flags & PCF_PANEL_IS_A) != 0, na, nb);
                if (c + S >= cend) break;
                c = c1;
                const int c2 = c + S < cend ? c + S : c;
                d0 = A.chunks[c2];
                load_chunk<PREC, XD>(F0, d0, A.packs, rp, rx, rxn, rw, xlo, wlo, (d0.flags & PCF_PANEL_IS_A) ? ncol : nrow, li, lk, lane);
                compute_chunk_n<PREC>(acc, F1, (d1.flags & PCF_PANEL_IS_A) != 0, na, nb);
                if (c + S >= cend) break;
                c = c2;
            }
        }
    }

    double* C = A.out ? A.out + A.out_stride * v + t.c_off : nullptr;
    void* wsplit = A.wout ? (void*)(A.wout + A.ws_stride * v) : nullptr;
    if constexpr (S == 1) {
        if (!active) return;
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
                if (a < na && b < nb) store_sub<PREC>(acc[a][b], C, wsplit, A.plw, t, a, b, li, lk);
    } else {
        __shared__ Acc partial[kPWaves][4][64];
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
            store_sub<PREC>(sum, C, wsplit, A.plw, t, a, b, li, lk);
        }
    }
}

// K-split 1, four vectors per workgroup: the four waves run one task (one chunk list) for four
// vectors, so the panel fragments of a chunk, identical for all of them, are loaded once per
// workgroup into a double-buffered LDS copy (16-byte units, fragment f at units [f*HU, (f+1)*HU))
// and every wave reads them from there; each wave loads only its own vector's raw fragments.
// Every wave, including those past the last vector, takes part in the loads and barriers.
template <int PREC, bool XD>
__global__ __launch_bounds__(64 * kPWaves) void panel_gemm_shared_kernel(const PArgs A) {
    using TR = PTraits<PREC>;
    using Acc = typename TR::Acc;
    constexpr int RUN = TR::RUN;
    constexpr unsigned EB = PREC == IMPS_FP64 ? 8 : PREC == IMPS_FP32 ? 4 : 2;
    constexpr unsigned FB = PREC == IMPS_FP32 ? 64 * 16 : 64 * 32;  // bytes of one panel fragment
    constexpr int HU = FB / 16;                                     // 16-byte units per fragment
    __shared__ pi4_t lds[2][2 * HU];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int tid = threadIdx.x, lane = tid & 63;
    const int li = lane & 15, lk = lane >> 4;
    const int ti = blockIdx.x / A.nvg, vg = blockIdx.x - ti * A.nvg;
    const int v = vg * kPWaves + wave;
    const bool active = v < A.nvec;
    const PTask t = A.tasks[ti];
    const int na = min(2, (t.M - t.m0 + 15) >> 4), nb = min(2, (t.N - t.n0 + 15) >> 4);

    Acc acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = Acc{0, 0, 0, 0};

    const int vs = active ? v : 0;
    const __amdgpu_buffer_rsrc_t rx = rsrc(A.xs + A.xs_stride * vs, A.xs_stride);
    const __amdgpu_buffer_rsrc_t rxn = A.xd ? rsrc(A.xd + A.xd_stride * vs, A.xd_stride) : rx;
    const __amdgpu_buffer_rsrc_t rw = rsrc(A.ws + A.ws_stride * vs, A.ws_stride);
    const __amdgpu_buffer_rsrc_t rp = rsrc(A.pstore, A.pstore_bytes);
    const unsigned xlo = (unsigned)(2 * A.plx) * EB, wlo = (unsigned)(2 * A.plw) * EB;  // lo planes (BF16X3)
    const int nrow = t.M - t.m0, ncol = t.N - t.n0;
    // this thread's 16-byte unit of a chunk's panel data (fragment rg0, then rg0 + 1)
    auto panel_unit = [&](const PChunk& c) -> pi4_t {
        if (tid >= 2 * HU) return pi4_t{0, 0, 0, 0};
        const int f = tid / HU, w = tid - f * HU;
        const unsigned off = f == 0 ? (unsigned)c.pan * FB + 16u * w
                                    : (c.pan_sub >= 0 ? (unsigned)(c.pan + c.pan_sub) * FB + 16u * w : kOOB);
        return ld16(rp, off);
    };
    // raw fragments of output-side indices idx0 + [0, 16) and idx0 + [16, 32) (zeros when inactive)
    auto load_raw = [&](Frag<PREC>& F, const PChunk& c) {
        load_raw_frags<PREC, XD>(F, c, A.packs, rx, rxn, rw, xlo, wlo, active ? ((c.flags & PCF_PANEL_IS_A) ? ncol : nrow) : 0, li, lk);
    };
    // panel fragments of the chunk in LDS buffer b, in the lane layout of load_frg
    auto read_panels = [&](Frag<PREC>& F, int b) {
#pragma unroll
        for (int f = 0; f < 2; ++f) {
            if constexpr (PREC == IMPS_BF16X3) {
                F.p[f].hi = __builtin_bit_cast(pbf8_t, lds[b][f * HU + lane]);
                F.p[f].lo = __builtin_bit_cast(pbf8_t, lds[b][f * HU + 64 + lane]);
            } else if constexpr (PREC == IMPS_FP32) {
                F.p[f].v = __builtin_bit_cast(pf4_t, lds[b][f * HU + lane]);
            } else {
                const pd2_t x0 = __builtin_bit_cast(pd2_t, lds[b][f * HU + 2 * lane]);
                const pd2_t x1 = __builtin_bit_cast(pd2_t, lds[b][f * HU + 2 * lane + 1]);
                F.p[f].v = pd4_t{x0[0], x0[1], x1[0], x1[1]};
            }
        }
    };

    const int cend = t.ch0 + t.nch;
    if (t.nch > 0) with_tile_shape(na, nb, [&](auto NA_, auto NB_) {
        constexpr int NA = decltype(NA_)::value, NB = decltype(NB_)::value;
        int c = t.ch0;
        PChunk d = A.chunks[c];
        Frag<PREC> F, G;
        pi4_t unit = panel_unit(d);
        load_raw(F, d);
        if (tid < 2 * HU) lds[0][tid] = unit;
        __syncthreads();
        int buf = 0;
        while (true) {
            const bool more = c + 1 < cend;
            const PChunk dn = A.chunks[more ? c + 1 : c];
            if (more) {
                unit = panel_unit(dn);
                load_raw(G, dn);
            }
            read_panels(F, buf);
            if (active) compute_chunk<PREC, NA, NB>(acc, F, (d.flags & PCF_PANEL_IS_A) != 0);
            if (!more) break;
            if (tid < 2 * HU) lds[buf ^ 1][tid] = unit;
            __syncthreads();
            buf ^= 1;
            ++c;
            d = dn;
            F.r[0] = G.r[0];
            F.r[1] = G.r[1];
        }
    });
    if (!active) return;
    double* C = A.out ? A.out + A.out_stride * v + t.c_off : nullptr;
    void* wsplit = A.wout ? (void*)(A.wout + A.ws_stride * v) : nullptr;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
            if (a < na && b < nb) store_sub<PREC>(acc[a][b], C, wsplit, A.plw, t, a, b, li, lk);
}

__global__ void zero_kernel(double* __restrict__ p, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = 0.0;
}

// Default schedule (without the autotuner): K-split 1 in BF16X3 uses the shared-panel kernel
// (measured faster there; in FP64 and FP32 the per-wave kernel is); IMPS_PANEL_SHARED=0 selects the
// per-wave kernel (comparison runs)
bool panel_shared_enabled() {
    static const bool on = [] {
        const char* e = std::getenv("IMPS_PANEL_SHARED");
        return !(e && e[0] == '0');
    }();
    return on;
}

template <int PREC, int S>
hipError_t launch_ps(const DevPStage& st, PArgs a, bool shared, hipStream_t stream) {
    constexpr int V = kPWaves / S;
    a.nvg = (a.nvec + V - 1) / V;
    a.tasks = st.tasks;
    a.chunks = st.chunks;
    a.packs = st.packs;
    const long nblk = (long)st.ntasks * a.nvg;
    if (nblk > 0x7fffffffL) return hipErrorInvalidValue;
    if constexpr (S == 1) {
        if (shared) {
            if (a.xt) hipLaunchKernelGGL((panel_gemm_shared_kernel<PREC, true>), dim3((unsigned)nblk), dim3(64 * kPWaves), 0, stream, a);
            else hipLaunchKernelGGL((panel_gemm_shared_kernel<PREC, false>), dim3((unsigned)nblk), dim3(64 * kPWaves), 0, stream, a);
            return hipGetLastError();
        }
    }
    if (a.xt) hipLaunchKernelGGL((panel_gemm_kernel<PREC, S, true>), dim3((unsigned)nblk), dim3(64 * kPWaves), 0, stream, a);
    else hipLaunchKernelGGL((panel_gemm_kernel<PREC, S, false>), dim3((unsigned)nblk), dim3(64 * kPWaves), 0, stream, a);
    return hipGetLastError();
}

template <int PREC>
hipError_t launch_p(const DevPStage& st, int ksplit, bool shared, const PArgs& a, hipStream_t stream) {
    switch (ksplit) {
        case 1: return launch_ps<PREC, 1>(st, a, shared, stream);
        case 2: return launch_ps<PREC, 2>(st, a, false, stream);
        case 4: return launch_ps<PREC, 4>(st, a, false, stream);
        default: return hipErrorInvalidValue;
    }
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

hipError_t launch_split(int prec, const SplitUnit* d_units, int nunits, const SplitSpan* d_spans, const double* x, long x_stride,
                        void* split, long split_stride, long pl, int nvec, hipStream_t stream, bool write_n) {
    if (nunits == 0 || nvec == 0) return hipSuccess;
    if (nvec > 65535) return hipErrorInvalidValue;
    if (!write_n && prec != IMPS_FP64) return hipErrorInvalidValue;
    const dim3 grid(nunits, nvec);
    const int wn = write_n ? 1 : 0;
    switch (prec) {
        case IMPS_FP64:
            hipLaunchKernelGGL(split_kernel<IMPS_FP64>, grid, dim3(256), 0, stream, d_units, d_spans, x, x_stride, (char*)split, split_stride,
                               pl, wn);
            break;
        case IMPS_FP32:
            hipLaunchKernelGGL(split_kernel<IMPS_FP32>, grid, dim3(256), 0, stream, d_units, d_spans, x, x_stride, (char*)split, split_stride,
                               pl, wn);
            break;
        case IMPS_BF16X3:
            hipLaunchKernelGGL(split_kernel<IMPS_BF16X3>, grid, dim3(256), 0, stream, d_units, d_spans, x, x_stride, (char*)split, split_stride,
                               pl, wn);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

bool panel_shared_default(int prec) { return prec == IMPS_BF16X3 && panel_shared_enabled(); }

hipError_t launch_panel_stage(int prec, const DevPStage& st, int ksplit, bool shared, const PanelLaunch& L, hipStream_t stream) {
    if (st.ntasks == 0 || L.nvec == 0) return hipSuccess;
    PArgs a{};
    a.pstore = (const char*)L.pstore;
    a.pstore_bytes = L.pstore_bytes;
    a.xs = (const char*)L.xs;
    a.xs_stride = L.xs_stride;
    a.xd = (const char*)L.xd;
    a.xd_stride = L.xd_stride * (long)sizeof(double);
    a.xt = L.xt_direct ? 1 : 0;
    a.ws = (const char*)L.ws;
    a.ws_stride = L.ws_stride;
    a.plx = L.plx;
    a.plw = L.plw;
    a.out = st.out_slot == SLOT_W ? nullptr : L.out;
    a.out_stride = L.out_stride;
    a.wout = st.out_slot == SLOT_W ? (char*)L.ws : nullptr;
    a.nvec = L.nvec;
    switch (prec) {
        case IMPS_FP64: return launch_p<IMPS_FP64>(st, ksplit, shared, a, stream);
        case IMPS_FP32: return launch_p<IMPS_FP32>(st, ksplit, shared, a, stream);
        case IMPS_BF16X3: return launch_p<IMPS_BF16X3>(st, ksplit, shared, a, stream);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace imps
