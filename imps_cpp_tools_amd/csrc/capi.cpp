// C ABI of libimps_hip.so: contexts, plans and the operator-callback entry points.
// Declarations and the reference interfaces each entry point replaces: include/imps_capi.h.
#include <hip/hip_runtime.h>

#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/imps_capi.h"
#include "engine.hpp"
#include "layout.hpp"

using namespace imps;

struct imps_context_s {
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
};

namespace {

thread_local std::string g_err;

struct HipError : std::runtime_error { using std::runtime_error::runtime_error; };

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}

template <class F>
imps_status guarded(F&& f) {
    try {
        return f();
    } catch (const InvalidArg& e) {
        g_err = e.what();
        return IMPS_INVALID_ARG;
    } catch (const HipError& e) {
        g_err = e.what();
        return IMPS_HIP_ERROR;
    } catch (const std::bad_alloc&) {
        g_err = "host allocation failed";
        return IMPS_INVALID_ARG;
    } catch (const std::exception& e) {
        g_err = e.what();
        return IMPS_INVALID_ARG;
    }
}

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    ~DevBuf() { if (p) (void)hipFree(p); }
    void ensure(size_t n) {
        if (n <= bytes) return;
        if (p) { hip_check(hipFree(p), "hipFree"); p = nullptr; bytes = 0; }
        if (n == 0) return;
        hip_check(hipMalloc(&p, n), "hipMalloc");
        bytes = n;
    }
};

enum Kind { K_TM = 0, K_HAC = 1, K_HC = 2, K_REDOP = 3 };

}  // namespace

struct imps_plan_s {
    imps_context ctx = nullptr;
    int kind = 0;
    int prec = IMPS_FP64;
    Program prog;
    Layout out_layout;           // derived output structure (redop)
    std::vector<long> op_parts;  // element counts of the operator-store pieces, in set order
    bool op_set = false;
    // device state
    DevBuf d_prog, d_op, d_w;
    std::vector<DevStage> dstages;
    size_t elem_bytes() const { return prec == IMPS_FP64 ? 8 : 4; }
};

namespace {

void upload_program(imps_plan_s* P) {
    if (!P->ctx) return;
    hip_check(hipSetDevice(P->ctx->device), "hipSetDevice");
    size_t bytes = 0;
    for (auto& st : P->prog.stages) bytes += st.tasks.size() * sizeof(Task) + st.segs.size() * sizeof(Seg) + 64;
    P->d_prog.ensure(bytes + 64);
    std::vector<char> host(bytes + 64, 0);
    size_t pos = 0;
    P->dstages.clear();
    for (auto& st : P->prog.stages) {
        DevStage ds;
        ds.ntasks = (int)st.tasks.size();
        ds.out_slot = st.out_slot;
        pos = (pos + 15) & ~size_t(15);
        std::memcpy(host.data() + pos, st.tasks.data(), st.tasks.size() * sizeof(Task));
        ds.tasks = (const Task*)((char*)P->d_prog.p + pos);
        pos += st.tasks.size() * sizeof(Task);
        pos = (pos + 15) & ~size_t(15);
        std::memcpy(host.data() + pos, st.segs.data(), st.segs.size() * sizeof(Seg));
        ds.segs = (const Seg*)((char*)P->d_prog.p + pos);
        pos += st.segs.size() * sizeof(Seg);
        P->dstages.push_back(ds);
    }
    if (pos) hip_check(hipMemcpyAsync(P->d_prog.p, host.data(), pos, hipMemcpyHostToDevice, P->ctx->stream), "upload program");
    hip_check(hipStreamSynchronize(P->ctx->stream), "upload program");
    P->d_op.ensure(std::max<long>(P->prog.op_elems, 1) * P->elem_bytes());
}

// Copy FP64 operator pieces into the plan's operator store at its working precision.
void set_operator(imps_plan_s* P, const std::vector<const double*>& parts) {
    if (!P->ctx) throw InvalidArg("plan was created without a context");
    if (parts.size() != P->op_parts.size()) throw InvalidArg("operator piece count mismatch");
    hip_check(hipSetDevice(P->ctx->device), "hipSetDevice");
    long off = 0;
    for (size_t i = 0; i < parts.size(); ++i) {
        const long n = P->op_parts[i];
        if (n > 0 && !parts[i]) throw InvalidArg("operator data pointer is NULL");
        if (n > 0) {
            if (P->prec == IMPS_FP64)
                hip_check(hipMemcpyAsync((double*)P->d_op.p + off, parts[i], n * sizeof(double), hipMemcpyDeviceToDevice, P->ctx->stream),
                          "copy operator");
            else
                hip_check(launch_d2f(parts[i], (float*)P->d_op.p + off, n, P->ctx->stream), "convert operator");
        }
        off += n;
    }
    P->op_set = true;
}

void run(imps_plan_s* P, const double* x, double* y, int nvec) {
    if (!P->ctx) throw InvalidArg("plan was created without a context");
    if (nvec < 1) throw InvalidArg("nvec must be >= 1");
    if ((P->prog.x_elems > 0 && !x) || (P->prog.y_elems > 0 && !y)) throw InvalidArg("x or y is NULL");
    if (!P->op_set && P->prog.op_elems > 0) throw InvalidArg("operator data not set");
    hip_check(hipSetDevice(P->ctx->device), "hipSetDevice");
    hipStream_t s = P->ctx->stream;
    const size_t eb = P->elem_bytes();
    P->d_w.ensure((size_t)std::max<long>(P->prog.w_elems, 1) * nvec * eb);
    Slots sl;
    sl.base[SLOT_X] = x;           sl.stride[SLOT_X] = P->prog.x_elems;  // FP64, converted in-kernel
    sl.base[SLOT_OP] = P->d_op.p;  sl.stride[SLOT_OP] = 0;
    sl.base[SLOT_W] = P->d_w.p;    sl.stride[SLOT_W] = P->prog.w_elems;
    for (auto& ds : P->dstages) {
        if (ds.out_slot == SLOT_W)
            hip_check(launch_stage(P->prec, ds, sl, P->d_w.p, P->prog.w_elems, P->prec == IMPS_FP64, nvec, s), "grouped gemm");
        else
            hip_check(launch_stage(P->prec, ds, sl, y, P->prog.y_elems, true, nvec, s), "grouped gemm");
    }
}

imps_plan_s* new_plan(imps_context ctx, int kind, int prec) {
    if (prec < IMPS_FP64 || prec > IMPS_BF16X3) throw InvalidArg("unknown precision");
    auto* P = new imps_plan_s;
    P->ctx = ctx;
    P->kind = kind;
    P->prec = prec;
    return P;
}

imps_status finish_plan(imps_plan* out, std::unique_ptr<imps_plan_s>& P) {
    upload_program(P.get());
    *out = P.release();
    return IMPS_OK;
}

void check_kind(imps_plan P, int kind) {
    if (!P) throw InvalidArg("plan is NULL");
    if (P->kind != kind) throw InvalidArg("plan has a different operator kind");
}

}  // namespace

extern "C" {

const char* imps_last_error(void) { return g_err.c_str(); }
const char* imps_version(void) { return "imps_hip 0.1 gfx950"; }

imps_status imps_create(imps_context* ctx, int device) {
    return guarded([&] {
        if (!ctx) throw InvalidArg("ctx is NULL");
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) {
            g_err = "no HIP device " + std::to_string(device);
            return IMPS_NO_DEVICE;
        }
        hipDeviceProp_t prop;
        hip_check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
            g_err = std::string("device is ") + prop.gcnArchName + ", this library is built for gfx950";
            return IMPS_NO_DEVICE;
        }
        hip_check(hipSetDevice(device), "hipSetDevice");
        auto* c = new imps_context_s;
        c->device = device;
        if (hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking) != hipSuccess) {
            delete c;
            throw HipError("hipStreamCreate failed");
        }
        c->stream = c->own;
        *ctx = c;
        return IMPS_OK;
    });
}

imps_status imps_destroy(imps_context ctx) {
    return guarded([&] {
        if (!ctx) return IMPS_OK;
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        if (ctx->own) (void)hipStreamDestroy(ctx->own);
        delete ctx;
        return IMPS_OK;
    });
}

imps_status imps_set_stream(imps_context ctx, void* stream) {
    return guarded([&] {
        if (!ctx) throw InvalidArg("ctx is NULL");
        ctx->stream = (hipStream_t)stream;  // NULL selects the legacy default stream
        return IMPS_OK;
    });
}

imps_status imps_synchronize(imps_context ctx) {
    return guarded([&] {
        if (!ctx) throw InvalidArg("ctx is NULL");
        hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        return IMPS_OK;
    });
}

imps_status imps_plan_destroy(imps_plan plan) {
    return guarded([&] {
        if (plan && plan->ctx) (void)hipSetDevice(plan->ctx->device);
        delete plan;
        return IMPS_OK;
    });
}

imps_status imps_plan_cost(imps_plan P, double* flops, double* bytes) {
    return guarded([&] {
        if (!P) throw InvalidArg("plan is NULL");
        if (flops) *flops = P->prog.flops;
        if (bytes) *bytes = (double)P->prog.op_elems * P->elem_bytes() + 8.0 * (P->prog.x_elems + P->prog.y_elems);
        return IMPS_OK;
    });
}

imps_status imps_plan_shape(imps_plan P, int* launches, int* tiles) {
    return guarded([&] {
        if (!P) throw InvalidArg("plan is NULL");
        int t = 0;
        for (auto& st : P->prog.stages) t += (int)st.tasks.size();
        if (launches) *launches = (int)P->prog.stages.size();
        if (tiles) *tiles = t;
        return IMPS_OK;
    });
}

imps_status imps_plan_out_layout(imps_plan P, int* nlist, int* nblk_total, long* nelem, int* nblk, int* qin, int* qout, int* rows,
                                 int* cols) {
    return guarded([&] {
        if (!P) throw InvalidArg("plan is NULL");
        const Layout& L = P->out_layout;
        if (nlist) *nlist = (int)L.l.size();
        if (nblk_total) *nblk_total = (int)L.nblocks();
        if (nelem) *nelem = L.nelem;
        if (!qin) return IMPS_OK;
        if (!nblk || !qout || !rows || !cols) throw InvalidArg("output arrays are NULL");
        int b = 0;
        for (size_t e = 0; e < L.l.size(); ++e) {
            nblk[e] = (int)L.l[e].b.size();
            for (auto& blk : L.l[e].b) {
                for (int i = 0; i < L.nsym; ++i) { qin[b * L.nsym + i] = blk.qin.q[i]; qout[b * L.nsym + i] = blk.qout.q[i]; }
                rows[b] = blk.rows;
                cols[b] = blk.cols;
                ++b;
            }
        }
        return IMPS_OK;
    });
}

// ---- transfer matrices ----
imps_status imps_tm_plan(imps_context ctx, imps_plan* plan, int dir, const imps_layout* A, const imps_layout* B, const imps_layout* x,
                         const imps_layout* y, int prec) {
    return guarded([&] {
        if (!plan) throw InvalidArg("plan is NULL");
        if (dir != IMPS_LEFT && dir != IMPS_RIGHT) throw InvalidArg("dir must be IMPS_LEFT or IMPS_RIGHT");
        std::unique_ptr<imps_plan_s> P(new_plan(ctx, K_TM, prec));
        Layout LA = layout_from_c(A, "A"), LX = layout_from_c(x, "x"), LY = layout_from_c(y, "y");
        Layout LB;
        if (B) LB = layout_from_c(B, "B");
        P->prog = build_tm(dir, LA, B ? &LB : nullptr, LX, LY);
        P->op_parts = {LA.nelem};
        if (B) P->op_parts.push_back(LB.nelem);
        return finish_plan(plan, P);
    });
}
imps_status imps_tm_set_mps(imps_plan P, const double* A, const double* B) {
    return guarded([&] {
        check_kind(P, K_TM);
        std::vector<const double*> parts{A};
        if (P->op_parts.size() == 2) {
            if (!B) throw InvalidArg("plan was built for a mixed TM: B is required");
            parts.push_back(B);
        }
        set_operator(P, parts);
        return IMPS_OK;
    });
}
imps_status imps_tm_apply(imps_plan P, const double* x, double* y, int nvec) {
    return guarded([&] { check_kind(P, K_TM); run(P, x, y, nvec); return IMPS_OK; });
}

// ---- H_AC ----
imps_status imps_hac_plan(imps_context ctx, imps_plan* plan, const imps_layout* AC, const imps_layout* EHL, const imps_layout* EHR,
                          const imps_layout* HL, const imps_layout* HR, int prec) {
    return guarded([&] {
        if (!plan) throw InvalidArg("plan is NULL");
        std::unique_ptr<imps_plan_s> P(new_plan(ctx, K_HAC, prec));
        Layout la = layout_from_c(AC, "AC"), l1 = layout_from_c(EHL, "EHL"), l2 = layout_from_c(EHR, "EHR"),
               l3 = layout_from_c(HL, "HL"), l4 = layout_from_c(HR, "HR");
        P->prog = build_hac(la, l1, l2, l3, l4);
        P->op_parts = {l1.nelem, l2.nelem, l3.nelem, l4.nelem};
        return finish_plan(plan, P);
    });
}
imps_status imps_hac_set_operator(imps_plan P, const double* EHL, const double* EHR, const double* HL, const double* HR) {
    return guarded([&] { check_kind(P, K_HAC); set_operator(P, {EHL, EHR, HL, HR}); return IMPS_OK; });
}
imps_status imps_hac_apply(imps_plan P, const double* x, double* y, int nvec) {
    return guarded([&] { check_kind(P, K_HAC); run(P, x, y, nvec); return IMPS_OK; });
}

// ---- H_C ----
imps_status imps_hc_plan(imps_context ctx, imps_plan* plan, const imps_layout* C, const imps_layout* EHL, const imps_layout* EHR,
                         const imps_layout* HL, const imps_layout* AR, int prec) {
    return guarded([&] {
        if (!plan) throw InvalidArg("plan is NULL");
        std::unique_ptr<imps_plan_s> P(new_plan(ctx, K_HC, prec));
        Layout lc = layout_from_c(C, "C"), l1 = layout_from_c(EHL, "EHL"), l2 = layout_from_c(EHR, "EHR"),
               l3 = layout_from_c(HL, "HL"), l4 = layout_from_c(AR, "AR");
        P->prog = build_hc(lc, l1, l2, l3, l4);
        P->op_parts = {l1.nelem, l2.nelem, l3.nelem, l4.nelem};
        return finish_plan(plan, P);
    });
}
imps_status imps_hc_set_operator(imps_plan P, const double* EHL, const double* EHR, const double* HL, const double* AR) {
    return guarded([&] { check_kind(P, K_HC); set_operator(P, {EHL, EHR, HL, AR}); return IMPS_OK; });
}
imps_status imps_hc_apply(imps_plan P, const double* x, double* y, int nvec) {
    return guarded([&] { check_kind(P, K_HC); run(P, x, y, nvec); return IMPS_OK; });
}

// ---- ReducedOp ----
imps_status imps_redop_plan(imps_context ctx, imps_plan* plan, const double* H, int d, const imps_layout* A, int dir, int prec) {
    return guarded([&] {
        if (!plan || !H) throw InvalidArg("plan or H is NULL");
        if (d < 1 || d > 64) throw InvalidArg("d out of range");
        if (dir != IMPS_LEFT && dir != IMPS_RIGHT) throw InvalidArg("dir must be IMPS_LEFT or IMPS_RIGHT");
        std::unique_ptr<imps_plan_s> P(new_plan(ctx, K_REDOP, prec));
        Layout la = layout_from_c(A, "A");
        P->prog = build_redop(H, d, la, dir, P->out_layout);
        P->op_set = true;  // no operator store: A is the input operand
        return finish_plan(plan, P);
    });
}
imps_status imps_redop_apply(imps_plan P, const double* A, double* y) {
    return guarded([&] { check_kind(P, K_REDOP); run(P, A, y, 1); return IMPS_OK; });
}

}  // extern "C"
