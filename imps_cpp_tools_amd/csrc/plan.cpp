// Plan construction helpers and execution of grouped-GEMM programs.
#include "plan.hpp"

#include <cstring>

namespace imps {

void ctx_release(imps_context ctx) {
    if (!ctx || --ctx->refs > 0) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->own) (void)hipStreamDestroy(ctx->own);
    delete ctx;
}

imps_plan_s* new_plan(imps_context ctx, int kind, int prec) {
    if (prec < IMPS_FP64 || prec > IMPS_BF16X3) throw InvalidArg("unknown precision");
    auto* P = new imps_plan_s;
    P->hold.hold(ctx);
    P->ctx = ctx;
    P->kind = kind;
    P->prec = prec;
    return P;
}

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

void run(imps_plan_s* P, const double* x, double* y, int nvec, const double* op_override) {
    if (!P->ctx) throw InvalidArg("plan was created without a context");
    if (nvec < 1) throw InvalidArg("nvec must be >= 1");
    if ((P->prog.x_elems > 0 && !x) || (P->prog.y_elems > 0 && !y)) throw InvalidArg("x or y is NULL");
    if (!P->op_set && !op_override && P->prog.op_elems > 0) throw InvalidArg("operator data not set");
    hip_check(hipSetDevice(P->ctx->device), "hipSetDevice");
    hipStream_t s = P->ctx->stream;
    P->d_w.ensure((size_t)std::max<long>(P->prog.w_elems, 1) * nvec * P->elem_bytes());
    Slots sl;
    sl.base[SLOT_X] = x;           sl.stride[SLOT_X] = P->prog.x_elems;  // FP64, converted in-kernel
    sl.base[SLOT_OP] = op_override ? (const void*)op_override : P->d_op.p;
    sl.stride[SLOT_OP] = 0;
    sl.base[SLOT_W] = P->d_w.p;    sl.stride[SLOT_W] = P->prog.w_elems;
    for (auto& ds : P->dstages) {
        if (ds.out_slot == SLOT_W)
            hip_check(launch_stage(P->prec, ds, sl, P->d_w.p, P->prog.w_elems, P->prec == IMPS_FP64, nvec, s), "grouped gemm");
        else
            hip_check(launch_stage(P->prec, ds, sl, y, P->prog.y_elems, true, nvec, s), "grouped gemm");
    }
}

}  // namespace imps
