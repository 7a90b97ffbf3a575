// C ABI of libimps_hip.so: contexts, plans and the operator-callback entry points.
// Declarations and the reference interfaces each entry point replaces: include/imps_capi.h.
#include <hip/hip_runtime.h>

#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/imps_capi.h"
#include "adjoint.hpp"
#include "heff.hpp"
#include "krylov.hpp"
#include "plan.hpp"
#include "static_scratch.hpp"

using namespace imps;

namespace {

thread_local std::string g_err;

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
    } catch (const NotConverged& e) {
        g_err = e.what();
        return IMPS_NOT_CONVERGED;
    } catch (const Numerical& e) {
        g_err = e.what();
        return IMPS_NUMERICAL;
    } catch (const std::bad_alloc&) {
        g_err = "host allocation failed";
        return IMPS_INVALID_ARG;
    } catch (const std::exception& e) {
        g_err = e.what();
        return IMPS_INVALID_ARG;
    }
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
        // set-up of the fixed scratch of imps_hac_apply_static on this device (its first call
        // retries and reports an error if this fails)
        try {
            (void)static_scratch_address(device);
        } catch (const std::exception&) {
        }
        *ctx = c;
        return IMPS_OK;
    });
}

imps_status imps_destroy(imps_context ctx) {
    return guarded([&] {
        if (!ctx) return IMPS_OK;
        if (ctx->refs > 1) {  // plans still alive: finish this context's queued work now
            hip_check(hipSetDevice(ctx->device), "hipSetDevice");
            hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        }
        ctx_release(ctx);
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

imps_status imps_plan_timing(imps_plan P, int enable, double* ms, int* launches) {
    return guarded([&] {
        if (!P) throw InvalidArg("plan is NULL");
        if (P->ctx) hip_check(hipSetDevice(P->ctx->device), "hipSetDevice");
        const double t = take_timing(P, launches);
        if (ms) *ms = t;
        P->timing = enable != 0;
        return IMPS_OK;
    });
}

imps_status imps_plan_schedule(imps_plan P, int nvec, int* ksplit, int* shared) {
    return guarded([&] {
        if (!P) throw InvalidArg("plan is NULL");
        if (nvec < 1) throw InvalidArg("nvec must be >= 1");
        const int sc = panel_schedule_of(P, nvec);
        static const int ks[4] = {1, 1, 2, 4};
        if (ksplit) *ksplit = P->use_panels ? ks[sc] : 0;
        if (shared) *shared = P->use_panels && sc == 1 ? 1 : 0;
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
        P->lays = {LA, B ? LB : LA, LX, LY};
        P->tm_dir = dir;
        P->tm_mixed = B != nullptr;
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
        P->lays = {la, l1, l2, l3, l4};
        return finish_plan(plan, P);
    });
}
imps_status imps_hac_set_operator(imps_plan P, const double* EHL, const double* EHR, const double* HL, const double* HR) {
    return guarded([&] { check_kind(P, K_HAC); set_operator(P, {EHL, EHR, HL, HR}); return IMPS_OK; });
}
imps_status imps_hac_apply(imps_plan P, const double* x, double* y, int nvec) {
    return guarded([&] { check_kind(P, K_HAC); run(P, x, y, nvec); return IMPS_OK; });
}
imps_status imps_hac_apply_static(imps_plan P, const double* x, double* y, int nvec) {
    return guarded([&] {
        check_kind(P, K_HAC);
        if (!P->ctx) throw InvalidArg("plan was created without a context");
        run_static(P, x, y, nvec, static_scratch_address(P->ctx->device), kStaticScratchBytes);
        return IMPS_OK;
    });
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
        P->lays = {lc, l1, l2, l3, l4};
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

// ---- environments (A14/A15) ----
imps_status imps_env_plan(imps_context ctx, imps_plan* plan, int dir, int ncell, const imps_layout* A, const imps_layout* x, int prec) {
    return guarded([&] {
        if (!plan || !A) throw InvalidArg("plan or A is NULL");
        if (dir != IMPS_LEFT && dir != IMPS_RIGHT) throw InvalidArg("dir must be IMPS_LEFT or IMPS_RIGHT");
        if (ncell < 1 || ncell > 64) throw InvalidArg("ncell out of range");
        std::unique_ptr<imps_plan_s> P(new_plan(ctx, K_ENV, prec));
        std::vector<Layout> LA;
        for (int k = 0; k < ncell; ++k) LA.push_back(layout_from_c(&A[k], "A"));
        Layout lx = layout_from_c(x, "x");
        for (auto& l : LA) require_same_nsym(LA[0], l, "env");
        require_same_nsym(LA[0], lx, "env");
        P->env = std::make_shared<EnvSolver>();
        P->env->build(ctx, dir, LA, lx, prec);
        P->op_set = true;
        *plan = P.release();
        return IMPS_OK;
    });
}
imps_status imps_env_set_mps(imps_plan P, const double* const* A) {
    return guarded([&] {
        check_kind(P, K_ENV);
        if (!A) throw InvalidArg("A is NULL");
        P->env->set_mps(std::vector<const double*>(A, A + P->env->ncell));
        return IMPS_OK;
    });
}
imps_status imps_env_solve(imps_plan P, const double* x, const double* L, const double* R, const double* y0, double* y, double tol,
                           int maxit, int krylov, int* iters, double* relres) {
    return guarded([&] {
        check_kind(P, K_ENV);
        if (!(tol > 0)) throw InvalidArg("tol must be > 0");
        KrylovStats st = P->env->solve(x, L, R, y0, y, tol, maxit, krylov);
        if (iters) *iters = st.iters;
        if (relres) *relres = st.relres;
        if (!st.converged) throw NotConverged("env solve: relative residual " + std::to_string(st.relres) + " after " +
                                              std::to_string(st.iters) + " applications");
        return IMPS_OK;
    });
}

// ---- lowest eigenpair of H_AC / H_C (A17) ----
imps_status imps_eigsh(imps_plan P, imps_plan P64, double* x, double tol, int maxit, int krylov, double* eval, int* iters, double* resid) {
    return guarded([&] {
        if (!P || !x) throw InvalidArg("plan or x is NULL");
        if (P->kind != K_HAC && P->kind != K_HC) throw InvalidArg("eigsh needs an H_AC or H_C plan");
        if (P64 && (P64->kind != P->kind || P64->prec != IMPS_FP64 || P64->prog.x_elems != P->prog.x_elems))
            throw InvalidArg("refinement plan must be an FP64 plan of the same operator");
        if (!P->ctx) throw InvalidArg("plan was created without a context");
        if (!(tol > 0)) throw InvalidArg("tol must be > 0");
        hipStream_t s = P->ctx->stream;
        hip_check(hipSetDevice(P->ctx->device), "hipSetDevice");
        if (maxit <= 0) maxit = 10000;
        if (krylov <= 0) krylov = 20;  // ARPACK NCV floor used by eigs_rs (eigs.cpp:400)
        auto op_of = [](imps_plan q) {
            LinOp op;
            op.n = q->prog.x_elems;
            op.apply = [q](const double* a, double* b) { run(q, a, b, 1); };
            return op;
        };
        KrylovStats st;
        if (P->prec == IMPS_FP64 || !P64) {
            st = lanczos_lowest(s, op_of(P), x, tol, maxit, krylov);
        } else {
            // The low-precision phase stops at about 100x the relative rounding of its products
            // (FP32 ~1e-7, BF16X3 ~1e-6); nearer that floor its Ritz vector carries rounding-
            // correlated error (handing over at 1e-6, FP32 VUMPS at D = 64 stalled at a stationary
            // point 2e-8 above the FP64 energy), and the FP64 phase finishes the solve.
            // A start vector that already meets the hand-over tolerance (the later iterations of a
            // converging VUMPS: the previous eigenvector) skips the low-precision phase, whose first
            // cycle would take `krylov` steps before its convergence is read.
            const double tlo = std::max(tol, P->prec == IMPS_FP32 ? 1e-5 : 1e-4);
            KrylovStats a;
            a.iters = 1;
            if (rayleigh_relres(s, op_of(P64), x) > tlo) {
                a = lanczos_lowest(s, op_of(P), x, tlo, maxit, krylov);
                a.iters += 1;
            }
            st = lanczos_lowest(s, op_of(P64), x, tol, std::max(1, maxit - a.iters), krylov);
            st.iters += a.iters;
        }
        if (eval) *eval = st.value;
        if (iters) *iters = st.iters;
        if (resid) *resid = st.relres;
        if (!st.converged) throw NotConverged("eigsh: relative residual " + std::to_string(st.relres));
        return IMPS_OK;
    });
}

// ---- block products (A6) ----
imps_status imps_prod_plan(imps_context ctx, imps_plan* plan, int kind, const imps_layout* X1, const imps_layout* X2) {
    return guarded([&] {
        if (!plan) throw InvalidArg("plan is NULL");
        std::unique_ptr<imps_plan_s> P(new_plan(ctx, K_PROD, IMPS_FP64));
        Layout a = layout_from_c(X1, "X1"), b = layout_from_c(X2, "X2");
        P->prog = build_prod(kind, a, b, P->out_layout);
        P->op_set = true;
        return finish_plan(plan, P);
    });
}
imps_status imps_prod_apply(imps_plan P, const double* x1, const double* x2, double* y) {
    return imps_prod_apply_n(P, x1, x2, y, 1);
}

imps_status imps_prod_apply_n(imps_plan P, const double* x1, const double* x2, double* y, int nvec) {
    return guarded([&] {
        check_kind(P, K_PROD);
        if (!x2 && P->prog.op_elems > 0) throw InvalidArg("x2 is NULL");
        run(P, x1, y, nvec, x2);
        return IMPS_OK;
    });
}

// ---- per-sector QR (A18) ----
imps_status imps_qr_plan(imps_context ctx, imps_plan* plan, const imps_layout* A, int dir) {
    return guarded([&] {
        if (!plan) throw InvalidArg("plan is NULL");
        if (dir != IMPS_LEFT && dir != IMPS_RIGHT) throw InvalidArg("dir must be IMPS_LEFT or IMPS_RIGHT");
        std::unique_ptr<imps_plan_s> P(new_plan(ctx, K_QR, IMPS_FP64));
        Layout a = layout_from_c(A, "A");
        P->qr = build_qr(a, dir);
        P->qr_dir = dir;
        P->out_layout = P->qr.R;
        P->prog.x_elems = a.nelem;
        P->prog.y_elems = a.nelem;
        P->op_set = true;
        if (ctx) {
            hip_check(hipSetDevice(ctx->device), "hipSetDevice");
            const size_t b1 = P->qr.secs.size() * sizeof(QrSector), b2 = P->qr.parts.size() * sizeof(QrPart);
            P->d_prog.ensure(b1 + b2 + 64);
            if (b1) hip_check(hipMemcpy(P->d_prog.p, P->qr.secs.data(), b1, hipMemcpyHostToDevice), "upload");
            if (b2) hip_check(hipMemcpy((char*)P->d_prog.p + ((b1 + 15) & ~size_t(15)), P->qr.parts.data(), b2, hipMemcpyHostToDevice), "upload");
            if (P->qr.gwork > 0) P->d_w.ensure((size_t)P->qr.gwork * sizeof(double));  // work areas of the sectors beyond LDS
        }
        *plan = P.release();
        return IMPS_OK;
    });
}
imps_status imps_qr_apply(imps_plan P, const double* A, double* Q, double* R) {
    return guarded([&] {
        check_kind(P, K_QR);
        if (!P->ctx) throw InvalidArg("plan was created without a context");
        if (!A || !Q || !R) throw InvalidArg("NULL buffer");
        hip_check(hipSetDevice(P->ctx->device), "hipSetDevice");
        const size_t b1 = P->qr.secs.size() * sizeof(QrSector);
        const QrSector* secs = (const QrSector*)P->d_prog.p;
        const QrPart* parts = (const QrPart*)((char*)P->d_prog.p + ((b1 + 15) & ~size_t(15)));
        hip_check(launch_sector_qr(secs, P->qr.nlds, (int)P->qr.secs.size() - P->qr.nlds, parts, A, Q, R, P->qr_dir == IMPS_RIGHT, P->qr.lds,
                                   P->qr.gwork > 0 ? (double*)P->d_w.p : nullptr, P->ctx->stream),
                  "sector qr");
        return IMPS_OK;
    });
}

imps_status imps_env_tm_apply(imps_plan P, const double* x, double* y) {
    return guarded([&] {
        check_kind(P, K_ENV);
        if (!x || !y) throw InvalidArg("NULL vector");
        P->env->tm_apply(x, y);
        return IMPS_OK;
    });
}
imps_status imps_env_fixed_point(imps_plan P, double* x, double tol, int maxit, int* iters, double* delta) {
    return guarded([&] {
        check_kind(P, K_ENV);
        if (!x) throw InvalidArg("NULL vector");
        double dl = 0;
        const int it = P->env->fixed_point(x, tol, maxit <= 0 ? 100000 : maxit, &dl);
        if (iters) *iters = it;
        if (delta) *delta = dl;
        if (dl >= tol) throw NotConverged("fixed point: last change " + std::to_string(dl));
        return IMPS_OK;
    });
}


// ---- excitations ---------------------------------------------------------------------------------
imps_status imps_heff_plan(imps_context ctx, imps_plan* plan, int ncell, int d, const imps_layout* AL, const imps_layout* AR,
                           const imps_layout* NL, const imps_layout* C, const double* H, double E0, const int* K, int nmom,
                           int prec) {
    return guarded([&] {
        if (!plan || !AL || !AR || !NL || !C || !K) throw InvalidArg("plan, layouts or K is NULL");
        if (prec != IMPS_FP64 && prec != IMPS_FP32 && prec != IMPS_BF16X3) throw InvalidArg("prec must be IMPS_FP64, IMPS_FP32 or IMPS_BF16X3");
        if (ncell < 1 || ncell > 64) throw InvalidArg("ncell out of range");
        std::unique_ptr<imps_plan_s> P(new_plan(ctx, K_HEFF, IMPS_FP64));
        std::vector<Layout> la, lr, ln, lc;
        for (int k = 0; k < ncell; ++k) {
            la.push_back(layout_from_c(&AL[k], "AL"));
            lr.push_back(layout_from_c(&AR[k], "AR"));
            ln.push_back(layout_from_c(&NL[k], "NL"));
            lc.push_back(layout_from_c(&C[k], "C"));
        }
        if (ctx) hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        P->heff = std::make_shared<HeffEngine>();
        P->heff->build(ctx, ncell, d, la, lr, ln, lc, H, E0, K, nmom, prec);
        P->out_layout = P->heff->Xin.lay;
        *plan = P.release();
        return IMPS_OK;
    });
}

imps_status imps_opeff_plan(imps_context ctx, imps_plan* plan, int ncell, int d, const imps_layout* AL, const imps_layout* AR,
                            const imps_layout* NL, const imps_layout* C, const double* O, double O0, const int* K, int nmom, int prec) {
    return guarded([&] {
        if (!plan || !AL || !AR || !NL || !C || !K || !O) throw InvalidArg("plan, layouts, O or K is NULL");
        if (ncell < 1 || ncell > 64) throw InvalidArg("ncell out of range");
        if (prec != IMPS_FP64 && prec != IMPS_FP32 && prec != IMPS_BF16X3) throw InvalidArg("prec must be IMPS_FP64, IMPS_FP32 or IMPS_BF16X3");
        std::unique_ptr<imps_plan_s> P(new_plan(ctx, K_HEFF, IMPS_FP64));
        std::vector<Layout> la, lr, ln, lc;
        for (int k = 0; k < ncell; ++k) {
            la.push_back(layout_from_c(&AL[k], "AL"));
            lr.push_back(layout_from_c(&AR[k], "AR"));
            ln.push_back(layout_from_c(&NL[k], "NL"));
            lc.push_back(layout_from_c(&C[k], "C"));
        }
        if (ctx) hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        P->heff = std::make_shared<HeffEngine>();
        P->heff->build(ctx, ncell, d, la, lr, ln, lc, O, O0, K, nmom, prec, 1);
        P->out_layout = P->heff->Xin.lay;
        *plan = P.release();
        return IMPS_OK;
    });
}

imps_status imps_heff_set_state(imps_plan P, const double* const* AL, const double* const* AR, const double* const* NL,
                                const double* const* C, double inv_tol) {
    return guarded([&] {
        check_kind(P, K_HEFF);
        if (!AL || !AR || !NL || !C) throw InvalidArg("tensor arrays are NULL");
        const int N = P->heff->N;
        P->heff->set_state(std::vector<const double*>(AL, AL + N), std::vector<const double*>(AR, AR + N),
                           std::vector<const double*>(NL, NL + N), std::vector<const double*>(C, C + N), inv_tol);
        P->op_set = true;
        return IMPS_OK;
    });
}

imps_status imps_heff_set_state_lr(imps_plan P, const double* const* AL, const double* const* AR, const double* const* NL,
                                   const double* const* C, const imps_layout* Llay, const double* L, const imps_layout* LMlay,
                                   const double* LM, const imps_layout* RMlay, const double* RM, double inv_tol) {
    return guarded([&] {
        check_kind(P, K_HEFF);
        if (!AL || !AR || !NL || !C) throw InvalidArg("tensor arrays are NULL");
        HeffEngine::BondInput b;
        b.Llay = layout_from_c(Llay, "heff L");
        b.LMlay = layout_from_c(LMlay, "heff LM");
        b.RMlay = layout_from_c(RMlay, "heff RM");
        b.L = L;
        b.LM = LM;
        b.RM = RM;
        const int N = P->heff->N;
        P->heff->set_state(std::vector<const double*>(AL, AL + N), std::vector<const double*>(AR, AR + N),
                           std::vector<const double*>(NL, NL + N), std::vector<const double*>(C, C + N), inv_tol, &b);
        P->op_set = true;
        return IMPS_OK;
    });
}

imps_status imps_heff_apply(imps_plan P, const double* p, const double* x, double* y, int* gmres_steps, double* gmres_relres) {
    return guarded([&] {
        check_kind(P, K_HEFF);
        P->heff->stats = HeffStats();
        int steps = 0;
        double worst = 0;
        // the engine records the statistics of each geometric sum; report their sum and worst residual
        P->heff->apply(p, x, y);
        steps = P->heff->stats_steps_total;
        worst = P->heff->stats_relres_worst;
        hip_check(hipStreamSynchronize(P->ctx->stream), "hipStreamSynchronize");
        if (gmres_steps) *gmres_steps = steps;
        if (gmres_relres) *gmres_relres = worst;
        return IMPS_OK;
    });
}

imps_status imps_heff_eigs(imps_plan P, const double* p, int nev, double tol, int maxit, int krylov, double* evals, double* evecs,
                           double* resid, int* napply) {
    return guarded([&] {
        check_kind(P, K_HEFF);
        int conv = 0;
        const int n = P->heff->eigs(p, nev, tol, maxit, krylov, evals, evecs, resid, &conv);
        if (napply) *napply = n;
        if (!conv) {
            g_err = "heff eigs: not converged within maxit applications";
            return IMPS_NOT_CONVERGED;
        }
        return IMPS_OK;
    });
}

imps_status imps_heff_check(imps_plan P, double* trHL_R, double* trL_HR, int* env_iters) {
    return guarded([&] {
        check_kind(P, K_HEFF);
        if (!P->heff->state_set) throw InvalidArg("heff: state not set");
        if (trHL_R) *trHL_R = P->heff->check_trL;
        if (trL_HR) *trL_HR = P->heff->check_trR;
        if (env_iters) *env_iters = P->heff->env_iters;
        return IMPS_OK;
    });
}


// ---- adjoints ----------------------------------------------------------------------------------
imps_status imps_tm_vjp(imps_plan P, const double* A, const double* B, const double* x, const double* ybar, double* xbar, double* Abar,
                        double* Bbar) {
    return guarded([&] {
        check_kind(P, K_TM);
        if (P->tm_mixed && !B) throw InvalidArg("plan was built for a mixed TM: B is required");
        if (!P->tm_mixed && (B || Bbar)) throw InvalidArg("regular TM: B and Bbar must be NULL (Abar holds both roles)");
        vjp_run(P, {A, P->tm_mixed ? B : A}, x, ybar, {xbar, Abar, Bbar});
        return IMPS_OK;
    });
}

imps_status imps_hac_vjp(imps_plan P, const double* EHL, const double* EHR, const double* HL, const double* HR, const double* x,
                         const double* ybar, double* xbar, double* EHLbar, double* EHRbar, double* HLbar, double* HRbar) {
    return guarded([&] {
        check_kind(P, K_HAC);
        vjp_run(P, {EHL, EHR, HL, HR}, x, ybar, {xbar, EHLbar, EHRbar, HLbar, HRbar});
        return IMPS_OK;
    });
}

imps_status imps_hc_vjp(imps_plan P, const double* EHL, const double* EHR, const double* HL, const double* AR, const double* x,
                        const double* ybar, double* xbar, double* EHLbar, double* EHRbar, double* HLbar, double* ARbar) {
    return guarded([&] {
        check_kind(P, K_HC);
        vjp_run(P, {EHL, EHR, HL, AR}, x, ybar, {xbar, EHLbar, EHRbar, HLbar, ARbar});
        return IMPS_OK;
    });
}

}  // extern "C"
