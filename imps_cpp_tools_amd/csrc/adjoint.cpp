// Adjoint tapes of the TM, H_AC and H_C plans (adjoint.hpp).
#include "adjoint.hpp"

namespace imps {

namespace {

// The gradient combination onto a fixed layout: entry e of the output sums the listed single-list
// objects (coefficient 1).
XObj gather_onto(BlockTape& T, const Layout& lay, const std::vector<std::vector<const XObj*>>& parts, const XObj& dest) {
    std::vector<std::vector<BlockTape::Src>> rows(lay.l.size());
    const int one = T.coef(1.0, 0);
    for (size_t e = 0; e < rows.size(); ++e)
        for (const XObj* o : parts[e]) rows[e].push_back(BlockTape::Src{one, o, 0});
    return T.comb(rows, &lay, &dest);
}

void record_tm(VjpTape& T, int dir, bool mixed) {
    const XObj &Ao = T.ins[0], &Bo = T.ins[1], &X = T.ins[2], &Y = T.ins[3];
    const Layout &Al = Ao.lay, &xl = X.lay;
    XObj xb, ab, bb;
    if (dir == IMPS_LEFT) {
        xb = T.tmr(Bo, Ao, Y);                                                       // sum_s B_s ybar A_s^T
        ab = T.prod(PROD_MPS_MAT, T.prod(PROD_MATT_MPS, X, Bo), Y);                 // x^T B_s ybar
        bb = T.prod(PROD_MPS_MATT, T.prod(PROD_MAT_MPS, X, Ao), Y);                 // x A_s ybar^T
    } else {
        xb = T.tml(Bo, Ao, Y);                                                       // sum_s A_s^T ybar B_s
        ab = T.prod(PROD_MPS_MATT, T.prod(PROD_MAT_MPS, Y, Bo), X);                 // ybar B_s x^T
        bb = T.prod(PROD_MPS_MAT, T.prod(PROD_MATT_MPS, Y, Ao), X);                 // ybar^T A_s x
    }
    T.onto(xb, xl, &T.outs[0]);
    if (mixed) {
        T.onto(ab, Al, &T.outs[1]);
        T.onto(bb, Bo.lay, &T.outs[2]);
    } else {
        // regular TM: A plays both roles
        std::vector<std::vector<BlockTape::Src>> rows(Al.l.size());
        const int one = T.coef(1.0, 0);
        for (size_t e = 0; e < rows.size(); ++e) rows[e] = {BlockTape::Src{one, &ab, (int)e}, BlockTape::Src{one, &bb, (int)e}};
        T.comb(rows, &Al, &T.outs[1]);
    }
}

void record_hac(VjpTape& T) {
    const XObj &EHL = T.ins[0], &EHR = T.ins[1], &HL = T.ins[2], &HR = T.ins[3], &X = T.ins[4], &Y = T.ins[5];
    const int d = (int)X.lay.l.size();
    const XObj t1 = T.prod(PROD_MATT_MPS, EHL, Y);  // EHL^T ybar[j]
    const XObj t2 = T.prod(PROD_MPS_MATT, Y, EHR);  // ybar[j] EHR^T
    std::vector<XObj> Ye, Xe;
    for (int i = 0; i < d; ++i) {
        Ye.push_back(T.view_entry(Y, i));
        Xe.push_back(T.view_entry(X, i));
    }
    std::vector<XObj> p1, p2;  // [i * d + j]
    for (int i = 0; i < d; ++i)
        for (int j = 0; j < d; ++j) {
            p1.push_back(T.prod(PROD_MATT_MPS, T.view_entry(HL, i * d + j), Ye[i]));  // HL[i][j]^T ybar[i]
            p2.push_back(T.prod(PROD_MPS_MATT, Ye[i], T.view_entry(HR, i * d + j)));  // ybar[i] HR[i][j]^T
        }
    {
        std::vector<std::vector<BlockTape::Src>> rows(d);
        const int one = T.coef(1.0, 0);
        for (int j = 0; j < d; ++j) {
            rows[j] = {BlockTape::Src{one, &t1, j}, BlockTape::Src{one, &t2, j}};
            for (int i = 0; i < d; ++i) {
                rows[j].push_back(BlockTape::Src{one, &p1[i * d + j], 0});
                rows[j].push_back(BlockTape::Src{one, &p2[i * d + j], 0});
            }
        }
        T.comb(rows, &X.lay, &T.outs[0]);
    }
    T.onto(T.prod(PROD_TM_RIGHT, Y, X), EHL.lay, &T.outs[1]);  // sum_i ybar[i] x[i]^T
    T.onto(T.prod(PROD_TM_LEFT, Y, X), EHR.lay, &T.outs[2]);   // sum_i x[i]^T ybar[i]
    std::vector<XObj> gl, gr;
    for (int i = 0; i < d; ++i)
        for (int j = 0; j < d; ++j) {
            gl.push_back(T.prod(PROD_MPS_MATT, Ye[i], Xe[j]));  // ybar[i] x[j]^T
            gr.push_back(T.prod(PROD_MATT_MPS, Xe[j], Ye[i]));  // x[j]^T ybar[i]
        }
    std::vector<std::vector<const XObj*>> pl(d * d), pr(d * d);
    for (int k = 0; k < d * d; ++k) {
        pl[k] = {&gl[k]};
        pr[k] = {&gr[k]};
    }
    gather_onto(T, HL.lay, pl, T.outs[3]);
    gather_onto(T, HR.lay, pr, T.outs[4]);
}

void record_hc(VjpTape& T) {
    const XObj &EHL = T.ins[0], &EHR = T.ins[1], &HL = T.ins[2], &AR = T.ins[3], &X = T.ins[4], &Y = T.ins[5];
    const int d = (int)AR.lay.l.size();
    std::vector<XObj> ARe, YAR, XAR;
    for (int i = 0; i < d; ++i) {
        ARe.push_back(T.view_entry(AR, i));
        YAR.push_back(T.prod(PROD_MAT_MPS, Y, ARe[i]));  // ybar AR[i]
        XAR.push_back(T.prod(PROD_MAT_MPS, X, ARe[i]));  // C AR[i] (recomputed)
    }
    std::vector<XObj> cb, hlb, arb_j, arb_i;
    std::vector<std::vector<const XObj*>> hlparts(d * d), arparts(d);
    cb.push_back(T.prod(PROD_MATT_MPS, EHL, Y));  // EHL^T ybar
    cb.push_back(T.prod(PROD_MPS_MATT, Y, EHR));  // ybar EHR^T
    std::vector<XObj> Z((size_t)d * d);
    for (int i = 0; i < d; ++i)
        for (int j = 0; j < d; ++j) Z[i * d + j] = T.prod(PROD_MPS_MATT, YAR[i], ARe[j]);  // ybar AR[i] AR[j]^T
    for (int i = 0; i < d; ++i)
        for (int j = 0; j < d; ++j) {
            const XObj h = T.view_entry(HL, i * d + j);
            cb.push_back(T.prod(PROD_MATT_MPS, h, Z[i * d + j]));                                   // HL^T ybar AR[i] AR[j]^T
            hlb.push_back(T.prod(PROD_MPS_MATT, Z[i * d + j], X));                                  // ybar AR[i] AR[j]^T C^T
            arb_j.push_back(T.prod(PROD_MATT_MPS, X, T.prod(PROD_MATT_MPS, h, YAR[i])));            // C^T HL^T ybar AR[i] -> AR[j]
            arb_i.push_back(T.prod(PROD_MATT_MPS, Y, T.prod(PROD_MAT_MPS, h, XAR[j])));             // ybar^T HL C AR[j] -> AR[i]
        }
    {
        std::vector<std::vector<BlockTape::Src>> rows(1);
        const int one = T.coef(1.0, 0);
        for (auto& o : cb) rows[0].push_back(BlockTape::Src{one, &o, 0});
        T.comb(rows, &X.lay, &T.outs[0]);
    }
    T.onto(T.prod(PROD_MPS_MATT, Y, X), EHL.lay, &T.outs[1]);  // ybar C^T
    T.onto(T.prod(PROD_MATT_MPS, X, Y), EHR.lay, &T.outs[2]);  // C^T ybar
    for (int k = 0; k < d * d; ++k) hlparts[k] = {&hlb[k]};
    gather_onto(T, HL.lay, hlparts, T.outs[3]);
    for (int i = 0; i < d; ++i)
        for (int j = 0; j < d; ++j) {
            arparts[j].push_back(&arb_j[i * d + j]);
            arparts[i].push_back(&arb_i[i * d + j]);
        }
    gather_onto(T, AR.lay, arparts, T.outs[4]);
}

std::shared_ptr<VjpTape> compile(imps_plan_s* P) {
    auto T = std::make_shared<VjpTape>();
    T->ctx = P->ctx;
    T->nmom = 1;
    const std::vector<Layout>& L = P->lays;
    T->nsym = L.empty() ? 1 : L[0].nsym;
    if (!L.empty()) for (int i = 0; i < kMaxSym; ++i) T->per[i] = L[0].per[i];
    T->coef(1.0, 0);
    T->rec = &T->tape;
    if (P->kind == K_TM) {
        // ins: A, B, x, ybar (y layout); outs: xbar, Abar, Bbar
        T->ins = {T->make(L[0], false), T->make(L[1], false), T->make(L[2], false), T->make(L[3], false)};
        T->outs = {T->make(L[2], false), T->make(L[0], false), T->make(L[1], false)};
        record_tm(*T, P->tm_dir, P->tm_mixed);
    } else if (P->kind == K_HAC) {
        // ins: EHL, EHR, HL, HR, x, ybar; outs: xbar, EHLbar, EHRbar, HLbar, HRbar
        T->ins = {T->make(L[1], false), T->make(L[2], false), T->make(L[3], false), T->make(L[4], false), T->make(L[0], false),
                  T->make(L[0], false)};
        T->outs = {T->make(L[0], false), T->make(L[1], false), T->make(L[2], false), T->make(L[3], false), T->make(L[4], false)};
        record_hac(*T);
    } else if (P->kind == K_HC) {
        T->ins = {T->make(L[1], false), T->make(L[2], false), T->make(L[3], false), T->make(L[4], false), T->make(L[0], false),
                  T->make(L[0], false)};
        T->outs = {T->make(L[0], false), T->make(L[1], false), T->make(L[2], false), T->make(L[3], false), T->make(L[4], false)};
        record_hc(*T);
    } else {
        throw InvalidArg("vjp: the plan has no adjoint");
    }
    T->rec = nullptr;
    T->finalize();
    T->upload_coefs();
    return T;
}

}  // namespace

void vjp_run(imps_plan_s* P, const std::vector<const double*>& ops, const double* x, const double* ybar, const std::vector<double*>& outs) {
    if (!P->ctx) throw InvalidArg("plan was created without a context");
    if (!x || !ybar) throw InvalidArg("vjp: x or ybar is NULL");
    hip_check(hipSetDevice(P->ctx->device), "hipSetDevice");
    if (!P->vjp) P->vjp = compile(P);
    VjpTape& T = *P->vjp;
    hipStream_t s = P->ctx->stream;
    std::vector<const double*> in(ops);
    in.push_back(x);
    in.push_back(ybar);
    if (in.size() != T.ins.size() || outs.size() != T.outs.size()) throw InvalidArg("vjp: wrong operand count");
    for (size_t k = 0; k < in.size(); ++k) {
        const long n = T.ins[k].nelem();
        if (n > 0 && !in[k]) throw InvalidArg("vjp: an operand pointer is NULL");
        if (n > 0) hip_check(hipMemcpyAsync(T.A() + T.ins[k].off, in[k], sizeof(double) * n, hipMemcpyDeviceToDevice, s), "vjp in");
    }
    T.run_tape(T.tape);
    for (size_t k = 0; k < outs.size(); ++k) {
        const long n = T.outs[k].nelem();
        if (outs[k] && n > 0) hip_check(hipMemcpyAsync(outs[k], T.A() + T.outs[k].off, sizeof(double) * n, hipMemcpyDeviceToDevice, s), "vjp out");
    }
}

}  // namespace imps
