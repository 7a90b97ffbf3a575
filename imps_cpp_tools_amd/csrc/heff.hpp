// Excitation engine: the momentum-k effective Hamiltonian of the quasiparticle ansatz (ApplyHeff,
// Excitations_helpers.hpp:20-257) with its constants (HeffConstants, :260-347), for a batch of
// momenta at once.
//
// Design.  Every block object of ApplyHeff has a structure that depends only on the ground state's
// sectors and the excitation charge K, so the whole operator is compiled once, when the plan is
// made, into a tape of device operations over one arena of doubles:
//   * products of a complex object (X and what is derived from it) with a real one (AL, AR, NL,
//     the environments HLtot / HRtot, the reduced operators) as grouped block-GEMM launches whose
//     vector index runs over the 2 * nmom real and imaginary planes of all momenta;
//   * sums over sector unions, restrictions, scalings by e^{+-ik} and ApplyOperator as block-level
//     gather-scale-sums (one launch each);
//   * the two geometric sums as one momentum-batched complex GMRES each, whose transfer-matrix
//     chains run on the panel kernels with 2 * nmom vectors per launch.
// The constants (HLtot, HRtot, the reduced operators, AL AL products) are a second tape, run once
// per state; their two environment solves use the FP64 environment solver.
#pragma once

#include <complex>
#include <functional>
#include <memory>
#include <vector>

#include "cplx.hpp"
#include "krylov.hpp"
#include "plan.hpp"

namespace imps {

using cplx = std::complex<double>;

// A block object in the arena: plane 0 at `off`; complex objects have 2 * nmom planes `ps` apart
// (momentum t: real plane 2 t, imaginary plane 2 t + 1), real objects one plane.
struct XObj {
    Layout lay;
    long off = 0;
    long ps = 0;
    bool cplx = false;
    long nelem() const { return lay.nelem; }
};

struct HeffEngine;

struct HeffStats {
    int gmres_steps = 0;  // inner GMRES operator applications (batched over momenta) of the last apply
    double gmres_relres = 0;
};

// A factorised spec of one ApplyHeff coefficient: value * kfac^pow (pow in {-1, 0, 1}).
struct CoefSpec {
    cplx v;
    int pow;
};

// A recorded program of block-object operations over one device arena (the base of the excitation
// engine and of the adjoint plans).  Objects are made while recording; `finalize` uploads the
// recorded tables and allocates the arena.
struct BlockTape {
    imps_context ctx = nullptr;
    int nsym = 1, d = 1, nmom = 1;
    int per[kMaxSym] = {0, 0, 0, 0};  // charge periods of the layouts (0: U(1), N > 1: Z(N))
    std::vector<double> H0;  // d^2 x d^2 column-major two-site operator used by apply_operator/redop

    long arena_elems = 0;
    DevBuf arena, coef_dev, progs_dev;
    std::vector<CoefSpec> coefs;
    double* A() const { return arena.d(); }

    struct ProdOp {
        Program prog;
        std::vector<DevStage> dst;
        XObj a, b, out;
    };
    struct CombOp {
        std::vector<CombTile> tiles;
        std::vector<CombTerm> terms;
        const CombTile* dtiles = nullptr;
        const CombTerm* dterms = nullptr;
        long out_ps = 0;
        bool out_cplx = false;
    };
    struct Op {
        int type;  // 0 prod, 1 comb, >= 2 engine-specific
        int idx;
    };
    std::vector<ProdOp> prods;
    std::vector<CombOp> combs;
    std::vector<Op>* rec = nullptr;
    std::vector<double> p_cur;  // momenta of the current batch (coefficient powers of e^{i p pi N})
    int ncell_phase = 1;        // N of the phase e^{i p pi N}

    virtual ~BlockTape() = default;
    long alloc(long nelem, bool cplx);
    XObj make(const Layout& lay, bool cplx);
    XObj view_entry(const XObj& x, int e) const;
    int coef(cplx v, int pow);
    XObj prod(int kind, const XObj& a, const XObj& b);
    struct Src {
        int coef;
        const XObj* obj;
        int entry;
    };
    // out entry e = sum over rows[e] of coef * entry of obj (sector union in term order; or onto
    // the fixed layout `fixed` when given, dropping source blocks outside it)
    XObj comb(const std::vector<std::vector<Src>>& rows, const Layout* fixed = nullptr, const XObj* dest = nullptr, bool force_cplx = false);
    XObj add(const XObj& a, const XObj& b, cplx ca = 1.0, cplx cb = 1.0, int powb = 0);
    XObj scale(const XObj& a, cplx c, int pow);
    XObj onto(const XObj& a, const Layout& lay, const XObj* dest = nullptr);
    XObj diag_part(const XObj& a);
    XObj apply_operator(const XObj& AA);
    XObj tml(const XObj& A, const XObj& B, const XObj& X) { return prod(PROD_TM_LEFT, prod(PROD_MAT_MPS, X, A), B); }
    XObj tmr(const XObj& A, const XObj& B, const XObj& X) { return prod(PROD_TM_RIGHT, prod(PROD_MPS_MAT, A, X), B); }
    XObj redop(const XObj& A, int dir);
    // uploads the recorded programs and combinations, allocates and zeroes the arena
    void finalize();
    void upload_coefs();
    void run_prod(const ProdOp& op);
    void run_comb(const CombOp& op);
    void run_tape(const std::vector<Op>& tape);
    virtual void run_other(const Op&) { throw InvalidArg("tape: unknown operation"); }
    hipStream_t stream() const { return ctx->stream; }
};

struct HeffEngine : BlockTape {
    // ---- structure ----
    int N = 0;
    Key K;
    std::vector<Layout> ALl, ARl, NLl, Cl;
    std::vector<Layout> xdims;
    long mtot = 0;
    double inv_tol = 1e-14;
    int krylov = 40;
    int gmres_maxit = 4000;
    // IMPS_FP64, or IMPS_FP32 / IMPS_BF16X3: the Arnoldi steps of the geometric sums apply their TM
    // chains at that precision; every restart takes the true FP64 residual (iterative refinement)
    int prec = IMPS_FP64;
    // 2: H_eff of the two-site Hamiltonian (H0 is d^2 x d^2); 1: the effective operator of a one-site
    // observable (ApplyOPeff with OPeffConstants; H0 is the d x d operator, HLtot / HRtot hold
    // OPLtot / OPRtot)
    int op_sites = 2;

    // ---- objects ----
    std::vector<XObj> AL, AR, NL, C;  // constants, filled by set_state
    std::vector<XObj> HLtot, HRtot, HLr, HRr;
    XObj Xin, Xout;                   // contiguous complex objects in XDims order (views per site)
    std::vector<XObj> Xin_v, Xout_v;
    // bond between the unit cells: bL on AL's side (right bond of AL[N-1]), bR on AR's side (left
    // bond of AR[0]); lmL / rmL: the charge-conserving blocks between them, (bR x bL) / (bL x bR)
    Layout bL, bR, lmL, rmL;
    XObj LMo, RMo;                    // dominant left / right eigenmatrices of the mixed TM (AL, AR)
    XObj oneL, oneR;                  // identities on bL and bR
    XObj Rb, Lb;                      // right fixed point C C^T of AL's TM, left fixed point of AR's TM
    std::vector<XObj> AAL;            // AL[n-1] AL[n] two-site products
    XObj checkHL, checkHR, checkL, checkR;  // operands of HeffConstants' trace check

    struct GeoOp {
        int dir = 0, pow = 1;
        XObj x, y;
        std::vector<std::unique_ptr<imps_plan_s>> chain;  // TM plans in application order
        std::vector<std::unique_ptr<imps_plan_s>> chain_lo;  // the same at the engine's low precision (mixed engines)
        std::vector<long> bond_elems;
        std::vector<std::pair<XObj, XObj>> chain_ops;     // (A, B) of each chain plan
        bool proj = false;
        XObj w, rout;         // real, on x's sectors: trace weights (rho_in^T) and rho_out
    };
    std::vector<std::unique_ptr<GeoOp>> geos;
    std::vector<std::function<void()>> calls;
    std::vector<Op> setup_tape, apply_tape;

    // ---- per-apply state ----
    DevBuf fac_dev, gam_dev;     // per geo op: kfac^pow per momentum (complex), projection flags
    DevBuf kry;                  // batched Krylov scratch of the geometric sums
    DevBuf eig_kry;              // Lanczos basis and scratch of the eigensolver
    int env_iters = 0;           // environment-solver operator applications of set_state
    bool state_set = false;
    HeffStats stats;             // of the last geometric sum
    int stats_steps_total = 0;   // GMRES operator applications of the last apply (both sums)
    double stats_relres_worst = 0;
    // environment solvers of the constants
    std::unique_ptr<EnvSolver> envL, envR;
    double check_trL = 0, check_trR = 0;  // tr(HL R), tr(L HR) of HeffConstants' check

    // build: layouts (AL, AR, NL per site; C[n] on bond n + 1), H (dense d^2 x d^2 column-major), E0, K
    void build(imps_context c, int ncell, int dphys, const std::vector<Layout>& al, const std::vector<Layout>& ar,
               const std::vector<Layout>& nl, const std::vector<Layout>& cl, const double* H, double E0, const int* Kq, int nmomenta,
               int precision = IMPS_FP64, int sites = 2);
    // Bond objects given by the caller in place of the closed forms L = C^T C, LM = C^T, RM = C (an AR
    // that is not C^-1 AL C, e.g. transformed or shifted): each with its own layout, mapped onto the
    // engine's by charge
    struct BondInput {
        Layout Llay, LMlay, RMlay;
        const double *L = nullptr, *LM = nullptr, *RM = nullptr;
    };
    void set_state(const std::vector<const double*>& al, const std::vector<const double*>& ar, const std::vector<const double*>& nl,
                   const std::vector<const double*>& cl, double tol, const BondInput* bond = nullptr);
    // y = H_eff(p_t) x for every momentum t (interleaved complex [nmom][mtot] device vectors)
    void apply(const double* p, const double* x, double* y);
    // lowest nev eigenpairs per momentum (Hermitian thick-restart Lanczos, batched; evecs
    // interleaved complex [nmom][nev][mtot] or NULL); returns the batched applications
    int eigs(const double* p, int nev, double tol, int maxit, int m, double* evals, double* evecs, double* resid, int* converged);
    // y = H_eff x on the engine's planar vectors (Xin / Xout of the arena)
    void apply_planar(const double* xin, double* xout);

    XObj redop_left_mixed(const XObj& Abra, const XObj& Bket);
    XObj redop_right_mixed(const XObj& Aket, const XObj& Bbra);
    XObj invert_e(const std::vector<XObj>& As, const std::vector<XObj>& Bs, const XObj& x, int dir, int pow, bool proj);
    void record_setup();
    void record_apply();
    void record_apply_observable();
    // sum_c O(r, c) A[c] per physical index r for the one-site operator H0 (ApplyOperator with a
    // one-site operator, MPSBlockUtilities.hpp:18-34)
    XObj apply_site_operator(const XObj& A);
    void run_geo(GeoOp& op);
    void run_other(const Op& op) override;
};

}  // namespace imps
