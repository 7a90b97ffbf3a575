// Device Krylov drivers: restarted GMRES for the environment (geometric-sum) equations and
// restarted Lanczos for the lowest eigenpair of H_AC / H_C.  Vectors stay in HBM in FP64; the
// small Hessenberg / tridiagonal problems are solved on the host in FP64.
#pragma once

#include <functional>
#include <memory>
#include <vector>

#include "plan.hpp"

namespace imps {

// y = A(x) on device FP64 vectors of length n, ordered on stream s
struct LinOp {
    long n = 0;
    std::function<void(const double* x, double* y)> apply;
};

struct KrylovStats {
    int iters = 0;          // operator applications
    double relres = 0;      // final relative residual (GMRES) or residual norm / |eigenvalue| (Lanczos)
    double value = 0;       // eigenvalue (Lanczos)
    bool converged = false;
};

// Restarted GMRES(m) with CGS2: solves A x = b from the initial x; stops at ||b - A x|| < tol ||b||.
// Mirrors the stopping rule of gmres (IterativeSolvers.hpp:119) with restarts added.
KrylovStats gmres(hipStream_t s, const LinOp& A, const double* b, double* x, double tol, int maxit, int m);

// Mixed-precision refinement: outer FP64 residuals with A64, inner GMRES on the low-precision Alo
// (relative tolerance inner_tol, at most 4 restart cycles), x += correction, until
// ||b - A64 x|| < tol ||b|| or maxit operator applications in total.
KrylovStats gmres_refined(hipStream_t s, const LinOp& A64, const LinOp& Alo, const double* b, double* x, double tol, int maxit,
                          int m, double inner_tol);

// Lowest eigenpair of a symmetric operator ("SA", nev = 1, as eigs_rs in VUMPS_main.cpp:186):
// restarted Lanczos with full FP64 reorthogonalisation (CGS2), restarting from the Ritz vector.
// x: start vector in, normalised eigenvector out.  Converged when ||A x - theta x|| <= tol |theta|.
KrylovStats lanczos_lowest(hipStream_t s, const LinOp& A, double* x, double tol, int maxit, int m);

// ||A x - theta x|| / (|theta| ||x||), theta = x.Ax / x.x (one application; resolves down to ~1e-8).
double rayleigh_relres(hipStream_t s, const LinOp& A, const double* x);

// Environment solver for one unit cell (InvertE_proj, InvEBlockFunctions.hpp:205-266):
//   dir l: (y|[1 - T + |R)(L|] = (x|[1 - |R)(L|],  T = T_L[A_{N-1}] ... T_L[A_0]
//   dir r: [1 - T + |R)(L|]|y) = [1 - |R)(L|]|x),  T = T_R[A_0] ... T_R[A_{N-1}]
struct EnvSolver {
    int dir = 0, prec = IMPS_FP64, ncell = 0;
    imps_context ctx = nullptr;
    Layout xlay;
    std::vector<std::unique_ptr<imps_plan_s>> tm64, tmlo;  // chain plans in application order
    std::vector<long> bond_elems;                           // elements of each intermediate bond
    DevBuf bonds, rhoT, rhs, tmp;
    bool mps_set = false;
    void build(imps_context c, int direction, const std::vector<Layout>& A, const Layout& x, int precision);
    void set_mps(const std::vector<const double*>& A);
    KrylovStats solve(const double* x, const double* L, const double* R, const double* y0, double* y, double tol, int maxit, int krylov);
    // y = T(x), the unit-cell transfer matrix of the chain (FP64)
    void tm_apply(const double* x, double* y);
    // dominant fixed point of the trace-preserving cell TM from the start x (overwritten), trace
    // normalised (TMDominantEig, EigsBlockFunctions.hpp:91-172), by a correction solve of the
    // environment form; returns operator applications, *delta = ||T(x) - x|| of the result
    int fixed_point(double* x, double tol, int maxit, double* delta);
};

}  // namespace imps
