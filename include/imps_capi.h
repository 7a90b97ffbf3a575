/* imps_capi.h -- C ABI of the MI355X block-sparse iMPS operator library (libimps_hip.so).
 *
 * The library replaces the operator callbacks that MPShelpers hands to its Krylov drivers
 * (ARPACK eigs_rs/eigs_n, MPShelpers/include/eigs.h:43-55, and gmres,
 * MPShelpers/include/IterativeSolvers.hpp:21) with device-resident, grouped HIP kernels for gfx950.
 *
 * Conventions shared by every entry point
 *  - Plain C types only: ints, doubles, raw pointers.  Vectors and operator data are FP64 device
 *    pointers (hipMalloc'ed by the caller), laid out exactly as the reference's Vectorize():
 *    blocks concatenated in ascending ingoing-charge order, each block column-major
 *    (BlockObj.hpp:764-780, :1709-1726; MPSBlockMat.hpp:446-465).  An MPS tensor is s-major.
 *  - Block structures are described by imps_layout (below), the counterpart of the dims vectors
 *    the reference passes to its callbacks (dimkeypair_vec / dimpair_vec, DimMaps.hpp:31-201).
 *  - `nvec` applies the operator to nvec vectors stored back to back (x[v*len_x + i]); each
 *    output vector is fully overwritten, as the reference callbacks memset `out` first
 *    (VUMPS_helpers.hpp:69,118; EigsBlockFunctions.hpp:55,73).
 *  - Every function returns an imps_status; no C++ exception crosses this boundary.  On error,
 *    imps_last_error() returns a thread-local message.
 *  - A plan is not thread-safe.  Work is ordered on the context's HIP stream.  Different plans may
 *    be driven from different host threads at once (Krylov scratch and error messages are per
 *    thread); give each thread's plans their own context and stream to overlap the solves.  The
 *    library never retains caller pointers beyond a call; *_set_* calls copy the operator data
 *    on the plan's stream, so their sources must stay valid until that stream reaches the copy.
 *  - Precision (imps_prec) selects the arithmetic of the contractions.  Inputs and outputs stay FP64.
 */
#ifndef IMPS_CAPI_H
#define IMPS_CAPI_H

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    IMPS_OK = 0,
    IMPS_INVALID_ARG = 1,   /* malformed layout, mismatched sizes, unknown enum value */
    IMPS_NOT_CONVERGED = 2, /* iterative solver hit maxit before tol */
    IMPS_NUMERICAL = 3,     /* breakdown / non-finite values */
    IMPS_HIP_ERROR = 4,     /* a HIP runtime call failed */
    IMPS_NO_DEVICE = 5      /* no usable gfx950 device */
} imps_status;

typedef enum {
    IMPS_FP64 = 0,   /* FP64 MFMA (v_mfma_f64_16x16x4f64), FP64 accumulation */
    IMPS_FP32 = 1,   /* FP32 MFMA (v_mfma_f32_16x16x4f32), FP32 accumulation */
    IMPS_BF16X3 = 2  /* split-BF16: a = a_hi + a_lo, a*b ~ hi*hi + hi*lo + lo*hi on bf16 MFMA, FP32 accumulation */
} imps_prec;

typedef enum { IMPS_LEFT = 0, IMPS_RIGHT = 1 } imps_dir;  /* dirtype l / r (arma_typedefs.h) */

/* Block structure of a BlockMat (nlist = 1), BlockDiagMat (nlist = 1, qin == qout, square blocks),
 * MPSBlockMat (nlist = d) or RedOp (nlist = d*d, row-major grid, OperatorTypes.hpp:160-187).
 * Within each list entry blocks are in strictly ascending qin order (lexicographic over the nsym
 * charges), which is the std::map order of the reference.  `periods` is the symmetry group
 * (symobj::init, symobj.hpp:80-161): per charge 0 or 1 for U(1), N > 1 for Z(N), whose charges
 * must lie in [0, N); NULL means U(1) for every charge.  All layouts of one plan share the group. */
typedef struct {
    int nsym;          /* number of abelian charges per sector key, 1..4 */
    int nlist;         /* list entries */
    const int* nblk;   /* [nlist] blocks per entry */
    const int* qin;    /* [sum(nblk) * nsym] ingoing charges */
    const int* qout;   /* [sum(nblk) * nsym] outgoing charges */
    const int* rows;   /* [sum(nblk)] */
    const int* cols;   /* [sum(nblk)] */
    const int* periods; /* [nsym] charge periods, or NULL (all U(1)) */
} imps_layout;

typedef struct imps_context_s* imps_context;
typedef struct imps_plan_s* imps_plan;

/* ---- runtime ---------------------------------------------------------------------------- */
const char* imps_last_error(void);
const char* imps_version(void);
/* Binds `device`, creates a default-priority stream.  Fails with IMPS_NO_DEVICE if absent. */
imps_status imps_create(imps_context* ctx, int device);
/* Releases the caller's handle.  Plans created on the context keep it alive until the last of
 * them is destroyed, so contexts and plans may be destroyed in any order. */
imps_status imps_destroy(imps_context ctx);
/* Order all later work of this context on a caller stream (hipStream_t passed as void*); NULL
 * selects the legacy default stream.  Until called, the context uses a stream of its own. */
imps_status imps_set_stream(imps_context ctx, void* stream);
imps_status imps_synchronize(imps_context ctx);
imps_status imps_plan_destroy(imps_plan plan);
/* Algorithmic cost of one single-vector apply: flops (sum of 2*M*N*K over executed block GEMMs)
 * and bytes (each operand block read once and each output block written once at the working
 * precision).  Either pointer may be NULL. */
imps_status imps_plan_cost(imps_plan plan, double* flops, double* bytes);
/* Measurement hook (no counterpart in the reference; for benchmarks and profiling): with timing
 * enabled, every apply of the plan records HIP events on the plan's stream just before its first
 * and just after its last GEMM kernel (the per-apply split copy of the input is outside the pair).
 * A call returns the summed GPU milliseconds and the number of timed applies since the previous
 * call (waiting for the last event), resets both and then sets timing on (enable != 0) or off. */
imps_status imps_plan_timing(imps_plan plan, int enable, double* ms, int* launches);
/* Introspection (no counterpart in the reference): the schedule of the panel GEMM kernels for applies
 * of nvec vectors -- the autotuned one once the plan has applied nvec vectors, the default rule
 * before that or with IMPS_AUTOTUNE=0.  ksplit = waves sharing one vector's output tile (1, 2, 4;
 * 0 for plans without panel kernels), shared = 1 when panel fragments are shared through LDS. */
imps_status imps_plan_schedule(imps_plan plan, int nvec, int* ksplit, int* shared);
/* Number of grouped kernel launches per apply and total work tiles per vector. */
imps_status imps_plan_shape(imps_plan plan, int* launches, int* tiles);
/* Output structure of a plan whose output layout the library derives (imps_redop_plan):
 * first call with qin == NULL to get nlist, total blocks and total elements. */
imps_status imps_plan_out_layout(imps_plan plan, int* nlist, int* nblk_total, long* nelem,
                                 int* nblk, int* qin, int* qout, int* rows, int* cols);

/* ---- transfer matrices (A7/A8) ------------------------------------------------------------
 * Replaces the preallocated ApplyTMmixedLeft / ApplyTMmixedRight (TMBlockFunctions.hpp:352-390,
 * :494-533) and, with B == NULL, the regular ApplyTMLeft / ApplyTMRight on BlockDiagMat and
 * BlockMat inputs (:60-84, :97-125, :184-205, :220-250) as called by TMMultDiag / TMMult
 * (EigsBlockFunctions.hpp:50-85):
 *   left : y[Qout B_s(qi)] += B_s(qi)^T * x(qi->qo) * A_s(qo)
 *   right: y[Qin A_s(qi-dA_s)] += A_s(qi-dA_s) * x(qi->qo) * B_s(qo-dB_s)^T
 * x and y are BlockMat/BlockDiagMat layouts; y must hold every sector that receives a
 * contribution it should keep (others are dropped, as in the reference). */
imps_status imps_tm_plan(imps_context ctx, imps_plan* plan, int dir, const imps_layout* A,
                         const imps_layout* B, const imps_layout* x, const imps_layout* y, int prec);
/* Copies (and converts to the plan precision) the MPS data; B may be NULL for the regular TM. */
imps_status imps_tm_set_mps(imps_plan plan, const double* A, const double* B);
imps_status imps_tm_apply(imps_plan plan, const double* x, double* y, int nvec);

/* ---- two-site effective Hamiltonians (A12/A13) -------------------------------------------
 * H_AC: ApplyHAvec (VUMPS_helpers.hpp:55-88) / ApplyHA (:38-53):
 *   out[i] = EHL*in[i] + in[i]*EHR + sum_j (HL[i][j]*in[j] + in[j]*HR[i][j])
 * AC: MPS layout (d lists); EHL, EHR: BlockDiagMat; HL, HR: RedOp layouts (d*d lists). */
imps_status imps_hac_plan(imps_context ctx, imps_plan* plan, const imps_layout* AC,
                          const imps_layout* EHL, const imps_layout* EHR,
                          const imps_layout* HL, const imps_layout* HR, int prec);
imps_status imps_hac_set_operator(imps_plan plan, const double* EHL, const double* EHR,
                                  const double* HL, const double* HR);
imps_status imps_hac_apply(imps_plan plan, const double* x, double* y, int nvec);
/* The same apply without any allocation.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  imps_hac_apply keeps the split copies of its input (the compute-format copy
 * of x, stored as is and transposed) in buffers the plan allocates on first use; this entry point
 * keeps them in a fixed 64 MiB scratch array compiled into the library, one per GPU (held until
 * the process ends), and calls no hipMalloc, hipFree, event or autotuning routine (the autotuned
 * schedule is used if imps_hac_apply already chose one for nvec, else the default rule).  It
 * fails with IMPS_INVALID_ARG when nvec split copies exceed the scratch: one vector's split copy takes
 * P = 64 * ceil((len + 64) / 64) elements per plane, two planes in FP64 (16 P bytes) and FP32 (8 P
 * bytes), four bf16 planes in BF16X3 (8 P bytes), len the H_AC vector length.  Every call on one GPU shares that GPU's scratch: calls that
 * overlap on one GPU (several streams or host threads) race.  The scratch is cleared by a fill
 * kernel at every call (its pads must read as zeros). */
imps_status imps_hac_apply_static(imps_plan plan, const double* x, double* y, int nvec);

/* H_C: ApplyHCvec (VUMPS_helpers.hpp:113-133) / ApplyHC (:91-111).  The reference evaluates
 * sum_{nnz H} h*AL[i0]^T*AL[j0]*C*AR[j1]*AR[i1]^T; this plan executes the algebraically identical
 * reduced form  EHL*C + C*EHR + sum_{i,j} HL[i][j]*(C*AR[j])*AR[i]^T  with HL = ReducedOp(H,AL,l)
 * (MPSBlockUtilities.hpp:39-74), keeping diagonal blocks only (BlockObj.hpp:1527-1536). */
imps_status imps_hc_plan(imps_context ctx, imps_plan* plan, const imps_layout* C,
                         const imps_layout* EHL, const imps_layout* EHR,
                         const imps_layout* HL, const imps_layout* AR, int prec);
imps_status imps_hc_set_operator(imps_plan plan, const double* EHL, const double* EHR,
                                 const double* HL, const double* AR);
imps_status imps_hc_apply(imps_plan plan, const double* x, double* y, int nvec);

/* ---- adjoints (vector-Jacobian products) ------------------------------------------------------
 * No counterpart in the reference (it has no derivatives); for gradient-based callers of the
 * operator applications above.  Given the plan of a forward apply y = f(x; operator data) and the
 * cotangent ybar (layout of y), each entry writes xbar = (df/dx)^T ybar and the gradient of
 * <ybar, y> with respect to every operator part, restricted onto that part's sector structure.
 * Arithmetic is FP64 whatever the plan's precision; the forward's intermediates are recomputed,
 * not stored.  Operator data are passed as in the plan's set call; any output pointer may be NULL
 * (not written); outputs are fully overwritten.  Formulas:
 *   TM left  (y = sum_s B_s^T x A_s): xbar = sum_s B_s ybar A_s^T, Abar_s = x^T B_s ybar, Bbar_s = x A_s ybar^T
 *   TM right (y = sum_s A_s x B_s^T): xbar = sum_s A_s^T ybar B_s, Abar_s = ybar B_s x^T, Bbar_s = ybar^T A_s x
 *   (regular TM: B and Bbar NULL, Abar the sum of both roles)
 *   H_AC: xbar[j] = EHL^T ybar[j] + ybar[j] EHR^T + sum_i HL[i][j]^T ybar[i] + ybar[i] HR[i][j]^T,
 *         EHLbar = sum_i ybar[i] x[i]^T, EHRbar = sum_i x[i]^T ybar[i], HLbar[i][j] = ybar[i] x[j]^T, HRbar[i][j] = x[j]^T ybar[i]
 *   H_C (reduced form): Cbar = EHL^T ybar + ybar EHR^T + sum_ij HL[i][j]^T ybar AR[i] AR[j]^T, EHLbar = ybar C^T,
 *         EHRbar = C^T ybar, HLbar[i][j] = ybar AR[i] AR[j]^T C^T, ARbar[j] += C^T HL[i][j]^T ybar AR[i],
 *         ARbar[i] += ybar^T HL[i][j] C AR[j] */
imps_status imps_tm_vjp(imps_plan plan, const double* A, const double* B, const double* x, const double* ybar, double* xbar,
                        double* Abar, double* Bbar);
imps_status imps_hac_vjp(imps_plan plan, const double* EHL, const double* EHR, const double* HL, const double* HR, const double* x,
                         const double* ybar, double* xbar, double* EHLbar, double* EHRbar, double* HLbar, double* HRbar);
imps_status imps_hc_vjp(imps_plan plan, const double* EHL, const double* EHR, const double* HL, const double* AR, const double* x,
                        const double* ybar, double* xbar, double* EHLbar, double* EHRbar, double* HLbar, double* ARbar);

/* ---- reduced operators (A11) ---------------------------------------------------------------
 * ReducedOp(H, A, dir) (MPSBlockUtilities.hpp:39-74); H is a dense d^2 x d^2 column-major host
 * array whose nonzeros define the sparse two-site operator.  The output layout (d*d lists) is
 * derived from A's structure; query it with imps_plan_out_layout.  apply: y = ReducedOp. */
imps_status imps_redop_plan(imps_context ctx, imps_plan* plan, const double* H, int d,
                            const imps_layout* A, int dir, int prec);
imps_status imps_redop_apply(imps_plan plan, const double* A, double* y);

/* ---- environments (A14/A15) ---------------------------------------------------------------
 * InvertE_proj (InvEBlockFunctions.hpp:205-266) with the geometric-sum map of InvEfun_proj (:13-47)
 * over a unit cell of ncell MPS tensors (chained TM, TMBlockFunctions.hpp:565-615):
 *   dir l: (y|[1 - T + |R)(L|] = (x|[1 - |R)(L|],   T(y) = T_L[A_{N-1}](... T_L[A_0](y))
 *   dir r: [1 - T + |R)(L|]|y) = [1 - |R)(L|]|x),   T(y) = T_R[A_0](... T_R[A_{N-1}](y))
 * x, L, R, y: BlockDiagMat vectors on the layout x.  Solved by restarted GMRES(krylov) with FP64
 * CGS2 orthogonalisation; with prec != IMPS_FP64 the TM applies of the inner solves run in that
 * precision inside FP64 residual refinement.  Stops at ||b - A y|| < tol ||b|| (the rule of gmres,
 * IterativeSolvers.hpp:119).  y0 (optional) is the initial guess.  Returns IMPS_NOT_CONVERGED
 * (with y, iters and relres written) when maxit operator applications do not suffice. */
imps_status imps_env_plan(imps_context ctx, imps_plan* plan, int dir, int ncell, const imps_layout* A,
                          const imps_layout* x, int prec);
imps_status imps_env_set_mps(imps_plan plan, const double* const* A);
imps_status imps_env_solve(imps_plan plan, const double* x, const double* L, const double* R,
                           const double* y0, double* y, double tol, int maxit, int krylov,
                           int* iters, double* relres);
/* y = T(x) with the plan's unit-cell chain (ApplyTMLeft/Right on an MPSBlockMatArray,
 * TMBlockFunctions.hpp:555-615), FP64. */
imps_status imps_env_tm_apply(imps_plan plan, const double* x, double* y);
/* Dominant fixed point of the unit-cell TM of isometric tensors (TMDominantEig,
 * EigsBlockFunctions.hpp:91-172, trace-normalised as at :162), from the start x: the TM preserves
 * the trace, so with x0 of unit trace the fixed point is x0 + dx, [1 - T + |x0)(1|] dx = T(x0) - x0,
 * solved by the environment GMRES; until ||T(x) - x|| < tol (IMPS_NOT_CONVERGED otherwise). */
imps_status imps_env_fixed_point(imps_plan plan, double* x, double tol, int maxit, int* iters, double* delta);

/* ---- lowest eigenpair (A17) ----------------------------------------------------------------
 * Replaces eigs_rs(MultOPx, N, vals, vecs, 1, "SA", tol, x0) (eigs.h:43-44; eigs.cpp:369-494) for
 * an H_AC or H_C plan: restarted Lanczos with full FP64 reorthogonalisation.  x: start vector in,
 * normalised eigenvector out; converged when ||H x - e x|| <= tol |e| (ARPACK's criterion).  With a
 * low-precision plan and an FP64 plan `plan_fp64` of the same operator, the Lanczos runs first in
 * low precision and is then refined with FP64 applies from its Ritz vector. */
imps_status imps_eigsh(imps_plan plan, imps_plan plan_fp64, double* x, double tol, int maxit,
                       int krylov, double* eval, int* iters, double* resid);

/* ---- block products (A6) ------------------------------------------------------------------
 * Two-operand products of block objects with the reference's matching rules (BlockMat operator*,
 * BlockObj.hpp:2557-2625; MPS products, MPSBlockMat.hpp:934-1043; unity mixed TMs,
 * TMBlockFunctions.hpp:310-333 and :444-470).  kind:
 *   0 y[s] = M A[s]   1 y[s] = A[s] M   2 y[s] = A[s] M^T   3 y[s] = M^T A[s]
 *   4 y = sum_s B_s^T A_s  (X1 = A, X2 = B)   5 y = sum_s A_s B_s^T  (X1 = A, X2 = B)
 *   6 y[i*d + j] = A[i] B[j] (two-site tensor; X1 = A, X2 = B)
 *   7 y[i] = sum_j R[i][j] A[j]   8 y[i] = sum_j A[j] R[i][j]   (X1 = R, a RedOp; X2 = A)
 * For kinds 0/3, X1 is M; for 1/2, X2 is M.  FP64.  The output layout is derived (query with
 * imps_plan_out_layout). */
imps_status imps_prod_plan(imps_context ctx, imps_plan* plan, int kind, const imps_layout* X1,
                           const imps_layout* X2);
imps_status imps_prod_apply(imps_plan plan, const double* x1, const double* x2, double* y);
/* The same product for nvec first operands stored back to back (x1[v*len_x1 + i], outputs
 * y[v*len_y + i]) with one second operand x2: the real and imaginary parts of a complex first
 * operand in one launch, as the reference's complex BlockMat products need them
 * (Excitations_helpers.hpp:20-228 with VTX = Complex). */
imps_status imps_prod_apply_n(imps_plan plan, const double* x1, const double* x2, double* y, int nvec);

/* ---- per-sector QR (A18) -------------------------------------------------------------------
 * qr(R, MPS, dir) (MPSBlockUtilities.hpp:181-244) with qr_pos (helpers.hpp:181-195): the dense
 * regrouped sector matrices (GetDenseMat, MPSBlockMat.hpp:468-526) are factorised in FP64 on the
 * GPU, one workgroup per sector; Q has the layout of A, R is a BlockDiagMat (out layout).  Every
 * sector must have at least as many stacked rows (l) / columns (r) as its size. */
imps_status imps_qr_plan(imps_context ctx, imps_plan* plan, const imps_layout* A, int dir);
imps_status imps_qr_apply(imps_plan plan, const double* A, double* Q, double* R);

/* ---- quasiparticle excitations (A21, A14 complex, A17 complex) ----------------------------
 * The momentum-k effective Hamiltonian of the excitation ansatz B[n] = NL[n] X[n] for a ground
 * state in mixed gauge, for nmom momenta at once.
 *
 * imps_heff_plan: the structure.  AL, AR, NL: ncell MPS layouts each (nlist = d; NL[n] the left
 * null space of AL[n], Nullspace(AL[n], l), MPSBlockUtilities.hpp:247-340); C: ncell BlockDiagMat
 * layouts, C[n] on the bond right of site n (C[ncell-1] on bond 0); H: dense d^2 x d^2 column-major
 * two-site Hamiltonian; E0: the ground-state energy per site, subtracted from H as
 * Excitations_main.cpp:191 does; K: excitation charge (nsym ints).  The parameter space X is XDims
 * (Excitations_helpers.hpp:639-678): query it with imps_plan_out_layout (nlist = ncell, entry n =
 * the blocks of X[n]; mtot = nelem).  Fails with IMPS_INVALID_ARG when XDims is empty ("no
 * excitations for this quantum number", Excitations_main.cpp:198).
 * imps_heff_set_state: copies the tensors (device FP64, Vectorize order) and computes the constants
 * of HeffConstants (Excitations_helpers.hpp:260-347: HLtot, HRtot through two environment solves
 * at tolerance inv_tol), the reduced operators and the dominant eigenmatrices of the mixed transfer
 * matrix in closed form, LM = C^T and RM = C on bond 0 (LRoverlaps, :915-946, for AL C = C AR).
 * imps_heff_check: the reference's check of HeffConstants (:287-303), tr(HL R) and tr(L HR), which
 * vanish when E0 is the state's energy; and the environment solver's operator applications.
 * prec: IMPS_FP64 (the reference's arithmetic), or IMPS_FP32 / IMPS_BF16X3: the Arnoldi steps of the
 * geometric sums apply their transfer-matrix chains in that precision and every GMRES restart
 * takes the FP64 residual (iterative refinement), so the sums still reach inv_tol in FP64; the
 * constants, the eigensolver and everything else stay FP64. */
imps_status imps_heff_plan(imps_context ctx, imps_plan* plan, int ncell, int d, const imps_layout* AL, const imps_layout* AR,
                           const imps_layout* NL, const imps_layout* C, const double* H, double E0, const int* K, int nmom,
                           int prec);
/* Replaces OPeffConstants + ApplyOPeff (Excitations_helpers.hpp:350-600), the excitation observables
 * of MeasureExcitations (:774-826): a plan of the same kind as imps_heff_plan's for the effective
 * operator of a one-site observable O (d x d column-major, charge-conserving: O(r, c) != 0 only
 * between physical indices whose blocks connect the same charges) with O0 (its ground-state
 * expectation, MeasureExcitations' eobs) subtracted.  imps_heff_set_state computes OPLtot and OPRtot
 * through two environment solves, imps_heff_check returns the reference's checks tr(OPL R) and
 * tr(L OPR), imps_heff_apply applies O_eff(p_t) to the nmom momenta at once; <X_n | (O_eff X)_n>
 * per site n is MeasureExcitations' value.  The reference skips the left geometric sum when its
 * source is below 10 InvETol in norm; this plan always solves it. */
imps_status imps_opeff_plan(imps_context ctx, imps_plan* plan, int ncell, int d, const imps_layout* AL, const imps_layout* AR,
                            const imps_layout* NL, const imps_layout* C, const double* O, double O0, const int* K, int nmom, int prec);
imps_status imps_heff_set_state(imps_plan plan, const double* const* AL, const double* const* AR, const double* const* NL,
                                const double* const* C, double inv_tol);
/* imps_heff_set_state for an AR that is not C^-1 AL C: the state of ApplyTransformation
 * (Excitations_helpers.hpp:680-771, spin or charge flips of AR and L) or of a relative unit-cell
 * shift (Excitations_main.cpp:153-163).  C still gives R = C C^T on AL's side
 * (Excitations_main.cpp:142-146); L (BlockDiagMat, the left fixed point of AR's transfer matrix on
 * AR's bond 0, Lvec.back() after the transformation, Excitations_main.cpp:173), and the dominant
 * left / right eigenmatrices LM (rows on AR's bond, columns on AL's) and RM (rows on AL's bond,
 * columns on AR's) of the mixed transfer matrix in K = 0 are given with their layouts (nlist = 1,
 * device FP64, Vectorize order; LRoverlaps, :915-946, normalised to tr(LM RM) = 1).  Blocks are
 * matched by charge; a block outside the bond structure of AL and AR is an error.  LM and RM only
 * project the sums in K = 0 (Excitations_main.cpp:264-268): for another K they are not read. */
imps_status imps_heff_set_state_lr(imps_plan plan, const double* const* AL, const double* const* AR, const double* const* NL,
                                   const double* const* C, const imps_layout* Llay, const double* L, const imps_layout* LMlay,
                                   const double* LM, const imps_layout* RMlay, const double* RM, double inv_tol);
imps_status imps_heff_check(imps_plan plan, double* trHL_R, double* trL_HR, int* env_iters);
/* Replaces the raw callback ApplyHeff(in, out, xdims, mtot, kfac, AL, AR, NL, LM, RM, H, HLtot, HRtot,
 * InvETol, ...) (Excitations_helpers.hpp:230-257) for nmom momenta at once: y_t = H_eff(p_t) x_t.
 * p[t]: momentum in units of pi per site; the unit-cell phase is kfac = e^{i p pi ncell}
 * (Excitations_main.cpp:258-259).  The two geometric sums (InvertE_fac, and InvertE_proj_fac when
 * K = 0 and |p pi ncell| < 10 inv_tol, Excitations_main.cpp:264-268) are solved for all momenta
 * by one batched complex GMRES each, to ||r|| < inv_tol ||b|| (IterativeSolvers.hpp:119).
 * x, y: device arrays of nmom * mtot interleaved complex doubles (std::complex<double>, the
 * reference's VTX = Complex), momentum-major; y is fully overwritten.  gmres_steps (optional):
 * batched GMRES operator applications of the call; gmres_relres: the worst final relative residual. */
imps_status imps_heff_apply(imps_plan plan, const double* p, const double* x, double* y, int* gmres_steps, double* gmres_relres);
/* Replaces eigs_n(Hfun, mtot, vals, vecs, nev, "SR", tol) (eigs.h:54-55, eigs.cpp:6-168) as called
 * per momentum by Excitations_main.cpp:270-279, for all nmom momenta at once.  H_eff(k) is
 * Hermitian: a thick-restart Lanczos with full CGS2 re-orthogonalisation (krylov basis vectors,
 * default max(20, 2 nev + 1), the reference's ncv), converged when every wanted Ritz pair has
 * ||H y - theta y|| <= tol max(1, |theta|).  evals: [nmom][nev] ascending; evecs (optional): device
 * [nmom][nev][mtot] interleaved complex, unit norm; resid (optional, host): [nmom][nev] residual
 * norms; napply: batched H_eff applications.  IMPS_NOT_CONVERGED when maxit applications do not
 * suffice (results written). */
imps_status imps_heff_eigs(imps_plan plan, const double* p, int nev, double tol, int maxit, int krylov, double* evals, double* evecs,
                           double* resid, int* napply);

/* ---- bookkeeping (A1/A2) -------------------------------------------------------------------
 * model (CreateModel's types, Models.cpp:881-1114): IMPS_MODEL_XXZ spin-1/2 (one U(1));
 * IMPS_MODEL_FHUB Fermi-Hubbard in (N,M) representation (U(1)xU(1)); IMPS_MODEL_QIM quantum Ising
 * (Z(2)); IMPS_MODEL_QPOTTS(q) q-state quantum Potts (Z(q)); IMPS_MODEL_BHUB(d) Bose-Hubbard with
 * local dimension d (U(1)).  An id is type | (d << 8); 2 <= d <= 16 for QPOTTS and BHUB.
 * imps_model_nsym / imps_model_d: charges per key and local dimension, -1 for an unknown id.
 * imps_model_periods: the symmetry group (GroupObj_.init of each model's Init): writes nsym
 * periods (0 = U(1), N = Z(N)) for imps_layout.periods; returns nsym or -1.
 * imps_model_i2k: GetI2K (Models.cpp:262-295, :372-376, :500-506, :699-733, :862-878); writes
 * d*nsym charges, returns d or -1.
 * imps_model_validpath: validpath (ItoKey.hpp:126-165); writes N indices.
 * imps_model_make_dims: MakeDims (Models.cpp:13-32) from an explicit bond-0 dim map
 * (keys0 ascending); writes N bonds of nsec (keys, sizes). Returns nsec or -1.
 * imps_model_dimmap: GetDimMap (Models.cpp:248-260, :357-370, :478-498, :680-697, :845-860) from
 * the per-sector dims list.  QN: XXZ {2 Sz per N sites}, FHUB {n, m}, BHUB {n}, QIM / QPOTTS unused.
 * imps_model_ham: local two-site Hamiltonian (Models.cpp:57-93, :411-430, :543-572, :768-777),
 * dense d^2 x d^2 column-major; params XXZ {J, Delta, hz}, FHUB {t, U, V, mu, hz}, QIM {J, h},
 * QPOTTS {h}, BHUB {t, U, mu}. */
#define IMPS_MODEL_XXZ 0
#define IMPS_MODEL_FHUB 1
#define IMPS_MODEL_QIM 2
#define IMPS_MODEL_QPOTTS(q) (3 | ((q) << 8))
#define IMPS_MODEL_BHUB(d) (4 | ((d) << 8))
int imps_model_nsym(int model);
int imps_model_d(int model);
int imps_model_periods(int model, int* periods);
int imps_model_i2k(int model, int N, const int* QN, int* i2k);
int imps_model_validpath(int model, int N, const int* QN, int* path);
int imps_model_dimmap(int model, int N, const int* QN, const int* dims, int ndims, int* keys, int* sizes);
int imps_model_make_dims(int model, int N, const int* QN, int nsec, const int* keys0, const int* sizes0,
                         int* keys, int* sizes);
int imps_model_ham(int model, const double* params, double* H);
/* MPSBlockMat(I2K, ml, mr) structure (MPSBlockMat.hpp:221-236): blocks q -> q + I2K[s] for every
 * left sector q whose shifted charge exists on the right.  Writes nblk[d] and the block arrays
 * (capacity d*nl blocks); returns the total block count or -1. */
int imps_mps_structure(int model, int N, const int* QN, int nl, const int* kl, const int* sl,
                       int nr, const int* kr, const int* sr, int* nblk, int* qin, int* qout,
                       int* rows, int* cols);

#ifdef __cplusplus
}
#endif
#endif /* IMPS_CAPI_H */
