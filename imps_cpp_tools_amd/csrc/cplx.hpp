// Device kernels of the excitation engine: batched complex Krylov vector algebra and the
// block-level gather-scale-sum of block objects.
//
// A batched complex vector of length n over nmom momenta is planar: momentum t keeps its real parts
// at base + 2 t n and its imaginary parts at base + (2 t + 1) n.  A basis of such vectors stores
// vector i at V + i * ldv.  All reductions sum chunk partials in a fixed order (deterministic).
#pragma once

#include <hip/hip_runtime.h>

namespace imps {

int cblas_chunks(long n);     // chunk count of the batched inner products over vectors of length n
int cblas_gemv_grid(long n, int nmom);  // workgroups per momentum (= norm partials) of cgemv_sub_partials

// part[((t * nb + i) * nch + c) * 2 + {0, 1}] = chunk c of <V_i(t), w(t)> = sum conj(V_i) w
hipError_t cdot_partials(const double* V, long ldv, int nb, const double* w, long n, int nmom, double* part, hipStream_t s);
// w(t) -= sum_i h_i(t) V_i(t), h_i(t) the summed partials; h stored at hout[t * hstride + 2 i + {0, 1}]
// when hout is given; with qpart, qpart[t * gch + g] = partials of ||w(t)||^2 after the update
hipError_t cgemv_sub_partials(const double* V, long ldv, int nb, const double* part, int nch, double* hout, long hstride, double* w,
                              long n, int nmom, double* qpart, hipStream_t s);
// w(t) /= sqrt(Q_t), Q_t the summed gch partials (stored at qout[t * qstride]); w(t) = 0 when Q_t <= 0
hipError_t cscale_inv_norm(const double* qpart, int gch, double* qout, long qstride, double* w, long n, int nmom, hipStream_t s);
// x(t) += sum_{i < nb} c_i(t) V_i(t), c_i(t) = coef[t * cstride + 2 i + {0, 1}]
hipError_t cgemv_add(const double* V, long ldv, int nb, const double* coef, long cstride, double* x, long n, int nmom, hipStream_t s);
// Y_i(t) = sum_{j < nb} S_ij(t) V_j(t) for i < nk, S_ij(t) = S[((t * nk + i) * nb + j) * 2 + {0, 1}]; Y_i at Y + i * ldy
hipError_t cgemm_basis(const double* V, long ldv, int nb, const double* S, int nk, double* Y, long ldy, long n, int nmom, hipStream_t s);
// part[(t * nch + c) * 2 + {0, 1}] = chunk c of sum_e wv[e] z(t)[e] (wv real)
hipError_t rdot_partials(const double* wv, const double* z, long n, int nmom, double* part, hipStream_t s);
// out(t) = a z(t) - fac_t Tz(t) + gam_t tr_t rout, fac_t = fac[2 t + {0, 1}] (complex), gam_t = gam[t],
// tr_t = the summed partials tpart[(t * nch + c) * 2 + {0, 1}]; Tz NULL: no T term; tpart NULL: no trace term
hipError_t geo_combine(double* out, double a, const double* z, const double* Tz, const double* fac, const double* gam, const double* tpart,
                       int nch, const double* rout, long n, int nmom, hipStream_t s);
// interleaved complex [nmom][n] <-> planar batched vectors
hipError_t cinterleave(const double* planar, double* inter, long n, int nmom, hipStream_t s);
hipError_t cdeinterleave(const double* inter, double* planar, long n, int nmom, hipStream_t s);

// Gather-scale-sum of whole blocks into the blocks of a new object, all inside one arena of doubles.
// Output piece k covers out_off .. out_off + len (plane 0) and sums the terms t0 .. t0 + nt in order:
//   out = sum_terms coef(t) * src,  src real (one plane) or complex (planes src_ps apart).
struct CombTile {
    long out_off;
    int len, t0, nt, pad;
};
struct CombTerm {
    long src_off;
    long src_ps;  // plane stride of a complex source; 0 for a real one
    int coef;     // row of the coefficient table (nmom complex values per row)
    int pad;
};
// grid (ntiles, out_cplx ? nmom : 1).  Real output: only plane 0, coefficients' real parts of
// momentum 0.  Complex output: planes 2 t and 2 t + 1, out_ps apart.
hipError_t comb_launch(const CombTile* tiles, int ntiles, const CombTerm* terms, const double* coef, double* arena, long out_ps, int out_cplx,
                       int nmom, hipStream_t s);

}  // namespace imps
