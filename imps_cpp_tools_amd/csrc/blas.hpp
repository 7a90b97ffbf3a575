// FP64 device vector kernels used by the Krylov drivers.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>

namespace imps {

// out[i] = <V_i, w> for the nb vectors V_i = V + i*ldv of length n; scratch holds
// blas_scratch_doubles(nb) doubles.  Deterministic (fixed two-stage reduction).
hipError_t blas_multidot(const double* V, long ldv, int nb, const double* w, long n, double* out, double* scratch, hipStream_t s);
// w -= sum_i h[i] V_i  (h on the device)
hipError_t blas_gemv_sub(const double* V, long ldv, int nb, const double* h, double* w, long n, hipStream_t s);
// How blas_axpby reads a device coefficient *da: as is, its reciprocal, or the reciprocal of its
// square root (a squared norm; 0 when *da <= 0, so a zero vector stays zero).
enum AxpbyScalar { AXPBY_DA = 0, AXPBY_INV_DA = 1, AXPBY_INV_SQRT_DA = 2 };
// y = a*x + b*y, with a replaced by the device coefficient *da (read as `mode` says) when da is given
hipError_t blas_axpby(double a, const double* da, int mode, const double* x, double b, double* y, long n, hipStream_t s);
// scratch doubles for blas_multidot with nb vectors, or for blas_multidot_partials with nb vectors
// followed by the blas_gemv_grid(n) norm partials of blas_gemv_sub_partials
size_t blas_scratch_doubles(int nb);

// The Krylov step's orthogonalisation without separate reduction launches: the chunk partials of
// the inner products are summed by the kernel that consumes them.
int blas_chunks(long n);     // chunk count of the inner products over vectors of length n
int blas_gemv_grid(long n);  // workgroups (= norm partials) of blas_gemv_sub_partials
// partials[c*nb + i] = chunk c of <V_i, w>, c < blas_chunks(n)
hipError_t blas_multidot_partials(const double* V, long ldv, int nb, const double* w, long n, double* partials, hipStream_t s);
// w -= sum_i h[i] V_i with h[i] = the summed partials (also stored at hout); with qpart, the
// blas_gemv_grid(n) partials of ||w||^2 after the update
hipError_t blas_gemv_sub_partials(const double* V, long ldv, int nb, const double* partials, int nch, double* hout, double* w, long n,
                                  double* qpart, hipStream_t s);
// out = x - y + t z, t = the summed partials of one inner product (blas_multidot_partials, nb = 1)
hipError_t blas_diff_plus_scaled(const double* x, const double* y, const double* z, const double* partials, int nch, double* out, long n,
                                 hipStream_t s);
// y /= sqrt(Q), Q = the summed nq partials (stored at qout; y = 0 when Q <= 0)
hipError_t blas_scale_inv_norm(const double* qpart, int nq, double* qout, double* y, long n, hipStream_t s);

}  // namespace imps
