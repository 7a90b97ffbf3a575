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
size_t blas_scratch_doubles(int nb);

}  // namespace imps
