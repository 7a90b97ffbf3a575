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
// y = a*x + b*y, with a replaced by *da (or 1 / *da when inv_a) when da is given
hipError_t blas_axpby(double a, const double* da, bool inv_a, const double* x, double b, double* y, long n, hipStream_t s);
size_t blas_scratch_doubles(int nb);

}  // namespace imps
