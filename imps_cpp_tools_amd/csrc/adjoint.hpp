// Adjoints (vector-Jacobian products) of the transfer-matrix, H_AC and H_C applications, compiled
// per plan into a block tape and run in FP64 whatever the forward precision.  The forward's
// intermediates (x A_s of a TM, C AR_j of H_C) are recomputed from the inputs, not stored.
//
//   left TM   y = sum_s B_s^T x A_s          xbar = sum_s B_s ybar A_s^T,  Abar_s = x^T B_s ybar,  Bbar_s = x A_s ybar^T
//   right TM  y = sum_s A_s x B_s^T          xbar = sum_s A_s^T ybar B_s,  Abar_s = ybar B_s x^T,  Bbar_s = ybar^T A_s x
//   H_AC      y[i] = EHL x[i] + x[i] EHR + sum_j HL[i][j] x[j] + x[j] HR[i][j]
//             xbar[j] = EHL^T ybar[j] + ybar[j] EHR^T + sum_i HL[i][j]^T ybar[i] + ybar[i] HR[i][j]^T
//             EHLbar = sum_i ybar[i] x[i]^T, EHRbar = sum_i x[i]^T ybar[i], HLbar[i][j] = ybar[i] x[j]^T, HRbar[i][j] = x[j]^T ybar[i]
//   H_C       y = EHL C + C EHR + sum_ij HL[i][j] C AR[j] AR[i]^T (diagonal blocks)
//             Cbar = EHL^T ybar + ybar EHR^T + sum_ij HL[i][j]^T ybar AR[i] AR[j]^T
//             EHLbar = ybar C^T, EHRbar = C^T ybar, HLbar[i][j] = ybar AR[i] AR[j]^T C^T,
//             ARbar[j] += C^T HL[i][j]^T ybar AR[i],  ARbar[i] += ybar^T HL[i][j] C AR[j]
// For the regular TM (B = A) Abar is the sum of both roles.  Every gradient is restricted onto the
// sector structure of its parameter.
#pragma once

#include "heff.hpp"

namespace imps {

struct VjpTape : BlockTape {
    std::vector<XObj> ins;   // operator parts, then x, then ybar
    std::vector<XObj> outs;  // xbar, then the operator gradients
    std::vector<Op> tape;
};

// Compiles (once) and runs the adjoint of plan P: ops are the operator parts in the order of the
// plan's set call; outs (xbar, gradient of each part; NULL skips) are fully overwritten.
void vjp_run(imps_plan_s* P, const std::vector<const double*>& ops, const double* x, const double* ybar, const std::vector<double*>& outs);

}  // namespace imps
