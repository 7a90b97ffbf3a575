// Device execution of grouped block-GEMM programs (see layout.hpp for the program model).
#pragma once

#include <hip/hip_runtime.h>

#include "layout.hpp"

namespace imps {

// Operand slots as seen by a kernel: base pointer and per-vector stride (elements) for the input
// vector (always FP64), the operator store and the intermediate workspace (compute type).
struct Slots {
    const void* base[3];
    long stride[3];
};

struct DevStage {
    const Task* tasks = nullptr;
    const Seg* segs = nullptr;
    int ntasks = 0;
    int out_slot = SLOT_OUT;
};

// One grouped launch: grid (ntasks, nvec), four waves per workgroup.  Compute type is double for IMPS_FP64 and float
// otherwise; the output is double when `out_double`, else the compute type.
hipError_t launch_stage(int prec, const DevStage& st, const Slots& slots, void* out, long out_stride,
                        bool out_double, int nvec, hipStream_t stream);

hipError_t launch_d2f(const double* in, float* out, long n, hipStream_t stream);

}  // namespace imps
