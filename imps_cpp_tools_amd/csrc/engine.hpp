// Device execution of grouped block-GEMM programs (see layout.hpp for the program model).
#pragma once

#include <hip/hip_runtime.h>

#include "layout.hpp"
#include "panel.hpp"

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

// Panel programs (panel.hpp): packing of the operator store and one grouped launch per stage.
struct DevPStage {
    const PTask* tasks = nullptr;
    const PChunk* chunks = nullptr;
    const RawPack* packs = nullptr;
    int ntasks = 0;
    int out_slot = SLOT_OUT;
};
hipError_t launch_panelize(int prec, const PanelDesc* d_panels, int npanels, const double* op, void* store, hipStream_t stream);
// Split copies (panel.hpp) of the input vectors' raw blocks: grid (nunits, nvec).  write_n = false
// (FP64 only) writes plane T alone: the kernels then read plane N from the input itself.
hipError_t launch_split(int prec, const SplitUnit* d_units, int nunits, const SplitSpan* d_spans, const double* x, long x_stride,
                        void* split, long split_stride, long pl, int nvec, hipStream_t stream, bool write_n = true);
// Buffers of one panel-program apply.  Strides and sizes of split copies are bytes per vector.
struct PanelLaunch {
    const void* pstore = nullptr;
    long pstore_bytes = 0;
    const void* xs = nullptr;
    long xs_stride = 0;
    const double* xd = nullptr;  // FP64 input read directly for plane-N chunks (nullptr: from the split copy)
    long xd_stride = 0;          // elements per vector
    bool xt_direct = false;      // with xd: plane-T chunks read the input directly too (then no split copy is made)
    void* ws = nullptr;
    long ws_stride = 0;
    long plx = 0, plw = 0;
    double* out = nullptr;
    long out_stride = 0;
    int nvec = 0;
};
// ksplit (1, 2, 4): waves sharing one vector's tile; the workgroup serves 4 / ksplit vectors.
// ksplit in {1, 2, 4}: waves sharing one vector's tile; shared (ksplit 1 only): the shared-panel kernel
hipError_t launch_panel_stage(int prec, const DevPStage& st, int ksplit, bool shared, const PanelLaunch& L, hipStream_t stream);
// the schedule used without the autotuner: shared-panel kernel for K-split 1 in BF16X3
bool panel_shared_default(int prec);
hipError_t launch_zero(double* p, long n, hipStream_t stream);

}  // namespace imps
