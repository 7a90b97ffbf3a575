// Plans: a compiled grouped-GEMM Program, its device copy and the operator store.
#pragma once

#include <atomic>
#include <map>
#include <memory>
#include <vector>

#include "../../include/imps_capi.h"
#include "common.hpp"
#include "engine.hpp"
#include "qr.hpp"

struct imps_context_s {
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    std::atomic<int> refs{1};  // the caller's handle + one per live plan; freed at zero (plans may be made and dropped on several threads)
};

namespace imps {

// Drops one reference to a context; the last one synchronises and frees it.
void ctx_release(imps_context ctx);

// Reference held by a plan, so that a context outlives every plan created on it whatever the
// order in which a caller destroys them.  Declared as the first member of imps_plan_s: it is
// destroyed last, after the plan's device buffers.
struct CtxHold {
    imps_context c = nullptr;
    CtxHold() = default;
    CtxHold(const CtxHold&) = delete;
    CtxHold& operator=(const CtxHold&) = delete;
    void hold(imps_context ctx) { c = ctx; if (c) c->refs.fetch_add(1); }
    ~CtxHold() { if (c) ctx_release(c); }
};

}  // namespace imps

namespace imps {

enum Kind { K_TM = 0, K_HAC = 1, K_HC = 2, K_REDOP = 3, K_ENV = 4, K_PROD = 5, K_QR = 6, K_HEFF = 7 };

struct EnvSolver;   // krylov.cpp
struct HeffEngine;  // heff.cpp
struct VjpTape;     // adjoint.cpp

}  // namespace imps

struct imps_plan_s {
    imps::CtxHold hold;
    imps_context ctx = nullptr;
    int kind = 0;
    int prec = IMPS_FP64;
    imps::Program prog;
    imps::Layout out_layout;      // derived output structure (redop)
    std::vector<long> op_parts;   // element counts of the operator-store pieces, in set order
    bool op_set = false;
    imps::DevBuf d_prog, d_op, d_w;
    std::vector<imps::DevStage> dstages;
    // panel execution (TM, H_AC, H_C): packed operator store, FP64 workspace
    bool use_panels = false;
    imps::PanelProgram pprog;
    imps::DevBuf d_pprog, d_pstore, d_xs, d_ws;  // program tables, panels, split copies of input and workspace
    const imps::PanelDesc* d_panels = nullptr;
    const imps::SplitSpan* d_xspans = nullptr;
    const imps::SplitUnit* d_xunits = nullptr;
    std::vector<imps::DevPStage> dpstages;
    long xs_ready = 0, ws_ready = 0;  // vectors of the split buffers already zero-filled
    std::map<int, int> sched;         // autotuned panel schedule per vector count (see panel_schedule)
    // opt-in GPU timing of the GEMM launches of each apply (events recorded on the plan's stream
    // around the grouped/panel kernels, split copies excluded); pairs are read by imps_plan_timing
    bool timing = false;
    std::vector<hipEvent_t> tev;
    size_t tev_used = 0;
    std::shared_ptr<imps::EnvSolver> env;  // K_ENV
    std::shared_ptr<imps::HeffEngine> heff;  // K_HEFF
    // layouts of the operands (TM: A, B, x, y; H_AC: AC, EHL, EHR, HL, HR; H_C: C, EHL, EHR, HL, AR),
    // the TM direction and whether it is mixed; the adjoint program, compiled on first use
    std::vector<imps::Layout> lays;
    int tm_dir = 0;
    bool tm_mixed = false;
    std::shared_ptr<imps::VjpTape> vjp;
    imps::QrPlanData qr;                   // K_QR
    int qr_dir = 0;
    size_t elem_bytes() const { return prec == IMPS_FP64 ? 8 : 4; }
    ~imps_plan_s();
};

namespace imps {

imps_plan_s* new_plan(imps_context ctx, int kind, int prec);
void upload_program(imps_plan_s* P);
void set_operator(imps_plan_s* P, const std::vector<const double*>& parts);
// y = op(x) for nvec vectors on the context stream; op_override (FP64 plans) replaces the operator
// store by a caller buffer (two-operand products)
void run(imps_plan_s* P, const double* x, double* y, int nvec, const double* op_override = nullptr);
// y = op(x) for a panel plan (H_AC) with the split copies of the input in the caller-given device
// scratch of `bytes` bytes instead of the plan's own buffers: no allocation, no event, no
// autotuning (the tuned schedule if one exists for nvec, else the default rule); the used part
// of the scratch is cleared by a fill kernel first (pads must read as zeros)
void run_static(imps_plan_s* P, const double* x, double* y, int nvec, void* scratch, size_t bytes);
// bytes of scratch run_static needs for nvec vectors
size_t static_bytes(const imps_plan_s* P, int nvec);
// GPU milliseconds of the timed GEMM launches since the last call (waits for them), then resets
double take_timing(imps_plan_s* P, int* launches);
// panel schedule for nvec vectors: 0 = K-split 1, 1 = K-split 1 with shared panels, 2 = K-split 2,
// 3 = K-split 4 (the tuned one if the plan has applied nvec vectors, else the default rule)
int panel_schedule_of(const imps_plan_s* P, int nvec);

}  // namespace imps
