// Plan construction helpers and execution of grouped-GEMM programs.
#include "plan.hpp"

#include "adjoint.hpp"
#include "heff.hpp"

#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>

imps_plan_s::~imps_plan_s() {
    for (hipEvent_t e : tev) (void)hipEventDestroy(e);
}

namespace imps {

void ctx_release(imps_context ctx) {
    if (!ctx || ctx->refs.fetch_sub(1) > 1) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->own) (void)hipStreamDestroy(ctx->own);
    delete ctx;
}

imps_plan_s* new_plan(imps_context ctx, int kind, int prec) {
    if (prec < IMPS_FP64 || prec > IMPS_BF16X3) throw InvalidArg("unknown precision");
    auto* P = new imps_plan_s;
    P->hold.hold(ctx);
    P->ctx = ctx;
    P->kind = kind;
    P->prec = prec;
    return P;
}

namespace {

// IMPS_PANELS=0 in the environment selects the segment kernels for every plan (comparison runs).
bool panels_enabled() {
    const char* e = std::getenv("IMPS_PANELS");
    return !(e && e[0] == '0');
}

// Panel schedules: K-split (waves sharing one vector's tile) and, for K-split 1, per-wave or shared
// (LDS) panel loads.  Encoded 0 = K-split 1, 1 = K-split 1 shared, 2 = K-split 2, 3 = K-split 4.
constexpr int kSchedKsplit[4] = {1, 1, 2, 4};
constexpr bool kSchedShared[4] = {false, true, false, false};

// IMPS_KSPLIT forces a K-split (comparison runs and tests); -1 when unset
int forced_ksplit() {
    if (const char* e = std::getenv("IMPS_KSPLIT")) {
        const int k = std::atoi(e);
        if (k == 1 || k == 2 || k == 4) return k;
    }
    return -1;
}

// The rule used without tuning: one vector -> all four waves on its tile (a single apply still
// spreads over the chip); 2-3 vectors -> two waves; four or more -> one wave per vector, with the
// shared-panel kernel in BF16X3.
int default_schedule(int prec, int nvec) {
    const int k = forced_ksplit() > 0 ? forced_ksplit() : nvec >= 4 ? 1 : nvec >= 2 ? 2 : 4;
    return k == 4 ? 3 : k == 2 ? 2 : (panel_shared_default(prec) ? 1 : 0);
}

// IMPS_SCHEDULE=0..3 forces one schedule (tests of every path the autotuner may pick); -1 when unset
int forced_schedule() {
    if (const char* e = std::getenv("IMPS_SCHEDULE")) {
        if (e[0] >= '0' && e[0] <= '3' && e[1] == '\0') return e[0] - '0';
    }
    return -1;
}

// IMPS_AUTOTUNE=0 keeps the default rule (bitwise-reproducible runs: schedules sum in different orders)
bool autotune_enabled() {
    const char* e = std::getenv("IMPS_AUTOTUNE");
    return !(e && e[0] == '0') && forced_ksplit() < 0 && forced_schedule() < 0;
}

void launch_panel_schedule(imps_plan_s* P, int sched, const PanelLaunch& L, hipStream_t s) {
    for (auto& ds : P->dpstages)
        hip_check(launch_panel_stage(P->prec, ds, kSchedKsplit[sched], kSchedShared[sched], L, s), "panel gemm");
}

// Autotuner: on the first apply of a plan with nvec vectors, every schedule runs the GEMM stages
// once to warm up and three times under HIP events on the plan's stream; the fastest is kept for
// the plan and that vector count.  Skipped while the stream is being captured into a graph.
// Tuned schedules kept in a text file (IMPS_TUNE_CACHE=path): one "key schedule" line per tuned
// (panel program, precision, vector count, input read), the key a 64-bit FNV-1a hash of the
// program's task and chunk lists.  A plan whose key is in the file takes the recorded schedule
// without timing, so a run (or a later one) repeats the first run's choices bitwise; a new key is
// timed and appended.
std::mutex g_tune_mu;

uint64_t fnv(uint64_t h, const void* p, size_t n) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; ++i) {
        h ^= b[i];
        h *= 1099511628211ull;
    }
    return h;
}

// field by field: the structs' padding bytes are not part of the program
uint64_t schedule_key(const imps_plan_s* P, int nvec, bool xt) {
    uint64_t h = 1469598103934665603ull;
    const int head[4] = {P->kind, P->prec, nvec, xt ? 1 : 0};
    h = fnv(h, head, sizeof(head));
    for (const PStage& st : P->pprog.stages) {
        const long hs[3] = {st.out_slot, (long)st.tasks.size(), (long)st.chunks.size()};
        h = fnv(h, hs, sizeof(hs));
        for (const PTask& t : st.tasks) {
            const long v[8] = {t.c_off, t.c_ld, t.m0, t.n0, t.M, t.N, t.ch0, t.nch};
            h = fnv(h, v, sizeof(v));
        }
        for (const PChunk& c : st.chunks) {
            const long v[8] = {c.raw_off, c.pan, c.pan_sub, c.s_idx, c.flags, c.krem, c.xoff, c.xks};
            h = fnv(h, v, sizeof(v));
        }
    }
    return h;
}

int cached_schedule(const char* path, uint64_t key) {
    FILE* f = std::fopen(path, "r");
    if (!f) return -1;
    unsigned long long k = 0;
    int sc = -1, found = -1;
    while (std::fscanf(f, "%llx %d", &k, &sc) == 2)
        if (k == key && sc >= 0 && sc < 4) found = sc;
    std::fclose(f);
    return found;
}

int choose_schedule(imps_plan_s* P, int nvec, const PanelLaunch& L, hipStream_t s) {
    if (forced_schedule() >= 0) return forced_schedule();
    auto it = P->sched.find(nvec);
    if (it != P->sched.end()) return it->second;
    const int def = default_schedule(P->prec, nvec);
    if (!autotune_enabled()) return def;
    const char* cache = std::getenv("IMPS_TUNE_CACHE");
    const uint64_t key = cache && cache[0] ? schedule_key(P, nvec, L.xt_direct) : 0;
    if (key) {
        std::lock_guard<std::mutex> g(g_tune_mu);
        const int sc = cached_schedule(cache, key);
        if (sc >= 0) {
            P->sched[nvec] = sc;
            return sc;
        }
    }
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    hip_check(hipStreamIsCapturing(s, &cap), "hipStreamIsCapturing");
    if (cap != hipStreamCaptureStatusNone) return def;
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "hipEventCreate");
    hip_check(hipEventCreate(&e1), "hipEventCreate");
    int best = def;
    float best_ms = 0.0f;
    for (int sc = 0; sc < 4; ++sc) {
        launch_panel_schedule(P, sc, L, s);
        hip_check(hipEventRecord(e0, s), "hipEventRecord");
        for (int r = 0; r < 3; ++r) launch_panel_schedule(P, sc, L, s);
        hip_check(hipEventRecord(e1, s), "hipEventRecord");
        hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
        float ms = 0.0f;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        if (sc == 0 || ms < best_ms) {
            best_ms = ms;
            best = sc;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    P->sched[nvec] = best;
    if (key) {
        std::lock_guard<std::mutex> g(g_tune_mu);
        if (cached_schedule(cache, key) < 0)
            if (FILE* f = std::fopen(cache, "a")) {
                std::fprintf(f, "%016" PRIx64 " %d\n", key, best);
                std::fclose(f);
            }
    }
    return best;
}

template <typename T>
size_t put(std::vector<char>& host, size_t pos, const std::vector<T>& v) {
    pos = (pos + 255) & ~size_t(255);
    if (host.size() < pos + v.size() * sizeof(T)) host.resize(pos + v.size() * sizeof(T));
    if (!v.empty()) std::memcpy(host.data() + pos, v.data(), v.size() * sizeof(T));
    return pos;
}

void upload_panel_program(imps_plan_s* P) {
    P->pprog = build_panel_program(P->prog, P->prec);
    std::vector<char> host;
    std::vector<size_t> tpos, cpos, kpos;
    size_t pos = 0;
    for (auto& st : P->pprog.stages) {
        tpos.push_back(pos = put(host, pos, st.tasks));
        pos += st.tasks.size() * sizeof(PTask);
        cpos.push_back(pos = put(host, pos, st.chunks));
        pos += st.chunks.size() * sizeof(PChunk);
        kpos.push_back(pos = put(host, pos, st.packs));
        pos += st.packs.size() * sizeof(RawPack);
    }
    const size_t ppos = put(host, pos, P->pprog.panels);
    pos = ppos + P->pprog.panels.size() * sizeof(PanelDesc);
    const size_t xpos = put(host, pos, P->pprog.xspans);
    pos = xpos + P->pprog.xspans.size() * sizeof(SplitSpan);
    const size_t upos = put(host, pos, P->pprog.xunits);
    pos = upos + P->pprog.xunits.size() * sizeof(SplitUnit);
    P->d_pprog.ensure(pos + 256);
    if (pos) hip_check(hipMemcpyAsync(P->d_pprog.p, host.data(), pos, hipMemcpyHostToDevice, P->ctx->stream), "upload panel program");
    hip_check(hipStreamSynchronize(P->ctx->stream), "upload panel program");
    char* base = (char*)P->d_pprog.p;
    P->dpstages.clear();
    for (size_t i = 0; i < P->pprog.stages.size(); ++i) {
        DevPStage ds;
        ds.tasks = (const PTask*)(base + tpos[i]);
        ds.chunks = (const PChunk*)(base + cpos[i]);
        ds.packs = (const RawPack*)(base + kpos[i]);
        ds.ntasks = (int)P->pprog.stages[i].tasks.size();
        ds.out_slot = P->pprog.stages[i].out_slot;
        P->dpstages.push_back(ds);
    }
    P->d_panels = (const PanelDesc*)(base + ppos);
    P->d_xspans = (const SplitSpan*)(base + xpos);
    P->d_xunits = (const SplitUnit*)(base + upos);
    P->d_pstore.ensure((size_t)std::max<long>(P->pprog.nfrag, 1) * panel_frag_bytes(P->prec));
    P->d_op.ensure(std::max<long>(P->prog.op_elems, 1) * sizeof(double));  // FP64 staging of the operator pieces
    P->xs_ready = P->ws_ready = 0;
}

}  // namespace

// FP64 applies of at most this many vectors read every raw block from x itself (no split copy)
constexpr int kDirectTMaxVectors = 2;

// true when the element ranges [a, a + na) and [b, b + nb) share memory
static bool overlaps(const double* a, long na, const double* b, long nb) {
    const uintptr_t a0 = (uintptr_t)a, a1 = a0 + sizeof(double) * (size_t)na;
    const uintptr_t b0 = (uintptr_t)b, b1 = b0 + sizeof(double) * (size_t)nb;
    return a0 < b1 && b0 < a1;
}

void upload_program(imps_plan_s* P) {
    if (!P->ctx) return;
    hip_check(hipSetDevice(P->ctx->device), "hipSetDevice");
    size_t bytes = 0;
    for (auto& st : P->prog.stages) bytes += st.tasks.size() * sizeof(Task) + st.segs.size() * sizeof(Seg) + 64;
    P->d_prog.ensure(bytes + 64);
    std::vector<char> host(bytes + 64, 0);
    size_t pos = 0;
    P->dstages.clear();
    for (auto& st : P->prog.stages) {
        DevStage ds;
        ds.ntasks = (int)st.tasks.size();
        ds.out_slot = st.out_slot;
        pos = (pos + 15) & ~size_t(15);
        std::memcpy(host.data() + pos, st.tasks.data(), st.tasks.size() * sizeof(Task));
        ds.tasks = (const Task*)((char*)P->d_prog.p + pos);
        pos += st.tasks.size() * sizeof(Task);
        pos = (pos + 15) & ~size_t(15);
        std::memcpy(host.data() + pos, st.segs.data(), st.segs.size() * sizeof(Seg));
        ds.segs = (const Seg*)((char*)P->d_prog.p + pos);
        pos += st.segs.size() * sizeof(Seg);
        P->dstages.push_back(ds);
    }
    if (pos) hip_check(hipMemcpyAsync(P->d_prog.p, host.data(), pos, hipMemcpyHostToDevice, P->ctx->stream), "upload program");
    hip_check(hipStreamSynchronize(P->ctx->stream), "upload program");
    P->use_panels = (P->kind == K_TM || P->kind == K_HAC || P->kind == K_HC) && panel_compatible(P->prog) && panels_enabled();
    if (P->use_panels) upload_panel_program(P);
    else P->d_op.ensure(std::max<long>(P->prog.op_elems, 1) * P->elem_bytes());
}

void set_operator(imps_plan_s* P, const std::vector<const double*>& parts) {
    if (!P->ctx) throw InvalidArg("plan was created without a context");
    if (parts.size() != P->op_parts.size()) throw InvalidArg("operator piece count mismatch");
    hip_check(hipSetDevice(P->ctx->device), "hipSetDevice");
    long off = 0;
    for (size_t i = 0; i < parts.size(); ++i) {
        const long n = P->op_parts[i];
        if (n > 0 && !parts[i]) throw InvalidArg("operator data pointer is NULL");
        if (n > 0) {
            if (P->use_panels || P->prec == IMPS_FP64)
                hip_check(hipMemcpyAsync((double*)P->d_op.p + off, parts[i], n * sizeof(double), hipMemcpyDeviceToDevice, P->ctx->stream),
                          "copy operator");
            else
                hip_check(launch_d2f(parts[i], (float*)P->d_op.p + off, n, P->ctx->stream), "convert operator");
        }
        off += n;
    }
    if (P->use_panels)
        hip_check(launch_panelize(P->prec, P->d_panels, (int)P->pprog.panels.size(), (const double*)P->d_op.p, P->d_pstore.p, P->ctx->stream),
                  "pack operator panels");
    P->op_set = true;
}

// records the next timing event of an apply on stream s (no-op unless the plan's timing is on)
static void time_mark(imps_plan_s* P, hipStream_t s) {
    if (!P->timing) return;
    if (P->tev_used == P->tev.size()) {
        hipEvent_t e;
        hip_check(hipEventCreate(&e), "hipEventCreate");
        P->tev.push_back(e);
    }
    hip_check(hipEventRecord(P->tev[P->tev_used++], s), "hipEventRecord");
}

void run(imps_plan_s* P, const double* x, double* y, int nvec, const double* op_override) {
    if (!P->ctx) throw InvalidArg("plan was created without a context");
    if (nvec < 1) throw InvalidArg("nvec must be >= 1");
    if ((P->prog.x_elems > 0 && !x) || (P->prog.y_elems > 0 && !y)) throw InvalidArg("x or y is NULL");
    if (!P->op_set && !op_override && P->prog.op_elems > 0) throw InvalidArg("operator data not set");
    hip_check(hipSetDevice(P->ctx->device), "hipSetDevice");
    hipStream_t s = P->ctx->stream;
    if (P->use_panels && !op_override) {
        const PanelProgram& Q = P->pprog;
        const long xsb = split_bytes(P->prec, P->prog.x_elems), wsb = split_bytes(P->prec, P->prog.w_elems);
        // split copies: zero-filled when (re)allocated, so that pads and unused places hold zeros
        if (P->xs_ready < nvec) {
            P->d_xs.ensure((size_t)xsb * nvec);
            hip_check(launch_zero((double*)P->d_xs.p, xsb * nvec / 8, s), "zero split copy");
            P->xs_ready = nvec;
        }
        if (P->prog.w_elems > 0 && P->ws_ready < nvec) {
            P->d_ws.ensure((size_t)wsb * nvec);
            hip_check(launch_zero((double*)P->d_ws.p, wsb * nvec / 8, s), "zero workspace");
            P->ws_ready = nvec;
        }
        // Unless y overlaps x (in-place applies keep the split copy): FP64 plane-N chunks read x
        // itself; at one or two vectors every input-vector chunk reads x (plane-T ones along its
        // stored columns; FP32 / BF16X3 convert on load) and no split copy is made (measured at
        // D = 1024 H_AC in FP64: 13.7 us per single-vector apply against 13.4 + 5 us with the copy;
        // with 16 vectors the copy pays: 60 + 9.5 us against 68 us)
        const bool apart = !overlaps(x, P->prog.x_elems * nvec, y, P->prog.y_elems * nvec);
        const bool xt = apart && nvec <= kDirectTMaxVectors;
        const bool direct = apart && (P->prec == IMPS_FP64 || xt);
        if (!xt)
            hip_check(launch_split(P->prec, P->d_xunits, (int)Q.xunits.size(), P->d_xspans, x, P->prog.x_elems, P->d_xs.p, xsb, Q.plx, nvec, s,
                                   !direct),
                      "split copy");
        PanelLaunch L;
        L.pstore = P->d_pstore.p;
        L.pstore_bytes = Q.nfrag * panel_frag_bytes(P->prec);
        L.xs = P->d_xs.p;
        L.xs_stride = xsb;
        L.xd = direct ? x : nullptr;
        L.xd_stride = P->prog.x_elems;
        L.xt_direct = xt;
        L.ws = P->d_ws.p;
        L.ws_stride = wsb;
        L.plx = Q.plx;
        L.plw = Q.plw;
        L.out = y;
        L.out_stride = P->prog.y_elems;
        L.nvec = nvec;
        const int sc = choose_schedule(P, nvec, L, s);
        time_mark(P, s);
        launch_panel_schedule(P, sc, L, s);
        time_mark(P, s);
        return;
    }
    P->d_w.ensure((size_t)std::max<long>(P->prog.w_elems, 1) * nvec * P->elem_bytes());
    Slots sl;
    sl.base[SLOT_X] = x;           sl.stride[SLOT_X] = P->prog.x_elems;  // FP64, converted in-kernel
    sl.base[SLOT_OP] = op_override ? (const void*)op_override : P->d_op.p;
    sl.stride[SLOT_OP] = 0;
    sl.base[SLOT_W] = P->d_w.p;    sl.stride[SLOT_W] = P->prog.w_elems;
    time_mark(P, s);
    for (auto& ds : P->dstages) {
        if (ds.out_slot == SLOT_W)
            hip_check(launch_stage(P->prec, ds, sl, P->d_w.p, P->prog.w_elems, P->prec == IMPS_FP64, nvec, s), "grouped gemm");
        else
            hip_check(launch_stage(P->prec, ds, sl, y, P->prog.y_elems, true, nvec, s), "grouped gemm");
    }
    time_mark(P, s);
}

size_t static_bytes(const imps_plan_s* P, int nvec) {
    return (size_t)split_bytes(P->prec, P->prog.x_elems) * (size_t)std::max(nvec, 0);
}

void run_static(imps_plan_s* P, const double* x, double* y, int nvec, void* scratch, size_t bytes) {
    if (!P->ctx) throw InvalidArg("plan was created without a context");
    if (nvec < 1) throw InvalidArg("nvec must be >= 1");
    if (!P->use_panels || P->prog.w_elems > 0) throw InvalidArg("static apply: needs a single-stage panel plan (H_AC)");
    if (!P->op_set) throw InvalidArg("operator data not set");
    if ((P->prog.x_elems > 0 && !x) || (P->prog.y_elems > 0 && !y)) throw InvalidArg("x or y is NULL");
    const long xsb = split_bytes(P->prec, P->prog.x_elems);
    if (static_bytes(P, nvec) > bytes)
        throw InvalidArg("static apply: " + std::to_string(nvec) + " vectors need " + std::to_string(static_bytes(P, nvec)) +
                         " bytes of scratch, the fixed scratch holds " + std::to_string(bytes));
    hip_check(hipSetDevice(P->ctx->device), "hipSetDevice");
    hipStream_t s = P->ctx->stream;
    const PanelProgram& Q = P->pprog;
    const bool apart = !overlaps(x, P->prog.x_elems * nvec, y, P->prog.y_elems * nvec);  // as in run()
    const bool xt = apart && nvec <= kDirectTMaxVectors;
    const bool direct = apart && (P->prec == IMPS_FP64 || xt);
    if (!xt) {
        hip_check(launch_zero((double*)scratch, xsb * nvec / 8, s), "clear scratch");
        hip_check(launch_split(P->prec, P->d_xunits, (int)Q.xunits.size(), P->d_xspans, x, P->prog.x_elems, scratch, xsb, Q.plx, nvec, s,
                               !direct),
                  "split copy");
    }
    PanelLaunch L;
    L.pstore = P->d_pstore.p;
    L.pstore_bytes = Q.nfrag * panel_frag_bytes(P->prec);
    L.xs = scratch;
    L.xs_stride = xsb;
    L.xd = direct ? x : nullptr;
    L.xd_stride = P->prog.x_elems;
    L.xt_direct = xt;
    L.ws = nullptr;
    L.ws_stride = 0;
    L.plx = Q.plx;
    L.plw = Q.plw;
    L.out = y;
    L.out_stride = P->prog.y_elems;
    L.nvec = nvec;
    launch_panel_schedule(P, panel_schedule_of(P, nvec), L, s);
}

int panel_schedule_of(const imps_plan_s* P, int nvec) {
    if (forced_schedule() >= 0) return forced_schedule();
    auto it = P->sched.find(nvec);
    return it != P->sched.end() ? it->second : default_schedule(P->prec, nvec);
}

double take_timing(imps_plan_s* P, int* launches) {
    double ms = 0.0;
    const size_t pairs = P->tev_used / 2;
    if (pairs > 0) {
        hip_check(hipEventSynchronize(P->tev[2 * pairs - 1]), "timing event wait");
        for (size_t i = 0; i < pairs; ++i) {
            float e = 0.0f;
            hip_check(hipEventElapsedTime(&e, P->tev[2 * i], P->tev[2 * i + 1]), "timing event elapsed");
            ms += e;
        }
    }
    if (launches) *launches = (int)pairs;
    P->tev_used = 0;
    return ms;
}

}  // namespace imps
