// Device Krylov drivers (see krylov.hpp).
#include "krylov.hpp"

#include <algorithm>
#include <cstdlib>
#include <deque>
#include <memory>
#include <cmath>
#include <map>
#include <mutex>

#include "blas.hpp"

namespace imps {

namespace {

// Pinned host staging for device-to-host reads (one buffer per host thread, grown on demand and
// freed when the thread exits): the copy is a direct DMA instead of a staged pageable copy.
struct PinnedStage {
    double* p = nullptr;
    size_t cap = 0;
    PinnedStage() = default;
    PinnedStage(const PinnedStage&) = delete;
    PinnedStage& operator=(const PinnedStage&) = delete;
    ~PinnedStage() { if (p) (void)hipHostFree(p); }
    double* get(size_t n) {
        if (n > cap) {
            if (p) hip_check(hipHostFree(p), "hipHostFree");
            p = nullptr;
            cap = 0;
            const size_t c = std::max<size_t>(n, 4096);
            hip_check(hipHostMalloc((void**)&p, c * sizeof(double), hipHostMallocDefault), "hipHostMalloc");
            cap = c;
        }
        return p;
    }
};
thread_local PinnedStage pinned_stage;

void d2h(hipStream_t s, double* h, const double* d, size_t n) {
    if (n == 0) return;
    double* p = pinned_stage.get(n);
    hip_check(hipMemcpyAsync(p, d, n * sizeof(double), hipMemcpyDeviceToHost, s), "D2H");
    hip_check(hipStreamSynchronize(s), "sync");
    std::copy(p, p + n, h);
}
// synchronous: the host source may be a short-lived temporary
void h2d(hipStream_t s, double* d, const double* h, size_t n) {
    hip_check(hipMemcpyAsync(d, h, n * sizeof(double), hipMemcpyHostToDevice, s), "H2D");
    hip_check(hipStreamSynchronize(s), "sync");
}
void d2d(hipStream_t s, double* d, const double* src, size_t n) {
    if (n) hip_check(hipMemcpyAsync(d, src, n * sizeof(double), hipMemcpyDeviceToDevice, s), "D2D");
}

// Device scratch of the Krylov drivers, reused across calls: one pool per device, shared by all
// host threads (a mutex guards it).  A call takes free buffers (the smallest that fits, else grows
// the largest free one) and gives them back when it ends, so repeated solves do not allocate and a
// thread that exits leaves its buffers to the next caller.  The pool holds at most as many buffers
// as were in use at once; they stay allocated for the process (never freed, which keeps the HIP
// runtime's teardown order out of the picture).
class BufPool {
  public:
    static BufPool& get() {
        static std::mutex m;
        static std::map<int, BufPool*> pools;
        int dev = 0;
        hip_check(hipGetDevice(&dev), "hipGetDevice");
        std::lock_guard<std::mutex> g(m);
        BufPool*& p = pools[dev];
        if (!p) p = new BufPool();
        return *p;
    }
    DevBuf* take(size_t bytes) {
        std::lock_guard<std::mutex> g(mu_);
        Slot* fit = nullptr;      // smallest free buffer that holds `bytes`
        Slot* largest = nullptr;  // else the largest free one, grown
        for (auto& s : slots_) {
            if (s.used) continue;
            if (s.buf->bytes >= bytes && (!fit || s.buf->bytes < fit->buf->bytes)) fit = &s;
            if (!largest || s.buf->bytes > largest->buf->bytes) largest = &s;
        }
        Slot* best = fit ? fit : largest;
        if (!best) {
            slots_.push_back(Slot{std::make_unique<DevBuf>(), false});
            best = &slots_.back();
        }
        best->used = true;
        try {
            best->buf->ensure(std::max<size_t>(bytes, 8));
        } catch (...) {
            best->used = false;
            throw;
        }
        return best->buf.get();
    }
    void give(DevBuf* b) {
        std::lock_guard<std::mutex> g(mu_);
        for (auto& s : slots_)
            if (s.buf.get() == b) s.used = false;
    }
    // buffers and bytes held (introspection for tests)
    void stats(int* nbuf, size_t* bytes) {
        std::lock_guard<std::mutex> g(mu_);
        size_t t = 0;
        for (auto& s : slots_) t += s.buf->bytes;
        *nbuf = (int)slots_.size();
        *bytes = t;
    }

  private:
    struct Slot {
        std::unique_ptr<DevBuf> buf;
        bool used;
    };
    std::mutex mu_;
    std::deque<Slot> slots_;  // stable addresses
};

// A pooled device buffer held for the lifetime of this object.
struct PoolBuf {
    DevBuf* b = nullptr;
    PoolBuf() = default;
    PoolBuf(const PoolBuf&) = delete;
    PoolBuf& operator=(const PoolBuf&) = delete;
    ~PoolBuf() { if (b) BufPool::get().give(b); }
    void ensure(size_t bytes) {
        if (b && b->bytes >= bytes) return;
        if (b) BufPool::get().give(b);
        b = BufPool::get().take(bytes);
    }
    double* d() const { return b->d(); }
};

// Scratch for one Krylov run: basis V (m+1 vectors), w, reduction buffers, and the raw CGS2
// coefficients of up to m + 1 steps (step j at coef + j * cstride).
struct Work {
    PoolBuf V, w, dh, scratch, coef;
    long n = 0;
    int m = 0;
    int cstride = 0;
    void init(long n_, int m_) {
        n = n_;
        m = m_;
        cstride = 2 * (m + 1) + 1;  // >= 2k + 1 for k <= m + 1
        V.ensure(sizeof(double) * (size_t)n * (m + 1));
        w.ensure(sizeof(double) * (size_t)std::max<long>(n, 1));
        dh.ensure(sizeof(double) * (size_t)(2 * (m + 2)));
        scratch.ensure(sizeof(double) * blas_scratch_doubles(m + 2));
        coef.ensure(sizeof(double) * (size_t)cstride * (m + 1));
    }
    double* Vi(int i) const { return V.d() + (size_t)i * n; }
    double* Ci(int j) const { return coef.d() + (size_t)j * cstride; }
};

double dev_norm(hipStream_t s, Work& W, const double* v) {
    double h = 0;
    hip_check(blas_multidot(v, 0, 1, v, W.n, W.dh.d(), W.scratch.d(), s), "norm");
    d2h(s, &h, W.dh.d(), 1);
    return std::sqrt(std::max(h, 0.0));
}

// One Krylov step's orthogonalisation, on the device only: w <- w - V[0..k) (V^T w), twice (CGS2),
// then w <- w / ||w|| (a zero w stays zero).  The 2k + 1 raw coefficients (first-pass projections,
// second-pass projections, ||w||^2 before the scaling) stay at dc for the host to collect later,
// so consecutive steps queue on the stream without a host round trip.
void cgs2_dev(hipStream_t s, Work& W, int k, double* w, double* dc) {
    // five launches: inner products (chunk partials), update (sums them), the same again with the
    // partials of ||w||^2, and the scaling (sums those)
    const int nch = blas_chunks(W.n);
    double* p = W.scratch.d();
    double* qp = p + (size_t)nch * k;
    for (int pass = 0; pass < 2; ++pass) {
        hip_check(blas_multidot_partials(W.V.d(), W.n, k, w, W.n, p, s), "multidot");
        hip_check(blas_gemv_sub_partials(W.V.d(), W.n, k, p, nch, dc + pass * k, w, W.n, pass ? qp : nullptr, s), "gemv");
    }
    hip_check(blas_scale_inv_norm(qp, blas_gemv_grid(W.n), dc + 2 * k, w, W.n, s), "normalise");
}

// A Krylov cycle's launch sequence as a HIP graph: captured on the first run() with a given key,
// replayed afterwards.  Graphs need a capturable stream (not the legacy default stream, not one
// already being captured); run() returns false when the sequence was not run (the caller then
// issues it directly).  Capture is thread-local, so other host threads may keep allocating.
// Opt-in (IMPS_KRYLOV_GRAPH=1): on the D = 1024 Hubbard VUMPS iteration, capturing per solve and
// replaying the restart cycles measured no faster than issuing the launches directly.
class CycleGraph {
  public:
    explicit CycleGraph(hipStream_t s) : s_(s) {
        const char* e = std::getenv("IMPS_KRYLOV_GRAPH");
        enabled_ = e && e[0] == '1' && s != nullptr;
        if (enabled_) {
            hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
            enabled_ = hipStreamIsCapturing(s, &cap) == hipSuccess && cap == hipStreamCaptureStatusNone;
        }
    }
    ~CycleGraph() { reset(); }
    template <class F>
    bool run(int key, F&& issue) {
        if (!enabled_) return false;
        if (!exec_ || key != key_) {
            reset();
            hipGraph_t g = nullptr;
            hip_check(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            try {
                issue();
            } catch (...) {
                (void)hipStreamEndCapture(s_, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            hip_check(hipStreamEndCapture(s_, &g), "hipStreamEndCapture");
            const hipError_t e = hipGraphInstantiate(&exec_, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                exec_ = nullptr;
                enabled_ = false;
                return false;
            }
            key_ = key;
        }
        hip_check(hipGraphLaunch(exec_, s_), "hipGraphLaunch");
        return true;
    }

  private:
    void reset() {
        if (exec_) (void)hipGraphExecDestroy(exec_);
        exec_ = nullptr;
    }
    hipStream_t s_;
    bool enabled_ = false;
    hipGraphExec_t exec_ = nullptr;
    int key_ = -1;
};

// Host side of cgs2_dev: h = the k projection coefficients (both passes summed) from the raw
// values t; returns ||w|| before the scaling.
double cgs2_coeffs(const double* t, int k, std::vector<double>& h) {
    h.resize(k);
    for (int i = 0; i < k; ++i) h[i] = t[i] + t[k + i];
    return std::sqrt(std::max(t[2 * k], 0.0));
}

// Cyclic Jacobi eigen-decomposition of a symmetric n x n matrix (column-major, FP64).
void jacobi_eig(int n, std::vector<double> A, std::vector<double>& w, std::vector<double>& Z) {
    Z.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) Z[i + (size_t)i * n] = 1.0;
    auto a = [&](int i, int j) -> double& { return A[i + (size_t)j * n]; };
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0, tot = 0;
        for (int j = 0; j < n; ++j) for (int i = 0; i < n; ++i) { tot += a(i, j) * a(i, j); if (i != j) off += a(i, j) * a(i, j); }
        if (off <= 1e-30 * tot || off == 0) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = a(p, q);
                if (apq == 0) continue;
                const double theta = (a(q, q) - a(p, p)) / (2 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::abs(theta) + std::sqrt(theta * theta + 1));
                const double c = 1 / std::sqrt(t * t + 1), sn = t * c;
                for (int k = 0; k < n; ++k) {
                    const double akp = a(k, p), akq = a(k, q);
                    a(k, p) = c * akp - sn * akq;
                    a(k, q) = sn * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    const double apk = a(p, k), aqk = a(q, k);
                    a(p, k) = c * apk - sn * aqk;
                    a(q, k) = sn * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    const double zkp = Z[k + (size_t)p * n], zkq = Z[k + (size_t)q * n];
                    Z[k + (size_t)p * n] = c * zkp - sn * zkq;
                    Z[k + (size_t)q * n] = sn * zkp + c * zkq;
                }
            }
    }
    w.resize(n);
    for (int i = 0; i < n; ++i) w[i] = a(i, i);
}

}  // namespace

KrylovStats gmres(hipStream_t s, const LinOp& A, const double* b, double* x, double tol, int maxit, int m) {
    KrylovStats st;
    const long n = A.n;
    if (n == 0) { st.converged = true; return st; }
    m = (int)std::max<long>(1, std::min<long>(m, n));
    Work W;
    W.init(n, m);
    const double bnorm = dev_norm(s, W, b);
    if (bnorm == 0) {
        hip_check(launch_zero(x, n, s), "zero");
        st.converged = true;
        return st;
    }
    std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), h, raw;
    // Steps are queued in batches and their coefficients read once per batch; the batch length
    // follows the convergence rate seen so far (about the steps still needed, 1 to 16), so few
    // steps are computed past convergence.  Those extra steps do not enter the solution.
    int batch = 8;
    while (st.iters < maxit) {
        // r = b - A x  -> V0
        A.apply(x, W.Vi(0));
        ++st.iters;
        hip_check(blas_axpby(1.0, nullptr, AXPBY_DA, b, -1.0, W.Vi(0), n, s), "residual");
        const double beta = dev_norm(s, W, W.Vi(0));
        st.relres = beta / bnorm;
        if (beta < tol * bnorm) { st.converged = true; break; }
        hip_check(blas_axpby(1.0 / beta, nullptr, AXPBY_DA, W.Vi(0), 0.0, W.Vi(0), n, s), "scale");
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int k = 0;
        bool stop = false;
        for (int j0 = 0; j0 < m && st.iters < maxit && !stop;) {
            const int q = std::min({batch, m - j0, maxit - st.iters});
            for (int j = j0; j < j0 + q; ++j) {
                A.apply(W.Vi(j), W.Vi(j + 1));
                ++st.iters;
                cgs2_dev(s, W, j + 1, W.Vi(j + 1), W.Ci(j - j0));
            }
            raw.resize((size_t)q * W.cstride);
            d2h(s, raw.data(), W.Ci(0), raw.size());
            const double rel0 = st.relres;
            for (int j = j0; j < j0 + q; ++j) {
                const double hn = cgs2_coeffs(raw.data() + (size_t)(j - j0) * W.cstride, j + 1, h);
                for (int i = 0; i <= j; ++i) H[i + (size_t)j * (m + 1)] = h[i];
                H[j + 1 + (size_t)j * (m + 1)] = hn;
                for (int i = 0; i < j; ++i) {  // previous Givens rotations
                    const double t1 = H[i + (size_t)j * (m + 1)], t2 = H[i + 1 + (size_t)j * (m + 1)];
                    H[i + (size_t)j * (m + 1)] = cs[i] * t1 + sn[i] * t2;
                    H[i + 1 + (size_t)j * (m + 1)] = -sn[i] * t1 + cs[i] * t2;
                }
                const double hjj = H[j + (size_t)j * (m + 1)];
                const double beta2 = std::hypot(hjj, hn);
                k = j + 1;
                if (beta2 == 0) { stop = true; break; }
                cs[j] = hjj / beta2;
                sn[j] = hn / beta2;
                H[j + (size_t)j * (m + 1)] = beta2;
                H[j + 1 + (size_t)j * (m + 1)] = 0;
                g[j + 1] = -sn[j] * g[j];
                g[j] *= cs[j];
                st.relres = std::abs(g[j + 1]) / bnorm;
                if (st.relres < tol || hn == 0) { stop = true; break; }
            }
            j0 += q;
            if (!stop) {
                const double rate = st.relres > 0 && rel0 > 0 ? std::pow(st.relres / rel0, 1.0 / q) : 1.0;
                const double need = rate < 1.0 ? std::ceil(std::log(tol / st.relres) / std::log(rate)) : 16.0;
                batch = (int)std::min(16.0, std::max(1.0, need));
            }
        }
        // x += V[0..k) * R^{-1} g
        std::vector<double> y(k, 0.0);
        for (int i = k - 1; i >= 0; --i) {
            double t = g[i];
            for (int c = i + 1; c < k; ++c) t -= H[i + (size_t)c * (m + 1)] * y[c];
            y[i] = t / H[i + (size_t)i * (m + 1)];
        }
        for (auto& v : y) v = -v;  // gemv_sub subtracts
        h2d(s, W.dh.d(), y.data(), k);
        hip_check(blas_gemv_sub(W.V.d(), n, k, W.dh.d(), x, n, s), "update");
        if (st.relres < tol) {
            // confirm with the true residual on the next pass
            continue;
        }
    }
    if (!st.converged && st.iters >= maxit) {
        // final true residual
        A.apply(x, W.Vi(0));
        hip_check(blas_axpby(1.0, nullptr, AXPBY_DA, b, -1.0, W.Vi(0), n, s), "residual");
        st.relres = dev_norm(s, W, W.Vi(0)) / bnorm;
        st.converged = st.relres < tol;
    }
    return st;
}

KrylovStats gmres_refined(hipStream_t s, const LinOp& A64, const LinOp& Alo, const double* b, double* x, double tol, int maxit, int m,
                          double inner_tol) {
    KrylovStats st;
    const long n = A64.n;
    Work W;
    W.init(n, 1);
    PoolBuf r, dx;
    r.ensure(sizeof(double) * std::max<long>(n, 1));
    dx.ensure(sizeof(double) * std::max<long>(n, 1));
    const double bnorm = dev_norm(s, W, b);
    if (bnorm == 0) {
        hip_check(launch_zero(x, n, s), "zero");
        st.converged = true;
        return st;
    }
    // Each inner solve only needs to reduce the residual by a constant factor; it is capped at one
    // Krylov cycle: in low precision the residual estimate of an ill-conditioned operator
    // stagnates near eps_low * cond, and only the FP64 residual of the next round can tell.
    while (st.iters < maxit) {
        A64.apply(x, r.d());
        ++st.iters;
        hip_check(blas_axpby(1.0, nullptr, AXPBY_DA, b, -1.0, r.d(), n, s), "residual");
        const double rn = dev_norm(s, W, r.d());
        st.relres = rn / bnorm;
        if (st.relres < tol) { st.converged = true; break; }
        if (st.iters >= maxit) break;
        hip_check(launch_zero(dx.d(), n, s), "zero");
        const int budget = std::min(maxit - st.iters, m + 1);
        KrylovStats in = gmres(s, Alo, r.d(), dx.d(), std::max(inner_tol, 0.5 * tol * bnorm / rn), budget, m);
        st.iters += in.iters;
        hip_check(blas_axpby(1.0, nullptr, AXPBY_DA, dx.d(), 1.0, x, n, s), "update");
    }
    return st;
}

double rayleigh_relres(hipStream_t s, const LinOp& A, const double* x) {
    const long n = A.n;
    Work W;
    W.init(n, 1);
    d2d(s, W.Vi(0), x, n);
    A.apply(x, W.Vi(1));
    double h[2] = {0, 0}, xx = 0;
    hip_check(blas_multidot(W.V.d(), n, 2, W.Vi(1), n, W.dh.d(), W.scratch.d(), s), "rayleigh");  // x.Ax, Ax.Ax
    d2h(s, h, W.dh.d(), 2);
    hip_check(blas_multidot(x, 0, 1, x, n, W.dh.d(), W.scratch.d(), s), "rayleigh");
    d2h(s, &xx, W.dh.d(), 1);
    if (!(xx > 0)) throw InvalidArg("lanczos: zero start vector");
    const double theta = h[0] / xx;
    // ||Ax - theta x||^2 = Ax.Ax - (x.Ax)^2 / x.x: cancellation limits this to ~1e-8 relative
    return std::sqrt(std::max(h[1] - h[0] * theta, 0.0)) / std::max(std::abs(theta) * std::sqrt(xx), 1e-300);
}

KrylovStats lanczos_lowest(hipStream_t s, const LinOp& A, double* x, double tol, int maxit, int m) {
    // Thick-restart Lanczos (Wu-Simon): with full CGS2 re-orthogonalisation the projected matrix
    // T = V^T A V is assembled column by column from the orthogonalisation coefficients; at a
    // restart the lowest `keep` Ritz vectors (T diagonal there) and the next Lanczos direction
    // form the new basis, so the search space is not thrown away.  A cycle's steps run on the
    // device back to back; the host reads their coefficients once, at the end of the cycle, and
    // assembles T there (steps after a breakdown are computed and then ignored).
    KrylovStats st;
    const long n = A.n;
    m = (int)std::max<long>(2, std::min<long>(m, n));
    const int keep = std::max(1, std::min(m / 3, m - 2));
    Work W;
    W.init(n, m);
    PoolBuf Y;
    Y.ensure(sizeof(double) * (size_t)n * keep);
    double nx = dev_norm(s, W, x);
    if (nx == 0) throw InvalidArg("lanczos: zero start vector");
    hip_check(blas_axpby(1.0 / nx, nullptr, AXPBY_DA, x, 0.0, W.Vi(0), n, s), "scale");
    std::vector<double> T((size_t)m * m, 0.0), h, w, Z, raw;
    int j0 = 0;  // basis vectors already present (and their T block) at the start of a cycle
    // From the first restart on, every cycle issues the same launches on the same buffers (steps
    // j0 = nk .. m-1): it is captured once into a graph and replayed.
    CycleGraph graph(s);
    while (true) {
        int jend = j0;  // every cycle takes at least one step (j0 <= m - 2)
        auto steps = [&](int upto) {
            for (int j = j0; j < upto; ++j) {
                A.apply(W.Vi(j), W.Vi(j + 1));
                cgs2_dev(s, W, j + 1, W.Vi(j + 1), W.Ci(j));
            }
        };
        if (j0 > 0 && maxit - st.iters >= m - j0 && graph.run(j0, [&] { steps(m); })) {
            st.iters += m - j0;
            jend = m;
        } else {
            for (int j = j0; j < m; ++j) {
                A.apply(W.Vi(j), W.Vi(j + 1));
                ++st.iters;
                cgs2_dev(s, W, j + 1, W.Vi(j + 1), W.Ci(j));
                jend = j + 1;
                if (st.iters >= maxit) break;
            }
        }
        raw.resize((size_t)(jend - j0) * W.cstride);
        d2h(s, raw.data(), W.Ci(j0), raw.size());
        int k = j0;
        double beta = 0;
        for (int j = j0; j < jend; ++j) {
            beta = cgs2_coeffs(raw.data() + (size_t)(j - j0) * W.cstride, j + 1, h);
            for (int i = 0; i <= j; ++i) { T[i + (size_t)j * m] = h[i]; T[j + (size_t)i * m] = h[i]; }
            k = j + 1;
            if (beta <= 1e-14 * std::max(std::abs(h[j]), 1e-300) || j + 1 == jend) break;
            T[j + 1 + (size_t)j * m] = beta;
            T[j + (size_t)(j + 1) * m] = beta;
        }
        // Ritz pairs of the k x k projection (ascending)
        std::vector<double> Tk((size_t)k * k);
        for (int c = 0; c < k; ++c) for (int r = 0; r < k; ++r) Tk[r + (size_t)c * k] = T[r + (size_t)c * m];
        jacobi_eig(k, Tk, w, Z);
        std::vector<int> order(k);
        for (int i = 0; i < k; ++i) order[i] = i;
        std::sort(order.begin(), order.end(), [&](int a, int b) { return w[a] < w[b]; });
        const int imin = order[0];
        const double theta = w[imin];
        st.value = theta;
        st.relres = std::abs(beta * Z[(k - 1) + (size_t)imin * k]) / std::max(std::abs(theta), 1e-300);
        const bool done = st.relres <= tol || k < 2 || st.iters >= maxit || beta <= 1e-14 * std::max(std::abs(theta), 1e-300);
        const int nk = done ? 1 : std::min(keep, k - 1);
        // Ritz vectors Y_i = V[0..k) Z[:, order[i]] for the lowest nk (coefficients in one transfer)
        {
            std::vector<double> z((size_t)nk * k);
            for (int i = 0; i < nk; ++i)
                for (int r = 0; r < k; ++r) z[(size_t)i * k + r] = -Z[r + (size_t)order[i] * k];
            h2d(s, W.coef.d(), z.data(), z.size());
            for (int i = 0; i < nk; ++i) {
                double* yi = Y.d() + (size_t)i * n;
                hip_check(launch_zero(yi, n, s), "zero");  // fresh memory: not 0 * (possibly NaN) garbage
                hip_check(blas_gemv_sub(W.V.d(), n, k, W.coef.d() + (size_t)i * k, yi, n, s), "ritz");
            }
        }
        if (done) {
            d2d(s, x, Y.d(), n);
            nx = dev_norm(s, W, x);
            hip_check(blas_axpby(1.0 / nx, nullptr, AXPBY_DA, x, 0.0, x, n, s), "scale");
            st.converged = st.relres <= tol;
            break;
        }
        // new basis: the nk Ritz vectors and the next Lanczos vector (normalised by its step)
        d2d(s, W.Vi(nk), W.Vi(k), n);
        for (int i = 0; i < nk; ++i) d2d(s, W.Vi(i), Y.d() + (size_t)i * n, n);
        std::fill(T.begin(), T.end(), 0.0);
        for (int i = 0; i < nk; ++i) T[i + (size_t)i * m] = w[order[i]];
        j0 = nk;
    }
    // true residual ||A x - theta x|| / |theta|
    A.apply(x, W.w.d());
    hip_check(blas_axpby(-st.value, nullptr, AXPBY_DA, x, 1.0, W.w.d(), n, s), "resid");
    st.relres = dev_norm(s, W, W.w.d()) / std::max(std::abs(st.value), 1e-300);
    st.converged = st.relres <= std::max(tol, 1e-14) * 10;
    return st;
}

// ------------------------------------------------------------------------------------------------
namespace {
// BlockDiagMat layout of the left (qin) or right (qout) bond of an MPS tensor
Layout bond_layout(const Layout& A, bool right) {
    std::map<Key, int> m;
    for (auto& bl : A.l)
        for (auto& b : bl.b) {
            const Key& k = right ? b.qout : b.qin;
            const int sz = right ? b.cols : b.rows;
            auto it = m.find(k);
            if (it == m.end()) m.emplace(k, sz);
            else if (it->second != sz) throw InvalidArg("env: inconsistent sector sizes on a bond");
        }
    Layout L;
    set_group(L, A);
    L.l.resize(1);
    long pos = 0;
    for (auto& kv : m) {
        Blk b{kv.first, kv.first, kv.second, kv.second, pos};
        pos += b.size();
        L.l[0].b.push_back(b);
    }
    L.nelem = pos;
    return L;
}

std::unique_ptr<imps_plan_s> tm_plan(imps_context ctx, int dir, const Layout& A, const Layout& x, const Layout& y, int prec) {
    std::unique_ptr<imps_plan_s> P(new_plan(ctx, K_TM, prec));
    P->prog = build_tm(dir, A, nullptr, x, y);
    P->op_parts = {A.nelem};
    upload_program(P.get());
    return P;
}
}  // namespace

void EnvSolver::build(imps_context c, int direction, const std::vector<Layout>& A, const Layout& x, int precision) {
    if (!c) throw InvalidArg("env: a device context is required");
    if (A.empty()) throw InvalidArg("env: empty unit cell");
    if (x.l.size() != 1) throw InvalidArg("env: x must be a BlockDiagMat layout");
    ctx = c;
    dir = direction;
    prec = precision;
    ncell = (int)A.size();
    xlay = x;
    // application order and bond layouts
    std::vector<int> order(ncell);
    for (int k = 0; k < ncell; ++k) order[k] = dir == IMPS_LEFT ? k : ncell - 1 - k;
    Layout in = x;
    bond_elems.clear();
    for (int k = 0; k < ncell; ++k) {
        const Layout& Ak = A[order[k]];
        Layout out = (k == ncell - 1) ? x : bond_layout(Ak, dir == IMPS_LEFT);
        tm64.push_back(tm_plan(ctx, dir, Ak, in, out, IMPS_FP64));
        if (prec != IMPS_FP64) tmlo.push_back(tm_plan(ctx, dir, Ak, in, out, prec));
        if (k < ncell - 1) bond_elems.push_back(out.nelem);
        in = out;
    }
    long tot = 0;
    for (long e : bond_elems) tot += e;
    bonds.ensure(sizeof(double) * std::max<long>(tot, 1));
    rhoT.ensure(sizeof(double) * std::max<long>(x.nelem, 1));
    rhs.ensure(sizeof(double) * std::max<long>(x.nelem, 1));
    tmp.ensure(sizeof(double) * std::max<long>(x.nelem, 1));
}

void EnvSolver::set_mps(const std::vector<const double*>& A) {
    if ((int)A.size() != ncell) throw InvalidArg("env: wrong number of MPS tensors");
    for (int k = 0; k < ncell; ++k) {
        const int idx = dir == IMPS_LEFT ? k : ncell - 1 - k;
        set_operator(tm64[k].get(), {A[idx]});
        if (!tmlo.empty()) set_operator(tmlo[k].get(), {A[idx]});
    }
    mps_set = true;
}

KrylovStats EnvSolver::solve(const double* x, const double* L, const double* R, const double* y0, double* y, double tol, int maxit,
                             int krylov) {
    if (!mps_set) throw InvalidArg("env: MPS data not set");
    if (!x || !L || !R || !y) throw InvalidArg("env: NULL vector");
    hipStream_t s = ctx->stream;
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    const long n = xlay.nelem;
    // rho_in / rho_out: l -> (R, L), r -> (L, R).  tr(v rho_in) = <v, rho_in^T> with the blockwise transpose.
    const double* rho_in = dir == IMPS_LEFT ? R : L;
    const double* rho_out = dir == IMPS_LEFT ? L : R;
    {
        std::vector<double> h(n), t(n);
        d2h(s, h.data(), rho_in, n);
        for (auto& b : xlay.l[0].b)
            for (int j = 0; j < b.cols; ++j)
                for (int i = 0; i < b.rows; ++i) t[b.off + j + (long)i * b.rows] = h[b.off + i + (long)j * b.rows];
        h2d(s, rhoT.d(), t.data(), n);
    }
    PoolBuf dd;
    dd.ensure(sizeof(double) * (blas_scratch_doubles(1) + 4));
    auto trace_in = [&](const double* v) {
        double t = 0;
        hip_check(blas_multidot(rhoT.d(), 0, 1, v, n, dd.d(), dd.d() + 4, s), "trace");
        d2h(s, &t, dd.d(), 1);
        return t;
    };
    auto make_op = [&](std::vector<std::unique_ptr<imps_plan_s>>& chain) {
        LinOp op;
        op.n = n;
        op.apply = [&, n](const double* v, double* out) {
            // out = v - T(v) + tr(v rho_in) rho_out
            const double* in = v;
            long off = 0;
            for (int k = 0; k < ncell; ++k) {
                double* o = (k == ncell - 1) ? tmp.d() : bonds.d() + off;
                run(chain[k].get(), in, o, 1);
                in = o;
                if (k < ncell - 1) off += bond_elems[k];
            }
            // tr(v rho_in) stays on the device as chunk partials, summed by the combining kernel
            hip_check(blas_multidot_partials(rhoT.d(), 0, 1, v, n, dd.d() + 4, s), "trace");
            hip_check(blas_diff_plus_scaled(v, tmp.d(), rho_out, dd.d() + 4, blas_chunks(n), out, n, s), "1 - T + projector");
        };
        return op;
    };
    // right-hand side: x - tr(x rho_in) rho_out  (InvEBlockFunctions.hpp:243, :252)
    const double trx = trace_in(x);
    d2d(s, rhs.d(), x, n);
    hip_check(blas_axpby(-trx, nullptr, AXPBY_DA, rho_out, 1.0, rhs.d(), n, s), "rhs");
    if (y0) d2d(s, y, y0, n);
    else hip_check(launch_zero(y, n, s), "zero y");
    if (maxit <= 0) maxit = (int)std::min<long>(std::max<long>(20 * n, 1000), 100000);
    if (krylov <= 0) krylov = 100;
    LinOp A64 = make_op(tm64);
    KrylovStats st;
    if (prec == IMPS_FP64) st = gmres(s, A64, rhs.d(), y, tol, maxit, krylov);
    else {
        LinOp Alo = make_op(tmlo);
        st = gmres_refined(s, A64, Alo, rhs.d(), y, tol, maxit, krylov, prec == IMPS_FP32 ? 1e-4 : 1e-3);
    }
    hip_check(hipStreamSynchronize(s), "sync");
    return st;
}

void EnvSolver::tm_apply(const double* x, double* y) {
    if (!mps_set) throw InvalidArg("env: MPS data not set");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    const double* in = x;
    long off = 0;
    for (int k = 0; k < ncell; ++k) {
        double* o = (k == ncell - 1) ? y : bonds.d() + off;
        run(tm64[k].get(), in, o, 1);
        in = o;
        if (k < ncell - 1) off += bond_elems[k];
    }
}

int EnvSolver::fixed_point(double* x, double tol, int maxit, double* delta) {
    // The cell TM of an isometric chain preserves the trace (sum_s A_s^T A_s = 1 or sum_s A_s A_s^T = 1),
    // so for x0 of unit trace the fixed point is x0 + dx with [1 - T + |x0)(1|] dx = T(x0) - x0:
    // tracing the equation gives tr(dx) = 0 and then (1 - T) dx = T(x0) - x0.  Each round is one
    // GMRES solve of the environment form; a second round only polishes rounding.
    hipStream_t s = ctx->stream;
    const long n = xlay.nelem;
    PoolBuf id, dd, r0, dx;
    id.ensure(sizeof(double) * std::max<long>(n, 1));
    r0.ensure(sizeof(double) * std::max<long>(n, 1));
    dx.ensure(sizeof(double) * std::max<long>(n, 1));
    dd.ensure(sizeof(double) * (blas_scratch_doubles(1) + 4));
    {
        std::vector<double> h(n, 0.0);
        for (auto& b : xlay.l[0].b) for (int i = 0; i < std::min(b.rows, b.cols); ++i) h[b.off + i + (long)i * b.rows] = 1.0;
        h2d(s, id.d(), h.data(), n);
    }
    auto dotd = [&](const double* a, const double* b) {
        double t = 0;
        hip_check(blas_multidot(a, 0, 1, b, n, dd.d(), dd.d() + 4, s), "dot");
        d2h(s, &t, dd.d(), 1);
        return t;
    };
    int its = 0;
    double dl = 1;
    for (int round = 0; round < 6; ++round) {
        const double tr = dotd(id.d(), x);
        if (tr == 0) throw Numerical("fixed point: iterate has zero trace");
        hip_check(blas_axpby(1.0 / tr, nullptr, AXPBY_DA, x, 0.0, x, n, s), "normalise");
        tm_apply(x, r0.d());
        ++its;
        hip_check(blas_axpby(-1.0, nullptr, AXPBY_DA, x, 1.0, r0.d(), n, s), "T(x) - x");
        dl = std::sqrt(std::max(dotd(r0.d(), r0.d()), 0.0));
        if (dl < tol || its >= maxit) break;
        // l chain: rho_in = R slot, rho_out = L slot; r chain: rho_in = L, rho_out = R
        const double* Lp = dir == IMPS_LEFT ? x : id.d();
        const double* Rp = dir == IMPS_LEFT ? id.d() : x;
        KrylovStats st = solve(r0.d(), Lp, Rp, nullptr, dx.d(), std::max(0.1 * tol / std::max(dl, 1e-300), 1e-13), maxit - its, 200);
        its += st.iters;
        hip_check(blas_axpby(1.0, nullptr, AXPBY_DA, dx.d(), 1.0, x, n, s), "update");
    }
    if (delta) *delta = dl;
    return its;
}

}  // namespace imps
