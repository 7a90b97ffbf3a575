// FP64 CPU oracle for the block-sparse iMPS operator path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the smoke() entry and the
// cpu_baseline leg of bench.py load it (through liboracle.so).  The product path never links it.
//
// It restates, in plain C++ with std::map containers keyed by charge vectors, the algorithms of
// Darkdragon84/IMPS_CPP_tools (MPShelpers), so that the map/find/insert semantics of the reference
// (which blocks exist, which are matched, which are dropped) are reproduced literally.  Every
// function cites the reference file:line it follows.  Dense GEMM is a plain loop kernel, FP64.
//
// Parity pins (see tests/test_oracle.py): dense brute-force embedding at tiny D, closed-form
// bookkeeping known answers, isometry/gauge/Hermiticity/adjoint identities, GMRES residuals and
// the geometric-sum identity of Excitations_helpers.hpp:855-877.
//
// Symmetry groups: abelian, U(1) or Z(N) per charge (symobj.hpp:80-161).  The group is process
// state here (g_per, set by orc_set_periods and by the model entry points), as every key of the
// reference refers to its model's one group object (KeyTypes.hpp:19-58).

#include <algorithm>
#include <cassert>
#include <cmath>
#include <complex>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <cstdint>
#include <cstring>
#include <functional>
#include <map>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

namespace orc {

using Key = std::vector<int>;  // std::vector<int> lexicographic order = KeyType order (KeyTypes.hpp:19-58)

// charge periods of the current group: 0 = U(1), N > 1 = Z(N) (symobj::init's Plus_ / Minus_:
// the sum or difference taken modulo N into [0, N))
static std::vector<int> g_per;
static int kmod(int x, size_t i) {
    const int p = i < g_per.size() ? g_per[i] : 0;
    return p > 1 ? (x % p + p) % p : x;
}
static Key kadd(const Key& a, const Key& b) { Key c(a); for (size_t i = 0; i < c.size(); ++i) c[i] = kmod(c[i] + b[i], i); return c; }
static Key ksub(const Key& a, const Key& b) { Key c(a); for (size_t i = 0; i < c.size(); ++i) c[i] = kmod(c[i] - b[i], i); return c; }

// ---------------------------------------------------------------------------------------------
// dense column-major matrix
struct Mat {
    int r = 0, c = 0;
    std::vector<double> v;
    Mat() {}
    Mat(int r_, int c_) : r(r_), c(c_), v((size_t)r_ * c_, 0.0) {}
    double& operator()(int i, int j) { return v[(size_t)i + (size_t)j * r]; }
    double operator()(int i, int j) const { return v[(size_t)i + (size_t)j * r]; }
    Mat t() const { Mat o(c, r); for (int j = 0; j < c; ++j) for (int i = 0; i < r; ++i) o(j, i) = (*this)(i, j); return o; }
    Mat& operator+=(const Mat& o) { assert(o.r == r && o.c == c); for (size_t i = 0; i < v.size(); ++i) v[i] += o.v[i]; return *this; }
    Mat& operator-=(const Mat& o) { assert(o.r == r && o.c == c); for (size_t i = 0; i < v.size(); ++i) v[i] -= o.v[i]; return *this; }
    Mat& operator*=(double a) { for (auto& x : v) x *= a; return *this; }
};

// C = A*B (plain i-k-j loop order over column-major storage)
static Mat mm(const Mat& A, const Mat& B) {
    assert(A.c == B.r);
    Mat C(A.r, B.c);
    for (int j = 0; j < B.c; ++j) {
        double* cj = &C.v[(size_t)j * C.r];
        for (int k = 0; k < A.c; ++k) {
            const double b = B(k, j);
            const double* ak = &A.v[(size_t)k * A.r];
            for (int i = 0; i < A.r; ++i) cj[i] += ak[i] * b;
        }
    }
    return C;
}
static Mat operator*(double a, Mat m) { m *= a; return m; }
static Mat operator+(Mat a, const Mat& b) { a += b; return a; }
static double trace(const Mat& m) { double s = 0; for (int i = 0; i < std::min(m.r, m.c); ++i) s += m(i, i); return s; }

// ---------------------------------------------------------------------------------------------
// Block containers (BlockObj.hpp:480-600 BlockMat; :1359-1467 BlockDiagMat).
// A BlockMat maps Qin -> (Qout, Mat).  A BlockDiagMat is held as a BlockMat with Qin == Qout,
// which reproduces the BlockDiagMat find-by-key semantics since the key is then the single charge.
struct QM { Key qout; Mat m; };
using BM = std::map<Key, QM>;
using MPS = std::vector<BM>;  // MPSBlockMat: std::vector<BlockMat> of length d^N (MPSBlockMat.hpp:40-175)

// BlockMat::operator+= (BlockObj.hpp:1044-1063): add into matching Qin, insert otherwise.
static void addto(BM& self, const BM& other, double fac = 1.0) {
    for (const auto& o : other) {
        auto it = self.lower_bound(o.first);
        if (it != self.end() && it->first == o.first) {
            assert(it->second.qout == o.second.qout);
            if (fac == 1.0) it->second.m += o.second.m; else it->second.m += fac * o.second.m;
        } else {
            self.emplace_hint(it, o.first, QM{o.second.qout, fac == 1.0 ? o.second.m : fac * o.second.m});
        }
    }
}
static BM scaled(const BM& a, double f) { BM o(a); for (auto& b : o) b.second.m *= f; return o; }

// products of block matrices (BlockObj.hpp:2557-2625): match lhs Qout with rhs Qin.
static BM bmul(const BM& lhs, const BM& rhs) {
    BM out;
    for (const auto& l : lhs) {
        auto r = rhs.find(l.second.qout);
        if (r != rhs.end()) out.emplace_hint(out.end(), l.first, QM{r->second.qout, mm(l.second.m, r->second.m)});
    }
    return out;
}
// BlockMat::t(): (Qin->Qout, M) becomes (Qout->Qin, M^T)
static BM btrans(const BM& a) {
    BM out;
    for (const auto& b : a) out.emplace(b.second.qout, QM{b.first, b.second.m.t()});
    return out;
}
// BlockDiagMat(const BlockMat&) conversion keeps diagonal blocks only (BlockObj.hpp:1527-1536)
static BM diagpart(const BM& a) {
    BM out;
    for (const auto& b : a) if (b.first == b.second.qout) out.emplace_hint(out.end(), b.first, b.second);
    return out;
}
static double bnorm2(const BM& a) { double s = 0; for (auto& b : a) for (double x : b.second.m.v) s += x * x; return s; }

// dK of an MPS slice: Qout - Qin of any block (BlockMat::dK)
static bool sliceDK(const BM& a, Key& dk) {
    if (a.empty()) return false;
    dk = ksub(a.begin()->second.qout, a.begin()->first);
    return true;
}

// MPS products (MPSBlockMat.hpp:934-1043)
static MPS mpsmul(const MPS& l, const MPS& r) {  // concatenation, out[i*|r|+j] = l[i]*r[j]
    MPS out(l.size() * r.size());
    for (size_t i = 0; i < l.size(); ++i) for (size_t j = 0; j < r.size(); ++j) out[i * r.size() + j] = bmul(l[i], r[j]);
    return out;
}
static MPS matmps(const BM& m, const MPS& a) { MPS o(a.size()); for (size_t s = 0; s < a.size(); ++s) o[s] = bmul(m, a[s]); return o; }
static MPS mpsmat(const MPS& a, const BM& m) { MPS o(a.size()); for (size_t s = 0; s < a.size(); ++s) o[s] = bmul(a[s], m); return o; }

// ---------------------------------------------------------------------------------------------
// Transfer matrices (TMBlockFunctions.hpp)

// preallocated mixed left TM, BlockMat in/out (TMBlockFunctions.hpp:352-390):
// out[Qout(B_s(Qin x))] += B_s(Qin x)^T * x * A_s(Qout x)
static void tm_mixed_left_into(const MPS& A, const MPS& B, const BM& in, BM& out) {
    for (size_t s = 0; s < A.size(); ++s) {
        if (A[s].empty() || B[s].empty()) continue;
        for (const auto& x : in) {
            auto lit = B[s].find(x.first);
            auto rit = A[s].find(x.second.qout);
            if (lit != B[s].end() && rit != A[s].end()) {
                auto o = out.find(lit->second.qout);
                if (o != out.end()) {
                    assert(o->second.qout == rit->second.qout);
                    o->second.m += mm(mm(lit->second.m.t(), x.second.m), rit->second.m);
                }
            }
        }
    }
}
// preallocated mixed right TM (TMBlockFunctions.hpp:494-533):
// out[Qin(A_s(Qin x - dKA))] += A_s * x * B_s(Qout x - dKB)^T
static void tm_mixed_right_into(const MPS& A, const MPS& B, const BM& in, BM& out) {
    for (size_t s = 0; s < A.size(); ++s) {
        if (A[s].empty() || B[s].empty()) continue;
        Key dKA, dKB;
        sliceDK(A[s], dKA);
        sliceDK(B[s], dKB);
        for (const auto& x : in) {
            auto lit = A[s].find(ksub(x.first, dKA));
            auto rit = B[s].find(ksub(x.second.qout, dKB));
            if (lit != A[s].end() && rit != B[s].end()) {
                auto o = out.find(lit->first);
                if (o != out.end()) {
                    assert(o->second.qout == rit->first);
                    o->second.m += mm(mm(lit->second.m, x.second.m), rit->second.m.t());
                }
            }
        }
    }
}
// returning regular left TM on a BlockDiagMat (TMBlockFunctions.hpp:33-56)
static BM tm_left_ret(const MPS& A, const BM& in) {
    BM out;
    for (const auto& As : A)
        for (const auto& a : As) {
            auto x = in.find(a.first);
            if (x == in.end()) continue;
            Mat r = mm(a.second.m.t(), mm(x->second.m, a.second.m));
            auto o = out.lower_bound(a.second.qout);
            if (o != out.end() && o->first == a.second.qout) o->second.m += r;
            else out.emplace_hint(o, a.second.qout, QM{a.second.qout, r});
        }
    return out;
}
// returning regular right TM on a BlockDiagMat (TMBlockFunctions.hpp:155-181)
static BM tm_right_ret(const MPS& A, const BM& in) {
    BM out;
    for (const auto& As : A)
        for (const auto& a : As) {
            auto x = in.find(a.second.qout);
            if (x == in.end()) continue;
            Mat r = mm(mm(a.second.m, x->second.m), a.second.m.t());
            auto o = out.lower_bound(a.first);
            if (o != out.end() && o->first == a.first) o->second.m += r;
            else out.emplace_hint(o, a.first, QM{a.first, r});
        }
    return out;
}
// ApplyTMmixedLeftDiag(A,B): sum_s B_s^T A_s keyed by Qout(B_s) (TMBlockFunctions.hpp:259-288)
static BM tm_mixed_left_diag_unity(const MPS& A, const MPS& B) {
    BM out;
    for (size_t s = 0; s < A.size(); ++s) {
        if (A[s].empty() || B[s].empty()) continue;
        for (const auto& a : A[s]) {
            auto b = B[s].find(a.first);
            if (b == B[s].end()) continue;
            Mat r = mm(b->second.m.t(), a.second.m);
            const Key& k = b->second.qout;
            auto o = out.lower_bound(k);
            if (o != out.end() && o->first == k) o->second.m += r;
            else out.emplace_hint(o, k, QM{k, r});
        }
    }
    return out;
}
// ApplyTMmixedRightDiag(A,B): sum_s A_s B_s^T keyed by Qin(A_s) (TMBlockFunctions.hpp:401-427)
static BM tm_mixed_right_diag_unity(const MPS& A, const MPS& B) {
    BM out;
    for (size_t s = 0; s < A.size(); ++s) {
        if (A[s].empty() || B[s].empty()) continue;
        for (const auto& a : A[s]) {
            auto b = B[s].find(a.first);
            if (b == B[s].end()) continue;
            Mat r = mm(a.second.m, b->second.m.t());
            const Key& k = a.first;
            auto o = out.lower_bound(k);
            if (o != out.end() && o->first == k) o->second.m += r;
            else out.emplace_hint(o, k, QM{k, r});
        }
    }
    return out;
}
// ApplyTMmixedRight(A,B) unity -> BlockMat: out[Qin A_s] = (Qin B_s, A_s B_s^T) (TMBlockFunctions.hpp:444-470)
static BM tm_mixed_right_unity(const MPS& A, const MPS& B) {
    BM out;
    for (size_t s = 0; s < A.size(); ++s) {
        if (A[s].empty() || B[s].empty()) continue;
        Key dKB;
        sliceDK(B[s], dKB);
        for (const auto& a : A[s]) {
            auto b = B[s].find(ksub(a.second.qout, dKB));
            if (b == B[s].end()) continue;
            Mat r = mm(a.second.m, b->second.m.t());
            auto o = out.lower_bound(a.first);
            if (o != out.end() && o->first == a.first && o->second.qout == b->first) o->second.m += r;
            else out.emplace_hint(o, a.first, QM{b->first, r});
        }
    }
    return out;
}
// ApplyTMmixedLeft(A,B) unity -> BlockMat: out[Qout B_s] = (Qout A_s, B_s^T A_s) (TMBlockFunctions.hpp:310-333)
static BM tm_mixed_left_unity(const MPS& A, const MPS& B) {
    BM out;
    for (size_t s = 0; s < A.size(); ++s) {
        if (A[s].empty() || B[s].empty()) continue;
        for (const auto& a : A[s]) {
            auto b = B[s].find(a.first);
            if (b == B[s].end()) continue;
            Mat r = mm(b->second.m.t(), a.second.m);
            const Key& k = b->second.qout;
            auto o = out.lower_bound(k);
            if (o != out.end() && o->first == k && o->second.qout == a.second.qout) o->second.m += r;
            else out.emplace_hint(o, k, QM{a.second.qout, r});
        }
    }
    return out;
}

// ---------------------------------------------------------------------------------------------
// two-site operators (MPSBlockUtilities.hpp:18-74; OperatorTypes.hpp; helpers.cpp:3-17)
// H is a dense (d^2 x d^2) column-major matrix; its nonzeros are visited column by column, which
// is the order in which an Armadillo SpMat iterator walks them.
struct Nz { int row, col; double val; };
static std::vector<Nz> nonzeros(const double* H, int D) {
    std::vector<Nz> nz;
    for (int j = 0; j < D; ++j) for (int i = 0; i < D; ++i) if (H[i + (size_t)j * D] != 0.0) nz.push_back({i, j, H[i + (size_t)j * D]});
    return nz;
}
static MPS apply_operator(const MPS& in, const double* H, int D) {  // ApplyOperator :18-34
    MPS out(in.size());
    for (const auto& z : nonzeros(H, D)) addto(out[z.row], scaled(in[z.col], z.val));
    return out;
}
static std::vector<BM> reduced_op(const double* H, int d, const MPS& A, int dir) {  // ReducedOp :39-74
    std::vector<BM> R((size_t)d * d);
    std::vector<bool> filled((size_t)d * d, false);
    for (const auto& z : nonzeros(H, d * d)) {
        int i0 = z.row / d, i1 = z.row % d, j0 = z.col / d, j1 = z.col % d;
        if (dir == 1) {  // r: Ored[i0][j0] += h * A[j1] * A[i1]^T
            BM t = bmul(A[j1], btrans(A[i1]));
            size_t idx = (size_t)i0 * d + j0;
            if (!filled[idx] && R[idx].empty()) { R[idx] = scaled(t, z.val); } else addto(R[idx], scaled(t, z.val));
            filled[idx] = true;
        } else {         // l: Ored[i1][j1] += h * A[i0]^T * A[j0]
            BM t = bmul(btrans(A[i0]), A[j0]);
            size_t idx = (size_t)i1 * d + j1;
            if (!filled[idx] && R[idx].empty()) { R[idx] = scaled(t, z.val); } else addto(R[idx], scaled(t, z.val));
            filled[idx] = true;
        }
    }
    return R;
}
// ApplyOpTMLeftDiag / RightDiag (TMBlockFunctions.hpp:797-864)
static BM op_tm_left_diag(const double* H, int d, const MPS& AA) { return tm_mixed_left_diag_unity(apply_operator(AA, H, d * d), AA); }
static BM op_tm_right_diag(const double* H, int d, const MPS& AA) { return tm_mixed_right_diag_unity(apply_operator(AA, H, d * d), AA); }

// BlockDiagMat sum that keeps the union (BlockDiagMat::operator+= BlockObj.hpp:1890-1906)
static BM bdsum(const BM& a, const BM& b) { BM o(a); addto(o, b); return o; }

// GetHL / GetHR (VUMPS_helpers.hpp:3-36)
static BM get_hl(const std::vector<MPS>& AL, const double* H, int d) {
    BM HL = op_tm_left_diag(H, d, mpsmul(AL.back(), AL.front()));
    for (size_t k = 1; k < AL.size(); ++k) HL = bdsum(tm_left_ret(AL[k], HL), op_tm_left_diag(H, d, mpsmul(AL[k - 1], AL[k])));
    return HL;
}
static BM get_hr(const std::vector<MPS>& AR, const double* H, int d) {
    BM HR = op_tm_right_diag(H, d, mpsmul(AR.back(), AR.front()));
    for (int k = (int)AR.size() - 2; k >= 0; --k) HR = bdsum(tm_right_ret(AR[k], HR), op_tm_right_diag(H, d, mpsmul(AR[k], AR[k + 1])));
    return HR;
}

// ---------------------------------------------------------------------------------------------
// H_AC and H_C (VUMPS_helpers.hpp:38-133).  `out` is a preallocated structure (zeroed here, as
// ApplyHAvec/ApplyHCvec memset it); contributions for blocks outside it are dropped, which is what
// the reference's += into a map of buffer views amounts to.
static void add_projected(BM& out, const BM& contrib) {
    for (const auto& c : contrib) {
        auto o = out.find(c.first);
        if (o != out.end()) { assert(o->second.qout == c.second.qout); o->second.m += c.second.m; }
    }
}
static void zero(BM& a) { for (auto& b : a) std::fill(b.second.m.v.begin(), b.second.m.v.end(), 0.0); }
static void apply_hac(const MPS& in, MPS& out, const BM& EHL, const BM& EHR, const std::vector<BM>& HL, const std::vector<BM>& HR) {
    const size_t d = in.size();
    for (auto& o : out) zero(o);
    // out += EHL*in + in*EHR
    for (size_t s = 0; s < d; ++s) {
        BM t = bmul(EHL, in[s]);
        addto(t, bmul(in[s], EHR));
        add_projected(out[s], t);
    }
    for (size_t i = 0; i < d; ++i)
        for (size_t j = 0; j < d; ++j) {
            BM t = bmul(HL[i * d + j], in[j]);
            addto(t, bmul(in[j], HR[i * d + j]));
            add_projected(out[i], t);
        }
}
static void apply_hc(const BM& in, BM& out, const BM& EHL, const BM& EHR, const double* H, int d, const MPS& AL, const MPS& AR) {
    zero(out);
    BM t = bmul(EHL, in);
    addto(t, bmul(in, EHR));
    add_projected(out, diagpart(t));
    for (const auto& z : nonzeros(H, d * d)) {
        int i0 = z.row / d, i1 = z.row % d, j0 = z.col / d, j1 = z.col % d;
        // (*opit)*AL[i0].t()*AL[j0]*in*AR[j1]*AR[i1].t(), evaluated left to right
        BM c = scaled(btrans(AL[i0]), z.val);
        c = bmul(c, AL[j0]);
        c = bmul(c, in);
        c = bmul(c, AR[j1]);
        c = bmul(c, btrans(AR[i1]));
        add_projected(out, diagpart(c));
    }
}

// ---------------------------------------------------------------------------------------------
// GMRES (IterativeSolvers.hpp:21-149), real scalars.  Unrestarted, modified Gram-Schmidt with
// Givens rotations, stop at |r| < tol*|b|.  Returns the number of Krylov steps or -1/-2.
using Vec = std::vector<double>;
static double vnorm(const Vec& x) { double s = 0; for (double v : x) s += v * v; return std::sqrt(s); }
static int gmres(const std::function<Vec(const Vec&)>& Afun, const Vec& b, Vec& x, const Vec& x0, double tol, unsigned maxit, unsigned krylovdim = 0) {
    int flag = -1;
    const unsigned dim = (unsigned)b.size();
    if (maxit > dim) maxit = dim;
    if (krylovdim == 0) krylovdim = std::min(dim, maxit);
    if (krylovdim > dim) krylovdim = dim;
    Vec Ax0 = Afun(x0), r(dim);
    for (unsigned i = 0; i < dim; ++i) r[i] = b[i] - Ax0[i];
    x = x0;
    const double bnorm = vnorm(b), thresh = 2e-16;
    double rnorm = vnorm(r);
    std::vector<Vec> V;
    V.reserve(krylovdim + 1);
    std::vector<Vec> Hc;  // columns of H
    Vec c(krylovdim + 1), s(krylovdim + 1), gamma(krylovdim + 1);
    gamma[0] = rnorm;
    V.push_back(r);
    for (auto& v : V[0]) v /= rnorm;
    unsigned kdim = 0;
    for (unsigned j = 0; j < krylovdim; ++j) {
        Vec w = Afun(V[j]);
        Vec h(j + 2, 0.0);
        for (unsigned i = 0; i <= j; ++i) {
            double dot = 0;
            for (unsigned k = 0; k < dim; ++k) dot += V[i][k] * w[k];
            h[i] = dot;
            for (unsigned k = 0; k < dim; ++k) w[k] -= dot * V[i][k];
        }
        h[j + 1] = vnorm(w);
        for (unsigned i = 0; i < j; ++i) {
            double t1 = h[i], t2 = h[i + 1];
            h[i] = c[i] * t1 + s[i] * t2;
            h[i + 1] = -s[i] * t1 + c[i] * t2;
        }
        double beta = std::sqrt(h[j] * h[j] + h[j + 1] * h[j + 1]);
        if (std::abs(beta) < thresh) { flag = -2; break; }
        Vec vn(w);
        for (auto& v : vn) v /= h[j + 1];
        V.push_back(vn);
        double habs = std::abs(h[j]);
        double phase = habs == 0 ? 1.0 : h[j] / habs;
        s[j] = phase * h[j + 1] / beta;
        c[j] = habs / beta;
        h[j] = phase * beta;
        h[j + 1] = 0;
        gamma[j + 1] = -s[j] * gamma[j];
        gamma[j] *= c[j];
        ++kdim;
        Hc.push_back(h);
        rnorm = std::abs(gamma[j + 1]);
        if (rnorm < tol * bnorm) { flag = (int)kdim; break; }
    }
    // back substitution R alpha = gamma
    Vec alpha(kdim, 0.0);
    for (int i = (int)kdim - 1; i >= 0; --i) {
        double t = gamma[i];
        for (unsigned k = i + 1; k < kdim; ++k) t -= Hc[k][i] * alpha[k];
        alpha[i] = t / Hc[i][i];
    }
    for (unsigned k = 0; k < kdim; ++k) for (unsigned i = 0; i < dim; ++i) x[i] += alpha[k] * V[k][i];
    return flag;
}

// vectorization in reference order (BlockObj.hpp:764-780, MPSBlockMat.hpp:446-465)
static Vec vectorize(const BM& a) { Vec o; for (auto& b : a) o.insert(o.end(), b.second.m.v.begin(), b.second.m.v.end()); return o; }
static void devectorize(const Vec& v, BM& a) { size_t p = 0; for (auto& b : a) { std::copy(v.begin() + p, v.begin() + p + b.second.m.v.size(), b.second.m.v.begin()); p += b.second.m.v.size(); } }

// chained regular TM over a unit cell, preallocated output (TMBlockFunctions.hpp:565-615)
static void tm_chain_into(const std::vector<MPS>& Avec, int dir, const BM& in, BM& out) {
    zero(out);
    if (dir == 0) {
        if (Avec.size() > 1) {
            BM tmp = tm_left_ret(Avec[0], in);
            for (size_t k = 1; k + 1 < Avec.size(); ++k) tmp = tm_left_ret(Avec[k], tmp);
            tm_mixed_left_into(Avec.back(), Avec.back(), tmp, out);
        } else tm_mixed_left_into(Avec[0], Avec[0], in, out);
    } else {
        if (Avec.size() > 1) {
            BM tmp = tm_right_ret(Avec.back(), in);
            for (int k = (int)Avec.size() - 2; k >= 1; --k) tmp = tm_right_ret(Avec[k], tmp);
            tm_mixed_right_into(Avec[0], Avec[0], tmp, out);
        } else tm_mixed_right_into(Avec[0], Avec[0], in, out);
    }
}
static double trace_prod(const BM& a, const BM& b) {  // sum over matched blocks of trace(a_q b_q)
    double s = 0;
    for (auto& x : a) { auto y = b.find(x.second.qout); if (y != b.end()) s += trace(mm(x.second.m, y->second.m)); }
    return s;
}
// InvertE_proj (InvEBlockFunctions.hpp:205-266) with InvEfun_proj (:13-47), BlockDiagMat form.
// dir 0 (l): (y|[1 - T + |R)(L|] = (x|[1 - |R)(L|];  dir 1 (r): mirrored.
static BM invert_e_proj(const std::vector<MPS>& Avec, const BM& x, const BM& L, const BM& R, int dir, double tol, unsigned maxit, const BM* y0, int* flag_out) {
    const size_t m_tot = vectorize(x).size();
    if (maxit == 0) maxit = (unsigned)m_tot;
    Vec xpv = vectorize(x);
    const BM& rho_in = dir == 0 ? R : L;
    const BM& rho_out = dir == 0 ? L : R;
    {   // xpv -= trace(x*R)*L  (l)  or  trace(L*x)*R  (r)
        double tr = dir == 0 ? trace_prod(x, R) : trace_prod(L, x);
        Vec lv = vectorize(rho_out);
        for (size_t i = 0; i < m_tot; ++i) xpv[i] -= tr * lv[i];
    }
    BM work = x;
    auto Afun = [&](const Vec& v) -> Vec {
        BM in = x;
        devectorize(v, in);
        double tr = 0;
        {   // trace(x * rho_in) walking both maps in order (InvEBlockFunctions.hpp:27-32)
            auto rit = rho_in.begin();
            for (auto it = in.begin(); it != in.end(); ++it, ++rit) tr += trace(mm(it->second.m, rit->second.m));
        }
        BM Tx = x;
        tm_chain_into(Avec, dir, in, Tx);
        Vec out(v);
        Vec tv = vectorize(Tx), rv = vectorize(rho_out);
        for (size_t i = 0; i < out.size(); ++i) out[i] += tr * rv[i] - tv[i];
        return out;
    };
    Vec y(m_tot), yinit = y0 ? vectorize(*y0) : Vec(m_tot, 0.0);
    if (!y0) {  // reference initializes a random y0 (InvEBlockFunctions.hpp:232-236); use a fixed seed here
        std::mt19937_64 g(12345);
        std::normal_distribution<double> nd;
        for (auto& v : yinit) v = nd(g);
    }
    int flag = gmres(Afun, xpv, y, yinit, tol, maxit, 0);
    if (flag_out) *flag_out = flag;
    BM out = x;
    devectorize(y, out);
    return out;
}

// ---------------------------------------------------------------------------------------------
// QR with non-negative R diagonal (helpers.hpp:181-195) by Householder reflections; economy size.
static void qr_pos(const Mat& X, Mat& Q, Mat& R) {
    const int m = X.r, n = X.c, k = std::min(m, n);
    Mat A = X;
    std::vector<Vec> vs;
    Vec betas;
    for (int j = 0; j < k; ++j) {
        double nrm = 0;
        for (int i = j; i < m; ++i) nrm += A(i, j) * A(i, j);
        nrm = std::sqrt(nrm);
        Vec v(m - j, 0.0);
        double beta = 0;
        if (nrm > 0) {
            double alpha = A(j, j) >= 0 ? -nrm : nrm;
            for (int i = j; i < m; ++i) v[i - j] = A(i, j);
            v[0] -= alpha;
            double vn = 0;
            for (double t : v) vn += t * t;
            if (vn > 0) {
                beta = 2.0 / vn;
                for (int c = j; c < n; ++c) {
                    double dot = 0;
                    for (int i = j; i < m; ++i) dot += v[i - j] * A(i, c);
                    for (int i = j; i < m; ++i) A(i, c) -= beta * dot * v[i - j];
                }
            }
        }
        vs.push_back(v);
        betas.push_back(beta);
    }
    R = Mat(k, n);
    for (int j = 0; j < n; ++j) for (int i = 0; i <= std::min(j, k - 1); ++i) R(i, j) = A(i, j);
    Q = Mat(m, k);
    for (int i = 0; i < k; ++i) Q(i, i) = 1.0;
    for (int j = k - 1; j >= 0; --j) {
        const Vec& v = vs[j];
        if (betas[j] == 0) continue;
        for (int c = 0; c < k; ++c) {
            double dot = 0;
            for (int i = j; i < m; ++i) dot += v[i - j] * Q(i, c);
            for (int i = j; i < m; ++i) Q(i, c) -= betas[j] * dot * v[i - j];
        }
    }
    for (int i = 0; i < k; ++i)
        if (R(i, i) < 0) {
            for (int r = 0; r < m; ++r) Q(r, i) = -Q(r, i);
            for (int c = 0; c < n; ++c) R(i, c) = -R(i, c);
        }
}
// full orthonormal complement of the column space of X (m x n, m > n): m x (m - rank) basis
static Mat complement(const Mat& X) {
    const int m = X.r;
    // orthonormalise [X | I] column by column (two passes of Gram-Schmidt) and keep the new directions
    std::vector<Vec> basis;
    auto try_add = [&](Vec v, double tolrel) {
        double n0 = vnorm(v);
        if (n0 == 0) return false;
        for (int pass = 0; pass < 2; ++pass)
            for (auto& b : basis) { double d = 0; for (int i = 0; i < m; ++i) d += b[i] * v[i]; for (int i = 0; i < m; ++i) v[i] -= d * b[i]; }
        double n1 = vnorm(v);
        if (n1 <= tolrel * n0) return false;
        for (auto& t : v) t /= n1;
        basis.push_back(v);
        return true;
    };
    double scale = 0;
    for (double t : X.v) scale = std::max(scale, std::abs(t));
    for (int j = 0; j < X.c; ++j) { Vec v(m); for (int i = 0; i < m; ++i) v[i] = X(i, j); try_add(v, 1e-12); }
    const size_t rank = basis.size();
    for (int e = 0; e < m && (int)basis.size() < m; ++e) { Vec v(m, 0.0); v[e] = 1.0; try_add(v, 1e-8); }
    Mat N(m, (int)(basis.size() - rank));
    for (size_t k = rank; k < basis.size(); ++k) for (int i = 0; i < m; ++i) N(i, (int)(k - rank)) = basis[k][i];
    return N;
}

// dense regrouping of an MPS slice per charge sector (MPSBlockMat.hpp:468-526, 636-719)
struct SecPart { int s; Key other; int off, size; };
struct Sector { Key q; int ml = 0, mr = 0; std::vector<SecPart> parts; Mat dense; };
static std::vector<Sector> dense_sectors(const MPS& A, int dir) {
    std::map<Key, Sector> secs;
    for (size_t s = 0; s < A.size(); ++s)
        for (const auto& b : A[s]) {
            const Key& q = dir == 0 ? b.second.qout : b.first;
            Sector& sec = secs[q];
            sec.q = q;
            if (dir == 0) { sec.mr = b.second.m.c; sec.parts.push_back({(int)s, b.first, 0, b.second.m.r}); }
            else { sec.ml = b.second.m.r; sec.parts.push_back({(int)s, b.second.qout, 0, b.second.m.c}); }
        }
    std::vector<Sector> out;
    for (auto& kv : secs) {
        Sector sec = kv.second;
        int tot = 0;
        for (auto& p : sec.parts) { p.off = tot; tot += p.size; }  // parts are in ascending s
        if (dir == 0) sec.ml = tot; else sec.mr = tot;
        sec.dense = Mat(sec.ml, sec.mr);
        for (auto& p : sec.parts) {
            const BM& As = A[p.s];
            const Mat& m = dir == 0 ? As.at(p.other).m : As.at(sec.q).m;
            for (int j = 0; j < m.c; ++j) for (int i = 0; i < m.r; ++i) {
                if (dir == 0) sec.dense(p.off + i, j) = m(i, j); else sec.dense(i, p.off + j) = m(i, j);
            }
        }
        out.push_back(sec);
    }
    return out;
}
// qr(R, MPS, dir) (MPSBlockUtilities.hpp:181-244)
static MPS mps_qr(const MPS& A, int dir, BM& R) {
    MPS Q(A.size());
    R.clear();
    for (auto& sec : dense_sectors(A, dir)) {
        Mat Qm, Rm;
        if (dir == 0) {
            qr_pos(sec.dense, Qm, Rm);
            R.emplace(sec.q, QM{sec.q, Rm});
            for (auto& p : sec.parts) {
                Mat blk(p.size, Qm.c);
                for (int j = 0; j < Qm.c; ++j) for (int i = 0; i < p.size; ++i) blk(i, j) = Qm(p.off + i, j);
                Q[p.s].emplace(p.other, QM{sec.q, blk});
            }
        } else {
            qr_pos(sec.dense.t(), Qm, Rm);
            R.emplace(sec.q, QM{sec.q, Rm.t()});
            for (auto& p : sec.parts) {
                Mat blk(Qm.c, p.size);  // Qmat(span,all).t()
                for (int j = 0; j < p.size; ++j) for (int i = 0; i < Qm.c; ++i) blk(i, j) = Qm(p.off + j, i);
                Q[p.s].emplace(sec.q, QM{p.other, blk});
            }
        }
    }
    return Q;
}
// Nullspace(MPS, dir) (MPSBlockUtilities.hpp:247-340).  The basis spans the same space as
// Armadillo's null() (an SVD basis); it differs from it by an orthogonal rotation, under which
// every quantity the hot path takes from it (norms of contractions with it) is invariant.
static MPS mps_nullspace(const MPS& A, int dir) {
    MPS NS(A.size());
    for (auto& sec : dense_sectors(A, dir)) {
        if (dir == 0) {
            if (sec.ml <= sec.mr) continue;
            Mat B = complement(sec.dense);
            if (B.c == 0) continue;
            for (auto& p : sec.parts) {
                Mat blk(p.size, B.c);
                for (int j = 0; j < B.c; ++j) for (int i = 0; i < p.size; ++i) blk(i, j) = B(p.off + i, j);
                NS[p.s].emplace(p.other, QM{sec.q, blk});
            }
        } else {
            if (sec.mr <= sec.ml) continue;
            Mat B = complement(sec.dense.t());
            if (B.c == 0) continue;
            for (auto& p : sec.parts) {
                Mat blk(B.c, p.size);
                for (int j = 0; j < p.size; ++j) for (int i = 0; i < B.c; ++i) blk(i, j) = B(p.off + j, i);
                NS[p.s].emplace(sec.q, QM{p.other, blk});
            }
        }
    }
    return NS;
}

// ---------------------------------------------------------------------------------------------
// bookkeeping (Models.cpp; ItoKey.hpp)
static int igcd(int u, int v) { u = std::abs(u); v = std::abs(v); while (v != 0) { int r = u % v; u = v; v = r; } return u; }
// model ids: type | (d << 8); type 0 = XXZ (d=2, one U(1)), 1 = FHUB in (N,M) representation
// (d=4, U(1)xU(1)), 2 = quantum Ising (d=2, Z(2)), 3 = quantum Potts (d=q, Z(q)), 4 = Bose-Hubbard
// (local dimension d, U(1))
static int mtype(int model) { return model & 0xff; }
static int model_d(int model) {
    const int t = mtype(model), d = model >> 8;
    if (t == 0 || t == 2) return (d == 0 || d == 2) ? 2 : -1;
    if (t == 1) return (d == 0 || d == 4) ? 4 : -1;
    if (t == 3 || t == 4) return (d >= 2 && d <= 16) ? d : -1;
    return -1;
}
static int model_nsym(int model) { return model_d(model) < 0 ? -1 : (mtype(model) == 1 ? 2 : 1); }
// the model's group (GroupObj_.init in each Init: Models.cpp:200, 335, 393, 521, 749) made current
static void use_group(int model) {
    const int t = mtype(model), d = model_d(model);
    if (d < 0) throw std::domain_error("unknown model");
    g_per = t == 1 ? std::vector<int>{0, 0} : t == 2 ? std::vector<int>{2} : t == 3 ? std::vector<int>{d} : std::vector<int>{0};
}
static std::vector<std::pair<int, int>> getdQ(int model, int N, const int* QN) {
    std::vector<std::pair<int, int>> dQ;
    const int t = mtype(model);
    if (t == 2 || t == 3) {  // QuantumIsingModel::GetdQ :351-355, QuantumPottsModel::GetdQ :472-476
        dQ.push_back({1, 0});
    } else if (t == 4) {     // BoseHubbardModel::GetdQ :830-843
        const int n = QN[0];
        if (n < 1 || n >= N * model_d(model)) throw std::domain_error("0 < n < N d");
        const int div = igcd(N, n);
        dQ.push_back({N / div, -n / div});
    } else if (t == 0) {  // Models.cpp:216-240
        int m = QN[0];
        if (std::abs(m) >= N) throw std::domain_error("|m| < N");
        int div = igcd(N - m, N + m);
        dQ.push_back({2 * N / div, -(N + m) / div});
    } else {           // Models.cpp:642-662
        if (t != 1) throw std::domain_error("unknown model");
        int n = QN[0], m = QN[1];
        if (n + m <= 0 || n + m >= 2 * N) throw std::domain_error("n range");
        if (n - m <= 0 || n - m >= 2 * N) throw std::domain_error("n range");
        int divn = igcd(igcd(N - n, 2 * N - n), n);
        int divm = igcd(igcd(N - m, N + m), m);
        dQ.push_back({N / divn, -n / divn});
        dQ.push_back({N / divm, -(N + m) / divm});
    }
    return dQ;
}
static std::vector<Key> get_i2k(int model, int N, const int* QN) {
    auto dQ = getdQ(model, N, QN);
    std::vector<Key> I2K;
    const int t = mtype(model), d = model_d(model);
    if (t == 0) {  // Models.cpp:270-277: s = d-1 ... 0 pushed in that order
        for (int s = 1; s >= 0; --s) I2K.push_back({dQ[0].second + s * dQ[0].first});
    } else if (t == 2) {  // QuantumIsingModel::GetI2K :372-376: {0}, {1}
        I2K.push_back({0});
        I2K.push_back({1});
    } else if (t == 3) {  // QuantumPottsModel::GetI2K :500-506: {i}
        for (int i = 0; i < d; ++i) I2K.push_back({i});
    } else if (t == 4) {  // BoseHubbardModel::GetI2K :862-878: {dQ.second + i dQ.first}
        for (int i = 0; i < d; ++i) I2K.push_back({dQ[0].second + i * dQ[0].first});
    } else {           // Models.cpp:712-718
        I2K.push_back({dQ[0].second, dQ[1].second + dQ[1].first});
        I2K.push_back({dQ[0].second + dQ[0].first, dQ[1].second});
        I2K.push_back({dQ[0].second + dQ[0].first, dQ[1].second + 2 * dQ[1].first});
        I2K.push_back({dQ[0].second + 2 * dQ[0].first, dQ[1].second + dQ[1].first});
    }
    return I2K;
}
static std::map<Key, int> get_dimmap(int model, const std::vector<int>& dims, int N, const int* QN) {
    auto dQ = getdQ(model, N, QN);
    int Nsec = (int)dims.size();
    std::map<Key, int> dm;
    const int t = mtype(model), d = model_d(model);
    if (t == 2 || t == 3) {  // QIM :357-370 (2 sectors), QPOTTS :478-498 (q sectors); extra dims dropped
        if (Nsec < d) throw std::domain_error("not enough symmetry sectors");
        for (int i = 0; i < d; ++i) dm[{i}] = dims[i];
    } else if (t == 4) {  // BHUB :845-860
        for (int i = 0; i < Nsec; ++i) dm[{i * dQ[0].first}] = dims[i];
    } else if (t == 0) {  // Models.cpp:248-260
        for (int i = 0; i < Nsec; ++i) dm[{dQ[0].first * (i - Nsec / 2)}] = dims[i];
    } else {           // Models.cpp:680-697 (N,M) representation
        for (int i = 0; i < Nsec; ++i)
            for (int j = 0; j < Nsec; ++j) dm[{dQ[0].first * (i + j - Nsec + 1), dQ[1].first * (i - j)}] = dims[i] + dims[j];
    }
    return dm;
}
// path_iter / validpath (ItoKey.hpp:126-165)
static bool path_iter(int N, const Key& in, const std::vector<Key>& I2K, const Key& zero, std::vector<int>& path) {
    if (N == 1) {
        for (size_t j = 0; j < I2K.size(); ++j) if (kadd(in, I2K[j]) == zero) { path = {(int)j}; return true; }
        return false;
    }
    for (size_t j = 0; j < I2K.size(); ++j) {
        std::vector<int> p;
        if (path_iter(N - 1, kadd(in, I2K[j]), I2K, zero, p)) { p.push_back((int)j); path = p; return true; }
    }
    return false;
}
static std::vector<int> validpath(int N, const std::vector<Key>& I2K) {
    std::vector<int> p;
    if (!path_iter(N, I2K.front(), I2K, I2K.front(), p)) throw std::domain_error("validpath: no valid path");
    return p;
}
// MakeDims (Models.cpp:13-32): bond k+1 = bond k shifted by I2K[path[N-k-1]]
static std::vector<std::map<Key, int>> make_dims(const std::map<Key, int>& dm0, int N, const std::vector<Key>& I2K) {
    auto path = validpath(N, I2K);
    std::vector<std::map<Key, int>> dv;
    std::map<Key, int> l = dm0;
    for (int k = 0; k < N; ++k) {
        dv.push_back(l);
        std::map<Key, int> r;
        for (auto& kv : l) r[kadd(kv.first, I2K[path[N - k - 1]])] = kv.second;
        l = r;
    }
    return dv;
}
// MPSBlockMat(I2K, ml, mr, fill) structure (MPSBlockMat.hpp:221-236)
static MPS mps_structure(const std::vector<Key>& I2K, const std::map<Key, int>& ml, const std::map<Key, int>& mr) {
    MPS A(I2K.size());
    for (size_t s = 0; s < I2K.size(); ++s)
        for (auto& l : ml) {
            auto r = mr.find(kadd(l.first, I2K[s]));
            if (r != mr.end()) A[s].emplace_hint(A[s].end(), l.first, QM{r->first, Mat(l.second, r->second)});
        }
    return A;
}
// local two-site Hamiltonians (Models.cpp:57-93 XXZ; :543-572 FHUB), dense d^2 x d^2 column-major
static std::vector<double> kron(const std::vector<double>& A, const std::vector<double>& B, int d) {
    const int D = d * d;
    std::vector<double> K((size_t)D * D, 0.0);
    for (int i = 0; i < d; ++i) for (int k = 0; k < d; ++k) for (int j = 0; j < d; ++j) for (int l = 0; l < d; ++l)
        K[(i * d + j) + (size_t)(k * d + l) * D] = A[i + k * d] * B[j + l * d];
    return K;
}
static std::vector<double> dmul(const std::vector<double>& A, const std::vector<double>& B, int d) {
    std::vector<double> C((size_t)d * d, 0.0);
    for (int j = 0; j < d; ++j) for (int k = 0; k < d; ++k) for (int i = 0; i < d; ++i) C[i + j * d] += A[i + k * d] * B[k + j * d];
    return C;
}
static std::vector<double> dtr(const std::vector<double>& A, int d) { std::vector<double> T(A.size()); for (int i = 0; i < d; ++i) for (int j = 0; j < d; ++j) T[j + i * d] = A[i + j * d]; return T; }
static std::vector<double> axpy(double a, const std::vector<double>& x, const std::vector<double>& y) { std::vector<double> z(y); for (size_t i = 0; i < z.size(); ++i) z[i] += a * x[i]; return z; }
static std::vector<double> model_ham(int model, const double* p) {
    const int mt = mtype(model);
    if (mt == 2) {  // QIM (Models.cpp:325-330 -> XYZAbstractModel::MakeOps :57-93 with Jx = J, Jy = Jz = 0, hz = h)
        const int d = 2;
        std::vector<double> sx(4, 0), sz(4, 0), id(4, 0);
        sx[0 + 1 * d] = 0.5; sx[1 + 0 * d] = 0.5;
        sz[0] = 0.5; sz[1 + 1 * d] = -0.5;
        id[0] = 1; id[3] = 1;
        std::vector<double> H((size_t)16, 0.0);
        H = axpy(-p[0], kron(sx, sx, d), H);
        H = axpy(-0.5 * p[1], kron(sz, id, d), H);
        H = axpy(-0.5 * p[1], kron(id, sz, d), H);
        return H;
    }
    if (mt == 3) {  // QPOTTS, QuantumPottsModel::MakeOps :411-430; p = {h}
        const int q = model_d(model);
        std::vector<double> X((size_t)q * q, 0), Z((size_t)q * q, 0), id((size_t)q * q, 0);
        X[(q - 1) + 0 * q] = 1;
        Z[0] = q - 1;
        for (int i = 1; i < q; ++i) { X[(i - 1) + i * q] = 1; Z[i + i * q] = -1; }
        for (int i = 0; i < q; ++i) id[i + i * q] = 1;
        std::vector<std::vector<double>> MXq{X};
        for (int i = 1; i < q - 1; ++i) MXq.push_back(dmul(MXq.back(), X, q));
        std::vector<double> H((size_t)q * q * q * q, 0.0);
        H = axpy(-0.5 * p[0], kron(Z, id, q), H);
        H = axpy(-0.5 * p[0], kron(id, Z, q), H);
        for (int i = 0; i < q - 1; ++i) H = axpy(-1.0, kron(MXq[i], MXq[q - 2 - i], q), H);
        return H;
    }
    if (mt == 4) {  // BHUB, BoseHubbardModel::MakeOps :768-777; p = {t, U, mu}
        const int d = model_d(model);
        std::vector<double> b((size_t)d * d, 0), n((size_t)d * d, 0), id((size_t)d * d, 0);
        for (int i = 1; i < d; ++i) { b[(i - 1) + i * d] = std::sqrt((double)i); n[i + i * d] = i; }
        for (int i = 0; i < d; ++i) id[i + i * d] = 1;
        auto onsite = axpy(-p[2], n, dmul(axpy(0.5 * p[1], n, std::vector<double>((size_t)d * d, 0.0)), axpy(-1.0, id, n), d));
        auto bt = dtr(b, d);
        std::vector<double> H((size_t)d * d * d * d, 0.0);
        H = axpy(-p[0], kron(b, bt, d), H);
        H = axpy(-p[0], kron(bt, b, d), H);
        H = axpy(0.5, kron(onsite, id, d), H);
        H = axpy(0.5, kron(id, onsite, d), H);
        return H;
    }
    if (mt == 0) {  // XXZ: p = {J, Delta, hz}; Jx = Jy = J, Jz = Delta
        const int d = 2;
        std::vector<double> sx(4, 0), syi(4, 0), sz(4, 0), id(4, 0);
        sx[0 + 1 * d] = 0.5; sx[1 + 0 * d] = 0.5;
        syi[0 + 1 * d] = -0.5; syi[1 + 0 * d] = 0.5;
        sz[0] = 0.5; sz[1 + 1 * d] = -0.5;
        id[0] = 1; id[3] = 1;
        const double J = p[0], Delta = p[1], hz = p[2];
        std::vector<double> H((size_t)16, 0.0);
        H = axpy(-J, kron(sx, sx, d), H);
        H = axpy(J, kron(syi, syi, d), H);
        H = axpy(-Delta, kron(sz, sz, d), H);
        H = axpy(-0.5 * hz, kron(sz, id, d), H);
        H = axpy(-0.5 * hz, kron(id, sz, d), H);
        return H;
    }
    // FHUB: p = {t, U, V, mu, hz}
    const int d = 4;
    std::vector<double> cup(16, 0), cdo(16, 0), nup(16, 0), ndo(16, 0), id(16, 0), F(16, 0);
    cup[0 + 2 * d] = 1; cup[1 + 3 * d] = 1;
    cdo[0 + 1 * d] = 1; cdo[2 + 3 * d] = -1;
    nup[2 + 2 * d] = 1; nup[3 + 3 * d] = 1;
    ndo[1 + 1 * d] = 1; ndo[3 + 3 * d] = 1;
    for (int i = 0; i < 4; ++i) id[i + i * d] = 1;
    F[0] = 1; F[1 + d] = -1; F[2 + 2 * d] = -1; F[3 + 3 * d] = 1;
    const double t = p[0], U = p[1], V = p[2], mu = p[3], hz = p[4];
    auto nup2 = axpy(-0.5, id, nup), ndo2 = axpy(-0.5, id, ndo);
    auto nund = dmul(nup2, ndo2, d);
    auto nupnd = axpy(1.0, ndo, nup);
    auto nupnd2 = axpy(1.0, ndo2, nup2);
    std::vector<double> sz(16);
    for (int i = 0; i < 16; ++i) sz[i] = 0.5 * (nup[i] - ndo[i]);
    std::vector<double> H(256, 0.0);
    auto hop = kron(dmul(dtr(cup, d), F, d), cup, d);
    hop = axpy(-1.0, kron(dmul(cup, F, d), dtr(cup, d), d), hop);
    hop = axpy(1.0, kron(dmul(dtr(cdo, d), F, d), cdo, d), hop);
    hop = axpy(-1.0, kron(dmul(cdo, F, d), dtr(cdo, d), d), hop);
    H = axpy(-t, hop, H);
    H = axpy(0.5 * U, kron(nund, id, d), H);
    H = axpy(0.5 * U, kron(id, nund, d), H);
    H = axpy(V, kron(nupnd2, nupnd2, d), H);
    H = axpy(-0.5 * mu, kron(nupnd, id, d), H);
    H = axpy(-0.5 * mu, kron(id, nupnd, d), H);
    H = axpy(-0.5 * hz, kron(sz, id, d), H);
    H = axpy(-0.5 * hz, kron(id, sz, d), H);
    return H;
}

}  // namespace orc

// =============================================================================================
// C interface (ctypes).  Objects are lists of BlockMats: 1 entry for a BlockMat/BlockDiagMat,
// d entries for an MPS tensor, d*d (row-major) for a reduced operator.
using namespace orc;

struct OList { int nsym; std::vector<BM> v; };

static thread_local std::string g_err;
#define ORC_TRY(...) try { __VA_ARGS__ } catch (const std::exception& e) { g_err = e.what(); return 0; }

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

void* orc_new(int nsym, int nlist, const int* nblk, const int* qin, const int* qout, const int* rows, const int* cols, const double* data) {
    auto* o = new OList{nsym, std::vector<BM>(nlist)};
    size_t b = 0, pos = 0;
    for (int l = 0; l < nlist; ++l)
        for (int k = 0; k < nblk[l]; ++k, ++b) {
            Key ki(qin + b * nsym, qin + (b + 1) * nsym), ko(qout + b * nsym, qout + (b + 1) * nsym);
            Mat m(rows[b], cols[b]);
            if (data) std::copy(data + pos, data + pos + m.v.size(), m.v.begin());
            pos += m.v.size();
            o->v[l].emplace(ki, QM{ko, m});
        }
    return o;
}
void orc_free(void* h) { delete (OList*)h; }
int orc_nlist(void* h) { return (int)((OList*)h)->v.size(); }
int orc_nsym(void* h) { return ((OList*)h)->nsym; }
long orc_counts(void* h, int* nblk) {
    auto* o = (OList*)h;
    long n = 0;
    for (size_t l = 0; l < o->v.size(); ++l) { nblk[l] = (int)o->v[l].size(); for (auto& b : o->v[l]) n += (long)b.second.m.v.size(); }
    return n;
}
void orc_export(void* h, int* qin, int* qout, int* rows, int* cols, double* data) {
    auto* o = (OList*)h;
    size_t b = 0, pos = 0;
    for (auto& bm : o->v)
        for (auto& kv : bm) {
            for (int i = 0; i < o->nsym; ++i) { qin[b * o->nsym + i] = kv.first[i]; qout[b * o->nsym + i] = kv.second.qout[i]; }
            rows[b] = kv.second.m.r; cols[b] = kv.second.m.c;
            if (data) std::copy(kv.second.m.v.begin(), kv.second.m.v.end(), data + pos);
            pos += kv.second.m.v.size();
            ++b;
        }
}
static MPS as_mps(void* h) { return ((OList*)h)->v; }
static void* wrap(int nsym, std::vector<BM> v) { return new OList{nsym, std::move(v)}; }

// ---- transfer matrices ----
// dir 0: out = sum_s B_s^T x A_s ; dir 1: out = sum_s A_s x B_s^T.  `y` supplies the output structure.
void* orc_tm_mixed(int dir, void* A, void* B, void* x, void* y) {
    ORC_TRY(
        OList* yo = (OList*)y;
        BM out = yo->v[0];
        zero(out);
        if (dir == 0) tm_mixed_left_into(as_mps(A), as_mps(B ? B : A), ((OList*)x)->v[0], out);
        else tm_mixed_right_into(as_mps(A), as_mps(B ? B : A), ((OList*)x)->v[0], out);
        return wrap(yo->nsym, {out});
    )
}
void* orc_tm_ret(int dir, void* A, void* x) {
    ORC_TRY(
        BM out = dir == 0 ? tm_left_ret(as_mps(A), ((OList*)x)->v[0]) : tm_right_ret(as_mps(A), ((OList*)x)->v[0]);
        return wrap(((OList*)x)->nsym, {out});
    )
}
void* orc_tm_chain(int dir, void** Avec, int n, void* x, void* y) {
    ORC_TRY(
        std::vector<MPS> As;
        for (int i = 0; i < n; ++i) As.push_back(as_mps(Avec[i]));
        BM out = ((OList*)y)->v[0];
        tm_chain_into(As, dir, ((OList*)x)->v[0], out);
        return wrap(((OList*)y)->nsym, {out});
    )
}
void* orc_tm_unity(int kind, void* A, void* B) {  // 0 mixedLeftDiag, 1 mixedRightDiag, 2 mixedLeft(BlockMat), 3 mixedRight(BlockMat)
    ORC_TRY(
        MPS a = as_mps(A), b = as_mps(B ? B : A);
        BM out = kind == 0 ? tm_mixed_left_diag_unity(a, b) : kind == 1 ? tm_mixed_right_diag_unity(a, b) : kind == 2 ? tm_mixed_left_unity(a, b) : tm_mixed_right_unity(a, b);
        return wrap(((OList*)A)->nsym, {out});
    )
}
// ---- block products ----
void* orc_mps_mul(void* l, void* r) { ORC_TRY(return wrap(((OList*)l)->nsym, mpsmul(as_mps(l), as_mps(r)));) }
void* orc_mat_mps(void* m, void* a) { ORC_TRY(return wrap(((OList*)a)->nsym, matmps(((OList*)m)->v[0], as_mps(a)));) }
void* orc_mps_mat(void* a, void* m) { ORC_TRY(return wrap(((OList*)a)->nsym, mpsmat(as_mps(a), ((OList*)m)->v[0]));) }
void* orc_bmul(void* l, void* r) { ORC_TRY(return wrap(((OList*)l)->nsym, {bmul(((OList*)l)->v[0], ((OList*)r)->v[0])});) }
void* orc_btrans(void* a) { ORC_TRY(std::vector<BM> v; for (auto& b : ((OList*)a)->v) v.push_back(btrans(b)); return wrap(((OList*)a)->nsym, v);) }
// ---- operators ----
void* orc_apply_operator(void* AA, const double* H, int D) { ORC_TRY(return wrap(((OList*)AA)->nsym, apply_operator(as_mps(AA), H, D));) }
void* orc_reduced_op(const double* H, int d, void* A, int dir) { ORC_TRY(return wrap(((OList*)A)->nsym, reduced_op(H, d, as_mps(A), dir));) }
void* orc_op_tm_diag(int dir, const double* H, int d, void* AA) {
    ORC_TRY(return wrap(((OList*)AA)->nsym, {dir == 0 ? op_tm_left_diag(H, d, as_mps(AA)) : op_tm_right_diag(H, d, as_mps(AA))});)
}
void* orc_get_h(int dir, void** Avec, int n, const double* H, int d) {
    ORC_TRY(
        std::vector<MPS> As;
        for (int i = 0; i < n; ++i) As.push_back(as_mps(Avec[i]));
        return wrap(((OList*)Avec[0])->nsym, {dir == 0 ? get_hl(As, H, d) : get_hr(As, H, d)});
    )
}
void* orc_hac(void* EHL, void* EHR, void* HL, void* HR, void* x) {
    ORC_TRY(
        MPS out = as_mps(x);
        apply_hac(as_mps(x), out, ((OList*)EHL)->v[0], ((OList*)EHR)->v[0], ((OList*)HL)->v, ((OList*)HR)->v);
        return wrap(((OList*)x)->nsym, out);
    )
}
void* orc_hc(void* EHL, void* EHR, const double* H, int d, void* AL, void* AR, void* x) {
    ORC_TRY(
        BM out = ((OList*)x)->v[0];
        apply_hc(((OList*)x)->v[0], out, ((OList*)EHL)->v[0], ((OList*)EHR)->v[0], H, d, as_mps(AL), as_mps(AR));
        return wrap(((OList*)x)->nsym, {out});
    )
}
// ---- H_AC on flat vectors, compiled once and threaded (the CPU baseline) ----
// The same products as apply_hac (EHL*in + in*EHR and HL[i][j]*in[j] + in[j]*HR[i][j], matched as
// bmul matches blocks, kept where the output has the block), listed once per output block as
// segments over flat Vectorize-order arrays; apply runs the output blocks in parallel (OpenMP,
// dynamic schedule, one output block per task, each summing its segments into a zeroed buffer with
// an i-k-j loop that the compiler vectorises along the contiguous rows).
struct HacSeg { long a, b; int K, lda, ldb; int a_op; };  // a_op 0: operator block in the op array, 1: x
struct HacOut { long off; int M, N; int s0, ns; };
struct HacPlan { std::vector<HacOut> outs; std::vector<HacSeg> segs; long nx = 0, nop = 0; std::vector<double> op; };

void* orc_hac_plan(void* EHL, void* EHR, void* HL, void* HR, void* x) {
    ORC_TRY(
        auto* P = new HacPlan;
        const MPS in = as_mps(x);
        const BM& ehl = ((OList*)EHL)->v[0];
        const BM& ehr = ((OList*)EHR)->v[0];
        const std::vector<BM>& hl = ((OList*)HL)->v;
        const std::vector<BM>& hr = ((OList*)HR)->v;
        const size_t d = in.size();
        // flat offsets: x in Vectorize order; operators [EHL | EHR | HL | HR]
        std::vector<std::map<Key, long>> xo(d);
        long pos = 0;
        for (size_t s = 0; s < d; ++s) for (auto& b : in[s]) { xo[s][b.first] = pos; pos += (long)b.second.m.v.size(); }
        P->nx = pos;
        std::map<const Mat*, long> oo;
        auto addop = [&](const BM& bm) { for (auto& b : bm) { oo[&b.second.m] = (long)P->op.size(); P->op.insert(P->op.end(), b.second.m.v.begin(), b.second.m.v.end()); } };
        addop(ehl); addop(ehr);
        for (auto& b : hl) addop(b);
        for (auto& b : hr) addop(b);
        P->nop = (long)P->op.size();
        for (size_t i = 0; i < d; ++i)
            for (auto& ob : in[i]) {
                HacOut o{xo[i][ob.first], ob.second.m.r, ob.second.m.c, (int)P->segs.size(), 0};
                auto push = [&](long a, int aop, long b, int K, int lda, int ldb) { P->segs.push_back(HacSeg{a, b, K, lda, ldb, aop}); };
                // EHL*in[i]: EHL block keyed by Qin of the out block
                if (auto e = ehl.find(ob.first); e != ehl.end() && e->second.qout == ob.first)
                    push(oo[&e->second.m], 0, xo[i][ob.first], e->second.m.c, e->second.m.r, ob.second.m.r);
                // in[i]*EHR: EHR block keyed by Qout of the out block
                if (auto e = ehr.find(ob.second.qout); e != ehr.end())
                    push(xo[i][ob.first], 1, oo[&e->second.m], ob.second.m.c, ob.second.m.r, e->second.m.r);
                for (size_t j = 0; j < d; ++j) {
                    // HL[i][j]*in[j]: HL block q -> p, in[j] block p -> Qout(out)
                    if (auto h = hl[i * d + j].find(ob.first); h != hl[i * d + j].end())
                        if (auto xb = in[j].find(h->second.qout); xb != in[j].end() && xb->second.qout == ob.second.qout)
                            push(oo[&h->second.m], 0, xo[j][xb->first], h->second.m.c, h->second.m.r, xb->second.m.r);
                    // in[j]*HR[i][j]: in[j] block q -> p, HR block p -> Qout(out)
                    if (auto xb = in[j].find(ob.first); xb != in[j].end())
                        if (auto h = hr[i * d + j].find(xb->second.qout); h != hr[i * d + j].end() && h->second.qout == ob.second.qout)
                            push(xo[j][ob.first], 1, oo[&h->second.m], xb->second.m.c, xb->second.m.r, h->second.m.r);
                }
                o.ns = (int)P->segs.size() - o.s0;
                P->outs.push_back(o);
            }
        return P;
    )
}
void orc_hac_plan_free(void* p) { delete (HacPlan*)p; }
long orc_hac_plan_nx(void* p) { return ((HacPlan*)p)->nx; }
// y = H_AC x on flat arrays with nthreads OpenMP threads (0: the runtime's default)
void orc_hac_plan_apply(void* p, const double* x, double* y, int nthreads) {
    const HacPlan* P = (const HacPlan*)p;
    const double* op = P->op.data();
#ifdef _OPENMP
    const int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
#endif
    for (long t = 0; t < (long)P->outs.size(); ++t) {
        const HacOut& o = P->outs[t];
        double* c = y + o.off;
        std::fill(c, c + (long)o.M * o.N, 0.0);
        for (int si = o.s0; si < o.s0 + o.ns; ++si) {
            const HacSeg& g = P->segs[si];
            // segment C += A B; operator block either left (a_op 0: A in op, B in x) or right (A in x, B in op)
            const double* A = g.a_op == 0 ? op + g.a : x + g.a;
            const double* B = g.a_op == 0 ? x + g.b : op + g.b;
            int j = 0;
            for (; j + 4 <= o.N; j += 4) {  // four output columns share each column of A
                double* c0 = c + (long)j * o.M;
                double* c1 = c0 + o.M;
                double* c2 = c1 + o.M;
                double* c3 = c2 + o.M;
                const double* b0 = B + (long)j * g.ldb;
                for (int k = 0; k < g.K; ++k) {
                    const double* ak = A + (long)k * g.lda;
                    const double x0 = b0[k], x1 = b0[k + g.ldb], x2 = b0[k + 2L * g.ldb], x3 = b0[k + 3L * g.ldb];
                    for (int i = 0; i < o.M; ++i) {
                        const double a = ak[i];
                        c0[i] += a * x0;
                        c1[i] += a * x1;
                        c2[i] += a * x2;
                        c3[i] += a * x3;
                    }
                }
            }
            for (; j < o.N; ++j) {
                double* cj = c + (long)j * o.M;
                for (int k = 0; k < g.K; ++k) {
                    const double bkj = B[k + (long)j * g.ldb];
                    const double* ak = A + (long)k * g.lda;
                    for (int i = 0; i < o.M; ++i) cj[i] += ak[i] * bkj;
                }
            }
        }
    }
    (void)nthreads;
}
int orc_max_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// ---- environments ----
void* orc_invert_e_proj(int dir, void** Avec, int n, void* x, void* L, void* R, double tol, int maxit, void* y0, int* flag) {
    ORC_TRY(
        std::vector<MPS> As;
        for (int i = 0; i < n; ++i) As.push_back(as_mps(Avec[i]));
        BM y0b;
        if (y0) y0b = ((OList*)y0)->v[0];
        BM out = invert_e_proj(As, ((OList*)x)->v[0], ((OList*)L)->v[0], ((OList*)R)->v[0], dir, tol, (unsigned)maxit, y0 ? &y0b : nullptr, flag);
        return wrap(((OList*)x)->nsym, {out});
    )
}
// ---- decompositions ----
void* orc_mps_qr(void* A, int dir, void** R) {
    ORC_TRY(
        BM r;
        MPS q = mps_qr(as_mps(A), dir, r);
        *R = wrap(((OList*)A)->nsym, {r});
        return wrap(((OList*)A)->nsym, q);
    )
}
void* orc_nullspace(void* A, int dir) { ORC_TRY(return wrap(((OList*)A)->nsym, mps_nullspace(as_mps(A), dir));) }
// GradNormRight (VUMPS_helpers.hpp:147-153): norm_scaled(ApplyTMmixedRight(H_AC(C*AR), Nullspace(AR,r)))
// norm_scaled divides by sqrt(number of sectors) of the BlockMat (helpers.hpp:143-146).
double orc_gradnorm_right(void* AR, void* C, void* EHL, void* EHR, void* HL, void* HR) {
    try {
        MPS AC = matmps(((OList*)C)->v[0], as_mps(AR));
        MPS HAC = AC;
        apply_hac(AC, HAC, ((OList*)EHL)->v[0], ((OList*)EHR)->v[0], ((OList*)HL)->v, ((OList*)HR)->v);
        BM g = tm_mixed_right_unity(HAC, mps_nullspace(as_mps(AR), 1));
        return std::sqrt(bnorm2(g)) / std::sqrt((double)g.size());
    } catch (const std::exception& e) { g_err = e.what(); return -1; }
}
double orc_gradnorm_left(void* AL, void* C, void* EHL, void* EHR, void* HL, void* HR) {
    try {
        MPS AC = mpsmat(as_mps(AL), ((OList*)C)->v[0]);
        MPS HAC = AC;
        apply_hac(AC, HAC, ((OList*)EHL)->v[0], ((OList*)EHR)->v[0], ((OList*)HL)->v, ((OList*)HR)->v);
        // ApplyTMmixedLeft(HAC, NL): out[Qout NL_s] = (Qout HAC_s, NL_s^T HAC_s)
        BM g = tm_mixed_left_unity(HAC, mps_nullspace(as_mps(AL), 0));
        return std::sqrt(bnorm2(g)) / std::sqrt((double)g.size());
    } catch (const std::exception& e) { g_err = e.what(); return -1; }
}
// ---- bookkeeping ----
void orc_set_periods(int nsym, const int* periods) {
    g_per.assign(periods, periods + nsym);
    for (auto& p : g_per) if (p == 1) p = 0;
}
int orc_model_d(int model) { return model_d(model); }
int orc_model_nsym(int model) { return model_nsym(model); }
int orc_i2k(int model, int N, const int* QN, int* out) {
    try { use_group(model); auto I = get_i2k(model, N, QN); int k = 0; for (auto& q : I) for (int v : q) out[k++] = v; return (int)I.size(); }
    catch (const std::exception& e) { g_err = e.what(); return -1; }
}
int orc_validpath(int model, int N, const int* QN, int* path) {
    try { use_group(model); auto p = validpath(N, get_i2k(model, N, QN)); for (size_t i = 0; i < p.size(); ++i) path[i] = p[i]; return (int)p.size(); }
    catch (const std::exception& e) { g_err = e.what(); return -1; }
}
// GetDimMap + MakeDims: writes N bonds of (keys, sizes); returns sectors per bond
int orc_make_dims_model(int model, const int* dims, int nsec, int N, const int* QN, int* keys, int* sizes) {
    try {
        use_group(model);
        auto dm0 = get_dimmap(model, std::vector<int>(dims, dims + nsec), N, QN);
        auto dv = make_dims(dm0, N, get_i2k(model, N, QN));
        int k = 0, b = 0;
        for (auto& bond : dv) for (auto& kv : bond) { for (int v : kv.first) keys[k++] = v; sizes[b++] = kv.second; }
        return (int)dv[0].size();
    } catch (const std::exception& e) { g_err = e.what(); return -1; }
}
// MakeDims from an explicit bond-0 dim map
int orc_make_dims(int model, int N, const int* QN, int nsym, int nsec, const int* keys0, const int* sizes0, int* keys, int* sizes) {
    try {
        use_group(model);
        std::map<Key, int> dm0;
        for (int i = 0; i < nsec; ++i) dm0[Key(keys0 + i * nsym, keys0 + (i + 1) * nsym)] = sizes0[i];
        auto dv = make_dims(dm0, N, get_i2k(model, N, QN));
        int k = 0, b = 0;
        for (auto& bond : dv) for (auto& kv : bond) { for (int v : kv.first) keys[k++] = v; sizes[b++] = kv.second; }
        return (int)dv[0].size();
    } catch (const std::exception& e) { g_err = e.what(); return -1; }
}
void* orc_mps_structure(int model, int N, const int* QN, int nsym, int nl, const int* kl, const int* sl, int nr, const int* kr, const int* sr) {
    ORC_TRY(
        use_group(model);
        std::map<Key, int> ml, mr;
        for (int i = 0; i < nl; ++i) ml[Key(kl + i * nsym, kl + (i + 1) * nsym)] = sl[i];
        for (int i = 0; i < nr; ++i) mr[Key(kr + i * nsym, kr + (i + 1) * nsym)] = sr[i];
        return wrap(nsym, mps_structure(get_i2k(model, N, QN), ml, mr));
    )
}
void orc_model_ham(int model, const double* params, double* H) {
    if (model_d(model) < 0) return;
    auto h = model_ham(model, params);
    std::copy(h.begin(), h.end(), H);
}

}  // extern "C"
