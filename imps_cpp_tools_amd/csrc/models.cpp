// Sector bookkeeping of the models: spin-1/2 XXZ with one U(1), the Fermi-Hubbard chain in (N, M)
// representation with U(1)xU(1), the quantum Ising chain with Z(2), the q-state quantum Potts
// chain with Z(q) and the Bose-Hubbard chain (local dimension d) with U(1).  Integer-exact
// restatement of Models.cpp (GetdQ, GetI2K, GetDimMap, MakeDims, the symmetry groups of Init) and
// ItoKey.hpp (validpath).  A model id is type | (d << 8): the local dimension is part of the id for
// the Potts (q) and Bose-Hubbard chains and is 0 (or the fixed value) for the others.
#include <cmath>
#include <cstdlib>
#include <map>
#include <stdexcept>
#include <vector>

#include "../../include/imps_capi.h"
#include "layout.hpp"

namespace imps {
namespace {

int igcd(int u, int v) {
    u = std::abs(u);
    v = std::abs(v);
    while (v != 0) { int r = u % v; u = v; v = r; }
    return u;
}

struct DQ { int period, offset; };

enum ModelType { XXZ = 0, FHUB = 1, QIM = 2, QPOTTS = 3, BHUB = 4 };
constexpr int kMaxLocalDim = 16;

int mtype(int model) { return model & 0xff; }

// local dimension of a model id, or -1 for an unknown id
int local_dim(int model) {
    const int t = mtype(model), dd = model >> 8;
    switch (t) {
        case XXZ: case QIM: return (dd == 0 || dd == 2) ? 2 : -1;
        case FHUB: return (dd == 0 || dd == 4) ? 4 : -1;
        case QPOTTS: case BHUB: return (dd >= 2 && dd <= kMaxLocalDim) ? dd : -1;
        default: return -1;
    }
}

int nsym_of(int model) {
    if (local_dim(model) < 0) return -1;
    return mtype(model) == FHUB ? 2 : 1;
}

// symmetry groups of the models' Init (Models.cpp: XXZ {0} :200, QIM {2} :335, QPOTTS {q} :393,
// FHUB {0, 0} :521, BHUB {0} :749); 0 = U(1)
std::vector<int> periods_of(int model) {
    const int t = mtype(model), dd = local_dim(model);
    if (dd < 0) throw InvalidArg("unknown model");
    if (t == FHUB) return {0, 0};
    if (t == QIM) return {2};
    if (t == QPOTTS) return {dd};
    return {0};
}

Key mkkey(std::initializer_list<int> v, const std::vector<int>& per) {
    Key k;
    k.n = (int)v.size();
    int i = 0;
    for (int x : v) { k.per[i] = per[i]; k.q[i] = x; ++i; }
    return k;
}

// GetdQ: XXZ Models.cpp:216-240 (spin-1/2 branch); FHUB (N,M) Models.cpp:642-662; QIM :351-355 and
// QPOTTS :472-476 ((1, 0), unused by their I2K); BHUB :830-843
std::vector<DQ> getdQ(int model, int N, const int* QN) {
    if (local_dim(model) < 0) throw InvalidArg("unknown model");
    const int t = mtype(model);
    if (t == QIM || t == QPOTTS) return {{1, 0}};
    if (t == BHUB) {
        const int n = QN[0], dd = local_dim(model);
        if (n < 1 || n >= N * dd) throw InvalidArg("BHUB: 0 < n < N d required");
        const int div = igcd(N, n);
        return {{N / div, -n / div}};
    }
    if (t == XXZ) {
        const int m = QN[0];
        if (std::abs(m) >= N) throw InvalidArg("XXZ: |m| must be < N");
        const int div = igcd(N - m, N + m);
        return {{2 * N / div, -(N + m) / div}};
    }
    if (t == FHUB) {
        const int n = QN[0], m = QN[1];
        if (n + m <= 0 || n + m >= 2 * N || n - m <= 0 || n - m >= 2 * N) throw InvalidArg("FHUB: (n, m) out of range");
        const int divn = igcd(igcd(N - n, 2 * N - n), n);
        const int divm = igcd(igcd(N - m, N + m), m);
        return {{N / divn, -n / divn}, {N / divm, -(N + m) / divm}};
    }
    throw InvalidArg("unknown model");
}

// GetI2K: XXZ Models.cpp:262-277 (s = d-1 ... 0 emplaced in that order); FHUB Models.cpp:699-718;
// QIM :372-376 (s -> s); QPOTTS :500-506 (s -> s); BHUB :862-878 (s -> offset + s period)
std::vector<Key> getI2K(int model, int N, const int* QN) {
    auto dQ = getdQ(model, N, QN);
    const auto per = periods_of(model);
    const int t = mtype(model), dd = local_dim(model);
    std::vector<Key> I;
    auto mk = [&](std::initializer_list<int> v) { return mkkey(v, per); };
    if (t == XXZ) {
        for (int s = 1; s >= 0; --s) I.push_back(mk({dQ[0].offset + s * dQ[0].period}));
    } else if (t == QIM || t == QPOTTS) {
        for (int s = 0; s < dd; ++s) I.push_back(mk({s}));
    } else if (t == BHUB) {
        for (int s = 0; s < dd; ++s) I.push_back(mk({dQ[0].offset + s * dQ[0].period}));
    } else {
        const int p0 = dQ[0].period, o0 = dQ[0].offset, p1 = dQ[1].period, o1 = dQ[1].offset;
        I.push_back(mk({o0, o1 + p1}));           // |0> empty
        I.push_back(mk({o0 + p0, o1}));           // |1> down
        I.push_back(mk({o0 + p0, o1 + 2 * p1}));  // |2> up
        I.push_back(mk({o0 + 2 * p0, o1 + p1}));  // |3> double
    }
    return I;
}

// path_iter / validpath (ItoKey.hpp:126-165): depth-first search, first match wins
bool path_iter(int N, const Key& in, const std::vector<Key>& I2K, const Key& zero, std::vector<int>& path) {
    for (size_t j = 0; j < I2K.size(); ++j) {
        if (N == 1) {
            if (in + I2K[j] == zero) { path = {(int)j}; return true; }
        } else {
            std::vector<int> p;
            if (path_iter(N - 1, in + I2K[j], I2K, zero, p)) { p.push_back((int)j); path = p; return true; }
        }
    }
    return false;
}

std::vector<int> validpath(int N, const std::vector<Key>& I2K) {
    std::vector<int> p;
    if (!path_iter(N, I2K.front(), I2K, I2K.front(), p)) throw InvalidArg("validpath: no valid path");
    return p;
}

template <class F>
int guard(F&& f) {
    try { return f(); } catch (...) { return -1; }
}

void kron(const double* A, const double* B, int d, double fac, double* K) {  // K += fac * kron(A, B)
    const int D = d * d;
    for (int i = 0; i < d; ++i) for (int k = 0; k < d; ++k) for (int j = 0; j < d; ++j) for (int l = 0; l < d; ++l)
        K[(i * d + j) + (size_t)(k * d + l) * D] += fac * A[i + k * d] * B[j + l * d];
}
void dmul(const double* A, const double* B, int d, double* C) {
    for (int i = 0; i < d * d; ++i) C[i] = 0;
    for (int j = 0; j < d; ++j) for (int k = 0; k < d; ++k) for (int i = 0; i < d; ++i) C[i + j * d] += A[i + k * d] * B[k + j * d];
}
void dtrans(const double* A, int d, double* T) { for (int i = 0; i < d; ++i) for (int j = 0; j < d; ++j) T[j + i * d] = A[i + j * d]; }

}  // namespace
}  // namespace imps

using namespace imps;

extern "C" {

int imps_model_nsym(int model) { return nsym_of(model); }
int imps_model_d(int model) { return local_dim(model); }

int imps_model_periods(int model, int* periods) {
    return guard([&] {
        if (!periods) throw InvalidArg("bad args");
        auto p = periods_of(model);
        for (size_t i = 0; i < p.size(); ++i) periods[i] = p[i];
        return (int)p.size();
    });
}

int imps_model_i2k(int model, int N, const int* QN, int* i2k) {
    return guard([&] {
        if (!QN || !i2k || N < 1) throw InvalidArg("bad args");
        auto I = getI2K(model, N, QN);
        int k = 0;
        for (auto& q : I) for (int i = 0; i < q.n; ++i) i2k[k++] = q.q[i];
        return (int)I.size();
    });
}

int imps_model_validpath(int model, int N, const int* QN, int* path) {
    return guard([&] {
        if (!QN || !path || N < 1) throw InvalidArg("bad args");
        auto p = validpath(N, getI2K(model, N, QN));
        for (size_t i = 0; i < p.size(); ++i) path[i] = p[i];
        return (int)p.size();
    });
}

// GetDimMap: XXZ Models.cpp:248-260; FHUB (N,M) Models.cpp:680-697; QIM :357-370 (exactly two
// sectors {0}, {1}: more are cut to two, fewer are an error); QPOTTS :478-498 (q sectors {0..q-1},
// likewise); BHUB :845-860 (sector i at charge i period).  Keys are written in ascending (map) order.
int imps_model_dimmap(int model, int N, const int* QN, const int* dims, int ndims, int* keys, int* sizes) {
    return guard([&] {
        if (!QN || !dims || !keys || !sizes || ndims < 1) throw InvalidArg("bad args");
        auto dQ = getdQ(model, N, QN);
        const int ns = nsym_of(model), t = mtype(model), dd = local_dim(model);
        const auto per = periods_of(model);
        std::map<Key, int> dm;
        if (t == QIM || t == QPOTTS) {
            if (ndims < dd) throw InvalidArg("Z(q) models need at least q sectors");
            ndims = dd;
        }
        for (int i = 0; i < ndims; ++i) {
            if (t == QIM || t == QPOTTS) {
                dm[mkkey({i}, per)] = dims[i];
            } else if (t == BHUB) {
                dm[mkkey({i * dQ[0].period}, per)] = dims[i];
            } else if (t == XXZ) {
                int q = dQ[0].period * (i - ndims / 2);
                dm[Key(&q, 1)] = dims[i];
            } else {
                for (int j = 0; j < ndims; ++j) {
                    int q[2] = {dQ[0].period * (i + j - ndims + 1), dQ[1].period * (i - j)};
                    dm[Key(q, 2)] = dims[i] + dims[j];
                }
            }
        }
        int b = 0;
        for (auto& kv : dm) { for (int i = 0; i < ns; ++i) keys[b * ns + i] = kv.first.q[i]; sizes[b] = kv.second; ++b; }
        return b;
    });
}

// MakeDims (Models.cpp:13-32): bond k+1 = bond k shifted by I2K[path[N-k-1]]
int imps_model_make_dims(int model, int N, const int* QN, int nsec, const int* keys0, const int* sizes0, int* keys, int* sizes) {
    return guard([&] {
        if (!QN || !keys0 || !sizes0 || !keys || !sizes || nsec < 1) throw InvalidArg("bad args");
        const int ns = nsym_of(model);
        const auto per = periods_of(model);
        auto I2K = getI2K(model, N, QN);
        auto path = validpath(N, I2K);
        std::map<Key, int> l;
        for (int i = 0; i < nsec; ++i) {
            Key k(keys0 + i * ns, ns, per.data());
            if (!k.canonical()) throw InvalidArg("a Z(N) charge outside [0, N)");
            l[k] = sizes0[i];
        }
        if ((int)l.size() != nsec) throw InvalidArg("duplicate sector keys");
        int b = 0;
        for (int k = 0; k < N; ++k) {
            for (auto& kv : l) { for (int i = 0; i < ns; ++i) keys[b * ns + i] = kv.first.q[i]; sizes[b] = kv.second; ++b; }
            std::map<Key, int> r;
            for (auto& kv : l) r[kv.first + I2K[path[N - k - 1]]] = kv.second;
            l.swap(r);
        }
        return nsec;
    });
}

// local two-site Hamiltonians, dense d^2 x d^2 column-major: XXZ Models.cpp:57-93, FHUB :543-572,
// QIM, QPOTTS and BHUB as below
int imps_model_ham(int model, const double* p, double* H) {
    return guard([&] {
        if (!p || !H) throw InvalidArg("bad args");
        const int t = mtype(model), dd = local_dim(model);
        if (dd < 0) throw InvalidArg("unknown model");
        if (t == QIM) {
            // XYZAbstractModel::MakeOps (Models.cpp:57-93) with Jx = J, Jy = Jz = 0, hz = h
            // (QuantumIsingModel, :325-330): -J sx sx - h/2 (sz 1 + 1 sz)
            const int d = 2;
            double sx[4] = {0, 0.5, 0.5, 0}, sz[4] = {0.5, 0, 0, -0.5}, id[4] = {1, 0, 0, 1};
            for (int i = 0; i < 16; ++i) H[i] = 0;
            kron(sx, sx, d, -p[0], H);
            kron(sz, id, d, -0.5 * p[1], H);
            kron(id, sz, d, -0.5 * p[1], H);
            return 16;
        }
        if (t == QPOTTS) {
            // QuantumPottsModel::MakeOps (Models.cpp:411-430): X the cyclic lowering X(i-1, i) = 1,
            // X(q-1, 0) = 1; Z = diag(q-1, -1, ..., -1);
            // H = -h/2 (Z 1 + 1 Z) - sum_{i=0}^{q-2} X^(i+1) X^(q-1-i)
            const int q = dd, D = q * q;
            std::vector<double> X((size_t)D, 0.0), Z((size_t)D, 0.0), id((size_t)D, 0.0);
            X[(q - 1) + 0 * q] = 1.0;
            Z[0] = q - 1;
            for (int i = 1; i < q; ++i) { X[(i - 1) + (size_t)i * q] = 1.0; Z[i + (size_t)i * q] = -1.0; }
            for (int i = 0; i < q; ++i) id[i + (size_t)i * q] = 1.0;
            std::vector<std::vector<double>> MX{X};
            for (int i = 1; i < q - 1; ++i) {
                std::vector<double> M((size_t)D);
                dmul(MX.back().data(), X.data(), q, M.data());
                MX.push_back(M);
            }
            for (int i = 0; i < D * D; ++i) H[i] = 0;
            kron(Z.data(), id.data(), q, -0.5 * p[0], H);
            kron(id.data(), Z.data(), q, -0.5 * p[0], H);
            for (int i = 0; i < q - 1; ++i) kron(MX[i].data(), MX[q - 2 - i].data(), q, -1.0, H);
            return D * D;
        }
        if (t == BHUB) {
            // BoseHubbardModel::MakeOps (Models.cpp:768-777): b(i-1, i) = sqrt(i), n = diag(0..d-1),
            // onsite = U/2 n (n - 1) - mu n; H = -t (b b^T + b^T b) + (onsite 1 + 1 onsite) / 2
            const int d = dd, D = d * d;
            std::vector<double> bop((size_t)D, 0.0), bT((size_t)D, 0.0), on((size_t)D, 0.0), id((size_t)D, 0.0);
            for (int i = 1; i < d; ++i) bop[(i - 1) + (size_t)i * d] = std::sqrt((double)i);
            dtrans(bop.data(), d, bT.data());
            for (int i = 0; i < d; ++i) {
                on[i + (size_t)i * d] = 0.5 * p[1] * i * (i - 1.0) - p[2] * i;
                id[i + (size_t)i * d] = 1.0;
            }
            for (int i = 0; i < D * D; ++i) H[i] = 0;
            kron(bop.data(), bT.data(), d, -p[0], H);
            kron(bT.data(), bop.data(), d, -p[0], H);
            kron(on.data(), id.data(), d, 0.5, H);
            kron(id.data(), on.data(), d, 0.5, H);
            return D * D;
        }
        if (t == XXZ) {
            const int d = 2;
            double sx[4] = {0, 0.5, 0.5, 0}, syi[4] = {0, 0.5, -0.5, 0}, sz[4] = {0.5, 0, 0, -0.5}, id[4] = {1, 0, 0, 1};
            for (int i = 0; i < 16; ++i) H[i] = 0;
            kron(sx, sx, d, -p[0], H);
            kron(syi, syi, d, p[0], H);
            kron(sz, sz, d, -p[1], H);
            kron(sz, id, d, -0.5 * p[2], H);
            kron(id, sz, d, -0.5 * p[2], H);
            return 16;
        }
        if (t == FHUB) {
            const int d = 4;
            double cup[16] = {0}, cdo[16] = {0}, nup[16] = {0}, ndo[16] = {0}, id[16] = {0}, F[16] = {0};
            cup[0 + 2 * d] = 1; cup[1 + 3 * d] = 1;
            cdo[0 + 1 * d] = 1; cdo[2 + 3 * d] = -1;
            nup[2 + 2 * d] = 1; nup[3 + 3 * d] = 1;
            ndo[1 + 1 * d] = 1; ndo[3 + 3 * d] = 1;
            for (int i = 0; i < 4; ++i) id[i + i * d] = 1;
            F[0] = 1; F[1 + d] = -1; F[2 + 2 * d] = -1; F[3 + 3 * d] = 1;
            const double t = p[0], U = p[1], V = p[2], mu = p[3], hz = p[4];
            double nup2[16], ndo2[16], nund[16], nupnd[16], nupnd2[16], sz[16], T1[16], T2[16];
            for (int i = 0; i < 16; ++i) {
                nup2[i] = nup[i] - 0.5 * id[i];
                ndo2[i] = ndo[i] - 0.5 * id[i];
                nupnd[i] = nup[i] + ndo[i];
                sz[i] = 0.5 * (nup[i] - ndo[i]);
            }
            for (int i = 0; i < 16; ++i) nupnd2[i] = nup2[i] + ndo2[i];
            dmul(nup2, ndo2, d, nund);
            for (int i = 0; i < 256; ++i) H[i] = 0;
            double cupT[16], cdoT[16];
            dtrans(cup, d, cupT);
            dtrans(cdo, d, cdoT);
            dmul(cupT, F, d, T1); kron(T1, cup, d, -t, H);
            dmul(cup, F, d, T2); kron(T2, cupT, d, t, H);
            dmul(cdoT, F, d, T1); kron(T1, cdo, d, -t, H);
            dmul(cdo, F, d, T2); kron(T2, cdoT, d, t, H);
            kron(nund, id, d, 0.5 * U, H);
            kron(id, nund, d, 0.5 * U, H);
            kron(nupnd2, nupnd2, d, V, H);
            kron(nupnd, id, d, -0.5 * mu, H);
            kron(id, nupnd, d, -0.5 * mu, H);
            kron(sz, id, d, -0.5 * hz, H);
            kron(id, sz, d, -0.5 * hz, H);
            return 256;
        }
        throw InvalidArg("unknown model");
    });
}

// MPSBlockMat(I2K, ml, mr) structure (MPSBlockMat.hpp:221-236)
int imps_mps_structure(int model, int N, const int* QN, int nl, const int* kl, const int* sl, int nr, const int* kr, const int* sr,
                       int* nblk, int* qin, int* qout, int* rows, int* cols) {
    return guard([&] {
        if (!QN || !kl || !sl || !kr || !sr || !nblk || !qin || !qout || !rows || !cols) throw InvalidArg("bad args");
        const int ns = nsym_of(model);
        const auto per = periods_of(model);
        auto I2K = getI2K(model, N, QN);
        std::map<Key, int> ml, mr;
        for (int i = 0; i < nl; ++i) ml[Key(kl + i * ns, ns, per.data())] = sl[i];
        for (int i = 0; i < nr; ++i) mr[Key(kr + i * ns, ns, per.data())] = sr[i];
        for (auto* m : {&ml, &mr})
            for (auto& kv : *m)
                if (!kv.first.canonical()) throw InvalidArg("a Z(N) charge outside [0, N)");
        int b = 0;
        for (size_t s = 0; s < I2K.size(); ++s) {
            nblk[s] = 0;
            for (auto& l : ml) {
                auto r = mr.find(l.first + I2K[s]);
                if (r == mr.end()) continue;
                for (int i = 0; i < ns; ++i) { qin[b * ns + i] = l.first.q[i]; qout[b * ns + i] = r->first.q[i]; }
                rows[b] = l.second;
                cols[b] = r->second;
                ++b;
                ++nblk[s];
            }
        }
        return b;
    });
}

}  // extern "C"
