// Sector bookkeeping for the two models of the hot path: spin-1/2 XXZ with one U(1) and the
// Fermi-Hubbard chain in (N, M) representation with U(1)xU(1).  Integer-exact restatement of
// Models.cpp (GetdQ, GetI2K, GetDimMap, MakeDims) and ItoKey.hpp (validpath).
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

// GetdQ: XXZ Models.cpp:216-240 (spin-1/2 branch); FHUB (N,M) Models.cpp:642-662
std::vector<DQ> getdQ(int model, int N, const int* QN) {
    if (model == 0) {
        const int m = QN[0];
        if (std::abs(m) >= N) throw InvalidArg("XXZ: |m| must be < N");
        const int div = igcd(N - m, N + m);
        return {{2 * N / div, -(N + m) / div}};
    }
    if (model == 1) {
        const int n = QN[0], m = QN[1];
        if (n + m <= 0 || n + m >= 2 * N || n - m <= 0 || n - m >= 2 * N) throw InvalidArg("FHUB: (n, m) out of range");
        const int divn = igcd(igcd(N - n, 2 * N - n), n);
        const int divm = igcd(igcd(N - m, N + m), m);
        return {{N / divn, -n / divn}, {N / divm, -(N + m) / divm}};
    }
    throw InvalidArg("unknown model");
}

// GetI2K: XXZ Models.cpp:262-277 (s = d-1 ... 0 emplaced in that order); FHUB Models.cpp:699-718
std::vector<Key> getI2K(int model, int N, const int* QN) {
    auto dQ = getdQ(model, N, QN);
    std::vector<Key> I;
    auto mk = [](std::initializer_list<int> v) { Key k; k.n = (int)v.size(); int i = 0; for (int x : v) k.q[i++] = x; return k; };
    if (model == 0) {
        for (int s = 1; s >= 0; --s) I.push_back(mk({dQ[0].offset + s * dQ[0].period}));
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

int nsym_of(int model) { return model == 0 ? 1 : model == 1 ? 2 : -1; }

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
int imps_model_d(int model) { return model == 0 ? 2 : model == 1 ? 4 : -1; }

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

// GetDimMap: XXZ Models.cpp:248-260; FHUB (N,M) Models.cpp:680-697.  Keys are written in
// ascending (map) order.
int imps_model_dimmap(int model, int N, const int* QN, const int* dims, int ndims, int* keys, int* sizes) {
    return guard([&] {
        if (!QN || !dims || !keys || !sizes || ndims < 1) throw InvalidArg("bad args");
        auto dQ = getdQ(model, N, QN);
        const int ns = nsym_of(model);
        std::map<Key, int> dm;
        for (int i = 0; i < ndims; ++i) {
            if (model == 0) {
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
        auto I2K = getI2K(model, N, QN);
        auto path = validpath(N, I2K);
        std::map<Key, int> l;
        for (int i = 0; i < nsec; ++i) l[Key(keys0 + i * ns, ns)] = sizes0[i];
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

// local two-site Hamiltonians, dense d^2 x d^2 column-major: XXZ Models.cpp:57-93, FHUB :543-572
int imps_model_ham(int model, const double* p, double* H) {
    return guard([&] {
        if (!p || !H) throw InvalidArg("bad args");
        if (model == 0) {
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
        if (model == 1) {
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
        auto I2K = getI2K(model, N, QN);
        std::map<Key, int> ml, mr;
        for (int i = 0; i < nl; ++i) ml[Key(kl + i * ns, ns)] = sl[i];
        for (int i = 0; i < nr; ++i) mr[Key(kr + i * ns, ns)] = sr[i];
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
