// Excitation engine (heff.hpp): recording of the HeffConstants and ApplyHeff tapes, their
// execution, and the momentum-batched complex GMRES and Lanczos drivers.
#include "heff.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <numeric>

namespace imps {

namespace {

constexpr int kCombTile = 4096;  // elements per gather-scale-sum tile

void d2h_sync(hipStream_t s, void* h, const void* d, size_t bytes) {
    if (!bytes) return;
    hip_check(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, s), "D2H");
    hip_check(hipStreamSynchronize(s), "sync");
}
void h2d_sync(hipStream_t s, void* d, const void* h, size_t bytes) {
    if (!bytes) return;
    hip_check(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s), "H2D");
    hip_check(hipStreamSynchronize(s), "sync");
}

bool same_layout(const Layout& a, const Layout& b) {
    if (a.nsym != b.nsym || a.l.size() != b.l.size() || a.nelem != b.nelem) return false;
    for (size_t e = 0; e < a.l.size(); ++e) {
        if (a.l[e].b.size() != b.l[e].b.size()) return false;
        for (size_t i = 0; i < a.l[e].b.size(); ++i) {
            const Blk &x = a.l[e].b[i], &y = b.l[e].b[i];
            if (x.qin != y.qin || x.qout != y.qout || x.rows != y.rows || x.cols != y.cols) return false;
        }
    }
    return true;
}

// Layout from per-entry maps qin -> block (offsets assigned in entry, then key order)
Layout from_maps(int nsym, const int* per, const std::vector<std::map<Key, Blk>>& ents) {
    Layout L;
    L.nsym = nsym;
    for (int i = 0; i < kMaxSym; ++i) L.per[i] = per[i];
    L.l.resize(ents.size());
    long pos = 0;
    for (size_t e = 0; e < ents.size(); ++e)
        for (auto& kv : ents[e]) {
            Blk b = kv.second;
            b.off = pos;
            pos += b.size();
            L.l[e].b.push_back(b);
        }
    L.nelem = pos;
    return L;
}

// Diagonal layout of one bond of an MPS layout: one square block per charge, from the blocks' right
// (outgoing) or left (ingoing) side
Layout bond_layout(const Layout& A, bool right_side) {
    std::vector<std::map<Key, Blk>> ent(1);
    for (const BList& e : A.l)
        for (const Blk& b : e.b) {
            const Key q = right_side ? b.qout : b.qin;
            const int m = right_side ? b.cols : b.rows;
            auto it = ent[0].find(q);
            if (it == ent[0].end()) ent[0].emplace(q, Blk{q, q, m, m, 0});
            else if (it->second.rows != m) throw InvalidArg("heff: a bond sector with two dimensions");
        }
    return from_maps(A.nsym, A.per, ent);
}

// Charge-conserving blocks (q, q) between two bonds: rows from bond a, columns from bond b
Layout cross_layout(const Layout& a, const Layout& b) {
    std::vector<std::map<Key, Blk>> ent(1);
    for (const Blk& x : a.l[0].b)
        if (const Blk* y = b.l[0].find(x.qin)) ent[0].emplace(x.qin, Blk{x.qin, x.qin, x.rows, y->rows, 0});
    return from_maps(a.nsym, a.per, ent);
}

// Host copy of `src` (layout sl) onto layout dl by charge: every block of sl must exist in dl with
// the same shape; blocks of dl without a source stay zero
std::vector<double> map_blocks(const Layout& sl, const std::vector<double>& src, const Layout& dl, const char* what) {
    std::vector<double> out(dl.nelem, 0.0);
    if (sl.l.size() != 1) throw InvalidArg(std::string("heff: ") + what + " needs nlist = 1");
    for (const Blk& b : sl.l[0].b) {
        const Blk* t = dl.l[0].find(b.qin);
        if (!t || t->qout != b.qout || t->rows != b.rows || t->cols != b.cols)
            throw InvalidArg(std::string("heff: ") + what + " has a block outside the bond structure of AL and AR");
        std::copy(src.begin() + b.off, src.begin() + b.off + b.size(), out.begin() + t->off);
    }
    return out;
}

long entry_start(const Layout& L, int e) {
    for (int k = e; k < (int)L.l.size(); ++k)
        if (!L.l[k].b.empty()) return L.l[k].b.front().off;
    return L.nelem;
}

Layout entry_layout(const Layout& L, int e) {
    Layout out;
    set_group(out, L);
    out.l.resize(1);
    const long base = entry_start(L, e);
    long pos = 0;
    for (Blk b : L.l[e].b) {
        b.off -= base;
        pos += b.size();
        out.l[0].b.push_back(b);
    }
    out.nelem = pos;
    return out;
}

// Cyclic Jacobi eigen-decomposition of a Hermitian n x n matrix (column-major complex): per pair
// (p, q) a phase makes a_pq real, then the real Jacobi rotation zeroes it.  Ascending eigenvalues.
void herm_eig(int n, std::vector<cplx> A, std::vector<double>& w, std::vector<cplx>& Z) {
    Z.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) Z[i + (size_t)i * n] = 1.0;
    auto a = [&](int i, int j) -> cplx& { return A[i + (size_t)j * n]; };
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0, tot = 0;
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const double v = std::norm(a(i, j));
                tot += v;
                if (i != j) off += v;
            }
        if (off <= 1e-30 * tot || off == 0) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double r = std::abs(a(p, q));
                if (r == 0) continue;
                const cplx ph = a(p, q) / r;  // e^{i phi}
                const double theta = (a(q, q).real() - a(p, p).real()) / (2 * r);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::abs(theta) + std::sqrt(theta * theta + 1));
                const double c = 1 / std::sqrt(t * t + 1), s = t * c;
                const cplx e = std::conj(ph);  // e^{-i phi}
                // columns: p' = c p - s e q, q' = s p + c e q
                for (int k = 0; k < n; ++k) {
                    const cplx akp = a(k, p), akq = a(k, q);
                    a(k, p) = c * akp - s * e * akq;
                    a(k, q) = s * akp + c * e * akq;
                }
                // rows: p' = c p - s conj(e) q, q' = s p + c conj(e) q
                for (int k = 0; k < n; ++k) {
                    const cplx apk = a(p, k), aqk = a(q, k);
                    a(p, k) = c * apk - s * ph * aqk;
                    a(q, k) = s * apk + c * ph * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    const cplx zkp = Z[k + (size_t)p * n], zkq = Z[k + (size_t)q * n];
                    Z[k + (size_t)p * n] = c * zkp - s * e * zkq;
                    Z[k + (size_t)q * n] = s * zkp + c * e * zkq;
                }
            }
    }
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int x, int y) { return a(x, x).real() < a(y, y).real(); });
    std::vector<cplx> Zs((size_t)n * n);
    w.resize(n);
    for (int k = 0; k < n; ++k) {
        w[k] = a(order[k], order[k]).real();
        for (int i = 0; i < n; ++i) Zs[i + (size_t)k * n] = Z[i + (size_t)order[k] * n];
    }
    Z.swap(Zs);
}

// IMPS_VERBOSE=1: progress lines of the excitation solvers on stderr
bool verbose() {
    static const bool v = [] {
        const char* e = std::getenv("IMPS_VERBOSE");
        return e && e[0] == '1';
    }();
    return v;
}

}  // namespace

// ================================================================================================
// recording
long BlockTape::alloc(long n, bool cplx) {
    const long off = arena_elems;
    arena_elems += (cplx ? 2L * nmom : 1L) * std::max<long>(n, 0);
    arena_elems = (arena_elems + 31) / 32 * 32;  // 256-byte aligned objects
    return off;
}

XObj BlockTape::make(const Layout& lay, bool cplx) {
    XObj o;
    o.lay = lay;
    o.cplx = cplx;
    o.ps = lay.nelem;
    o.off = alloc(lay.nelem, cplx);
    return o;
}

XObj BlockTape::view_entry(const XObj& x, int e) const {
    XObj v;
    v.lay = entry_layout(x.lay, e);
    v.off = x.off + entry_start(x.lay, e);
    v.ps = x.ps;
    v.cplx = x.cplx;
    return v;
}

int BlockTape::coef(cplx v, int pow) {
    for (size_t i = 0; i < coefs.size(); ++i)
        if (coefs[i].v == v && coefs[i].pow == pow) return (int)i;
    coefs.push_back(CoefSpec{v, pow});
    return (int)coefs.size() - 1;
}

XObj BlockTape::prod(int kind, const XObj& a, const XObj& b) {
    if (a.cplx && b.cplx) throw InvalidArg("heff: product of two complex objects");
    ProdOp op;
    Layout out;
    op.prog = build_prod(kind, a.lay, b.lay, out);
    XObj y = make(out, a.cplx || b.cplx);
    if (out.nelem > 0) {
        op.a = a;
        op.b = b;
        op.out = y;
        prods.push_back(std::move(op));
        rec->push_back(Op{0, (int)prods.size() - 1});
    }
    return y;
}

XObj BlockTape::redop(const XObj& A, int dir) {
    ProdOp op;
    Layout out;
    op.prog = build_redop(H0.data(), d, A.lay, dir, out);
    XObj y = make(out, false);
    if (out.nelem > 0) {
        op.a = A;
        op.b = A;
        op.out = y;
        prods.push_back(std::move(op));
        rec->push_back(Op{0, (int)prods.size() - 1});
    }
    return y;
}

XObj BlockTape::comb(const std::vector<std::vector<Src>>& rows, const Layout* fixed, const XObj* dest, bool force_cplx) {
    const size_t nl = fixed ? fixed->l.size() : rows.size();
    if (rows.size() != nl) throw InvalidArg("heff: combination rows do not match the output entries");
    bool cp = force_cplx;
    for (auto& r : rows)
        for (auto& s : r) {
            if (s.obj->cplx) cp = true;
            const CoefSpec& c = coefs[s.coef];
            if (c.pow != 0 || c.v.imag() != 0) cp = true;
        }
    Layout lay;
    if (fixed) {
        lay = *fixed;
    } else {
        std::vector<std::map<Key, Blk>> ents(nl);
        for (size_t e = 0; e < nl; ++e)
            for (auto& s : rows[e])
                for (const Blk& b : s.obj->lay.l[s.entry].b) {
                    auto it = ents[e].find(b.qin);
                    if (it == ents[e].end()) ents[e].emplace(b.qin, b);
                    else if (it->second.qout != b.qout || it->second.rows != b.rows || it->second.cols != b.cols)
                        throw InvalidArg("heff: sum of blocks with different shapes or outgoing charges");
                }
        lay = from_maps(nsym, per, ents);
    }
    XObj y;
    if (dest) {
        if (!same_layout(dest->lay, lay) || dest->cplx != cp) throw InvalidArg("heff: destination does not match the combination");
        y = *dest;
    } else {
        y = make(lay, cp);
    }
    if (lay.nelem == 0) return y;
    CombOp op;
    op.out_ps = y.ps;
    op.out_cplx = cp;
    for (size_t e = 0; e < nl; ++e)
        for (const Blk& ob : lay.l[e].b) {
            std::vector<CombTerm> ts;
            for (auto& s : rows[e]) {
                const Blk* b = s.obj->lay.l[s.entry].find(ob.qin);
                if (!b) continue;
                if (b->qout != ob.qout || b->rows != ob.rows || b->cols != ob.cols) {
                    if (fixed) continue;  // restriction: a block of another shape is dropped
                    throw InvalidArg("heff: block shapes disagree in a combination");
                }
                ts.push_back(CombTerm{s.obj->off + b->off, s.obj->cplx ? s.obj->ps : 0, s.coef, 0});
            }
            for (long e0 = 0; e0 < ob.size(); e0 += kCombTile) {
                CombTile t;
                t.out_off = y.off + ob.off + e0;
                t.len = (int)std::min<long>(kCombTile, ob.size() - e0);
                t.t0 = (int)op.terms.size();
                t.nt = (int)ts.size();
                t.pad = 0;
                for (CombTerm c : ts) {
                    c.src_off += e0;
                    op.terms.push_back(c);
                }
                op.tiles.push_back(t);
            }
        }
    combs.push_back(std::move(op));
    rec->push_back(Op{1, (int)combs.size() - 1});
    return y;
}

XObj BlockTape::add(const XObj& a, const XObj& b, cplx ca, cplx cb, int powb) {
    if (a.lay.l.size() != b.lay.l.size()) throw InvalidArg("heff: sum of objects with different list lengths");
    std::vector<std::vector<Src>> rows(a.lay.l.size());
    const int c1 = coef(ca, 0), c2 = coef(cb, powb);
    for (size_t e = 0; e < rows.size(); ++e) rows[e] = {Src{c1, &a, (int)e}, Src{c2, &b, (int)e}};
    return comb(rows);
}

XObj BlockTape::scale(const XObj& a, cplx c, int pow) {
    if (pow == 0 && c == cplx(1.0) && a.cplx) return a;
    std::vector<std::vector<Src>> rows(a.lay.l.size());
    const int k = coef(c, pow);
    for (size_t e = 0; e < rows.size(); ++e) rows[e] = {Src{k, &a, (int)e}};
    return comb(rows, nullptr, nullptr, true);
}

XObj BlockTape::onto(const XObj& a, const Layout& lay, const XObj* dest) {
    if (!dest && same_layout(a.lay, lay)) return a;
    if (a.lay.l.size() != lay.l.size()) throw InvalidArg("heff: restriction onto a layout with another list length");
    std::vector<std::vector<Src>> rows(lay.l.size());
    const int k = coef(1.0, 0);
    for (size_t e = 0; e < rows.size(); ++e) rows[e] = {Src{k, &a, (int)e}};
    return comb(rows, &lay, dest, dest ? dest->cplx : false);
}

XObj BlockTape::diag_part(const XObj& a) {
    std::vector<std::map<Key, Blk>> ents(a.lay.l.size());
    for (size_t e = 0; e < ents.size(); ++e)
        for (const Blk& b : a.lay.l[e].b)
            if (b.qin == b.qout) ents[e].emplace(b.qin, b);
    return onto(a, from_maps(nsym, per, ents));
}

XObj BlockTape::apply_operator(const XObj& AA) {
    const int D = d * d;
    if ((int)AA.lay.l.size() != D) throw InvalidArg("heff: ApplyOperator needs a two-site object");
    std::vector<std::vector<Src>> rows(D);
    for (int r = 0; r < D; ++r)
        for (int c = 0; c < D; ++c) {
            const double h = H0[r + (size_t)c * D];
            if (h != 0.0) rows[r].push_back(Src{coef(h, 0), &AA, c});
        }
    return comb(rows);
}

XObj HeffEngine::redop_left_mixed(const XObj& Abra, const XObj& Bket) {
    // M[s][j] = sum_{k,i} H(k d + s, i d + j) Abra[k]^T Bket[i]
    const int D = d * d;
    std::vector<XObj> T;
    for (int k = 0; k < d; ++k) T.push_back(prod(PROD_MATT_MPS, view_entry(Abra, k), Bket));
    std::vector<std::vector<Src>> rows(D);
    for (int s = 0; s < d; ++s)
        for (int j = 0; j < d; ++j)
            for (int k = 0; k < d; ++k)
                for (int i = 0; i < d; ++i) {
                    const double h = H0[(k * d + s) + (size_t)(i * d + j) * D];
                    if (h != 0.0) rows[s * d + j].push_back(Src{coef(h, 0), &T[k], i});
                }
    return comb(rows);
}

XObj HeffEngine::redop_right_mixed(const XObj& Aket, const XObj& Bbra) {
    // M[s][i] = sum_{k,j} H(s d + k, i d + j) Aket[j] Bbra[k]^T
    const int D = d * d;
    std::vector<XObj> U;
    for (int k = 0; k < d; ++k) U.push_back(prod(PROD_MPS_MATT, Aket, view_entry(Bbra, k)));
    std::vector<std::vector<Src>> rows(D);
    for (int s = 0; s < d; ++s)
        for (int i = 0; i < d; ++i)
            for (int k = 0; k < d; ++k)
                for (int j = 0; j < d; ++j) {
                    const double h = H0[(s * d + k) + (size_t)(i * d + j) * D];
                    if (h != 0.0) rows[s * d + i].push_back(Src{coef(h, 0), &U[k], j});
                }
    return comb(rows);
}

XObj HeffEngine::invert_e(const std::vector<XObj>& As, const std::vector<XObj>& Bs, const XObj& x, int dir, int pow, bool proj) {
    auto g = std::make_unique<GeoOp>();
    g->dir = dir;
    g->pow = pow;
    g->x = x;
    g->y = make(x.lay, true);
    Layout in = x.lay;
    for (int k = 0; k < N; ++k) {
        const int idx = dir == IMPS_LEFT ? k : N - 1 - k;
        const XObj &A = As[idx], &B = Bs[idx];
        Layout out;
        if (k == N - 1) {
            out = x.lay;
        } else {
            Layout l1, l2;
            if (dir == IMPS_LEFT) {
                build_prod(PROD_MAT_MPS, in, A.lay, l1);
                build_prod(PROD_TM_LEFT, l1, B.lay, l2);
            } else {
                build_prod(PROD_MPS_MAT, A.lay, in, l1);
                build_prod(PROD_TM_RIGHT, l1, B.lay, l2);
            }
            out = l2;
            g->bond_elems.push_back(out.nelem);
        }
        for (int lo = 0; lo < (prec == IMPS_FP64 ? 1 : 2); ++lo) {
            std::unique_ptr<imps_plan_s> P(new_plan(ctx, K_TM, lo ? prec : IMPS_FP64));
            P->prog = build_tm(dir, A.lay, &B.lay, in, out);
            P->op_parts = {A.lay.nelem, B.lay.nelem};
            upload_program(P.get());
            (lo ? g->chain_lo : g->chain).push_back(std::move(P));
        }
        g->chain_ops.push_back({A, B});
        in = out;
    }
    if (proj) {
        g->proj = true;
        g->w = make(x.lay, false);
        g->rout = make(x.lay, false);
    }
    geos.push_back(std::move(g));
    rec->push_back(Op{2, (int)geos.size() - 1});
    return geos.back()->y;
}

void HeffEngine::build(imps_context c, int ncell, int dphys, const std::vector<Layout>& al, const std::vector<Layout>& ar,
                       const std::vector<Layout>& nl, const std::vector<Layout>& cl, const double* H, double E0, const int* Kq, int nmomenta,
                       int precision, int sites) {
    if (!c) throw InvalidArg("heff: a device context is required");
    if (sites != 1 && sites != 2) throw InvalidArg("heff: the operator acts on one or two sites");
    prec = precision;
    op_sites = sites;
    if (ncell < 1 || dphys < 1) throw InvalidArg("heff: bad cell size or physical dimension");
    if ((int)al.size() != ncell || (int)ar.size() != ncell || (int)nl.size() != ncell || (int)cl.size() != ncell)
        throw InvalidArg("heff: one layout per site is needed for AL, AR, NL and C");
    if (nmomenta < 1 || nmomenta > 4096) throw InvalidArg("heff: nmom must be 1..4096");
    if (!H) throw InvalidArg("heff: H is NULL");
    ctx = c;
    if (const char* e = std::getenv("IMPS_HEFF_GMRES_M")) krylov = std::max(2, std::min(1000, std::atoi(e)));
    N = ncell;
    ncell_phase = ncell;
    d = dphys;
    nmom = nmomenta;
    nsym = al[0].nsym;
    for (int i = 0; i < kMaxSym; ++i) per[i] = al[0].per[i];
    for (int n = 0; n < N; ++n) {
        if ((int)al[n].l.size() != d || (int)ar[n].l.size() != d || (int)nl[n].l.size() != d)
            throw InvalidArg("heff: MPS layouts need nlist = d");
        if (cl[n].l.size() != 1) throw InvalidArg("heff: C layouts need nlist = 1");
        require_same_nsym(al[0], al[n], "heff"); require_same_nsym(al[0], ar[n], "heff");
        require_same_nsym(al[0], nl[n], "heff"); require_same_nsym(al[0], cl[n], "heff");
    }
    ALl = al;
    ARl = ar;
    NLl = nl;
    Cl = cl;
    {
        // the excitation's charge, as a residue for Z(N) charges
        int kq[kMaxSym] = {0, 0, 0, 0};
        for (int i = 0; i < nsym; ++i) kq[i] = kred(Kq[i], per[i]);
        K = Key(kq, nsym, per);
    }
    const int D = op_sites == 2 ? d * d : d;
    H0.assign(H, H + (size_t)D * D);
    for (int i = 0; i < D; ++i) H0[i + (size_t)i * D] -= E0;
    if (op_sites == 1) {
        // a one-site observable must keep the block structure: O(r, c) != 0 only between physical
        // indices whose blocks connect the same charges
        auto keys = [](const Layout& L, int e) {
            std::vector<std::pair<Key, Key>> k;
            for (const Blk& b : L.l[e].b) k.push_back({b.qin, b.qout});
            return k;
        };
        for (int r = 0; r < d; ++r)
            for (int c2 = 0; c2 < d; ++c2)
                if (r != c2 && H0[r + (size_t)c2 * d] != 0.0)
                    for (int n = 0; n < N; ++n)
                        if (keys(al[n], r) != keys(al[n], c2) || keys(ar[n], r) != keys(ar[n], c2))
                            throw InvalidArg("heff: the observable couples physical indices of different charge (it must conserve the charges)");
    }
    // XDims (Excitations_helpers.hpp:639-678): X[n] blocks Qout(NL[n]) -> + K with sizes from NL's
    // columns and AR[n+1]'s rows
    xdims.clear();
    mtot = 0;
    for (int n = 0; n < N; ++n) {
        std::map<Key, int> ml, mr;
        for (auto& bl : NLl[n].l)
            for (auto& b : bl.b) ml[b.qout] = b.cols;
        for (auto& bl : ARl[(n + 1) % N].l)
            for (auto& b : bl.b) mr[b.qin] = b.rows;
        std::vector<std::map<Key, Blk>> ent(1);
        for (auto& kv : ml) {
            const Key qk = kv.first + K;
            auto it = mr.find(qk);
            if (it != mr.end() && kv.second > 0 && it->second > 0) ent[0].emplace(kv.first, Blk{kv.first, qk, kv.second, it->second, 0});
        }
        Layout L = from_maps(nsym, per, ent);
        mtot += L.nelem;
        xdims.push_back(L);
    }
    if (mtot == 0) throw InvalidArg("heff: no excitations for this quantum number (empty XDims)");
    // constants
    for (int n = 0; n < N; ++n) {
        AL.push_back(make(ALl[n], false));
        AR.push_back(make(ARl[n], false));
        NL.push_back(make(NLl[n], false));
        C.push_back(make(Cl[n], false));
    }
    // contiguous excitation vectors (XDims order), one view per site
    {
        std::vector<std::map<Key, Blk>> ents(N);
        Layout all;
        set_group(all, al[0]);
        all.l.resize(N);
        long pos = 0;
        for (int n = 0; n < N; ++n)
            for (Blk b : xdims[n].l[0].b) {
                b.off = pos;
                pos += b.size();
                all.l[n].b.push_back(b);
            }
        all.nelem = pos;
        Xin = make(all, true);
        Xout = make(all, true);
        for (int n = 0; n < N; ++n) {
            Xin_v.push_back(view_entry(Xin, n));
            Xout_v.push_back(view_entry(Xout, n));
        }
    }
    bL = bond_layout(ALl[N - 1], true);
    bR = bond_layout(ARl[0], false);
    lmL = cross_layout(bR, bL);
    rmL = cross_layout(bL, bR);
    coef(1.0, 0);
    rec = &setup_tape;
    record_setup();
    rec = &apply_tape;
    if (op_sites == 2) record_apply();
    else record_apply_observable();
    rec = nullptr;
    finalize();
    fac_dev.ensure(sizeof(double) * 2 * nmom * std::max<size_t>(geos.size(), 1));
    gam_dev.ensure(sizeof(double) * 2 * nmom * std::max<size_t>(geos.size(), 1));
    hip_check(hipStreamSynchronize(stream()), "sync");
}

void HeffEngine::record_setup() {
    // HeffConstants (Excitations_helpers.hpp:260-347) with E0 already subtracted from H, or, for a
    // one-site observable, OPeffConstants (:531-600) with its expectation subtracted: the same
    // environment solves on the one-site sums OPL / OPR
    auto P = [this](int x) { return ((x % N) + N) % N; };
    std::vector<XObj> HL(N), HR(N);
    if (op_sites == 2) {
        auto opl = [this](const XObj& AA) { return diag_part(prod(PROD_TM_LEFT, apply_operator(AA), AA)); };
        auto opr = [this](const XObj& AA) { return diag_part(prod(PROD_TM_RIGHT, apply_operator(AA), AA)); };
        HL[0] = opl(prod(PROD_MPS_MPS, AL[N - 1], AL[0]));
        for (int n = 1; n < N; ++n) HL[n] = add(opl(prod(PROD_MPS_MPS, AL[n - 1], AL[n])), diag_part(tml(AL[n], AL[n], HL[n - 1])));
        HR[N - 1] = opr(prod(PROD_MPS_MPS, AR[N - 1], AR[0]));
        for (int n = N - 2; n >= 0; --n) HR[n] = add(opr(prod(PROD_MPS_MPS, AR[n], AR[n + 1])), diag_part(tmr(AR[n], AR[n], HR[n + 1])));
    } else {
        auto opl = [this](const XObj& A1) { return diag_part(prod(PROD_TM_LEFT, apply_site_operator(A1), A1)); };
        auto opr = [this](const XObj& A1) { return diag_part(prod(PROD_TM_RIGHT, apply_site_operator(A1), A1)); };
        HL[0] = opl(AL[0]);
        for (int n = 1; n < N; ++n) HL[n] = add(opl(AL[n]), diag_part(tml(AL[n], AL[n], HL[n - 1])));
        HR[N - 1] = opr(AR[N - 1]);
        for (int n = N - 2; n >= 0; --n) HR[n] = add(opr(AR[n]), diag_part(tmr(AR[n], AR[n], HR[n + 1])));
    }
    // bond objects (filled by set_state): identities, R = C C^T on AL's side, L on AR's side, and the
    // dominant eigenmatrices of the mixed TM
    oneL = make(bL, false);
    oneR = make(bR, false);
    Rb = make(bL, false);
    Lb = make(bR, false);
    LMo = make(lmL, false);
    RMo = make(rmL, false);
    const XObj xl = onto(HL[N - 1], bL), xr = onto(HR[0], bR);
    const XObj yl = make(bL, false), yr = make(bR, false);
    envL = std::make_unique<EnvSolver>();
    envL->build(ctx, IMPS_LEFT, ALl, bL, IMPS_FP64);
    envR = std::make_unique<EnvSolver>();
    envR->build(ctx, IMPS_RIGHT, ARl, bR, IMPS_FP64);
    checkHL = HL[N - 1];
    checkHR = HR[0];
    checkR = Rb;
    checkL = Lb;
    calls.push_back([this, xl, yl]() {
        KrylovStats st = envL->solve(A() + xl.off, A() + oneL.off, A() + Rb.off, nullptr, A() + yl.off, inv_tol, 20000, 100);
        env_iters += st.iters;
    });
    rec->push_back(Op{3, (int)calls.size() - 1});
    calls.push_back([this, xr, yr]() {
        KrylovStats st = envR->solve(A() + xr.off, A() + Lb.off, A() + oneR.off, nullptr, A() + yr.off, inv_tol, 20000, 100);
        env_iters += st.iters;
    });
    rec->push_back(Op{3, (int)calls.size() - 1});
    std::vector<XObj> EHL(N), EHR(N);
    HLtot.assign(N, XObj());
    HRtot.assign(N, XObj());
    EHL[N - 1] = yl;
    HLtot[N - 1] = yl;
    for (int n = 0; n < N - 1; ++n) {
        EHL[n] = tml(AL[n], AL[n], EHL[P(n - 1)]);
        HLtot[n] = add(HL[n], EHL[n]);
    }
    EHR[0] = yr;
    HRtot[0] = yr;
    for (int n = N - 1; n > 0; --n) {
        EHR[n] = tmr(AR[n], AR[n], EHR[P(n + 1)]);
        HRtot[n] = add(EHR[n], HR[n]);
    }
    HLr.clear();
    HRr.clear();
    AAL.clear();
    if (op_sites == 1) return;
    for (int n = 0; n < N; ++n) {
        HLr.push_back(redop(AL[n], IMPS_LEFT));
        HRr.push_back(redop(AR[n], IMPS_RIGHT));
    }
    for (int n = 0; n < N; ++n) AAL.push_back(prod(PROD_MPS_MPS, AL[P(n - 1)], AL[n]));
}

XObj HeffEngine::apply_site_operator(const XObj& A1) {
    if ((int)A1.lay.l.size() != d) throw InvalidArg("heff: a one-site operator needs a one-site object");
    std::vector<std::vector<Src>> rows(d);
    for (int r = 0; r < d; ++r)
        for (int c2 = 0; c2 < d; ++c2) {
            const double o = H0[r + (size_t)c2 * d];
            if (o != 0.0) rows[r].push_back(Src{coef(o, 0), &A1, c2});
        }
    return comb(rows);
}

void HeffEngine::record_apply_observable() {
    // ApplyOPeff (Excitations_helpers.hpp:350-528) for B[n] = NL[n] X[n] and the one-site operator
    // H0 (expectation subtracted); OPLtot / OPRtot are in HLtot / HRtot.  The reference skips the
    // left geometric sum when the norm of its source is below 10 InvETol; here it always runs (a
    // zero source gives a zero solution at once), which changes the result by less than that.
    auto P = [this](int x) { return ((x % N) + N) % N; };
    std::vector<XObj> Bin(N);
    for (int n = 0; n < N; ++n) Bin[n] = prod(PROD_MPS_MAT, NL[n], Xin_v[n]);
    // terms where Bin is right of Bout (as in ApplyHeff)
    std::vector<XObj> ABR(N);
    ABR[N - 1] = prod(PROD_TM_RIGHT, Bin[N - 1], AR[N - 1]);
    for (int n = N - 1; n > 0; --n) ABR[n - 1] = prod(PROD_TM_RIGHT, add(Bin[n - 1], prod(PROD_MPS_MAT, AL[n - 1], ABR[n])), AR[n - 1]);
    XObj EBR = invert_e(AL, AR, ABR[0], IMPS_RIGHT, +1, true);
    std::vector<XObj> ABRtot(N);
    ABRtot[0] = scale(EBR, 1.0, +1);
    for (int n = N - 1; n > 0; --n) {
        EBR = tmr(AL[n], AR[n], EBR);
        ABRtot[n] = add(ABR[n], EBR, 1.0, 1.0, +1);
    }
    // terms where Bin is left of Bout: OPBL and its geometric sum
    std::vector<XObj> OPBL(N);
    OPBL[0] = add(prod(PROD_TM_LEFT, prod(PROD_MAT_MPS, HLtot[N - 1], Bin[0]), AL[0]), prod(PROD_TM_LEFT, apply_site_operator(Bin[0]), AL[0]));
    for (int n = 1; n < N; ++n)
        OPBL[n] = add(prod(PROD_TM_LEFT, add(prod(PROD_MAT_MPS, OPBL[n - 1], AR[n]), prod(PROD_MAT_MPS, HLtot[n - 1], Bin[n])), AL[n]),
                      prod(PROD_TM_LEFT, apply_site_operator(Bin[n]), AL[n]));
    XObj EOPBL = invert_e(AR, AL, OPBL[N - 1], IMPS_LEFT, -1, true);
    // Bin and Bout on the same site
    std::vector<XObj> Bout(N);
    for (int n = 0; n < N; ++n) {
        XObj b = add(prod(PROD_MAT_MPS, HLtot[P(n - 1)], Bin[n]), prod(PROD_MPS_MAT, Bin[n], HRtot[P(n + 1)]));
        Bout[n] = add(b, apply_site_operator(Bin[n]));
    }
    // Bin left of Bout
    Bout[0] = add(Bout[0], prod(PROD_MAT_MPS, scale(EOPBL, 1.0, -1), AR[0]));
    for (int n = 1; n < N; ++n) {
        EOPBL = tml(AR[n - 1], AL[n - 1], EOPBL);
        Bout[n] = add(Bout[n], prod(PROD_MAT_MPS, add(OPBL[n - 1], EOPBL, 1.0, 1.0, -1), AR[n]));
    }
    // Bin right of Bout
    for (int n = 0; n < N; ++n) {
        Bout[n] = add(Bout[n], prod(PROD_MAT_MPS, HLtot[P(n - 1)], prod(PROD_MPS_MAT, AL[n], ABRtot[P(n + 1)])));
        Bout[n] = add(Bout[n], prod(PROD_MPS_MAT, apply_site_operator(AL[n]), ABRtot[P(n + 1)]));
    }
    for (int n = 0; n < N; ++n) onto(prod(PROD_TM_LEFT, Bout[n], NL[n]), Xin_v[n].lay, &Xout_v[n]);
}

void HeffEngine::record_apply() {
    // ApplyHeff (Excitations_helpers.hpp:20-228) for B[n] = NL[n] X[n]; kfac = e^{i p pi N} per
    // momentum (coefficient power +1), ck = conj(kfac) (power -1)
    auto P = [this](int x) { return ((x % N) + N) % N; };
    std::vector<XObj> Bin(N);
    for (int n = 0; n < N; ++n) Bin[n] = prod(PROD_MPS_MAT, NL[n], Xin_v[n]);
    // terms where Bin is right of Bout
    std::vector<XObj> ABR(N);
    ABR[N - 1] = prod(PROD_TM_RIGHT, Bin[N - 1], AR[N - 1]);
    for (int n = N - 1; n > 0; --n) ABR[n - 1] = prod(PROD_TM_RIGHT, add(Bin[n - 1], prod(PROD_MPS_MAT, AL[n - 1], ABR[n])), AR[n - 1]);
    XObj EBR = invert_e(AL, AR, ABR[0], IMPS_RIGHT, +1, true);
    std::vector<XObj> ABRtot(N);
    ABRtot[0] = scale(EBR, 1.0, +1);
    for (int n = N - 1; n > 0; --n) {
        EBR = tmr(AL[n], AR[n], EBR);
        ABRtot[n] = add(ABR[n], EBR, 1.0, 1.0, +1);
    }
    // terms where Bin is left of Bout
    std::vector<XObj> HBL(N);
    HBL[0] = add(prod(PROD_TM_LEFT, prod(PROD_MAT_MPS, HLtot[N - 1], Bin[0]), AL[0]),
                 prod(PROD_TM_LEFT, apply_operator(add(prod(PROD_MPS_MPS, AL[N - 1], Bin[0]), prod(PROD_MPS_MPS, Bin[N - 1], AR[0]), 1.0, 1.0, -1)),
                      AAL[0]));
    for (int n = 1; n < N; ++n)
        HBL[n] = add(prod(PROD_TM_LEFT, add(prod(PROD_MAT_MPS, HBL[n - 1], AR[n]), prod(PROD_MAT_MPS, HLtot[n - 1], Bin[n])), AL[n]),
                     prod(PROD_TM_LEFT, apply_operator(add(prod(PROD_MPS_MPS, AL[n - 1], Bin[n]), prod(PROD_MPS_MPS, Bin[n - 1], AR[n]))), AAL[n]));
    XObj EHBL = invert_e(AR, AL, HBL[N - 1], IMPS_LEFT, -1, true);
    std::vector<XObj> HBLtot(N);
    HBLtot[N - 1] = scale(EHBL, 1.0, -1);
    for (int n = 0; n < N - 1; ++n) {
        EHBL = tml(AR[n], AL[n], EHBL);
        HBLtot[n] = add(HBL[n], EHBL, 1.0, 1.0, -1);
    }
    // Bin and Bout on top of each other
    std::vector<XObj> Bout(N);
    for (int n = 0; n < N; ++n) {
        XObj b = add(prod(PROD_MAT_MPS, HLtot[P(n - 1)], Bin[n]), prod(PROD_MPS_MAT, Bin[n], HRtot[P(n + 1)]));
        b = add(b, prod(PROD_REDOP_MPS, HLr[P(n - 1)], Bin[n]));
        Bout[n] = add(b, prod(PROD_MPS_REDOP, HRr[P(n + 1)], Bin[n]));
    }
    // Bin left of Bout
    Bout[0] = add(Bout[0], prod(PROD_MAT_MPS, HBLtot[N - 1], AR[0]));
    Bout[0] = add(Bout[0], prod(PROD_REDOP_MPS, redop_left_mixed(AL[N - 1], Bin[N - 1]), AR[0]), 1.0, 1.0, -1);
    for (int n = 1; n < N; ++n) {
        Bout[n] = add(Bout[n], prod(PROD_MAT_MPS, HBLtot[n - 1], AR[n]));
        Bout[n] = add(Bout[n], prod(PROD_REDOP_MPS, redop_left_mixed(AL[n - 1], Bin[n - 1]), AR[n]));
    }
    // Bin right of Bout
    for (int n = 0; n < N; ++n) {
        XObj Atmp = add(Bin[P(n + 1)], prod(PROD_MPS_MAT, AL[P(n + 1)], ABRtot[P(n + 2)]));
        if (n == N - 1) Atmp = scale(Atmp, 1.0, +1);
        const XObj Y = prod(PROD_MPS_MAT, AL[n], ABRtot[P(n + 1)]);
        Bout[n] = add(Bout[n], prod(PROD_MAT_MPS, HLtot[P(n - 1)], Y));
        Bout[n] = add(Bout[n], prod(PROD_REDOP_MPS, HLr[P(n - 1)], Y));
        Bout[n] = add(Bout[n], prod(PROD_MPS_REDOP, redop_right_mixed(Atmp, AR[P(n + 1)]), AL[n]));
    }
    // X_out = NL^T B_out on the structure of X
    for (int n = 0; n < N; ++n) onto(prod(PROD_TM_LEFT, Bout[n], NL[n]), Xin_v[n].lay, &Xout_v[n]);
}

// ================================================================================================
// execution
void BlockTape::finalize() {
    // device tables: programs and combinations
    {
        std::vector<char> host;
        auto put = [&](const void* p, size_t bytes) {
            size_t pos = (host.size() + 255) & ~size_t(255);
            host.resize(pos + bytes);
            if (bytes) std::memcpy(host.data() + pos, p, bytes);
            return pos;
        };
        std::vector<std::vector<std::pair<size_t, size_t>>> ppos(prods.size());
        for (size_t i = 0; i < prods.size(); ++i)
            for (auto& st : prods[i].prog.stages) {
                const size_t tp = put(st.tasks.data(), st.tasks.size() * sizeof(Task));
                const size_t sp = put(st.segs.data(), st.segs.size() * sizeof(Seg));
                ppos[i].push_back({tp, sp});
            }
        std::vector<std::pair<size_t, size_t>> cpos(combs.size());
        for (size_t i = 0; i < combs.size(); ++i) {
            const size_t tp = put(combs[i].tiles.data(), combs[i].tiles.size() * sizeof(CombTile));
            const size_t sp = put(combs[i].terms.data(), combs[i].terms.size() * sizeof(CombTerm));
            cpos[i] = {tp, sp};
        }
        progs_dev.ensure(std::max<size_t>(host.size(), 256));
        h2d_sync(stream(), progs_dev.p, host.data(), host.size());
        char* base = (char*)progs_dev.p;
        for (size_t i = 0; i < prods.size(); ++i) {
            prods[i].dst.clear();
            for (size_t k = 0; k < prods[i].prog.stages.size(); ++k) {
                DevStage ds;
                ds.tasks = (const Task*)(base + ppos[i][k].first);
                ds.segs = (const Seg*)(base + ppos[i][k].second);
                ds.ntasks = (int)prods[i].prog.stages[k].tasks.size();
                ds.out_slot = prods[i].prog.stages[k].out_slot;
                prods[i].dst.push_back(ds);
            }
        }
        for (size_t i = 0; i < combs.size(); ++i) {
            combs[i].dtiles = (const CombTile*)(base + cpos[i].first);
            combs[i].dterms = (const CombTerm*)(base + cpos[i].second);
        }
    }
    arena.ensure(sizeof(double) * (size_t)std::max<long>(arena_elems, 1));
    hip_check(launch_zero(arena.d(), arena_elems, stream()), "zero arena");
    coef_dev.ensure(sizeof(double) * 2 * coefs.size() * nmom);
}

void BlockTape::upload_coefs() {
    std::vector<double> h(2 * coefs.size() * nmom);
    for (size_t i = 0; i < coefs.size(); ++i)
        for (int t = 0; t < nmom; ++t) {
            const double pa = (t < (int)p_cur.size() ? p_cur[t] : 0.0) * M_PI * ncell_phase;
            const cplx kf(std::cos(pa), std::sin(pa));
            cplx v = coefs[i].v;
            if (coefs[i].pow == 1) v *= kf;
            else if (coefs[i].pow == -1) v *= std::conj(kf);
            h[(i * nmom + t) * 2] = v.real();
            h[(i * nmom + t) * 2 + 1] = v.imag();
        }
    h2d_sync(stream(), coef_dev.p, h.data(), h.size() * sizeof(double));
}

void BlockTape::run_prod(const ProdOp& op) {
    Slots sl;
    sl.base[SLOT_X] = A() + op.a.off;
    sl.stride[SLOT_X] = op.a.cplx ? op.a.ps : 0;
    sl.base[SLOT_OP] = A() + op.b.off;
    sl.stride[SLOT_OP] = op.b.cplx ? op.b.ps : 0;
    sl.base[SLOT_W] = nullptr;
    sl.stride[SLOT_W] = 0;
    const int nv = op.out.cplx ? 2 * nmom : 1;
    for (auto& ds : op.dst) hip_check(launch_stage(IMPS_FP64, ds, sl, A() + op.out.off, op.out.ps, true, nv, stream()), "heff product");
}

void BlockTape::run_comb(const CombOp& op) {
    hip_check(comb_launch(op.dtiles, (int)op.tiles.size(), op.dterms, coef_dev.d(), A(), op.out_ps, op.out_cplx ? 1 : 0, nmom, stream()),
              "heff combination");
}

void BlockTape::run_tape(const std::vector<Op>& tape) {
    for (const Op& op : tape) {
        if (op.type == 0) run_prod(prods[op.idx]);
        else if (op.type == 1) run_comb(combs[op.idx]);
        else run_other(op);
    }
}

void HeffEngine::run_other(const Op& op) {
    if (op.type == 2) {
        run_geo(*geos[op.idx]);
        stats_steps_total += stats.gmres_steps;
        stats_relres_worst = std::max(stats_relres_worst, stats.gmres_relres);
    } else if (op.type == 3) {
        calls[op.idx]();
    } else {
        throw InvalidArg("heff: bad tape entry");
    }
}

void HeffEngine::set_state(const std::vector<const double*>& al, const std::vector<const double*>& ar, const std::vector<const double*>& nl,
                           const std::vector<const double*>& cl, double tol, const BondInput* bond) {
    if ((int)al.size() != N || (int)ar.size() != N || (int)nl.size() != N || (int)cl.size() != N)
        throw InvalidArg("heff: one tensor per site is needed");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    hipStream_t s = stream();
    inv_tol = tol > 0 ? tol : 1e-14;
    for (int n = 0; n < N; ++n) {
        const std::pair<const double*, const XObj*> parts[4] = {{al[n], &AL[n]}, {ar[n], &AR[n]}, {nl[n], &NL[n]}, {cl[n], &C[n]}};
        for (auto& pp : parts) {
            if (pp.second->nelem() > 0 && !pp.first) throw InvalidArg("heff: tensor data pointer is NULL");
            if (pp.second->nelem() > 0)
                hip_check(hipMemcpyAsync(A() + pp.second->off, pp.first, sizeof(double) * pp.second->nelem(), hipMemcpyDeviceToDevice, s),
                          "copy state");
        }
    }
    // host-side bond objects: identities; R = C C^T (the right fixed point of AL's TM); in closed form
    // (AL C = C AR) L = C^T C and the dominant eigenmatrices LM = C^T, RM = C of the mixed TM in K = 0
    // (eigenvalue 1, tr(LM RM) = 1), else L, LM, RM as given; then the projection weights of the
    // geometric sums
    const Layout& cl0 = Cl[N - 1];
    std::vector<double> ch(cl0.nelem);
    d2h_sync(s, ch.data(), A() + C[N - 1].off, sizeof(double) * cl0.nelem);
    auto eye = [](const Layout& b) {
        std::vector<double> id(b.nelem, 0.0);
        for (const Blk& x : b.l[0].b)
            for (int i = 0; i < x.rows; ++i) id[x.off + i + (long)i * x.rows] = 1.0;
        return id;
    };
    std::vector<double> rh(bL.nelem, 0.0), lh(bR.nelem, 0.0), lm(lmL.nelem, 0.0), rm(rmL.nelem, 0.0);
    for (const Blk& b : cl0.l[0].b) {
        if (b.qin != b.qout) throw InvalidArg("heff: C must be block diagonal");
        const double* c = ch.data() + b.off;  // rows x cols, column-major
        const Blk* tb = bL.l[0].find(b.qin);
        if (!tb || tb->rows != b.rows) throw InvalidArg("heff: a block of C does not match the right bond of AL");
        for (int j = 0; j < b.rows; ++j)
            for (int i = 0; i < b.rows; ++i) {
                double a = 0;
                for (int k = 0; k < b.cols; ++k) a += c[i + (long)k * b.rows] * c[j + (long)k * b.rows];
                rh[tb->off + i + (long)j * b.rows] = a;
            }
        if (bond) continue;
        const Blk* ub = bR.l[0].find(b.qin);
        const Blk* lb = lmL.l[0].find(b.qin);
        const Blk* rb = rmL.l[0].find(b.qin);
        if (!ub || ub->rows != b.cols || !lb || !rb) throw InvalidArg("heff: a block of C does not match the left bond of AR");
        for (int j = 0; j < b.cols; ++j)
            for (int i = 0; i < b.cols; ++i) {
                double a = 0;
                for (int k = 0; k < b.rows; ++k) a += c[k + (long)i * b.rows] * c[k + (long)j * b.rows];
                lh[ub->off + i + (long)j * b.cols] = a;
            }
        for (int j = 0; j < b.cols; ++j)
            for (int i = 0; i < b.rows; ++i) {
                rm[rb->off + i + (long)j * b.rows] = c[i + (long)j * b.rows];
                lm[lb->off + j + (long)i * b.cols] = c[i + (long)j * b.rows];
            }
    }
    if (bond) {
        auto fetch = [&](const Layout& l, const double* p, const Layout& dl, const char* what) {
            std::vector<double> h(l.nelem);
            if (l.nelem > 0 && !p) throw InvalidArg(std::string("heff: ") + what + " data pointer is NULL");
            d2h_sync(s, h.data(), p, sizeof(double) * l.nelem);
            return map_blocks(l, h, dl, what);
        };
        lh = fetch(bond->Llay, bond->L, bR, "L");
        // LM and RM project the geometric sums in K = 0 only (Excitations_main.cpp:264-268); for
        // another charge they are not used
        bool zeroK = true;
        for (int i = 0; i < nsym; ++i) zeroK = zeroK && K.q[i] == 0;
        if (zeroK) {
            lm = fetch(bond->LMlay, bond->LM, lmL, "LM");
            rm = fetch(bond->RMlay, bond->RM, rmL, "RM");
        }
    }
    h2d_sync(s, A() + oneL.off, eye(bL).data(), sizeof(double) * bL.nelem);
    h2d_sync(s, A() + oneR.off, eye(bR).data(), sizeof(double) * bR.nelem);
    h2d_sync(s, A() + Rb.off, rh.data(), sizeof(double) * rh.size());
    h2d_sync(s, A() + Lb.off, lh.data(), sizeof(double) * lh.size());
    h2d_sync(s, A() + LMo.off, lm.data(), sizeof(double) * lm.size());
    h2d_sync(s, A() + RMo.off, rm.data(), sizeof(double) * rm.size());
    for (auto& gp : geos) {
        GeoOp& g = *gp;
        for (size_t k = 0; k < g.chain.size(); ++k) {
            set_operator(g.chain[k].get(), {A() + g.chain_ops[k].first.off, A() + g.chain_ops[k].second.off});
            if (!g.chain_lo.empty()) set_operator(g.chain_lo[k].get(), {A() + g.chain_ops[k].first.off, A() + g.chain_ops[k].second.off});
        }
        if (!g.proj) continue;
        // both sums act on objects between AL's bond (rows) and AR's (columns); on x's sectors,
        // lmt = LM^T and rmx = RM.  The right sum is InvertE_proj_fac(AL, AR, x, LM, RM, r): trace
        // tr(LM x) (weights LM^T), rho_out = RM; the left sum InvertE_proj_fac(AR, AL, x, LM^T, RM^T, l)
        // (Excitations_helpers.hpp:83, :124): trace tr(x RM^T) (weights RM), rho_out = LM^T
        // (InvEBlockFunctions.hpp:268-331)
        std::vector<double> lmt(g.x.nelem(), 0.0), rmx(g.x.nelem(), 0.0);
        for (const Blk& xb : g.x.lay.l[0].b) {
            if (const Blk* rb = lmL.l[0].find(xb.qout))
                if (rb->qout == xb.qin && rb->rows == xb.cols && rb->cols == xb.rows)
                    for (int j = 0; j < xb.cols; ++j)
                        for (int i = 0; i < xb.rows; ++i) lmt[xb.off + i + (long)j * xb.rows] = lm[rb->off + j + (long)i * rb->rows];
            if (const Blk* ob = rmL.l[0].find(xb.qin))
                if (ob->qout == xb.qout && ob->rows == xb.rows && ob->cols == xb.cols)
                    for (long e = 0; e < xb.size(); ++e) rmx[xb.off + e] = rm[ob->off + e];
        }
        const bool right = g.dir == IMPS_RIGHT;
        h2d_sync(s, A() + g.w.off, (right ? lmt : rmx).data(), sizeof(double) * lmt.size());
        h2d_sync(s, A() + g.rout.off, (right ? rmx : lmt).data(), sizeof(double) * lmt.size());
    }
    envL->set_mps([&] { std::vector<const double*> v; for (auto& o : AL) v.push_back(A() + o.off); return v; }());
    envR->set_mps([&] { std::vector<const double*> v; for (auto& o : AR) v.push_back(A() + o.off); return v; }());
    p_cur.assign(nmom, 0.0);
    upload_coefs();
    env_iters = 0;
    run_tape(setup_tape);
    hip_check(hipStreamSynchronize(s), "sync");
    // HeffConstants' check (Excitations_helpers.hpp:287-291, :299-303): tr(HL R), tr(L HR)
    auto host = [&](const XObj& o) {
        std::vector<double> h(o.nelem());
        d2h_sync(s, h.data(), A() + o.off, sizeof(double) * h.size());
        return h;
    };
    auto trace = [](const XObj& a, const std::vector<double>& ha, const XObj& b, const std::vector<double>& hb) {
        double t = 0;
        for (size_t e = 0; e < a.lay.l.size(); ++e)
            for (const Blk& x : a.lay.l[e].b) {
                const Blk* y = b.lay.l[e].find(x.qout);
                if (!y || y->qout != x.qin) continue;
                for (int j = 0; j < x.cols; ++j)
                    for (int i = 0; i < x.rows; ++i) t += ha[x.off + i + (long)j * x.rows] * hb[y->off + j + (long)i * y->rows];
            }
        return t;
    };
    check_trL = trace(checkHL, host(checkHL), checkR, host(checkR));
    check_trR = trace(checkL, host(checkL), checkHR, host(checkHR));
    state_set = true;
}

// ---- momentum-batched complex GMRES for one geometric sum ---------------------------------------
void HeffEngine::run_geo(GeoOp& g) {
    hipStream_t s = stream();
    const long n = g.x.nelem();
    if (n == 0) return;
    const int nv = 2 * nmom;
    const long vl = (long)nv * n;  // one batched vector
    const int m = std::max(2, std::min<int>(krylov, (int)std::min<long>(n, 1000)));
    const int nch = cblas_chunks(n), gch = cblas_gemv_grid(n, nmom);
    // scratch layout
    long tot_bond = 0;
    for (long e : g.bond_elems) tot_bond += e;
    const long cst = 4L * (m + 1) + 1;       // per momentum per step: h (pass 1), h (pass 2), ||w||^2
    const long sst = cst * nmom;             // per step
    const size_t need = (size_t)vl * (m + 1) + 2 * (size_t)vl + (size_t)tot_bond * nv + (size_t)nch * (m + 1) * nmom * 2 +
                        (size_t)gch * nmom + (size_t)sst * (m + 1) + 4 * (size_t)nmom * (m + 1) + 64;
    kry.ensure(sizeof(double) * need);
    double* V = kry.d();
    double* Tb = V + (size_t)vl * (m + 1);
    double* bb = Tb + vl;
    double* bonds = bb + vl;
    double* part = bonds + (size_t)tot_bond * nv;
    double* qpart = part + (size_t)nch * (m + 1) * nmom * 2;
    double* cbuf = qpart + (size_t)gch * nmom;
    double* ycoef = cbuf + (size_t)sst * (m + 1);
    // per-momentum factors: fac = kfac^pow, projection flag
    std::vector<double> fac(2 * nmom), gam(nmom), gneg(nmom), ones(2 * nmom, 0.0);
    for (int t = 0; t < nmom; ++t) {
        const double pa = p_cur[t] * M_PI * N;
        const cplx kf = g.pow > 0 ? cplx(std::cos(pa), std::sin(pa)) : cplx(std::cos(pa), -std::sin(pa));
        fac[2 * t] = kf.real();
        fac[2 * t + 1] = kf.imag();
        bool zeroK = true;
        for (int i = 0; i < nsym; ++i) zeroK = zeroK && K.q[i] == 0;
        // Excitations_main.cpp:264-268 projects when K = 0 and |p_act| < 10 InvETol; here whenever
        // e^{i p_act} = 1 (p_act = 0 mod 2 pi), where 1 - fac T is singular in the same way
        const bool pr = g.proj && zeroK && std::abs(kf - cplx(1.0)) < 10 * inv_tol;
        gam[t] = pr ? 1.0 : 0.0;
        gneg[t] = -gam[t];
        ones[2 * t] = 1.0;
    }
    double* dfac = fac_dev.d();
    double* dgam = gam_dev.d();
    h2d_sync(s, dfac, fac.data(), sizeof(double) * 2 * nmom);
    std::vector<double> gh(4 * nmom);
    for (int t = 0; t < nmom; ++t) {
        gh[t] = gam[t];
        gh[nmom + t] = gneg[t];
        gh[2 * nmom + 2 * t] = 1.0;
        gh[2 * nmom + 2 * t + 1] = 0.0;
    }
    h2d_sync(s, dgam, gh.data(), sizeof(double) * 4 * nmom);
    const double* dgpos = dgam;
    const double* dgneg = dgam + nmom;
    const double* dones = dgam + 2 * nmom;
    bool anyproj = false;
    for (int t = 0; t < nmom; ++t) anyproj = anyproj || gam[t] != 0;
    // operator: out = v - fac T(v) + gam tr(v rho_in) rho_out, T from the FP64 chain or (Arnoldi
    // steps of a mixed engine) the low-precision one
    const bool mixed = !g.chain_lo.empty();
    auto op_chain = [&](const std::vector<std::unique_ptr<imps_plan_s>>& chain, const double* v, double* out) {
        const double* in = v;
        long off = 0;
        for (size_t k = 0; k < chain.size(); ++k) {
            double* o = (k + 1 == chain.size()) ? Tb : bonds + off * nv;
            run(chain[k].get(), in, o, nv);
            in = o;
            if (k + 1 < chain.size()) off += g.bond_elems[k];
        }
        if (anyproj) hip_check(rdot_partials(A() + g.w.off, v, n, nmom, part, s), "trace");
        hip_check(geo_combine(out, 1.0, v, Tb, dfac, dgpos, anyproj ? part : nullptr, nch, A() + g.rout.off, n, nmom, s), "1 - fac T");
    };
    auto op = [&](const double* v, double* out) { op_chain(g.chain, v, out); };
    auto op_step = [&](const double* v, double* out) { op_chain(mixed ? g.chain_lo : g.chain, v, out); };
    // a mixed cycle reduces its starting residual at most this far: below, the low-precision
    // products' rounding dominates and the next restart's FP64 residual takes over
    const double inner_tol = prec == IMPS_FP32 ? 1e-5 : 1e-4;
    // right-hand side b = x - gam tr(x rho_in) rho_out
    const double* x = A() + g.x.off;
    double* y = A() + g.y.off;
    if (g.x.ps != n || g.y.ps != n) throw InvalidArg("heff: geometric sums need contiguous planes");
    if (anyproj) {
        hip_check(rdot_partials(A() + g.w.off, x, n, nmom, part, s), "trace");
        hip_check(geo_combine(bb, 1.0, x, nullptr, nullptr, dgneg, part, nch, A() + g.rout.off, n, nmom, s), "rhs");
    } else {
        hip_check(hipMemcpyAsync(bb, x, sizeof(double) * vl, hipMemcpyDeviceToDevice, s), "rhs");
    }
    hip_check(launch_zero(y, vl, s), "zero solution");
    // norms of b
    std::vector<double> bnorm(nmom), beta(nmom), qh(nmom);
    auto norms = [&](double* v, std::vector<double>& out) {  // normalises v per momentum, returns the norms
        hip_check(cgemv_sub_partials(V, vl, 0, part, nch, nullptr, 0, v, n, nmom, qpart, s), "norm");
        hip_check(cscale_inv_norm(qpart, gch, cbuf, 1, v, n, nmom, s), "normalise");
        d2h_sync(s, qh.data(), cbuf, sizeof(double) * nmom);
        for (int t = 0; t < nmom; ++t) out[t] = std::sqrt(std::max(qh[t], 0.0));
    };
    hip_check(hipMemcpyAsync(V, bb, sizeof(double) * vl, hipMemcpyDeviceToDevice, s), "copy");
    norms(V, bnorm);
    beta = bnorm;
    std::vector<char> done(nmom, 0);
    std::vector<double> prev(nmom, 1e300);  // relative residual at the previous cycle start
    for (int t = 0; t < nmom; ++t) done[t] = bnorm[t] == 0.0;
    const double tol = inv_tol;
    stats.gmres_steps = 0;
    stats.gmres_relres = 0;
    bool first = true;
    std::vector<std::vector<cplx>> Hm(nmom), cs(nmom), sn(nmom), gg(nmom);
    std::vector<int> kdim(nmom);
    std::vector<double> raw;
    while (true) {
        if (!first) {  // true residual r = b - A y (V slot m is free at a cycle start)
            double* Ay = V + (size_t)vl * m;
            op(y, Ay);
            hip_check(geo_combine(V, 1.0, bb, Ay, dones, dgpos, nullptr, nch, nullptr, n, nmom, s), "residual");
            stats.gmres_steps += 1;
            norms(V, beta);
        }
        first = false;
        bool any = false;
        double worst = 0;
        for (int t = 0; t < nmom; ++t) {
            if (done[t]) continue;
            const double rel = beta[t] / bnorm[t];
            worst = std::max(worst, rel);
            if (rel < tol || rel > 0.5 * prev[t]) done[t] = 1;  // converged, or no longer improving
            prev[t] = rel;
            if (!done[t]) any = true;
        }
        stats.gmres_relres = worst;
        if (!any || stats.gmres_steps >= gmres_maxit) break;
        for (int t = 0; t < nmom; ++t) {
            Hm[t].assign((size_t)(m + 1) * m, 0.0);
            cs[t].assign(m, 0.0);
            sn[t].assign(m, 0.0);
            gg[t].assign(m + 1, 0.0);
            gg[t][0] = beta[t];
            kdim[t] = 0;
        }
        std::vector<char> cyc(nmom);  // still stepping in this cycle
        for (int t = 0; t < nmom; ++t) cyc[t] = !done[t];
        int j0 = 0, batch = 8;
        while (j0 < m) {
            const int q = std::min(batch, m - j0);
            for (int j = j0; j < j0 + q; ++j) {
                double* w = V + (size_t)vl * (j + 1);
                op_step(V + (size_t)vl * j, w);
                for (int pass = 0; pass < 2; ++pass) {
                    hip_check(cdot_partials(V, vl, j + 1, w, n, nmom, part, s), "cgs2 dots");
                    hip_check(cgemv_sub_partials(V, vl, j + 1, part, nch, cbuf + (size_t)(j - j0) * sst + pass * 2 * (m + 1), cst, w, n, nmom,
                                                 pass ? qpart : nullptr, s),
                              "cgs2 update");
                }
                hip_check(cscale_inv_norm(qpart, gch, cbuf + (size_t)(j - j0) * sst + 4 * (m + 1), cst, w, n, nmom, s), "cgs2 norm");
            }
            stats.gmres_steps += q;
            raw.resize((size_t)q * sst);
            d2h_sync(s, raw.data(), cbuf, sizeof(double) * raw.size());
            bool stepping = false;
            for (int t = 0; t < nmom; ++t) {
                if (!cyc[t]) continue;
                for (int j = j0; j < j0 + q && cyc[t]; ++j) {
                    const double* r = raw.data() + (size_t)(j - j0) * sst + (size_t)t * cst;
                    std::vector<cplx>& H = Hm[t];
                    for (int i = 0; i <= j; ++i) H[i + (size_t)j * (m + 1)] = cplx(r[2 * i] + r[2 * (m + 1) + 2 * i], r[2 * i + 1] + r[2 * (m + 1) + 2 * i + 1]);
                    const double hn = std::sqrt(std::max(r[4 * (m + 1)], 0.0));
                    H[j + 1 + (size_t)j * (m + 1)] = hn;
                    for (int i = 0; i < j; ++i) {
                        const cplx h1 = H[i + (size_t)j * (m + 1)], h2 = H[i + 1 + (size_t)j * (m + 1)];
                        H[i + (size_t)j * (m + 1)] = cs[t][i] * h1 + sn[t][i] * h2;
                        H[i + 1 + (size_t)j * (m + 1)] = -std::conj(sn[t][i]) * h1 + cs[t][i] * h2;
                    }
                    const cplx hjj = H[j + (size_t)j * (m + 1)];
                    const double b2 = std::sqrt(std::norm(hjj) + hn * hn);
                    kdim[t] = j + 1;
                    if (b2 == 0) {
                        cyc[t] = 0;
                        kdim[t] = j;
                        break;
                    }
                    const double ha = std::abs(hjj);
                    const cplx ph = ha > 0 ? hjj / ha : cplx(1.0);
                    sn[t][j] = ph * hn / b2;
                    cs[t][j] = ha / b2;
                    H[j + (size_t)j * (m + 1)] = ph * b2;
                    H[j + 1 + (size_t)j * (m + 1)] = 0.0;
                    gg[t][j + 1] = -std::conj(sn[t][j]) * gg[t][j];
                    gg[t][j] *= cs[t][j];
                    const double stop = mixed ? std::max(tol * bnorm[t], inner_tol * beta[t]) : tol * bnorm[t];
                    if (std::abs(gg[t][j + 1]) < stop || hn == 0) cyc[t] = 0;
                }
                if (cyc[t]) stepping = true;
            }
            j0 += q;
            if (!stepping) break;
        }
        // y += V R^-1 g per momentum (coefficients zero past each momentum's steps)
        std::vector<double> yc((size_t)nmom * 2 * m, 0.0);
        int kmax = 0;
        for (int t = 0; t < nmom; ++t) {
            const int k = kdim[t];
            kmax = std::max(kmax, k);
            std::vector<cplx> z(k);
            for (int i = k - 1; i >= 0; --i) {
                cplx acc = gg[t][i];
                for (int c = i + 1; c < k; ++c) acc -= Hm[t][i + (size_t)c * (m + 1)] * z[c];
                z[i] = acc / Hm[t][i + (size_t)i * (m + 1)];
            }
            for (int i = 0; i < k; ++i) {
                yc[(size_t)t * 2 * m + 2 * i] = z[i].real();
                yc[(size_t)t * 2 * m + 2 * i + 1] = z[i].imag();
            }
        }
        h2d_sync(s, ycoef, yc.data(), sizeof(double) * yc.size());
        hip_check(cgemv_add(V, vl, kmax, ycoef, 2 * m, y, n, nmom, s), "solution update");
    }
}

void HeffEngine::apply_planar(const double* xin, double* xout) {
    const long vl = 2L * nmom * mtot;
    if (xin != A() + Xin.off) hip_check(hipMemcpyAsync(A() + Xin.off, xin, sizeof(double) * vl, hipMemcpyDeviceToDevice, stream()), "copy in");
    run_tape(apply_tape);
    if (xout != A() + Xout.off)
        hip_check(hipMemcpyAsync(xout, A() + Xout.off, sizeof(double) * vl, hipMemcpyDeviceToDevice, stream()), "copy out");
}

// ---- momentum-batched Hermitian thick-restart Lanczos for the lowest eigenpairs -----------------
// Replaces eigs_n(Hfun, mtot, vals, vecs, nev, "SR", tol) (eigs.cpp:6-168) for every momentum at
// once: H_eff(k) is Hermitian, so the Krylov projection is a Hermitian matrix (assembled from the
// CGS2 coefficients of full re-orthogonalisation) and the restart keeps the lowest Ritz vectors and
// the next Lanczos vector (Wu-Simon).  Converged when every wanted Ritz pair of a momentum has
// residual <= tol max(1, |theta|) (ARPACK's criterion with the scale floored at 1).
int HeffEngine::eigs(const double* p, int nev, double tol, int maxit, int m, double* evals, double* evecs, double* resid, int* converged) {
    if (!state_set) throw InvalidArg("heff: state not set");
    if (!p || !evals) throw InvalidArg("heff: NULL argument");
    if (nev < 1) throw InvalidArg("heff: nev must be >= 1");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    hipStream_t s = stream();
    const long n = mtot;
    if (nev >= n) throw InvalidArg("heff: nev must be below the excitation dimension");
    if (m <= 0) m = std::max(20, 2 * nev + 1);
    m = (int)std::min<long>(std::max(m, nev + 3), n);
    const int keep = std::min(m - 2, nev + std::max(2, (m - nev) / 2));
    if (keep < nev) throw InvalidArg("heff: Krylov dimension too small for nev");
    if (maxit <= 0) maxit = 100 * m;
    p_cur.assign(p, p + nmom);
    upload_coefs();
    stats_steps_total = 0;
    stats_relres_worst = 0;
    const int nv = 2 * nmom;
    const long vl = (long)nv * n;
    const int nch = cblas_chunks(n), gch = cblas_gemv_grid(n, nmom);
    const long cst = 4L * (m + 1) + 1, sst = cst * nmom;
    const size_t need = (size_t)vl * (m + 1) + (size_t)vl * keep + (size_t)nch * (m + 1) * nmom * 2 + (size_t)gch * nmom +
                        (size_t)sst * (m + 1) + (size_t)nmom * keep * m * 2 + 64;
    eig_kry.ensure(sizeof(double) * need);
    double* V = eig_kry.d();
    double* Y = V + (size_t)vl * (m + 1);
    double* part = Y + (size_t)vl * keep;
    double* qpart = part + (size_t)nch * (m + 1) * nmom * 2;
    double* cbuf = qpart + (size_t)gch * nmom;
    double* Sd = cbuf + (size_t)sst * (m + 1);
    // seeded start vector: SplitMix64 + Box-Muller normal entries per momentum
    {
        std::vector<double> h((size_t)vl);
        uint64_t st = 0x9E3779B97F4A7C15ull;
        auto next = [&]() {
            uint64_t z = (st += 0x9E3779B97F4A7C15ull);
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            z ^= z >> 31;
            return ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        };
        for (size_t i = 0; i < h.size(); i += 2) {
            const double u1 = next(), u2 = next(), r = std::sqrt(-2 * std::log(u1));
            h[i] = r * std::cos(2 * M_PI * u2);
            if (i + 1 < h.size()) h[i + 1] = r * std::sin(2 * M_PI * u2);
        }
        h2d_sync(s, V, h.data(), sizeof(double) * h.size());
        hip_check(cgemv_sub_partials(V, vl, 0, part, nch, nullptr, 0, V, n, nmom, qpart, s), "norm");
        hip_check(cscale_inv_norm(qpart, gch, cbuf, 1, V, n, nmom, s), "normalise");
    }
    std::vector<std::vector<cplx>> T(nmom, std::vector<cplx>((size_t)m * m, 0.0));
    std::vector<std::vector<double>> theta(nmom);
    std::vector<std::vector<cplx>> Sv(nmom);
    std::vector<double> res((size_t)nmom * nev, 0.0), raw;
    int j0 = 0, napply = 0;
    bool all_ok = false;
    while (true) {
        for (int j = j0; j < m; ++j) {
            double* w = V + (size_t)vl * (j + 1);
            apply_planar(V + (size_t)vl * j, w);
            ++napply;
            for (int pass = 0; pass < 2; ++pass) {
                hip_check(cdot_partials(V, vl, j + 1, w, n, nmom, part, s), "cgs2 dots");
                hip_check(cgemv_sub_partials(V, vl, j + 1, part, nch, cbuf + (size_t)j * sst + pass * 2 * (m + 1), cst, w, n, nmom,
                                             pass ? qpart : nullptr, s),
                          "cgs2 update");
            }
            hip_check(cscale_inv_norm(qpart, gch, cbuf + (size_t)j * sst + 4 * (m + 1), cst, w, n, nmom, s), "cgs2 norm");
        }
        raw.resize((size_t)(m - j0) * sst);
        d2h_sync(s, raw.data(), cbuf + (size_t)j0 * sst, sizeof(double) * raw.size());
        all_ok = true;
        std::vector<double> beta(nmom);
        for (int t = 0; t < nmom; ++t) {
            std::vector<cplx>& Tt = T[t];
            for (int j = j0; j < m; ++j) {
                const double* r = raw.data() + (size_t)(j - j0) * sst + (size_t)t * cst;
                for (int i = 0; i <= j; ++i) {
                    const cplx h(r[2 * i] + r[2 * (m + 1) + 2 * i], r[2 * i + 1] + r[2 * (m + 1) + 2 * i + 1]);
                    Tt[i + (size_t)j * m] = i == j ? cplx(h.real(), 0.0) : h;
                    if (i != j) Tt[j + (size_t)i * m] = std::conj(h);
                }
                const double b = std::sqrt(std::max(r[4 * (m + 1)], 0.0));
                if (j + 1 < m) {
                    Tt[j + 1 + (size_t)j * m] = b;
                    Tt[j + (size_t)(j + 1) * m] = b;
                } else {
                    beta[t] = b;
                }
            }
            herm_eig(m, Tt, theta[t], Sv[t]);
            for (int i = 0; i < nev; ++i) {
                res[(size_t)t * nev + i] = beta[t] * std::abs(Sv[t][(m - 1) + (size_t)i * m]);
                if (res[(size_t)t * nev + i] > tol * std::max(1.0, std::abs(theta[t][i]))) all_ok = false;
            }
        }
        if (verbose()) {
            double worst = 0;
            int open = 0;
            for (int t = 0; t < nmom; ++t) {
                bool ok = true;
                for (int i = 0; i < nev; ++i) {
                    const double r = res[(size_t)t * nev + i] / std::max(1.0, std::abs(theta[t][i]));
                    worst = std::max(worst, r);
                    ok = ok && r <= tol;
                }
                open += ok ? 0 : 1;
            }
            std::fprintf(stderr, "heff eigs: %d applications, worst relative residual %.3e, %d of %d momenta open (geometric sums: worst GMRES residual %.1e)\n",
                         napply, worst, open, nmom, stats_relres_worst);
        }
        const bool stop = all_ok || napply >= maxit;
        const int nk = stop ? nev : keep;
        // Ritz vectors Y_i = sum_j S_ji V_j for the lowest nk
        std::vector<double> S((size_t)nmom * nk * m * 2);
        for (int t = 0; t < nmom; ++t)
            for (int i = 0; i < nk; ++i)
                for (int j = 0; j < m; ++j) {
                    const cplx z = Sv[t][j + (size_t)i * m];
                    S[(((size_t)t * nk + i) * m + j) * 2] = z.real();
                    S[(((size_t)t * nk + i) * m + j) * 2 + 1] = z.imag();
                }
        h2d_sync(s, Sd, S.data(), sizeof(double) * S.size());
        hip_check(cgemm_basis(V, vl, m, Sd, nk, Y, vl, n, nmom, s), "ritz vectors");
        if (stop) break;
        // new basis: the keep Ritz vectors and the last Lanczos vector
        hip_check(hipMemcpyAsync(V + (size_t)vl * keep, V + (size_t)vl * m, sizeof(double) * vl, hipMemcpyDeviceToDevice, s), "restart");
        hip_check(hipMemcpyAsync(V, Y, sizeof(double) * vl * keep, hipMemcpyDeviceToDevice, s), "restart");
        for (int t = 0; t < nmom; ++t) {
            std::fill(T[t].begin(), T[t].end(), cplx(0.0));
            for (int i = 0; i < keep; ++i) T[t][i + (size_t)i * m] = theta[t][i];
        }
        j0 = keep;
    }
    for (int t = 0; t < nmom; ++t)
        for (int i = 0; i < nev; ++i) evals[(size_t)t * nev + i] = theta[t][i];
    if (resid) std::copy(res.begin(), res.end(), resid);
    if (converged) *converged = all_ok ? 1 : 0;
    if (evecs) {
        DevBuf tmp;
        tmp.ensure(sizeof(double) * vl);
        for (int i = 0; i < nev; ++i) {
            hip_check(cinterleave(Y + (size_t)vl * i, tmp.d(), n, nmom, s), "interleave");
            hip_check(hipMemcpy2DAsync(evecs + (size_t)2 * n * i, sizeof(double) * 2 * n * nev, tmp.d(), sizeof(double) * 2 * n,
                                       sizeof(double) * 2 * n, nmom, hipMemcpyDeviceToDevice, s),
                      "copy eigenvectors");
        }
        hip_check(hipStreamSynchronize(s), "sync");
    }
    return napply;
}

void HeffEngine::apply(const double* p, const double* x, double* y) {
    if (!state_set) throw InvalidArg("heff: state not set");
    if (!p || !x || !y) throw InvalidArg("heff: NULL argument");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    p_cur.assign(p, p + nmom);
    upload_coefs();
    hip_check(cdeinterleave(x, A() + Xin.off, mtot, nmom, stream()), "deinterleave");
    stats_steps_total = 0;
    stats_relres_worst = 0;
    run_tape(apply_tape);
    hip_check(cinterleave(A() + Xout.off, y, mtot, nmom, stream()), "interleave");
}

}  // namespace imps
