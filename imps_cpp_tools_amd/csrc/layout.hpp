// Host-side sector bookkeeping and grouped block-GEMM programs.
//
// A block-sparse object (BlockMat, BlockDiagMat, MPSBlockMat, RedOp) lives in one flat buffer in
// the reference's Vectorize order; its structure is a Layout: per list entry, the blocks sorted by
// ingoing charge with their element offsets.  Charge lookups replace the reference's
// std::map::find (BlockObj.hpp:2557-2625, TMBlockFunctions.hpp) by binary search over the sorted
// block table, done once per plan on the host.
//
// Every operator application is compiled, once per sector structure, into a Program: a short
// sequence of Stages, each one grouped launch.  A Stage is a list of Tasks, each owning one output
// tile (<= kTile x kTile) of one output block, and each Task sums a list of Segments
// C_tile += alpha * op(A) * op(B) taken from the operand slots (input vector, operator store,
// intermediate workspace).  Because a task owns its tile and sums its segments in a fixed order,
// results are deterministic and need no atomics.
#pragma once

#include <algorithm>
#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/imps_capi.h"

namespace imps {

constexpr int kMaxSym = 4;
constexpr int kTile = 32;  // output tile edge of the grouped GEMM kernel (one workgroup per tile)

struct InvalidArg : std::runtime_error { using std::runtime_error::runtime_error; };

// Residue of x modulo a period p > 1 in [0, p); p = 0 (U(1)) leaves x as it is.
inline int kred(int x, int p) { return p > 1 ? ((x % p) + p) % p : x; }

// Charge key: nsym abelian charges compared lexicographically (std::vector<int> order,
// KeyTypes.hpp:19-58).  Each charge has a period (symobj.hpp:80-161): 0 for U(1), N > 1 for Z(N),
// whose charges live in [0, N).  Sums and differences reduce with the left operand's periods, as
// the reference's KeyType operators use the left key's group object (KeyTypes.hpp:133-151).
struct Key {
    int n = 0;
    int q[kMaxSym] = {0, 0, 0, 0};
    int per[kMaxSym] = {0, 0, 0, 0};
    Key() {}
    Key(const int* p, int nsym) : n(nsym) { for (int i = 0; i < nsym; ++i) q[i] = p[i]; }
    Key(const int* p, int nsym, const int* periods) : n(nsym) {
        for (int i = 0; i < nsym; ++i) { per[i] = periods ? periods[i] : 0; q[i] = p[i]; }
    }
    bool operator<(const Key& o) const { for (int i = 0; i < n; ++i) if (q[i] != o.q[i]) return q[i] < o.q[i]; return false; }
    bool operator==(const Key& o) const { for (int i = 0; i < n; ++i) if (q[i] != o.q[i]) return false; return true; }
    bool operator!=(const Key& o) const { return !(*this == o); }
    Key operator+(const Key& o) const { Key r(*this); for (int i = 0; i < n; ++i) r.q[i] = kred(q[i] + o.q[i], per[i]); return r; }
    Key operator-(const Key& o) const { Key r(*this); for (int i = 0; i < n; ++i) r.q[i] = kred(q[i] - o.q[i], per[i]); return r; }
    bool canonical() const { for (int i = 0; i < n; ++i) if (per[i] > 1 && (q[i] < 0 || q[i] >= per[i])) return false; return true; }
};

struct Blk {
    Key qin, qout;
    int rows = 0, cols = 0;
    long off = 0;  // element offset in the object's flat buffer
    long size() const { return (long)rows * cols; }
};

// One BlockMat: blocks in ascending qin.
struct BList {
    std::vector<Blk> b;
    const Blk* find(const Key& k) const {
        auto it = std::lower_bound(b.begin(), b.end(), k, [](const Blk& x, const Key& y) { return x.qin < y; });
        return (it != b.end() && it->qin == k) ? &*it : nullptr;
    }
    // dK = Qout - Qin of the slice (BlockMat::dK); false for an empty slice
    bool dK(Key& k) const { if (b.empty()) return false; k = b.front().qout - b.front().qin; return true; }
};

struct Layout {
    int nsym = 1;
    int per[kMaxSym] = {0, 0, 0, 0};  // charge periods (0: U(1), N > 1: Z(N))
    std::vector<BList> l;
    long nelem = 0;
    size_t nblocks() const { size_t n = 0; for (auto& x : l) n += x.b.size(); return n; }
};

inline Layout layout_from_c(const imps_layout* c, const char* what) {
    if (!c) throw InvalidArg(std::string(what) + ": layout is NULL");
    if (c->nsym < 1 || c->nsym > kMaxSym) throw InvalidArg(std::string(what) + ": nsym must be 1..4");
    if (c->nlist < 1) throw InvalidArg(std::string(what) + ": nlist must be >= 1");
    if (!c->nblk) throw InvalidArg(std::string(what) + ": nblk is NULL");
    Layout L;
    L.nsym = c->nsym;
    for (int i = 0; i < c->nsym; ++i) {
        const int p = c->periods ? c->periods[i] : 0;
        if (p < 0) throw InvalidArg(std::string(what) + ": negative charge period");
        L.per[i] = p > 1 ? p : 0;  // period 1 means U(1), as in symobj::init
    }
    L.l.resize(c->nlist);
    long pos = 0;
    size_t k = 0;
    for (int e = 0; e < c->nlist; ++e) {
        if (c->nblk[e] < 0) throw InvalidArg(std::string(what) + ": negative block count");
        if (c->nblk[e] > 0 && (!c->qin || !c->qout || !c->rows || !c->cols)) throw InvalidArg(std::string(what) + ": NULL block arrays");
        for (int i = 0; i < c->nblk[e]; ++i, ++k) {
            Blk b;
            b.qin = Key(c->qin + k * c->nsym, c->nsym, L.per);
            b.qout = Key(c->qout + k * c->nsym, c->nsym, L.per);
            if (!b.qin.canonical() || !b.qout.canonical()) throw InvalidArg(std::string(what) + ": a Z(N) charge outside [0, N)");
            b.rows = c->rows[k];
            b.cols = c->cols[k];
            if (b.rows < 0 || b.cols < 0) throw InvalidArg(std::string(what) + ": negative block size");
            if (!L.l[e].b.empty() && !(L.l[e].b.back().qin < b.qin))
                throw InvalidArg(std::string(what) + ": blocks must be in strictly ascending ingoing-charge order");
            b.off = pos;
            pos += b.size();
            L.l[e].b.push_back(b);
        }
    }
    L.nelem = pos;
    return L;
}

inline bool same_group(const Layout& a, const Layout& b) {
    if (a.nsym != b.nsym) return false;
    for (int i = 0; i < a.nsym; ++i) if (a.per[i] != b.per[i]) return false;
    return true;
}

// dst takes src's symmetry group (charge count and periods)
inline void set_group(Layout& dst, const Layout& src) {
    dst.nsym = src.nsym;
    for (int i = 0; i < kMaxSym; ++i) dst.per[i] = src.per[i];
}

// ---------------------------------------------------------------------------------------------
// Grouped GEMM programs
enum Slot : int { SLOT_X = 0, SLOT_OP = 1, SLOT_W = 2, SLOT_OUT = 3 };

// C_tile += alpha * op(A)[m0:, :K] * op(B)[:K, n0:]; op(A)(i,k) = tr ? S[k + i*ld] : S[i + k*ld]
struct Seg {
    long a_off, b_off;
    int a_slot, a_ld, a_tr;
    int b_slot, b_ld, b_tr;
    int K;
    double alpha;
};
struct Task {
    long c_off;
    int c_ld, m0, n0, M, N;
    int seg0, nseg;
};
struct Stage {
    int out_slot = SLOT_OUT;   // SLOT_OUT: final output; SLOT_W: intermediate workspace
    std::vector<Task> tasks;   // tiles of at most kTile x kTile
    std::vector<Seg> segs;
};
struct Program {
    std::vector<Stage> stages;
    long w_elems = 0;    // per-vector workspace elements
    double flops = 0;    // algorithmic flops per single-vector apply
    long x_elems = 0, y_elems = 0, op_elems = 0;
    long operand_elems_read = 0;  // algorithmic operand reads (each operand block counted once)
};

// An output block being assembled from segments.
struct OutBlock {
    long c_off;
    int M, N;
    std::vector<Seg> segs;
};

struct Operand {
    int slot;
    long off;
    int rows, cols;  // stored shape
    bool tr;         // use the transpose of the stored block
    int orows() const { return tr ? cols : rows; }
    int ocols() const { return tr ? rows : cols; }
};
inline Operand opnd(int slot, const Blk& b, long base, bool tr = false) { return Operand{slot, base + b.off, b.rows, b.cols, tr}; }

// Append alpha*op(A)*op(B) to an output block.
inline void add_seg(OutBlock& ob, const Operand& A, const Operand& B, double alpha, double& flops) {
    if (A.orows() != ob.M || B.ocols() != ob.N || A.ocols() != B.orows())
        throw InvalidArg("block sizes of matched sectors disagree");
    Seg s;
    s.a_off = A.off; s.a_slot = A.slot; s.a_ld = std::max(A.rows, 1); s.a_tr = A.tr ? 1 : 0;
    s.b_off = B.off; s.b_slot = B.slot; s.b_ld = std::max(B.rows, 1); s.b_tr = B.tr ? 1 : 0;
    s.K = A.ocols();
    s.alpha = alpha;
    if (s.K > 0 && ob.M > 0 && ob.N > 0) {
        ob.segs.push_back(s);
        flops += 2.0 * ob.M * ob.N * s.K;
    }
}

// Tile the assembled output blocks into tasks of a stage.
inline void emit_stage(Program& P, std::vector<OutBlock>& obs, int out_slot) {
    Stage st;
    st.out_slot = out_slot;
    for (auto& ob : obs) {
        if (ob.M == 0 || ob.N == 0) continue;
        int seg0 = (int)st.segs.size();
        st.segs.insert(st.segs.end(), ob.segs.begin(), ob.segs.end());
        for (int m0 = 0; m0 < ob.M; m0 += kTile)
            for (int n0 = 0; n0 < ob.N; n0 += kTile)
                st.tasks.push_back(Task{ob.c_off, ob.M, m0, n0, ob.M, ob.N, seg0, (int)ob.segs.size()});
    }
    P.stages.push_back(std::move(st));
}

inline void require_same_nsym(const Layout& a, const Layout& b, const char* what) {
    if (!same_group(a, b)) throw InvalidArg(std::string(what) + ": layouts disagree on the symmetry group (nsym or charge periods)");
}

// ---------------------------------------------------------------------------------------------
// H_AC (VUMPS_helpers.hpp:38-53): out[i] = EHL*in[i] + in[i]*EHR + sum_j HL[i][j]*in[j] + in[j]*HR[i][j]
// Operator store layout: [EHL | EHR | HL | HR] in the caller's Vectorize order.
inline Program build_hac(const Layout& AC, const Layout& EHL, const Layout& EHR, const Layout& HL, const Layout& HR) {
    const int d = (int)AC.l.size();
    require_same_nsym(AC, EHL, "hac"); require_same_nsym(AC, EHR, "hac");
    require_same_nsym(AC, HL, "hac"); require_same_nsym(AC, HR, "hac");
    if (EHL.l.size() != 1 || EHR.l.size() != 1) throw InvalidArg("hac: EHL/EHR must have nlist = 1");
    if ((int)HL.l.size() != d * d || (int)HR.l.size() != d * d) throw InvalidArg("hac: HL/HR must have nlist = d*d");
    Program P;
    const long ehl0 = 0, ehr0 = EHL.nelem, hl0 = ehr0 + EHR.nelem, hr0 = hl0 + HL.nelem;
    P.op_elems = hr0 + HR.nelem;
    P.x_elems = P.y_elems = AC.nelem;
    std::vector<OutBlock> obs;
    for (int i = 0; i < d; ++i)
        for (const Blk& ob : AC.l[i].b) {
            OutBlock o{ob.off, ob.rows, ob.cols, {}};
            // EHL*in + in*EHR (BlockDiagMat x BlockMat products, BlockObj.hpp:2583-2606)
            if (const Blk* e = EHL.l[0].find(ob.qin)) add_seg(o, opnd(SLOT_OP, *e, ehl0), opnd(SLOT_X, ob, 0), 1.0, P.flops);
            if (const Blk* e = EHR.l[0].find(ob.qout)) add_seg(o, opnd(SLOT_X, ob, 0), opnd(SLOT_OP, *e, ehr0), 1.0, P.flops);
            for (int j = 0; j < d; ++j) {
                // HL[i][j]*in[j]: block of HL with Qin = qin, matched with in[j] at its Qout
                if (const Blk* h = HL.l[i * d + j].find(ob.qin))
                    if (const Blk* x = AC.l[j].find(h->qout))
                        if (x->qout == ob.qout) add_seg(o, opnd(SLOT_OP, *h, hl0), opnd(SLOT_X, *x, 0), 1.0, P.flops);
                // in[j]*HR[i][j]: block of in[j] with Qin = qin, matched with HR at its Qout
                if (const Blk* x = AC.l[j].find(ob.qin))
                    if (const Blk* h = HR.l[i * d + j].find(x->qout))
                        if (h->qout == ob.qout) add_seg(o, opnd(SLOT_X, *x, 0), opnd(SLOT_OP, *h, hr0), 1.0, P.flops);
            }
            obs.push_back(std::move(o));
        }
    emit_stage(P, obs, SLOT_OUT);
    P.operand_elems_read = P.op_elems + AC.nelem;
    return P;
}

// ---------------------------------------------------------------------------------------------
// Mixed transfer matrix (TMBlockFunctions.hpp:352-390 left, :494-533 right), two stages:
//   left : W = x(qi->qo) * A_s(qo) ;  y[Qout B_s(qi)] += B_s(qi)^T * W
//   right: W = x(qi->qo) * B_s(qo-dB)^T ;  y[Qin A_s(qi-dA)] += A_s(qi-dA) * W
// Operator store: [A | B] (B omitted when B is A).
inline Program build_tm(int dir, const Layout& A, const Layout* Bp, const Layout& x, const Layout& y) {
    const Layout& B = Bp ? *Bp : A;
    require_same_nsym(A, B, "tm"); require_same_nsym(A, x, "tm"); require_same_nsym(A, y, "tm");
    if (A.l.size() != B.l.size()) throw InvalidArg("tm: A and B need the same physical dimension");
    if (x.l.size() != 1 || y.l.size() != 1) throw InvalidArg("tm: x and y must have nlist = 1");
    Program P;
    const long a0 = 0, b0 = Bp ? A.nelem : 0;
    P.op_elems = A.nelem + (Bp ? B.nelem : 0);
    P.x_elems = x.nelem;
    P.y_elems = y.nelem;
    std::vector<OutBlock> st1;
    std::map<long, OutBlock> st2;  // keyed by output offset, filled in (s, x-block) order
    for (const Blk& yb : y.l[0].b) st2.emplace(yb.off, OutBlock{yb.off, yb.rows, yb.cols, {}});
    long w = 0;
    const int d = (int)A.l.size();
    for (int s = 0; s < d; ++s) {
        if (A.l[s].b.empty() || B.l[s].b.empty()) continue;
        Key dKA, dKB;
        A.l[s].dK(dKA);
        B.l[s].dK(dKB);
        for (const Blk& xb : x.l[0].b) {
            const Blk *lit, *rit, *yb;
            if (dir == IMPS_LEFT) {
                lit = B.l[s].find(xb.qin);
                rit = A.l[s].find(xb.qout);
                if (!lit || !rit) continue;
                yb = y.l[0].find(lit->qout);
                if (!yb) continue;
                if (yb->qout != rit->qout) throw InvalidArg("tm: output block has a different outgoing charge");
                OutBlock wb{w, xb.rows, rit->cols, {}};
                add_seg(wb, opnd(SLOT_X, xb, 0), opnd(SLOT_OP, *rit, a0), 1.0, P.flops);
                Operand W{SLOT_W, w, xb.rows, rit->cols, false};
                add_seg(st2.at(yb->off), opnd(SLOT_OP, *lit, b0, true), W, 1.0, P.flops);
                w += (long)xb.rows * rit->cols;
                st1.push_back(std::move(wb));
            } else {
                lit = A.l[s].find(xb.qin - dKA);
                rit = B.l[s].find(xb.qout - dKB);
                if (!lit || !rit) continue;
                yb = y.l[0].find(lit->qin);
                if (!yb) continue;
                if (yb->qout != rit->qin) throw InvalidArg("tm: output block has a different outgoing charge");
                OutBlock wb{w, xb.rows, rit->rows, {}};
                add_seg(wb, opnd(SLOT_X, xb, 0), opnd(SLOT_OP, *rit, b0, true), 1.0, P.flops);
                Operand W{SLOT_W, w, xb.rows, rit->rows, false};
                add_seg(st2.at(yb->off), opnd(SLOT_OP, *lit, a0), W, 1.0, P.flops);
                w += (long)xb.rows * rit->rows;
                st1.push_back(std::move(wb));
            }
        }
    }
    P.w_elems = w;
    emit_stage(P, st1, SLOT_W);
    std::vector<OutBlock> v2;
    for (auto& kv : st2) v2.push_back(std::move(kv.second));
    emit_stage(P, v2, SLOT_OUT);
    P.operand_elems_read = P.op_elems + x.nelem;
    return P;
}

// ---------------------------------------------------------------------------------------------
// H_C in reduced form (VUMPS_helpers.hpp:91-111 restated with HL = ReducedOp(H,AL,l)):
//   stage 1: W1_j = C * AR[j]                         (per AR[j] block p->p', C block p)
//   stage 2: W2_i = sum_j HL[i][j] * W1_j             (blocks q -> q + dI2K[i])
//   stage 3: out_q = EHL_q C_q + C_q EHR_q + sum_i W2_i(q -> q') * AR[i](q -> q')^T
// Operator store: [EHL | EHR | HL | AR].
inline Program build_hc(const Layout& C, const Layout& EHL, const Layout& EHR, const Layout& HL, const Layout& AR) {
    const int d = (int)AR.l.size();
    require_same_nsym(C, EHL, "hc"); require_same_nsym(C, EHR, "hc"); require_same_nsym(C, HL, "hc"); require_same_nsym(C, AR, "hc");
    if (C.l.size() != 1 || EHL.l.size() != 1 || EHR.l.size() != 1) throw InvalidArg("hc: C/EHL/EHR must have nlist = 1");
    if ((int)HL.l.size() != d * d) throw InvalidArg("hc: HL must have nlist = d*d");
    Program P;
    const long ehl0 = 0, ehr0 = EHL.nelem, hl0 = ehr0 + EHR.nelem, ar0 = hl0 + HL.nelem;
    P.op_elems = ar0 + AR.nelem;
    P.x_elems = P.y_elems = C.nelem;
    long w = 0;
    // stage 1
    std::vector<OutBlock> s1;
    std::vector<std::map<Key, Blk>> W1(d);  // per j: qin -> workspace block
    for (int j = 0; j < d; ++j)
        for (const Blk& a : AR.l[j].b) {
            const Blk* c = C.l[0].find(a.qin);
            if (!c) continue;
            Blk wb{a.qin, a.qout, c->rows, a.cols, w};
            OutBlock o{w, wb.rows, wb.cols, {}};
            add_seg(o, opnd(SLOT_X, *c, 0), opnd(SLOT_OP, a, ar0), 1.0, P.flops);
            s1.push_back(std::move(o));
            W1[j].emplace(a.qin, wb);
            w += wb.size();
        }
    emit_stage(P, s1, SLOT_W);
    // stage 2: W2_i[q] = sum_j HL[i][j](q -> p) * W1_j(p -> p')
    std::vector<OutBlock> s2;
    std::vector<std::map<Key, Blk>> W2(d);
    for (int i = 0; i < d; ++i) {
        std::map<Key, OutBlock> acc;
        std::map<Key, Blk> shape;
        for (int j = 0; j < d; ++j)
            for (const Blk& h : HL.l[i * d + j].b) {
                auto it = W1[j].find(h.qout);
                if (it == W1[j].end()) continue;
                const Blk& y1 = it->second;
                auto sh = shape.find(h.qin);
                if (sh == shape.end()) {
                    Blk wb{h.qin, y1.qout, h.rows, y1.cols, 0};
                    shape.emplace(h.qin, wb);
                    acc.emplace(h.qin, OutBlock{0, h.rows, y1.cols, {}});
                } else if (sh->second.qout != y1.qout) {
                    throw InvalidArg("hc: inconsistent charge flow in HL");
                }
                Operand Y1{SLOT_W, y1.off, y1.rows, y1.cols, false};
                add_seg(acc.at(h.qin), opnd(SLOT_OP, h, hl0), Y1, 1.0, P.flops);
            }
        for (auto& kv : shape) {
            kv.second.off = w;
            OutBlock ob = std::move(acc.at(kv.first));
            ob.c_off = w;
            w += kv.second.size();
            s2.push_back(std::move(ob));
        }
        W2[i] = std::move(shape);
    }
    emit_stage(P, s2, SLOT_W);
    // stage 3
    std::vector<OutBlock> s3;
    for (const Blk& cb : C.l[0].b) {
        OutBlock o{cb.off, cb.rows, cb.cols, {}};
        if (const Blk* e = EHL.l[0].find(cb.qin)) add_seg(o, opnd(SLOT_OP, *e, ehl0), opnd(SLOT_X, cb, 0), 1.0, P.flops);
        if (const Blk* e = EHR.l[0].find(cb.qout)) add_seg(o, opnd(SLOT_X, cb, 0), opnd(SLOT_OP, *e, ehr0), 1.0, P.flops);
        for (int i = 0; i < d; ++i) {
            auto it = W2[i].find(cb.qin);
            if (it == W2[i].end()) continue;
            const Blk& z = it->second;
            const Blk* a = AR.l[i].find(cb.qin);  // AR[i] block q -> q'; its transpose maps q' -> q
            if (!a || a->qout != z.qout) continue;
            Operand Z{SLOT_W, z.off, z.rows, z.cols, false};
            add_seg(o, Z, opnd(SLOT_OP, *a, ar0, true), 1.0, P.flops);
        }
        s3.push_back(std::move(o));
    }
    emit_stage(P, s3, SLOT_OUT);
    P.w_elems = w;
    P.operand_elems_read = P.op_elems + C.nelem;
    return P;
}

// ---------------------------------------------------------------------------------------------
// Two-operand block products (BlockObj.hpp:2557-2625, MPSBlockMat.hpp:934-1043,
// TMBlockFunctions.hpp:259-470).  Operand 1 is the X slot, operand 2 the OP slot (FP64 plans bind
// it to the caller's buffer).  The output layout is derived from the operands' structure.
enum ProdKind : int {
    PROD_MAT_MPS = 0,    // y[s] = M * A[s]           (M: BlockMat/BlockDiagMat, A: MPS)
    PROD_MPS_MAT = 1,    // y[s] = A[s] * M
    PROD_MPS_MATT = 2,   // y[s] = A[s] * M^T
    PROD_MATT_MPS = 3,   // y[s] = M^T * A[s]
    PROD_TM_LEFT = 4,    // y = sum_s B_s^T A_s, keyed by Qout(B_s) (ApplyTMmixedLeft(A,B) unity)
    PROD_TM_RIGHT = 5,   // y = sum_s A_s B_s^T, keyed by Qin(A_s)  (ApplyTMmixedRight(A,B) unity)
    PROD_MPS_MPS = 6,    // y[i*d2 + j] = A[i] * B[j] (two-site concatenation)
    PROD_REDOP_MPS = 7,  // y[i] = sum_j R[i][j] * A[j]   (X1 = R: d*d lists, X2 = A)
    PROD_MPS_REDOP = 8   // y[i] = sum_j A[j] * R[i][j]   (X1 = R, X2 = A)
};

// Per-entry product lhs * rhs of two BlockMats with the BlockMat operator* matching (lhs Qout =
// rhs Qin); lhs/rhs may be used transposed.  Output blocks keyed by the lhs row charge.
inline void prod_entry(const BList& L, long l0, int lslot, bool ltr, const BList& R, long r0, int rslot, bool rtr, BList& out,
                       std::vector<OutBlock>& obs, long& pos, double& flops) {
    // rows of op(L) are keyed by Qin (no transpose) or Qout (transpose); index op(R) blocks by row key
    std::map<Key, const Blk*> rrow;
    for (const Blk& b : R.b) rrow[rtr ? b.qout : b.qin] = &b;
    std::map<Key, std::pair<Blk, OutBlock>> acc;
    for (const Blk& a : L.b) {
        const Key arow = ltr ? a.qout : a.qin, acol = ltr ? a.qin : a.qout;
        auto it = rrow.find(acol);
        if (it == rrow.end()) continue;
        const Blk& b = *it->second;
        const Key bcol = rtr ? b.qin : b.qout;
        const int M = ltr ? a.cols : a.rows, N = rtr ? b.rows : b.cols;
        auto ins = acc.find(arow);
        if (ins == acc.end()) ins = acc.emplace(arow, std::make_pair(Blk{arow, bcol, M, N, 0}, OutBlock{0, M, N, {}})).first;
        else if (ins->second.first.qout != bcol) throw InvalidArg("product: inconsistent charge flow");
        add_seg(ins->second.second, opnd(lslot, a, l0, ltr), opnd(rslot, b, r0, rtr), 1.0, flops);
    }
    for (auto& kv : acc) {
        Blk b = kv.second.first;
        b.off = pos;
        OutBlock ob = std::move(kv.second.second);
        ob.c_off = pos;
        pos += b.size();
        out.b.push_back(b);
        obs.push_back(std::move(ob));
    }
}

inline Program build_prod(int kind, const Layout& X1, const Layout& X2, Layout& out) {
    require_same_nsym(X1, X2, "prod");
    Program P;
    P.x_elems = X1.nelem;
    P.op_elems = X2.nelem;
    out = Layout();
    set_group(out, X1);
    long pos = 0;
    std::vector<OutBlock> obs;
    auto entry_offset = [](const Layout& L, size_t e) { return L.l[e].b.empty() ? 0L : L.l[e].b.front().off; };
    (void)entry_offset;
    switch (kind) {
        case PROD_MAT_MPS:
        case PROD_MATT_MPS:
            if (X1.l.size() != 1) throw InvalidArg("prod: M must have nlist = 1");
            out.l.resize(X2.l.size());
            for (size_t s = 0; s < X2.l.size(); ++s)
                prod_entry(X1.l[0], 0, SLOT_X, kind == PROD_MATT_MPS, X2.l[s], 0, SLOT_OP, false, out.l[s], obs, pos, P.flops);
            break;
        case PROD_MPS_MAT:
        case PROD_MPS_MATT:
            if (X2.l.size() != 1) throw InvalidArg("prod: M must have nlist = 1");
            out.l.resize(X1.l.size());
            for (size_t s = 0; s < X1.l.size(); ++s)
                prod_entry(X1.l[s], 0, SLOT_X, false, X2.l[0], 0, SLOT_OP, kind == PROD_MPS_MATT, out.l[s], obs, pos, P.flops);
            break;
        case PROD_MPS_MPS:
            out.l.resize(X1.l.size() * X2.l.size());
            for (size_t i = 0; i < X1.l.size(); ++i)
                for (size_t j = 0; j < X2.l.size(); ++j)
                    prod_entry(X1.l[i], 0, SLOT_X, false, X2.l[j], 0, SLOT_OP, false, out.l[i * X2.l.size() + j], obs, pos, P.flops);
            break;
        case PROD_TM_LEFT:
        case PROD_TM_RIGHT: {
            if (X1.l.size() != X2.l.size()) throw InvalidArg("prod: A and B need the same physical dimension");
            // accumulate over s into blocks keyed by the output row charge
            std::map<Key, std::pair<Blk, OutBlock>> acc;
            for (size_t s = 0; s < X1.l.size(); ++s) {
                if (kind == PROD_TM_LEFT) {  // B_s^T A_s: match Qin(B) with Qin(A)
                    for (const Blk& a : X1.l[s].b) {
                        const Blk* b = X2.l[s].find(a.qin);
                        if (!b) continue;
                        auto it = acc.find(b->qout);
                        if (it == acc.end()) it = acc.emplace(b->qout, std::make_pair(Blk{b->qout, a.qout, b->cols, a.cols, 0}, OutBlock{0, b->cols, a.cols, {}})).first;
                        else if (it->second.first.qout != a.qout) throw InvalidArg("prod: inconsistent charge flow");
                        add_seg(it->second.second, opnd(SLOT_OP, *b, 0, true), opnd(SLOT_X, a, 0), 1.0, P.flops);
                    }
                } else {  // A_s B_s^T: match Qout(A) with Qout(B)
                    Key dKB;
                    if (!X2.l[s].dK(dKB)) continue;
                    for (const Blk& a : X1.l[s].b) {
                        const Blk* b = X2.l[s].find(a.qout - dKB);
                        if (!b) continue;
                        auto it = acc.find(a.qin);
                        if (it == acc.end()) it = acc.emplace(a.qin, std::make_pair(Blk{a.qin, b->qin, a.rows, b->rows, 0}, OutBlock{0, a.rows, b->rows, {}})).first;
                        else if (it->second.first.qout != b->qin) throw InvalidArg("prod: inconsistent charge flow");
                        add_seg(it->second.second, opnd(SLOT_X, a, 0), opnd(SLOT_OP, *b, 0, true), 1.0, P.flops);
                    }
                }
            }
            out.l.resize(1);
            for (auto& kv : acc) {
                Blk b = kv.second.first;
                b.off = pos;
                OutBlock ob = std::move(kv.second.second);
                ob.c_off = pos;
                pos += b.size();
                out.l[0].b.push_back(b);
                obs.push_back(std::move(ob));
            }
            break;
        }
        case PROD_REDOP_MPS:
        case PROD_MPS_REDOP: {
            const size_t d = X2.l.size();
            if (X1.l.size() != d * d) throw InvalidArg("prod: R must have nlist = d*d");
            out.l.resize(d);
            for (size_t i = 0; i < d; ++i) {
                std::map<Key, std::pair<Blk, OutBlock>> acc;
                for (size_t j = 0; j < d; ++j) {
                    const BList& R = X1.l[i * d + j];
                    if (kind == PROD_REDOP_MPS) {  // R block (q -> p) times A[j] block (p -> p')
                        for (const Blk& h : R.b) {
                            const Blk* a = X2.l[j].find(h.qout);
                            if (!a) continue;
                            auto it = acc.find(h.qin);
                            if (it == acc.end()) it = acc.emplace(h.qin, std::make_pair(Blk{h.qin, a->qout, h.rows, a->cols, 0}, OutBlock{0, h.rows, a->cols, {}})).first;
                            else if (it->second.first.qout != a->qout) throw InvalidArg("prod: inconsistent charge flow");
                            add_seg(it->second.second, opnd(SLOT_X, h, 0), opnd(SLOT_OP, *a, 0), 1.0, P.flops);
                        }
                    } else {  // A[j] block (q -> p) times R block (p -> p')
                        for (const Blk& a : X2.l[j].b) {
                            const Blk* h = R.find(a.qout);
                            if (!h) continue;
                            auto it = acc.find(a.qin);
                            if (it == acc.end()) it = acc.emplace(a.qin, std::make_pair(Blk{a.qin, h->qout, a.rows, h->cols, 0}, OutBlock{0, a.rows, h->cols, {}})).first;
                            else if (it->second.first.qout != h->qout) throw InvalidArg("prod: inconsistent charge flow");
                            add_seg(it->second.second, opnd(SLOT_OP, a, 0), opnd(SLOT_X, *h, 0), 1.0, P.flops);
                        }
                    }
                }
                for (auto& kv : acc) {
                    Blk b = kv.second.first;
                    b.off = pos;
                    OutBlock ob = std::move(kv.second.second);
                    ob.c_off = pos;
                    pos += b.size();
                    out.l[i].b.push_back(b);
                    obs.push_back(std::move(ob));
                }
            }
            break;
        }
        default:
            throw InvalidArg("prod: unknown kind");
    }
    out.nelem = pos;
    P.y_elems = pos;
    emit_stage(P, obs, SLOT_OUT);
    P.operand_elems_read = X1.nelem + X2.nelem;
    return P;
}

// ---------------------------------------------------------------------------------------------
// ReducedOp (MPSBlockUtilities.hpp:39-74).  Single stage; the A tensor is the "x" operand.
// l: R[i1][j1] = sum h * A[i0]^T A[j0] ; r: R[i0][j0] = sum h * A[j1] A[i1]^T.
// The nonzeros of H are visited column by column (the order of an Armadillo SpMat iterator).
inline Program build_redop(const double* H, int d, const Layout& A, int dir, Layout& out) {
    if ((int)A.l.size() != d) throw InvalidArg("redop: A must have nlist = d");
    Program P;
    P.x_elems = A.nelem;
    const int D = d * d;
    struct Acc { Key qout; int rows, cols; OutBlock ob; };
    std::vector<std::map<Key, Acc>> R((size_t)D);
    for (int c = 0; c < D; ++c)
        for (int r = 0; r < D; ++r) {
            const double h = H[r + (size_t)c * D];
            if (h == 0.0) continue;
            const int i0 = r / d, i1 = r % d, j0 = c / d, j1 = c % d;
            if (dir == IMPS_LEFT) {
                auto& Rm = R[(size_t)i1 * d + j1];
                for (const Blk& a : A.l[i0].b) {          // A[i0]^T block: qout(a) -> qin(a)
                    const Blk* b = A.l[j0].find(a.qin);    // A[j0] block qin(a) -> ...
                    if (!b) continue;
                    auto it = Rm.find(a.qout);
                    if (it == Rm.end()) it = Rm.emplace(a.qout, Acc{b->qout, a.cols, b->cols, OutBlock{0, a.cols, b->cols, {}}}).first;
                    else if (it->second.qout != b->qout) throw InvalidArg("redop: inconsistent charge flow");
                    add_seg(it->second.ob, opnd(SLOT_X, a, 0, true), opnd(SLOT_X, *b, 0), h, P.flops);
                }
            } else {
                auto& Rm = R[(size_t)i0 * d + j0];
                Key dK1;
                if (!A.l[i1].dK(dK1)) continue;
                for (const Blk& a : A.l[j1].b) {          // A[j1] block p -> p'
                    const Blk* b = A.l[i1].find(a.qout - dK1);  // A[i1] block with qout = p'
                    if (!b) continue;
                    auto it = Rm.find(a.qin);
                    if (it == Rm.end()) it = Rm.emplace(a.qin, Acc{b->qin, a.rows, b->rows, OutBlock{0, a.rows, b->rows, {}}}).first;
                    else if (it->second.qout != b->qin) throw InvalidArg("redop: inconsistent charge flow");
                    add_seg(it->second.ob, opnd(SLOT_X, a, 0), opnd(SLOT_X, *b, 0, true), h, P.flops);
                }
            }
        }
    out = Layout();
    set_group(out, A);
    out.l.resize(D);
    long pos = 0;
    std::vector<OutBlock> obs;
    for (int e = 0; e < D; ++e)
        for (auto& kv : R[e]) {
            Blk b{kv.first, kv.second.qout, kv.second.rows, kv.second.cols, pos};
            out.l[e].b.push_back(b);
            OutBlock ob = std::move(kv.second.ob);
            ob.c_off = pos;
            pos += b.size();
            obs.push_back(std::move(ob));
        }
    out.nelem = pos;
    P.y_elems = pos;
    emit_stage(P, obs, SLOT_OUT);
    P.operand_elems_read = A.nelem;
    return P;
}

}  // namespace imps
