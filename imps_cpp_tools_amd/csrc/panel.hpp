// Panel programs: the execution form of a grouped-GEMM Program whose every segment multiplies
// one operator-store block by one vector (or workspace) block, which is the case for the
// transfer matrices, H_AC and H_C.
//
// Operator blocks are pre-packed once per set_operator into *panels*: zero-padded, pre-converted
// (FP64, FP32, or split BF16 hi/lo), already in the MFMA register layout, so that a wave loads a
// 16 x KC operator fragment with fully coalesced 16-byte loads and no conversion or masking.
// A panel P(r, k) holds alpha * op(S) with r the output-side index (rows of C when the block is
// the left factor, columns of C when it is the right factor) and k the contracted index; rows
// >= R and k >= K are zero.  Fragment (rg, kc) covers r in [16 rg, 16 rg + 16) and
// k in [KC kc, KC kc + KC) and is stored as 64 lanes x RUN elements: lane l holds
// r = 16 rg + l % 16 and the RUN consecutive k starting at KC kc + RUN (l / 16).
//
// The other factor (the "raw" operand: a block of the input vector or of the workspace) is read
// from a *split copy*: the vector converted once per apply to the compute format and stored
// twice, as is (plane N) and with every block transposed (plane T), so that whichever way a
// segment uses the block, the k index is contiguous and a lane's RUN consecutive k come in one
// or two 16-byte loads.  BF16X3 keeps separate hi and lo planes.  Split copy of one vector:
//   FP64/FP32: [N | T], each PL elements;   BF16X3: [N_hi | T_hi | N_lo | T_lo], each PL elements,
// PL = element count + a zero pad of at least KC, so that a k-run past the end of the last block
// stays inside the buffer (such elements meet zero panel entries).  Workspace stages write their
// output directly in split form.
//
// Every task (output tile <= 32 x 32) owns an ordered list of chunks (segment, K-range of KC).
// Tasks are ordered by decreasing chunk count so that the longest ones start first.
#pragma once

#include <algorithm>
#include <array>
#include <cstring>
#include <map>
#include <numeric>
#include <tuple>
#include <vector>

#include "layout.hpp"

namespace imps {

// Contracted-index step of one chunk and k-run length per lane, by precision.
inline int panel_kc(int prec) { return prec == IMPS_BF16X3 ? 32 : 16; }
inline int panel_run(int prec) { return prec == IMPS_BF16X3 ? 8 : 4; }
// Bytes of one 64-lane fragment: BF16X3 stores a hi plane and a lo plane of 64 x 8 bf16 each.
inline long panel_frag_bytes(int prec) { return prec == IMPS_FP64 ? 64 * 4 * 8 : prec == IMPS_FP32 ? 64 * 4 * 4 : 2 * 64 * 8 * 2; }
// Plane length of a split copy of `len` elements, and bytes of one vector's split copy.
inline long split_plane(long len) { return (len + 64 + 63) / 64 * 64; }
inline long split_elem_bytes(int prec) { return prec == IMPS_FP64 ? 8 : prec == IMPS_FP32 ? 4 : 2; }
inline long split_bytes(int prec, long len) { return (prec == IMPS_BF16X3 ? 4 : 2) * split_plane(len) * split_elem_bytes(prec); }

// PCF_RAW_XN: the raw factor is an input-vector block read k-contiguously (plane N).  In FP64 plane N
// equals the input itself, so such chunks read the caller's vector directly when it is passed to the
// kernel (no copy of plane N), else the split copy's plane N.
// PCF_RAW_XT: the raw factor is an input-vector block read along its stored columns (plane T of the
// split copy).  When the apply passes the caller's vector for these chunks too (few vectors), they
// read it directly, element (idx, k) at xoff + idx + k * xks: each lane loads its k one by one, and
// the sixteen lanes of one k read sixteen consecutive elements; no split copy is made then.  FP32
// and BF16X3 convert the FP64 elements as the split copy does, plane-N chunks included.
// PCF_PACKED: the chunk's k-slots hold up to four short products of the task (K <= the k-run of a
// lane group, one product per lane group: the sectors of dimension 1..4, or 1..8 in BF16X3), whose
// panel is packed the same way; raw_off indexes the stage's RawPack table, which gives each lane
// group its own raw block.
enum PChunkFlag : int { PCF_PANEL_IS_A = 1, PCF_RAW_W = 4, PCF_RAW_XN = 8, PCF_RAW_XT = 16, PCF_PACKED = 32 };

// Raw side of a packed chunk, per lane group g: raw(idx0, 0) of product g (raw_off, s_idx as in
// PChunk), its K (krem; 0: no product, the group reads zeros) and, PCF_RAW_XT, its input offset
// and k stride.
struct RawPack {      // 96 bytes
    long raw_off[4];
    int s_idx[4];
    int krem[4];
    int xoff[4];
    int xks[4];
};

struct PChunk {       // 40 bytes
    long raw_off;     // element index, within the (hi) N|T planes, of raw(idx0, k0)
    int pan;          // fragment index of the panel fragment (rg0, kc)
    int pan_sub;      // fragment index step to the second 16-row group (-1 if the panel has none)
    int s_idx;        // raw element stride along the output-side index (k is contiguous)
    int flags;
    int krem;         // valid k in this chunk and after it (K - k0)
    int xoff;         // PCF_RAW_XT: element offset in the input vector of raw(idx0, k0)
    int xks;          // PCF_RAW_XT: element stride along k in the input vector (the block's row count)
    int pad;
};

struct PTask {        // 40 bytes
    long c_off;
    int c_ld, m0, n0, M, N;
    int ch0, nch;
};

struct PanelDesc {    // one packed operator block, or (npk > 0) one packed chunk's operator blocks
    long src;         // element offset in the FP64 operator store
    long frag0;       // first fragment index
    int ld, tr;       // P(r, k) = alpha * (tr ? S[k + r*ld] : S[r + k*ld])
    int R, K;
    int nrg, nkc;
    double alpha;
    // npk > 0: one chunk (nkc = 1) of the nrg (<= 2) row groups from row r0, whose lane group
    // g < npk holds block g: P(r, RUN g + q) = palpha[g] * (ptr[g] ? S[q + r*pld[g]] :
    // S[r + q*pld[g]]) at psrc[g], q < pK[g], r0 <= r < R
    int npk, r0;
    long psrc[4];
    double palpha[4];
    int pld[4], ptr[4], pK[4];
};

struct SplitSpan {    // one input-vector block read as a raw factor: offset and shape
    long off;
    int rows, cols;
};
// A split unit is one workgroup's share of the split copy: the flat element range [o0, o0 + n) of
// the vector, covered by the spans b0 .. b0 + nb - 1 (several whole small blocks, or a slice of one
// large block).  Each thread takes elements o of the range and writes plane T at o (the element of
// block (r, c) = divmod(o - off, cols) read from N position off + r + c rows) and plane N at o.
struct SplitUnit {
    long o0;
    int n, b0, nb, pad;
};
constexpr int SPLIT_UNIT = 1024;      // elements per unit at most
constexpr int SPLIT_UNIT_SPANS = 64;  // spans per unit at most (staged in LDS)

struct PStage {
    int out_slot = SLOT_OUT;
    std::vector<PTask> tasks;
    std::vector<PChunk> chunks;
    std::vector<RawPack> packs;
};

struct PanelProgram {
    int prec = IMPS_FP64;
    std::vector<PStage> stages;
    std::vector<PanelDesc> panels;
    std::vector<SplitSpan> xspans;    // input-vector blocks to split per apply, by offset
    std::vector<SplitUnit> xunits;    // their packing into workgroup units
    long nfrag = 0;                   // panel fragments in total
    long plx = 0, plw = 0;            // split plane lengths of the input vector and the workspace
};

// True when every segment of the program has exactly one operator-store operand.
inline bool panel_compatible(const Program& P) {
    for (auto& st : P.stages)
        for (auto& s : st.segs)
            if ((s.a_slot == SLOT_OP) == (s.b_slot == SLOT_OP)) return false;
    return P.op_elems > 0;
}

// Packs the split spans (by ascending offset; disjoint) into units: consecutive small blocks share
// a unit, a block longer than SPLIT_UNIT is cut into slices, and the last slice of a block stays
// open for the small blocks after it.
inline void pack_split_units(const std::map<long, SplitSpan>& xb, std::vector<SplitSpan>& spans, std::vector<SplitUnit>& units) {
    for (auto& kv : xb) {
        const SplitSpan& b = kv.second;
        const long len = (long)b.rows * b.cols;
        if (len == 0) continue;
        const int bi = (int)spans.size();
        spans.push_back(b);
        SplitUnit* u = units.empty() ? nullptr : &units.back();
        if (u && u->nb < SPLIT_UNIT_SPANS && b.off + len - u->o0 <= SPLIT_UNIT) {
            u->n = (int)(b.off + len - u->o0);
            ++u->nb;
            continue;
        }
        for (long o = 0; o < len; o += SPLIT_UNIT) units.push_back(SplitUnit{b.off + o, (int)std::min<long>(SPLIT_UNIT, len - o), bi, 1, 0});
    }
}

inline PanelProgram build_panel_program(const Program& P, int prec) {
    PanelProgram Q;
    Q.prec = prec;
    Q.plx = split_plane(P.x_elems);
    Q.plw = split_plane(P.w_elems);
    // buffer loads use 32-bit byte offsets within one vector's split copy
    if (split_bytes(prec, P.x_elems) >= (1L << 31) || split_bytes(prec, P.w_elems) >= (1L << 31))
        throw InvalidArg("panel: vector too long for 32-bit offsets");
    const int KC = panel_kc(prec);
    const int RUN = panel_run(prec);
    std::map<std::tuple<long, int, double>, int> index;  // (src, tr, alpha) -> panel
    std::map<std::array<long, 22>, int> pindex;            // packed chunk panel -> panel
    std::map<long, SplitSpan> xb;
    for (auto& st : P.stages) {
        PStage ps;
        ps.out_slot = st.out_slot;
        std::vector<std::pair<int, PTask>> order;
        std::vector<std::vector<PChunk>> tchunks;
        for (auto& t : st.tasks) {
            std::vector<PChunk> chs;
            // short products (K <= RUN) of one kind (panel role, raw slot, raw orientation) wait
            // here to be packed four to a chunk
            std::map<int, std::vector<int>> pending;
            for (int si = t.seg0; si < t.seg0 + t.nseg; ++si) {
                const Seg& s = st.segs[si];
                const bool pa = s.a_slot == SLOT_OP;
                // panel side: P(r, k) = op(A)(r, k)      if the operator block is A
                //             P(r, k) = op(B)(k, r)      if it is B
                const long src = pa ? s.a_off : s.b_off;
                const int ld = pa ? s.a_ld : s.b_ld;
                const int ptr = pa ? s.a_tr : (s.b_tr ? 0 : 1);
                const int R = pa ? t.M : t.N;
                // raw side.  The stored raw block S (rows Rb, cols Cb, at boff) is used as
                //   raw = A: op(A)(i, k) = a_tr ? S(k, i) : S(i, k)     (i runs over the output rows)
                //   raw = B: op(B)(k, j) = b_tr ? S(j, k) : S(k, j)     (j runs over the output columns)
                // so raw(idx, k) is S(idx, k) ("idx-major") or S(k, idx) ("k-major").  A k-major
                // element sits in plane N at boff + k + idx*Rb; an idx-major one in plane T, where
                // S(r, c) is stored at boff + c + r*Cb, at boff + k + idx*Cb.
                const int rslot = pa ? s.b_slot : s.a_slot;
                if (rslot != SLOT_X && rslot != SLOT_W) throw InvalidArg("panel: raw operand must be the vector or the workspace");
                const long boff = pa ? s.b_off : s.a_off;
                const int rtr = pa ? s.b_tr : s.a_tr;
                const int Rb = pa ? (rtr ? t.N : s.K) : (rtr ? s.K : t.M);
                const int Cb = pa ? (rtr ? s.K : t.N) : (rtr ? t.M : s.K);
                const bool kmajor = pa ? (rtr == 0) : (rtr != 0);
                if (rslot == SLOT_X) xb[boff] = SplitSpan{boff, Rb, Cb};
                if (s.K <= RUN) {
                    pending[(pa ? 1 : 0) | (rslot == SLOT_W ? 2 : 0) | (kmajor ? 4 : 0)].push_back(si);
                    continue;
                }
                auto key = std::make_tuple(src, ptr, s.alpha);
                auto it = index.find(key);
                if (it == index.end()) {
                    PanelDesc pd{};
                    pd.src = src; pd.ld = ld; pd.tr = ptr; pd.R = R; pd.K = s.K; pd.alpha = s.alpha;
                    pd.nrg = (R + 15) / 16;
                    pd.nkc = (s.K + KC - 1) / KC;
                    pd.frag0 = Q.nfrag;
                    Q.nfrag += (long)pd.nrg * pd.nkc;
                    it = index.emplace(key, (int)Q.panels.size()).first;
                    Q.panels.push_back(pd);
                }
                const PanelDesc& pd = Q.panels[it->second];
                if (pd.R != R || pd.K != s.K) throw InvalidArg("panel: operator block used with two shapes");
                const int idx0 = pa ? t.n0 : t.m0;
                const long pl = rslot == SLOT_X ? Q.plx : Q.plw;
                const int s_idx = kmajor ? Rb : Cb;
                const long base = kmajor ? boff : pl + boff;
                const int rg0 = (pa ? t.m0 : t.n0) / 16;
                for (int kc = 0; kc < pd.nkc; ++kc) {
                    PChunk c;
                    c.raw_off = base + (long)kc * KC + (long)idx0 * s_idx;
                    c.pan = (int)(pd.frag0 + (long)rg0 * pd.nkc + kc);
                    c.pan_sub = rg0 + 1 < pd.nrg ? pd.nkc : -1;  // -1: no second row group
                    c.s_idx = s_idx;
                    const bool fx = rslot == SLOT_X;
                    c.flags = (pa ? PCF_PANEL_IS_A : 0) | (rslot == SLOT_W ? PCF_RAW_W : 0) | (fx && kmajor ? PCF_RAW_XN : 0) |
                              (fx && !kmajor ? PCF_RAW_XT : 0);
                    c.krem = s.K - kc * KC;
                    // idx-major: S(idx, k) of the stored block (Rb rows, column-major) at boff + idx + k * Rb
                    c.xoff = fx && !kmajor ? (int)(boff + idx0 + (long)kc * KC * Rb) : 0;
                    c.xks = fx && !kmajor ? Rb : 0;
                    c.pad = 0;
                    chs.push_back(c);
                }
            }
            for (auto& kv : pending) {
                const bool pa = (kv.first & 1) != 0, fw = (kv.first & 2) != 0, kmajor = (kv.first & 4) != 0;
                const bool fx = !fw;
                const std::vector<int>& sis = kv.second;
                for (size_t g0 = 0; g0 < sis.size(); g0 += 4) {
                    const int ng = (int)std::min<size_t>(4, sis.size() - g0);
                    const int rg0 = (pa ? t.m0 : t.n0) / 16;
                    PanelDesc pd{};
                    pd.R = pa ? t.M : t.N;
                    pd.K = KC;
                    pd.r0 = 16 * rg0;
                    pd.nrg = std::min(2, (pd.R + 15) / 16 - rg0);  // the task's row groups only
                    pd.nkc = 1;
                    pd.alpha = 1.0;
                    pd.npk = ng;
                    RawPack rp{};
                    const int idx0 = pa ? t.n0 : t.m0;
                    for (int g = 0; g < ng; ++g) {
                        const Seg& s = st.segs[sis[g0 + g]];
                        pd.psrc[g] = pa ? s.a_off : s.b_off;
                        pd.pld[g] = pa ? s.a_ld : s.b_ld;
                        pd.ptr[g] = pa ? s.a_tr : (s.b_tr ? 0 : 1);
                        pd.pK[g] = s.K;
                        pd.palpha[g] = s.alpha;
                        const long boff = pa ? s.b_off : s.a_off;
                        const int rtr = pa ? s.b_tr : s.a_tr;
                        const int Rb = pa ? (rtr ? t.N : s.K) : (rtr ? s.K : t.M);
                        const int Cb = pa ? (rtr ? s.K : t.N) : (rtr ? t.M : s.K);
                        const long pl = fx ? Q.plx : Q.plw;
                        const int s_idx = kmajor ? Rb : Cb;
                        rp.raw_off[g] = (kmajor ? boff : pl + boff) + (long)idx0 * s_idx;
                        rp.s_idx[g] = s_idx;
                        rp.krem[g] = s.K;
                        rp.xoff[g] = fx && !kmajor ? (int)(boff + idx0) : 0;
                        rp.xks[g] = fx && !kmajor ? Rb : 0;
                    }
                    // tasks that pack the same operator blocks over the same rows share the panel
                    std::array<long, 22> key{};
                    for (int g = 0; g < 4; ++g) {
                        long abits;
                        std::memcpy(&abits, &pd.palpha[g], sizeof abits);
                        key[g] = pd.psrc[g]; key[4 + g] = pd.pld[g]; key[8 + g] = pd.ptr[g]; key[12 + g] = pd.pK[g]; key[16 + g] = abits;
                    }
                    key[20] = pd.r0;
                    key[21] = pd.R;
                    auto pit = pindex.find(key);
                    if (pit == pindex.end()) {
                        pd.frag0 = Q.nfrag;
                        Q.nfrag += pd.nrg;
                        pit = pindex.emplace(key, (int)Q.panels.size()).first;
                        Q.panels.push_back(pd);
                    }
                    const PanelDesc& pq = Q.panels[pit->second];
                    PChunk c;
                    c.raw_off = (long)ps.packs.size();
                    c.pan = (int)pq.frag0;
                    c.pan_sub = pq.nrg > 1 ? 1 : -1;
                    c.s_idx = 0;
                    c.flags = PCF_PACKED | (pa ? PCF_PANEL_IS_A : 0) | (fw ? PCF_RAW_W : 0) | (fx && kmajor ? PCF_RAW_XN : 0) |
                              (fx && !kmajor ? PCF_RAW_XT : 0);
                    c.krem = KC;
                    c.xoff = 0;
                    c.xks = 0;
                    c.pad = 0;
                    ps.packs.push_back(rp);
                    chs.push_back(c);
                }
            }
            PTask pt{t.c_off, t.c_ld, t.m0, t.n0, t.M, t.N, 0, (int)chs.size()};
            order.emplace_back((int)tchunks.size(), pt);
            tchunks.push_back(std::move(chs));
        }
        std::stable_sort(order.begin(), order.end(), [](const std::pair<int, PTask>& a, const std::pair<int, PTask>& b) {
            return a.second.nch > b.second.nch;
        });
        for (auto& o : order) {
            PTask t = o.second;
            t.ch0 = (int)ps.chunks.size();
            ps.chunks.insert(ps.chunks.end(), tchunks[o.first].begin(), tchunks[o.first].end());
            ps.tasks.push_back(t);
        }
        Q.stages.push_back(std::move(ps));
    }
    pack_split_units(xb, Q.xspans, Q.xunits);
    if (Q.nfrag * panel_frag_bytes(prec) >= (1L << 31)) throw InvalidArg("panel store too large for 32-bit offsets");
    return Q;
}

}  // namespace imps
