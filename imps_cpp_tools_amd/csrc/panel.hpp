// Panel programs: the execution form of a grouped-GEMM Program whose every segment multiplies
// one operator-store block by one vector (or workspace) block, which is the case for the
// transfer matrices, H_AC and H_C.
//
// Operator blocks are pre-packed once per set_operator into *panels*: zero-padded, pre-converted
// (FP64, FP32, or split BF16 hi/lo), already in the MFMA register layout, so that a wave loads a
// 16 x KC operator fragment with one or two fully coalesced 16-byte loads per lane and no
// conversion, clamping or masking.  A panel P(r, k) holds alpha * op(S) with r the output-side
// index (rows of C when the block is the left factor, columns of C when it is the right factor)
// and k the contracted index; rows >= R and k >= K are zero.  Fragment (rg, kc) covers
// r in [16 rg, 16 rg + 16), k in [KC kc, KC kc + KC) and is stored as 64 lanes x LANE elements:
// lane l holds r = 16 rg + l % 16 and the run of consecutive k starting at KC kc + RUN (l / 16).
//
// The other factor (the "raw" operand) is read from the FP64 input vector or workspace in its
// natural layout and converted in registers.  Reads are buffer loads; the elements of the padded
// 32 x KC region that lie outside the block (rows/columns past the block edge, k >= K) are given
// an out-of-range offset, for which the hardware returns zero without touching memory, so that
// small blocks cost only the cache lines they occupy.
//
// Every task (output tile <= 32 x 32) owns an ordered list of chunks (segment, K-range of KC).
// Tasks are ordered by decreasing chunk count so that the longest ones start first.
#pragma once

#include <algorithm>
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

// PCF_STAGE: the raw factor is k-contiguous in memory (s_k == 1).  Its fragment would then need
// one cache line per output-side index for every load; instead the wave reads the whole 32 x KC
// region with consecutive lanes on consecutive k (coalesced), parks it in LDS and reads it back
// in the MFMA layout.
enum PChunkFlag : int { PCF_PANEL_IS_A = 1, PCF_STAGE = 2, PCF_RAW_W = 4 };

struct PChunk {       // 32 bytes
    long raw_off;     // element offset of raw(idx0, k0) within one vector's slot
    int pan;          // fragment index of the panel fragment (rg0, kc)
    int pan_sub;      // fragment index step to the second 16-row group (-1 if the panel has none)
    int s_idx, s_k;   // raw element strides along the output-side index and along k
    int flags;
    int krem;         // valid k in this chunk and after it (K - k0)
};

struct PTask {        // 40 bytes
    long c_off;
    int c_ld, m0, n0, M, N;
    int ch0, nch;
};

struct PanelDesc {    // one packed operator block
    long src;         // element offset in the FP64 operator store
    long frag0;       // first fragment index
    int ld, tr;       // P(r, k) = alpha * (tr ? S[k + r*ld] : S[r + k*ld])
    int R, K;
    int nrg, nkc;
    double alpha;
};

struct PStage {
    int out_slot = SLOT_OUT;
    std::vector<PTask> tasks;
    std::vector<PChunk> chunks;
};

struct PanelProgram {
    int prec = IMPS_FP64;
    std::vector<PStage> stages;
    std::vector<PanelDesc> panels;
    long nfrag = 0;   // panel fragments in total
};

// True when every segment of the program has exactly one operator-store operand.
inline bool panel_compatible(const Program& P) {
    for (auto& st : P.stages)
        for (auto& s : st.segs)
            if ((s.a_slot == SLOT_OP) == (s.b_slot == SLOT_OP)) return false;
    return P.op_elems > 0;
}

inline PanelProgram build_panel_program(const Program& P, int prec) {
    // raw reads use 32-bit byte offsets within one vector; buffer sizes are 32-bit byte counts
    if (P.x_elems >= (1L << 28) || P.w_elems >= (1L << 28)) throw InvalidArg("panel: vector too long for 32-bit offsets");
    PanelProgram Q;
    Q.prec = prec;
    const int KC = panel_kc(prec);
    std::map<std::tuple<long, int, double>, int> index;  // (src, tr, alpha) -> panel
    for (auto& st : P.stages) {
        PStage ps;
        ps.out_slot = st.out_slot;
        std::vector<std::pair<int, PTask>> order;
        std::vector<std::vector<PChunk>> tchunks;
        for (auto& t : st.tasks) {
            std::vector<PChunk> chs;
            for (int si = t.seg0; si < t.seg0 + t.nseg; ++si) {
                const Seg& s = st.segs[si];
                const bool pa = s.a_slot == SLOT_OP;
                // panel side: P(r, k) = op(A)(r, k)      if the operator block is A
                //             P(r, k) = op(B)(k, r)      if it is B
                const long src = pa ? s.a_off : s.b_off;
                const int ld = pa ? s.a_ld : s.b_ld;
                const int ptr = pa ? s.a_tr : (s.b_tr ? 0 : 1);
                const int R = pa ? t.M : t.N;
                auto key = std::make_tuple(src, ptr, s.alpha);
                auto it = index.find(key);
                if (it == index.end()) {
                    PanelDesc pd;
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
                // raw side: element (idx, k) of the raw factor as stored
                //   raw = A (pa false): op(A)(i, k) = a_tr ? S[k + i*ld] : S[i + k*ld]
                //   raw = B (pa true):  op(B)(k, j) = b_tr ? S[j + k*ld] : S[k + j*ld]
                int s_idx, s_k;
                long roff;
                int rslot, idx0;
                if (pa) {
                    rslot = s.b_slot; roff = s.b_off; idx0 = t.n0;
                    if (s.b_tr) { s_idx = 1; s_k = s.b_ld; } else { s_idx = s.b_ld; s_k = 1; }
                } else {
                    rslot = s.a_slot; roff = s.a_off; idx0 = t.m0;
                    if (s.a_tr) { s_idx = s.a_ld; s_k = 1; } else { s_idx = 1; s_k = s.a_ld; }
                }
                if (rslot != SLOT_X && rslot != SLOT_W) throw InvalidArg("panel: raw operand must be the vector or the workspace");
                const int rg0 = (pa ? t.m0 : t.n0) / 16;
                for (int kc = 0; kc < pd.nkc; ++kc) {
                    PChunk c;
                    c.raw_off = roff + (long)idx0 * s_idx + (long)kc * KC * s_k;
                    c.pan = (int)(pd.frag0 + (long)rg0 * pd.nkc + kc);
                    c.pan_sub = rg0 + 1 < pd.nrg ? pd.nkc : -1;  // -1: no second row group
                    c.s_idx = s_idx;
                    c.s_k = s_k;
                    c.flags = (pa ? PCF_PANEL_IS_A : 0) | (rslot == SLOT_W ? PCF_RAW_W : 0) | (s_k == 1 && s_idx != 1 ? PCF_STAGE : 0);
                    c.krem = s.K - kc * KC;
                    if (c.raw_off + 31L * s_idx + (long)(KC - 1) * s_k >= (1L << 28))
                        throw InvalidArg("panel: raw operand offsets exceed 32-bit range");
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
    if (Q.nfrag * panel_frag_bytes(prec) >= (1L << 31)) throw InvalidArg("panel store too large for 32-bit offsets");
    return Q;
}

}  // namespace imps
