// Batched per-sector QR of MPS tensors (see qr_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <vector>

#include "layout.hpp"

namespace imps {

// One part of a dense sector matrix: the block of MPS slice s whose rows (left form) or columns
// (right form) fill rows [off, off + len) of the factorised matrix.
struct QrPart {
    long src;   // element offset of the block in the MPS buffer
    long qdst;  // element offset of the Q block in the output buffer (same layout as the input)
    int ld;     // leading dimension (rows) of the block
    int len;    // rows (left) or columns (right) of the block
    int off;    // first row in the factorised matrix
    int pad;
};
struct QrSector {
    int m, n;          // factorised matrix: m x n (left: stacked rows x sector size; right: transposed)
    int part0, nparts;
    long rdst;         // element offset of the sector's R block
};

struct QrPlanData {
    std::vector<QrSector> secs;
    std::vector<QrPart> parts;
    Layout R;          // BlockDiagMat layout of R
    size_t lds = 0;    // dynamic LDS bytes (largest sector)
};

size_t sector_qr_lds_bytes(int m, int n);
hipError_t launch_sector_qr(const QrSector* secs, int nsec, const QrPart* parts, const double* A, double* Q, double* R, bool right,
                            size_t lds_bytes, hipStream_t s);

// Sector regrouping of GetSingleSecData (MPSBlockMat.hpp:636-719): sectors keyed by the right
// charge (left form) or left charge (right form), parts in ascending physical index.
inline QrPlanData build_qr(const Layout& A, int dir) {
    struct Part { int s; const Blk* b; };
    std::map<Key, std::vector<Part>> secs;
    for (size_t s = 0; s < A.l.size(); ++s)
        for (const Blk& b : A.l[s].b) secs[dir == IMPS_LEFT ? b.qout : b.qin].push_back(Part{(int)s, &b});
    QrPlanData P;
    P.R.nsym = A.nsym;
    P.R.l.resize(1);
    long rpos = 0;
    for (auto& kv : secs) {
        QrSector sc;
        sc.part0 = (int)P.parts.size();
        sc.nparts = (int)kv.second.size();
        int off = 0, width = -1;
        for (auto& p : kv.second) {
            const int len = dir == IMPS_LEFT ? p.b->rows : p.b->cols;
            const int w = dir == IMPS_LEFT ? p.b->cols : p.b->rows;
            if (width >= 0 && w != width) throw InvalidArg("qr: inconsistent sector size");
            width = w;
            P.parts.push_back(QrPart{p.b->off, p.b->off, p.b->rows, len, off, 0});
            off += len;
        }
        sc.m = off;
        sc.n = width;
        if (sc.m < sc.n) throw InvalidArg("qr: a sector has fewer stacked rows than its size; Q would change the block structure");
        sc.rdst = rpos;
        Blk rb{kv.first, kv.first, sc.n, sc.n, rpos};
        rpos += rb.size();
        P.R.l[0].b.push_back(rb);
        P.secs.push_back(sc);
        P.lds = std::max(P.lds, sector_qr_lds_bytes(sc.m, sc.n));
    }
    P.R.nelem = rpos;
    if (P.lds > 160 * 1024) throw InvalidArg("qr: sector matrix too large for LDS");
    return P;
}

}  // namespace imps
