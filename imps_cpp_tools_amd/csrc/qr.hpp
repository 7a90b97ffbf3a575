// Batched per-sector QR of MPS tensors (see qr_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
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
    long woff;         // element offset of the sector's work area in device memory (-1: factorised in LDS)
};

// Sectors whose work area (matrix, taus, scales, R diagonal) fits kQrLdsMax are factorised in LDS, larger ones
// in a device work buffer, so that no sector size is out of reach; 16-wave workgroups either way.
// secs lists the LDS sectors first (nlds of them).
constexpr size_t kQrLdsMax = 160 * 1024 - 256;  // leaves room for the kernel's static LDS
constexpr int kQrWaves = 16;

struct QrPlanData {
    std::vector<QrSector> secs;
    std::vector<QrPart> parts;
    Layout R;          // BlockDiagMat layout of R
    size_t lds = 0;    // dynamic LDS bytes (largest LDS sector)
    int nlds = 0;      // sectors factorised in LDS (the first nlds of secs)
    long gwork = 0;    // doubles of the device work buffer of the other sectors
};

// doubles of one sector's work area: m x n matrix, and min(m, n) each of taus, scales and R's diagonal
inline long sector_qr_work(int m, int n) { return (long)m * n + 3L * std::min(m, n); }
hipError_t launch_sector_qr(const QrSector* secs, int nlds, int nglobal, const QrPart* parts, const double* A, double* Q, double* R,
                            bool right, size_t lds_bytes, double* gwork, hipStream_t s);

// Sector regrouping of GetSingleSecData (MPSBlockMat.hpp:636-719): sectors keyed by the right
// charge (left form) or left charge (right form), parts in ascending physical index.
inline QrPlanData build_qr(const Layout& A, int dir) {
    struct Part { int s; const Blk* b; };
    std::map<Key, std::vector<Part>> secs;
    for (size_t s = 0; s < A.l.size(); ++s)
        for (const Blk& b : A.l[s].b) secs[dir == IMPS_LEFT ? b.qout : b.qin].push_back(Part{(int)s, &b});
    QrPlanData P;
    set_group(P.R, A);
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
        const size_t lds = sizeof(double) * (size_t)sector_qr_work(sc.m, sc.n);
        if (lds <= kQrLdsMax) {
            sc.woff = -1;
            P.lds = std::max(P.lds, lds);
        } else {
            sc.woff = P.gwork;
            P.gwork += (sector_qr_work(sc.m, sc.n) + 31) / 32 * 32;  // 256-byte aligned areas
        }
        P.secs.push_back(sc);
    }
    P.R.nelem = rpos;
    std::stable_partition(P.secs.begin(), P.secs.end(), [](const QrSector& s) { return s.woff < 0; });
    P.nlds = (int)std::count_if(P.secs.begin(), P.secs.end(), [](const QrSector& s) { return s.woff < 0; });
    return P;
}

}  // namespace imps
