// Host check of the panel program builder (imps_cpp_tools_amd/csrc/panel.hpp): for an H_AC
// program read from a layout file, every task's chunks cover each of its block products' k exactly
// as often as the product has k (no product lost or doubled, short products packed at most four
// to a chunk, one k-run per lane group), and packed chunks keep to the task's rows.  Prints "ok"
// and the chunk counts per precision, or the first violation.
#include <cstdio>
#include <map>
#include <vector>

#include "panel.hpp"

using namespace imps;

static Layout read_layout(FILE* f) {
    int nsym = 0, nl = 0;
    if (fscanf(f, "%d %d", &nsym, &nl) != 2) throw InvalidArg("layout file: header");
    std::vector<int> nblk(nl);
    int nb = 0;
    for (auto& x : nblk) {
        if (fscanf(f, "%d", &x) != 1) throw InvalidArg("layout file: nblk");
        nb += x;
    }
    std::vector<int> qin(nb * nsym), qout(nb * nsym), rows(nb), cols(nb);
    for (int b = 0; b < nb; ++b) {
        for (int s = 0; s < nsym; ++s)
            if (fscanf(f, "%d", &qin[b * nsym + s]) != 1) throw InvalidArg("layout file: qin");
        for (int s = 0; s < nsym; ++s)
            if (fscanf(f, "%d", &qout[b * nsym + s]) != 1) throw InvalidArg("layout file: qout");
        if (fscanf(f, "%d %d", &rows[b], &cols[b]) != 2) throw InvalidArg("layout file: shape");
    }
    imps_layout c{};
    c.nsym = nsym; c.nlist = nl; c.nblk = nblk.data(); c.qin = qin.data(); c.qout = qout.data(); c.rows = rows.data(); c.cols = cols.data();
    return layout_from_c(&c, "layout file");
}

static int fail(const char* what, int prec, long task) {
    printf("FAIL prec %d task %ld: %s\n", prec, task, what);
    return 1;
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    FILE* f = fopen(argv[1], "r");
    if (!f) return 2;
    Layout AC = read_layout(f), EHL = read_layout(f), EHR = read_layout(f), HL = read_layout(f), HR = read_layout(f);
    fclose(f);
    const Program P = build_hac(AC, EHL, EHR, HL, HR);
    for (int prec : {IMPS_FP64, IMPS_FP32, IMPS_BF16X3}) {
        const PanelProgram Q = build_panel_program(P, prec);
        const int KC = panel_kc(prec), RUN = panel_run(prec);
        std::map<long, const PanelDesc*> packed;  // first fragment -> packed panel
        for (const PanelDesc& pd : Q.panels) {
            if (pd.frag0 < 0 || pd.frag0 + (long)pd.nrg * pd.nkc > Q.nfrag) return fail("panel fragments outside the store", prec, -1);
            if (pd.npk > 0) packed[pd.frag0] = &pd;
        }
        long npacked = 0, nchunks = 0;
        for (size_t si = 0; si < P.stages.size(); ++si) {
            const auto& st = P.stages[si];
            const PStage& ps = Q.stages[si];
            if (st.tasks.size() != ps.tasks.size()) return fail("task count", prec, -1);
            // the builder reorders tasks; match them by output tile
            std::map<std::tuple<long, int, int>, long> want;
            for (const auto& t : st.tasks) {
                long k = 0;
                for (int s = t.seg0; s < t.seg0 + t.nseg; ++s) k += st.segs[s].K;
                want[std::make_tuple(t.c_off, t.m0, t.n0)] = k;
            }
            for (size_t ti = 0; ti < ps.tasks.size(); ++ti) {
                const PTask& t = ps.tasks[ti];
                long got = 0;
                for (int c = t.ch0; c < t.ch0 + t.nch; ++c) {
                    const PChunk& ch = ps.chunks[c];
                    ++nchunks;
                    if (ch.flags & PCF_PACKED) {
                        ++npacked;
                        auto it = packed.find(ch.pan);
                        if (it == packed.end()) return fail("packed chunk without a packed panel", prec, (long)ti);
                        const PanelDesc& pd = *it->second;
                        if (ch.raw_off < 0 || ch.raw_off >= (long)ps.packs.size()) return fail("pack index", prec, (long)ti);
                        const RawPack& rp = ps.packs[ch.raw_off];
                        if (pd.npk < 1 || pd.npk > 4 || pd.nrg < 1 || pd.nrg > 2) return fail("pack shape", prec, (long)ti);
                        if (pd.r0 != 16 * (((ch.flags & PCF_PANEL_IS_A) ? t.m0 : t.n0) / 16)) return fail("pack rows", prec, (long)ti);
                        if ((ch.pan_sub >= 0) != (pd.nrg > 1)) return fail("pack second row group", prec, (long)ti);
                        for (int g = 0; g < 4; ++g) {
                            if (g < pd.npk) {
                                if (pd.pK[g] < 1 || pd.pK[g] > RUN || rp.krem[g] < pd.pK[g]) return fail("run length", prec, (long)ti);
                                got += pd.pK[g];
                            } else if (pd.pK[g] != 0 || rp.krem[g] != 0) {
                                return fail("unused lane group not empty", prec, (long)ti);
                            }
                        }
                    } else {
                        if (ch.krem < 1) return fail("empty chunk", prec, (long)ti);
                        got += ch.krem < KC ? ch.krem : KC;
                    }
                }
                auto w = want.find(std::make_tuple(t.c_off, t.m0, t.n0));
                if (w == want.end()) return fail("task not in the program", prec, (long)ti);
                if (w->second != got) return fail("k covered != sum of the products' K", prec, (long)ti);
            }
        }
        printf("prec %d chunks %ld packed %ld\n", prec, nchunks, npacked);
        if (npacked == 0) return fail("no packed chunk in this workload", prec, -1);
    }
    printf("ok\n");
    return 0;
}
