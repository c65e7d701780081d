// Resident solver objects shared by the single-system and the batched drivers: device slice layouts, the
// level-scheduled / multicolour block ILU(0), the CPR-AMG hierarchy, the context and the resident system.
#pragma once
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <tuple>

#include "../../../include/ressim_hip.h"
#include "rsh_internal.hpp"

namespace rsh {
inline thread_local std::string g_err;

inline int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

template <typename F>
int guarded(F&& f) {
    try {
        return f();
    } catch (const Error& e) {
        return fail(e.code, e.what());
    } catch (const std::bad_alloc&) {
        return fail(RSH_ERR_DEVICE, "host allocation failed");
    } catch (const std::exception& e) {
        return fail(RSH_ERR_DEVICE, e.what());
    }
}

// owner system of every slice / long row (-1 when a slice mixes systems); rowSys maps original rows to systems
inline void sliceOwners(SliceHost& h, const std::vector<int>& rowSys) {
    h.owner.assign(size_t(h.nslices), -1);
    for (int sl = 0; sl < h.nslices; ++sl) {
        int o = -2;
        for (int lane = 0; lane < kSliceRows; ++lane) {
            const int r = h.orig.empty() ? h.row[size_t(sl) * kSliceRows + size_t(lane)] : h.orig[size_t(sl) * kSliceRows + size_t(lane)];
            if (r < 0) continue;
            const int q = rowSys[size_t(r)];
            o = (o == -2 || o == q) ? q : -1;
        }
        h.owner[size_t(sl)] = o < 0 ? -1 : o;
    }
    h.lowner.assign(h.lrow.size(), -1);
    for (size_t q = 0; q < h.lrow.size(); ++q) h.lowner[q] = rowSys[size_t(h.lorig.empty() ? h.lrow[q] : h.lorig[q])];
}

template <typename TA>
void uploadSlices(const SliceHost& h, SliceDev<TA>& d, int nb, hipStream_t s) {
    d.nslices = h.nslices;
    d.nlong = int(h.lrow.size());
    d.nb = nb;
    std::vector<int2> info(size_t(h.nslices));
    for (int i = 0; i < h.nslices; ++i) info[size_t(i)] = make_int2(h.width[size_t(i)], h.off[size_t(i)]);
    d.info.upload(info, s);
    d.row.upload(h.row, s);
    d.orig.upload(h.orig, s);
    d.col.upload(h.col, s);
    d.dcol.upload(h.dcol, s);
    d.src.upload(h.src, s);
    d.dsrc.upload(h.dsrc, s);
    d.lrow.upload(h.lrow, s);
    d.lorig.upload(h.lorig, s);
    d.lptr.upload(h.lptr, s);
    d.lcol.upload(h.lcol, s);
    d.lsrc.upload(h.lsrc, s);
    d.ldsrc.upload(h.ldsrc, s);
    d.owner.upload(h.owner, s);
    d.lowner.upload(h.lowner, s);
    const size_t b2 = size_t(nb) * nb;
    d.val.alloc(size_t(h.nslots) * b2 * kWave);
    d.lval.alloc(h.lcol.size() * b2);
    if (!h.dsrc.empty()) d.dval.alloc(size_t(h.nslices) * b2 * kWave);
    if (!h.ldsrc.empty()) d.ldval.alloc(h.ldsrc.size() * b2);
}

template <typename TA>
void packSlices(hipStream_t s, int nb, const double* src, SliceDev<TA>& d, long long nslots, bool f32) {
    launchPackSlices(s, nb, src, d.src.p, nslots, d.val.p, f32);
    if (d.lsrc.n) launchPackLong(s, nb, src, d.lsrc.p, int(d.lsrc.n), d.lval.p, f32);
    if (d.dsrc.n) launchPackSlices(s, nb, src, d.dsrc.p, d.nslices, d.dval.p, f32);
    if (d.ldsrc.n) launchPackLong(s, nb, src, d.ldsrc.p, int(d.ldsrc.n), d.ldval.p, f32);
}

// ------------------------------------------------------------------ dense groups of fused runs
// Per sweep direction: the runs of small levels whose inside couplings split into groups of at most kDenseMaxRows
// rows, with the dense inverse of each group (recomputed after every factorisation).
struct DenseSide {
    std::vector<int> runTask0, runTask1;   // per run: its range in tasks, or -1 when swept level by level
    std::vector<int> runGroup0, runGroup1; // per run: its range of groups, or -1
    std::vector<char> noOutside;           // per run: no entry of its rows couples to a row outside the run
    std::vector<int> hStart, hRows;        // host copies of start / rows (working rows, ascending per group)
    int ngroups = 0, nmap = 0, maxRows = 0;   // maxRows: rows of the largest group
    long long elems = 0;
    DBuf<int> start, rows;
    DBuf<long long> off;
    DBuf<int4> map;
    DBuf<int2> tasks;
    DBuf<double> raw, inv;
    DBuf<float> inv32;
    DBuf<int> gowner;                      // batched systems: the system of each group

    // E: the sweep's entries per working row (src = factor block; dsrc = pivot inverse for the upper sweep);
    // rowSys (optional): original row -> system
    void build(const std::vector<LevelRun>& runs, const RowEntries& E, const std::vector<int>& lstart, int nb, bool upper, hipStream_t s,
               const std::vector<int>* rowSys = nullptr) {
        std::vector<int> hstart(1, 0), hrows, hown;
        std::vector<long long> hoff;
        std::vector<int4> hmap;
        std::vector<int2> htasks;
        ngroups = 0;
        elems = 0;
        runTask0.assign(runs.size(), -1);
        runTask1.assign(runs.size(), -1);
        runGroup0.assign(runs.size(), -1);
        runGroup1.assign(runs.size(), -1);
        noOutside.assign(runs.size(), 0);
        constexpr long long kBudget = 1ll << 26;   // dense elements per direction (512 MiB in FP64)
        for (size_t ri = 0; ri < runs.size(); ++ri) {
            const LevelRun& run = runs[ri];
            if (!run.fused) continue;
            const int wb = lstart[size_t(run.l0)], we = lstart[size_t(run.l1)], m = we - wb;
            std::vector<int> par(static_cast<size_t>(m));
            for (int a = 0; a < m; ++a) par[size_t(a)] = a;
            auto find = [&](int a) {
                while (par[size_t(a)] != a) { par[size_t(a)] = par[size_t(par[size_t(a)])]; a = par[size_t(a)]; }
                return a;
            };
            for (int w = wb; w < we; ++w)
                for (int p = E.ptr[size_t(w)]; p < E.ptr[size_t(w) + 1]; ++p) {
                    const int c = E.col[size_t(p)];
                    if (c < wb || c >= we) continue;
                    const int a = find(w - wb), b = find(c - wb);
                    if (a != b) par[size_t(std::max(a, b))] = std::min(a, b);
                }
            std::vector<int> gid(static_cast<size_t>(m), -1), loc(static_cast<size_t>(m));
            std::vector<std::vector<int>> groups;
            for (int a = 0; a < m; ++a) {
                const int r = find(a);
                if (gid[size_t(r)] < 0) { gid[size_t(r)] = int(groups.size()); groups.emplace_back(); }
                loc[size_t(a)] = int(groups[size_t(gid[size_t(r)])].size());
                groups[size_t(gid[size_t(r)])].push_back(wb + a);
            }
            size_t maxRows = 0;
            long long add = 0;
            for (const auto& g : groups) {
                maxRows = std::max(maxRows, g.size());
                add += (long long)g.size() * nb * (long long)g.size() * nb;
            }
            // one dense product per group pays off against a dependent step per level once the run is a few
            // levels deep (a level step costs about 2.5 us, the dense product about 4 us + 10 ns per scalar row)
            const double levelCost = 2.5 * (run.l1 - run.l0), denseCost = 4.0 + 0.005 * double(maxRows) * nb;
            if (maxRows > size_t(kDenseMaxRows) || elems + add > kBudget || denseCost >= levelCost) continue;
            bool inside = true;
            for (int w = wb; inside && w < we; ++w)
                for (int p = E.ptr[size_t(w)]; p < E.ptr[size_t(w) + 1]; ++p)
                    if (E.col[size_t(p)] < wb || E.col[size_t(p)] >= we) { inside = false; break; }
            noOutside[ri] = inside ? 1 : 0;
            runTask0[ri] = int(htasks.size());
            runGroup0[ri] = ngroups;
            for (const auto& g : groups) {
                const int gi = ngroups++;
                const int K = int(g.size()) * nb;
                hoff.push_back(elems);
                elems += (long long)K * K;
                for (size_t i = 0; i < g.size(); ++i) {
                    const int w = g[i];
                    hrows.push_back(w);
                    for (int p = E.ptr[size_t(w)]; p < E.ptr[size_t(w) + 1]; ++p) {
                        const int c = E.col[size_t(p)];
                        if (c < wb || c >= we) continue;
                        hmap.push_back(make_int4(gi, int(i), loc[size_t(c - wb)], E.src[size_t(p)]));
                    }
                    if (upper) hmap.push_back(make_int4(gi, int(i), int(i), E.dsrc[size_t(w)]));
                }
                hstart.push_back(int(hrows.size()));
                if (rowSys) hown.push_back((*rowSys)[size_t(E.orig[size_t(g[0])])]);
                for (int a0 = 0; a0 < K; a0 += kWave) htasks.push_back(make_int2(gi, a0));
            }
            runTask1[ri] = int(htasks.size());
            runGroup1[ri] = ngroups;
        }
        nmap = int(hmap.size());
        maxRows = 0;
        for (size_t g = 0; g + 1 < hstart.size(); ++g) maxRows = std::max(maxRows, hstart[g + 1] - hstart[g]);
        hStart = hstart;
        hRows = hrows;
        start.upload(hstart, s);
        rows.upload(hrows, s);
        off.upload(hoff, s);
        map.upload(hmap, s);
        tasks.upload(htasks, s);
        gowner.upload(hown, s);
        raw.alloc(size_t(elems));
        inv.alloc(size_t(elems));
        inv32.release();
    }
    void factor(int nb, bool upper, const double* lu, bool f32, hipStream_t s) {
        if (!ngroups) return;
        launchDenseGroups(s, nb, upper, ngroups, maxRows, start.p, off.p, nmap, map.p, lu, raw.p, elems, inv.p);
        if (f32) {
            inv32.alloc(size_t(elems));
            launchToFloat(s, elems, inv.p, inv32.p);
        }
    }
    size_t bytes() const { return raw.bytes() + inv.bytes() + inv32.bytes(); }
};

// a new generation number for a rebuilt schedule or hierarchy: unique over the process, so that an apply captured
// against an object that has since been replaced (a new IluDev after an option change) is never replayed
inline int nextGeneration() {
    static std::atomic<int> g{0};
    return ++g;
}

// ------------------------------------------------------------------ level-scheduled block ILU(0)
struct IluDev {
    bool mc = false;
    int n = 0, nb = 1, colours = 0;
    std::vector<int> lstart;            // level (or colour) starts in working numbering
    DBuf<int> perm;                     // working row -> original row
    DBuf<int> iperm;                    // original row -> working row
    bool firstCopy = false;             // the first level has no lower entries: its forward sweep copies r
    // factorisation matrix: natural mode = the system's pattern (rows visited in level order),
    // multicolour mode = the colour-permuted pattern P A P^T
    HostBsr fpat;
    DBuf<int> frp, fci, fdiag, fmap;
    const int *rp = nullptr, *ci = nullptr, *diag = nullptr;
    DBuf<double> lu;
    SliceHost hL, hU;
    LevelSched sL, sU;
    SliceDev<double> L64, U64;
    SliceDev<float> L32, U32;
    bool have32 = false;
    DBuf<double> y64;
    DBuf<float> y32;
    DBuf<int> nreg;
    int regularised = 0;
    int generation = 0;
    // factorisation schedule: per level the rows eliminated one per thread, then the long rows one per wavefront
    DBuf<int> fShort, fLong;
    std::vector<int> fShortStart, fLongStart;
    // multicolour order: the forward runs swept through dense groups are factorised group by group (the rows before
    // the run first, then kIluGroupFactor); per run its range of factor groups, or -1
    std::vector<int> fgRun0, fgRun1;
    DBuf<int> fgStart, fgRows, fgIdx;
    DBuf<long long> fgIdxOff;
    // sweep schedule: runs of levels (fused small levels or single launches) and device copies of level starts
    std::vector<LevelRun> runsL, runsU;
    DBuf<int> dLs, dLq, dUs, dUq;
    DenseSide denseL, denseU;
    DBuf<double> rt64;                  // outside contributions of the run being swept
    DBuf<float> rt32;
    // the last forward and the first backward step are the same run of dense groups whose rows have no upper
    // couplings outside it (the well-clique colours at the end of the multicolour order): the forward product
    // writes its result in the run's numbering (tail64 / tail32) and the backward product starts from it, without
    // the backward outside pass (which would only copy y)
    bool tailCopy = false;
    DBuf<double> tail64;
    DBuf<float> tail32;
    // level chunking: levels with at most fuseMax tasks are swept in runs (one workgroup or dense groups) of at most
    // runRowsMax rows (the LDS tile of a run's working vector)
    int fuseMax = 16;
    int runRowsMax = kRunRowsMax;
    // batched systems: original row -> system, and the systems' BiCGSTAB states (slices of stopped systems skipped)
    const std::vector<int>* rowSys = nullptr;
    const KScalars* gate = nullptr;
    // natural order: the sweeps tiled by TileSched (one launch per tile level) instead of one launch per level.
    // schedule: 0 tiles where the cost model prefers them, 1 level launches, 2 tiles whenever the pattern admits them;
    // tileDims: the tile extents along the pattern's three dominant lower offsets (bx | by << 8 | bz << 16, 0: chosen)
    int schedule = 0, tileDims = 0;
    bool tiled = false;
    TileSched tiles;
    DBuf<int> tUnit, tLs, tLq, tUs, tUq, tLvPtr, tLv;
    int tLongMax = 0, tLongCap = 0;
    // multicolour order: some rows finish their backward step in the forward launch (their lower slice lanes carry the
    // pivot block, dsrc >= 0) and are left out of the upper slices
    bool fuseBwd = false;
    // red / black tiles instead of the cell colouring (rbDims = bx | by << 8 | bz << 16, 0: off; rbTileOrder): level 0
    // holds the red tiles' rows, level 1 the black ones', the levels after them the rows behind the tiles; levels 0 and
    // 1 are swept tile by tile (kIluTileWave: tilesL / tilesU = each tile's slices, red tiles first) and factorised by
    // local level (factor levels fLevel0[l] .. fLevel0[l + 1] of sweep level l)
    int rbDims = 0;
    bool rbt = false;
    RbTiles rb;
    DBuf<int4> tilesL, tilesU;
    int rbTiles = 0, rbRedTiles = 0, rbMaxRows = 0;
    std::vector<int> fLevel0;
    // natural order: the long rows' update maps (maxLen 0: none, the long rows use kBilu0LongRows)
    DBuf<int> lmRow, lmPtr;
    DBuf<int2> lmMap;
    LongRowMap longMap{nullptr, nullptr, nullptr, 0};
    // kept when keepEntries is set (batched systems re-chunk the levels when tuning)
    bool keepEntries = false;
    RowEntries keptL, keptU;
    std::vector<int> workL, workU, rowsPerLevel;
    void build(const HostBsr& S, const int* drp, const int* dci, const int* ddiag, bool multicolour, hipStream_t s) {
        generation = nextGeneration();
        mc = multicolour;
        n = S.n;
        nb = S.nb;
        std::vector<int> ord, colour;
        RowEntries EL, EU;
        EL.ptr.assign(1, 0);
        EU.ptr.assign(1, 0);
        // sweep numbering of the natural order: level order, or the tile order when the sweeps are tiled
        std::vector<int> sweepOrd;
        auto naturalEntries = [&](const std::vector<int>& order) {
            EL = RowEntries{};
            EU = RowEntries{};
            EL.ptr.assign(1, 0);
            EU.ptr.assign(1, 0);
            std::vector<int> iperm(static_cast<size_t>(n));
            for (int w = 0; w < n; ++w) iperm[size_t(order[size_t(w)])] = w;
            for (int w = 0; w < n; ++w) {
                const int i = order[size_t(w)];
                for (int p = S.rp[size_t(i)]; p < S.diag[size_t(i)]; ++p) { EL.col.push_back(iperm[size_t(S.ci[size_t(p)])]); EL.src.push_back(p); }
                for (int p = S.diag[size_t(i)] + 1; p < S.rp[size_t(i) + 1]; ++p) { EU.col.push_back(iperm[size_t(S.ci[size_t(p)])]); EU.src.push_back(p); }
                EL.ptr.push_back(int(EL.col.size()));
                EU.ptr.push_back(int(EU.col.size()));
                EU.dsrc.push_back(S.diag[size_t(i)]);
            }
            EL.orig = order;
            EU.orig = order;
        };
        tiled = false;
        if (!mc) {
            levelOrder(S, ord, lstart);
            rp = drp; ci = dci; diag = ddiag;
            if (!rowSys && schedule != 1 && chooseTiles(S, int(lstart.size()) - 1)) {
                naturalEntries(tiles.ord);
                tiled = tileSlices(EL, EU);
            }
            if (!tiled) naturalEntries(ord);
            sweepOrd = tiled ? tiles.ord : ord;
        } else {
            rbt = false;
            if (rbDims > 0 && !rowSys) {
                const int dims[3] = {rbDims & 255, (rbDims >> 8) & 255, (rbDims >> 16) & 255};
                rbt = rbTileOrder(S, dims, rb);
            }
            if (rbt) {
                ord = rb.ord;
                lstart.assign({0, rb.nRed, rb.nTile});
                for (size_t k = 1; k < rb.farStart.size(); ++k) lstart.push_back(rb.nTile + rb.farStart[k]);
                colours = int(lstart.size()) - 1;
            } else {
                colours = colourGraph(S, colour, ord, lstart);
            }
            std::vector<int> bmap;
            permutePattern(S, ord, fpat, bmap);
            frp.upload(fpat.rp, s);
            fci.upload(fpat.ci, s);
            fdiag.upload(fpat.diag, s);
            fmap.upload(bmap, s);
            rp = frp.p; ci = fci.p; diag = fdiag.p;
            for (int w = 0; w < n; ++w) {
                for (int p = fpat.rp[size_t(w)]; p < fpat.diag[size_t(w)]; ++p) { EL.col.push_back(fpat.ci[size_t(p)]); EL.src.push_back(p); }
                for (int p = fpat.diag[size_t(w)] + 1; p < fpat.rp[size_t(w) + 1]; ++p) { EU.col.push_back(fpat.ci[size_t(p)]); EU.src.push_back(p); }
                EL.ptr.push_back(int(EL.col.size()));
                EU.ptr.push_back(int(EU.col.size()));
                EU.dsrc.push_back(fpat.diag[size_t(w)]);
            }
        }
        if (mc) {
            EL.orig = ord;
            EU.orig = ord;
            sweepOrd = ord;
        }
        perm.upload(sweepOrd, s);
        {
            std::vector<int> ip(static_cast<size_t>(n));
            for (int w = 0; w < n; ++w) ip[size_t(sweepOrd[size_t(w)])] = w;
            iperm.upload(ip, s);
        }
        fuseBwd = false;
        if (!tiled) {
            // multicolour order: rows of a large level (beyond any fused run) with no upper entries that no later row
            // reads finish their backward step in the forward launch (z = D^-1 y, the pivot block carried by the lower
            // slices), and are left out of the upper slices
            std::vector<char> skipU;
            const int capL = entriesCap(EL);
            if (mc && !rowSys && !rbt) {
                std::vector<char> read(static_cast<size_t>(n), 0);
                for (int c : EL.col) if (c >= 0 && c < n) read[size_t(c)] = 1;
                skipU.assign(size_t(n), 0);
                EL.dsrc.assign(size_t(n), -1);
                for (int l = 1; l + 1 < int(lstart.size()); ++l) {
                    if (lstart[size_t(l) + 1] - lstart[size_t(l)] <= kRunRowsMax) continue;
                    for (int w = lstart[size_t(l)]; w < lstart[size_t(l) + 1]; ++w) {
                        if (EU.ptr[size_t(w) + 1] != EU.ptr[size_t(w)] || read[size_t(w)] || EL.ptr[size_t(w) + 1] - EL.ptr[size_t(w)] > capL) continue;
                        skipU[size_t(w)] = 1;
                        EL.dsrc[size_t(w)] = EU.dsrc[size_t(w)];
                        fuseBwd = true;
                    }
                }
                if (!fuseBwd) {
                    EL.dsrc.clear();
                    skipU.clear();
                }
            }
            if (rbt) {
                // slices per (tile, local level) unit, then per level of the rows behind the tiles; a tile row is never a
                // long row (the caps cover every tile row); the sweep levels' slice ranges are the units' ranges joined
                std::vector<int> bounds = rb.unitStart;
                for (size_t k = 1; k < rb.farStart.size(); ++k) bounds.push_back(rb.nTile + rb.farStart[k]);
                auto tileCap = [&](const RowEntries& E) {
                    int c = entriesCap(E);
                    for (int w = 0; w < rb.nTile; ++w) c = std::max(c, E.ptr[size_t(w) + 1] - E.ptr[size_t(w)]);
                    return c;
                };
                LevelSched uL, uU;
                buildSlices(EL, bounds, tileCap(EL), hL, uL);
                buildSlices(EU, bounds, tileCap(EU), hU, uU);
                const int nu = int(rb.unitStart.size()) - 1, redU = rb.tileUnit[size_t(rb.redTiles)];
                for (int u = 0; u < nu; ++u)
                    if (uL.qStart[size_t(u)] != uL.qStart[size_t(u) + 1] || uU.qStart[size_t(u)] != uU.qStart[size_t(u) + 1])
                        throw Error(RSH_ERR_DEVICE, "red / black tiles: a tile row became a long row");
                auto join = [&](const LevelSched& u, LevelSched& out) {
                    out.sStart = {u.sStart[0], u.sStart[size_t(redU)], u.sStart[size_t(nu)]};
                    out.qStart = {u.qStart[0], u.qStart[size_t(redU)], u.qStart[size_t(nu)]};
                    for (size_t k = size_t(nu) + 1; k < u.sStart.size(); ++k) {
                        out.sStart.push_back(u.sStart[k]);
                        out.qStart.push_back(u.qStart[k]);
                    }
                };
                join(uL, sL);
                join(uU, sU);
                rbTiles = int(rb.tileUnit.size()) - 1;
                rbRedTiles = rb.redTiles;
                std::vector<int4> tl, tu;
                rbMaxRows = 0;
                for (int t = 0; t < rbTiles; ++t) {
                    const int u0 = rb.tileUnit[size_t(t)], u1 = rb.tileUnit[size_t(t) + 1];
                    const int r0 = rb.unitStart[size_t(u0)], r1 = rb.unitStart[size_t(u1)];
                    rbMaxRows = std::max(rbMaxRows, r1 - r0);
                    tl.push_back(make_int4(uL.sStart[size_t(u0)], uL.sStart[size_t(u1)], r0, r1));
                    tu.push_back(make_int4(uU.sStart[size_t(u0)], uU.sStart[size_t(u1)], r0, r1));
                }
                for (int sl = 0; sl < uL.sStart[size_t(nu)]; ++sl)
                    if (hL.width[size_t(sl)] > kTileWaveCap) throw Error(RSH_ERR_DEVICE, "red / black tiles: a tile row has too many slots");
                for (int sl = 0; sl < uU.sStart[size_t(nu)]; ++sl)
                    if (hU.width[size_t(sl)] > kTileWaveCap) throw Error(RSH_ERR_DEVICE, "red / black tiles: a tile row has too many slots");
                tilesL.upload(tl, s);
                tilesU.upload(tu, s);
            } else {
                buildSlices(EL, lstart, capL, hL, sL);
                buildSlices(EU, lstart, entriesCap(EU), hU, sU, nullptr, skipU.empty() ? nullptr : &skipU);
            }
        }
        if (rowSys) {
            sliceOwners(hL, *rowSys);
            sliceOwners(hU, *rowSys);
        }
        uploadSlices(hL, L64, nb, s);
        uploadSlices(hU, U64, nb, s);
        {
            const int nl = int(lstart.size()) - 1;
            std::vector<int> wL(static_cast<size_t>(nl)), wU(static_cast<size_t>(nl));
            for (int l = 0; l < nl && !tiled; ++l) {
                wL[size_t(l)] = (sL.sStart[size_t(l) + 1] - sL.sStart[size_t(l)]) + (sL.qStart[size_t(l) + 1] - sL.qStart[size_t(l)]);
                wU[size_t(l)] = (sU.sStart[size_t(l) + 1] - sU.sStart[size_t(l)]) + (sU.qStart[size_t(l) + 1] - sU.qStart[size_t(l)]);
            }
            std::vector<int> rowsL(static_cast<size_t>(nl));
            for (int l = 0; l < nl; ++l) rowsL[size_t(l)] = lstart[size_t(l) + 1] - lstart[size_t(l)];
            runsL.clear();
            runsU.clear();
            if (!tiled && rbt) {
                // the two tile levels launch alone; the rows behind the tiles in runs as the multicolour order's
                std::vector<int> wL2(wL.begin() + 2, wL.end()), wU2(wU.begin() + 2, wU.end()), rows2(rowsL.begin() + 2, rowsL.end());
                for (auto* pr : {&runsL, &runsU}) {
                    const std::vector<LevelRun> rest = levelRuns(pr == &runsL ? wL2 : wU2, fuseMax, &rows2, runRowsMax);
                    pr->assign({{0, 1, false}, {1, 2, false}});
                    for (const LevelRun& r2 : rest) pr->push_back({r2.l0 + 2, r2.l1 + 2, r2.fused});
                }
            } else if (!tiled) {
                runsL = levelRuns(wL, fuseMax, &rowsL, runRowsMax);
                runsU = levelRuns(wU, fuseMax, &rowsL, runRowsMax);
            }
            denseL.build(runsL, EL, lstart, nb, false, s, rowSys);
            denseU.build(runsU, EU, lstart, nb, true, s, rowSys);
            buildFactorGroups(s);
            if (keepEntries) {
                keptL = EL;
                keptU = EU;
                workL = wL;
                workU = wU;
                rowsPerLevel = rowsL;
            }
            dLs.upload(sL.sStart, s);
            dLq.upload(sL.qStart, s);
            dUs.upload(sU.sStart, s);
            dUq.upload(sU.qStart, s);
            if (tiled) {
                tUnit.upload(tiles.unitStart, s);
                tLvPtr.upload(tiles.lvPtr, s);
                tLv.upload(tiles.lvStart, s);
            }
            updateFirstCopy();
            // factor rows (F numbering) per level, long rows apart
            const HostBsr& F = mc ? fpat : S;
            std::vector<int> sh, lo;
            fShortStart.assign(1, 0);
            fLongStart.assign(1, 0);
            // factorisation levels: the sweep levels, except that a red / black tile level is factorised by local level
            std::vector<std::vector<int>> flev;
            fLevel0.assign(1, 0);
            for (int l = 0; l < nl; ++l) {
                if (rbt && l < 2) {
                    const int L0 = l == 0 ? 0 : rb.redLevels, L1 = l == 0 ? rb.redLevels : rb.redLevels + rb.blackLevels;
                    const size_t base = flev.size();
                    flev.resize(base + size_t(L1 - L0));
                    for (size_t u = 0; u + 1 < rb.unitStart.size(); ++u) {
                        const int fl = rb.unitLevel[u];
                        if (fl < L0 || fl >= L1) continue;
                        for (int w = rb.unitStart[u]; w < rb.unitStart[u + 1]; ++w) flev[base + size_t(fl - L0)].push_back(w);
                    }
                } else {
                    std::vector<int> rows;
                    for (int w = lstart[size_t(l)]; w < lstart[size_t(l) + 1]; ++w) rows.push_back(w);
                    flev.push_back(std::move(rows));
                }
                fLevel0.push_back(int(flev.size()));
            }
            for (const std::vector<int>& fr : flev) {
                for (int w : fr) {
                    const int i = mc ? w : ord[size_t(w)];
                    const int lower = F.diag[size_t(i)] - F.rp[size_t(i)];
                    // natural order: a row whose pivot rows carry long upper parts (the neighbours of a well's perforated
                    // cells walk every perforation's U entries) is eliminated like a long row, through its update map
                    int walk = 0;
                    if (!mc)
                        for (int p = F.rp[size_t(i)]; p < F.diag[size_t(i)]; ++p) {
                            const int k = F.ci[size_t(p)];
                            walk += F.rp[size_t(k) + 1] - F.diag[size_t(k)] - 1;
                        }
                    (lower > 24 || walk > 48 ? lo : sh).push_back(i);
                }
                fShortStart.push_back(int(sh.size()));
                fLongStart.push_back(int(lo.size()));
            }
            fShort.upload(sh, s);
            fLong.upload(lo, s);
            // natural order: the long rows' update maps (per lower entry, the pivot row's U blocks that meet the row and
            // the entry each updates), so that their elimination runs from LDS (kBilu0LongMapped)
            longMap = LongRowMap{nullptr, nullptr, nullptr, 0};
            if (!mc && !lo.empty()) {
                std::vector<int> mrow, mptr(1, 0);
                std::vector<int2> mp;
                int maxLen = 0;
                for (size_t t = 0; t < lo.size(); ++t) {
                    const int i = lo[t];
                    const int pb = F.rp[size_t(i)], pe = F.rp[size_t(i) + 1], pd = F.diag[size_t(i)];
                    mrow.push_back(int(mptr.size()) - 1);
                    maxLen = std::max(maxLen, pe - pb);
                    for (int p = pb; p < pd; ++p) {
                        const int k = F.ci[size_t(p)];
                        for (int q = F.diag[size_t(k)] + 1; q < F.rp[size_t(k) + 1]; ++q) {
                            const int* b = F.ci.data() + p + 1;
                            const int* e = F.ci.data() + pe;
                            const int* f = std::lower_bound(b, e, F.ci[size_t(q)]);
                            if (f != e && *f == F.ci[size_t(q)]) mp.push_back(make_int2(q, int(f - F.ci.data()) - pb));
                        }
                        mptr.push_back(int(mp.size()));
                    }
                }
                if (size_t(maxLen) * size_t(nb) * size_t(nb) * sizeof(double) <= size_t(kLongRowLdsMax)) {
                    lmRow.upload(mrow, s);
                    lmPtr.upload(mptr, s);
                    lmMap.upload(mp, s);
                    longMap = LongRowMap{lmRow.p, lmPtr.p, lmMap.p, maxLen};
                }
            }
        }
        lu.alloc(size_t(S.rp[size_t(n)]) * nb * nb);
        y64.alloc(size_t(n + 1) * nb);
        y64.zero(s);
        rt64.alloc(size_t(kRunRowsMax) * nb);
        tail64.alloc(size_t(kRunRowsMax) * nb);
        nreg.alloc(1);
    }
    static int entriesCap(const RowEntries& E) {
        const int m = int(E.ptr.size()) - 1;
        std::vector<int> len(static_cast<size_t>(m));
        for (int w = 0; w < m; ++w) len[size_t(w)] = E.ptr[size_t(w) + 1] - E.ptr[size_t(w)];
        return rowCap(len);
    }
    // Tile extents for the natural-order sweeps from a cost model of one sweep: per tile level a launch and the
    // unit's loads, then one barrier step per local level (tileLaunchUs + tileStepUs * depth + tileRowUs * rows), against
    // levelLaunchUs per level launch (constants measured on MI355X with in-kernel clock stamps at 1.1M cells: dispatch
    // gap + prologue + write-back about 8 us per launch, phase A about 5 ns per row, 0.35-0.55 us per local level).
    // Tiles are used when the model makes them at least a quarter cheaper (schedule 0) or whenever the pattern admits
    // them (schedule 2).
    static constexpr double tileLaunchUs = 8.0, tileStepUs = 0.45, tileRowUs = 0.005, levelLaunchUs = 4.85;
    bool chooseTiles(const HostBsr& S, int nlevels) {
        int off[3] = {0, 0, 0};
        std::vector<int> coord[3];
        const int nd = gridCoords(S, off, coord);
        if (nd == 0) return false;
        int ext[3] = {1, 1, 1};
        for (int a = 0; a < nd; ++a)
            for (int i = 0; i < S.n; ++i) ext[a] = std::max(ext[a], coord[a][size_t(i)] + 1);
        int best[3] = {1, 1, 1};
        double bestCost = 1e300;
        if (tileDims > 0) {
            best[0] = tileDims & 255;
            best[1] = (tileDims >> 8) & 255;
            best[2] = (tileDims >> 16) & 255;
            bestCost = 0;
        } else {
            static const int cand[] = {1, 2, 3, 4, 5, 6, 8, 10, 12, 15, 16, 20, 24, 30, 32, 40, 48, 64};
            for (int b0 : cand)
                for (int b1 : cand)
                    for (int b2 : cand) {
                        const int b[3] = {b0, b1, b2};
                        bool ok = true;
                        long long rows = 1;
                        int ntl = 1, depth = 1;
                        for (int a = 0; a < 3; ++a) {
                            if (a >= nd) { ok = ok && b[a] == 1; continue; }
                            const int e = std::min(b[a], ext[a]);
                            rows *= e;
                            ntl += (ext[a] + b[a] - 1) / b[a] - 1;
                            depth += e - 1;
                        }
                        if (!ok || rows > (kTileWaves - 1) * kWave) continue;
                        const double cost = ntl * (tileLaunchUs + tileStepUs * depth + tileRowUs * double(rows));
                        if (cost < bestCost) {
                            bestCost = cost;
                            best[0] = b0; best[1] = b1; best[2] = b2;
                        }
                    }
            if (schedule != 2 && bestCost > 0.75 * levelLaunchUs * nlevels) return false;
        }
        tileOrder(S, coord, nd, best, tiles);
        return tiles.tileLevels() > 0;
    }
    // slices of the tiled sweeps (the units as the slice layout's levels) and the launch limits they must meet; false
    // (level launches instead) when a unit needs more wavefronts or LDS than a workgroup has
    bool tileSlices(const RowEntries& EL, const RowEntries& EU) {
        buildSlices(EL, tiles.unitStart, entriesCap(EL), hL, sL);
        buildSlices(EU, tiles.unitStart, entriesCap(EU), hU, sU);
        tLongMax = 0;
        tLongCap = 0;
        for (int u = 0; u < tiles.units(); ++u) {
            const int w0 = tiles.unitStart[size_t(u)], w1 = tiles.unitStart[size_t(u) + 1];
            for (int side = 0; side < 2; ++side) {
                const SliceHost& h = side ? hU : hL;
                const LevelSched& sc = side ? sU : sL;
                const int ns = sc.sStart[size_t(u) + 1] - sc.sStart[size_t(u)], nq = sc.qStart[size_t(u) + 1] - sc.qStart[size_t(u)];
                if (ns + (nq > 0 ? 1 : 0) > kTileWaves) return false;
                tLongMax = std::max(tLongMax, nq);
                for (int q = sc.qStart[size_t(u)]; q < sc.qStart[size_t(u) + 1]; ++q) {
                    int in = 0;
                    for (int p = h.lptr[size_t(q)]; p < h.lptr[size_t(q) + 1]; ++p) in += h.lcol[size_t(p)] >= w0 && h.lcol[size_t(p)] < w1;
                    tLongCap = std::max(tLongCap, in);
                }
            }
        }
        tLongCap = std::max(tLongCap, 1);
        if (tLongMax > kWave) return false;   // the long-row wavefront gives each long row of a unit at least one lane
        return tileLdsBytes(nb, kIluFp64, tiles.maxRows, tiles.maxLevels, tLongMax, tLongCap) <= kTileLdsMax;
    }
    void ensure32(hipStream_t s) {
        if (have32) return;
        uploadSlices(hL, L32, nb, s);
        uploadSlices(hU, U32, nb, s);
        y32.alloc(size_t(n + 1) * nb);
        y32.zero(s);
        rt32.alloc(size_t(kRunRowsMax) * nb);
        tail32.alloc(size_t(kRunRowsMax) * nb);
        have32 = true;
    }
    // new level chunking on the built schedule (needs keepEntries at build); the dense groups are refactored from
    // the current factors
    // the member-pair block tables of the dense groups of the forward runs (multicolour order, where working rows
    // are the factorisation rows and every row before a run precedes all of its rows)
    void buildFactorGroups(hipStream_t s) {
        fgRun0.assign(runsL.size(), -1);
        fgRun1.assign(runsL.size(), -1);
        if (!mc) return;
        std::vector<int> st(1, 0), rw, idx;
        std::vector<long long> off;
        std::vector<int> loc(static_cast<size_t>(n), -1);
        int ng = 0;
        for (size_t ri = 0; ri < runsL.size(); ++ri) {
            if (denseL.runGroup0[ri] < 0) continue;
            fgRun0[ri] = ng;
            for (int g = denseL.runGroup0[ri]; g < denseL.runGroup1[ri]; ++g) {
                const int m0 = denseL.hStart[size_t(g)], K = denseL.hStart[size_t(g) + 1] - m0;
                for (int a = 0; a < K; ++a) loc[size_t(denseL.hRows[size_t(m0 + a)])] = a;
                const size_t base = idx.size();
                off.push_back((long long)base);
                idx.resize(base + size_t(K) * K, -1);
                for (int a = 0; a < K; ++a) {
                    const int i = denseL.hRows[size_t(m0 + a)];
                    for (int p = fpat.rp[size_t(i)]; p < fpat.rp[size_t(i) + 1]; ++p) {
                        const int b = loc[size_t(fpat.ci[size_t(p)])];
                        if (b >= 0) idx[base + size_t(a) * K + size_t(b)] = p;
                    }
                }
                for (int a = 0; a < K; ++a) {
                    rw.push_back(denseL.hRows[size_t(m0 + a)]);
                    loc[size_t(denseL.hRows[size_t(m0 + a)])] = -1;
                }
                st.push_back(int(rw.size()));
                ++ng;
            }
            fgRun1[ri] = ng;
        }
        fgStart.upload(st, s);
        fgRows.upload(rw, s);
        fgIdx.upload(idx, s);
        fgIdxOff.upload(off, s);
    }
    // the first level's forward sweep is a copy when it runs alone and none of its rows has lower entries
    void updateFirstCopy() {
        firstCopy = lstart.size() > 1 && !runsL.empty() && !runsL[0].fused && runsL[0].l1 == 1 && sL.qStart[1] == sL.qStart[0];
        for (int sl = sL.sStart[0]; firstCopy && sl < sL.sStart[1]; ++sl) firstCopy = hL.width[size_t(sl)] == 0;
        tailCopy = false;
        if (!runsL.empty() && !runsU.empty()) {
            const size_t a = runsL.size() - 1, b = runsU.size() - 1;
            tailCopy = runsL[a].fused && runsU[b].fused && runsL[a].l0 == runsU[b].l0 && runsL[a].l1 == runsU[b].l1 && denseL.runTask0[a] >= 0 &&
                       denseU.runTask0[b] >= 0 && denseU.noOutside[b];
        }
    }
    // newRunRows > 0: also the rows a fused run may hold (its working vector's LDS tile, at most kRunRowsMax)
    void retune(int newFuseMax, hipStream_t s, int newRunRows = 0) {
        if (!keepEntries || rbt) throw Error(RSH_ERR_ARG, "level chunking can only be retuned on a batch member's schedule");
        fuseMax = newFuseMax;
        if (newRunRows > 0) runRowsMax = std::min(newRunRows, kRunRowsMax);
        runsL = levelRuns(workL, fuseMax, &rowsPerLevel, runRowsMax);
        runsU = levelRuns(workU, fuseMax, &rowsPerLevel, runRowsMax);
        denseL.build(runsL, keptL, lstart, nb, false, s, rowSys);
        denseU.build(runsU, keptU, lstart, nb, true, s, rowSys);
        buildFactorGroups(s);
        updateFirstCopy();
        denseL.factor(nb, false, lu.p, have32, s);
        denseU.factor(nb, true, lu.p, have32, s);
        generation = nextGeneration();
    }
    // numeric factorisation from the system's values (natural layout), then the sweep layouts
    // lean: the C45 lean factorisation (input and every finished row rounded to float32; the factors are then
    // float32 values, applied from the FP32 slices)
    bool lean = false;
    void factor(const double* val, bool f32, hipStream_t s, bool leanFactors = false) {
        const size_t nvals = lu.n;
        lean = leanFactors;
        if (lean) f32 = true;
        if (mc) launchGatherBlocks(s, nb, val, fmap.p, int(nvals / (size_t(nb) * nb)), lu.p);
        else RSH_HIP(hipMemcpyAsync(lu.p, val, nvals * sizeof(double), hipMemcpyDeviceToDevice, s));
        if (lean) launchRoundToFloat(s, (long long)nvals, lu.p);
        nreg.zero(s);
        if (!mc) {
            for (int l = 0; l + 1 < int(fShortStart.size()); ++l) {
                launchBilu0Level(s, nb, fShort.p, fShortStart[size_t(l)], fShortStart[size_t(l) + 1], rp, ci, diag, lu.p, nreg.p, -1, lean);
                launchBilu0LongRows(s, nb, fLong.p, fLongStart[size_t(l)], fLongStart[size_t(l) + 1], rp, ci, diag, lu.p, nreg.p, lean,
                                    longMap.maxLen > 0 ? &longMap : nullptr);
            }
        }
        for (size_t ri = 0; mc && ri < runsL.size(); ++ri) {
            const LevelRun& run = runsL[ri];
            if (ri < fgRun0.size() && fgRun0[ri] >= 0) {
                // the run's rows: the contributions of the rows before the run, then the groups right-looking
                const int w0 = lstart[size_t(run.l0)], w1 = lstart[size_t(run.l1)];
                launchBilu0Level(s, nb, nullptr, w0, w1, rp, ci, diag, lu.p, nreg.p, w0, lean);
                launchIluGroupFactor(s, nb, fgRun0[ri], fgRun1[ri], fgStart.p, fgRows.p, fgIdxOff.p, fgIdx.p, diag, lu.p, nreg.p, rp, lean);
                continue;
            }
            for (int l = run.l0; l < run.l1; ++l)
                for (int fl = fLevel0[size_t(l)]; fl < fLevel0[size_t(l) + 1]; ++fl) {
                    launchBilu0Level(s, nb, fShort.p, fShortStart[size_t(fl)], fShortStart[size_t(fl) + 1], rp, ci, diag, lu.p, nreg.p, -1, lean);
                    launchBilu0LongRows(s, nb, fLong.p, fLongStart[size_t(fl)], fLongStart[size_t(fl) + 1], rp, ci, diag, lu.p, nreg.p, lean);
                }
        }
        packSlices(s, nb, lu.p, L64, hL.nslots, false);
        packSlices(s, nb, lu.p, U64, hU.nslots, false);
        if (f32) {
            ensure32(s);
            packSlices(s, nb, lu.p, L32, hL.nslots, true);
            packSlices(s, nb, lu.p, U32, hU.nslots, true);
        }
        denseL.factor(nb, false, lu.p, f32, s);
        denseU.factor(nb, true, lu.p, f32, s);
        RSH_HIP(hipMemcpyAsync(&regularised, nreg.p, sizeof(int), hipMemcpyDeviceToHost, s));
    }
    // fp: IluFp (kIluFp32 and kIluStore32 need the FP32 factors of a factor(..., true))
    // firstDone: y of the first level (firstCopy) is already r in the working numbering
    // inv64: with FP32-stored factors (kIluStore32), the dense group inverses in FP64 (lean factors in FP64 precision)
    void apply(const double* r, double* z, const double* xp, int fp, hipStream_t s, bool firstDone = false, bool inv64 = false) const {
        const bool f32 = fp != kIluFp64;          // FP32-stored factors
        const bool v32 = fp == kIluFp32;          // FP32 working vectors
        const void* Lp = f32 ? static_cast<const void*>(&L32) : static_cast<const void*>(&L64);
        const void* Up = f32 ? static_cast<const void*>(&U32) : static_cast<const void*>(&U64);
        void* y = v32 ? static_cast<void*>(y32.p) : static_cast<void*>(y64.p);
        void* rt = v32 ? static_cast<void*>(rt32.p) : static_cast<void*>(rt64.p);
        void* tail = v32 ? static_cast<void*>(tail32.p) : static_cast<void*>(tail64.p);
        auto sweepRun = [&](const LevelRun& run, size_t ri, bool fwd) {
            const DenseSide& D = fwd ? denseL : denseU;
            const LevelSched& sc = fwd ? sL : sU;
            const void* sl = fwd ? Lp : Up;
            const int wb = lstart[size_t(run.l0)], we = lstart[size_t(run.l1)];
            if (D.runTask0[ri] >= 0) {
                const int t0 = sc.sStart[size_t(run.l0)], u0 = sc.qStart[size_t(run.l0)];
                const bool i64 = !f32 || (inv64 && fp == kIluStore32);
                const void* inv = i64 ? static_cast<const void*>(D.inv.p) : static_cast<const void*>(D.inv32.p);
                const bool tailStep = tailCopy && ri + 1 == (fwd ? runsL.size() : runsU.size());
                const bool skipOutside = tailStep && !fwd;
                launchRunDense(s, nb, fp, sl, fwd, t0, skipOutside ? 0 : sc.sStart[size_t(run.l1)] - t0, u0,
                               skipOutside ? 0 : sc.qStart[size_t(run.l1)] - u0, wb, we, D.runTask1[ri] - D.runTask0[ri],
                               D.tasks.p + D.runTask0[ri], D.start.p, D.rows.p, D.off.p, inv, perm.p, r, skipOutside ? tail : rt, y, z, xp,
                               D.gowner.p, gate, tailStep && fwd ? tail : nullptr, i64);
                return;
            }
            launchIluFused(s, nb, fp, sl, fwd, run.l0, run.l1, fwd ? dLs.p : dUs.p, fwd ? dLq.p : dUq.p, wb, we - wb, perm.p, r, y, z, xp);
        };
        if (tiled) {
            const int ntl = tiles.tileLevels();
            for (int t = 0; t < ntl; ++t)
                launchIluTile(s, nb, fp, Lp, true, tiles.tlStart[size_t(t)], tiles.tlStart[size_t(t) + 1] - tiles.tlStart[size_t(t)], tUnit.p, dLs.p,
                              dLq.p, tLvPtr.p, tLv.p, tiles.maxRows, tiles.maxLevels, tLongMax, tLongCap, perm.p, r, y, z, xp);
            for (int t = ntl; t-- > 0;)
                launchIluTile(s, nb, fp, Up, false, tiles.tlStart[size_t(t)], tiles.tlStart[size_t(t) + 1] - tiles.tlStart[size_t(t)], tUnit.p, dUs.p,
                              dUq.p, tLvPtr.p, tLv.p, tiles.maxRows, tiles.maxLevels, tLongMax, tLongCap, perm.p, r, y, z, xp);
            return;
        }
        for (size_t ri = 0; ri < runsL.size(); ++ri) {
            const LevelRun& run = runsL[ri];
            if (run.fused) { sweepRun(run, ri, true); continue; }
            const int l = run.l0;
            if (l == 0 && firstDone) continue;   // the caller wrote y of the first level
            if (rbt && l < 2) {
                const int t0 = l == 0 ? 0 : rbRedTiles, t1 = l == 0 ? rbRedTiles : rbTiles;
                launchIluTileWave(s, nb, fp, Lp, true, t1 - t0, tilesL.p + t0, r, y, nullptr, nullptr);
                continue;
            }
            launchIluForward(s, nb, fp, Lp, sL.sStart[size_t(l)], sL.sStart[size_t(l) + 1], sL.qStart[size_t(l)], sL.qStart[size_t(l) + 1], r, y,
                             gate, fuseBwd ? z : nullptr, xp);
        }
        for (size_t k = runsU.size(); k-- > 0;) {
            const LevelRun& run = runsU[k];
            if (run.fused) { sweepRun(run, k, false); continue; }
            const int l = run.l0;
            if (rbt && l < 2) {
                const int t0 = l == 0 ? 0 : rbRedTiles, t1 = l == 0 ? rbRedTiles : rbTiles;
                launchIluTileWave(s, nb, fp, Up, false, t1 - t0, tilesU.p + t0, nullptr, y, z, xp);
                continue;
            }
            launchIluBackward(s, nb, fp, Up, sU.sStart[size_t(l)], sU.sStart[size_t(l) + 1], sU.qStart[size_t(l)], sU.qStart[size_t(l) + 1], y, z,
                              xp, gate, k != 0);
        }
    }
    size_t bytes() const {
        return lu.bytes() + L64.bytes() + U64.bytes() + L32.bytes() + U32.bytes() + y64.bytes() + y32.bytes() + denseL.bytes() + denseU.bytes() +
               lmRow.bytes() + lmPtr.bytes() + lmMap.bytes() + tUnit.bytes() + tLvPtr.bytes() + tLv.bytes();
    }
};

// ------------------------------------------------------------------ AMG hierarchy on the device
struct AmgLevelDev {
    int n = 0, nnz = 0, nc = 0;
    const int *rp = nullptr, *ci = nullptr, *diag = nullptr;   // level 0: the system's pattern
    DBuf<int> rpB, ciB, diagB;
    DBuf<double> v;
    DBuf<int> agg, rrp, rci;
    DBuf<int> apPtr, apSrc, cPtr, cSrc;
    DBuf<double> ap;
    DBuf<int> gsF, gsB;                  // rows in forward / backward level order
    std::vector<int> gsFs, gsBs;         // level starts
    DBuf<int> dgsFs, dgsBs;
    std::vector<LevelRun> runsF, runsB;
    DBuf<double> x1, x2, b, r;
    DBuf<double> x3;                     // further Jacobi steps (ping-pong)
    DBuf<double> dinv;                   // 1 / a_ii (Jacobi)
    DBuf<float> v32, dinv32;             // FP32 copies of the level's values and dinv (mixed-precision hierarchy)
    const double* result = nullptr;
    // the level's values in a row-slice layout for the streamed Jacobi steps: level 0 uses the SpMV layout of S (same
    // pattern); coarse levels (nullptr) stay CSR with G lanes per row, which keeps more loads in flight on a few
    // hundred thousand rows than one row per lane does
    const SliceDev<double>* plan = nullptr;
    SliceHost planHost;                  // a coarse level's own row-slice layout (when it has one)
    std::unique_ptr<SliceDev<double>> ownPlan;
    DBuf<double> pS, pL;
    DBuf<float> pS32, pL32;
    // repack the slice values from the CSR values v (after every numeric set-up)
    void packPlan(hipStream_t s, bool f32) {
        if (!plan) return;
        pS.alloc(plan->src.n);
        pL.alloc(plan->lsrc.n);
        launchPackSlices(s, 1, v.p, plan->src.p, (long long)(pS.n / kWave), pS.p, false);
        if (plan->lsrc.n) launchPackLong(s, 1, v.p, plan->lsrc.p, int(plan->lsrc.n), pL.p, false);
        if (f32) {
            pS32.alloc(pS.n);
            pL32.alloc(pL.n);
            launchToFloat(s, (long long)pS.n, pS.p, pS32.p);
            if (pL.n) launchToFloat(s, (long long)pL.n, pL.p, pL32.p);
        }
    }
};

struct CprDev {
    static constexpr double kOmegaDefault = 0.6667;   // damped-Jacobi weight of the reference's device V-cycle (bsr.h:173)
    double omega = kOmegaDefault;
    int n = 0, nb = 1;
    DBuf<double> w, xp, t;
    std::vector<std::unique_ptr<AmgLevelDev>> lv;
    std::vector<int> hostAggSig;          // sizes per level, for reporting
    int m = 0;                            // coarsest size
    bool dense = false;
    DBuf<double> M0, M1;                  // [C | I] ping-pong; the inverse ends in Minv
    double* Minv = nullptr;
    int smoother = RSH_SMOOTH_GS;
    int steps = 1;                        // Jacobi pre- and post-smoothing steps
    int coarseSteps = 0;                  // ... on the levels below the finest (0: steps)
    bool built = false;
    // the SpMV row-slice layout of S (same pattern as level 0), for level 0's streamed Jacobi steps
    const SliceDev<double>* plan = nullptr;
    bool have32 = false;
    std::vector<int> patRp, patCi;        // pattern of the hierarchy (reuse check)
    // coarse levels of at least this many rows are smoothed in row slices (one row per lane, which reads each
    // slot's values and columns as whole cache lines; smaller levels have too few wavefronts for it).  The sums are
    // the same either way; RSH_AMG_SLICE_ROWS sets another threshold (tests compare the two layouts)
    int sliceLevelRows = [] { const char* e = std::getenv("RSH_AMG_SLICE_ROWS"); return e ? std::max(1, std::atoi(e)) : 100000; }();

    int generation = 0;
    void buildHierarchy(const HostBsr& S, const int* drp, const int* dci, const int* ddiag, const double* apHost, hipStream_t s) {
        generation = nextGeneration();
        HostCsr A0;
        A0.n = S.n;
        A0.rp = S.rp;
        A0.ci = S.ci;
        A0.v.assign(apHost, apHost + S.ci.size());
        std::vector<AmgLevelHost> H;
        amgSetup(std::move(A0), H);
        lv.clear();
        for (size_t l = 0; l < H.size(); ++l) {
            auto L = std::make_unique<AmgLevelDev>();
            AmgLevelHost& h = H[l];
            L->n = h.A.n;
            L->nnz = int(h.A.ci.size());
            L->nc = h.nc;
            if (l == 0) {
                L->rp = drp; L->ci = dci; L->diag = ddiag;
            } else {
                L->rpB.upload(h.A.rp, s);
                L->ciB.upload(h.A.ci, s);
                L->diagB.upload(h.A.diag, s);
                L->rp = L->rpB.p; L->ci = L->ciB.p; L->diag = L->diagB.p;
            }
            L->v.upload(h.A.v, s);
            const int lanes = groupLanes(double(L->nnz) / std::max(L->n, 1));
            if (l > 0 && !h.agg.empty() && h.A.n >= sliceLevelRows && lanes >= 8 && lanes <= 32) {
                // same sums as the G-lane CSR step (bit-identical), read through row slices
                rowSliceLayout(h.A.n, h.A.rp, h.A.ci, L->planHost, true);
                L->ownPlan = std::make_unique<SliceDev<double>>();
                uploadSlices(L->planHost, *L->ownPlan, 1, s);
                L->ownPlan->chains = lanes;
                L->plan = L->ownPlan.get();
            }
            if (!h.agg.empty()) {
                L->agg.upload(h.agg, s);
                std::vector<int> rrp(size_t(h.nc) + 1, 0), rci(h.agg.size());
                for (int a : h.agg) rrp[size_t(a) + 1]++;
                for (int a = 0; a < h.nc; ++a) rrp[size_t(a) + 1] += rrp[size_t(a)];
                std::vector<int> fill(rrp.begin(), rrp.end() - 1);
                for (size_t i = 0; i < h.agg.size(); ++i) rci[size_t(fill[size_t(h.agg[i])]++)] = int(i);
                L->rrp.upload(rrp, s);
                L->rci.upload(rci, s);
                L->apPtr.upload(h.apPtr, s);
                L->apSrc.upload(h.apSrc, s);
                L->cPtr.upload(h.cPtr, s);
                L->cSrc.upload(h.cSrc, s);
                L->ap.alloc(h.apPtr.size() - 1);
            }
            std::vector<int> ord;
            csrLevels(h.A, true, ord, L->gsFs);
            L->gsF.upload(ord, s);
            csrLevels(h.A, false, ord, L->gsBs);
            L->gsB.upload(ord, s);
            L->dgsFs.upload(L->gsFs, s);
            L->dgsBs.upload(L->gsBs, s);
            auto work = [](const std::vector<int>& st) {
                std::vector<int> w(st.size() - 1);
                for (size_t q = 0; q + 1 < st.size(); ++q) w[q] = st[q + 1] - st[q];
                return w;
            };
            L->runsF = levelRuns(work(L->gsFs), 2048);
            L->runsB = levelRuns(work(L->gsBs), 2048);
            L->x1.alloc(size_t(L->n));
            L->x2.alloc(size_t(L->n));
            L->x3.alloc(size_t(L->n));
            L->dinv.alloc(size_t(L->n));
            L->b.alloc(size_t(L->n));
            L->r.alloc(size_t(L->n));
            lv.push_back(std::move(L));
        }
        m = lv.back()->n;
        dense = m <= 800;
        if (dense) {
            M0.alloc(size_t(m) * 2 * m);
            M1.alloc(size_t(m) * 2 * m);
        }
        patRp = S.rp;
        patCi = S.ci;
        built = true;
    }
    void factorCoarsest(hipStream_t s) {
        if (!dense) return;
        AmgLevelDev& C = *lv.back();
        launchDenseFromCsr(s, m, C.rp, C.ci, C.v.p, M0.p);
        double* a = M0.p;
        double* b = M1.p;
        for (int k = 0; k < m;) {
            const int steps = k + 1 < m ? 2 : 1;
            launchGaussJordanStep(s, m, k, steps, a, b);
            std::swap(a, b);
            k += steps;
        }
        Minv = a;
    }
    // numeric set-up on the device values of S: weights, pressure matrix, hierarchy (fresh or refreshed)
    // acc: n accumulation blocks for true-IMPES weights (nullptr: quasi-IMPES); the cells that fell back to
    // quasi-IMPES are counted in fallbackDev
    DBuf<int> fallbackDev;
    void setup(const HostBsr& S, const int* drp, const int* dci, const int* ddiag, const double* val, bool reuse, hipStream_t s,
               bool f32 = false, const double* acc = nullptr) {
        n = S.n;
        nb = S.nb;
        w.alloc(size_t(n) * nb);
        xp.alloc(size_t(n));
        t.alloc(size_t(n) * nb);
        fallbackDev.alloc(1);
        fallbackDev.zero(s);
        launchCprWeights(s, nb, n, ddiag, val, w.p, acc, fallbackDev.p);
        const bool refresh = reuse && built && patRp == S.rp && patCi == S.ci;
        DBuf<double> apTmp;
        double* ap0;
        if (refresh) ap0 = lv[0]->v.p;
        else { apTmp.alloc(S.ci.size()); ap0 = apTmp.p; }
        launchCprPressure(s, nb, n, drp, val, w.p, ap0);
        launchFixDiag(s, n, ddiag, ap0);
        if (!refresh) {
            std::vector<double> h(S.ci.size());
            RSH_HIP(hipMemcpyAsync(h.data(), ap0, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
            RSH_HIP(hipStreamSynchronize(s));
            buildHierarchy(S, drp, dci, ddiag, h.data(), s);
        } else {
            for (size_t l = 0; l + 1 < lv.size(); ++l) {
                AmgLevelDev& L = *lv[l];
                AmgLevelDev& Cn = *lv[l + 1];
                launchGalerkin(s, int(L.ap.n), L.apPtr.p, L.apSrc.p, L.v.p, L.ap.p, Cn.nnz, L.cPtr.p, L.cSrc.p, Cn.v.p);
                launchFixDiag(s, Cn.n, Cn.diag, Cn.v.p);
            }
        }
        factorCoarsest(s);
        for (auto& L : lv) launchDinv(s, L->n, L->diag, L->v.p, L->dinv.p);
        lv[0]->plan = plan;
        for (auto& L : lv) L->packPlan(s, f32);
        have32 = f32;
        if (f32) {
            // the mixed-precision hierarchy: FP32 level values and dinv (the coarsest inverse stays FP64)
            for (auto& L : lv) {
                L->v32.alloc(L->v.n);
                L->dinv32.alloc(L->dinv.n);
                launchToFloat(s, (long long)L->v.n, L->v.p, L->v32.p);
                launchToFloat(s, (long long)L->dinv.n, L->dinv.p, L->dinv32.p);
            }
        }
    }
    // one Gauss-Seidel sweep of level L over its level schedule (fused runs of small levels)
    static void gsSweep(hipStream_t s, AmgLevelDev& L, const double* xin, double* xout, bool fwd) {
        const std::vector<LevelRun>& runs = fwd ? L.runsF : L.runsB;
        const std::vector<int>& st = fwd ? L.gsFs : L.gsBs;
        const int* rows = fwd ? L.gsF.p : L.gsB.p;
        const int* dst = fwd ? L.dgsFs.p : L.dgsBs.p;
        for (const LevelRun& run : runs) {
            if (run.fused) { launchGsFused(s, run.l0, run.l1, dst, rows, L.rp, L.ci, L.diag, L.v.p, L.b.p, xin, xout, fwd); continue; }
            const int q = run.l0;
            launchGsLevel(s, st[size_t(q) + 1] - st[size_t(q)], rows + st[size_t(q)], L.rp, L.ci, L.diag, L.v.p, L.b.p, xin, xout, fwd);
        }
    }
    // V-cycle from level l.  pfine (Jacobi smoothing): the finer level's iterate, to which this level's last kernel
    // adds P x_l (the prolongation); returns whether it did (else the caller prolongs)
    bool cycle(size_t l, hipStream_t s, bool f32 = false, double* pfine = nullptr) {
        AmgLevelDev& L = *lv[l];
        const int* prp = pfine ? lv[l - 1]->rrp.p : nullptr;   // aggregate members in the finer level
        const int* pci = pfine ? lv[l - 1]->rci.p : nullptr;
        if (l + 1 == lv.size()) {
            if (dense) {
                launchDenseGemv2(s, m, 2 * m, Minv + m, L.b.p, L.x1.p, prp, pci, pfine);
                L.result = L.x1.p;
                return pfine != nullptr;
            } else {
                // 20 symmetric Gauss-Seidel sweeps from zero (amg.cpp:315-316)
                launchFill(s, L.x1.p, 0.0, L.n);
                double* a = L.x1.p;
                double* b = L.x2.p;
                for (int sw = 0; sw < 40; ++sw) {
                    gsSweep(s, L, a, b, (sw % 2) == 0);
                    std::swap(a, b);
                }
                L.result = a;
            }
            return false;
        }
        AmgLevelDev& C = *lv[l + 1];
        if (smoother == RSH_SMOOTH_JACOBI) {
            // the first iterate x1 = w D^-1 b is in L.x1 (written with b); the pre-smoothing steps end with the
            // residual; the coarse level's last kernel adds P x_c to the pre-smoothed iterate; the post-smoothing
            // steps ping-pong between it and L.x2, the last one adding P x_l to the finer level's iterate
            const SliceDev<double>* sl = L.plan;
            if (sl) pfine = nullptr;   // a sliced level's caller adds its prolongation (the same one add per fine row)
            const double mean = double(L.nnz) / std::max(L.n, 1);
            const bool h = use32(f32);
            const void* sv = h ? static_cast<const void*>(L.pS32.p) : static_cast<const void*>(L.pS.p);
            const void* lvv = h ? static_cast<const void*>(L.pL32.p) : static_cast<const void*>(L.pL.p);
            const void* vv = h ? static_cast<const void*>(L.v32.p) : static_cast<const void*>(L.v.p);
            const void* dv = dinvOf(l, f32);
            const int nst = l > 0 && coarseSteps > 0 ? coarseSteps : steps;
            double *xa = L.x1.p, *xb = L.x3.p;
            for (int k = 1; k < nst; ++k) {
                launchJacobiFused(s, 1, h, sl, sv, lvv, L.n, L.rp, L.ci, vv, dv, L.b.p, xa, omega, xb, mean);
                std::swap(xa, xb);
            }
            launchJacobiFused(s, 0, h, sl, sv, lvv, L.n, L.rp, L.ci, vv, dv, L.b.p, xa, omega, L.r.p, mean);
            const bool smoothedC = l + 2 < lv.size();   // the coarse level is smoothed too: its x1 comes with its b
            launchRestrict(s, L.nc, L.rrp.p, L.rci.p, L.r.p, C.b.p, double(L.n) / std::max(L.nc, 1), smoothedC ? dinvOf(l + 1, f32) : nullptr,
                           h, omega, smoothedC ? C.x1.p : nullptr);
            if (!cycle(l + 1, s, f32, xa)) launchProlongAdd(s, L.n, L.agg.p, C.result, xa);
            double *pa = xa, *pb = L.x2.p;
            for (int k = 0; k < nst; ++k) {
                const bool last = k + 1 == nst;
                launchJacobiFused(s, 1, h, sl, sv, lvv, L.n, L.rp, L.ci, vv, dv, L.b.p, pa, omega, pb, mean, last ? prp : nullptr,
                                  last ? pci : nullptr, last ? pfine : nullptr);
                std::swap(pa, pb);
            }
            L.result = pa;
            return pfine != nullptr && !sl;
        }
        gsSweep(s, L, nullptr, L.x1.p, true);
        launchCsrSpmvResidual(s, L.n, L.rp, L.ci, L.v.p, L.x1.p, L.b.p, L.r.p);
        launchRestrict(s, L.nc, L.rrp.p, L.rci.p, L.r.p, C.b.p, double(L.n) / std::max(L.nc, 1));
        cycle(l + 1, s, f32);
        launchProlongAdd(s, L.n, L.agg.p, C.result, L.x1.p);
        gsSweep(s, L, L.x1.p, L.x2.p, false);
        L.result = L.x2.p;
        return false;
    }
    bool use32(bool f32) const { return f32 && have32; }
    const void* dinvOf(size_t l, bool f32) const {
        return use32(f32) ? static_cast<const void*>(lv[l]->dinv32.p) : static_cast<const void*>(lv[l]->dinv.p);
    }
    // one damped-Jacobi step of level l (x1 -> x3), the kernel the V-cycle smooths with; returns its algorithmic bytes:
    // values and column indices per entry, the row pointers, and D^-1, b, the iterate and the new iterate per row
    double smoothStep(size_t l, hipStream_t s, bool f32) {
        if (l + 1 >= lv.size()) throw Error(RSH_ERR_ARG, "no smoothed level " + std::to_string(l));
        AmgLevelDev& L = *lv[l];
        const bool h = use32(f32);
        const void* sv = h ? static_cast<const void*>(L.pS32.p) : static_cast<const void*>(L.pS.p);
        const void* lvv = h ? static_cast<const void*>(L.pL32.p) : static_cast<const void*>(L.pL.p);
        const void* vv = h ? static_cast<const void*>(L.v32.p) : static_cast<const void*>(L.v.p);
        launchJacobiFused(s, 1, h, L.plan, sv, lvv, L.n, L.rp, L.ci, vv, dinvOf(l, f32), L.b.p, L.x1.p, omega, L.x3.p,
                          double(L.nnz) / std::max(L.n, 1));
        const double pa = h ? 4.0 : 8.0;
        return double(L.nnz) * (pa + 4.0) + (L.n + 1) * 4.0 + double(L.n) * (pa + 3 * 8.0);
    }
    // level 0 is smoothed by damped Jacobi: its first iterate is needed with r_p
    bool jacobi0() const { return smoother == RSH_SMOOTH_JACOBI && lv.size() > 1; }
    // one V-cycle on r_p (in lv[0].b) -> result pointer; x1Ready: level 0's first Jacobi iterate was written with r_p
    const double* vcycle(hipStream_t s, bool f32 = false, bool x1Ready = false) {
        if (jacobi0() && !x1Ready) launchJacobiInit(s, lv[0]->n, use32(f32), dinvOf(0, f32), lv[0]->b.p, omega, lv[0]->x1.p);
        cycle(0, s, f32);
        return lv[0]->result;
    }
    size_t bytes() const {
        size_t b = w.bytes() + xp.bytes() + t.bytes() + M0.bytes() + M1.bytes();
        for (auto& L : lv)
            b += L->rpB.bytes() + L->ciB.bytes() + L->diagB.bytes() + L->v.bytes() + L->agg.bytes() + L->rrp.bytes() + L->rci.bytes() +
                 L->apPtr.bytes() + L->apSrc.bytes() + L->cPtr.bytes() + L->cSrc.bytes() + L->ap.bytes() + L->gsF.bytes() + L->gsB.bytes() +
                 L->x1.bytes() + L->x2.bytes() + L->x3.bytes() + L->b.bytes() + L->r.bytes() + L->v32.bytes() + L->dinv32.bytes() +
                 L->pS.bytes() + L->pL.bytes() + L->pS32.bytes() + L->pL32.bytes() + (L->ownPlan ? L->ownPlan->bytes() : 0);
        return b;
    }
};

// a member system of a batch (rsh_batch.cpp), for the serving loop's batched rounds: its block-size group's value
// array and row pointers from the member's first row (block indices into that array), its first value there, its first
// entry in the input-order vectors, and its sizes
struct BatchMember {
    double* groupVal;
    const int* rp;
    long long valOff, inOff;
    int n, nb, nnzb;
};
BatchMember batchMember(rsh_batch* b, int i);
// after writing members' values: every group refactors at the next rsh_batch_setup
void batchValuesChanged(rsh_batch* b);
hipStream_t batchStream(rsh_batch* b);
size_t batchDeviceBytes(rsh_batch* b);

inline double nowMs() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace rsh

using namespace rsh;   // internal header: the library's own translation units only

// ------------------------------------------------------------------ context and system
struct rsh_context {
    Ctx c;
    std::mutex mu;
};

struct rsh_system {
    rsh_context* ctx = nullptr;
    int n = 0, nb = 1, nnzb = 0;
    long long N = 0;
    HostBsr pat;
    DBuf<int> rp, ci, diag;
    DBuf<double> val;
    SliceHost spmvHost;
    SliceDev<double> spmv64;
    SliceDev<float> spmv32;
    bool have32 = false;
    // the FP64 operator in slices of several lanes per row (setSpmvLanes; empty: the one-lane slices above)
    SliceHost spmvWHost;
    SliceDev<double> spmvW64;
    DBuf<double> acc;                   // accumulation blocks for true-IMPES CPR weights (n*nb*nb; empty: quasi-IMPES)
    int impesFallbacks = 0;             // cells of the last set-up whose accumulation block fell back to quasi-IMPES
    DBuf<double> colP;                  // pressure column of S in the SpMV slice layout (CPR second stage)
    DBuf<float> colP32;                 // its FP32 copy (mixed precision)
    DBuf<double> bjInv;                 // block-Jacobi second stage: D_b^-1 per block row
    DBuf<float> bjInv32;
    rsh_options opt{};
    bool ready = false;
    bool valuesSet = false;
    std::unique_ptr<IluDev> ilu;
    std::unique_ptr<CprDev> cpr;
    bool keepIluEntries = false;        // member of a batch: the ILU level chunking can be retuned
    int iluFuseMax = 16;
    std::vector<int> rowSys;            // member of a batch: block row -> system of the batch
    const KScalars* gate = nullptr;     // member of a batch: the systems' states (stopped systems are skipped)
    // Krylov workspace
    DBuf<double> vb, vx, r, rh, p, v, sv, t, ph, sh, w, irR, irD, irX;
    DBuf<double> V, Z, hcol, yv;
    int restartAlloc = 0;
    // last solve
    int lastIts = 0, lastOuter = 0;
    double lastRel = 0, lastSetupMs = 0, lastSolveMs = 0, lastIluMs = 0, lastCprMs = 0;
    std::string lastStop;
    bool lastFallback = false;          // the last refinement fell back to the FP64 preconditioner
    std::vector<int> lastInner;         // the last refinement's inner iterations per outer step
    bool fp32Stalled = false;           // a refinement fell back to the FP64 preconditioner on this system

    hipStream_t s() const { return ctx->c.stream; }
    // the FP32 preconditioner of the mixed modes (asked for by the solve; FP32 ILU arithmetic)
    bool precF32() const { return opt.precision == RSH_MIXED || opt.precision == RSH_MIXED_IR; }
    // FP32-stored CPR pressure stage / block-Jacobi inverses for an apply in the given precision
    bool amgF32(bool f32) const { return f32 || opt.precision == RSH_MIXED_STORE; }
    // the ILU apply's IluFp for an apply in the given precision
    // (lean factors are float32 values: the FP32 slices give the same apply as the FP64 ones at half the bytes)
    int iluFp(bool f32) const {
        return f32 ? kIluFp32 : (opt.precision == RSH_MIXED_STORE || opt.ilu_lean) ? kIluStore32 : kIluFp64;
    }
    // lean factors in FP64 precision: FP32 factor slices, but the dense group inverses (formed in FP64 from the
    // rounded factors) stay FP64, so the apply is the reference's float-factor sweep up to FP64 rounding
    bool iluInv64(bool f32) const { return !f32 && opt.ilu_lean && opt.precision == RSH_FP64; }

    void ensureVec(DBuf<double>& b) { if (b.n != size_t(N)) b.alloc(size_t(N)); }
    // device memory the system holds: pattern, values and their slice layouts, preconditioner, Krylov vectors
    size_t deviceBytes() const {
        size_t b = val.bytes() + spmv64.bytes() + spmv32.bytes() + spmvW64.bytes() + colP.bytes() + colP32.bytes() + bjInv.bytes() + bjInv32.bytes() + rp.bytes() +
                   ci.bytes() + diag.bytes();
        if (ilu) b += ilu->bytes();
        if (cpr) b += cpr->bytes();
        for (const DBuf<double>* v : {&vb, &vx, &r, &rh, &p, &this->v, &sv, &t, &ph, &sh, &w, &irR, &irD, &irX, &V, &Z}) b += v->bytes();
        return b;
    }

    // new values are in val: repack the SpMV layouts; the preconditioner needs a new setup
    void valuesChanged() {
        packSlices(s(), nb, val.p, spmv64, spmvHost.nslots, false);
        if (have32) packSlices(s(), nb, val.p, spmv32, spmvHost.nslots, true);
        if (spmvW64.lanes > 1) packSlices(s(), nb, val.p, spmvW64, spmvWHost.nslots, false);
        valuesSet = true;
        ready = false;
    }
    int spmvLanes() const { return spmvW64.lanes > 1 ? spmvW64.lanes : 1; }
    // The FP64 operator's SpMV with `lanes` lanes per row (1, 2 or 4): 64 / lanes rows per wavefront, each row's
    // blocks split over its lanes (another summation order than one lane per row; the FP32 operator and the CPR
    // second stage keep the one-lane slices).  Batch members' slices carry their systems for the gating.
    void setSpmvLanes(int lanes) {
        if (lanes == spmvLanes()) return;
        spmvWHost = SliceHost{};
        spmvW64.release();
        if (lanes == 1) return;
        laneSliceLayout(n, pat.rp, pat.ci, lanes, spmvWHost);
        if (!rowSys.empty()) sliceOwners(spmvWHost, rowSys);
        uploadSlices(spmvWHost, spmvW64, nb, s());
        spmvW64.lanes = lanes;
        if (valuesSet) packSlices(s(), nb, val.p, spmvW64, spmvWHost.nslots, false);
    }

    // Preconditioner applies are captured once per (input, output, precision) into a HIP graph and replayed: a
    // level-scheduled apply is hundreds of small dependent launches, which the graph issues without host work.
    // captured applies keyed by their vectors and every option their kernels bake in (weights and step counts too)
    std::map<std::tuple<const void*, const void*, int, double, int>, hipGraphExec_t> graphs;
    std::pair<int, int> graphGen{-1, -1};
    bool useGraphs = [] { const char* e = std::getenv("RSH_GRAPHS"); return !(e && e[0] == '0'); }();
    void clearGraphs() {
        for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
        graphs.clear();
    }
    ~rsh_system() { clearGraphs(); }
    void precondG(const double* rr, double* z, bool f32) {
        if (!useGraphs) { precond(rr, z, f32); return; }
        const std::pair<int, int> gen{ilu ? ilu->generation : 0, cpr ? cpr->generation : 0};
        if (gen != graphGen) { clearGraphs(); graphGen = gen; }
        const auto key = std::make_tuple(static_cast<const void*>(rr), static_cast<const void*>(z),
                                         (f32 ? 1 : 0) | (opt.precond << 1) | (opt.smoother << 5) | (amgF32(f32) ? 1 << 6 : 0) |
                                             (opt.smooth_steps << 7),
                                         cpr ? cpr->omega : 0.0, cpr ? cpr->coarseSteps : 0);
        auto it = graphs.find(key);
        if (it == graphs.end()) {
            hipGraph_t g = nullptr;
            RSH_HIP(hipStreamBeginCapture(s(), hipStreamCaptureModeThreadLocal));
            try {
                precond(rr, z, f32);
            } catch (...) {
                (void)hipStreamEndCapture(s(), &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            RSH_HIP(hipStreamEndCapture(s(), &g));
            hipGraphExec_t ex = nullptr;
            const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) throw Error(RSH_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
            it = graphs.emplace(key, ex).first;
        }
        RSH_HIP(hipGraphLaunch(it->second, s()));
    }

    // gated: skip the rows of stopped batched systems (their outputs are then stale)
    void spmv(const double* x, double* y, bool f32, bool gated = true) {
        launchSpmvSliced(s(), nb, f32, spmvW64.lanes > 1 ? &spmvW64 : &spmv64, &spmv32, x, y, gated ? gate : nullptr);
    }
    // preconditioner z = M^-1 r (device vectors)
    void precond(const double* rr, double* z, bool f32) {
        const int pc = opt.precond;
        if (pc == RSH_PREC_BILU0 || pc == RSH_PREC_MCBILU0) {
            ilu->apply(rr, z, nullptr, iluFp(f32), s(), false, iluInv64(f32));
            return;
        }
        // CPR (bsr.cpp:1061-1084): r_p = w^T r -> V-cycle -> t = r - S(:,0) x_p -> BILU0 -> z_0 += x_p; DeviceCprAmg
        // (bsr.cpp:1319-1371) ends with the block-Jacobi second stage z = E x_p + D_b^-1 t instead of the BILU0
        CprDev& C = *cpr;
        const bool a32 = amgF32(f32);
        const bool j0 = C.jacobi0();
        launchCprRestrict(s(), nb, n, C.w.p, rr, C.lv[0]->b.p, j0 ? C.dinvOf(0, a32) : nullptr, C.use32(a32), C.omega,
                          j0 ? C.lv[0]->x1.p : nullptr);
        const double* xp = C.vcycle(s(), a32, j0);
        const bool p32 = a32 && colP32.n > 0;
        // the ILU's first level without lower entries: its forward copy of t is done here (FP64 working vector)
        const bool pre = pc != RSH_PREC_CPR_BJ && ilu && ilu->firstCopy && iluFp(f32) != kIluFp32;
        launchCprSecond(s(), nb, &spmv64, p32 ? static_cast<const void*>(colP32.p) : static_cast<const void*>(colP.p), p32, rr, xp, C.t.p,
                        pre ? ilu->iperm.p : nullptr, pre ? ilu->lstart[1] : 0, pre ? ilu->y64.p : nullptr);
        if (pc == RSH_PREC_CPR_BJ) {
            const bool b32 = a32 && bjInv32.n > 0;
            launchBlockJacobi(s(), nb, n, b32 ? static_cast<const void*>(bjInv32.p) : static_cast<const void*>(bjInv.p), b32, C.t.p, xp, z);
            return;
        }
        ilu->apply(C.t.p, z, xp, iluFp(f32), s(), pre, iluInv64(f32));
    }
};

