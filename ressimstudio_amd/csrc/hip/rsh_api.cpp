// C-ABI and solver drivers of the MI355X linear-solve library (include/ressim_hip.h).
//
// A system keeps its pattern, values, slice layouts, ILU schedule and AMG hierarchy resident on the device, so
// a Newton iteration only moves new values (set_values), refreshes the preconditioner numerically (setup) and
// runs the Krylov loop (solve).  All Krylov scalars live on the device; BiCGSTAB reads them back once per
// iteration, FGMRES once per Arnoldi step.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <tuple>

#include "../../../include/ressim_hip.h"
#include "rsh_internal.hpp"

using namespace rsh;

namespace {
thread_local std::string g_err;

int fail(int code, const std::string& msg) {
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
    d.src.upload(h.src, s);
    d.dsrc.upload(h.dsrc, s);
    d.lrow.upload(h.lrow, s);
    d.lorig.upload(h.lorig, s);
    d.lptr.upload(h.lptr, s);
    d.lcol.upload(h.lcol, s);
    d.lsrc.upload(h.lsrc, s);
    d.ldsrc.upload(h.ldsrc, s);
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
    int ngroups = 0, nmap = 0;
    long long elems = 0;
    DBuf<int> start, rows;
    DBuf<long long> off;
    DBuf<int4> map;
    DBuf<int2> tasks;
    DBuf<double> raw, inv;
    DBuf<float> inv32;

    // E: the sweep's entries per working row (src = factor block; dsrc = pivot inverse for the upper sweep)
    void build(const std::vector<LevelRun>& runs, const RowEntries& E, const std::vector<int>& lstart, int nb, bool upper, hipStream_t s) {
        std::vector<int> hstart(1, 0), hrows;
        std::vector<long long> hoff;
        std::vector<int4> hmap;
        std::vector<int2> htasks;
        ngroups = 0;
        elems = 0;
        runTask0.assign(runs.size(), -1);
        runTask1.assign(runs.size(), -1);
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
            runTask0[ri] = int(htasks.size());
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
                for (int a0 = 0; a0 < K; a0 += kWave) htasks.push_back(make_int2(gi, a0));
            }
            runTask1[ri] = int(htasks.size());
        }
        nmap = int(hmap.size());
        start.upload(hstart, s);
        rows.upload(hrows, s);
        off.upload(hoff, s);
        map.upload(hmap, s);
        tasks.upload(htasks, s);
        raw.alloc(size_t(elems));
        inv.alloc(size_t(elems));
        inv32.release();
    }
    void factor(int nb, bool upper, const double* lu, bool f32, hipStream_t s) {
        if (!ngroups) return;
        launchDenseGroups(s, nb, upper, ngroups, start.p, off.p, nmap, map.p, lu, raw.p, elems, inv.p);
        if (f32) {
            inv32.alloc(size_t(elems));
            launchToFloat(s, elems, inv.p, inv32.p);
        }
    }
    size_t bytes() const { return raw.bytes() + inv.bytes() + inv32.bytes(); }
};

// ------------------------------------------------------------------ level-scheduled block ILU(0)
struct IluDev {
    bool mc = false;
    int n = 0, nb = 1, colours = 0;
    std::vector<int> lstart;            // level (or colour) starts in working numbering
    DBuf<int> perm;                     // working row -> original row
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
    // sweep schedule: runs of levels (fused small levels or single launches) and device copies of level starts
    std::vector<LevelRun> runsL, runsU;
    DBuf<int> dLs, dLq, dUs, dUq;
    DenseSide denseL, denseU;
    DBuf<double> rt64;                  // outside contributions of the run being swept
    DBuf<float> rt32;
    void build(const HostBsr& S, const int* drp, const int* dci, const int* ddiag, bool multicolour, hipStream_t s) {
        ++generation;
        mc = multicolour;
        n = S.n;
        nb = S.nb;
        std::vector<int> ord, colour;
        RowEntries EL, EU;
        EL.ptr.assign(1, 0);
        EU.ptr.assign(1, 0);
        if (!mc) {
            levelOrder(S, ord, lstart);
            std::vector<int> iperm(static_cast<size_t>(n));
            for (int w = 0; w < n; ++w) iperm[size_t(ord[size_t(w)])] = w;
            rp = drp; ci = dci; diag = ddiag;
            for (int w = 0; w < n; ++w) {
                const int i = ord[size_t(w)];
                for (int p = S.rp[size_t(i)]; p < S.diag[size_t(i)]; ++p) { EL.col.push_back(iperm[size_t(S.ci[size_t(p)])]); EL.src.push_back(p); }
                for (int p = S.diag[size_t(i)] + 1; p < S.rp[size_t(i) + 1]; ++p) { EU.col.push_back(iperm[size_t(S.ci[size_t(p)])]); EU.src.push_back(p); }
                EL.ptr.push_back(int(EL.col.size()));
                EU.ptr.push_back(int(EU.col.size()));
                EU.dsrc.push_back(S.diag[size_t(i)]);
            }
        } else {
            colours = colourGraph(S, colour, ord, lstart);
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
        EL.orig = ord;
        EU.orig = ord;
        perm.upload(ord, s);
        std::vector<int> lenL(static_cast<size_t>(n)), lenU(static_cast<size_t>(n));
        for (int w = 0; w < n; ++w) { lenL[size_t(w)] = EL.ptr[size_t(w) + 1] - EL.ptr[size_t(w)]; lenU[size_t(w)] = EU.ptr[size_t(w) + 1] - EU.ptr[size_t(w)]; }
        buildSlices(EL, lstart, rowCap(lenL), hL, sL);
        buildSlices(EU, lstart, rowCap(lenU), hU, sU);
        uploadSlices(hL, L64, nb, s);
        uploadSlices(hU, U64, nb, s);
        {
            const int nl = int(lstart.size()) - 1;
            std::vector<int> wL(static_cast<size_t>(nl)), wU(static_cast<size_t>(nl));
            for (int l = 0; l < nl; ++l) {
                wL[size_t(l)] = (sL.sStart[size_t(l) + 1] - sL.sStart[size_t(l)]) + (sL.qStart[size_t(l) + 1] - sL.qStart[size_t(l)]);
                wU[size_t(l)] = (sU.sStart[size_t(l) + 1] - sU.sStart[size_t(l)]) + (sU.qStart[size_t(l) + 1] - sU.qStart[size_t(l)]);
            }
            std::vector<int> rowsL(static_cast<size_t>(nl));
            for (int l = 0; l < nl; ++l) rowsL[size_t(l)] = lstart[size_t(l) + 1] - lstart[size_t(l)];
            runsL = levelRuns(wL, 16, &rowsL, kRunRowsMax);
            runsU = levelRuns(wU, 16, &rowsL, kRunRowsMax);
            denseL.build(runsL, EL, lstart, nb, false, s);
            denseU.build(runsU, EU, lstart, nb, true, s);
            dLs.upload(sL.sStart, s);
            dLq.upload(sL.qStart, s);
            dUs.upload(sU.sStart, s);
            dUq.upload(sU.qStart, s);
            // factor rows (F numbering) per level, long rows apart
            const HostBsr& F = mc ? fpat : S;
            std::vector<int> sh, lo;
            fShortStart.assign(1, 0);
            fLongStart.assign(1, 0);
            for (int l = 0; l < nl; ++l) {
                for (int w = lstart[size_t(l)]; w < lstart[size_t(l) + 1]; ++w) {
                    const int i = mc ? w : ord[size_t(w)];
                    const int lower = F.diag[size_t(i)] - F.rp[size_t(i)];
                    (lower > 24 ? lo : sh).push_back(i);
                }
                fShortStart.push_back(int(sh.size()));
                fLongStart.push_back(int(lo.size()));
            }
            fShort.upload(sh, s);
            fLong.upload(lo, s);
        }
        lu.alloc(size_t(S.rp[size_t(n)]) * nb * nb);
        y64.alloc(size_t(n + 1) * nb);
        y64.zero(s);
        rt64.alloc(size_t(kRunRowsMax) * nb);
        nreg.alloc(1);
    }
    void ensure32(hipStream_t s) {
        if (have32) return;
        uploadSlices(hL, L32, nb, s);
        uploadSlices(hU, U32, nb, s);
        y32.alloc(size_t(n + 1) * nb);
        y32.zero(s);
        rt32.alloc(size_t(kRunRowsMax) * nb);
        have32 = true;
    }
    // numeric factorisation from the system's values (natural layout), then the sweep layouts
    void factor(const double* val, bool f32, hipStream_t s) {
        const size_t nvals = lu.n;
        if (mc) launchGatherBlocks(s, nb, val, fmap.p, int(nvals / (size_t(nb) * nb)), lu.p);
        else RSH_HIP(hipMemcpyAsync(lu.p, val, nvals * sizeof(double), hipMemcpyDeviceToDevice, s));
        nreg.zero(s);
        const int nl = int(lstart.size()) - 1;
        for (int l = 0; l < nl; ++l) {
            launchBilu0Level(s, nb, fShort.p, fShortStart[size_t(l)], fShortStart[size_t(l) + 1], rp, ci, diag, lu.p, nreg.p);
            launchBilu0LongRows(s, nb, fLong.p, fLongStart[size_t(l)], fLongStart[size_t(l) + 1], rp, ci, diag, lu.p, nreg.p);
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
    void apply(const double* r, double* z, const double* xp, bool f32, hipStream_t s) const {
        const void* Lp = f32 ? static_cast<const void*>(&L32) : static_cast<const void*>(&L64);
        const void* Up = f32 ? static_cast<const void*>(&U32) : static_cast<const void*>(&U64);
        void* y = f32 ? static_cast<void*>(y32.p) : static_cast<void*>(y64.p);
        void* rt = f32 ? static_cast<void*>(rt32.p) : static_cast<void*>(rt64.p);
        auto sweepRun = [&](const LevelRun& run, size_t ri, bool fwd) {
            const DenseSide& D = fwd ? denseL : denseU;
            const LevelSched& sc = fwd ? sL : sU;
            const void* sl = fwd ? Lp : Up;
            const int wb = lstart[size_t(run.l0)], we = lstart[size_t(run.l1)];
            if (D.runTask0[ri] >= 0) {
                const int t0 = sc.sStart[size_t(run.l0)], u0 = sc.qStart[size_t(run.l0)];
                const void* inv = f32 ? static_cast<const void*>(D.inv32.p) : static_cast<const void*>(D.inv.p);
                launchRunDense(s, nb, f32, sl, fwd, t0, sc.sStart[size_t(run.l1)] - t0, u0, sc.qStart[size_t(run.l1)] - u0, wb, we,
                               D.runTask1[ri] - D.runTask0[ri], D.tasks.p + D.runTask0[ri], D.start.p, D.rows.p, D.off.p, inv, perm.p, r,
                               rt, y, z, xp);
                return;
            }
            launchIluFused(s, nb, f32, sl, fwd, run.l0, run.l1, fwd ? dLs.p : dUs.p, fwd ? dLq.p : dUq.p, wb, we - wb, perm.p, r, y, z, xp);
        };
        for (size_t ri = 0; ri < runsL.size(); ++ri) {
            const LevelRun& run = runsL[ri];
            if (run.fused) { sweepRun(run, ri, true); continue; }
            const int l = run.l0;
            launchIluForward(s, nb, f32, Lp, sL.sStart[size_t(l)], sL.sStart[size_t(l) + 1], sL.qStart[size_t(l)], sL.qStart[size_t(l) + 1], r, y);
        }
        for (size_t k = runsU.size(); k-- > 0;) {
            const LevelRun& run = runsU[k];
            if (run.fused) { sweepRun(run, k, false); continue; }
            const int l = run.l0;
            launchIluBackward(s, nb, f32, Up, sU.sStart[size_t(l)], sU.sStart[size_t(l) + 1], sU.qStart[size_t(l)], sU.qStart[size_t(l) + 1], y, z, xp);
        }
    }
    size_t bytes() const {
        return lu.bytes() + L64.bytes() + U64.bytes() + L32.bytes() + U32.bytes() + y64.bytes() + y32.bytes() + denseL.bytes() + denseU.bytes();
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
    DBuf<double> dinv;                   // 1 / a_ii (Jacobi)
    const double* result = nullptr;
};

struct CprDev {
    static constexpr double kOmega = 0.6667;   // damped-Jacobi weight of the reference's device V-cycle (bsr.h:173)
    int n = 0, nb = 1;
    DBuf<double> w, xp, t;
    std::vector<std::unique_ptr<AmgLevelDev>> lv;
    std::vector<int> hostAggSig;          // sizes per level, for reporting
    int m = 0;                            // coarsest size
    bool dense = false;
    DBuf<double> M0, M1;                  // [C | I] ping-pong; the inverse ends in Minv
    double* Minv = nullptr;
    int smoother = RSH_SMOOTH_GS;
    bool built = false;
    // level 0 values in the SpMV row-slice layout of S (same pattern) for the streamed Jacobi steps
    const SliceDev<double>* plan = nullptr;
    DBuf<double> apS, apL;
    std::vector<int> patRp, patCi;        // pattern of the hierarchy (reuse check)

    int generation = 0;
    void buildHierarchy(const HostBsr& S, const int* drp, const int* dci, const int* ddiag, const double* apHost, hipStream_t s) {
        ++generation;
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
        for (int k = 0; k < m; ++k) {
            launchGaussJordanStep(s, m, k, a, b);
            std::swap(a, b);
        }
        Minv = a;
    }
    // numeric set-up on the device values of S: weights, pressure matrix, hierarchy (fresh or refreshed)
    void setup(const HostBsr& S, const int* drp, const int* dci, const int* ddiag, const double* val, bool reuse, hipStream_t s) {
        n = S.n;
        nb = S.nb;
        w.alloc(size_t(n) * nb);
        xp.alloc(size_t(n));
        t.alloc(size_t(n) * nb);
        launchCprWeights(s, nb, n, ddiag, val, w.p);
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
        if (plan) {
            apS.alloc(size_t(plan->val.n / (size_t(nb) * nb)));
            apL.alloc(plan->lsrc.n);
            launchPackSlices(s, 1, lv[0]->v.p, plan->src.p, (long long)(apS.n / kWave), apS.p, false);
            if (plan->lsrc.n) launchPackLong(s, 1, lv[0]->v.p, plan->lsrc.p, int(plan->lsrc.n), apL.p, false);
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
    void cycle(size_t l, hipStream_t s) {
        AmgLevelDev& L = *lv[l];
        if (l + 1 == lv.size()) {
            if (dense) {
                launchDenseGemvLd(s, L, L.b.p, L.x1.p);
                L.result = L.x1.p;
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
            return;
        }
        AmgLevelDev& C = *lv[l + 1];
        if (smoother == RSH_SMOOTH_JACOBI) {
            // pre-smoothing from zero fused with the residual, prolongation fused with the post-smoothing
            const SliceDev<double>* sl = (l == 0 && plan) ? plan : nullptr;
            launchJacobiFused(s, 0, sl, apS.p, apL.p, L.n, L.rp, L.ci, L.v.p, L.dinv.p, L.b.p, nullptr, nullptr, nullptr, kOmega, L.x1.p, L.r.p,
                              double(L.nnz) / std::max(L.n, 1));
            launchRestrict(s, L.nc, L.rrp.p, L.rci.p, L.r.p, C.b.p, double(L.n) / std::max(L.nc, 1));
            cycle(l + 1, s);
            launchJacobiFused(s, 1, sl, apS.p, apL.p, L.n, L.rp, L.ci, L.v.p, L.dinv.p, L.b.p, L.x1.p, L.agg.p, C.result, kOmega, L.x2.p, nullptr,
                              double(L.nnz) / std::max(L.n, 1));
            L.result = L.x2.p;
            return;
        }
        gsSweep(s, L, nullptr, L.x1.p, true);
        launchCsrSpmvResidual(s, L.n, L.rp, L.ci, L.v.p, L.x1.p, L.b.p, L.r.p);
        launchRestrict(s, L.nc, L.rrp.p, L.rci.p, L.r.p, C.b.p, double(L.n) / std::max(L.nc, 1));
        cycle(l + 1, s);
        launchProlongAdd(s, L.n, L.agg.p, C.result, L.x1.p);
        gsSweep(s, L, L.x1.p, L.x2.p, false);
        L.result = L.x2.p;
    }
    void launchDenseGemvLd(hipStream_t s, AmgLevelDev& L, const double* bv, double* out);
    // one V-cycle on r_p (in lv[0].b) -> result pointer
    const double* vcycle(hipStream_t s) {
        cycle(0, s);
        return lv[0]->result;
    }
    size_t bytes() const {
        size_t b = w.bytes() + xp.bytes() + t.bytes() + M0.bytes() + M1.bytes();
        for (auto& L : lv)
            b += L->rpB.bytes() + L->ciB.bytes() + L->diagB.bytes() + L->v.bytes() + L->agg.bytes() + L->rrp.bytes() + L->rci.bytes() +
                 L->apPtr.bytes() + L->apSrc.bytes() + L->cPtr.bytes() + L->cSrc.bytes() + L->ap.bytes() + L->gsF.bytes() + L->gsB.bytes() +
                 L->x1.bytes() + L->x2.bytes() + L->b.bytes() + L->r.bytes();
        return b;
    }
};

}  // namespace

// dense y = Minv b with the inverse in the right half of [I | C^-1] (leading dimension 2m)
void CprDev::launchDenseGemvLd(hipStream_t s, AmgLevelDev& L, const double* bv, double* out) {
    launchDenseGemv2(s, m, 2 * m, Minv + m, bv, out);
    (void)L;
}

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
    DBuf<double> colP;                  // pressure column of S in the SpMV slice layout (CPR second stage)
    rsh_options opt{};
    bool ready = false;
    bool valuesSet = false;
    std::unique_ptr<IluDev> ilu;
    std::unique_ptr<CprDev> cpr;
    // Krylov workspace
    DBuf<double> vb, vx, r, rh, p, v, sv, t, ph, sh, w, irR, irD;
    DBuf<double> V, Z, hcol, yv;
    int restartAlloc = 0;
    // last solve
    int lastIts = 0, lastOuter = 0;
    double lastRel = 0, lastSetupMs = 0, lastSolveMs = 0, lastIluMs = 0, lastCprMs = 0;
    std::string lastStop;

    hipStream_t s() const { return ctx->c.stream; }

    void ensureVec(DBuf<double>& b) { if (b.n != size_t(N)) b.alloc(size_t(N)); }

    // Preconditioner applies are captured once per (input, output, precision) into a HIP graph and replayed: a
    // level-scheduled apply is hundreds of small dependent launches, which the graph issues without host work.
    std::map<std::tuple<const void*, const void*, int>, hipGraphExec_t> graphs;
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
        const auto key = std::make_tuple(static_cast<const void*>(rr), static_cast<const void*>(z), (f32 ? 1 : 0) | (opt.precond << 1) | (opt.smoother << 5));
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

    void spmv(const double* x, double* y, bool f32) {
        launchSpmvSliced(s(), nb, f32, &spmv64, &spmv32, x, y);
    }
    // preconditioner z = M^-1 r (device vectors)
    void precond(const double* rr, double* z, bool f32) {
        const int pc = opt.precond;
        if (pc == RSH_PREC_BILU0 || pc == RSH_PREC_MCBILU0) {
            ilu->apply(rr, z, nullptr, f32, s());
            return;
        }
        // CPR (bsr.cpp:1061-1084): r_p = w^T r -> V-cycle -> t = r - S(:,0) x_p -> BILU0 -> z_0 += x_p
        CprDev& C = *cpr;
        launchCprRestrict(s(), nb, n, C.w.p, rr, C.lv[0]->b.p);
        const double* xp = C.vcycle(s());
        launchCprSecond(s(), nb, &spmv64, colP.p, rr, xp, C.t.p);
        ilu->apply(C.t.p, z, xp, f32, s());
    }
};

namespace {

double nowMs() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void syncScalars(rsh_system& S) {
    Ctx& c = S.ctx->c;
    RSH_HIP(hipMemcpyAsync(c.hks, c.ks.p, sizeof(KScalars), hipMemcpyDeviceToHost, c.stream));
    RSH_HIP(hipStreamSynchronize(c.stream));
}

// Right-preconditioned BiCGSTAB from x0 = 0 (linsolve.cpp:159-216); the loop structure, breakdown tests and
// the two convergence tests (on ||s|| and on the recursive ||r||) are the reference's.
// precF32: the preconditioner's factors, hierarchy and sweeps in FP32; opF32: the operator too (inner solves of the
// refinement loop).  Vectors, dot products and the Krylov recurrences stay FP64.
int bicgstab(rsh_system& S, const double* b, double* x, double tol, int maxIt, int& its, double& rel, bool precF32 = false,
             bool opF32 = false) {
    Ctx& c = S.ctx->c;
    hipStream_t st = c.stream;
    const long long N = S.N;
    for (auto* bf : {&S.r, &S.rh, &S.p, &S.v, &S.sv, &S.t, &S.ph, &S.sh}) S.ensureVec(*bf);
    launchFill(st, x, 0.0, N);
    launchCopy(st, S.r.p, b, N);
    launchCopy(st, S.rh.p, b, N);
    launchFill(st, S.p.p, 0.0, N);
    launchFill(st, S.v.p, 0.0, N);
    KScalars init{};
    init.rho = 1; init.alpha = 1; init.omega = 1; init.tol = tol; init.state = kRun; init.it = 0;
    *c.hks = init;
    RSH_HIP(hipMemcpyAsync(c.ks.p, c.hks, sizeof(KScalars), hipMemcpyHostToDevice, st));
    launchDots(st, c, N, 1, b, b, nullptr, nullptr, c.scal.p, kEpBn, c.ks.p);
    syncScalars(S);
    its = 0;
    rel = 0;
    if (c.hks->bn == 0) { S.lastStop = "zero rhs"; return RSH_OK; }
    for (int it = 1; it <= maxIt; ++it) {
        launchDots(st, c, N, 1, S.rh.p, S.r.p, nullptr, nullptr, c.scal.p, kEpRho, c.ks.p);
        launchBicgP(st, N, c.ks.p, S.r.p, S.v.p, S.p.p);
        S.precondG(S.p.p, S.ph.p, precF32);
        S.spmv(S.ph.p, S.v.p, opF32);
        launchDots(st, c, N, 1, S.rh.p, S.v.p, nullptr, nullptr, c.scal.p, kEpAlpha, c.ks.p);
        launchBicgS(st, c, N, c.ks.p, S.r.p, S.v.p, S.sv.p);
        launchBicgXs(st, N, c.ks.p, S.ph.p, x);
        syncScalars(S);
        if (c.hks->state != kRun) break;
        S.precondG(S.sv.p, S.sh.p, precF32);
        S.spmv(S.sh.p, S.t.p, opF32);
        launchDots(st, c, N, 2, S.t.p, S.t.p, S.t.p, S.sv.p, c.scal.p, kEpOmega, c.ks.p);
        launchBicgXR(st, c, N, c.ks.p, S.ph.p, S.sh.p, S.sv.p, S.t.p, x, S.r.p);
        syncScalars(S);
        if (c.hks->state != kRun) break;
    }
    const KScalars& k = *c.hks;
    its = k.it;
    if (k.state == kConvS) { rel = k.sn / k.bn; S.lastStop = "converged (s)"; return RSH_OK; }
    rel = k.bn > 0 ? k.rn / k.bn : 0;
    if (k.state == kConv) { S.lastStop = "converged"; return RSH_OK; }
    if (k.state == kBreak) { S.lastStop = "breakdown"; return RSH_ERR_BREAKDOWN; }
    S.lastStop = k.state == kOmegaZero ? "omega = 0" : "max iterations";
    return RSH_ERR_NOT_CONVERGED;
}

// Flexible restarted GMRES (linsolve.cpp:218-273): x0 = 0 (or the given x when warm), Z_j = M V_j stored, the
// true residual recomputed at each restart, the estimate |g_{j+1}| / ||b|| tested each step.  Orthogonalisation:
// classical Gram-Schmidt applied twice (CGS2, one multi-dot and one multi-axpy per pass) in FP64.
int fgmres(rsh_system& S, const double* b, double* x, double tol, int maxIt, int m, bool precF32, bool opF32, bool warm, int& its,
           double& rel) {
    Ctx& c = S.ctx->c;
    hipStream_t st = c.stream;
    const long long N = S.N;
    if (m < 1) m = 1;
    if (m > 31) m = 31;
    if (S.restartAlloc < m) {
        S.V.alloc(size_t(m + 1) * size_t(N));
        S.Z.alloc(size_t(m) * size_t(N));
        S.restartAlloc = m;
    }
    S.ensureVec(S.r);
    S.ensureVec(S.w);
    S.hcol.alloc(64);
    S.yv.alloc(64);
    if (!warm) launchFill(st, x, 0.0, N);
    launchDots(st, c, N, 1, b, b, nullptr, nullptr, c.scal.p, kEpNone, nullptr);
    RSH_HIP(hipMemcpyAsync(c.hscal, c.scal.p, sizeof(double), hipMemcpyDeviceToHost, st));
    RSH_HIP(hipStreamSynchronize(st));
    const double bn = std::sqrt(c.hscal[0]);
    its = 0;
    rel = 0;
    if (bn == 0) { launchFill(st, x, 0.0, N); return RSH_OK; }
    std::vector<double> H(static_cast<size_t>(m + 1) * m, 0.0), cs(static_cast<size_t>(m)), sn(static_cast<size_t>(m)), g(static_cast<size_t>(m) + 1), y(static_cast<size_t>(m));
    auto h = [&](int i, int j) -> double& { return H[size_t(i) * m + j]; };
    int total = 0;
    while (total < maxIt) {
        S.spmv(x, S.r.p, false);                                   // true residual in FP64
        launchSub(st, N, b, S.r.p, S.r.p);
        launchDots(st, c, N, 1, S.r.p, S.r.p, nullptr, nullptr, c.scal.p, kEpNone, nullptr);
        RSH_HIP(hipMemcpyAsync(c.hscal, c.scal.p, sizeof(double), hipMemcpyDeviceToHost, st));
        RSH_HIP(hipStreamSynchronize(st));
        const double beta = std::sqrt(c.hscal[0]);
        rel = beta / bn;
        if (rel < tol) { its = total; return RSH_OK; }
        c.hscal[1] = beta;
        RSH_HIP(hipMemcpyAsync(c.scal.p + 1, c.hscal + 1, sizeof(double), hipMemcpyHostToDevice, st));
        launchScaleTo(st, N, c.scal.p + 1, S.r.p, S.V.p, true);     // V_0 = r / beta
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int j = 0;
        bool done = false;
        while (j < m && total < maxIt) {
            double* Vj = S.V.p + size_t(j) * N;
            double* Zj = S.Z.p + size_t(j) * N;
            S.precondG(Vj, Zj, precF32);
            S.spmv(Zj, S.w.p, opF32);
            // CGS2: h = V^T w; w -= V h; h2 = V^T w; w -= V h2
            launchMultiDot(st, c, N, j + 1, S.V.p, N, S.w.p, S.hcol.p);
            launchMultiAxpy(st, N, j + 1, S.V.p, N, S.hcol.p, S.w.p);
            launchMultiDot(st, c, N, j + 1, S.V.p, N, S.w.p, S.hcol.p + 32);
            launchMultiAxpy(st, N, j + 1, S.V.p, N, S.hcol.p + 32, S.w.p);
            launchDots(st, c, N, 1, S.w.p, S.w.p, nullptr, nullptr, S.hcol.p + 63, kEpNone, nullptr);
            std::vector<double> hh(64);
            RSH_HIP(hipMemcpyAsync(hh.data(), S.hcol.p, 64 * sizeof(double), hipMemcpyDeviceToHost, st));
            RSH_HIP(hipStreamSynchronize(st));
            for (int i = 0; i <= j; ++i) h(i, j) = hh[size_t(i)] + hh[size_t(32 + i)];
            const double hn = std::sqrt(std::max(hh[63], 0.0));
            h(j + 1, j) = hn;
            c.hscal[2] = hn;
            RSH_HIP(hipMemcpyAsync(c.scal.p + 2, c.hscal + 2, sizeof(double), hipMemcpyHostToDevice, st));
            if (hn > 0) launchScaleTo(st, N, c.scal.p + 2, S.w.p, S.V.p + size_t(j + 1) * N, true);
            for (int i = 0; i < j; ++i) {
                const double tmp = cs[size_t(i)] * h(i, j) + sn[size_t(i)] * h(i + 1, j);
                h(i + 1, j) = -sn[size_t(i)] * h(i, j) + cs[size_t(i)] * h(i + 1, j);
                h(i, j) = tmp;
            }
            const double den = std::hypot(h(j, j), h(j + 1, j));
            cs[size_t(j)] = den > 0 ? h(j, j) / den : 1;
            sn[size_t(j)] = den > 0 ? h(j + 1, j) / den : 0;
            h(j, j) = den;
            h(j + 1, j) = 0;
            g[size_t(j) + 1] = -sn[size_t(j)] * g[size_t(j)];
            g[size_t(j)] = cs[size_t(j)] * g[size_t(j)];
            rel = std::fabs(g[size_t(j) + 1]) / bn;
            ++j;
            ++total;
            if (rel < tol) { done = true; break; }
        }
        for (int i = j - 1; i >= 0; --i) {
            double sacc = g[size_t(i)];
            for (int k = i + 1; k < j; ++k) sacc -= h(i, k) * y[size_t(k)];
            y[size_t(i)] = sacc / h(i, i);
        }
        RSH_HIP(hipMemcpyAsync(S.yv.p, y.data(), size_t(j) * sizeof(double), hipMemcpyHostToDevice, st));
        launchMultiAxpyAdd(st, N, j, S.Z.p, N, S.yv.p, x);
        RSH_HIP(hipStreamSynchronize(st));
        if (done) { its = total; return RSH_OK; }
    }
    its = total;
    return RSH_ERR_NOT_CONVERGED;
}

// RSH_MIXED_IR: FP64 iterative refinement x += d, r = b - A x (FP64 SpMV, as refineWithDevice bsr.cpp:1373-1396)
// around inner solves (BiCGSTAB, as the reference's float device solve bsr.cpp:1130-1215, or FGMRES) on the FP32
// matrix and the FP32 preconditioner; their dot products and Krylov vectors stay FP64.  A correction that does not
// reduce the FP64 residual (a diverged inner solve) is taken back and the loop stops.
int refine(rsh_system& S, const double* b, double* x, double tol, int maxIt, int& its, double& rel, int& outer) {
    Ctx& c = S.ctx->c;
    hipStream_t st = c.stream;
    const long long N = S.N;
    S.ensureVec(S.irR);  // residual
    S.ensureVec(S.irD);  // correction
    launchFill(st, x, 0.0, N);
    launchCopy(st, S.irR.p, b, N);
    launchDots(st, c, N, 1, b, b, nullptr, nullptr, c.scal.p, kEpNone, nullptr);
    RSH_HIP(hipMemcpyAsync(c.hscal, c.scal.p, sizeof(double), hipMemcpyDeviceToHost, st));
    RSH_HIP(hipStreamSynchronize(st));
    const double bn = std::sqrt(c.hscal[0]);
    its = 0;
    outer = 0;
    rel = 0;
    if (bn == 0) return RSH_OK;
    const double innerTol = S.opt.inner_tol > 0 ? S.opt.inner_tol : 1e-5;
    const int maxOuter = S.opt.max_outer > 0 ? S.opt.max_outer : 10;
    double rn = bn;
    for (outer = 0; outer < maxOuter; ++outer) {
        rel = rn / bn;
        if (rel <= tol) return RSH_OK;
        int inIts = 0;
        double inRel = 0;
        const double itol = std::max(innerTol, tol * bn / rn * 0.5);
        const int left = std::max(1, maxIt - its);
        const int rc = S.opt.krylov == RSH_KRYLOV_FGMRES
                           ? fgmres(S, S.irR.p, S.irD.p, itol, left, S.opt.restart > 0 ? S.opt.restart : 30, true, true, false, inIts, inRel)
                           : bicgstab(S, S.irR.p, S.irD.p, itol, left, inIts, inRel, true, true);
        its += inIts;
        if (rc != RSH_OK && rc != RSH_ERR_NOT_CONVERGED && rc != RSH_ERR_BREAKDOWN) return rc;
        launchAxpy(st, N, 1.0, S.irD.p, x);
        S.spmv(x, S.irR.p, false);
        launchSub(st, N, b, S.irR.p, S.irR.p);
        launchDots(st, c, N, 1, S.irR.p, S.irR.p, nullptr, nullptr, c.scal.p, kEpNone, nullptr);
        RSH_HIP(hipMemcpyAsync(c.hscal, c.scal.p, sizeof(double), hipMemcpyDeviceToHost, st));
        RSH_HIP(hipStreamSynchronize(st));
        const double rnew = std::sqrt(c.hscal[0]);
        if (!(rnew < rn)) {
            // no progress (or a non-finite correction): take the step back, keep the last residual
            launchAxpy(st, N, -1.0, S.irD.p, x);
            if (!std::isfinite(rnew)) {
                S.spmv(x, S.irR.p, false);
                launchSub(st, N, b, S.irR.p, S.irR.p);
            }
            S.lastStop = "refinement stalled";
            break;
        }
        rn = rnew;
        if (its >= maxIt) break;
    }
    rel = rn / bn;
    return rel <= tol ? RSH_OK : RSH_ERR_NOT_CONVERGED;
}

int solveDevice(rsh_system& S, const double* b, double* x, double tol, int maxIt, int* iterations, double* relRes) {
    if (!S.ready) return fail(RSH_ERR_ARG, "rsh_solve: call rsh_system_setup first");
    const double t0 = nowMs();
    int its = 0, outer = 0, rc;
    double rel = 0;
    const bool precF32 = S.opt.precision == RSH_MIXED;
    if (S.opt.precision == RSH_MIXED_IR) rc = refine(S, b, x, tol, maxIt, its, rel, outer);
    else if (S.opt.krylov == RSH_KRYLOV_FGMRES)
        rc = fgmres(S, b, x, tol, maxIt, S.opt.restart > 0 ? S.opt.restart : 30, precF32, false, false, its, rel);
    else rc = bicgstab(S, b, x, tol, maxIt, its, rel, precF32, false);
    RSH_HIP(hipStreamSynchronize(S.s()));
    S.lastSolveMs = nowMs() - t0;
    S.lastIts = its;
    S.lastOuter = outer;
    S.lastRel = rel;
    if (iterations) *iterations = its;
    if (relRes) *relRes = rel;
    if (!std::isfinite(rel)) return fail(RSH_ERR_NONFINITE, "non-finite residual");
    if (rc == RSH_ERR_NOT_CONVERGED) return fail(rc, "not converged: " + S.lastStop);
    if (rc == RSH_ERR_BREAKDOWN) return fail(rc, "Krylov breakdown");
    return rc;
}

std::once_flag g_devOnce;
int g_devCount = -1;

}  // namespace

// ================================================================== C-ABI
extern "C" {

int rsh_version(void) { return 1; }
const char* rsh_last_error(void) { return g_err.c_str(); }

int rsh_device_count(void) {
    std::call_once(g_devOnce, [] {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
        g_devCount = n;
    });
    return g_devCount;
}

void rsh_default_options(rsh_options* o) {
    std::memset(o, 0, sizeof(*o));
    o->precond = RSH_PREC_CPR;
    o->krylov = RSH_KRYLOV_BICGSTAB;
    o->precision = RSH_FP64;
    o->smoother = RSH_SMOOTH_GS;
    o->restart = 30;
    o->max_outer = 10;
    o->inner_tol = 1e-5;
    o->reuse_amg = 1;
}

int rsh_context_create(int device, rsh_context** out) {
    return guarded([&] {
        if (!out) return fail(RSH_ERR_ARG, "null output pointer");
        *out = nullptr;
        if (rsh_device_count() <= 0) return fail(RSH_ERR_DEVICE, "no HIP device");
        if (device < 0 || device >= rsh_device_count()) return fail(RSH_ERR_ARG, "device index out of range");
        auto ctx = std::make_unique<rsh_context>();
        ctx->c.device = device;
        RSH_HIP(hipSetDevice(device));
        RSH_HIP(hipStreamCreateWithFlags(&ctx->c.stream, hipStreamNonBlocking));
        ctx->c.red.alloc(size_t(kRedBlocks) * 32);
        ctx->c.ticket.alloc(1);
        ctx->c.ticket.zero(ctx->c.stream);
        ctx->c.ks.alloc(1);
        ctx->c.scal.alloc(64);
        RSH_HIP(hipHostMalloc(reinterpret_cast<void**>(&ctx->c.hks), sizeof(KScalars), hipHostMallocDefault));
        RSH_HIP(hipHostMalloc(reinterpret_cast<void**>(&ctx->c.hscal), 64 * sizeof(double), hipHostMallocDefault));
        RSH_HIP(hipStreamSynchronize(ctx->c.stream));
        *out = ctx.release();
        return RSH_OK;
    });
}

int rsh_context_destroy(rsh_context* ctx) {
    if (!ctx) return RSH_OK;
    return guarded([&] {
        RSH_HIP(hipSetDevice(ctx->c.device));
        (void)hipStreamSynchronize(ctx->c.stream);
        ctx->c.red.release();
        ctx->c.ticket.release();
        ctx->c.ks.release();
        ctx->c.scal.release();
        if (ctx->c.hks) (void)hipHostFree(ctx->c.hks);
        if (ctx->c.hscal) (void)hipHostFree(ctx->c.hscal);
        (void)hipStreamDestroy(ctx->c.stream);
        delete ctx;
        return RSH_OK;
    });
}

int rsh_context_synchronize(rsh_context* ctx) {
    return guarded([&] {
        RSH_HIP(hipStreamSynchronize(ctx->c.stream));
        return RSH_OK;
    });
}

int rsh_system_create(rsh_context* ctx, int n, int nb, const int* row_ptr, const int* col, rsh_system** out) {
    return guarded([&] {
        if (!ctx || !row_ptr || !col || !out) return fail(RSH_ERR_ARG, "null argument");
        if (n <= 0) return fail(RSH_ERR_ARG, "n must be positive");
        if (nb < 1 || nb > kMaxNb) return fail(RSH_ERR_UNSUPPORTED, "block size must be 1..4");
        if (row_ptr[0] != 0) return fail(RSH_ERR_ARG, "row_ptr[0] must be 0");
        const int nnzb = row_ptr[n];
        if (nnzb < n) return fail(RSH_ERR_ARG, "pattern has fewer blocks than rows");
        auto S = std::make_unique<rsh_system>();
        S->ctx = ctx;
        S->n = n;
        S->nb = nb;
        S->nnzb = nnzb;
        S->N = (long long)n * nb;
        S->pat.n = n;
        S->pat.nb = nb;
        S->pat.rp.assign(row_ptr, row_ptr + n + 1);
        S->pat.ci.assign(col, col + nnzb);
        S->pat.diag.assign(size_t(n), -1);
        for (int i = 0; i < n; ++i) {
            if (row_ptr[i + 1] < row_ptr[i]) return fail(RSH_ERR_ARG, "row_ptr not monotone");
            for (int p = row_ptr[i]; p < row_ptr[i + 1]; ++p) {
                const int j = col[p];
                if (j < 0 || j >= n) return fail(RSH_ERR_ARG, "column index out of range");
                if (p > row_ptr[i] && col[p - 1] >= j) return fail(RSH_ERR_ARG, "columns must be sorted and unique within a row");
                if (j == i) S->pat.diag[size_t(i)] = p;
            }
            if (S->pat.diag[size_t(i)] < 0) return fail(RSH_ERR_ARG, "every row needs its diagonal block");
        }
        RSH_HIP(hipSetDevice(ctx->c.device));
        hipStream_t st = ctx->c.stream;
        S->rp.upload(S->pat.rp, st);
        S->ci.upload(S->pat.ci, st);
        S->diag.upload(S->pat.diag, st);
        S->val.alloc(size_t(nnzb) * nb * nb);
        // SpMV slices over all rows in natural order; padding points at the lane's own row
        RowEntries E;
        E.ptr.assign(S->pat.rp.begin(), S->pat.rp.end());
        E.col = S->pat.ci;
        E.src.resize(size_t(nnzb));
        for (int p = 0; p < nnzb; ++p) E.src[size_t(p)] = p;
        E.orig.resize(size_t(n));
        for (int i = 0; i < n; ++i) E.orig[size_t(i)] = i;
        std::vector<int> len(static_cast<size_t>(n));
        for (int i = 0; i < n; ++i) len[size_t(i)] = row_ptr[i + 1] - row_ptr[i];
        LevelSched one;
        buildSlices(E, {0, n}, rowCap(len), S->spmvHost, one);
        for (int sl = 0; sl < S->spmvHost.nslices; ++sl)
            for (int k = 0; k < S->spmvHost.width[size_t(sl)]; ++k)
                for (int lane = 0; lane < kSliceRows; ++lane) {
                    const size_t at = (size_t(S->spmvHost.off[size_t(sl)]) + size_t(k)) * kSliceRows + size_t(lane);
                    if (S->spmvHost.src[at] < 0) {
                        const int rr = S->spmvHost.row[size_t(sl) * kSliceRows + size_t(lane)];
                        S->spmvHost.col[at] = rr >= 0 ? rr : 0;
                    }
                }
        uploadSlices(S->spmvHost, S->spmv64, nb, st);
        RSH_HIP(hipStreamSynchronize(st));
        *out = S.release();
        return RSH_OK;
    });
}

int rsh_system_destroy(rsh_system* sys) {
    if (!sys) return RSH_OK;
    return guarded([&] {
        RSH_HIP(hipSetDevice(sys->ctx->c.device));
        (void)hipStreamSynchronize(sys->s());
        delete sys;
        return RSH_OK;
    });
}

static int afterValues(rsh_system& S) {
    hipStream_t st = S.s();
    packSlices(st, S.nb, S.val.p, S.spmv64, S.spmvHost.nslots, false);
    if (S.have32) packSlices(st, S.nb, S.val.p, S.spmv32, S.spmvHost.nslots, true);
    S.valuesSet = true;
    S.ready = false;
    return RSH_OK;
}

int rsh_system_set_values(rsh_system* sys, const double* val) {
    return guarded([&] {
        if (!sys || !val) return fail(RSH_ERR_ARG, "null argument");
        RSH_HIP(hipSetDevice(sys->ctx->c.device));
        RSH_HIP(hipMemcpyAsync(sys->val.p, val, sys->val.bytes(), hipMemcpyHostToDevice, sys->s()));
        return afterValues(*sys);
    });
}

int rsh_system_set_values_device(rsh_system* sys, const double* val_dev) {
    return guarded([&] {
        if (!sys || !val_dev) return fail(RSH_ERR_ARG, "null argument");
        RSH_HIP(hipSetDevice(sys->ctx->c.device));
        RSH_HIP(hipMemcpyAsync(sys->val.p, val_dev, sys->val.bytes(), hipMemcpyDeviceToDevice, sys->s()));
        return afterValues(*sys);
    });
}

int rsh_system_setup(rsh_system* sys, const rsh_options* opt) {
    return guarded([&] {
        if (!sys) return fail(RSH_ERR_ARG, "null system");
        if (!sys->valuesSet) return fail(RSH_ERR_ARG, "set values before setup");
        rsh_options o;
        rsh_default_options(&o);
        if (opt) o = *opt;
        if (o.precond < RSH_PREC_BILU0 || o.precond > RSH_PREC_CPR_MC) return fail(RSH_ERR_ARG, "unknown preconditioner");
        if (o.precision != RSH_FP64 && o.precision != RSH_MIXED && o.precision != RSH_MIXED_IR) return fail(RSH_ERR_ARG, "unknown precision");
        RSH_HIP(hipSetDevice(sys->ctx->c.device));
        rsh_system& S = *sys;
        hipStream_t st = S.s();
        const double t0 = nowMs();
        const bool f32 = o.precision != RSH_FP64;             // FP32 preconditioner
        if (o.precision == RSH_MIXED_IR && !S.have32) {       // FP32 operator of the inner solves
            uploadSlices(S.spmvHost, S.spmv32, S.nb, st);
            packSlices(st, S.nb, S.val.p, S.spmv32, S.spmvHost.nslots, true);
            S.have32 = true;
        }
        const bool wantMc = o.precond == RSH_PREC_MCBILU0 || o.precond == RSH_PREC_CPR_MC;
        if (!S.ilu || S.ilu->mc != wantMc) {
            S.ilu = std::make_unique<IluDev>();
            S.ilu->build(S.pat, S.rp.p, S.ci.p, S.diag.p, wantMc, st);
        }
        if (o.precond == RSH_PREC_CPR || o.precond == RSH_PREC_CPR_MC) {
            if (!S.cpr) S.cpr = std::make_unique<CprDev>();
            S.cpr->smoother = o.smoother;
            S.cpr->plan = &S.spmv64;
            if (S.colP.n == 0) {
                S.colP.alloc(size_t(S.spmvHost.nslots) * S.nb * kWave);
            }
            launchPackColumn0(st, S.nb, S.val.p, S.spmv64.src.p, S.spmvHost.nslots, S.colP.p);
            S.cpr->setup(S.pat, S.rp.p, S.ci.p, S.diag.p, S.val.p, o.reuse_amg != 0, st);
        }
        RSH_HIP(hipStreamSynchronize(st));
        const double t1 = nowMs();
        S.ilu->factor(S.val.p, f32, st);
        RSH_HIP(hipStreamSynchronize(st));
        S.lastIluMs = nowMs() - t1;
        S.lastCprMs = t1 - t0;
        S.opt = o;
        S.ready = true;
        S.lastSetupMs = nowMs() - t0;
        return RSH_OK;
    });
}

int rsh_solve_device(rsh_system* sys, const double* b_dev, double* x_dev, double tol, int max_it, int* iterations, double* rel_res) {
    return guarded([&] {
        if (!sys || !b_dev || !x_dev) return fail(RSH_ERR_ARG, "null argument");
        RSH_HIP(hipSetDevice(sys->ctx->c.device));
        return solveDevice(*sys, b_dev, x_dev, tol, max_it, iterations, rel_res);
    });
}

int rsh_solve(rsh_system* sys, const double* b, double* x, double tol, int max_it, int* iterations, double* rel_res) {
    return guarded([&] {
        if (!sys || !b || !x) return fail(RSH_ERR_ARG, "null argument");
        RSH_HIP(hipSetDevice(sys->ctx->c.device));
        rsh_system& S = *sys;
        S.ensureVec(S.vb);
        S.ensureVec(S.vx);
        RSH_HIP(hipMemcpyAsync(S.vb.p, b, size_t(S.N) * sizeof(double), hipMemcpyHostToDevice, S.s()));
        const int rc = solveDevice(S, S.vb.p, S.vx.p, tol, max_it, iterations, rel_res);
        RSH_HIP(hipMemcpyAsync(x, S.vx.p, size_t(S.N) * sizeof(double), hipMemcpyDeviceToHost, S.s()));
        RSH_HIP(hipStreamSynchronize(S.s()));
        if (rc == RSH_OK) {
            for (long long i = 0; i < S.N; ++i)
                if (!std::isfinite(x[i])) return fail(RSH_ERR_NONFINITE, "non-finite solution");
        }
        return rc;
    });
}

int rsh_spmv(rsh_system* sys, const double* x, double* y) {
    return guarded([&] {
        if (!sys || !x || !y) return fail(RSH_ERR_ARG, "null argument");
        if (!sys->valuesSet) return fail(RSH_ERR_ARG, "set values first");
        RSH_HIP(hipSetDevice(sys->ctx->c.device));
        rsh_system& S = *sys;
        S.ensureVec(S.vb);
        S.ensureVec(S.vx);
        RSH_HIP(hipMemcpyAsync(S.vb.p, x, size_t(S.N) * sizeof(double), hipMemcpyHostToDevice, S.s()));
        S.spmv(S.vb.p, S.vx.p, false);
        RSH_HIP(hipMemcpyAsync(y, S.vx.p, size_t(S.N) * sizeof(double), hipMemcpyDeviceToHost, S.s()));
        RSH_HIP(hipStreamSynchronize(S.s()));
        return RSH_OK;
    });
}

int rsh_precond_apply(rsh_system* sys, const double* r, double* z) {
    return guarded([&] {
        if (!sys || !r || !z) return fail(RSH_ERR_ARG, "null argument");
        if (!sys->ready) return fail(RSH_ERR_ARG, "call rsh_system_setup first");
        RSH_HIP(hipSetDevice(sys->ctx->c.device));
        rsh_system& S = *sys;
        S.ensureVec(S.vb);
        S.ensureVec(S.vx);
        RSH_HIP(hipMemcpyAsync(S.vb.p, r, size_t(S.N) * sizeof(double), hipMemcpyHostToDevice, S.s()));
        S.precondG(S.vb.p, S.vx.p, S.opt.precision != RSH_FP64);
        RSH_HIP(hipMemcpyAsync(z, S.vx.p, size_t(S.N) * sizeof(double), hipMemcpyDeviceToHost, S.s()));
        RSH_HIP(hipStreamSynchronize(S.s()));
        return RSH_OK;
    });
}

int rsh_amg_apply(rsh_system* sys, const double* r, double* z) {
    return guarded([&] {
        if (!sys || !r || !z) return fail(RSH_ERR_ARG, "null argument");
        if (!sys->ready || !sys->cpr) return fail(RSH_ERR_ARG, "no CPR preconditioner set up");
        RSH_HIP(hipSetDevice(sys->ctx->c.device));
        rsh_system& S = *sys;
        CprDev& C = *S.cpr;
        RSH_HIP(hipMemcpyAsync(C.lv[0]->b.p, r, size_t(S.n) * sizeof(double), hipMemcpyHostToDevice, S.s()));
        const double* xp = C.vcycle(S.s());
        RSH_HIP(hipMemcpyAsync(z, xp, size_t(S.n) * sizeof(double), hipMemcpyDeviceToHost, S.s()));
        RSH_HIP(hipStreamSynchronize(S.s()));
        return RSH_OK;
    });
}

int rsh_cpr_weights(rsh_system* sys, double* w) {
    return guarded([&] {
        if (!sys || !w) return fail(RSH_ERR_ARG, "null argument");
        if (!sys->ready || !sys->cpr) return fail(RSH_ERR_ARG, "no CPR preconditioner set up");
        RSH_HIP(hipMemcpy(w, sys->cpr->w.p, size_t(sys->N) * sizeof(double), hipMemcpyDeviceToHost));
        return RSH_OK;
    });
}

int rsh_bilu0_factors(rsh_system* sys, double* lu) {
    return guarded([&] {
        if (!sys || !lu) return fail(RSH_ERR_ARG, "null argument");
        if (!sys->ready || !sys->ilu) return fail(RSH_ERR_ARG, "call rsh_system_setup first");
        IluDev& I = *sys->ilu;
        const size_t b2 = size_t(sys->nb) * sys->nb;
        std::vector<double> h(I.lu.n);
        RSH_HIP(hipMemcpy(h.data(), I.lu.p, I.lu.bytes(), hipMemcpyDeviceToHost));
        if (!I.mc) { std::memcpy(lu, h.data(), I.lu.bytes()); return RSH_OK; }
        // multicolour factors are stored for P A P^T: return them in that (permuted) layout
        std::memcpy(lu, h.data(), I.lu.bytes());
        (void)b2;
        return RSH_OK;
    });
}

int rsh_amg_levels(rsh_system* sys, int* nlevels, int* sizes, int* nnz, int cap) {
    return guarded([&] {
        if (!sys || !nlevels) return fail(RSH_ERR_ARG, "null argument");
        if (!sys->cpr) { *nlevels = 0; return RSH_OK; }
        const auto& lv = sys->cpr->lv;
        *nlevels = int(lv.size());
        for (int l = 0; l < int(lv.size()) && l < cap; ++l) {
            if (sizes) sizes[l] = lv[size_t(l)]->n;
            if (nnz) nnz[l] = lv[size_t(l)]->nnz;
        }
        return RSH_OK;
    });
}

int rsh_system_info(rsh_system* sys, char* json, int cap) {
    return guarded([&] {
        if (!sys || !json || cap <= 0) return fail(RSH_ERR_ARG, "null argument");
        rsh_system& S = *sys;
        std::ostringstream o;
        o << "{\"n\":" << S.n << ",\"nb\":" << S.nb << ",\"nnzb\":" << S.nnzb << ",\"spmv_slices\":" << S.spmvHost.nslices
          << ",\"spmv_slots\":" << S.spmvHost.nslots << ",\"spmv_long_rows\":" << S.spmvHost.lrow.size();
        if (S.ilu) {
            o << ",\"ilu_levels\":" << S.ilu->sL.levels() << ",\"ilu_multicolour\":" << (S.ilu->mc ? 1 : 0) << ",\"ilu_regularised\":" << S.ilu->regularised
              << ",\"ilu_long_rows_L\":" << S.ilu->hL.lrow.size() << ",\"ilu_long_rows_U\":" << S.ilu->hU.lrow.size() << ",\"ilu_bytes\":" << S.ilu->bytes();
            int fusedL = 0, fusedU = 0, denseRunsL = 0, denseRunsU = 0;
            for (size_t i = 0; i < S.ilu->runsL.size(); ++i) { fusedL += S.ilu->runsL[i].fused; denseRunsL += S.ilu->denseL.runTask0[i] >= 0; }
            for (size_t i = 0; i < S.ilu->runsU.size(); ++i) { fusedU += S.ilu->runsU[i].fused; denseRunsU += S.ilu->denseU.runTask0[i] >= 0; }
            o << ",\"ilu_fused_runs\":[" << fusedL << "," << fusedU << "],\"ilu_dense_runs\":[" << denseRunsL << "," << denseRunsU
              << "],\"ilu_dense_groups\":[" << S.ilu->denseL.ngroups << "," << S.ilu->denseU.ngroups << "]";
        }
        if (S.cpr) {
            o << ",\"amg_levels\":[";
            for (size_t l = 0; l < S.cpr->lv.size(); ++l) o << (l ? "," : "") << S.cpr->lv[l]->n;
            o << "],\"amg_nnz\":[";
            for (size_t l = 0; l < S.cpr->lv.size(); ++l) o << (l ? "," : "") << S.cpr->lv[l]->nnz;
            o << "],\"gs_levels\":[";
            for (size_t l = 0; l < S.cpr->lv.size(); ++l) o << (l ? "," : "") << (S.cpr->lv[l]->gsFs.size() - 1);
            o << "],\"coarse_dense\":" << (S.cpr->dense ? 1 : 0) << ",\"cpr_bytes\":" << S.cpr->bytes();
        }
        o << ",\"last_iterations\":" << S.lastIts << ",\"last_outer\":" << S.lastOuter << ",\"last_rel_res\":" << S.lastRel
 << ",\"last_setup_ms\":" << S.lastSetupMs << ",\"last_setup_ilu_ms\":" << S.lastIluMs << ",\"last_setup_other_ms\":" << S.lastCprMs << ",\"last_solve_ms\":" << S.lastSolveMs << ",\"last_stop\":\"" << S.lastStop << "\"}";
        const std::string str = o.str();
        if (int(str.size()) + 1 > cap) return fail(RSH_ERR_ARG, "info buffer too small");
        std::memcpy(json, str.c_str(), str.size() + 1);
        return int(str.size());
    });
}

int rsh_device_alloc(rsh_context* ctx, size_t bytes, void** ptr) {
    return guarded([&] {
        if (!ctx || !ptr) return fail(RSH_ERR_ARG, "null argument");
        RSH_HIP(hipSetDevice(ctx->c.device));
        RSH_HIP(hipMalloc(ptr, bytes));
        return RSH_OK;
    });
}
int rsh_device_free(rsh_context* ctx, void* ptr) {
    return guarded([&] {
        if (!ctx) return fail(RSH_ERR_ARG, "null context");
        RSH_HIP(hipSetDevice(ctx->c.device));
        RSH_HIP(hipStreamSynchronize(ctx->c.stream));
        if (ptr) RSH_HIP(hipFree(ptr));
        return RSH_OK;
    });
}
int rsh_memcpy_h2d(rsh_context* ctx, void* dst, const void* src, size_t bytes) {
    return guarded([&] {
        RSH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->c.stream));
        RSH_HIP(hipStreamSynchronize(ctx->c.stream));
        return RSH_OK;
    });
}
int rsh_memcpy_d2h(rsh_context* ctx, void* dst, const void* src, size_t bytes) {
    return guarded([&] {
        RSH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->c.stream));
        RSH_HIP(hipStreamSynchronize(ctx->c.stream));
        return RSH_OK;
    });
}

int rsh_time_kernel(rsh_system* sys, int kind, int reps, double* avg_ms, double* bytes_per_launch) {
    return guarded([&] {
        if (!sys || !avg_ms || reps < 1) return fail(RSH_ERR_ARG, "bad argument");
        rsh_system& S = *sys;
        if ((kind == 1 || kind == 2 || kind == 3) && !S.ready) return fail(RSH_ERR_ARG, "call rsh_system_setup first");
        if (kind == 3 && !S.cpr) return fail(RSH_ERR_ARG, "no AMG");
        RSH_HIP(hipSetDevice(S.ctx->c.device));
        hipStream_t st = S.s();
        S.ensureVec(S.vb);
        S.ensureVec(S.vx);
        launchFill(st, S.vb.p, 1.0, S.N);
        const bool f32 = S.ready && S.opt.precision != RSH_FP64;          // preconditioner precision
        const bool op32 = S.ready && S.opt.precision == RSH_MIXED_IR;     // operator precision
        const double nb2 = double(S.nb) * S.nb;
        const double pa = f32 ? 4.0 : 8.0;
        double bytes = 0;
        auto once = [&]() {
            switch (kind) {
                case 0: S.spmv(S.vb.p, S.vx.p, op32); break;
                case 1: S.precondG(S.vb.p, S.vx.p, f32); break;
                case 2: S.ilu->apply(S.vb.p, S.vx.p, nullptr, f32, st); break;
                case 3: S.cpr->vcycle(st); break;
                case 4: launchCopyKernel(st, S.vb.p, S.vx.p, S.N); break;
                default: throw Error(RSH_ERR_ARG, "unknown kernel kind");
            }
        };
        // algorithmic bytes: values + indices + vectors
        if (kind == 0) bytes = S.nnzb * (nb2 * (op32 ? 4.0 : 8.0) + 4.0) + (S.n + 1) * 4.0 + 2.0 * S.N * 8.0;
        else if (kind == 2) bytes = S.nnzb * (nb2 * pa + 4.0) + 2.0 * (S.n + 1) * 4.0 + 4.0 * S.N * (f32 ? 4.0 : 8.0);
        else if (kind == 4) bytes = 2.0 * S.N * 8.0;
        once();   // warm-up
        hipEvent_t e0, e1;
        RSH_HIP(hipEventCreate(&e0));
        RSH_HIP(hipEventCreate(&e1));
        RSH_HIP(hipEventRecord(e0, st));
        for (int r = 0; r < reps; ++r) once();
        RSH_HIP(hipEventRecord(e1, st));
        RSH_HIP(hipEventSynchronize(e1));
        float ms = 0;
        RSH_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *avg_ms = double(ms) / reps;
        if (bytes_per_launch) *bytes_per_launch = bytes;
        return RSH_OK;
    });
}

}  // extern "C"
