// The library's host-side set-up, built with g++ on its own (no device code) so that it runs under the host
// sanitizers (AddressSanitizer, UndefinedBehaviorSanitizer): level schedules, colouring and the colour-permuted
// pattern, row-slice layouts and level runs (csrc/hip/rsh_setup.cpp), the aggregation AMG set-up with its Galerkin maps,
// and the well Schur plan (csrc/host/well_plan.hpp).  Every result is also checked for what it must satisfy.
//
// usage: host_check <systems file>
// systems file: int32 count; per system int32 n, nb, row_ptr[n+1], col[nnzb], float64 val[nnzb nb nb], then int32
// has_jacobian and, when 1, the scalar Jacobian it was reduced from: int32 rows, na, nv, nw, row_ptr[rows+1],
// col[nnz], float64 val[nnz].  Prints one line per system and exits non-zero on the first failed check.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../hip/rsh_internal.hpp"
#include "well_plan.hpp"

using namespace rsh;

namespace {
struct Input {
    int n = 0, nb = 1;
    std::vector<int> rp, ci;
    std::vector<double> val;
    bool jac = false;
    int rows = 0, na = 0, nv = 0, nw = 0;
    std::vector<int> jrp, jci;
    std::vector<double> jval;
};

FILE* g_in = nullptr;
bool rd(void* p, size_t bytes) { return bytes == 0 || std::fread(p, 1, bytes, g_in) == bytes; }

int fails = 0;
void check(bool ok, const char* what, int sys) {
    if (ok) return;
    std::printf("system %d: FAILED %s\n", sys, what);
    ++fails;
}

void run(const Input& in, int id) {
    HostBsr A;
    A.n = in.n;
    A.nb = in.nb;
    A.rp = in.rp;
    A.ci = in.ci;
    A.diag.assign(size_t(A.n), -1);
    for (int i = 0; i < A.n; ++i)
        for (int p = A.rp[size_t(i)]; p < A.rp[size_t(i) + 1]; ++p)
            if (A.ci[size_t(p)] == i) A.diag[size_t(i)] = p;
    // level schedule: a permutation in which every coupled pair keeps its order and sits on different levels
    std::vector<int> ord, start;
    levelOrder(A, ord, start);
    std::vector<int> level(size_t(A.n), -1);
    for (int l = 0; l + 1 < int(start.size()); ++l)
        for (int w = start[size_t(l)]; w < start[size_t(l) + 1]; ++w) level[size_t(ord[size_t(w)])] = l;
    bool okLev = int(ord.size()) == A.n;
    for (int i = 0; i < A.n && okLev; ++i) {
        okLev = level[size_t(i)] >= 0;
        for (int p = A.rp[size_t(i)]; p < A.rp[size_t(i) + 1] && okLev; ++p) {
            const int j = A.ci[size_t(p)];
            if (j < i) okLev = level[size_t(j)] < level[size_t(i)];
        }
    }
    check(okLev, "level order", id);
    // colouring: neighbours differ; perm lists the rows colour by colour
    std::vector<int> colour, perm, cstart;
    const int ncol = colourGraph(A, colour, perm, cstart);
    bool okCol = int(perm.size()) == A.n && int(cstart.size()) == ncol + 1 && cstart.back() == A.n;
    for (int i = 0; i < A.n && okCol; ++i)
        for (int p = A.rp[size_t(i)]; p < A.rp[size_t(i) + 1] && okCol; ++p)
            if (A.ci[size_t(p)] != i) okCol = colour[size_t(i)] != colour[size_t(A.ci[size_t(p)])];
    for (int c = 0; c < ncol && okCol; ++c)
        for (int k = cstart[size_t(c)]; k < cstart[size_t(c) + 1] && okCol; ++k) okCol = colour[size_t(perm[size_t(k)])] == c;
    check(okCol, "colouring", id);
    HostBsr P;
    std::vector<int> bmap;
    permutePattern(A, perm, P, bmap);
    bool okPerm = P.n == A.n && P.rp.back() == A.rp.back() && int(bmap.size()) == A.rp.back();
    for (int i = 0; i < P.n && okPerm; ++i)
        for (int p = P.rp[size_t(i)] + 1; p < P.rp[size_t(i) + 1] && okPerm; ++p) okPerm = P.ci[size_t(p) - 1] < P.ci[size_t(p)];
    check(okPerm, "permuted pattern", id);
    // slices of the lower and upper entries in level order, as the sweeps build them
    std::vector<int> iperm(size_t(A.n));
    for (int w = 0; w < A.n; ++w) iperm[size_t(ord[size_t(w)])] = w;
    RowEntries EL, EU;
    EL.ptr.assign(1, 0);
    EU.ptr.assign(1, 0);
    std::vector<int> lenL, lenU;
    for (int w = 0; w < A.n; ++w) {
        const int i = ord[size_t(w)];
        for (int p = A.rp[size_t(i)]; p < A.diag[size_t(i)]; ++p) { EL.col.push_back(iperm[size_t(A.ci[size_t(p)])]); EL.src.push_back(p); }
        for (int p = A.diag[size_t(i)] + 1; p < A.rp[size_t(i) + 1]; ++p) { EU.col.push_back(iperm[size_t(A.ci[size_t(p)])]); EU.src.push_back(p); }
        EL.ptr.push_back(int(EL.col.size()));
        EU.ptr.push_back(int(EU.col.size()));
        EU.dsrc.push_back(A.diag[size_t(i)]);
        lenL.push_back(EL.ptr[size_t(w) + 1] - EL.ptr[size_t(w)]);
        lenU.push_back(EU.ptr[size_t(w) + 1] - EU.ptr[size_t(w)]);
    }
    EL.orig = ord;
    EU.orig = ord;
    SliceHost hL, hU;
    LevelSched sL, sU;
    buildSlices(EL, start, rowCap(lenL), hL, sL);
    buildSlices(EU, start, rowCap(lenU), hU, sU);
    long long placed = 0;
    for (int v : hL.src) placed += v >= 0;
    for (int v : hL.lsrc) placed += v >= 0;
    check(placed == (long long)EL.col.size() && sL.levels() == int(start.size()) - 1, "lower slices hold every entry once", id);
    // red / black tiles: a permutation; a bulk row's lower entries lie in its own tile on an earlier unit (red and
    // black) or in a red tile (black rows); far rows come last; every lower entry of a bulk row has a smaller
    // factorisation level; the tile units' slices hold no long rows
    static const int rbDims[3][3] = {{4, 4, 4}, {8, 8, 2}, {3, 5, 2}};
    int rbOrders = 0;
    for (const int* dims : rbDims) {
        RbTiles R;
        if (!rbTileOrder(A, dims, R)) continue;
        ++rbOrders;
        const int n = A.n;
        bool ok = int(R.ord.size()) == n && R.nRed > 0 && R.nRed < R.nTile && R.nTile <= n && R.unitStart.front() == 0 &&
                  R.unitStart.back() == R.nTile && int(R.unitLevel.size()) == int(R.unitStart.size()) - 1 &&
                  R.tileUnit.back() == int(R.unitStart.size()) - 1 && R.farStart.back() == n - R.nTile;
        std::vector<int> seen(size_t(n), 0), wOf(size_t(n), -1), unitOf(size_t(n), -1), tileOf(size_t(n), -1);
        for (int w = 0; w < n && ok; ++w) {
            ok = R.ord[size_t(w)] >= 0 && R.ord[size_t(w)] < n && ++seen[size_t(R.ord[size_t(w)])] == 1;
            if (ok) wOf[size_t(R.ord[size_t(w)])] = w;
        }
        for (size_t t = 0; t + 1 < R.tileUnit.size() && ok; ++t)
            for (int u = R.tileUnit[t]; u < R.tileUnit[t + 1]; ++u)
                for (int w = R.unitStart[size_t(u)]; w < R.unitStart[size_t(u) + 1]; ++w) {
                    unitOf[size_t(w)] = u;
                    tileOf[size_t(w)] = int(t);
                }
        for (int w = 0; w < R.nTile && ok; ++w) {
            const int i = R.ord[size_t(w)];
            ok = unitOf[size_t(w)] >= 0 && (w < R.nRed) == (tileOf[size_t(w)] < R.redTiles);
            for (int p = A.rp[size_t(i)]; p < A.rp[size_t(i) + 1] && ok; ++p) {
                const int c = wOf[size_t(A.ci[size_t(p)])];
                if (c >= w) continue;
                const bool sameTile = tileOf[size_t(c)] == tileOf[size_t(w)];
                ok = sameTile ? unitOf[size_t(c)] < unitOf[size_t(w)] : (w >= R.nRed && c < R.nRed);
                ok = ok && R.unitLevel[size_t(unitOf[size_t(c)])] < R.unitLevel[size_t(unitOf[size_t(w)])];
            }
        }
        check(ok, "red / black tile order", id);
        // the tile sweeps' slices as the ILU builds them: one range of slices per tile, every row of the range in the
        // tile's working rows, no long rows among the tile units, every slot's column inside the vector (or the zero row n)
        HostBsr PR;
        std::vector<int> bmR;
        permutePattern(A, R.ord, PR, bmR);
        RowEntries RL, RU;
        RL.ptr.assign(1, 0);
        RU.ptr.assign(1, 0);
        for (int w = 0; w < n; ++w) {
            for (int p = PR.rp[size_t(w)]; p < PR.diag[size_t(w)]; ++p) { RL.col.push_back(PR.ci[size_t(p)]); RL.src.push_back(p); }
            for (int p = PR.diag[size_t(w)] + 1; p < PR.rp[size_t(w) + 1]; ++p) { RU.col.push_back(PR.ci[size_t(p)]); RU.src.push_back(p); }
            RL.ptr.push_back(int(RL.col.size()));
            RU.ptr.push_back(int(RU.col.size()));
            RU.dsrc.push_back(PR.diag[size_t(w)]);
        }
        RL.orig = R.ord;
        RU.orig = R.ord;
        std::vector<int> bounds = R.unitStart;
        for (size_t k = 1; k < R.farStart.size(); ++k) bounds.push_back(R.nTile + R.farStart[k]);
        for (const RowEntries* E : {&RL, &RU}) {
            std::vector<int> lens(static_cast<size_t>(n));
            for (int w = 0; w < n; ++w) lens[size_t(w)] = E->ptr[size_t(w) + 1] - E->ptr[size_t(w)];
            int cap = rowCap(lens);
            for (int w = 0; w < R.nTile; ++w) cap = std::max(cap, lens[size_t(w)]);
            SliceHost h;
            LevelSched u;
            buildSlices(*E, bounds, cap, h, u);
            const int nu = int(R.unitStart.size()) - 1;
            bool okS = int(u.sStart.size()) == int(bounds.size());
            for (int k = 0; k < nu && okS; ++k) okS = u.qStart[size_t(k)] == u.qStart[size_t(k) + 1];
            for (size_t t = 0; t + 1 < R.tileUnit.size() && okS; ++t) {
                const int s0 = u.sStart[size_t(R.tileUnit[t])], s1 = u.sStart[size_t(R.tileUnit[t + 1])];
                const int r0 = R.unitStart[size_t(R.tileUnit[t])], r1 = R.unitStart[size_t(R.tileUnit[t + 1])];
                int rows = 0;
                okS = s0 >= 0 && s0 <= s1 && s1 <= h.nslices;
                for (int sl = s0; sl < s1 && okS; ++sl)
                    for (int lane = 0; lane < kSliceRows && okS; ++lane) {
                        const int w = h.row[size_t(sl) * kSliceRows + size_t(lane)];
                        if (w < 0) continue;
                        ++rows;
                        okS = w >= r0 && w < r1;
                        for (int k = 0; k < h.width[size_t(sl)] && okS; ++k) {
                            const int c = h.col[(size_t(h.off[size_t(sl)]) + size_t(k)) * kSliceRows + size_t(lane)];
                            okS = c >= 0 && c <= n;
                        }
                    }
                okS = okS && rows == r1 - r0;
            }
            check(okS, "red / black tile slices", id);
        }
    }
    std::vector<int> work(size_t(sL.levels())), rows(size_t(sL.levels()));
    for (int l = 0; l < sL.levels(); ++l) {
        work[size_t(l)] = sL.sStart[size_t(l) + 1] - sL.sStart[size_t(l)] + sL.qStart[size_t(l) + 1] - sL.qStart[size_t(l)];
        rows[size_t(l)] = start[size_t(l) + 1] - start[size_t(l)];
    }
    const std::vector<LevelRun> runs = levelRuns(work, 16, &rows, kRunRowsMax);
    bool okRuns = !runs.empty() && runs.front().l0 == 0 && runs.back().l1 == sL.levels();
    for (size_t k = 1; k < runs.size() && okRuns; ++k) okRuns = runs[k].l0 == runs[k - 1].l1;
    check(okRuns, "level runs cover the levels in order", id);
    // tiled schedules of the natural-order sweeps (several tile shapes): a permutation, units contiguous, and every
    // coupled pair j < i either in an earlier tile level or in the same unit on an earlier local level
    int off[3] = {0, 0, 0};
    std::vector<int> coord[3];
    const int nd = gridCoords(A, off, coord);
    int ntl = 0;
    static const int shapes[4][3] = {{1, 1, 1}, {4, 4, 4}, {8, 3, 2}, {64, 64, 64}};
    for (const int* dims : shapes) {
        TileSched T;
        tileOrder(A, coord, nd, dims, T);
        std::vector<int> unitOf(size_t(A.n), -1), tlOf(size_t(A.n), -1), lvOf(size_t(A.n), -1);
        bool okT = int(T.ord.size()) == A.n && T.unitStart.front() == 0 && T.unitStart.back() == A.n && T.tlStart.back() == T.units() &&
                   int(T.lvPtr.size()) == T.units() + 1;
        for (int t = 0; t < T.tileLevels() && okT; ++t)
            for (int u = T.tlStart[size_t(t)]; u < T.tlStart[size_t(t) + 1] && okT; ++u) {
                const int w0 = T.unitStart[size_t(u)], w1 = T.unitStart[size_t(u) + 1];
                const int l0 = T.lvPtr[size_t(u)], l1 = T.lvPtr[size_t(u) + 1];
                okT = w1 > w0 && w1 - w0 <= T.maxRows && l1 - l0 >= 2 && T.lvStart[size_t(l0)] == 0 && T.lvStart[size_t(l1) - 1] == w1 - w0 &&
                      l1 - l0 - 1 <= T.maxLevels;
                for (int l = l0; l + 1 < l1 && okT; ++l)
                    for (int w = w0 + T.lvStart[size_t(l)]; w < w0 + T.lvStart[size_t(l) + 1] && okT; ++w) {
                        const int i = T.ord[size_t(w)];
                        okT = i >= 0 && i < A.n && unitOf[size_t(i)] < 0;
                        if (!okT) break;
                        unitOf[size_t(i)] = u;
                        tlOf[size_t(i)] = t;
                        lvOf[size_t(i)] = l - l0;
                    }
            }
        for (int i = 0; i < A.n && okT; ++i)
            for (int p = A.rp[size_t(i)]; p < A.rp[size_t(i) + 1] && okT; ++p) {
                const int j = A.ci[size_t(p)];
                const int lo = std::min(i, j), hi = std::max(i, j);
                if (lo == hi) continue;
                okT = tlOf[size_t(lo)] < tlOf[size_t(hi)] || (unitOf[size_t(lo)] == unitOf[size_t(hi)] && lvOf[size_t(lo)] < lvOf[size_t(hi)]);
            }
        check(okT, "tiled sweep schedule", id);
        if (dims[0] == 4) ntl = T.tileLevels();
    }
    // the pressure-pressure coefficients as a scalar matrix: AMG set-up (strength, aggregation, Galerkin maps)
    HostCsr C;
    C.n = A.n;
    C.rp = A.rp;
    C.ci = A.ci;
    for (int p = 0; p < A.rp.back(); ++p) C.v.push_back(in.val[size_t(p) * A.nb * A.nb]);
    std::vector<int> fo, fs, bo, bs;
    csrLevels(C, true, fo, fs);
    csrLevels(C, false, bo, bs);
    check(int(fo.size()) == C.n && int(bo.size()) == C.n, "Gauss-Seidel level orders", id);
    std::vector<AmgLevelHost> H;
    amgSetup(C, H);
    bool okAmg = !H.empty() && H.front().A.n == C.n;
    for (size_t l = 0; l + 1 < H.size() && okAmg; ++l) {
        okAmg = H[l].nc == H[l + 1].A.n && H[l + 1].A.n < H[l].A.n && int(H[l].agg.size()) == H[l].A.n;
        for (int a : H[l].agg) okAmg = okAmg && a >= 0 && a < H[l].nc;
        okAmg = okAmg && int(H[l].cPtr.size()) == H[l + 1].A.rp.back() + 1;
    }
    check(okAmg, "AMG hierarchy", id);
    // row slices of a coarse level as its Jacobi steps read them (far rows long): every row once, every entry once in
    // its row's order, columns inside the vector, 16-bit offsets equal to the columns
    for (size_t l = 1; l < H.size(); ++l) {
        const HostCsr& M = H[l].A;
        SliceHost sh;
        rowSliceLayout(M.n, M.rp, M.ci, sh, true);
        std::vector<int> seen(size_t(M.n), 0);
        std::vector<double> xv(size_t(M.n)), ys(size_t(M.n), 0.0), yc(size_t(M.n), 0.0);
        for (int i = 0; i < M.n; ++i) xv[size_t(i)] = 1.0 + 0.001 * (i % 97);
        bool okS = true;
        for (int sl = 0; sl < sh.nslices && okS; ++sl)
            for (int lane = 0; lane < kSliceRows && okS; ++lane) {
                const int i = sh.row[size_t(sl) * kSliceRows + size_t(lane)];
                if (i < 0) continue;
                okS = i < M.n && seen[size_t(i)]++ == 0;
                double acc = 0;
                int p = M.rp[size_t(i)];
                for (int k = 0; k < sh.width[size_t(sl)] && okS; ++k) {
                    const size_t at = (size_t(sh.off[size_t(sl)]) + size_t(k)) * kSliceRows + size_t(lane);
                    const int c = sh.col[at];
                    okS = c >= 0 && c < M.n && (sh.dcol.empty() || i + sh.dcol[at] == c);
                    if (sh.src[at] < 0) continue;
                    okS = okS && sh.src[at] == p && c == M.ci[size_t(p)];
                    acc += M.v[size_t(sh.src[at])] * xv[size_t(c)];
                    ++p;
                }
                okS = okS && p == M.rp[size_t(i) + 1];
                ys[size_t(i)] = acc;
            }
        for (size_t q = 0; q < sh.lrow.size() && okS; ++q) {
            const int i = sh.lrow[q];
            okS = i >= 0 && i < M.n && seen[size_t(i)]++ == 0 && sh.lptr[q + 1] - sh.lptr[q] == M.rp[size_t(i) + 1] - M.rp[size_t(i)];
            double acc = 0;
            for (int u = sh.lptr[q]; u < sh.lptr[q + 1] && okS; ++u) acc += M.v[size_t(sh.lsrc[size_t(u)])] * xv[size_t(sh.lcol[size_t(u)])];
            ys[size_t(i)] = acc;
        }
        for (int i = 0; i < M.n && okS; ++i) {
            okS = seen[size_t(i)] == 1;
            for (int p = M.rp[size_t(i)]; p < M.rp[size_t(i) + 1]; ++p) yc[size_t(i)] += M.v[size_t(p)] * xv[size_t(M.ci[size_t(p)])];
        }
        check(okS && ys == yc, "coarse-level row slices", id);
    }
    // SpMV slices with 1, 2 or 4 lanes per row: each bulk row on `lanes` aligned lanes of one slice, entry e of the row
    // on lane part e % lanes at slot e / lanes, padding a zero block at the row's own column, long rows whole
    for (int lanes : {1, 2, 4}) {
        SliceHost sh;
        laneSliceLayout(A.n, A.rp, A.ci, lanes, sh);
        std::vector<int> seen(size_t(A.n), 0);
        bool okW = sh.nslices == int(sh.width.size()) && sh.row.size() == size_t(sh.nslices) * kSliceRows;
        for (int sl = 0; sl < sh.nslices && okW; ++sl)
            for (int g = 0; g < kSliceRows / lanes && okW; ++g) {
                const int i = sh.row[size_t(sl) * kSliceRows + size_t(g * lanes)];
                for (int part = 0; part < lanes && okW; ++part) {
                    const int lane = g * lanes + part;
                    okW = sh.row[size_t(sl) * kSliceRows + size_t(lane)] == i;
                    for (int k = 0; k < sh.width[size_t(sl)] && okW; ++k) {
                        const size_t at = (size_t(sh.off[size_t(sl)]) + size_t(k)) * kSliceRows + size_t(lane);
                        const int c = sh.col[at];
                        okW = c >= 0 && c < A.n && (sh.dcol.empty() || (i < 0 ? -1 : i) + sh.dcol[at] == c);
                        if (i < 0) { okW = okW && sh.src[at] < 0; continue; }
                        const int e = k * lanes + part, len = A.rp[size_t(i) + 1] - A.rp[size_t(i)];
                        okW = okW && (e < len ? sh.src[at] == A.rp[size_t(i)] + e && c == A.ci[size_t(A.rp[size_t(i)] + e)] : sh.src[at] < 0 && c == i);
                    }
                }
                if (i >= 0 && okW) {
                    okW = seen[size_t(i)]++ == 0 && A.rp[size_t(i) + 1] - A.rp[size_t(i)] <= sh.width[size_t(sl)] * lanes;
                }
            }
        for (size_t q = 0; q < sh.lrow.size() && okW; ++q) {
            const int i = sh.lrow[q];
            okW = i >= 0 && i < A.n && seen[size_t(i)]++ == 0 && sh.lptr[q + 1] - sh.lptr[q] == A.rp[size_t(i) + 1] - A.rp[size_t(i)];
            for (int u = sh.lptr[q]; u < sh.lptr[q + 1] && okW; ++u)
                okW = sh.lsrc[size_t(u)] == A.rp[size_t(i)] + (u - sh.lptr[q]) && sh.lcol[size_t(u)] == A.ci[size_t(sh.lsrc[size_t(u)])];
        }
        for (int i = 0; i < A.n && okW; ++i) okW = seen[size_t(i)] == 1;
        check(okW, lanes == 1 ? "SpMV slices, 1 lane per row" : lanes == 2 ? "SpMV slices, 2 lanes per row" : "SpMV slices, 4 lanes per row", id);
    }
    // the well Schur plan of the Jacobian the system was reduced from: the same pattern and values
    int fill = 0;
    if (in.jac) {
        const double* jv = in.jval.data();
        wellplan::Plan W;
        std::string why = wellplan::split(in.rows, in.jrp.data(), in.jci.data(), in.na, in.nv, in.nw, [jv](int p) { return jv[p] != 0.0; }, W);
        check(why.empty(), "well plan split", id);
        wellplan::pattern(in.rows, in.jrp.data(), in.jci.data(), W);
        check(W.sPtr == in.rp && W.sCol == in.ci, "well plan pattern equals the reduced system's", id);
        std::vector<double> dInv;
        why = wellplan::wellInverses(W, jv, dInv);
        check(why.empty(), "well inverses", id);
        std::vector<double> s(W.sCol.size() * size_t(in.nv) * in.nv);
        wellplan::reduceValues(in.jrp.data(), in.jci.data(), jv, W, dInv, s.data());
        check(s == in.val, "well plan values equal the reduced system's", id);
        wellplan::ValueMap M;
        check(wellplan::valueMap(in.jrp.data(), in.jci.data(), W, M).empty() && M.cellSource.size() == s.size(), "well value map", id);
        std::vector<double> b(size_t(in.rows)), bc(size_t(W.nc)), x(size_t(in.rows));
        for (int r = 0; r < in.rows; ++r) b[size_t(r)] = std::cos(0.1 * r);
        wellplan::reduceRhs(W, jv, dInv, b.data(), bc.data());
        wellplan::recover(W, in.jci.data(), jv, dInv, b.data(), bc.data(), x.data());
        fill = W.fill();
    }
    std::printf("system %d: n %d nb %d levels %d colours %d L slices %d AMG levels %zu well fill %d grid offsets %d tile levels (4^3) %d "
                "red/black tile orders %d\n", id, A.n, A.nb, sL.levels(), ncol, hL.nslices, H.size(), fill, nd, ntl, rbOrders);
}
}  // namespace

int main(int argc, char** argv) {
    if (argc < 2 || !(g_in = std::fopen(argv[1], "rb"))) {
        std::fprintf(stderr, "usage: host_check <systems file>\n");
        return 2;
    }
    int count = 0;
    if (!rd(&count, 4)) return 2;
    for (int k = 0; k < count; ++k) {
        Input in;
        int hdr[2];
        if (!rd(hdr, 8)) return 2;
        in.n = hdr[0];
        in.nb = hdr[1];
        in.rp.resize(size_t(in.n) + 1);
        if (!rd(in.rp.data(), 4 * in.rp.size())) return 2;
        in.ci.resize(size_t(in.rp.back()));
        in.val.resize(in.ci.size() * size_t(in.nb) * in.nb);
        int has = 0;
        if (!rd(in.ci.data(), 4 * in.ci.size()) || !rd(in.val.data(), 8 * in.val.size()) || !rd(&has, 4)) return 2;
        if (has) {
            int jh[4];
            if (!rd(jh, 16)) return 2;
            in.jac = true;
            in.rows = jh[0];
            in.na = jh[1];
            in.nv = jh[2];
            in.nw = jh[3];
            in.jrp.resize(size_t(in.rows) + 1);
            if (!rd(in.jrp.data(), 4 * in.jrp.size())) return 2;
            in.jci.resize(size_t(in.jrp.back()));
            in.jval.resize(in.jci.size());
            if (!rd(in.jci.data(), 4 * in.jci.size()) || !rd(in.jval.data(), 8 * in.jval.size())) return 2;
        }
        run(in, k);
    }
    std::fclose(g_in);
    std::printf("%d failed checks\n", fails);
    return fails ? 1 : 0;
}
