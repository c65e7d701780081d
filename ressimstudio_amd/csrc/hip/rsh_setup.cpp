// Host-side symbolic set-up of the device solver: dependency levels, the multicolour ordering, the row-sliced
// layouts streamed by the kernels, and the aggregation AMG hierarchy.  These steps are sequential and
// order-dependent in the reference (greedy aggregation amg.cpp:165-194, greedy colouring bsr.cpp:612-622), so they
// run once per sparsity pattern on the host and their results stay resident on the device; the numeric work
// (factorisation, Galerkin products, sweeps) runs on the GPU.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <unordered_map>

#include "rsh_internal.hpp"

namespace rsh {

void levelOrder(const HostBsr& A, std::vector<int>& ord, std::vector<int>& start) {
    // level(i) = 1 + max level(j) over j < i coupled to i in either direction (the symmetrised lower graph): a
    // topological order of the L dependencies that also keeps every coupled pair in its relative order, so the
    // reversed levels are valid levels of the U sweep
    const int n = A.n;
    std::vector<int> lev(static_cast<size_t>(n), 0);
    for (int j = 0; j < n; ++j) {
        int l = lev[size_t(j)];
        for (int p = A.rp[size_t(j)]; p < A.rp[size_t(j) + 1]; ++p) {
            const int c = A.ci[size_t(p)];
            if (c < j) l = std::max(l, lev[size_t(c)] + 1);
        }
        lev[size_t(j)] = l;
        // push j's level to its later neighbours (transposed lower entries)
        for (int p = A.rp[size_t(j)]; p < A.rp[size_t(j) + 1]; ++p) {
            const int c = A.ci[size_t(p)];
            if (c > j) lev[size_t(c)] = std::max(lev[size_t(c)], l + 1);
        }
    }
    int nl = 0;
    for (int i = 0; i < n; ++i) nl = std::max(nl, lev[size_t(i)] + 1);
    start.assign(size_t(nl) + 1, 0);
    for (int i = 0; i < n; ++i) start[size_t(lev[size_t(i)]) + 1]++;
    for (int l = 0; l < nl; ++l) start[size_t(l) + 1] += start[size_t(l)];
    ord.assign(size_t(n), 0);
    std::vector<int> fill(start.begin(), start.end() - 1);
    for (int i = 0; i < n; ++i) ord[size_t(fill[size_t(lev[size_t(i)])]++)] = i;
}

void csrLevels(const HostCsr& A, bool forward, std::vector<int>& ord, std::vector<int>& start) {
    const int n = A.n;
    std::vector<int> lev(static_cast<size_t>(n), 0);
    int nl = 0;
    for (int t = 0; t < n; ++t) {
        const int i = forward ? t : n - 1 - t;
        int l = 0;
        for (int p = A.rp[size_t(i)]; p < A.rp[size_t(i) + 1]; ++p) {
            const int j = A.ci[size_t(p)];
            if (forward ? j < i : j > i) l = std::max(l, lev[size_t(j)] + 1);
        }
        lev[size_t(i)] = l;
        nl = std::max(nl, l + 1);
    }
    start.assign(size_t(nl) + 1, 0);
    for (int i = 0; i < n; ++i) start[size_t(lev[size_t(i)]) + 1]++;
    for (int l = 0; l < nl; ++l) start[size_t(l) + 1] += start[size_t(l)];
    ord.assign(size_t(n), 0);
    std::vector<int> fill(start.begin(), start.end() - 1);
    for (int i = 0; i < n; ++i) ord[size_t(fill[size_t(lev[size_t(i)])]++)] = i;
}

int colourGraph(const HostBsr& A, std::vector<int>& colour, std::vector<int>& perm, std::vector<int>& start) {
    // the reference's greedy: symmetrised adjacency, cells visited by decreasing degree (stable), smallest colour
    // not used by a neighbour (bsr.cpp:599-630)
    const int n = A.n;
    std::vector<int> deg(static_cast<size_t>(n), 0);
    std::vector<std::vector<int>> adj(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i)
        for (int p = A.rp[size_t(i)]; p < A.rp[size_t(i) + 1]; ++p) {
            const int j = A.ci[size_t(p)];
            if (j != i) { adj[size_t(i)].push_back(j); adj[size_t(j)].push_back(i); }
        }
    for (auto& a : adj) { std::sort(a.begin(), a.end()); a.erase(std::unique(a.begin(), a.end()), a.end()); }
    std::vector<int> order(static_cast<size_t>(n));
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return adj[size_t(x)].size() > adj[size_t(y)].size(); });
    colour.assign(size_t(n), -1);
    int colours = 0;
    std::vector<char> used;
    for (int i : order) {
        used.assign(adj[size_t(i)].size() + 2, 0);
        for (int j : adj[size_t(i)]) {
            const int c = colour[size_t(j)];
            if (c >= 0 && c < int(used.size())) used[size_t(c)] = 1;
        }
        int c = 0;
        while (used[size_t(c)]) ++c;
        colour[size_t(i)] = c;
        colours = std::max(colours, c + 1);
    }
    start.assign(size_t(colours) + 1, 0);
    for (int i = 0; i < n; ++i) start[size_t(colour[size_t(i)]) + 1]++;
    for (int c = 0; c < colours; ++c) start[size_t(c) + 1] += start[size_t(c)];
    perm.assign(size_t(n), 0);
    std::vector<int> fill(start.begin(), start.end() - 1);
    for (int i = 0; i < n; ++i) perm[size_t(fill[size_t(colour[size_t(i)])]++)] = i;
    return colours;
}

void permutePattern(const HostBsr& A, const std::vector<int>& perm, HostBsr& P, std::vector<int>& blockMap) {
    const int n = A.n;
    std::vector<int> iperm(static_cast<size_t>(n));
    for (int k = 0; k < n; ++k) iperm[size_t(perm[size_t(k)])] = k;
    P = HostBsr{};
    P.n = n; P.nb = A.nb;
    P.rp.assign(size_t(n) + 1, 0);
    P.ci.reserve(A.ci.size());
    blockMap.clear();
    blockMap.reserve(A.ci.size());
    std::vector<std::pair<int, int>> row;
    for (int k = 0; k < n; ++k) {
        const int i = perm[size_t(k)];
        row.clear();
        for (int p = A.rp[size_t(i)]; p < A.rp[size_t(i) + 1]; ++p) row.push_back({iperm[size_t(A.ci[size_t(p)])], p});
        std::sort(row.begin(), row.end());
        for (auto& e : row) { P.ci.push_back(e.first); blockMap.push_back(e.second); }
        P.rp[size_t(k) + 1] = int(P.ci.size());
    }
    P.diag.assign(size_t(n), -1);
    for (int k = 0; k < n; ++k)
        for (int p = P.rp[size_t(k)]; p < P.rp[size_t(k) + 1]; ++p) if (P.ci[size_t(p)] == k) { P.diag[size_t(k)] = p; break; }
}

std::vector<LevelRun> levelRuns(const std::vector<int>& workPerLevel, int smallMax, const std::vector<int>* rowsPerLevel, int maxRunRows) {
    std::vector<LevelRun> runs;
    const int nl = int(workPerLevel.size());
    int l = 0;
    while (l < nl) {
        if (workPerLevel[size_t(l)] > smallMax || (rowsPerLevel && (*rowsPerLevel)[size_t(l)] > maxRunRows)) {
            runs.push_back({l, l + 1, false});
            ++l;
            continue;
        }
        int e = l, rows = 0;
        while (e < nl && workPerLevel[size_t(e)] <= smallMax) {
            if (rowsPerLevel) {
                if (rows + (*rowsPerLevel)[size_t(e)] > maxRunRows) break;
                rows += (*rowsPerLevel)[size_t(e)];
            }
            ++e;
        }
        // a single small level is cheaper as a plain launch with its own grid
        runs.push_back({l, e, e - l > 1});
        l = e;
    }
    return runs;
}

int rowCap(const std::vector<int>& lengths) {
    if (lengths.empty()) return 1;
    std::vector<int> s(lengths);
    std::sort(s.begin(), s.end());
    const int mx = s.back();
    const int p = s[size_t(std::min<double>(double(s.size() - 1), std::floor(0.998 * double(s.size()))))];
    const int cap = std::max(p, 1);
    return mx <= cap + cap / 2 + 1 ? mx : cap;   // no long rows unless some rows are well beyond the bulk
}

void buildSlices(const RowEntries& E, const std::vector<int>& levelStart, int cap, SliceHost& S, LevelSched& L, const std::vector<char>* longRow,
                 const std::vector<char>* skipRow) {
    const int nl = int(levelStart.size()) - 1;
    const int n = int(E.ptr.size()) - 1;
    const bool withDiag = !E.dsrc.empty();
    S = SliceHost{};
    L.sStart.assign(size_t(nl) + 1, 0);
    L.qStart.assign(size_t(nl) + 1, 0);
    S.lptr.push_back(0);
    std::vector<int> shortRows;
    for (int l = 0; l < nl; ++l) {
        shortRows.clear();
        for (int w = levelStart[size_t(l)]; w < levelStart[size_t(l) + 1]; ++w) {
            if (skipRow && (*skipRow)[size_t(w)]) continue;
            const int len = E.ptr[size_t(w) + 1] - E.ptr[size_t(w)];
            if (len > cap || (longRow && (*longRow)[size_t(w)])) {
                S.lrow.push_back(w);
                S.lorig.push_back(E.orig[size_t(w)]);
                for (int p = E.ptr[size_t(w)]; p < E.ptr[size_t(w) + 1]; ++p) { S.lcol.push_back(E.col[size_t(p)]); S.lsrc.push_back(E.src[size_t(p)]); }
                S.lptr.push_back(int(S.lcol.size()));
                if (withDiag) S.ldsrc.push_back(E.dsrc[size_t(w)]);
            } else shortRows.push_back(w);
        }
        for (size_t c0 = 0; c0 < shortRows.size(); c0 += kSliceRows) {
            const size_t c1 = std::min(shortRows.size(), c0 + kSliceRows);
            int width = 0;
            for (size_t t = c0; t < c1; ++t) { const int w = shortRows[t]; width = std::max(width, E.ptr[size_t(w) + 1] - E.ptr[size_t(w)]); }
            S.width.push_back(width);
            S.off.push_back(int(S.nslots));
            const size_t base = S.row.size();
            S.row.resize(base + kSliceRows, -1);
            S.orig.resize(base + kSliceRows, -1);
            if (withDiag) S.dsrc.resize(base + kSliceRows, -1);
            const size_t cbase = S.col.size();
            S.col.resize(cbase + size_t(width) * kSliceRows, n);     // padding: the dummy row n (zero in the sweeps)
            S.src.resize(cbase + size_t(width) * kSliceRows, -1);
            for (size_t t = c0; t < c1; ++t) {
                const int lane = int(t - c0), w = shortRows[t];
                S.row[base + size_t(lane)] = w;
                S.orig[base + size_t(lane)] = E.orig[size_t(w)];
                if (withDiag) S.dsrc[base + size_t(lane)] = E.dsrc[size_t(w)];
                const int p0 = E.ptr[size_t(w)], len = E.ptr[size_t(w) + 1] - p0;
                for (int k = 0; k < width; ++k) {
                    const size_t at = cbase + size_t(k) * kSliceRows + size_t(lane);
                    if (k < len) { S.col[at] = E.col[size_t(p0 + k)]; S.src[at] = E.src[size_t(p0 + k)]; }
                }
            }
            S.nslots += width;
            S.nslices++;
        }
        L.sStart[size_t(l) + 1] = S.nslices;
        L.qStart[size_t(l) + 1] = int(S.lrow.size());
    }
}

void rowSliceLayout(int n, const std::vector<int>& rp, const std::vector<int>& ci, SliceHost& h, bool coarse) {
    std::vector<int> len(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) len[size_t(i)] = rp[size_t(i) + 1] - rp[size_t(i)];
    int cap = rowCap(len);
    std::vector<char> longRow;
    if (coarse) {
        // rows beyond the bulk (a quarter above the 99th percentile: the aggregates of perforated cells) and rows
        // reaching further than 16-bit offsets are long rows
        std::vector<int> srt(len);
        std::sort(srt.begin(), srt.end());
        const int p99 = n ? srt[size_t(std::floor(0.99 * double(n - 1)))] : 1;
        cap = std::max(4, p99 + (p99 + 3) / 4);
        longRow.assign(size_t(n), 0);
        for (int i = 0; i < n; ++i)
            for (int p = rp[size_t(i)]; p < rp[size_t(i) + 1]; ++p)
                if (std::abs((long long)ci[size_t(p)] - i) > 32767) { longRow[size_t(i)] = 1; break; }
    }
    RowEntries E;
    E.ptr.assign(rp.begin(), rp.end());
    E.col = ci;
    E.src.resize(ci.size());
    for (size_t p = 0; p < ci.size(); ++p) E.src[p] = int(p);
    E.orig.resize(size_t(n));
    for (int i = 0; i < n; ++i) E.orig[size_t(i)] = i;
    LevelSched one;
    buildSlices(E, {0, n}, cap, h, one, longRow.empty() ? nullptr : &longRow);
    // padding points at the lane's own row (idle lanes at row 0): every gather stays inside the vector
    for (int sl = 0; sl < h.nslices; ++sl)
        for (int k = 0; k < h.width[size_t(sl)]; ++k)
            for (int lane = 0; lane < kSliceRows; ++lane) {
                const size_t at = (size_t(h.off[size_t(sl)]) + size_t(k)) * kSliceRows + size_t(lane);
                if (h.src[at] < 0) {
                    const int rr = h.row[size_t(sl) * kSliceRows + size_t(lane)];
                    h.col[at] = rr >= 0 ? rr : 0;
                }
            }
    // the columns relative to the lane's row fit 16 bits when no slice row reaches further than 32767 rows (the
    // 7-point stencil of grids up to nx * ny = 32767 cells per layer; far couplings are long rows)
    std::vector<short> d(h.col.size());
    for (int sl = 0; sl < h.nslices; ++sl)
        for (int k = 0; k < h.width[size_t(sl)]; ++k)
            for (int lane = 0; lane < kSliceRows; ++lane) {
                const size_t at = (size_t(h.off[size_t(sl)]) + size_t(k)) * kSliceRows + size_t(lane);
                const int rr = h.row[size_t(sl) * kSliceRows + size_t(lane)];
                const long long delta = (long long)h.col[at] - (rr >= 0 ? rr : -1);
                if (delta < -32768 || delta > 32767) return;
                d[at] = short(delta);
            }
    h.dcol = std::move(d);
}

// Row slices with `lanes` lanes per row (1, 2 or 4): a slice holds 64 / lanes rows in natural order, lane
// j * lanes + p of the slice takes entries p, p + lanes, p + 2 lanes, ... of the slice's j-th row (slot k holds entry
// k * lanes + p), so the slice is ceil(longest row / lanes) slots wide; the lanes' partial sums are added afterwards.
// Rows longer than the bulk are long rows, as in rowSliceLayout; padding points at the lane's own row (idle lanes at
// row 0) with source -1 (a zero block).
void laneSliceLayout(int n, const std::vector<int>& rp, const std::vector<int>& ci, int lanes, SliceHost& h) {
    if (lanes != 1 && lanes != 2 && lanes != 4) throw Error(-1, "slice layout: lanes per row must be 1, 2 or 4");
    std::vector<int> len(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) len[size_t(i)] = rp[size_t(i) + 1] - rp[size_t(i)];
    const int cap = rowCap(len);
    const int per = kSliceRows / lanes;
    h = SliceHost{};
    h.lptr.push_back(0);
    std::vector<int> bulk;
    bulk.reserve(size_t(n));
    for (int i = 0; i < n; ++i) {
        if (len[size_t(i)] > cap) {
            h.lrow.push_back(i);
            h.lorig.push_back(i);
            for (int q = rp[size_t(i)]; q < rp[size_t(i) + 1]; ++q) { h.lcol.push_back(ci[size_t(q)]); h.lsrc.push_back(q); }
            h.lptr.push_back(int(h.lcol.size()));
        } else bulk.push_back(i);
    }
    bool rel16 = true;
    std::vector<short> d;
    for (size_t c0 = 0; c0 < bulk.size(); c0 += size_t(per)) {
        const size_t c1 = std::min(bulk.size(), c0 + size_t(per));
        int width = 0;
        for (size_t t = c0; t < c1; ++t) width = std::max(width, (len[size_t(bulk[t])] + lanes - 1) / lanes);
        h.width.push_back(width);
        h.off.push_back(int(h.nslots));
        const size_t base = h.row.size();
        h.row.resize(base + kSliceRows, -1);
        h.orig.resize(base + kSliceRows, -1);
        const size_t cbase = h.col.size();
        h.col.resize(cbase + size_t(width) * kSliceRows, 0);
        h.src.resize(cbase + size_t(width) * kSliceRows, -1);
        d.resize(cbase + size_t(width) * kSliceRows, short(1));    // idle lanes: column 0 relative to row -1
        for (size_t t = c0; t < c1; ++t) {
            const int w = bulk[t], p0 = rp[size_t(w)];
            for (int part = 0; part < lanes; ++part) {
                const int lane = int(t - c0) * lanes + part;
                h.row[base + size_t(lane)] = w;
                h.orig[base + size_t(lane)] = w;
                for (int k = 0; k < width; ++k) {
                    const size_t at = cbase + size_t(k) * kSliceRows + size_t(lane);
                    const int e = k * lanes + part;
                    if (e < len[size_t(w)]) { h.col[at] = ci[size_t(p0 + e)]; h.src[at] = p0 + e; }
                    else h.col[at] = w;
                    const long long delta = (long long)h.col[at] - w;
                    if (delta < -32768 || delta > 32767) rel16 = false;
                    else d[at] = short(delta);
                }
            }
        }
        h.nslots += width;
        h.nslices++;
    }
    if (rel16) h.dcol = std::move(d);
}

// ------------------------------------------------------------------ aggregation AMG set-up (amg.cpp:137-252)
namespace {
void csrDiag(HostCsr& A) {
    A.diag.assign(size_t(A.n), -1);
    for (int i = 0; i < A.n; ++i)
        for (int p = A.rp[size_t(i)]; p < A.rp[size_t(i) + 1]; ++p) if (A.ci[size_t(p)] == i) { A.diag[size_t(i)] = p; break; }
}

// strength of connection: |a_ij| >= 0.08 sqrt|a_ii a_jj| or >= 0.25 max_k |a_ik| (amg.cpp:150-164)
void strength(const HostCsr& A, std::vector<int>& sp, std::vector<int>& sc) {
    const int n = A.n;
    sp.assign(size_t(n) + 1, 0);
    sc.clear();
    sc.reserve(A.ci.size());
    for (int i = 0; i < n; ++i) {
        double rowMax = 0;
        for (int p = A.rp[size_t(i)]; p < A.rp[size_t(i) + 1]; ++p) if (A.ci[size_t(p)] != i) rowMax = std::max(rowMax, std::fabs(A.v[size_t(p)]));
        const double di = A.v[size_t(A.diag[size_t(i)])];
        for (int p = A.rp[size_t(i)]; p < A.rp[size_t(i) + 1]; ++p) {
            const int j = A.ci[size_t(p)];
            if (j == i) continue;
            const double a = std::fabs(A.v[size_t(p)]);
            if (a >= 0.08 * std::sqrt(std::fabs(di * A.v[size_t(A.diag[size_t(j)])])) || a >= 0.25 * rowMax) sc.push_back(j);
        }
        sp[size_t(i) + 1] = int(sc.size());
    }
}

// greedy aggregation in three passes (amg.cpp:165-193); returns the number of aggregates
int aggregate(const HostCsr& A, const std::vector<int>& sp, const std::vector<int>& sc, std::vector<int>& agg) {
    const int n = A.n;
    agg.assign(size_t(n), -1);
    int na = 0;
    for (int i = 0; i < n; ++i) {   // roots whose strong neighbourhood is still free
        if (agg[size_t(i)] >= 0 || sp[size_t(i)] == sp[size_t(i) + 1]) continue;
        bool freeNb = true;
        for (int p = sp[size_t(i)]; p < sp[size_t(i) + 1] && freeNb; ++p) freeNb = agg[size_t(sc[size_t(p)])] < 0;
        if (!freeNb) continue;
        agg[size_t(i)] = na;
        for (int p = sp[size_t(i)]; p < sp[size_t(i) + 1]; ++p) agg[size_t(sc[size_t(p)])] = na;
        ++na;
    }
    const std::vector<int> first = agg;
    for (int i = 0; i < n; ++i) {   // join the neighbouring first-pass aggregate with the largest |a_ij|
        if (agg[size_t(i)] >= 0 || sp[size_t(i)] == sp[size_t(i) + 1]) continue;
        double best = 0;
        int ba = -1;
        for (int p = A.rp[size_t(i)]; p < A.rp[size_t(i) + 1]; ++p) {
            const int j = A.ci[size_t(p)];
            if (j != i && first[size_t(j)] >= 0 && std::fabs(A.v[size_t(p)]) > best) { best = std::fabs(A.v[size_t(p)]); ba = first[size_t(j)]; }
        }
        if (ba >= 0) agg[size_t(i)] = ba;
    }
    for (int i = 0; i < n; ++i) {   // leftovers start new aggregates with their free strong neighbours
        if (agg[size_t(i)] >= 0) continue;
        agg[size_t(i)] = na;
        for (int p = sp[size_t(i)]; p < sp[size_t(i) + 1]; ++p) if (agg[size_t(sc[size_t(p)])] < 0) agg[size_t(sc[size_t(p)])] = na;
        ++na;
    }
    return na;
}

// Galerkin operator of plain aggregation with the summation order of R (A P) formed as two sparse products
// (amg.cpp:36-68, 232): entries of A P summed per fine row in column order, then per aggregate over its fine rows
// in ascending order; rows sorted by column.
void galerkin(AmgLevelHost& L, HostCsr& C) {
    const HostCsr& A = L.A;
    const int n = A.n, nc = L.nc;
    // A P: per fine row, the distinct aggregates of its columns, sorted; entry lists of source positions
    std::vector<int> apRp(size_t(n) + 1, 0), apCol;
    L.apPtr.assign(1, 0);
    L.apSrc.clear();
    std::vector<int> mark(size_t(nc), -1), slot(size_t(nc), -1);
    std::vector<std::vector<int>> lists;
    std::vector<int> cols;
    for (int i = 0; i < n; ++i) {
        cols.clear();
        for (int p = A.rp[size_t(i)]; p < A.rp[size_t(i) + 1]; ++p) {
            const int b = L.agg[size_t(A.ci[size_t(p)])];
            if (mark[size_t(b)] != i) { mark[size_t(b)] = i; slot[size_t(b)] = int(cols.size()); cols.push_back(b); if (lists.size() < cols.size()) lists.emplace_back(); lists[cols.size() - 1].clear(); }
            lists[size_t(slot[size_t(b)])].push_back(p);
        }
        std::vector<int> idx(cols.size());
        std::iota(idx.begin(), idx.end(), 0);
        std::sort(idx.begin(), idx.end(), [&](int x, int y) { return cols[size_t(x)] < cols[size_t(y)]; });
        for (int t : idx) {
            apCol.push_back(cols[size_t(t)]);
            L.apSrc.insert(L.apSrc.end(), lists[size_t(t)].begin(), lists[size_t(t)].end());
            L.apPtr.push_back(int(L.apSrc.size()));
        }
        apRp[size_t(i) + 1] = int(apCol.size());
    }
    std::vector<double> ap(apCol.size());
    for (size_t t = 0; t < apCol.size(); ++t) {
        double s = A.v[size_t(L.apSrc[size_t(L.apPtr[t])])];
        for (int u = L.apPtr[t] + 1; u < L.apPtr[t + 1]; ++u) s += A.v[size_t(L.apSrc[size_t(u)])];
        ap[t] = s;
    }
    // R (A P): aggregates' fine rows ascending
    std::vector<int> rrp(size_t(nc) + 1, 0), rci(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) rrp[size_t(L.agg[size_t(i)]) + 1]++;
    for (int a = 0; a < nc; ++a) rrp[size_t(a) + 1] += rrp[size_t(a)];
    {
        std::vector<int> fill(rrp.begin(), rrp.end() - 1);
        for (int i = 0; i < n; ++i) rci[size_t(fill[size_t(L.agg[size_t(i)])]++)] = i;
    }
    C = HostCsr{};
    C.n = nc;
    C.rp.assign(size_t(nc) + 1, 0);
    L.cPtr.assign(1, 0);
    L.cSrc.clear();
    std::fill(mark.begin(), mark.end(), -1);
    for (int a = 0; a < nc; ++a) {
        cols.clear();
        for (int q = rrp[size_t(a)]; q < rrp[size_t(a) + 1]; ++q) {
            const int i = rci[size_t(q)];
            for (int t = apRp[size_t(i)]; t < apRp[size_t(i) + 1]; ++t) {
                const int b = apCol[size_t(t)];
                if (mark[size_t(b)] != a) { mark[size_t(b)] = a; slot[size_t(b)] = int(cols.size()); cols.push_back(b); if (lists.size() < cols.size()) lists.emplace_back(); lists[cols.size() - 1].clear(); }
                lists[size_t(slot[size_t(b)])].push_back(t);
            }
        }
        std::vector<int> idx(cols.size());
        std::iota(idx.begin(), idx.end(), 0);
        std::sort(idx.begin(), idx.end(), [&](int x, int y) { return cols[size_t(x)] < cols[size_t(y)]; });
        for (int t : idx) {
            C.ci.push_back(cols[size_t(t)]);
            L.cSrc.insert(L.cSrc.end(), lists[size_t(t)].begin(), lists[size_t(t)].end());
            L.cPtr.push_back(int(L.cSrc.size()));
            double s = ap[size_t(lists[size_t(t)][0])];
            for (size_t u = 1; u < lists[size_t(t)].size(); ++u) s += ap[size_t(lists[size_t(t)][u])];
            C.v.push_back(s);
        }
        C.rp[size_t(a) + 1] = int(C.ci.size());
    }
    csrDiag(C);
    for (int a = 0; a < nc; ++a) {
        if (C.diag[size_t(a)] < 0) throw Error(-6, "coarse operator without a diagonal entry");
        if (std::fabs(C.v[size_t(C.diag[size_t(a)])]) < 1e-300) C.v[size_t(C.diag[size_t(a)])] = 1.0;
    }
}
}  // namespace

void amgSetup(HostCsr A0, std::vector<AmgLevelHost>& H) {
    H.clear();
    csrDiag(A0);
    for (int i = 0; i < A0.n; ++i) {
        if (A0.diag[size_t(i)] < 0) throw Error(-6, "pressure matrix row without a diagonal entry");
        if (std::fabs(A0.v[size_t(A0.diag[size_t(i)])]) < 1e-300) A0.v[size_t(A0.diag[size_t(i)])] = 1.0;
    }
    H.emplace_back();
    H[0].A = std::move(A0);
    std::vector<int> sp, sc, agg;
    while (true) {
        AmgLevelHost& L = H.back();
        const int n = L.A.n;
        if (n <= 400 || H.size() >= 14) break;
        strength(L.A, sp, sc);
        const int na = aggregate(L.A, sp, sc, agg);
        if (na >= int(0.85 * n) || na < 1) break;
        L.agg = agg;
        L.nc = na;
        AmgLevelHost next;
        galerkin(L, next.A);
        H.push_back(std::move(next));
    }
}

// ------------------------------------------------------------------ tiled schedule of the natural-order sweeps
int gridCoords(const HostBsr& A, int off[3], std::vector<int> coord[3]) {
    const int n = A.n;
    // the most frequent offsets i - j of the lower entries; those carried by at least an eighth of the rows are the
    // grid's steps (1, nx, nx ny for a natural-ordered grid)
    std::unordered_map<int, long long> cnt;
    for (int i = 0; i < n; ++i)
        for (int p = A.rp[size_t(i)]; p < A.diag[size_t(i)]; ++p) cnt[i - A.ci[size_t(p)]]++;
    std::vector<std::pair<long long, int>> v;
    for (const auto& kv : cnt)
        if (kv.second * 8 >= n) v.push_back({kv.second, kv.first});
    std::sort(v.begin(), v.end(), [](const std::pair<long long, int>& a, const std::pair<long long, int>& b) {
        return a.first != b.first ? a.first > b.first : a.second < b.second;
    });
    const int nd = int(std::min<size_t>(3, v.size()));
    for (int a = 0; a < nd; ++a) off[a] = v[size_t(a)].second;
    std::sort(off, off + nd);
    for (int a = 0; a < 3; ++a) {
        coord[a].assign(size_t(n), 0);
        if (a >= nd) continue;
        const int d = off[a];
        for (int i = d; i < n; ++i) {
            // columns are sorted: look for i - d among the lower entries
            const int* b = A.ci.data() + A.rp[size_t(i)];
            const int* e = A.ci.data() + A.diag[size_t(i)];
            const int* f = std::lower_bound(b, e, i - d);
            if (f != e && *f == i - d) coord[a][size_t(i)] = coord[a][size_t(i - d)] + 1;
        }
    }
    return nd;
}

void tileOrder(const HostBsr& A, const std::vector<int> coord[3], int ndim, const int dims[3], TileSched& T) {
    const int n = A.n;
    T = TileSched{};
    long long ext[3] = {1, 1, 1};
    for (int a = 0; a < 3; ++a) {
        T.dims[a] = a < ndim ? std::max(1, dims[a]) : 1;
        int mx = 0;
        if (a < ndim)
            for (int i = 0; i < n; ++i) mx = std::max(mx, coord[a][size_t(i)]);
        ext[a] = mx / T.dims[a] + 1;
    }
    std::vector<long long> tile(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) {
        long long t[3] = {0, 0, 0};
        for (int a = 0; a < ndim; ++a) t[a] = coord[a][size_t(i)] / T.dims[a];
        tile[size_t(i)] = t[0] + ext[0] * (t[1] + ext[1] * t[2]);
    }
    // tile level on the symmetrised lower graph: a coupled earlier row of the same tile may share the level, one of
    // another tile forces the next level (pushed forward along the upper entries, as levelOrder does)
    std::vector<int> tl(static_cast<size_t>(n), 0), ll(static_cast<size_t>(n), 0);
    for (int j = 0; j < n; ++j) {
        int l = tl[size_t(j)];
        for (int p = A.rp[size_t(j)]; p < A.rp[size_t(j) + 1]; ++p) {
            const int c = A.ci[size_t(p)];
            if (c < j) l = std::max(l, tl[size_t(c)] + (tile[size_t(c)] != tile[size_t(j)] ? 1 : 0));
        }
        tl[size_t(j)] = l;
        for (int p = A.rp[size_t(j)]; p < A.rp[size_t(j) + 1]; ++p) {
            const int c = A.ci[size_t(p)];
            if (c > j) tl[size_t(c)] = std::max(tl[size_t(c)], l + (tile[size_t(c)] != tile[size_t(j)] ? 1 : 0));
        }
    }
    // local levels inside a unit (same tile and tile level)
    auto same = [&](int a, int b) { return tile[size_t(a)] == tile[size_t(b)] && tl[size_t(a)] == tl[size_t(b)]; };
    for (int j = 0; j < n; ++j) {
        int l = ll[size_t(j)];
        for (int p = A.rp[size_t(j)]; p < A.rp[size_t(j) + 1]; ++p) {
            const int c = A.ci[size_t(p)];
            if (c < j && same(c, j)) l = std::max(l, ll[size_t(c)] + 1);
        }
        ll[size_t(j)] = l;
        for (int p = A.rp[size_t(j)]; p < A.rp[size_t(j) + 1]; ++p) {
            const int c = A.ci[size_t(p)];
            if (c > j && same(c, j)) ll[size_t(c)] = std::max(ll[size_t(c)], l + 1);
        }
    }
    T.ord.resize(size_t(n));
    std::iota(T.ord.begin(), T.ord.end(), 0);
    std::sort(T.ord.begin(), T.ord.end(), [&](int a, int b) {
        if (tl[size_t(a)] != tl[size_t(b)]) return tl[size_t(a)] < tl[size_t(b)];
        if (tile[size_t(a)] != tile[size_t(b)]) return tile[size_t(a)] < tile[size_t(b)];
        if (ll[size_t(a)] != ll[size_t(b)]) return ll[size_t(a)] < ll[size_t(b)];
        return a < b;
    });
    T.tlStart.assign(1, 0);
    T.lvPtr.assign(1, 0);
    for (int w = 0; w < n;) {
        const int i0 = T.ord[size_t(w)];
        if (w > 0 && tl[size_t(T.ord[size_t(w - 1)])] != tl[size_t(i0)]) T.tlStart.push_back(int(T.unitStart.size()));
        T.unitStart.push_back(w);
        int e = w;
        while (e < n && same(T.ord[size_t(e)], i0)) {
            if (e == w || ll[size_t(T.ord[size_t(e)])] != ll[size_t(T.ord[size_t(e - 1)])]) T.lvStart.push_back(e - w);
            ++e;
        }
        T.lvStart.push_back(e - w);
        T.lvPtr.push_back(int(T.lvStart.size()));
        T.maxRows = std::max(T.maxRows, e - w);
        T.maxLevels = std::max(T.maxLevels, ll[size_t(T.ord[size_t(e - 1)])] + 1);
        w = e;
    }
    T.unitStart.push_back(n);
    T.tlStart.push_back(int(T.unitStart.size()) - 1);
}

bool rbTileOrder(const HostBsr& S, const int dims[3], RbTiles& T) {
    T = RbTiles{};
    const int n = S.n;
    int off[3] = {0, 0, 0};
    std::vector<int> coord[3];
    const int nd = gridCoords(S, off, coord);
    if (nd == 0 || n == 0) return false;
    std::vector<int> len(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) len[size_t(i)] = S.rp[size_t(i) + 1] - S.rp[size_t(i)];
    const int cap = rowCap(len);
    std::vector<char> far(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) far[size_t(i)] = len[size_t(i)] > cap || len[size_t(i)] - 1 > kTileWaveCap;   // the tile sweep's slots per row
    // tile key (tile coordinates, the slowest axis first) and colour (parity of the tile coordinates' sum)
    std::vector<long long> tile(static_cast<size_t>(n));
    std::vector<char> colour(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) {
        long long key = 0;
        int par = 0;
        for (int a = nd; a-- > 0;) {
            const int t = coord[a][size_t(i)] / std::max(1, dims[a]);
            key = key * 65536 + t;
            par += t;
        }
        tile[size_t(i)] = key;
        colour[size_t(i)] = char(par & 1);
    }
    // bulk rows of different tiles of one colour must not be coupled (a 7-point stencil couples face neighbours only):
    // a row with such a coupling (a perforated cell coupled to the well's other perforations) goes behind the tiles too
    std::vector<char> cross(static_cast<size_t>(n), 0);
    for (int i = 0; i < n; ++i) {
        if (far[size_t(i)]) continue;
        for (int p = S.rp[size_t(i)]; p < S.rp[size_t(i) + 1]; ++p) {
            const int j = S.ci[size_t(p)];
            if (j != i && !far[size_t(j)] && tile[size_t(j)] != tile[size_t(i)] && colour[size_t(j)] == colour[size_t(i)]) cross[size_t(i)] = 1;
        }
    }
    long long nfar = 0;
    for (int i = 0; i < n; ++i) {
        far[size_t(i)] = far[size_t(i)] || cross[size_t(i)];
        nfar += far[size_t(i)];
    }
    if (10 * nfar > n) return false;   // not a grid with a few wells
    // local level inside a tile: natural order, one more than the coupled earlier rows of the same tile
    std::vector<int> lvl(static_cast<size_t>(n), 0);
    for (int i = 0; i < n; ++i) {
        if (far[size_t(i)]) continue;
        int l = 0;
        for (int p = S.rp[size_t(i)]; p < S.rp[size_t(i) + 1]; ++p) {
            const int j = S.ci[size_t(p)];
            if (j < i && !far[size_t(j)] && tile[size_t(j)] == tile[size_t(i)]) l = std::max(l, lvl[size_t(j)] + 1);
        }
        lvl[size_t(i)] = l;
    }
    std::vector<int> bulk;
    std::vector<int> farRows;
    for (int i = 0; i < n; ++i) (far[size_t(i)] ? farRows : bulk).push_back(i);
    std::sort(bulk.begin(), bulk.end(), [&](int a, int b) {
        if (colour[size_t(a)] != colour[size_t(b)]) return colour[size_t(a)] < colour[size_t(b)];
        if (tile[size_t(a)] != tile[size_t(b)]) return tile[size_t(a)] < tile[size_t(b)];
        if (lvl[size_t(a)] != lvl[size_t(b)]) return lvl[size_t(a)] < lvl[size_t(b)];
        return a < b;
    });
    T.nTile = int(bulk.size());
    T.nRed = 0;
    while (T.nRed < T.nTile && colour[size_t(bulk[size_t(T.nRed)])] == 0) ++T.nRed;
    if (T.nRed == 0 || T.nRed == T.nTile) return false;
    // units (tile, local level) and tiles over the bulk rows; factor levels: red local levels, then black ones
    int redLevels = 0;
    for (int k = 0; k < T.nRed; ++k) redLevels = std::max(redLevels, lvl[size_t(bulk[size_t(k)])] + 1);
    T.redLevels = redLevels;
    int blackLevels = 0;
    for (int k = T.nRed; k < T.nTile; ++k) blackLevels = std::max(blackLevels, lvl[size_t(bulk[size_t(k)])] + 1);
    T.blackLevels = blackLevels;
    for (int k = 0; k < T.nTile; ++k) {
        const int i = bulk[size_t(k)];
        const bool newTile = k == 0 || tile[size_t(bulk[size_t(k - 1)])] != tile[size_t(i)];
        const bool newUnit = newTile || lvl[size_t(bulk[size_t(k - 1)])] != lvl[size_t(i)];
        if (newTile) {
            T.tileUnit.push_back(int(T.unitStart.size()));
            if (k < T.nRed) T.redTiles = int(T.tileUnit.size());
        }
        if (newUnit) {
            T.unitStart.push_back(k);
            T.unitLevel.push_back(colour[size_t(i)] ? redLevels + lvl[size_t(i)] : lvl[size_t(i)]);
        }
    }
    T.unitStart.push_back(T.nTile);
    T.tileUnit.push_back(int(T.unitStart.size()) - 1);
    // the far rows behind the tiles, in the level order of their own couplings
    HostBsr W;
    W.n = int(farRows.size());
    W.nb = S.nb;
    W.rp.assign(1, 0);
    std::vector<int> at(static_cast<size_t>(n), -1);
    for (size_t k = 0; k < farRows.size(); ++k) at[size_t(farRows[k])] = int(k);
    for (int i : farRows) {
        for (int p = S.rp[size_t(i)]; p < S.rp[size_t(i) + 1]; ++p)
            if (at[size_t(S.ci[size_t(p)])] >= 0) W.ci.push_back(at[size_t(S.ci[size_t(p)])]);
        W.rp.push_back(int(W.ci.size()));
    }
    W.diag.assign(size_t(W.n), -1);
    for (int i = 0; i < W.n; ++i)
        for (int p = W.rp[size_t(i)]; p < W.rp[size_t(i) + 1]; ++p)
            if (W.ci[size_t(p)] == i) W.diag[size_t(i)] = p;
    std::vector<int> word;
    T.farStart.assign(1, 0);
    if (W.n > 0) levelOrder(W, word, T.farStart);
    T.ord = bulk;
    for (int k : word) T.ord.push_back(farRows[size_t(k)]);
    return true;
}

}  // namespace rsh
