// Exact elimination of the well (BHP) unknowns from a scalar reservoir Jacobian: the host-side plan shared by the
// workload generator (csrc/gen/rs_gen.cpp, values on the host) and the device reduction (csrc/hip/rsh_wells.cpp,
// values in HBM).  Plain C++17, no HIP: it builds and runs under the host sanitizers.
//
// The Jacobian J is scalar CSR with strictly increasing columns per row: cells first, in blocks of nv unknowns
// (na blocks, nc = na * nv rows), then one BHP row per well (nw rows).  Writing J in the 2 x 2 partition
//     J = [ A  B ]    A: nc x nc cell block,  B: nc x nw cell -> well columns,
//         [ C  D ]    C: nw x nc well -> cell columns,  D: nw x nw, diagonal (a well group is not supported),
// the reduced cell system is S = A - B D^-1 C.  With D diagonal, B D^-1 C is a sum of one outer product per well,
//     S = A - sum_w (B[:, w] / d_w) C[w, :],
// which is how the values are formed here: the cell entries first, then one rank-1 update per well in ascending
// well order.  Every entry of S therefore sees the same sequence of floating-point operations as the engine's
// row-by-row reduction (ressim::WellSchur::build, engine/src/bsr.cpp:733-840: the cell entry, then the wells of
// the scalar row in column order, each term (J_rw * (1 / d_w)) * J_wc), so S matches it bit for bit.
//
// The block pattern of S is the cell pattern of every block row, its diagonal block, and, for every well that a row
// of the block is coupled to, all the cell blocks the well's BHP row touches (its perforations).  The engine builds
// that pattern from the couplings whose value is nonzero (bsr.cpp:755-757); `keepCoupling` lets the caller apply that
// rule (the generator) or keep every structural coupling (the device, whose pattern is fixed before any values).
#pragma once

#include <algorithm>
#include <cmath>
#include <functional>
#include <string>
#include <vector>

namespace wellplan {

struct Plan {
    int na = 0, nv = 1, nw = 0, nc = 0;
    // per well: the J index of its diagonal (-1: absent), the J indices of its BHP row on cell columns (column
    // order), and of its entries on other wells' columns (they must be zero: checked on values)
    std::vector<int> bhpDiag;
    std::vector<int> toCellStart, toCellEntry;
    std::vector<int> toWellStart, toWellEntry;
    // per well: the scalar cell rows coupled to it (ascending) and the J index of each coupling
    std::vector<int> fromStart, fromRow, fromEntry;
    // per well: the cell blocks of its BHP row (sorted, unique)
    std::vector<int> perfStart, perfBlock;
    // the reduced block pattern (sorted columns per block row) and its size without the well fill
    std::vector<int> sPtr, sCol;
    long long cellBlocks = 0;
    int fill() const { return int((long long)sCol.size() - cellBlocks); }
    // the block of S at (block row i, block column j): its index, or -1
    int blockAt(int i, int j) const {
        const auto b = sCol.begin() + sPtr[size_t(i)], e = sCol.begin() + sPtr[size_t(i) + 1];
        const auto it = std::lower_bound(b, e, j);
        return it != e && *it == j ? int(it - sCol.begin()) : -1;
    }
};

// Checks J's shape and ordering and splits it by partition; returns "" or the reason it is refused.
inline std::string split(int n, const int* rowPtr, const int* col, int na, int nv, int nw, const std::function<bool(int)>& keepCoupling, Plan& P) {
    if (na <= 0 || nv < 1 || nw < 0 || n != na * nv + nw) return "n must equal na * nv + nw";
    P = Plan{};
    P.na = na; P.nv = nv; P.nw = nw; P.nc = na * nv;
    const int nc = P.nc;
    if (rowPtr[0] != 0) return "row_ptr[0] must be 0";
    for (int r = 0; r < n; ++r) {
        if (rowPtr[r + 1] < rowPtr[r]) return "row_ptr not monotone";
        for (int p = rowPtr[r]; p < rowPtr[r + 1]; ++p) {
            if (col[p] < 0 || col[p] >= n) return "column index out of range";
            if (p > rowPtr[r] && col[p] <= col[p - 1]) return "columns must be strictly increasing in every row";
        }
    }
    // well rows
    P.bhpDiag.assign(size_t(nw), -1);
    P.toCellStart.assign(1, 0);
    P.toWellStart.assign(1, 0);
    for (int w = 0; w < nw; ++w) {
        const int r = nc + w;
        for (int p = rowPtr[r]; p < rowPtr[r + 1]; ++p) {
            if (col[p] == r) P.bhpDiag[size_t(w)] = p;
            else if (col[p] < nc) P.toCellEntry.push_back(p);
            else P.toWellEntry.push_back(p);
        }
        P.toCellStart.push_back(int(P.toCellEntry.size()));
        P.toWellStart.push_back(int(P.toWellEntry.size()));
    }
    // cell -> well couplings, bucketed by well (rows visited in ascending order, so each bucket is sorted)
    std::vector<int> count(size_t(nw) + 1, 0);
    for (int r = 0; r < nc; ++r)
        for (int p = rowPtr[r]; p < rowPtr[r + 1]; ++p)
            if (col[p] >= nc && keepCoupling(p)) ++count[size_t(col[p] - nc) + 1];
    for (int w = 0; w < nw; ++w) count[size_t(w) + 1] += count[size_t(w)];
    P.fromStart = count;
    P.fromRow.assign(size_t(count[size_t(nw)]), 0);
    P.fromEntry.assign(size_t(count[size_t(nw)]), 0);
    for (int r = 0; r < nc; ++r)
        for (int p = rowPtr[r]; p < rowPtr[r + 1]; ++p)
            if (col[p] >= nc && keepCoupling(p)) {
                const int slot = count[size_t(col[p] - nc)]++;
                P.fromRow[size_t(slot)] = r;
                P.fromEntry[size_t(slot)] = p;
            }
    // perforated blocks of every well
    P.perfStart.assign(1, 0);
    for (int w = 0; w < nw; ++w) {
        std::vector<int> blocks;
        for (int q = P.toCellStart[size_t(w)]; q < P.toCellStart[size_t(w) + 1]; ++q) blocks.push_back(col[P.toCellEntry[size_t(q)]] / nv);
        std::sort(blocks.begin(), blocks.end());
        blocks.erase(std::unique(blocks.begin(), blocks.end()), blocks.end());
        P.perfBlock.insert(P.perfBlock.end(), blocks.begin(), blocks.end());
        P.perfStart.push_back(int(P.perfBlock.size()));
    }
    return "";
}

// The reduced block pattern: per block row, the sorted union of its cell columns, itself, and the perforations of
// the wells it is coupled to.
inline void pattern(int n, const int* rowPtr, const int* col, Plan& P) {
    (void)n;
    const int na = P.na, nv = P.nv, nc = P.nc;
    // wells coupled to each block row, ascending (wells are visited in order; a row meets a well at most once)
    std::vector<int> wPtr(size_t(na) + 1, 0), lastWell(size_t(na), -1);
    for (int w = 0; w < P.nw; ++w)
        for (int u = P.fromStart[size_t(w)]; u < P.fromStart[size_t(w) + 1]; ++u) {
            const int i = P.fromRow[size_t(u)] / nv;
            if (lastWell[size_t(i)] != w) { lastWell[size_t(i)] = w; ++wPtr[size_t(i) + 1]; }
        }
    for (int i = 0; i < na; ++i) wPtr[size_t(i) + 1] += wPtr[size_t(i)];
    std::vector<int> wList(static_cast<size_t>(wPtr[size_t(na)]));
    std::vector<int> next(wPtr.begin(), wPtr.end() - 1);
    std::fill(lastWell.begin(), lastWell.end(), -1);
    for (int w = 0; w < P.nw; ++w)
        for (int u = P.fromStart[size_t(w)]; u < P.fromStart[size_t(w) + 1]; ++u) {
            const int i = P.fromRow[size_t(u)] / nv;
            if (lastWell[size_t(i)] != w) { lastWell[size_t(i)] = w; wList[size_t(next[size_t(i)]++)] = w; }
        }
    P.sPtr.assign(1, 0);
    P.sCol.clear();
    P.cellBlocks = 0;
    std::vector<int> own, merged;
    for (int i = 0; i < na; ++i) {
        own.assign(1, i);
        for (int r = i * nv; r < (i + 1) * nv; ++r)
            for (int p = rowPtr[r]; p < rowPtr[r + 1] && col[p] < nc; ++p) own.push_back(col[p] / nv);
        std::sort(own.begin(), own.end());
        own.erase(std::unique(own.begin(), own.end()), own.end());
        P.cellBlocks += (long long)own.size();
        for (int k = wPtr[size_t(i)]; k < wPtr[size_t(i) + 1]; ++k) {
            const int w = wList[size_t(k)];
            merged.clear();
            std::set_union(own.begin(), own.end(), P.perfBlock.begin() + P.perfStart[size_t(w)], P.perfBlock.begin() + P.perfStart[size_t(w) + 1],
                           std::back_inserter(merged));
            own.swap(merged);
        }
        P.sCol.insert(P.sCol.end(), own.begin(), own.end());
        P.sPtr.push_back(int(P.sCol.size()));
    }
}

// Where each value of S comes from: `cellSource[t]` is the J index of the cell entry landing in S value t (-1: none),
// and the rank-1 terms of value t, in the order they are subtracted (ascending well), are
// terms[termStart[t] .. termStart[t+1]) = {J index of the coupling J_rw, well w, J index of J_wc}.
struct ValueMap {
    std::vector<int> cellSource, termStart;
    std::vector<int> termCoupling, termWell, termEntry;
};

// S value index of scalar (row r, column c), both cell unknowns
inline long long valueIndex(const Plan& P, int r, int c) {
    const int blk = P.blockAt(r / P.nv, c / P.nv);
    return blk < 0 ? -1 : (long long)blk * P.nv * P.nv + (r % P.nv) * P.nv + c % P.nv;
}

inline std::string valueMap(const int* rowPtr, const int* col, const Plan& P, ValueMap& M) {
    const long long nvals = (long long)P.sCol.size() * P.nv * P.nv;
    M.cellSource.assign(size_t(nvals), -1);
    for (int r = 0; r < P.nc; ++r)
        for (int p = rowPtr[r]; p < rowPtr[r + 1] && col[p] < P.nc; ++p) M.cellSource[size_t(valueIndex(P, r, col[p]))] = p;
    // count the terms per value, then place them well by well (ascending), so each value's terms are in well order
    std::vector<int> cnt(size_t(nvals) + 1, 0);
    for (int w = 0; w < P.nw; ++w)
        for (int u = P.fromStart[size_t(w)]; u < P.fromStart[size_t(w) + 1]; ++u)
            for (int q = P.toCellStart[size_t(w)]; q < P.toCellStart[size_t(w) + 1]; ++q) {
                const long long t = valueIndex(P, P.fromRow[size_t(u)], col[P.toCellEntry[size_t(q)]]);
                if (t < 0) return "internal: a rank-1 term outside the reduced pattern";
                ++cnt[size_t(t) + 1];
            }
    for (long long t = 0; t < nvals; ++t) cnt[size_t(t) + 1] += cnt[size_t(t)];
    M.termStart = cnt;
    const size_t nterms = size_t(cnt[size_t(nvals)]);
    M.termCoupling.assign(nterms, 0);
    M.termWell.assign(nterms, 0);
    M.termEntry.assign(nterms, 0);
    for (int w = 0; w < P.nw; ++w)
        for (int u = P.fromStart[size_t(w)]; u < P.fromStart[size_t(w) + 1]; ++u)
            for (int q = P.toCellStart[size_t(w)]; q < P.toCellStart[size_t(w) + 1]; ++q) {
                const long long t = valueIndex(P, P.fromRow[size_t(u)], col[P.toCellEntry[size_t(q)]]);
                const int k = cnt[size_t(t)]++;
                M.termCoupling[size_t(k)] = P.fromEntry[size_t(u)];
                M.termWell[size_t(k)] = w;
                M.termEntry[size_t(k)] = P.toCellEntry[size_t(q)];
            }
    return "";
}

// 1 / d_w for every well, with the engine's refusal of an unusable BHP derivative (|d_w| <= 1e-12 max |C[w, :]|, or
// not finite, bsr.cpp:747-748) and of a nonzero well-well coupling (bsr.cpp:745-746); returns "" or the reason
inline std::string wellInverses(const Plan& P, const double* val, std::vector<double>& dInv) {
    dInv.assign(size_t(P.nw), 0.0);
    for (int w = 0; w < P.nw; ++w) {
        for (int q = P.toWellStart[size_t(w)]; q < P.toWellStart[size_t(w) + 1]; ++q)
            if (val[P.toWellEntry[size_t(q)]] != 0.0) return "well-well coupling is not supported by the well Schur reduction";
        const double d = P.bhpDiag[size_t(w)] >= 0 ? val[P.bhpDiag[size_t(w)]] : 0.0;
        double big = 0;
        for (int q = P.toCellStart[size_t(w)]; q < P.toCellStart[size_t(w) + 1]; ++q) big = std::max(big, std::fabs(val[P.toCellEntry[size_t(q)]]));
        if (!std::isfinite(d) || !(std::fabs(d) > 1e-12 * std::max(big, 1e-300))) return "well row without usable BHP derivative";
        dInv[size_t(w)] = 1.0 / d;
    }
    return "";
}

// S's values: the cell entries, then one rank-1 update per well (ascending)
inline void reduceValues(const int* rowPtr, const int* col, const double* val, const Plan& P, const std::vector<double>& dInv, double* s) {
    const long long nvals = (long long)P.sCol.size() * P.nv * P.nv;
    std::fill(s, s + nvals, 0.0);
    for (int r = 0; r < P.nc; ++r)
        for (int p = rowPtr[r]; p < rowPtr[r + 1] && col[p] < P.nc; ++p) s[valueIndex(P, r, col[p])] += val[p];
    for (int w = 0; w < P.nw; ++w)
        for (int u = P.fromStart[size_t(w)]; u < P.fromStart[size_t(w) + 1]; ++u) {
            const int r = P.fromRow[size_t(u)];
            const double scale = val[P.fromEntry[size_t(u)]] * dInv[size_t(w)];
            for (int q = P.toCellStart[size_t(w)]; q < P.toCellStart[size_t(w) + 1]; ++q) {
                const int e = P.toCellEntry[size_t(q)];
                s[valueIndex(P, r, col[e])] -= scale * val[e];
            }
        }
}

// Couplings regrouped by cell row (row-major view of B): per scalar cell row, its (well, J index) pairs in ascending
// well order, for the right-hand-side reduction
inline void couplingsByRow(const Plan& P, std::vector<int>& start, std::vector<int>& well, std::vector<int>& entry) {
    start.assign(size_t(P.nc) + 1, 0);
    for (int u = 0; u < P.fromStart[size_t(P.nw)]; ++u) ++start[size_t(P.fromRow[size_t(u)]) + 1];
    for (int r = 0; r < P.nc; ++r) start[size_t(r) + 1] += start[size_t(r)];
    std::vector<int> next(start.begin(), start.end() - 1);
    well.assign(size_t(start[size_t(P.nc)]), 0);
    entry.assign(well.size(), 0);
    for (int w = 0; w < P.nw; ++w)
        for (int u = P.fromStart[size_t(w)]; u < P.fromStart[size_t(w) + 1]; ++u) {
            const int k = next[size_t(P.fromRow[size_t(u)])]++;
            well[size_t(k)] = w;
            entry[size_t(k)] = P.fromEntry[size_t(u)];
        }
}

// b_c = b_cells - B D^-1 b_wells, each row's wells ascending (as WellSchur::reduceRhs, bsr.cpp:973-980)
inline void reduceRhs(const Plan& P, const double* val, const std::vector<double>& dInv, const double* b, double* bc) {
    std::vector<int> start, well, entry;
    couplingsByRow(P, start, well, entry);
    for (int r = 0; r < P.nc; ++r) {
        double acc = b[r];
        for (int k = start[size_t(r)]; k < start[size_t(r) + 1]; ++k) acc -= val[entry[size_t(k)]] * dInv[size_t(well[size_t(k)])] * b[P.nc + well[size_t(k)]];
        bc[r] = acc;
    }
}

// x_cells = x_c, x_w = (b_w - C[w, :] x_c) / d_w (as WellSchur::recover, bsr.cpp:982-990)
inline void recover(const Plan& P, const int* col, const double* val, const std::vector<double>& dInv, const double* b, const double* xc, double* x) {
    if (x != xc) std::copy(xc, xc + P.nc, x);
    for (int w = 0; w < P.nw; ++w) {
        double acc = b[P.nc + w];
        for (int q = P.toCellStart[size_t(w)]; q < P.toCellStart[size_t(w) + 1]; ++q) {
            const int e = P.toCellEntry[size_t(q)];
            acc -= val[e] * xc[col[e]];
        }
        x[P.nc + w] = acc * dInv[size_t(w)];
    }
}

}  // namespace wellplan
