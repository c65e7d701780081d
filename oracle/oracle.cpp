// ORACLE - test infrastructure only.  A clean-room CPU restatement of the reference linear-solve path, used by the
// tests, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker.  The product library never links
// or calls it.  Every function cites the reference file:line whose algorithm it restates.
//
// Pinning (what checks each part, see tests/test_oracle_*.py):
//   spmv            closed form; equals the compiled reference (oracle/_ref) bit for bit; equals scipy CSR spmv
//   invert_block    A * A^-1 = I; singular detection; equals reference
//   BILU0 factor    exact on block-tridiagonal (run_tests.cpp:1606); nb = 1 equals textbook scalar ILU(0) (scipy
//                   products on the pattern); apply equals the compiled reference to 1e-13
//   MCBILU0         valid colouring; 5-point -> 2 colours; equals BILU0 of P A P^T; colouring equals reference
//   BiCGSTAB/GMRES  iteration counts and residuals equal the compiled reference; solution vs scipy spsolve
//   BlockCPR / AMG  weights closed form (first row of D^-1); pressure matrix closed form; aggregates, level sizes and
//                   V-cycle output equal the compiled reference; refresh equals a fresh setup
//   lean BILU0      factors are float32 values and differ from rounding the FP64 factors; apply, BiCGSTAB iterations,
//                   residuals and solutions equal the compiled reference (setLean) bit for bit
//   true-IMPES      weights from accumulation blocks (the engine's own, dumped with the demo waterflood, and
//                   ill-conditioned variants) and the fallback count equal the compiled reference bit for bit
// Built with -ffp-contract=off (the reference's MinGW flags, build_mingw.sh:18) so products are not fused.
#include "rs_check.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

namespace oracle {

constexpr int kMaxB = 16;

struct Bsr {
    int n = 0, nb = 1;
    std::vector<int> rp, ci, diag;
    std::vector<double> v;
    double* blk(int p) { return v.data() + size_t(p) * nb * nb; }
    const double* blk(int p) const { return v.data() + size_t(p) * nb * nb; }
};
struct Csr {
    int n = 0;
    std::vector<int> rp, ci, diag;
    std::vector<double> v;
};

static void findDiag(Bsr& A) {
    A.diag.assign(size_t(A.n), -1);
    for (int i = 0; i < A.n; ++i)
        for (int p = A.rp[size_t(i)]; p < A.rp[size_t(i) + 1]; ++p) if (A.ci[size_t(p)] == i) { A.diag[size_t(i)] = p; break; }
}
static void findDiag(Csr& A) {
    A.diag.assign(size_t(A.n), -1);
    for (int i = 0; i < A.n; ++i)
        for (int p = A.rp[size_t(i)]; p < A.rp[size_t(i) + 1]; ++p) if (A.ci[size_t(p)] == i) { A.diag[size_t(i)] = p; break; }
}
static Bsr makeBsr(int n, int nb, const int* rp, const int* ci, const double* v) {
    Bsr A; A.n = n; A.nb = nb;
    A.rp.assign(rp, rp + n + 1);
    A.ci.assign(ci, ci + rp[n]);
    A.v.assign(v, v + size_t(rp[n]) * nb * nb);
    findDiag(A);
    return A;
}

// ---- dense block kernels (bsr.cpp:14-56): per output row, partial sums over the inner index in ascending order
template <typename T>
static void blockMulAdd(const T* a, const double* x, double* y, int nb, double sign) {   // y +/-= a x
    for (int i = 0; i < nb; ++i) { double s = 0; for (int j = 0; j < nb; ++j) s += double(a[i * nb + j]) * x[j]; y[i] += sign * s; }
}
static void blockMulSubY(const double* a, const double* x, double* y, int nb) {
    for (int i = 0; i < nb; ++i) { double s = 0; for (int j = 0; j < nb; ++j) s += a[i * nb + j] * x[j]; y[i] -= s; }
}
static void blockGemm(const double* a, const double* b, double* c, int nb) {           // c = a b
    for (int i = 0; i < nb; ++i) {
        double r[kMaxB] = {};
        for (int k = 0; k < nb; ++k) { const double f = a[i * nb + k]; for (int j = 0; j < nb; ++j) r[j] += f * b[k * nb + j]; }
        for (int j = 0; j < nb; ++j) c[i * nb + j] = r[j];
    }
}
static void blockGemmSub(const double* a, const double* b, double* c, int nb) {        // c -= a b
    for (int i = 0; i < nb; ++i) {
        double r[kMaxB] = {};
        for (int k = 0; k < nb; ++k) { const double f = a[i * nb + k]; for (int j = 0; j < nb; ++j) r[j] += f * b[k * nb + j]; }
        for (int j = 0; j < nb; ++j) c[i * nb + j] -= r[j];
    }
}

// Gauss-Jordan inverse with partial pivoting on [A | I]; singular when a pivot <= 1e-14 * max|a| (bsr.cpp:67-87)
static bool invertBlock(double* a, int nb) {
    double m[kMaxB][2 * kMaxB];
    double scale = 0;
    for (int i = 0; i < nb * nb; ++i) scale = std::max(scale, std::fabs(a[i]));
    for (int i = 0; i < nb; ++i) for (int j = 0; j < nb; ++j) { m[i][j] = a[i * nb + j]; m[i][nb + j] = i == j ? 1.0 : 0.0; }
    if (!(scale > 0) || !std::isfinite(scale)) return false;
    for (int k = 0; k < nb; ++k) {
        int piv = k;
        for (int i = k + 1; i < nb; ++i) if (std::fabs(m[i][k]) > std::fabs(m[piv][k])) piv = i;
        if (std::fabs(m[piv][k]) <= 1e-14 * scale) return false;
        if (piv != k) for (int j = 0; j < 2 * nb; ++j) std::swap(m[k][j], m[piv][j]);
        const double inv = 1.0 / m[k][k];
        for (int j = 0; j < 2 * nb; ++j) m[k][j] *= inv;
        for (int i = 0; i < nb; ++i) {
            if (i == k || m[i][k] == 0.0) continue;
            const double f = m[i][k];
            for (int j = 0; j < 2 * nb; ++j) m[i][j] -= f * m[k][j];
        }
    }
    for (int i = 0; i < nb; ++i) for (int j = 0; j < nb; ++j) a[i * nb + j] = m[i][nb + j];
    return true;
}

// pivot block of a factorised row: invert, else shift the diagonal by 1e-8 * scale and retry, else identity / scale
// (bsr.cpp:202-212); returns 1 when the block was regularised
static int invertPivot(double* d, int nb) {
    double save[kMaxB * kMaxB];
    const int b2 = nb * nb;
    std::memcpy(save, d, sizeof(double) * size_t(b2));
    if (invertBlock(d, nb)) return 0;
    double sc = 0;
    for (int e = 0; e < b2; ++e) sc = std::max(sc, std::fabs(save[e]));
    if (!(sc > 0) || !std::isfinite(sc)) sc = 1.0;
    std::memcpy(d, save, sizeof(double) * size_t(b2));
    for (int e = 0; e < nb; ++e) {
        double& x = d[e * nb + e];
        x += (x >= 0 ? 1e-8 : -1e-8) * sc + (x == 0 ? sc * 1e-8 : 0);
    }
    if (!invertBlock(d, nb)) { for (int e = 0; e < b2; ++e) d[e] = 0; for (int e = 0; e < nb; ++e) d[e * nb + e] = 1.0 / sc; }
    return 1;
}

// BSR SpMV: y_i = sum_p A_p x_col(p), rows zeroed first (bsr.cpp:109-126)
static void spmv(const Bsr& A, const double* x, double* y) {
    const int b = A.nb;
    for (int i = 0; i < A.n; ++i) {
        double* yi = y + size_t(i) * b;
        for (int e = 0; e < b; ++e) yi[e] = 0;
        for (int p = A.rp[size_t(i)]; p < A.rp[size_t(i) + 1]; ++p) blockMulAdd(A.blk(p), x + size_t(A.ci[size_t(p)]) * b, yi, b, 1.0);
    }
}

// ---- BILU0 (bsr.cpp:183-215 row elimination, 437-596 sweeps)
// mode 1: FP64; 2: FP64 factorisation, float32 factors in the sweeps (C38, bsr.cpp:413-419); 3: BILU0f, all float
// (bsr.cpp:1087-1121); 4: lean (C45, bilu0RowF bsr.cpp:216-263 with setLean, :358-389): the matrix is rounded to
// float32, each row is eliminated in double against the already-rounded factors of the earlier rows, its pivot
// block inverted in double, and the finished row stored rounded; swept like mode 2 (float factors, double vectors)
struct Bilu0 {
    Bsr LU;                          // L strictly lower (unit diag implied), U_ii^-1 on the diagonal, U strictly upper
    std::vector<float> F;            // float32 copy of the factors (modes 2, 3 and 4)
    int mode = 1, nReg = 0;
    static double toF(double x) { return double(float(x)); }
    void factor(const Bsr& A) {
        LU = A;
        const int n = LU.n, b = LU.nb, b2 = b * b;
        if (mode == 4) for (double& x : LU.v) x = toF(x);
        std::vector<int> pos(size_t(n), -1);
        double L[kMaxB * kMaxB];
        nReg = 0;
        for (int i = 0; i < n; ++i) {
            const int r0 = LU.rp[size_t(i)], r1 = LU.rp[size_t(i) + 1];
            for (int p = r0; p < r1; ++p) pos[size_t(LU.ci[size_t(p)])] = p;
            for (int p = r0; p < r1; ++p) {
                const int k = LU.ci[size_t(p)];
                if (k >= i) break;
                blockGemm(LU.blk(p), LU.blk(LU.diag[size_t(k)]), L, b);          // L_ik = A_ik U_kk^-1
                std::memcpy(LU.blk(p), L, sizeof(double) * size_t(b2));
                for (int q = LU.diag[size_t(k)] + 1; q < LU.rp[size_t(k) + 1]; ++q) {
                    const int j = LU.ci[size_t(q)];
                    const int t = pos[size_t(j)];
                    if (t >= r0 && t < r1) blockGemmSub(L, LU.blk(q), LU.blk(t), b);   // A_ij -= L_ik U_kj
                }
            }
            for (int p = r0; p < r1; ++p) pos[size_t(LU.ci[size_t(p)])] = -1;
            nReg += invertPivot(LU.blk(LU.diag[size_t(i)]), b);
            if (mode == 4)
                for (size_t e = size_t(r0) * b2; e < size_t(r1) * b2; ++e) LU.v[e] = toF(LU.v[e]);
        }
        if (mode >= 2) { F.resize(LU.v.size()); for (size_t e = 0; e < F.size(); ++e) F[e] = float(LU.v[e]); }
    }
    void apply(const double* r, double* z) const {
        const int n = LU.n, b = LU.nb;
        const size_t b2 = size_t(b) * b;
        if (mode == 3) {   // all-float sweeps (bsr.cpp:1098-1121)
            std::vector<float> y(size_t(n) * b);
            for (int i = 0; i < n; ++i) {
                float* yi = y.data() + size_t(i) * b;
                for (int e = 0; e < b; ++e) yi[e] = float(r[size_t(i) * b + e]);
                for (int p = LU.rp[size_t(i)]; p < LU.diag[size_t(i)]; ++p) {
                    const float* a = F.data() + size_t(p) * b2; const float* yk = y.data() + size_t(LU.ci[size_t(p)]) * b;
                    for (int e = 0; e < b; ++e) { float s = 0; for (int f = 0; f < b; ++f) s += a[e * b + f] * yk[f]; yi[e] -= s; }
                }
            }
            float t[kMaxB];
            for (int i = n - 1; i >= 0; --i) {
                float* yi = y.data() + size_t(i) * b;
                for (int p = LU.diag[size_t(i)] + 1; p < LU.rp[size_t(i) + 1]; ++p) {
                    const float* a = F.data() + size_t(p) * b2; const float* yj = y.data() + size_t(LU.ci[size_t(p)]) * b;
                    for (int e = 0; e < b; ++e) { float s = 0; for (int f = 0; f < b; ++f) s += a[e * b + f] * yj[f]; yi[e] -= s; }
                }
                const float* d = F.data() + size_t(LU.diag[size_t(i)]) * b2;
                for (int e = 0; e < b; ++e) { float s = 0; for (int f = 0; f < b; ++f) s += d[e * b + f] * yi[f]; t[e] = s; }
                for (int e = 0; e < b; ++e) yi[e] = t[e];
            }
            for (size_t e = 0; e < size_t(n) * b; ++e) z[e] = double(y[e]);
            return;
        }
        // forward: y_i = r_i - sum_{k<i} L_ik y_k (y in z)
        for (int i = 0; i < n; ++i) {
            double* zi = z + size_t(i) * b;
            for (int e = 0; e < b; ++e) zi[e] = r[size_t(i) * b + e];
            for (int p = LU.rp[size_t(i)]; p < LU.diag[size_t(i)]; ++p) {
                const double* zk = z + size_t(LU.ci[size_t(p)]) * b;
                if (mode == 2 || mode == 4) blockMulAdd(F.data() + size_t(p) * b2, zk, zi, b, -1.0);
                else blockMulSubY(LU.blk(p), zk, zi, b);
            }
        }
        // backward: z_i = U_ii^-1 (y_i - sum_{j>i} U_ij z_j)
        double t[kMaxB];
        for (int i = n - 1; i >= 0; --i) {
            double* zi = z + size_t(i) * b;
            for (int p = LU.diag[size_t(i)] + 1; p < LU.rp[size_t(i) + 1]; ++p) {
                const double* zj = z + size_t(LU.ci[size_t(p)]) * b;
                if (mode == 2 || mode == 4) blockMulAdd(F.data() + size_t(p) * b2, zj, zi, b, -1.0);
                else blockMulSubY(LU.blk(p), zj, zi, b);
            }
            for (int e = 0; e < b; ++e) t[e] = 0;
            if (mode == 2 || mode == 4) blockMulAdd(F.data() + size_t(LU.diag[size_t(i)]) * b2, zi, t, b, 1.0);
            else blockMulAdd(LU.blk(LU.diag[size_t(i)]), zi, t, b, 1.0);
            for (int e = 0; e < b; ++e) zi[e] = t[e];
        }
    }
};

// ---- multicolour BILU0 (bsr.cpp:599-730): greedy distance-1 colouring, largest degree first (stable), of the
// symmetrised block graph; ILU(0) of P A P^T
struct McBilu0 {
    std::vector<int> colour, perm, iperm, start;
    int colours = 0;
    Bilu0 ilu;
    void factor(const Bsr& A) {
        const int n = A.n;
        std::vector<std::vector<int>> adj(static_cast<size_t>(n));
        for (int i = 0; i < n; ++i)
            for (int p = A.rp[size_t(i)]; p < A.rp[size_t(i) + 1]; ++p) {
                const int j = A.ci[size_t(p)];
                if (j != i) { adj[size_t(i)].push_back(j); adj[size_t(j)].push_back(i); }
            }
        for (auto& a : adj) { std::sort(a.begin(), a.end()); a.erase(std::unique(a.begin(), a.end()), a.end()); }
        std::vector<int> order(static_cast<size_t>(n));
        for (int i = 0; i < n; ++i) order[size_t(i)] = i;
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return adj[size_t(x)].size() > adj[size_t(y)].size(); });
        colour.assign(size_t(n), -1);
        colours = 0;
        std::vector<char> used;
        for (int i : order) {
            used.assign(adj[size_t(i)].size() + 2, 0);
            for (int j : adj[size_t(i)]) { const int c = colour[size_t(j)]; if (c >= 0 && c < int(used.size())) used[size_t(c)] = 1; }
            int c = 0;
            while (used[size_t(c)]) ++c;
            colour[size_t(i)] = c;
            colours = std::max(colours, c + 1);
        }
        start.assign(size_t(colours) + 1, 0);
        for (int i = 0; i < n; ++i) start[size_t(colour[size_t(i)]) + 1]++;
        for (int c = 0; c < colours; ++c) start[size_t(c) + 1] += start[size_t(c)];
        perm.assign(size_t(n), 0); iperm.assign(size_t(n), 0);
        std::vector<int> fill(start.begin(), start.end() - 1);
        for (int i = 0; i < n; ++i) { const int k = fill[size_t(colour[size_t(i)])]++; perm[size_t(k)] = i; iperm[size_t(i)] = k; }
        // P A P^T with sorted columns
        Bsr P; P.n = n; P.nb = A.nb; P.rp.assign(size_t(n) + 1, 0);
        const size_t b2 = size_t(A.nb) * A.nb;
        P.ci.reserve(A.ci.size()); P.v.reserve(A.v.size());
        std::vector<std::pair<int, int>> row;
        for (int k = 0; k < n; ++k) {
            const int i = perm[size_t(k)];
            row.clear();
            for (int p = A.rp[size_t(i)]; p < A.rp[size_t(i) + 1]; ++p) row.push_back({iperm[size_t(A.ci[size_t(p)])], p});
            std::sort(row.begin(), row.end());
            for (auto& e : row) { P.ci.push_back(e.first); P.v.insert(P.v.end(), A.v.begin() + long(size_t(e.second) * b2), A.v.begin() + long(size_t(e.second + 1) * b2)); }
            P.rp[size_t(k) + 1] = int(P.ci.size());
        }
        findDiag(P);
        ilu.mode = 1;
        ilu.factor(P);
    }
    void apply(const double* r, double* z) const {
        const int n = ilu.LU.n, b = ilu.LU.nb;
        std::vector<double> rp(size_t(n) * b), zp(size_t(n) * b);
        for (int k = 0; k < n; ++k) std::memcpy(&rp[size_t(k) * b], r + size_t(perm[size_t(k)]) * b, sizeof(double) * size_t(b));
        ilu.apply(rp.data(), zp.data());
        for (int k = 0; k < n; ++k) std::memcpy(z + size_t(perm[size_t(k)]) * b, &zp[size_t(k) * b], sizeof(double) * size_t(b));
    }
};

// ---- aggregation AMG (amg.cpp:10-337)
static Csr sparseProduct(const Csr& A, int acols, const std::vector<int>& brp, const std::vector<int>& bci, const std::vector<double>& bv,
                         int bcols, int arows) {
    // Gustavson product with rows sorted by column afterwards (amg.cpp:36-68); A is arows x acols
    (void)acols;
    Csr C; C.n = arows; C.rp.assign(size_t(arows) + 1, 0);
    std::vector<int> mark(size_t(bcols), -1);
    for (int i = 0; i < arows; ++i) {
        int cnt = 0;
        for (int p = A.rp[size_t(i)]; p < A.rp[size_t(i) + 1]; ++p)
            for (int q = brp[size_t(A.ci[size_t(p)])]; q < brp[size_t(A.ci[size_t(p)]) + 1]; ++q)
                if (mark[size_t(bci[size_t(q)])] != i) { mark[size_t(bci[size_t(q)])] = i; ++cnt; }
        C.rp[size_t(i) + 1] = C.rp[size_t(i)] + cnt;
    }
    C.ci.resize(size_t(C.rp[size_t(arows)])); C.v.assign(size_t(C.rp[size_t(arows)]), 0.0);
    std::fill(mark.begin(), mark.end(), -1);
    std::vector<std::pair<int, double>> row;
    for (int i = 0; i < arows; ++i) {
        const int start = C.rp[size_t(i)];
        int pos = start;
        for (int p = A.rp[size_t(i)]; p < A.rp[size_t(i) + 1]; ++p) {
            const double a = A.v[size_t(p)];
            for (int q = brp[size_t(A.ci[size_t(p)])]; q < brp[size_t(A.ci[size_t(p)]) + 1]; ++q) {
                const int j = bci[size_t(q)];
                if (mark[size_t(j)] < start) { mark[size_t(j)] = pos; C.ci[size_t(pos)] = j; C.v[size_t(pos)] = a * bv[size_t(q)]; ++pos; }
                else C.v[size_t(mark[size_t(j)])] += a * bv[size_t(q)];
            }
        }
        row.clear();
        for (int p = start; p < pos; ++p) row.push_back({C.ci[size_t(p)], C.v[size_t(p)]});
        std::sort(row.begin(), row.end());
        for (int p = start; p < pos; ++p) { C.ci[size_t(p)] = row[size_t(p - start)].first; C.v[size_t(p)] = row[size_t(p - start)].second; }
    }
    return C;
}

// every row gets a diagonal entry (inserted as 0 before the first larger column), then |d| < 1e-300 -> 1 (amg.cpp:70-96)
static Csr withDiagonal(Csr A) {
    bool missing = false;
    for (int i = 0; i < A.n && !missing; ++i) {
        bool f = false;
        for (int p = A.rp[size_t(i)]; p < A.rp[size_t(i) + 1]; ++p) if (A.ci[size_t(p)] == i) { f = true; break; }
        if (!f) missing = true;
    }
    if (missing) {
        Csr B; B.n = A.n; B.rp.assign(size_t(A.n) + 1, 0);
        for (int i = 0; i < A.n; ++i) {
            bool f = false;
            for (int p = A.rp[size_t(i)]; p < A.rp[size_t(i) + 1]; ++p) {
                if (!f && A.ci[size_t(p)] > i) { B.ci.push_back(i); B.v.push_back(0.0); f = true; }
                if (A.ci[size_t(p)] == i) f = true;
                B.ci.push_back(A.ci[size_t(p)]); B.v.push_back(A.v[size_t(p)]);
            }
            if (!f) { B.ci.push_back(i); B.v.push_back(0.0); }
            B.rp[size_t(i) + 1] = int(B.ci.size());
        }
        A = std::move(B);
    }
    findDiag(A);
    for (int i = 0; i < A.n; ++i) if (std::fabs(A.v[size_t(A.diag[size_t(i)])]) < 1e-300) A.v[size_t(A.diag[size_t(i)])] = 1.0;
    return A;
}

struct Amg {
    struct Level {
        Csr A;
        std::vector<int> agg;                // fine row -> aggregate (plain aggregation: P = 1.0 at (i, agg[i]))
        int nc = 0;
        mutable std::vector<double> x, b, r;
    };
    std::vector<Level> lv;
    std::vector<double> lu; std::vector<int> piv;

    static void gaussSeidel(const Csr& A, const double* b, double* x, bool forward) {   // amg.cpp:98-113
        const int n = A.n;
        for (int t = 0; t < n; ++t) {
            const int i = forward ? t : n - 1 - t;
            double s = b[i];
            for (int p = A.rp[size_t(i)]; p < A.rp[size_t(i) + 1]; ++p) if (A.ci[size_t(p)] != i) s -= A.v[size_t(p)] * x[A.ci[size_t(p)]];
            x[i] = s / A.v[size_t(A.diag[size_t(i)])];
        }
    }
    static void csrSpmv(const Csr& A, const double* x, double* y) {
        for (int i = 0; i < A.n; ++i) { double s = 0; for (int p = A.rp[size_t(i)]; p < A.rp[size_t(i) + 1]; ++p) s += A.v[size_t(p)] * x[A.ci[size_t(p)]]; y[i] = s; }
    }
    // Galerkin coarse operator R (A P) for plain aggregation, as the generic sparse products (amg.cpp:232, 265)
    static Csr galerkin(const Csr& A, const std::vector<int>& agg, int nc) {
        const int n = A.n;
        std::vector<int> prp(size_t(n) + 1), pci(static_cast<size_t>(n)); std::vector<double> pv(size_t(n), 1.0);
        for (int i = 0; i <= n; ++i) prp[size_t(i)] = i;
        for (int i = 0; i < n; ++i) pci[size_t(i)] = agg[size_t(i)];
        Csr AP = sparseProduct(A, n, prp, pci, pv, nc, n);
        // R = P^T: rows = aggregates, fine rows ascending (amg.cpp:25-34)
        Csr R; R.n = nc; R.rp.assign(size_t(nc) + 1, 0);
        for (int i = 0; i < n; ++i) R.rp[size_t(agg[size_t(i)]) + 1]++;
        for (int a = 0; a < nc; ++a) R.rp[size_t(a) + 1] += R.rp[size_t(a)];
        R.ci.resize(size_t(n)); R.v.assign(size_t(n), 1.0);
        std::vector<int> fill(R.rp.begin(), R.rp.end() - 1);
        for (int i = 0; i < n; ++i) R.ci[size_t(fill[size_t(agg[size_t(i)])]++)] = i;
        Csr C = sparseProduct(R, n, AP.rp, AP.ci, AP.v, nc, nc);
        return withDiagonal(std::move(C));
    }
    void setup(Csr A0) {                                                               // amg.cpp:137-252
        lv.clear();
        lv.emplace_back();
        findDiag(A0);
        bool noDiag = false;
        for (int i = 0; i < A0.n; ++i) if (A0.diag[size_t(i)] < 0) noDiag = true;
        if (noDiag) A0 = withDiagonal(std::move(A0));
        for (int i = 0; i < A0.n; ++i) if (std::fabs(A0.v[size_t(A0.diag[size_t(i)])]) < 1e-300) A0.v[size_t(A0.diag[size_t(i)])] = 1.0;
        lv[0].A = std::move(A0);
        while (true) {
            Level& L = lv.back();
            const Csr& A = L.A;
            const int n = A.n;
            if (n <= 400 || lv.size() >= 14) break;
            // strength: |a_ij| >= 0.08 sqrt|d_i d_j| or >= 0.25 max_k |a_ik|
            std::vector<int> sp(size_t(n) + 1, 0), sc;
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
            // aggregation: pass 1 roots with free strong neighbourhoods, pass 2 join the strongest neighbouring
            // pass-1 aggregate, pass 3 leftovers (with their unaggregated strong neighbours) form new aggregates
            std::vector<int> agg(size_t(n), -1);
            int na = 0;
            for (int i = 0; i < n; ++i) {
                if (agg[size_t(i)] >= 0 || sp[size_t(i)] == sp[size_t(i) + 1]) continue;
                bool isFree = true;
                for (int p = sp[size_t(i)]; p < sp[size_t(i) + 1]; ++p) if (agg[size_t(sc[size_t(p)])] >= 0) { isFree = false; break; }
                if (!isFree) continue;
                agg[size_t(i)] = na;
                for (int p = sp[size_t(i)]; p < sp[size_t(i) + 1]; ++p) agg[size_t(sc[size_t(p)])] = na;
                ++na;
            }
            const std::vector<int> pass1 = agg;
            for (int i = 0; i < n; ++i) {
                if (agg[size_t(i)] >= 0) continue;
                double best = 0; int ba = -1;
                for (int p = A.rp[size_t(i)]; p < A.rp[size_t(i) + 1]; ++p) {
                    const int j = A.ci[size_t(p)];
                    if (j != i && pass1[size_t(j)] >= 0 && std::fabs(A.v[size_t(p)]) > best) { best = std::fabs(A.v[size_t(p)]); ba = pass1[size_t(j)]; }
                }
                if (ba >= 0 && sp[size_t(i)] != sp[size_t(i) + 1]) agg[size_t(i)] = ba;
            }
            for (int i = 0; i < n; ++i) {
                if (agg[size_t(i)] >= 0) continue;
                agg[size_t(i)] = na;
                for (int p = sp[size_t(i)]; p < sp[size_t(i) + 1]; ++p) if (agg[size_t(sc[size_t(p)])] < 0) agg[size_t(sc[size_t(p)])] = na;
                ++na;
            }
            if (na >= int(0.85 * n) || na < 1) break;
            L.agg = agg; L.nc = na;
            Level next; next.A = galerkin(L.A, L.agg, na);
            lv.push_back(std::move(next));
        }
        for (auto& L : lv) { L.x.assign(size_t(L.A.n), 0.0); L.b.assign(size_t(L.A.n), 0.0); L.r.assign(size_t(L.A.n), 0.0); }
        factorCoarsest();
    }
    bool refresh(const Csr& A0) {                                                      // amg.cpp:257-271
        if (lv.empty() || A0.n != lv[0].A.n || A0.ci != lv[0].A.ci) return false;
        lv[0].A.v = A0.v;
        for (int i = 0; i < lv[0].A.n; ++i) if (std::fabs(lv[0].A.v[size_t(lv[0].A.diag[size_t(i)])]) < 1e-300) lv[0].A.v[size_t(lv[0].A.diag[size_t(i)])] = 1.0;
        for (size_t l = 0; l + 1 < lv.size(); ++l) {
            Csr C = galerkin(lv[l].A, lv[l].agg, lv[l].nc);
            if (C.ci != lv[l + 1].A.ci) return false;
            lv[l + 1].A.v = C.v;
        }
        factorCoarsest();
        return true;
    }
    void factorCoarsest() {                                                            // amg.cpp:273-295
        const Csr& C = lv.back().A;
        lu.clear(); piv.clear();
        if (C.n > 800) return;
        const int m = C.n;
        lu.assign(size_t(m) * m, 0.0); piv.resize(size_t(m));
        for (int i = 0; i < m; ++i) for (int p = C.rp[size_t(i)]; p < C.rp[size_t(i) + 1]; ++p) lu[size_t(i) * m + size_t(C.ci[size_t(p)])] = C.v[size_t(p)];
        for (int k = 0; k < m; ++k) {
            int pr = k; double mx = std::fabs(lu[size_t(k) * m + k]);
            for (int i = k + 1; i < m; ++i) if (std::fabs(lu[size_t(i) * m + k]) > mx) { mx = std::fabs(lu[size_t(i) * m + k]); pr = i; }
            piv[size_t(k)] = pr;
            if (pr != k) for (int j = 0; j < m; ++j) std::swap(lu[size_t(k) * m + j], lu[size_t(pr) * m + j]);
            double& dk = lu[size_t(k) * m + k];
            if (std::fabs(dk) < 1e-300) dk = 1e-300;
            for (int i = k + 1; i < m; ++i) {
                const double f = lu[size_t(i) * m + k] / dk;
                lu[size_t(i) * m + k] = f;
                if (f != 0) for (int j = k + 1; j < m; ++j) lu[size_t(i) * m + j] -= f * lu[size_t(k) * m + j];
            }
        }
    }
    void cycle(size_t l) const {                                                       // amg.cpp:303-329
        const Level& L = lv[l];
        const int n = L.A.n;
        if (l + 1 == lv.size()) {
            if (!lu.empty()) {
                std::vector<double>& x = L.x;
                x = L.b;
                for (int k = 0; k < n; ++k) if (piv[size_t(k)] != k) std::swap(x[size_t(k)], x[size_t(piv[size_t(k)])]);
                for (int i = 0; i < n; ++i) { double s = x[size_t(i)]; for (int j = 0; j < i; ++j) s -= lu[size_t(i) * n + j] * x[size_t(j)]; x[size_t(i)] = s; }
                for (int i = n - 1; i >= 0; --i) { double s = x[size_t(i)]; for (int j = i + 1; j < n; ++j) s -= lu[size_t(i) * n + j] * x[size_t(j)]; x[size_t(i)] = s / lu[size_t(i) * n + i]; }
            } else {
                std::fill(L.x.begin(), L.x.end(), 0.0);
                for (int s = 0; s < 20; ++s) { gaussSeidel(L.A, L.b.data(), L.x.data(), true); gaussSeidel(L.A, L.b.data(), L.x.data(), false); }
            }
            return;
        }
        std::fill(L.x.begin(), L.x.end(), 0.0);
        gaussSeidel(L.A, L.b.data(), L.x.data(), true);
        csrSpmv(L.A, L.x.data(), L.r.data());
        for (int i = 0; i < n; ++i) L.r[size_t(i)] = L.b[size_t(i)] - L.r[size_t(i)];
        const Level& C = lv[l + 1];
        // restriction: b_c[a] = sum over fine rows of a, ascending
        std::fill(C.b.begin(), C.b.end(), 0.0);
        {
            std::vector<double> acc(size_t(L.nc), 0.0);
            for (int i = 0; i < n; ++i) acc[size_t(L.agg[size_t(i)])] += L.r[size_t(i)];
            for (int a = 0; a < L.nc; ++a) C.b[size_t(a)] = acc[size_t(a)];
        }
        cycle(l + 1);
        for (int i = 0; i < n; ++i) L.x[size_t(i)] += C.x[size_t(L.agg[size_t(i)])];
        gaussSeidel(L.A, L.b.data(), L.x.data(), false);
    }
    void apply(const double* r, double* z) const {
        if (lv.empty()) return;
        std::copy(r, r + lv[0].A.n, lv[0].b.begin());
        cycle(0);
        std::copy(lv[0].x.begin(), lv[0].x.end(), z);
    }
};

// ---- BlockCPR (bsr.cpp:993-1084), quasi-IMPES weights (no accumulation blocks)
struct BlockCpr {
    Bsr S;
    std::vector<double> w;
    Amg amg;
    Bilu0 ilu;
    // acc (optional, n nb x nb blocks): true-IMPES (bsr.cpp:1005-1021): the accumulation block B is inverted and,
    // if its 1-norm condition estimate |B|_1 |B^-1|_1 is at most 1e10 (or nb = 1), its inverse replaces D^-1; a
    // singular scalar block is used as it is; otherwise (a fallback) the quasi-IMPES diagonal block
    std::vector<double> acc;
    int fallbacks = 0;
    void weights() {
        const int n = S.n, b = S.nb, b2 = b * b;
        w.assign(size_t(n) * b, 1.0);
        fallbacks = 0;
        for (int i = 0; i < n; ++i) {
            double D[kMaxB * kMaxB];
            std::memcpy(D, S.blk(S.diag[size_t(i)]), sizeof(double) * size_t(b2));
            double* wi = w.data() + size_t(i) * b;
            bool inverted = false;
            if (!acc.empty()) {
                const double* B = acc.data() + size_t(i) * b2;
                double inv[kMaxB * kMaxB];
                std::memcpy(inv, B, sizeof(double) * size_t(b2));
                if (invertBlock(inv, b)) {
                    double n1 = 0, ni = 0;
                    for (int j = 0; j < b; ++j) {
                        double c = 0, d = 0;
                        for (int r = 0; r < b; ++r) { c += std::fabs(B[r * b + j]); d += std::fabs(inv[r * b + j]); }
                        n1 = std::max(n1, c); ni = std::max(ni, d);
                    }
                    if (b == 1 || n1 * ni <= 1e10) { std::memcpy(D, inv, sizeof(double) * size_t(b2)); inverted = true; }
                    else ++fallbacks;
                } else if (b == 1) std::memcpy(D, B, sizeof(double) * size_t(b2));
                else ++fallbacks;
            }
            if (inverted || invertBlock(D, b)) { for (int e = 0; e < b; ++e) wi[e] = D[e]; }          // first row of D^-1
            else {
                double s = 0; for (int e = 0; e < b; ++e) s += S.blk(S.diag[size_t(i)])[e * b];
                for (int e = 0; e < b; ++e) wi[e] = std::fabs(s) > 1e-300 ? 1.0 / s : 1.0;
            }
            double wm = 0; for (int e = 0; e < b; ++e) wm = std::max(wm, std::fabs(wi[e]));
            if (!(wm > 0) || !std::isfinite(wm)) for (int e = 0; e < b; ++e) wi[e] = 1.0;
        }
    }
    Csr pressure() const {                       // Ap_ij = sum_e w_ie S_ij[e][0] on the pattern of S
        Csr Ap; Ap.n = S.n; Ap.rp = S.rp; Ap.ci = S.ci; Ap.v.assign(S.ci.size(), 0.0);
        const int b = S.nb;
        for (int i = 0; i < S.n; ++i)
            for (int p = S.rp[size_t(i)]; p < S.rp[size_t(i) + 1]; ++p) {
                double s = 0; for (int e = 0; e < b; ++e) s += w[size_t(i) * b + e] * S.blk(p)[e * b];
                Ap.v[size_t(p)] = s;
            }
        findDiag(Ap);
        return Ap;
    }
    bool lean = false;               // C45: the BILU0 stage factorised lean (setLean, bsr.cpp:1055)
    void setup(const Bsr& A, bool reuse) {
        S = A;
        weights();
        Csr Ap = pressure();
        if (!(reuse && amg.refresh(Ap))) amg.setup(std::move(Ap));
        ilu.mode = lean ? 4 : 1;
        ilu.factor(S);
    }
    void apply(const double* r, double* z) const {
        const int n = S.n, b = S.nb;
        std::vector<double> rp(static_cast<size_t>(n)), xp(static_cast<size_t>(n)), t(static_cast<size_t>(n) * b);
        for (int i = 0; i < n; ++i) { double s = 0; for (int e = 0; e < b; ++e) s += w[size_t(i) * b + e] * r[size_t(i) * b + e]; rp[size_t(i)] = s; }
        amg.apply(rp.data(), xp.data());
        for (int i = 0; i < n; ++i) {
            double acc[kMaxB];
            for (int e = 0; e < b; ++e) acc[e] = r[size_t(i) * b + e];
            for (int p = S.rp[size_t(i)]; p < S.rp[size_t(i) + 1]; ++p) {
                const double xk = xp[size_t(S.ci[size_t(p)])];
                for (int e = 0; e < b; ++e) acc[e] -= S.blk(p)[e * b] * xk;
            }
            for (int e = 0; e < b; ++e) t[size_t(i) * b + e] = acc[e];
        }
        ilu.apply(t.data(), z);
        for (int i = 0; i < n; ++i) z[size_t(i) * b] += xp[size_t(i)];
    }
};

// ---- DeviceCprAmg (bsr.h:158-179, build bsr.cpp:1235-1317, apply :1319-1371): the reference's device CPR in float32
// - quasi-IMPES restriction W, the host AMG hierarchy with damped Jacobi (omega 0.6667f, one pre / one post step), the
// explicit inverse of the coarsest operator (Gauss-Jordan with partial pivoting in double, m <= 1500, else 30 Jacobi
// sweeps), E = injection into the pressure slot, block-Jacobi second stage z = z1 + D_b^-1 (r - A z1)
struct DevCprAmg {
    struct FCsr { int rows = 0; std::vector<int> rp, ci; std::vector<float> v;
        void spmv(const float* x, float* y) const {                                    // float accumulation (bsr.cpp:1218-1224)
            for (int i = 0; i < rows; ++i) { float acc = 0; for (int k = rp[size_t(i)]; k < rp[size_t(i) + 1]; ++k) acc += v[size_t(k)] * x[size_t(ci[size_t(k)])]; y[i] = acc; }
        }
    };
    struct Lvl { FCsr A, P, R; std::vector<float> dinv; };
    Bsr S;
    std::vector<float> W, Av, bjInv;
    std::vector<Lvl> lv;
    FCsr coarseInv;
    float omega = 0.6667f;
    bool build(const Bsr& A) {
        BlockCpr C;
        C.S = A;
        C.weights();
        C.amg.setup(C.pressure());
        S = A;
        const int n = S.n, nb = S.nb, b2 = nb * nb;
        W.assign(C.w.begin(), C.w.end());
        Av.assign(S.v.begin(), S.v.end());
        const size_t L = C.amg.lv.size();
        lv.assign(L, Lvl{});
        for (size_t l = 0; l < L; ++l) {
            const Csr& H = C.amg.lv[l].A;
            Lvl& V = lv[l];
            V.A.rows = H.n; V.A.rp = H.rp; V.A.ci = H.ci; V.A.v.assign(H.v.begin(), H.v.end());
            V.dinv.resize(size_t(H.n));
            for (int i = 0; i < H.n; ++i) {
                double d = 0;
                for (int k = H.rp[size_t(i)]; k < H.rp[size_t(i) + 1]; ++k) if (H.ci[size_t(k)] == i) d = H.v[size_t(k)];
                V.dinv[size_t(i)] = std::fabs(d) > 1e-300 ? float(1.0 / d) : 0.0f;
            }
            if (l + 1 < L) {
                const auto& agg = C.amg.lv[l].agg;
                const int nc = C.amg.lv[l].nc;
                V.P.rows = H.n; V.P.rp.resize(size_t(H.n) + 1); V.P.ci.resize(size_t(H.n)); V.P.v.assign(size_t(H.n), 1.0f);
                for (int i = 0; i <= H.n; ++i) V.P.rp[size_t(i)] = i;
                for (int i = 0; i < H.n; ++i) V.P.ci[size_t(i)] = agg[size_t(i)];
                V.R.rows = nc; V.R.rp.assign(size_t(nc) + 1, 0);
                for (int i = 0; i < H.n; ++i) V.R.rp[size_t(agg[size_t(i)]) + 1]++;
                for (int a = 0; a < nc; ++a) V.R.rp[size_t(a) + 1] += V.R.rp[size_t(a)];
                V.R.ci.resize(size_t(H.n)); V.R.v.assign(size_t(H.n), 1.0f);
                std::vector<int> fill(V.R.rp.begin(), V.R.rp.end() - 1);
                for (int i = 0; i < H.n; ++i) V.R.ci[size_t(fill[size_t(agg[size_t(i)])]++)] = i;
            }
        }
        const Csr& Cc = C.amg.lv.back().A;
        const int m = Cc.n;
        coarseInv = FCsr{};
        if (m > 0 && m <= 1500) {
            std::vector<double> M(size_t(m) * 2 * m, 0.0);
            for (int i = 0; i < m; ++i) {
                for (int k = Cc.rp[size_t(i)]; k < Cc.rp[size_t(i) + 1]; ++k) M[size_t(i) * 2 * m + size_t(Cc.ci[size_t(k)])] += Cc.v[size_t(k)];
                M[size_t(i) * 2 * m + size_t(m + i)] = 1.0;
            }
            bool ok = true;
            for (int k = 0; k < m && ok; ++k) {
                int pr = k;
                for (int i = k + 1; i < m; ++i) if (std::fabs(M[size_t(i) * 2 * m + k]) > std::fabs(M[size_t(pr) * 2 * m + k])) pr = i;
                if (std::fabs(M[size_t(pr) * 2 * m + k]) < 1e-300) { ok = false; break; }
                if (pr != k) for (int j = 0; j < 2 * m; ++j) std::swap(M[size_t(k) * 2 * m + j], M[size_t(pr) * 2 * m + j]);
                const double piv = 1.0 / M[size_t(k) * 2 * m + k];
                for (int j = 0; j < 2 * m; ++j) M[size_t(k) * 2 * m + j] *= piv;
                for (int i = 0; i < m; ++i) {
                    if (i == k) continue;
                    const double f = M[size_t(i) * 2 * m + k];
                    if (f == 0) continue;
                    for (int j = 0; j < 2 * m; ++j) M[size_t(i) * 2 * m + j] -= f * M[size_t(k) * 2 * m + j];
                }
            }
            if (ok) {
                coarseInv.rows = m; coarseInv.rp.resize(size_t(m) + 1); coarseInv.ci.resize(size_t(m) * m); coarseInv.v.resize(size_t(m) * m);
                for (int i = 0; i < m; ++i) {
                    coarseInv.rp[size_t(i)] = i * m;
                    for (int j = 0; j < m; ++j) { coarseInv.ci[size_t(i) * m + j] = j; coarseInv.v[size_t(i) * m + j] = float(M[size_t(i) * 2 * m + size_t(m + j)]); }
                }
                coarseInv.rp[size_t(m)] = m * m;
            }
        }
        bjInv.resize(size_t(n) * b2);
        for (int i = 0; i < n; ++i) {
            double blk[kMaxB * kMaxB];
            const double* D = S.blk(S.diag[size_t(i)]);
            std::memcpy(blk, D, sizeof(double) * size_t(b2));
            if (!invertBlock(blk, nb)) { for (int e = 0; e < b2; ++e) blk[e] = 0; for (int e = 0; e < nb; ++e) blk[e * nb + e] = 1.0 / std::max(1e-30, std::fabs(D[e * nb + e])); }
            for (int e = 0; e < b2; ++e) bjInv[size_t(i) * b2 + e] = float(blk[e]);
        }
        return true;
    }
    void apply(const float* r, float* z) const {
        const int n = S.n, nb = S.nb, b2 = nb * nb;
        const size_t L = lv.size();
        std::vector<std::vector<float>> x(L), b(L), t(L);
        for (size_t l = 0; l < L; ++l) { x[l].assign(size_t(lv[l].A.rows), 0.f); b[l] = x[l]; t[l] = x[l]; }
        for (int i = 0; i < n; ++i) { float acc = 0; for (int e = 0; e < nb; ++e) acc += W[size_t(i) * nb + e] * r[size_t(i) * nb + e]; b[0][size_t(i)] = acc; }
        for (size_t l = 0; l + 1 < L; ++l) {
            const Lvl& V = lv[l];
            const size_t nl = x[l].size();
            for (size_t i = 0; i < nl; ++i) x[l][i] = omega * V.dinv[i] * b[l][i];
            V.A.spmv(x[l].data(), t[l].data());
            for (size_t i = 0; i < nl; ++i) t[l][i] = b[l][i] - t[l][i];
            V.R.spmv(t[l].data(), b[l + 1].data());
        }
        {
            const Lvl& V = lv[L - 1];
            const size_t nl = x[L - 1].size();
            if (!coarseInv.rp.empty()) coarseInv.spmv(b[L - 1].data(), x[L - 1].data());
            else {
                for (size_t i = 0; i < nl; ++i) x[L - 1][i] = omega * V.dinv[i] * b[L - 1][i];
                for (int s2 = 0; s2 < 30; ++s2) {
                    V.A.spmv(x[L - 1].data(), t[L - 1].data());
                    for (size_t i = 0; i < nl; ++i) x[L - 1][i] += omega * V.dinv[i] * (b[L - 1][i] - t[L - 1][i]);
                }
            }
        }
        for (size_t l = L - 1; l-- > 0;) {
            const Lvl& V = lv[l];
            const size_t nl = x[l].size();
            V.P.spmv(x[l + 1].data(), t[l].data());
            for (size_t i = 0; i < nl; ++i) x[l][i] += t[l][i];
            V.A.spmv(x[l].data(), t[l].data());
            for (size_t i = 0; i < nl; ++i) x[l][i] += omega * V.dinv[i] * (b[l][i] - t[l][i]);
        }
        std::vector<float> z1(size_t(n) * nb, 0.f), tt(size_t(n) * nb);
        for (int i = 0; i < n; ++i) z1[size_t(i) * nb] = 1.0f * x[0][size_t(i)];
        for (int i = 0; i < n; ++i) {
            float acc[kMaxB] = {};
            for (int k = S.rp[size_t(i)]; k < S.rp[size_t(i) + 1]; ++k) {
                const float* B = Av.data() + size_t(k) * b2;
                const float* xk = z1.data() + size_t(S.ci[size_t(k)]) * nb;
                for (int e = 0; e < nb; ++e) for (int f = 0; f < nb; ++f) acc[e] += B[e * nb + f] * xk[f];
            }
            for (int e = 0; e < nb; ++e) tt[size_t(i) * nb + e] = r[size_t(i) * nb + e] - acc[e];
        }
        for (int i = 0; i < n; ++i) {
            const float* D = bjInv.data() + size_t(i) * b2;
            for (int e = 0; e < nb; ++e) {
                float s2 = 0;
                for (int f = 0; f < nb; ++f) s2 += D[e * nb + f] * tt[size_t(i) * nb + f];
                z[size_t(i) * nb + e] = z1[size_t(i) * nb + e] + s2;
            }
        }
    }
};

// ---- Krylov (linsolve.cpp:147-273)
struct Prec { virtual ~Prec() = default; virtual void apply(const double* r, double* z) const = 0; };
template <typename P> struct PrecOf : Prec { const P& p; explicit PrecOf(const P& x) : p(x) {} void apply(const double* r, double* z) const override { p.apply(r, z); } };

static double dot(const std::vector<double>& a, const std::vector<double>& b) { double s = 0; for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i]; return s; }

static void bicgstab(const Bsr& A, const Prec& M, const std::vector<double>& b, std::vector<double>& x, double tol, int maxIt, int& its,
                     double& rel, bool& conv) {
    const int n = A.n * A.nb;
    its = 0; rel = 0; conv = false;
    x.assign(size_t(n), 0.0);
    std::vector<double> r(b), rh(b), p(size_t(n), 0), v(size_t(n), 0), s(static_cast<size_t>(n)), t(static_cast<size_t>(n)), ph(static_cast<size_t>(n)), sh(static_cast<size_t>(n));
    const double bn = std::sqrt(dot(b, b));
    if (bn == 0) { conv = true; return; }
    double rho = 1, alpha = 1, omega = 1;
    for (int it = 1; it <= maxIt; ++it) {
        const double rhoN = dot(rh, r);
        if (std::fabs(rhoN) < 1e-300) break;
        const double beta = (rhoN / rho) * (alpha / omega);
        rho = rhoN;
        for (int i = 0; i < n; ++i) p[size_t(i)] = r[size_t(i)] + beta * (p[size_t(i)] - omega * v[size_t(i)]);
        M.apply(p.data(), ph.data());
        spmv(A, ph.data(), v.data());
        const double rv = dot(rh, v);
        if (std::fabs(rv) < 1e-300) break;
        alpha = rho / rv;
        for (int i = 0; i < n; ++i) s[size_t(i)] = r[size_t(i)] - alpha * v[size_t(i)];
        const double sn = std::sqrt(dot(s, s));
        if (sn / bn < tol) {
            for (int i = 0; i < n; ++i) x[size_t(i)] += alpha * ph[size_t(i)];
            conv = true; its = it; rel = sn / bn; return;
        }
        M.apply(s.data(), sh.data());
        spmv(A, sh.data(), t.data());
        const double tt = dot(t, t);
        omega = tt > 0 ? dot(t, s) / tt : 0;
        for (int i = 0; i < n; ++i) { x[size_t(i)] += alpha * ph[size_t(i)] + omega * sh[size_t(i)]; r[size_t(i)] = s[size_t(i)] - omega * t[size_t(i)]; }
        const double rn = std::sqrt(dot(r, r));
        its = it; rel = rn / bn;
        if (rn / bn < tol) { conv = true; return; }
        if (omega == 0) break;
    }
}

static void gmres(const Bsr& A, const Prec& M, const std::vector<double>& b, std::vector<double>& x, double tol, int maxIt, int m, int& its,
                  double& rel, bool& conv) {
    const int n = A.n * A.nb;
    its = 0; rel = 0; conv = false;
    x.assign(size_t(n), 0.0);
    const double bn = std::sqrt(dot(b, b));
    if (bn == 0) { conv = true; return; }
    std::vector<std::vector<double>> V(static_cast<size_t>(m) + 1, std::vector<double>(static_cast<size_t>(n))), Z(static_cast<size_t>(m), std::vector<double>(static_cast<size_t>(n)));
    std::vector<double> H(static_cast<size_t>(m + 1) * m), cs(static_cast<size_t>(m)), sn(static_cast<size_t>(m)), g(static_cast<size_t>(m) + 1), r(static_cast<size_t>(n)), w(static_cast<size_t>(n));
    auto h = [&](int i, int j) -> double& { return H[size_t(i) * m + j]; };
    int total = 0;
    while (total < maxIt) {
        spmv(A, x.data(), r.data());
        for (int i = 0; i < n; ++i) r[size_t(i)] = b[size_t(i)] - r[size_t(i)];
        const double beta = std::sqrt(dot(r, r));
        rel = beta / bn;
        if (beta / bn < tol) { conv = true; its = total; return; }
        for (int i = 0; i < n; ++i) V[0][size_t(i)] = r[size_t(i)] / beta;
        std::fill(g.begin(), g.end(), 0.0); g[0] = beta;
        int j = 0;
        bool done = false;
        while (j < m && total < maxIt) {
            M.apply(V[size_t(j)].data(), Z[size_t(j)].data());
            spmv(A, Z[size_t(j)].data(), w.data());
            for (int i = 0; i <= j; ++i) {                                  // modified Gram-Schmidt
                const double hij = dot(w, V[size_t(i)]);
                h(i, j) = hij;
                for (int q = 0; q < n; ++q) w[size_t(q)] -= hij * V[size_t(i)][size_t(q)];
            }
            const double hn = std::sqrt(dot(w, w));
            h(j + 1, j) = hn;
            if (hn > 0) for (int q = 0; q < n; ++q) V[size_t(j) + 1][size_t(q)] = w[size_t(q)] / hn;
            for (int i = 0; i < j; ++i) {                                   // earlier Givens rotations
                const double t = cs[size_t(i)] * h(i, j) + sn[size_t(i)] * h(i + 1, j);
                h(i + 1, j) = -sn[size_t(i)] * h(i, j) + cs[size_t(i)] * h(i + 1, j);
                h(i, j) = t;
            }
            const double den = std::hypot(h(j, j), h(j + 1, j));
            cs[size_t(j)] = den > 0 ? h(j, j) / den : 1; sn[size_t(j)] = den > 0 ? h(j + 1, j) / den : 0;
            h(j, j) = den; h(j + 1, j) = 0;
            g[size_t(j) + 1] = -sn[size_t(j)] * g[size_t(j)]; g[size_t(j)] = cs[size_t(j)] * g[size_t(j)];
            rel = std::fabs(g[size_t(j) + 1]) / bn;
            ++j; ++total;
            if (rel < tol) { done = true; break; }
        }
        std::vector<double> y(static_cast<size_t>(j));
        for (int i = j - 1; i >= 0; --i) {
            double s = g[size_t(i)];
            for (int k = i + 1; k < j; ++k) s -= h(i, k) * y[size_t(k)];
            y[size_t(i)] = s / h(i, i);
        }
        for (int i = 0; i < j; ++i) for (int q = 0; q < n; ++q) x[size_t(q)] += y[size_t(i)] * Z[size_t(i)][size_t(q)];
        if (done) { conv = true; its = total; return; }
    }
    its = total;
}

}  // namespace oracle

using namespace oracle;

// ---- the float32 device solve (CpuFloatBlockSolver::solve, bsr.cpp:1130-1215) and the FP64 refinement around it
// (refineWithDevice, bsr.cpp:1373-1396).  Float32 matrix, vectors and preconditioner: the SpMV accumulates each
// block row in float (entry by entry), the preconditioner is BILU0f (double factorisation, float factors and
// sweeps), the multicolour factors of P A P^T rounded to float and swept colour by colour in float, or DeviceCprAmg;
// dot products accumulate double products in double; every vector update is formed in double and rounded to float.
struct FloatDeviceSolve {
    enum Kind { kBilu0f = 0, kMcBilu0 = 1, kDevCpr = 2 };
    int kind = kBilu0f;
    Bsr A;
    std::vector<float> Af;
    Bilu0 ilu;                    // kBilu0f: mode 3
    McBilu0 mc;                   // kMcBilu0: double factors of P A P^T ...
    std::vector<float> mcF;       // ... and their float copy
    DevCprAmg dev;
    bool ready = false;
    bool setup(const Bsr& M, int k) {
        kind = k;
        A = M;
        Af.assign(A.v.begin(), A.v.end());
        if (kind == kBilu0f) { ilu.mode = 3; ilu.factor(A); }
        else if (kind == kMcBilu0) { mc.factor(A); mcF.assign(mc.ilu.LU.v.begin(), mc.ilu.LU.v.end()); }
        else if (!dev.build(A)) return false;
        ready = true;
        return true;
    }
    void spmv(const std::vector<float>& x, std::vector<float>& y) const {
        const int b = A.nb;
        const size_t b2 = size_t(b) * b;
        for (int i = 0; i < A.n; ++i) {
            float acc[kMaxB] = {};
            for (int p = A.rp[size_t(i)]; p < A.rp[size_t(i) + 1]; ++p) {
                const float* B = Af.data() + size_t(p) * b2;
                const float* xk = x.data() + size_t(A.ci[size_t(p)]) * b;
                for (int e = 0; e < b; ++e)
                    for (int f = 0; f < b; ++f) acc[e] += B[e * b + f] * xk[f];
            }
            for (int e = 0; e < b; ++e) y[size_t(i) * b + e] = acc[e];
        }
    }
    void precond(const std::vector<float>& r, std::vector<float>& z) const {
        const size_t N = r.size();
        if (kind == kDevCpr) { dev.apply(r.data(), z.data()); return; }
        if (kind == kBilu0f) {
            std::vector<double> rd(r.begin(), r.end()), zd(N);
            ilu.apply(rd.data(), zd.data());
            for (size_t e = 0; e < N; ++e) z[e] = float(zd[e]);
            return;
        }
        // colour order: gather, the float sweeps over the permuted factors (colours in order, rows ascending), scatter
        const Bsr& F = mc.ilu.LU;
        const int n = F.n, b = F.nb;
        const size_t b2 = size_t(b) * b;
        std::vector<float> y(N);
        for (int k = 0; k < n; ++k)
            for (int e = 0; e < b; ++e) y[size_t(k) * b + e] = r[size_t(mc.perm[size_t(k)]) * b + e];
        for (int i = 0; i < n; ++i) {
            float* yi = y.data() + size_t(i) * b;
            for (int p = F.rp[size_t(i)]; p < F.diag[size_t(i)]; ++p) {
                const float* a = mcF.data() + size_t(p) * b2;
                const float* yk = y.data() + size_t(F.ci[size_t(p)]) * b;
                for (int e = 0; e < b; ++e) { float t = 0; for (int f = 0; f < b; ++f) t += a[e * b + f] * yk[f]; yi[e] -= t; }
            }
        }
        for (int i = n - 1; i >= 0; --i) {
            float* yi = y.data() + size_t(i) * b;
            for (int p = F.diag[size_t(i)] + 1; p < F.rp[size_t(i) + 1]; ++p) {
                const float* a = mcF.data() + size_t(p) * b2;
                const float* yj = y.data() + size_t(F.ci[size_t(p)]) * b;
                for (int e = 0; e < b; ++e) { float t = 0; for (int f = 0; f < b; ++f) t += a[e * b + f] * yj[f]; yi[e] -= t; }
            }
            const float* d = mcF.data() + size_t(F.diag[size_t(i)]) * b2;
            float u[kMaxB];
            for (int e = 0; e < b; ++e) { float t = 0; for (int f = 0; f < b; ++f) t += d[e * b + f] * yi[f]; u[e] = t; }
            for (int e = 0; e < b; ++e) yi[e] = u[e];
        }
        for (int k = 0; k < n; ++k)
            for (int e = 0; e < b; ++e) z[size_t(mc.perm[size_t(k)]) * b + e] = y[size_t(k) * b + e];
    }
    static double fdot(const std::vector<float>& u, const std::vector<float>& v) {
        double acc = 0;
        for (size_t e = 0; e < u.size(); ++e) acc += double(u[e]) * v[e];
        return acc;
    }
    // false: a non-finite solution
    bool solve(const double* bd, double* xd, double tol, int maxIt, int& its) const {
        const size_t N = size_t(A.n) * A.nb;
        its = 0;
        std::vector<float> x(N, 0.f), r(N), rh(N), p(N, 0.f), v(N, 0.f), ph(N), sv(N), sh(N), t(N);
        for (size_t e = 0; e < N; ++e) r[e] = float(bd[e]);
        rh = r;
        const double bn = std::sqrt(fdot(r, r));
        std::fill(xd, xd + N, 0.0);
        if (!(bn > 0)) return true;
        double rho = 1, alpha = 1, omega = 1;
        for (int it = 1; it <= maxIt; ++it) {
            its = it;
            const double rhoN = fdot(rh, r);
            if (rhoN == 0 || !std::isfinite(rhoN)) break;
            const double beta = (rhoN / rho) * (alpha / omega);
            for (size_t e = 0; e < N; ++e) p[e] = float(r[e] + beta * (p[e] - omega * v[e]));
            precond(p, ph);
            spmv(ph, v);
            const double rv = fdot(rh, v);
            if (rv == 0 || !std::isfinite(rv)) break;
            alpha = rhoN / rv;
            for (size_t e = 0; e < N; ++e) sv[e] = float(r[e] - alpha * v[e]);
            if (std::sqrt(fdot(sv, sv)) <= tol * bn) {
                for (size_t e = 0; e < N; ++e) x[e] += float(alpha * ph[e]);
                break;
            }
            precond(sv, sh);
            spmv(sh, t);
            const double tt = fdot(t, t);
            omega = tt > 0 ? fdot(t, sv) / tt : 0;
            for (size_t e = 0; e < N; ++e) {
                x[e] += float(alpha * ph[e] + omega * sh[e]);
                r[e] = float(sv[e] - omega * t[e]);
            }
            rho = rhoN;
            if (std::sqrt(fdot(r, r)) <= tol * bn || omega == 0) break;
        }
        bool finite = true;
        for (size_t e = 0; e < N; ++e) { xd[e] = x[e]; finite = finite && std::isfinite(xd[e]); }
        return finite;
    }
    // FP64 refinement: x = sum of float corrections, r = b - A x in FP64, inner tolerance max(1e-4, min(1e-2, 10 tol))
    bool refine(const double* b, double* x, double tol, int maxOuter, int innerIt, int& its, double& rel) const {
        const size_t N = size_t(A.n) * A.nb;
        std::vector<double> r(b, b + N), d(N), Ax(N);
        std::fill(x, x + N, 0.0);
        its = 0;
        rel = 0;
        double bn = 0;
        for (size_t e = 0; e < N; ++e) bn += b[e] * b[e];
        bn = std::sqrt(bn);
        if (!(bn > 0)) return true;
        const double innerTol = std::max(1e-4, std::min(1e-2, tol * 10));
        for (int outer = 0; outer < maxOuter; ++outer) {
            double rn = 0;
            for (double q : r) rn += q * q;
            rel = std::sqrt(rn) / bn;
            if (rel <= tol) return true;
            int k = 0;
            if (!solve(r.data(), d.data(), innerTol, innerIt, k)) return false;
            its += k;
            for (size_t e = 0; e < N; ++e) x[e] += d[e];
            oracle::spmv(A, x, Ax.data());
            for (size_t e = 0; e < N; ++e) r[e] = b[e] - Ax[e];
        }
        double rn = 0;
        for (double q : r) rn += q * q;
        rel = std::sqrt(rn) / bn;
        return rel <= tol;
    }
};

extern "C" {

int chk_spmv(int n, int nb, const int* rp, const int* ci, const double* v, const double* x, double* y) {
    Bsr A = makeBsr(n, nb, rp, ci, v);
    spmv(A, x, y);
    return 0;
}

void* chk_bilu0_new(int n, int nb, const int* rp, const int* ci, const double* v, int mode) {
    auto* h = new Bilu0;
    h->mode = mode;
    h->factor(makeBsr(n, nb, rp, ci, v));
    return h;
}
int chk_bilu0_apply(void* h, const double* r, double* z) { static_cast<Bilu0*>(h)->apply(r, z); return 0; }
int chk_bilu0_regularised(void* h) { return static_cast<Bilu0*>(h)->nReg; }
void chk_bilu0_free(void* h) { delete static_cast<Bilu0*>(h); }
int chk_bilu0_factors(void* h, double* lu) {
    const auto* b = static_cast<Bilu0*>(h);
    std::copy(b->LU.v.begin(), b->LU.v.end(), lu);
    return 0;
}

void* chk_mcbilu0_new(int n, int nb, const int* rp, const int* ci, const double* v) {
    auto* h = new McBilu0;
    h->factor(makeBsr(n, nb, rp, ci, v));
    return h;
}
int chk_mcbilu0_apply(void* h, const double* r, double* z) { static_cast<McBilu0*>(h)->apply(r, z); return 0; }
int chk_mcbilu0_colours(void* h) { return static_cast<McBilu0*>(h)->colours; }
int chk_mcbilu0_colouring(void* h, int* colour, int* perm) {
    auto* m = static_cast<McBilu0*>(h);
    std::copy(m->colour.begin(), m->colour.end(), colour);
    std::copy(m->perm.begin(), m->perm.end(), perm);
    return m->colours;
}
void chk_mcbilu0_free(void* h) { delete static_cast<McBilu0*>(h); }

void* chk_cpr_new(int n, int nb, const int* rp, const int* ci, const double* v) {
    auto* h = new BlockCpr;
    h->setup(makeBsr(n, nb, rp, ci, v), false);
    return h;
}
void* chk_cpr_new_acc(int n, int nb, const int* rp, const int* ci, const double* v, const double* acc) {
    auto* h = new BlockCpr;
    h->acc.assign(acc, acc + size_t(n) * nb * nb);
    h->setup(makeBsr(n, nb, rp, ci, v), false);
    return h;
}
int chk_cpr_fallbacks(void* h) { return static_cast<BlockCpr*>(h)->fallbacks; }
int chk_cpr_refresh(void* h, const double* v) {
    auto* c = static_cast<BlockCpr*>(h);
    Bsr A = c->S;
    std::copy(v, v + A.v.size(), A.v.begin());
    c->setup(A, true);
    return 0;
}
int chk_cpr_apply(void* h, const double* r, double* z) { static_cast<BlockCpr*>(h)->apply(r, z); return 0; }
int chk_cpr_weights(void* h, double* w) { auto* c = static_cast<BlockCpr*>(h); std::copy(c->w.begin(), c->w.end(), w); return 0; }
int chk_cpr_nlevels(void* h) { return int(static_cast<BlockCpr*>(h)->amg.lv.size()); }
int chk_cpr_level(void* h, int l, int* n, int* nnz) {
    const auto& L = static_cast<BlockCpr*>(h)->amg.lv[size_t(l)];
    *n = L.A.n; *nnz = int(L.A.ci.size());
    return 0;
}
int chk_cpr_level_matrix(void* h, int l, int* rp, int* ci, double* v) {
    const auto& L = static_cast<BlockCpr*>(h)->amg.lv[size_t(l)];
    std::copy(L.A.rp.begin(), L.A.rp.end(), rp); std::copy(L.A.ci.begin(), L.A.ci.end(), ci); std::copy(L.A.v.begin(), L.A.v.end(), v);
    return 0;
}
int chk_cpr_aggregates(void* h, int l, int* agg) {
    const auto& L = static_cast<BlockCpr*>(h)->amg.lv[size_t(l)];
    if (L.agg.empty()) return -1;
    std::copy(L.agg.begin(), L.agg.end(), agg);
    return L.nc;
}
int chk_cpr_amg_apply(void* h, const double* r, double* z) { static_cast<BlockCpr*>(h)->amg.apply(r, z); return 0; }
void chk_cpr_free(void* h) { delete static_cast<BlockCpr*>(h); }

int chk_krylov(int n, int nb, const int* rp, const int* ci, const double* v, const double* b, double* x, int precond, int method,
               int restart, double tol, int maxit, int* its, double* relres) {
    Bsr A = makeBsr(n, nb, rp, ci, v);
    std::unique_ptr<Prec> M;
    Bilu0 ilu; McBilu0 mc; BlockCpr cpr;
    switch (precond) {
        case 1: case 2: case 3: ilu.mode = precond; ilu.factor(A); M.reset(new PrecOf<Bilu0>(ilu)); break;
        case 4: mc.factor(A); M.reset(new PrecOf<McBilu0>(mc)); break;
        case 5: cpr.setup(A, false); M.reset(new PrecOf<BlockCpr>(cpr)); break;
        case 6: ilu.mode = 4; ilu.factor(A); M.reset(new PrecOf<Bilu0>(ilu)); break;
        case 7: cpr.lean = true; cpr.setup(A, false); M.reset(new PrecOf<BlockCpr>(cpr)); break;
        default: return -1;
    }
    std::vector<double> bb(b, b + size_t(n) * nb), xx;
    bool conv = false;
    if (method == 0) bicgstab(A, *M, bb, xx, tol, maxit, *its, *relres, conv);
    else gmres(A, *M, bb, xx, tol, maxit, restart, *its, *relres, conv);
    std::copy(xx.begin(), xx.end(), x);
    return conv ? 1 : 0;
}

// the steady state of the Newton loop: BlockCPR::setup with reuseAmg (bsr.cpp:993-1059) after a first set-up
int chk_krylov_cpr_steady(int n, int nb, const int* rp, const int* ci, const double* v, const double* b, double* x, double tol,
                          int maxit, int* its, double* relres, double* secs) {
    Bsr A = makeBsr(n, nb, rp, ci, v);
    BlockCpr cpr;
    cpr.setup(A, false);
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    cpr.setup(A, true);
    const auto t1 = clk::now();
    PrecOf<BlockCpr> M(cpr);
    std::vector<double> bb(b, b + size_t(n) * nb), xx;
    bool conv = false;
    bicgstab(A, M, bb, xx, tol, maxit, *its, *relres, conv);
    const auto t2 = clk::now();
    std::copy(xx.begin(), xx.end(), x);
    secs[0] = std::chrono::duration<double>(t1 - t0).count();
    secs[1] = std::chrono::duration<double>(t2 - t1).count();
    return conv ? 1 : 0;
}

void* chk_devcpr_new(int n, int nb, const int* rp, const int* ci, const double* v) {
    auto* d = new DevCprAmg;
    if (!d->build(makeBsr(n, nb, rp, ci, v))) { delete d; return nullptr; }
    return d;
}
int chk_devcpr_apply(void* h, const double* r, double* z) {
    auto* d = static_cast<DevCprAmg*>(h);
    const size_t N = size_t(d->S.n) * d->S.nb;
    std::vector<float> rf(r, r + N), zf(N);
    d->apply(rf.data(), zf.data());
    for (size_t i = 0; i < N; ++i) z[i] = zf[i];
    return 0;
}
void chk_devcpr_free(void* h) { delete static_cast<DevCprAmg*>(h); }

int chk_float_solve(int n, int nb, const int* rp, const int* ci, const double* v, const double* b, double* x, int kind, double tol, int maxit,
                    int* its) {
    FloatDeviceSolve F;
    if (!F.setup(makeBsr(n, nb, rp, ci, v), kind)) return -1;
    return F.solve(b, x, tol, maxit, *its) ? 1 : 0;
}
int chk_refine_float(int n, int nb, const int* rp, const int* ci, const double* v, const double* b, double* x, int kind, double tol, int max_outer,
                     int inner_it, int* its, double* relres) {
    FloatDeviceSolve F;
    if (!F.setup(makeBsr(n, nb, rp, ci, v), kind)) return -1;
    return F.refine(b, x, tol, max_outer, inner_it, *its, *relres) ? 1 : 0;
}

}  // extern "C"
