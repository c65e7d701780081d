// Seeded synthetic reservoir Jacobians and the engine-side well Schur reduction.
//
// This is the workload side of the linear-solve path: it assembles the scalar CSR Jacobian a fully-implicit
// black-oil engine hands to its linear solver (cells in blocks of nv unknowns, pressure first, followed by one
// BHP unknown per well), then eliminates the BHP unknowns exactly (S = Acc - Acw Dww^-1 Awc) to obtain the
// block (BSR) system the solver sees, as the engine does before it calls its solver (ressim::WellSchur,
// engine/src/bsr.cpp:733-990).  The reduction is the rank-1-per-well plan of csrc/host/well_plan.hpp, whose per-entry
// order of operations makes the reduced system match the engine's bit for bit.
//
// All random numbers come from a counter-based SplitMix64 hash (seed, stream, index), so every value is
// independent of generation order and identical on every platform.  Two reference-test generators that use
// std::mt19937 + std::uniform_real_distribution (libstdc++) are provided for the reference's own test inputs.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <random>
#include <string>
#include <vector>

#include "../host/well_plan.hpp"

namespace {

thread_local std::string g_err;

// ------------------------------------------------------------------ counter-based RNG
inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
inline uint64_t hash4(uint64_t seed, uint64_t stream, uint64_t a, uint64_t b) {
    return mix64(mix64(mix64(seed ^ (stream * 0xD1342543DE82EF95ULL)) ^ a) ^ (b * 0x9E3779B97F4A7C15ULL));
}
inline double u01(uint64_t seed, uint64_t stream, uint64_t a, uint64_t b = 0) {
    return double(hash4(seed, stream, a, b) >> 11) * (1.0 / 9007199254740992.0);   // [0, 1)
}
inline double normal(uint64_t seed, uint64_t stream, uint64_t a) {
    const double u1 = 1.0 - u01(seed, stream, a, 0), u2 = u01(seed, stream, a, 1);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

struct Csr { int n = 0; std::vector<int> rowPtr, col; std::vector<double> val; };
struct Bsr { int n = 0, nb = 1; std::vector<int> rowPtr, col, diag; std::vector<double> val; };

void finalizeDiag(Bsr& S) {
    S.diag.assign(size_t(S.n), -1);
    for (int i = 0; i < S.n; ++i)
        for (int p = S.rowPtr[size_t(i)]; p < S.rowPtr[size_t(i) + 1]; ++p) if (S.col[size_t(p)] == i) { S.diag[size_t(i)] = p; break; }
}

// ------------------------------------------------------------------ well Schur complement (engine side)
// The reduction itself (pattern, values, right-hand side, recovery) is csrc/host/well_plan.hpp: one rank-1 update
// per well, in the engine's per-entry order.  Here it is applied to host values, with the engine's value rule for
// the pattern: a cell -> well entry that is exactly zero couples nothing (bsr.cpp:755-757).
struct WellReduction {
    wellplan::Plan plan;
    std::vector<double> dInv;
    std::vector<double> jval;      // J's values (the right-hand-side reduction and recovery read them)
    std::vector<int> jcol;
};

bool reduceWells(const Csr& J, int na, int nv, int nw, Bsr& S, WellReduction& W) {
    const double* v = J.val.data();
    std::string why = wellplan::split(J.n, J.rowPtr.data(), J.col.data(), na, nv, nw, [v](int p) { return v[p] != 0.0; }, W.plan);
    if (why.empty()) why = wellplan::wellInverses(W.plan, v, W.dInv);
    if (!why.empty()) { g_err = why; return false; }
    wellplan::pattern(J.n, J.rowPtr.data(), J.col.data(), W.plan);
    S = Bsr{};
    S.n = na; S.nb = nv;
    S.rowPtr = W.plan.sPtr;
    S.col = W.plan.sCol;
    S.val.assign(S.col.size() * size_t(nv) * nv, 0.0);
    wellplan::reduceValues(J.rowPtr.data(), J.col.data(), v, W.plan, W.dInv, S.val.data());
    finalizeDiag(S);
    W.jval = J.val;
    W.jcol = J.col;
    return true;
}

void reduceWellRhs(const WellReduction& W, const double* b, double* bc) { wellplan::reduceRhs(W.plan, W.jval.data(), W.dInv, b, bc); }
void recoverWells(const WellReduction& W, const double* b, const double* xc, double* x) {
    wellplan::recover(W.plan, W.jcol.data(), W.jval.data(), W.dInv, b, xc, x);
}

struct System {
    Bsr S; std::vector<double> bc;          // reduced block system and right-hand side
    Csr J; std::vector<double> b;           // full scalar system (cells + BHP rows)
    WellReduction W; bool hasFull = false;
    std::vector<double> perm;               // cell permeability (mD), for diagnostics
};

// Assemble a full scalar Jacobian from cell blocks (7-point pattern) and well couplings.
//   cellBlk(c, slot) gives the nb x nb block of neighbour slot; wells are perforation lists.
struct Perf { int cell; double wi; double m[4]; };

enum Stream : uint64_t { kLnK = 1, kPoro = 2, kMob = 3, kRhs = 4, kChan = 5, kLayer = 6, kMode = 7, kWellMob = 8, kWellPos = 9 };

// Permeability field.
//  model 0: iid lognormal, ln k ~ N(ln 100, sigma^2)
//  model 1: layered; the top nSmooth layers smooth (per-layer offset + a sum of four random cosine modes, unit
//           variance, scaled by sigma); the layers below channelised: a binary mask of three sinuous channels
//           covering ~20 % of the cells with a 1000:1 permeability contrast over a lognormal background.
void permeability(int nx, int ny, int nz, int model, double sigma, uint64_t seed, std::vector<double>& k) {
    const size_t n = size_t(nx) * ny * nz;
    k.resize(n);
    const double lnMean = std::log(100.0);
    if (model == 0) {
        for (size_t c = 0; c < n; ++c) k[c] = std::exp(lnMean + sigma * normal(seed, kLnK, c));
        return;
    }
    const int nSmooth = std::min(nz, int(std::lround(nz * 35.0 / 85.0)));
    // smooth layers
    double ka[4], kb[4], ph[4];
    for (int m = 0; m < 4; ++m) { ka[m] = 0.5 + 3.0 * u01(seed, kMode, m, 0); kb[m] = 0.5 + 3.0 * u01(seed, kMode, m, 1); ph[m] = 6.283185307179586 * u01(seed, kMode, m, 2); }
    for (int z = 0; z < nSmooth; ++z) {
        const double off = normal(seed, kLayer, uint64_t(z));
        for (int j = 0; j < ny; ++j) for (int i = 0; i < nx; ++i) {
            double g = 0;
            for (int m = 0; m < 4; ++m) g += std::cos(6.283185307179586 * (ka[m] * i / nx + kb[m] * j / ny) + ph[m] + 0.3 * z);
            g *= 0.7071067811865476;                        // four modes of variance 1/2 -> unit variance
            const double zz = 0.6 * off + 0.8 * g;
            k[(size_t(z) * ny + j) * nx + i] = std::exp(lnMean + sigma * zz);
        }
    }
    // channelised layers: 3 channels, each ~6.7 % of nx wide, sinuous along y
    const double lnBack = std::log(10.0), lnContrast = std::log(1000.0);
    for (int z = nSmooth; z < nz; ++z) {
        double x0[3], amp[3], lam[3], phs[3];
        for (int c = 0; c < 3; ++c) {
            x0[c] = nx * u01(seed, kChan, uint64_t(z) * 8 + c, 0);
            amp[c] = 0.05 * nx + 0.1 * nx * u01(seed, kChan, uint64_t(z) * 8 + c, 1);
            lam[c] = ny * (0.25 + 0.75 * u01(seed, kChan, uint64_t(z) * 8 + c, 2));
            phs[c] = 6.283185307179586 * u01(seed, kChan, uint64_t(z) * 8 + c, 3);
        }
        const double halfW = 0.2 * nx / 3.0 / 2.0;
        for (int j = 0; j < ny; ++j) for (int i = 0; i < nx; ++i) {
            bool in = false;
            for (int c = 0; c < 3 && !in; ++c) {
                double xc = x0[c] + amp[c] * std::sin(6.283185307179586 * j / lam[c] + phs[c]);
                xc = std::fmod(std::fmod(xc, double(nx)) + nx, double(nx));
                double d = std::fabs(i + 0.5 - xc); d = std::min(d, nx - d);
                in = d < halfW;
            }
            const size_t c = (size_t(z) * ny + j) * nx + i;
            k[c] = std::exp(lnBack + 0.5 * normal(seed, kLnK, c) + (in ? lnContrast : 0.0));
        }
    }
}

// The reservoir Jacobian.  Per cell c: porosity phi ~ U[0.1, 0.3]; accumulation block diag(a_p, a_s, .., a_s) with
// a_p = phi accP (PV c_t / dt: compressibility, absolute - the pressure part stays near-elliptic) and
// a_s = (phi / 0.2) accS sum_d T_cd (PV / dt for the transported unknowns, expressed through the local throughput:
// accS is the inverse of a cell CFL number).  Per connection (c, d) of the 7-point stencil:
// T_cd = harmonic mean of the permeabilities / 100 (times kv/kh vertically); the off-diagonal block of row c is
// -T_cd M_cd with M_cd an nb x nb mobility block, entries U[0.1, 1] (independent for each ordered pair, i.e. upwinded
// and non-symmetric).  Diagonal block = accumulation + sum_d T_cd M_cd (so its pressure column carries the sum of the
// neighbour pressure-column magnitudes).  Each well perforates its whole (i, j) column: WI = 2 k / 100, component
// mobilities m_e ~ U[0.1, 1]; cell rows get +WI m_e on their pressure column and -WI m_e in the BHP column, the BHP
// row -WI sum_e m_e on the cell pressure and 1.1 sum WI sum_e m_e on its diagonal.  rhs ~ U[-1, 1].
System* makeReservoir(int nx, int ny, int nz, int nb, int model, double sigma, double kvkh, int nwells, uint64_t seed,
                      double accP, double accS) {
    auto* sys = new System;
    const int na = nx * ny * nz, nv = nb, nc = na * nv, nw = nwells, n = nc + nw;
    permeability(nx, ny, nz, model, sigma, seed, sys->perm);
    const std::vector<double>& k = sys->perm;
    std::vector<double> phi(static_cast<size_t>(na));
    for (int c = 0; c < na; ++c) phi[size_t(c)] = 0.1 + 0.2 * u01(seed, kPoro, uint64_t(c));
    auto id = [&](int i, int j, int z) { return (z * ny + j) * nx + i; };
    auto harm = [](double a, double b) { return 2.0 * a * b / (a + b); };
    // wells: fixed fractional positions for the first five, hashed positions after
    static const double pos5[5][2] = {{0.25, 0.25}, {0.75, 0.75}, {0.75, 0.25}, {0.25, 0.75}, {0.5, 0.5}};
    std::vector<std::vector<Perf>> wells(static_cast<size_t>(nw));
    std::vector<int> wellOfCell(size_t(na), -1);
    for (int w = 0; w < nw; ++w) {
        double fx, fy;
        if (w < 5) { fx = pos5[w][0]; fy = pos5[w][1]; }
        else { fx = u01(seed, kWellPos, uint64_t(w), 0); fy = u01(seed, kWellPos, uint64_t(w), 1); }
        int wi = std::min(nx - 1, int(fx * nx)), wj = std::min(ny - 1, int(fy * ny));
        // keep wells in distinct columns
        while (wellOfCell[size_t(id(wi, wj, 0))] >= 0) { wi = (wi + 1) % nx; }
        for (int z = 0; z < nz; ++z) {
            Perf p; p.cell = id(wi, wj, z); p.wi = 2.0 * k[size_t(p.cell)] / 100.0;
            for (int e = 0; e < 4; ++e) p.m[e] = 0.1 + 0.9 * u01(seed, kWellMob, uint64_t(p.cell), uint64_t(e));
            wells[size_t(w)].push_back(p);
            wellOfCell[size_t(p.cell)] = w;
        }
    }
    const size_t b2 = size_t(nb) * nb;
    // scalar CSR, built row by row: sorted neighbour cells (block columns), then the well column
    Csr& J = sys->J;
    J.n = n; J.rowPtr.assign(size_t(n) + 1, 0);
    // count
    std::vector<int> ncell(static_cast<size_t>(na));
    for (int z = 0; z < nz; ++z) for (int j = 0; j < ny; ++j) for (int i = 0; i < nx; ++i) {
        int m = 1 + (z > 0) + (j > 0) + (i > 0) + (i < nx - 1) + (j < ny - 1) + (z < nz - 1);
        ncell[size_t(id(i, j, z))] = m;
    }
    for (int c = 0; c < na; ++c) for (int e = 0; e < nv; ++e) J.rowPtr[size_t(c) * nv + e + 1] = ncell[size_t(c)] * nv + (wellOfCell[size_t(c)] >= 0 ? 1 : 0);
    for (int w = 0; w < nw; ++w) J.rowPtr[size_t(nc + w) + 1] = int(wells[size_t(w)].size()) + 1;
    for (int r = 0; r < n; ++r) J.rowPtr[size_t(r) + 1] += J.rowPtr[size_t(r)];
    J.col.resize(size_t(J.rowPtr[size_t(n)])); J.val.resize(size_t(J.rowPtr[size_t(n)]));
    #pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; ++z) {
        double diag[16 * 16];
        int nbr[7]; double T[7];
        for (int j = 0; j < ny; ++j) for (int i = 0; i < nx; ++i) {
            const int c = id(i, j, z);
            int m = 0;
            auto add = [&](int d, double t) { nbr[m] = d; T[m] = t; ++m; };
            if (z > 0) add(id(i, j, z - 1), kvkh * harm(k[size_t(c)], k[size_t(id(i, j, z - 1))]) / 100.0);
            if (j > 0) add(id(i, j - 1, z), harm(k[size_t(c)], k[size_t(id(i, j - 1, z))]) / 100.0);
            if (i > 0) add(id(i - 1, j, z), harm(k[size_t(c)], k[size_t(id(i - 1, j, z))]) / 100.0);
            add(c, 0.0);
            if (i < nx - 1) add(id(i + 1, j, z), harm(k[size_t(c)], k[size_t(id(i + 1, j, z))]) / 100.0);
            if (j < ny - 1) add(id(i, j + 1, z), harm(k[size_t(c)], k[size_t(id(i, j + 1, z))]) / 100.0);
            if (z < nz - 1) add(id(i, j, z + 1), kvkh * harm(k[size_t(c)], k[size_t(id(i, j, z + 1))]) / 100.0);
            // diagonal: accumulation + sum T M
            for (size_t q = 0; q < b2; ++q) diag[q] = 0.0;
            double tsum = 0;
            for (int s2 = 0; s2 < m; ++s2) tsum += T[s2];
            for (int e = 0; e < nv; ++e) diag[size_t(e) * nv + e] = e == 0 ? phi[size_t(c)] * accP : phi[size_t(c)] / 0.2 * accS * tsum;
            for (int s = 0; s < m; ++s) {
                if (nbr[s] == c) continue;
                for (int e = 0; e < nv; ++e) for (int f = 0; f < nv; ++f)
                    diag[size_t(e) * nv + f] += T[s] * (0.1 + 0.9 * u01(seed, kMob, uint64_t(c) * 7 + uint64_t(s), uint64_t(e * nv + f)));
            }
            const int w = wellOfCell[size_t(c)];
            const Perf* pf = nullptr;
            if (w >= 0) for (const Perf& p : wells[size_t(w)]) if (p.cell == c) pf = &p;
            if (pf) for (int e = 0; e < nv; ++e) diag[size_t(e) * nv] += pf->wi * pf->m[e];
            for (int e = 0; e < nv; ++e) {
                const int r = c * nv + e;
                int q = J.rowPtr[size_t(r)];
                for (int s = 0; s < m; ++s) {
                    const int d = nbr[s];
                    for (int f = 0; f < nv; ++f) {
                        J.col[size_t(q)] = d * nv + f;
                        J.val[size_t(q)] = d == c ? diag[size_t(e) * nv + f]
                                                  : -T[s] * (0.1 + 0.9 * u01(seed, kMob, uint64_t(c) * 7 + uint64_t(s), uint64_t(e * nv + f)));
                        ++q;
                    }
                }
                if (pf) { J.col[size_t(q)] = nc + w; J.val[size_t(q)] = -pf->wi * pf->m[e]; ++q; }
            }
        }
    }
    for (int w = 0; w < nw; ++w) {
        const int r = nc + w;
        int q = J.rowPtr[size_t(r)];
        double d = 0;
        for (const Perf& p : wells[size_t(w)]) {       // perforations are in ascending cell order (z ascending)
            double msum = 0; for (int e = 0; e < nv; ++e) msum += p.m[e];
            J.col[size_t(q)] = p.cell * nv; J.val[size_t(q)] = -p.wi * msum; ++q;
            d += p.wi * msum;
        }
        J.col[size_t(q)] = r; J.val[size_t(q)] = 1.1 * d; ++q;
    }
    sys->b.resize(size_t(n));
    for (int r = 0; r < n; ++r) sys->b[size_t(r)] = 2.0 * u01(seed, kRhs, uint64_t(r)) - 1.0;
    sys->hasFull = true;
    if (!reduceWells(J, na, nv, nw, sys->S, sys->W)) { delete sys; return nullptr; }
    sys->bc.resize(size_t(nc));
    reduceWellRhs(sys->W, sys->b.data(), sys->bc.data());
    return sys;
}

// The reference tests' 2-D block Laplacian (run_tests.cpp:3694-3724): 5-point, two-decade permeabilities from
// std::mt19937(seed) + uniform_real_distribution(0.5, 2) (libstdc++), diagonal 1.02 sum T, pressure-coupled
// diagonal block.  Regenerated here so the reference's own checks can be re-asserted on the same inputs.
System* makeBlockLaplace(int nx, int ny, int nb, unsigned seed) {
    auto* sys = new System;
    Bsr& A = sys->S;
    std::mt19937 rng(seed); std::uniform_real_distribution<double> U(0.5, 2.0);
    A.n = nx * ny; A.nb = nb; A.rowPtr.assign(1, 0);
    std::vector<double> kx(size_t(nx) * ny), ky(size_t(nx) * ny);
    for (auto& v : kx) v = std::pow(10.0, 2 * (U(rng) - 1.25));
    for (auto& v : ky) v = std::pow(10.0, 2 * (U(rng) - 1.25));
    for (int j = 0; j < ny; ++j) for (int i = 0; i < nx; ++i) {
        const int c = j * nx + i;
        std::vector<std::pair<int, double>> row;
        double diag = 0;
        if (j > 0) row.push_back({c - nx, std::sqrt(ky[size_t(c)] * ky[size_t(c - nx)])});
        if (i > 0) row.push_back({c - 1, std::sqrt(kx[size_t(c)] * kx[size_t(c - 1)])});
        if (i < nx - 1) row.push_back({c + 1, std::sqrt(kx[size_t(c)] * kx[size_t(c + 1)])});
        if (j < ny - 1) row.push_back({c + nx, std::sqrt(ky[size_t(c)] * ky[size_t(c + nx)])});
        for (auto& q : row) diag += q.second;
        diag *= 1.02;
        row.push_back({c, diag});
        std::sort(row.begin(), row.end());
        for (auto& q : row) {
            A.col.push_back(q.first);
            for (int e = 0; e < nb; ++e) for (int f = 0; f < nb; ++f) {
                double v;
                if (q.first == c) v = (e == f) ? diag * (1.0 + 0.3 * e) : 0.05 * diag * (e == 0 ? 1.0 : 0.2) * (f + 1);
                else v = (e == f) ? -q.second * (e == 0 ? 1.0 : 0.6) : (e == 0 ? -0.02 * q.second : 0.0);
                A.val.push_back(v);
            }
        }
        A.rowPtr.push_back(int(A.col.size()));
    }
    finalizeDiag(A);
    // the reference tests use b = A x_t with x_t = sin / cos patterns; the caller chooses; here rhs = 1
    sys->bc.assign(size_t(A.n) * nb, 1.0);
    return sys;
}

// The reference's coupled 1-D chain with two wells (run_tests.cpp:1534-1556): 300 cells x 3, a vanishing third
// diagonal (1e-13), BHP rows; std::mt19937(11) + uniform_real_distribution(-1, 1).  x_t = sin(0.1 i) + 2, b = J x_t.
System* makeChain() {
    auto* sys = new System;
    const int na = 300, nv = 3, nw = 2, nc = na * nv, n = nc + nw;
    std::vector<std::map<int, double>> rows(static_cast<size_t>(n));
    std::mt19937 rng(11); std::uniform_real_distribution<double> U(-1, 1);
    for (int a = 0; a < na; ++a)
        for (int e = 0; e < nv; ++e) {
            const int r = a * nv + e;
            for (int v = 0; v < nv; ++v) rows[size_t(r)][a * nv + v] = (e == v ? 8.0 : 0.0) + U(rng) * (e == 0 || v == 0 ? 2.0 : 0.3);
            if (e == 2) rows[size_t(r)][a * nv + 2] = 1e-13;
            for (int nbc : {a - 1, a + 1}) if (nbc >= 0 && nbc < na) for (int v = 0; v < nv; ++v) rows[size_t(r)][nbc * nv + v] = -1.0 + 0.2 * U(rng);
        }
    const int perf[2][3] = {{10, 11, 12}, {200, 250, 290}};
    for (int w = 0; w < nw; ++w) {
        double d = 0;
        for (int q = 0; q < 3; ++q) {
            const int a = perf[w][q];
            for (int e = 0; e < nv; ++e) { rows[size_t(a) * nv + e][nc + w] = -0.7; rows[size_t(a) * nv + e][a * nv + e] += 0.7; }
            rows[size_t(nc + w)][a * nv] = -0.7; d += 0.7;
        }
        rows[size_t(nc + w)][nc + w] = d * 1.1;
    }
    Csr& J = sys->J;
    J.n = n; J.rowPtr.assign(size_t(n) + 1, 0);
    for (int r = 0; r < n; ++r) { for (auto& cv : rows[size_t(r)]) { J.col.push_back(cv.first); J.val.push_back(cv.second); } J.rowPtr[size_t(r) + 1] = int(J.col.size()); }
    std::vector<double> xt(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) xt[size_t(i)] = std::sin(0.1 * i) + 2.0;
    sys->b.assign(size_t(n), 0.0);
    for (int r = 0; r < n; ++r) { double s = 0; for (int p = J.rowPtr[size_t(r)]; p < J.rowPtr[size_t(r) + 1]; ++p) s += J.val[size_t(p)] * xt[size_t(J.col[size_t(p)])]; sys->b[size_t(r)] = s; }
    sys->hasFull = true;
    if (!reduceWells(J, na, nv, nw, sys->S, sys->W)) { delete sys; return nullptr; }
    sys->bc.resize(size_t(nc));
    reduceWellRhs(sys->W, sys->b.data(), sys->bc.data());
    return sys;
}

}  // namespace

extern "C" {

const char* rsg_last_error() { return g_err.c_str(); }

void* rsg_reservoir(int nx, int ny, int nz, int nb, int model, double sigma, double kvkh, int nwells, unsigned long long seed,
                    double accP, double accS) {
    g_err.clear();
    if (nx < 1 || ny < 1 || nz < 1 || nb < 1 || nb > 16 || nwells < 0) { g_err = "bad reservoir dimensions"; return nullptr; }
    return makeReservoir(nx, ny, nz, nb, model, sigma, kvkh, nwells, seed, accP, accS);
}
void* rsg_block_laplace(int nx, int ny, int nb, unsigned seed) { g_err.clear(); return makeBlockLaplace(nx, ny, nb, seed); }
void* rsg_chain() { g_err.clear(); return makeChain(); }
void rsg_free(void* h) { delete static_cast<System*>(h); }

// dims: n, nb, nnzb, has_full, n_full, nnz_full, na, nv, nw, fill
int rsg_dims(void* h, long long* out) {
    const System* s = static_cast<const System*>(h);
    out[0] = s->S.n; out[1] = s->S.nb; out[2] = s->S.rowPtr.empty() ? 0 : s->S.rowPtr[size_t(s->S.n)];
    out[3] = s->hasFull; out[4] = s->J.n; out[5] = s->J.rowPtr.empty() ? 0 : s->J.rowPtr[size_t(s->J.n)];
    out[6] = s->W.plan.na; out[7] = s->W.plan.nv; out[8] = s->W.plan.nw; out[9] = s->W.plan.fill();
    return 0;
}
int rsg_get_bsr(void* h, int* rowPtr, int* col, double* val, double* rhs) {
    const System* s = static_cast<const System*>(h);
    std::memcpy(rowPtr, s->S.rowPtr.data(), sizeof(int) * s->S.rowPtr.size());
    std::memcpy(col, s->S.col.data(), sizeof(int) * s->S.col.size());
    std::memcpy(val, s->S.val.data(), sizeof(double) * s->S.val.size());
    if (rhs) std::memcpy(rhs, s->bc.data(), sizeof(double) * s->bc.size());
    return 0;
}
int rsg_get_full(void* h, int* rowPtr, int* col, double* val, double* rhs) {
    const System* s = static_cast<const System*>(h);
    if (!s->hasFull) { g_err = "no full system"; return -1; }
    std::memcpy(rowPtr, s->J.rowPtr.data(), sizeof(int) * s->J.rowPtr.size());
    std::memcpy(col, s->J.col.data(), sizeof(int) * s->J.col.size());
    std::memcpy(val, s->J.val.data(), sizeof(double) * s->J.val.size());
    if (rhs) std::memcpy(rhs, s->b.data(), sizeof(double) * s->b.size());
    return 0;
}
int rsg_get_perm(void* h, double* k) {
    const System* s = static_cast<const System*>(h);
    std::memcpy(k, s->perm.data(), sizeof(double) * s->perm.size());
    return int(s->perm.size());
}
int rsg_reduce_rhs(void* h, const double* b, double* bc) {
    const System* s = static_cast<const System*>(h);
    if (!s->hasFull) { g_err = "no well Schur data"; return -1; }
    reduceWellRhs(s->W, b, bc);
    return 0;
}
int rsg_recover(void* h, const double* b, const double* xc, double* x) {
    const System* s = static_cast<const System*>(h);
    if (!s->hasFull) { g_err = "no well Schur data"; return -1; }
    recoverWells(s->W, b, xc, x);
    return 0;
}
// Reduce an arbitrary scalar CSR (cells + BHP rows) into a handle (no generation): the engine-side step above the
// solver boundary, for callers that assemble their own Jacobian.
void* rsg_from_csr(int n, const int* rowPtr, const int* col, const double* val, const double* b, int na, int nv, int nw) {
    g_err.clear();
    auto* sys = new System;
    sys->J.n = n; sys->J.rowPtr.assign(rowPtr, rowPtr + n + 1);
    sys->J.col.assign(col, col + rowPtr[n]); sys->J.val.assign(val, val + rowPtr[n]);
    sys->b.assign(b, b + n);
    sys->hasFull = true;
    if (!reduceWells(sys->J, na, nv, nw, sys->S, sys->W)) { delete sys; return nullptr; }
    sys->bc.resize(size_t(na) * nv);
    reduceWellRhs(sys->W, sys->b.data(), sys->bc.data());
    return sys;
}

}  // extern "C"
