// TEST INFRASTRUCTURE: the checker API (rs_check.h) implemented on top of the reference's own classes, compiled
// from the reference sources where they lie (engine/src/{bsr,linsolve,amg}.cpp, see oracle/Makefile) into
// oracle/_ref/.  Used only to pin the oracle restatement and as the CPU baseline; never linked by the product.
#include "rs_check.h"
#include "ressim/bsr.h"

#include <chrono>
#include <cstdlib>
#include <memory>
#include <vector>

using namespace ressim;

namespace {
BSR makeBsr(int n, int nb, const int* rp, const int* ci, const double* v) {
    BSR A; A.n = n; A.nb = nb;
    A.rowPtr.assign(rp, rp + n + 1);
    A.col.assign(ci, ci + rp[n]);
    A.val.assign(v, v + size_t(rp[n]) * nb * nb);
    A.finalizeDiag();
    return A;
}
struct Bilu0H { int mode = 1; BILU0 d; BILU0f f; };
struct CprH { BSR S; BlockCPR P; };
}  // namespace

extern "C" {

int chk_spmv(int n, int nb, const int* rp, const int* ci, const double* v, const double* x, double* y) {
    BSR A = makeBsr(n, nb, rp, ci, v);
    A.spmv(x, y);
    return 0;
}

void* chk_bilu0_new(int n, int nb, const int* rp, const int* ci, const double* v, int mode) {
    BSR A = makeBsr(n, nb, rp, ci, v);
    auto* h = new Bilu0H; h->mode = mode;
    if (mode == 3) { h->f.factor(A); return h; }
    if (mode == 4) { h->d.setLean(true); h->d.factor(A); return h; }
    if (mode == 2) {
        // float factors exist only in the contiguous level-ordered layout (bsr.cpp:413-419): force it on
        setenv("RESSIM_LEVEL_ILU", "2", 1);
        h->d.setFloatFactors(true);
        h->d.factor(A);
        unsetenv("RESSIM_LEVEL_ILU");
        if (!h->d.floatFactors()) { delete h; return nullptr; }
        return h;
    }
    h->d.factor(A);
    return h;
}
int chk_bilu0_apply(void* h, const double* r, double* z) {
    auto* b = static_cast<Bilu0H*>(h);
    if (b->mode == 3) b->f.apply(r, z); else b->d.apply(r, z);
    return 0;
}
int chk_bilu0_regularised(void* h) { auto* b = static_cast<Bilu0H*>(h); return b->mode == 3 ? b->f.regularisedBlocks() : b->d.regularisedBlocks(); }
void chk_bilu0_free(void* h) { delete static_cast<Bilu0H*>(h); }

void* chk_mcbilu0_new(int n, int nb, const int* rp, const int* ci, const double* v) {
    auto* m = new MCBILU0;
    m->factor(makeBsr(n, nb, rp, ci, v));
    return m;
}
int chk_mcbilu0_apply(void* h, const double* r, double* z) { static_cast<MCBILU0*>(h)->apply(r, z); return 0; }
int chk_mcbilu0_colours(void* h) { return static_cast<MCBILU0*>(h)->colours(); }
int chk_mcbilu0_colouring(void* h, int* colour, int* perm) {
    const Colouring& C = static_cast<MCBILU0*>(h)->colouring();
    for (size_t i = 0; i < C.colour.size(); ++i) { colour[i] = C.colour[i]; perm[i] = C.perm[i]; }
    return C.colours;
}
void chk_mcbilu0_free(void* h) { delete static_cast<MCBILU0*>(h); }

void* chk_cpr_new(int n, int nb, const int* rp, const int* ci, const double* v) {
    auto* c = new CprH;
    c->S = makeBsr(n, nb, rp, ci, v);
    c->P.setup(c->S);
    return c;
}
void* chk_cpr_new_acc(int n, int nb, const int* rp, const int* ci, const double* v, const double* acc) {
    auto* c = new CprH;
    c->S = makeBsr(n, nb, rp, ci, v);
    c->P.setup(c->S, acc);
    return c;
}
int chk_cpr_fallbacks(void* h) { return static_cast<CprH*>(h)->P.trueImpesFallbacks(); }
int chk_cpr_refresh(void* h, const double* v) {
    auto* c = static_cast<CprH*>(h);
    std::copy(v, v + c->S.val.size(), c->S.val.begin());
    c->P.setup(c->S, nullptr, true);
    return c->P.amgReused() ? 1 : 0;
}
int chk_cpr_apply(void* h, const double* r, double* z) { static_cast<CprH*>(h)->P.apply(r, z); return 0; }
int chk_cpr_weights(void* h, double* w) {
    const auto& W = static_cast<CprH*>(h)->P.weights();
    for (size_t i = 0; i < W.size(); ++i) w[i] = W[i];
    return 0;
}
int chk_cpr_nlevels(void* h) { return static_cast<CprH*>(h)->P.amg().levels(); }
int chk_cpr_level(void* h, int l, int* n, int* nnz) {
    const auto& L = static_cast<CprH*>(h)->P.amg().hierarchy()[size_t(l)];
    *n = L.A.n; *nnz = int(L.A.col.size());
    return 0;
}
int chk_cpr_level_matrix(void* h, int l, int* rp, int* ci, double* v) {
    const auto& L = static_cast<CprH*>(h)->P.amg().hierarchy()[size_t(l)];
    for (size_t i = 0; i < L.A.rowPtr.size(); ++i) rp[i] = L.A.rowPtr[i];
    for (size_t i = 0; i < L.A.col.size(); ++i) { ci[i] = L.A.col[i]; v[i] = L.A.val[i]; }
    return 0;
}
int chk_cpr_aggregates(void* h, int l, int* agg) {
    const auto& H = static_cast<CprH*>(h)->P.amg().hierarchy();
    const auto& L = H[size_t(l)];
    if (size_t(l) + 1 >= H.size()) return -1;
    for (int i = 0; i < L.A.n; ++i) {
        if (L.pRow[size_t(i) + 1] - L.pRow[size_t(i)] != 1) return -2;   // plain aggregation: one entry per row
        agg[i] = L.pCol[size_t(L.pRow[size_t(i)])];
    }
    return L.nc;
}
int chk_cpr_amg_apply(void* h, const double* r, double* z) { static_cast<CprH*>(h)->P.amg().apply(r, z); return 0; }
void chk_cpr_free(void* h) { delete static_cast<CprH*>(h); }

struct DevH { BSR S; DeviceCprAmg D; std::vector<float> Av; };
void* chk_devcpr_new(int n, int nb, const int* rp, const int* ci, const double* v) {
    auto* d = new DevH;
    d->S = makeBsr(n, nb, rp, ci, v);
    if (!d->D.build(d->S)) { delete d; return nullptr; }
    d->Av.assign(d->S.val.begin(), d->S.val.end());
    return d;
}
int chk_devcpr_apply(void* h, const double* r, double* z) {
    auto* d = static_cast<DevH*>(h);
    const size_t N = size_t(d->S.n) * d->S.nb;
    std::vector<float> rf(r, r + N), zf(N);
    d->D.apply(d->S, d->Av, rf.data(), zf.data());
    for (size_t i = 0; i < N; ++i) z[i] = zf[i];
    return 0;
}
void chk_devcpr_free(void* h) { delete static_cast<DevH*>(h); }

int chk_krylov(int n, int nb, const int* rp, const int* ci, const double* v, const double* b, double* x, int precond, int method,
               int restart, double tol, int maxit, int* its, double* relres) {
    BSR A = makeBsr(n, nb, rp, ci, v);
    std::unique_ptr<Preconditioner> M;
    switch (precond) {
        case 1: { auto* p = new BILU0; p->factor(A); M.reset(p); break; }
        case 2: {
            setenv("RESSIM_LEVEL_ILU", "2", 1);
            auto* p = new BILU0; p->setFloatFactors(true); p->factor(A); M.reset(p);
            unsetenv("RESSIM_LEVEL_ILU");
            break;
        }
        case 3: { auto* p = new BILU0f; p->factor(A); M.reset(p); break; }
        case 4: { auto* p = new MCBILU0; p->factor(A); M.reset(p); break; }
        case 5: { auto* p = new BlockCPR; p->setup(A); M.reset(p); break; }
        case 6: { auto* p = new BILU0; p->setLean(true); p->factor(A); M.reset(p); break; }
        case 7: { auto* p = new BlockCPR; p->setLean(true); p->setup(A); M.reset(p); break; }
        default: return -1;
    }
    std::vector<double> bb(b, b + size_t(n) * nb), xx;
    BsrOperator op(A);
    const LinResult r = method == 0 ? bicgstab(op, *M, bb, xx, tol, maxit) : gmres(op, *M, bb, xx, tol, maxit, restart);
    for (size_t i = 0; i < xx.size(); ++i) x[i] = xx[i];
    *its = r.iterations; *relres = r.relResidual;
    return r.converged ? 1 : 0;
}

int chk_krylov_cpr_steady(int n, int nb, const int* rp, const int* ci, const double* v, const double* b, double* x, double tol,
                          int maxit, int* its, double* relres, double* secs) {
    BSR A = makeBsr(n, nb, rp, ci, v);
    BlockCPR P;
    P.setup(A);
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    P.setup(A, nullptr, true);
    const auto t1 = clk::now();
    std::vector<double> bb(b, b + size_t(n) * nb), xx;
    BsrOperator op(A);
    const LinResult r = bicgstab(op, P, bb, xx, tol, maxit);
    const auto t2 = clk::now();
    for (size_t i = 0; i < xx.size(); ++i) x[i] = xx[i];
    *its = r.iterations; *relres = r.relResidual;
    secs[0] = std::chrono::duration<double>(t1 - t0).count();
    secs[1] = std::chrono::duration<double>(t2 - t1).count();
    return r.converged ? 1 : 0;
}

// WellSchur::build / reduceRhs / recover (bsr.cpp:733-990) on a scalar CSR Jacobian: returns a handle (null when the
// reference refuses the system); sizes[0] = reduced block rows, sizes[1] = reduced blocks, sizes[2] = fill blocks
void* chk_wellschur_new(int n, const int* rp, const int* ci, const double* v, int na, int nv, int nw, int* sizes) {
    CSR J; J.n = n;
    J.rowPtr.assign(rp, rp + n + 1);
    J.col.assign(ci, ci + rp[n]);
    J.val.assign(v, v + rp[n]);
    J.finalizeDiag();
    auto* w = new WellSchur;
    if (!w->build(J, na, nv, nw)) { delete w; return nullptr; }
    sizes[0] = w->reduced().n; sizes[1] = int(w->reduced().col.size()); sizes[2] = w->fillBlocks();
    return w;
}
int chk_wellschur_reduced(void* h, int* rp, int* ci, double* v) {
    const BSR& S = static_cast<WellSchur*>(h)->reduced();
    for (size_t i = 0; i < S.rowPtr.size(); ++i) rp[i] = S.rowPtr[i];
    for (size_t i = 0; i < S.col.size(); ++i) ci[i] = S.col[i];
    for (size_t i = 0; i < S.val.size(); ++i) v[i] = S.val[i];
    return 0;
}
int chk_wellschur_reduce_rhs(void* h, const double* b, double* bc) { static_cast<WellSchur*>(h)->reduceRhs(b, bc); return 0; }
int chk_wellschur_recover(void* h, const double* b, const double* xc, double* x) { static_cast<WellSchur*>(h)->recover(b, xc, x); return 0; }
void chk_wellschur_free(void* h) { delete static_cast<WellSchur*>(h); }

int chk_float_solve(int n, int nb, const int* rp, const int* ci, const double* v, const double* b, double* x, int kind, double tol, int maxit,
                    int* its) {
    BSR A = makeBsr(n, nb, rp, ci, v);
    CpuFloatBlockSolver dev;
    dev.multicolour = kind == 1;
    dev.cprAmg = kind == 2;
    std::vector<double> bb(b, b + size_t(n) * nb), xx;
    const bool ok = dev.solve(A, bb, xx, tol, maxit, *its);
    for (size_t i = 0; i < xx.size(); ++i) x[i] = xx[i];
    return ok ? 1 : 0;
}
int chk_refine_float(int n, int nb, const int* rp, const int* ci, const double* v, const double* b, double* x, int kind, double tol, int max_outer,
                     int inner_it, int* its, double* relres) {
    BSR A = makeBsr(n, nb, rp, ci, v);
    CpuFloatBlockSolver dev;
    dev.multicolour = kind == 1;
    dev.cprAmg = kind == 2;
    std::vector<double> bb(b, b + size_t(n) * nb), xx;
    const LinResult r = refineWithDevice(dev, A, bb, xx, tol, max_outer, inner_it);
    for (size_t i = 0; i < xx.size(); ++i) x[i] = xx[i];
    *its = r.iterations;
    *relres = r.relResidual;
    return r.converged ? 1 : 0;
}

// number of OpenMP threads the reference will use (setNumThreads(0) = physical cores unless OMP_NUM_THREADS)
int chk_set_threads(int n) { setNumThreads(n); return numThreads(); }

}  // extern "C"
