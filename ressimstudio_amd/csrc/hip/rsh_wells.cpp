// Well Schur reduction on the device (include/ressim_hip.h, rsh_wells_*): the engine's scalar Jacobian (cells in
// blocks of nv unknowns, then one BHP row per well, the layout the engine reduces with WellSchur, bsr.cpp:733-990) is
// handed over as it is assembled.  The host plan of csrc/host/well_plan.hpp (S = A - sum_w (B[:, w] / d_w) C[w, :],
// one rank-1 update per well) gives the reduced block pattern and, per value of S, its cell entry and its ordered
// rank-1 terms, once per pattern; every Newton iteration then moves only J's values: 1 / d_w, S's values, the
// reduced right-hand side and the recovered BHPs are computed in HBM (kWellDinv, kSchurValues, kSchurRhs,
// kSchurRecover) with the engine's per-entry order of operations.  The reduced system is an ordinary resident
// rsh_system (setup / solve as usual).
//
// The pattern is fixed before any values exist, so every structural cell -> well entry couples its cell to the well
// (the engine drops exactly-zero couplings from its pattern, bsr.cpp:755-757, and would then build a smaller S).
// Structural well-well entries are accepted and must hold zeros (checked on every set of values).
#include "../host/well_plan.hpp"
#include "rsh_solver.hpp"

using namespace rsh;

struct rsh_wells {
    rsh_context* ctx = nullptr;
    int n = 0, na = 0, nv = 1, nw = 0, nnzJ = 0;
    rsh_system* S = nullptr;
    DBuf<int> wdiag, wcPtr, wcCol, wcSrc, wwPtr, wwSrc, cwPtr, cwWell, cwSrc, direct, termPtr, bad;
    DBuf<int4> terms;
    DBuf<double> jval, dinv, bfull, bc, xc, xfull;
    int* hbad = nullptr;
    bool valuesSet = false;
    long long nvals = 0, nterms = 0;
    hipStream_t s() const { return ctx->c.stream; }
    ~rsh_wells() {
        if (S) rsh_system_destroy(S);
        if (hbad) (void)hipHostFree(hbad);
    }
};

extern "C" {

int rsh_wells_create(rsh_context* ctx, int n, const int* row_ptr, const int* col, int na, int nv, int nw, rsh_wells** out) {
    return guarded([&] {
        if (!ctx || !row_ptr || !col || !out) return fail(RSH_ERR_ARG, "null argument");
        *out = nullptr;
        if (nv > kMaxNb) return fail(RSH_ERR_UNSUPPORTED, "block size above 4");
        wellplan::Plan P;
        std::string why = wellplan::split(n, row_ptr, col, na, nv, nw, [](int) { return true; }, P);
        if (!why.empty()) return fail(RSH_ERR_ARG, why);
        wellplan::pattern(n, row_ptr, col, P);
        wellplan::ValueMap M;
        why = wellplan::valueMap(row_ptr, col, P, M);
        if (!why.empty()) return fail(RSH_ERR_ARG, why);
        std::vector<int> byRowStart, byRowWell, byRowEntry;
        wellplan::couplingsByRow(P, byRowStart, byRowWell, byRowEntry);
        std::vector<int> toCellCol(P.toCellEntry.size());
        for (size_t q = 0; q < toCellCol.size(); ++q) toCellCol[q] = col[P.toCellEntry[q]];
        std::vector<int4> terms(M.termWell.size());
        for (size_t k = 0; k < terms.size(); ++k) terms[k] = make_int4(M.termCoupling[k], M.termWell[k], M.termEntry[k], 0);

        auto W = std::make_unique<rsh_wells>();
        W->ctx = ctx;
        W->n = n; W->na = na; W->nv = nv; W->nw = nw; W->nnzJ = row_ptr[n];
        W->nvals = (long long)M.cellSource.size();
        W->nterms = (long long)terms.size();
        rsh_system* S = nullptr;
        const int rc = rsh_system_create(ctx, na, nv, P.sPtr.data(), P.sCol.data(), &S);
        if (rc != RSH_OK) return rc;
        W->S = S;
        hipStream_t st = ctx->c.stream;
        W->wdiag.upload(P.bhpDiag, st);
        W->wcPtr.upload(P.toCellStart, st);
        W->wcCol.upload(toCellCol, st);
        W->wcSrc.upload(P.toCellEntry, st);
        W->wwPtr.upload(P.toWellStart, st);
        W->wwSrc.upload(P.toWellEntry, st);
        W->cwPtr.upload(byRowStart, st);
        W->cwWell.upload(byRowWell, st);
        W->cwSrc.upload(byRowEntry, st);
        W->direct.upload(M.cellSource, st);
        W->termPtr.upload(M.termStart, st);
        W->terms.upload(terms, st);
        W->jval.alloc(size_t(std::max(W->nnzJ, 1)));
        W->dinv.alloc(size_t(std::max(nw, 1)));
        W->bfull.alloc(size_t(n));
        W->xfull.alloc(size_t(n));
        W->bc.alloc(size_t(na) * nv);
        W->xc.alloc(size_t(na) * nv);
        W->bad.alloc(2);
        RSH_HIP(hipHostMalloc(reinterpret_cast<void**>(&W->hbad), 2 * sizeof(int), hipHostMallocDefault));
        RSH_HIP(hipStreamSynchronize(st));
        *out = W.release();
        return RSH_OK;
    });
}

int rsh_wells_destroy(rsh_wells* w) {
    if (!w) return RSH_OK;
    return guarded([&] {
        RSH_HIP(hipSetDevice(w->ctx->c.device));
        (void)hipStreamSynchronize(w->s());
        delete w;
        return RSH_OK;
    });
}

int rsh_wells_system(rsh_wells* w, rsh_system** sys) {
    if (!w || !sys) return fail(RSH_ERR_ARG, "null argument");
    *sys = w->S;
    return RSH_OK;
}

// J's values (nnz doubles, host or device): S's values, reduced in HBM
int rsh_wells_set_values(rsh_wells* w, const double* jval, int on_device) {
    return guarded([&] {
        if (!w || !jval) return fail(RSH_ERR_ARG, "null argument");
        RSH_HIP(hipSetDevice(w->ctx->c.device));
        hipStream_t st = w->s();
        w->S->ctx = w->ctx;
        RSH_HIP(hipMemcpyAsync(w->jval.p, jval, sizeof(double) * size_t(w->nnzJ), on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
        RSH_HIP(hipMemsetAsync(w->bad.p, 0, 2 * sizeof(int), st));
        launchWellDinv(st, w->nw, w->wdiag.p, w->wcPtr.p, w->wcSrc.p, w->wwPtr.p, w->wwSrc.p, w->jval.p, w->dinv.p, w->bad.p);
        RSH_HIP(hipMemcpyAsync(w->hbad, w->bad.p, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
        launchSchurValues(st, w->nvals, w->direct.p, w->termPtr.p, w->terms.p, w->jval.p, w->dinv.p, w->S->val.p);
        w->S->valuesChanged();
        RSH_HIP(hipStreamSynchronize(st));
        if (w->hbad[1]) return fail(RSH_ERR_UNSUPPORTED, "well-well coupling (a nonzero entry between two BHP unknowns) is not supported");
        if (w->hbad[0]) return fail(RSH_ERR_UNSUPPORTED, "a well row has no usable diagonal (|J_ww| <= 1e-12 max |J_w,cells|)");
        w->valuesSet = true;
        return RSH_OK;
    });
}

// b_full (n doubles) -> reduced solve on the set-up system -> x_full (cells and BHPs); host or device vectors
int rsh_wells_solve(rsh_wells* w, const double* b_full, double* x_full, int on_device, double tol, int max_it, int* iterations, double* rel_res) {
    return guarded([&] {
        if (!w || !b_full || !x_full) return fail(RSH_ERR_ARG, "null argument");
        if (!w->valuesSet) return fail(RSH_ERR_ARG, "rsh_wells_set_values first");
        RSH_HIP(hipSetDevice(w->ctx->c.device));
        hipStream_t st = w->s();
        const int nc = w->na * w->nv;
        const double* b = b_full;
        if (!on_device) {
            RSH_HIP(hipMemcpyAsync(w->bfull.p, b_full, sizeof(double) * size_t(w->n), hipMemcpyHostToDevice, st));
            b = w->bfull.p;
        }
        launchSchurRhs(st, nc, w->cwPtr.p, w->cwWell.p, w->cwSrc.p, w->jval.p, w->dinv.p, b, w->bc.p);
        const int rc = rsh_solve_device(w->S, w->bc.p, w->xc.p, tol, max_it, iterations, rel_res);
        if (rc != RSH_OK && rc != RSH_ERR_NOT_CONVERGED && rc != RSH_ERR_BREAKDOWN) return rc;
        double* x = on_device ? x_full : w->xfull.p;
        launchSchurRecover(st, nc, w->nw, w->wcPtr.p, w->wcCol.p, w->wcSrc.p, w->jval.p, w->dinv.p, b, w->xc.p, x);
        if (!on_device) RSH_HIP(hipMemcpyAsync(x_full, w->xfull.p, sizeof(double) * size_t(w->n), hipMemcpyDeviceToHost, st));
        RSH_HIP(hipStreamSynchronize(st));
        return rc;
    });
}

// S's values (nnzb * nv * nv doubles, the reduced system's block layout) to the host, for checks
int rsh_wells_reduced_values(rsh_wells* w, double* out) {
    return guarded([&] {
        if (!w || !out) return fail(RSH_ERR_ARG, "null argument");
        RSH_HIP(hipSetDevice(w->ctx->c.device));
        RSH_HIP(hipMemcpyAsync(out, w->S->val.p, w->S->val.bytes(), hipMemcpyDeviceToHost, w->s()));
        RSH_HIP(hipStreamSynchronize(w->s()));
        return RSH_OK;
    });
}

}  // extern "C"
