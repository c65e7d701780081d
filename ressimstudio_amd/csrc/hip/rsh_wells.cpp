// Well Schur reduction on the device (include/ressim_hip.h, rsh_wells_*): the engine's scalar Jacobian (cells in
// blocks of nv unknowns, then one BHP row per well, the layout of WellSchur::build, bsr.cpp:733-840) is handed over as
// is; the block pattern of S = A_cc - A_cw D_ww^-1 A_wc and the maps from J's entries to S's values are built once per
// pattern on the host, then every Newton iteration moves only J's values: S's values, the reduced right-hand side and
// the recovered BHPs are computed in HBM (kWellDinv, kSchurValues, kSchurRhs, kSchurRecover) in the reference's order
// of operations.  The reduced system is an ordinary resident rsh_system (setup / solve as usual).
#include "rsh_solver.hpp"

using namespace rsh;

struct rsh_wells {
    rsh_context* ctx = nullptr;
    int n = 0, na = 0, nv = 1, nw = 0, nnzJ = 0;
    rsh_system* S = nullptr;
    DBuf<int> wdiag, wcPtr, wcCol, wcSrc, cwPtr, cwWell, cwSrc, direct, termPtr, bad;
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
        if (na <= 0 || nv < 1 || nv > kMaxNb || nw < 0 || n != na * nv + nw) return fail(RSH_ERR_ARG, "n must equal na * nv + nw");
        const int nc = na * nv, nnz = row_ptr[n];
        for (int r = 0; r < n; ++r) {
            if (row_ptr[r + 1] < row_ptr[r]) return fail(RSH_ERR_ARG, "row_ptr not monotone");
            for (int p = row_ptr[r]; p < row_ptr[r + 1]; ++p)
                if (col[p] < 0 || col[p] >= n) return fail(RSH_ERR_ARG, "column index out of range");
        }
        auto W = std::make_unique<rsh_wells>();
        W->ctx = ctx;
        W->n = n; W->na = na; W->nv = nv; W->nw = nw; W->nnzJ = nnz;
        // well rows: diagonal and cell couplings (a well-well coupling is not supported, bsr.cpp:745)
        std::vector<int> wdiag(size_t(nw), -1), wcPtr(1, 0), wcCol, wcSrc;
        for (int w = 0; w < nw; ++w) {
            const int r = nc + w;
            for (int p = row_ptr[r]; p < row_ptr[r + 1]; ++p) {
                const int c = col[p];
                if (c == r) wdiag[size_t(w)] = p;
                else if (c < nc) { wcCol.push_back(c); wcSrc.push_back(p); }
                else return fail(RSH_ERR_UNSUPPORTED, "well-well coupling (well groups) is not supported");
            }
            wcPtr.push_back(int(wcCol.size()));
        }
        // cell rows: well columns
        std::vector<int> cwPtr(1, 0), cwWell, cwSrc;
        for (int r = 0; r < nc; ++r) {
            for (int p = row_ptr[r]; p < row_ptr[r + 1]; ++p)
                if (col[p] >= nc) { cwWell.push_back(col[p] - nc); cwSrc.push_back(p); }
            cwPtr.push_back(int(cwWell.size()));
        }
        // block pattern: the cell pattern plus the couplings through wells (bsr.cpp:775-807)
        std::vector<std::vector<int>> wellCells(static_cast<size_t>(nw));
        for (int w = 0; w < nw; ++w) {
            auto& v = wellCells[size_t(w)];
            for (int q = wcPtr[size_t(w)]; q < wcPtr[size_t(w) + 1]; ++q) v.push_back(wcCol[size_t(q)] / nv);
            std::sort(v.begin(), v.end());
            v.erase(std::unique(v.begin(), v.end()), v.end());
        }
        std::vector<int> srp(1, 0), sci, mark(static_cast<size_t>(na), -1), cols, wl;
        for (int i = 0; i < na; ++i) {
            cols.clear();
            wl.clear();
            for (int e = 0; e < nv; ++e) {
                const int r = i * nv + e;
                for (int p = row_ptr[r]; p < row_ptr[r + 1]; ++p) {
                    const int c = col[p];
                    if (c < nc && mark[size_t(c / nv)] != i) { mark[size_t(c / nv)] = i; cols.push_back(c / nv); }
                }
                for (int p = cwPtr[size_t(r)]; p < cwPtr[size_t(r) + 1]; ++p) wl.push_back(cwWell[size_t(p)]);
            }
            if (mark[size_t(i)] != i) { mark[size_t(i)] = i; cols.push_back(i); }
            std::sort(wl.begin(), wl.end());
            wl.erase(std::unique(wl.begin(), wl.end()), wl.end());
            for (int w : wl)
                for (int j : wellCells[size_t(w)])
                    if (mark[size_t(j)] != i) { mark[size_t(j)] = i; cols.push_back(j); }
            std::sort(cols.begin(), cols.end());
            sci.insert(sci.end(), cols.begin(), cols.end());
            srp.push_back(int(sci.size()));
        }
        // maps: per S value (block p, entry e f) the J cell entry landing there and the ordered fill terms
        const int nv2 = nv * nv;
        const long long nvals = (long long)sci.size() * nv2;
        std::vector<int> direct(static_cast<size_t>(nvals), -1), pos(static_cast<size_t>(na), -1);
        std::vector<std::vector<int4>> tl(static_cast<size_t>(nvals));
        for (int i = 0; i < na; ++i) {
            for (int p = srp[size_t(i)]; p < srp[size_t(i) + 1]; ++p) pos[size_t(sci[size_t(p)])] = p;
            for (int e = 0; e < nv; ++e) {
                const int r = i * nv + e;
                for (int p = row_ptr[r]; p < row_ptr[r + 1]; ++p) {
                    const int c = col[p];
                    if (c >= nc) continue;
                    const long long t = (long long)pos[size_t(c / nv)] * nv2 + e * nv + c % nv;
                    if (direct[size_t(t)] >= 0) return fail(RSH_ERR_ARG, "duplicate entries in a row of the Jacobian");
                    direct[size_t(t)] = p;
                }
                for (int p = cwPtr[size_t(r)]; p < cwPtr[size_t(r) + 1]; ++p) {
                    const int w = cwWell[size_t(p)];
                    for (int q = wcPtr[size_t(w)]; q < wcPtr[size_t(w) + 1]; ++q) {
                        const int c = wcCol[size_t(q)];
                        const long long t = (long long)pos[size_t(c / nv)] * nv2 + e * nv + c % nv;
                        tl[size_t(t)].push_back(make_int4(cwSrc[size_t(p)], w, wcSrc[size_t(q)], 0));
                    }
                }
            }
        }
        std::vector<int> termPtr(1, 0);
        std::vector<int4> terms;
        for (auto& v : tl) {
            terms.insert(terms.end(), v.begin(), v.end());
            termPtr.push_back(int(terms.size()));
        }
        W->nvals = nvals;
        W->nterms = (long long)terms.size();
        rsh_system* S = nullptr;
        const int rc = rsh_system_create(ctx, na, nv, srp.data(), sci.data(), &S);
        if (rc != RSH_OK) return rc;
        W->S = S;
        hipStream_t st = ctx->c.stream;
        W->wdiag.upload(wdiag, st);
        W->wcPtr.upload(wcPtr, st);
        W->wcCol.upload(wcCol, st);
        W->wcSrc.upload(wcSrc, st);
        W->cwPtr.upload(cwPtr, st);
        W->cwWell.upload(cwWell, st);
        W->cwSrc.upload(cwSrc, st);
        W->direct.upload(direct, st);
        W->termPtr.upload(termPtr, st);
        W->terms.upload(terms, st);
        W->jval.alloc(size_t(nnz));
        W->dinv.alloc(size_t(std::max(nw, 1)));
        W->bfull.alloc(size_t(n));
        W->xfull.alloc(size_t(n));
        W->bc.alloc(size_t(nc));
        W->xc.alloc(size_t(nc));
        W->bad.alloc(1);
        RSH_HIP(hipHostMalloc(reinterpret_cast<void**>(&W->hbad), sizeof(int), hipHostMallocDefault));
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
        RSH_HIP(hipMemsetAsync(w->bad.p, 0, sizeof(int), st));
        launchWellDinv(st, w->nw, w->wdiag.p, w->wcPtr.p, w->wcSrc.p, w->jval.p, w->dinv.p, w->bad.p);
        RSH_HIP(hipMemcpyAsync(w->hbad, w->bad.p, sizeof(int), hipMemcpyDeviceToHost, st));
        launchSchurValues(st, w->nvals, w->direct.p, w->termPtr.p, w->terms.p, w->jval.p, w->dinv.p, w->S->val.p);
        w->S->valuesChanged();
        RSH_HIP(hipStreamSynchronize(st));
        if (*w->hbad) return fail(RSH_ERR_UNSUPPORTED, "a well row has no usable diagonal (|J_ww| <= 1e-12 max |J_w,cells|)");
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
