// Batched solves of many independent block systems (include/ressim_hip.h, rsh_batch_*).
//
// Systems of one block size are concatenated into one block-diagonal resident system, so the SpMV slices, the
// level-scheduled (or multicolour) block ILU(0) and its dense well groups cover every system of that block size in
// one launch per level.  The multicolour ordering of a block-diagonal matrix colours every block exactly as the
// block alone (the greedy visits vertices by degree with stable ties and only looks at a vertex's own block), and
// the natural-order factors are order-independent, so each system's preconditioner is the one it gets alone.
// BiCGSTAB keeps one scalar set per system: every vector kernel runs one workgroup per task (a chunk of one
// system's entries), the reductions finish per system in task order, and a converged system freezes while the
// others iterate.  One BiCGSTAB iteration over all systems is captured once as a HIP graph and replayed; the host
// reads the per-system states back once per iteration.
#include <functional>
#include <thread>

#include "rsh_solver.hpp"

using namespace rsh;

struct rsh_batch {
    rsh_context* ctx = nullptr;
    struct Sys {
        int n = 0, nb = 1, nnzb = 0, group = 0;
        int rowBase = 0;             // first block row inside its group
        long long off = 0;           // first entry in the flat (group-major) vectors
        long long inOff = 0;         // first entry in the caller's input-order concatenation
        long long valOff = 0;        // first value inside its group's value array
        long long inValOff = 0;      // first value in the caller's input-order value concatenation
    };
    struct Group {
        int nb = 1;
        long long off = 0, N = 0;
        rsh_system* S = nullptr;
        std::vector<int> members;
        std::vector<double> lanesMs;   // the autotuner's SpMV time per candidate (1, 2, 4 lanes per row), ms
    };
    std::vector<Sys> sys;
    std::vector<Group> groups;
    long long N = 0, totalVals = 0;
    DBuf<double> x, b, r, rh, p, v, sv, t, ph, sh, tolD, xo, rt, ax, normD;
    DBuf<double> xPrev, rPrev;    // refinement: iterate and residual before the last correction
    DBuf<int> keepD;
    double* hnorm = nullptr;
    DBuf<int4> task;
    DBuf<int> tStart;
    int T = 0, chunk = 8192;
    DBuf<double> partial;
    DBuf<KScalars> ks;
    KScalars* hks = nullptr;
    rsh_options opt{};
    bool ready = false, tuned = false;
    std::vector<int> fuseChosen;
    int lanesReq = 0;             // SpMV lanes per row: 0 = tuned per group at the next set-up, else 1, 2 or 4
    bool lanesDone = false;       // the groups' SpMV layouts follow lanesReq
    hipGraphExec_t iter = nullptr;
    std::vector<int> graphKey;
    // last solve
    std::vector<int> lastIts, lastStatus;
    std::vector<double> lastRel, lastDoneMs;
    double lastSolveMs = 0, lastSetupMs = 0;
    int lastSweeps = 0, lastOuter = 0;
    bool prec32 = false;          // the preconditioner precision of the running solve
    bool lastFallback = false;    // the last refinement fell back to the FP64 preconditioner

    hipStream_t s() const { return ctx->c.stream; }
    // one context (stream) per group for the numeric set-ups, which run concurrently (one host thread each)
    std::vector<rsh_context*> setupCtx;
    ~rsh_batch() {
        if (iter) (void)hipGraphExecDestroy(iter);
        for (auto& g : groups) delete g.S;
        for (auto* c : setupCtx) rsh_context_destroy(c);
        if (hks) (void)hipHostFree(hks);
        if (hnorm) (void)hipHostFree(hnorm);
    }
    void buildTasks(int ch) {
        chunk = ch;
        std::vector<int4> h;
        std::vector<int> st(1, 0);
        for (size_t i = 0; i < sys.size(); ++i) {
            const long long e0 = sys[i].off, e1 = e0 + (long long)sys[i].n * sys[i].nb;
            for (long long a = e0; a < e1; a += ch) h.push_back(make_int4(int(i), int(a), int(std::min(e1, a + ch)), 0));
            st.push_back(int(h.size()));
        }
        T = int(h.size());
        task.upload(h, s());
        tStart.upload(st, s());
        partial.alloc(size_t(2) * size_t(T));
        dropGraph();
    }
    void dropGraph() {
        if (iter) (void)hipGraphExecDestroy(iter);
        iter = nullptr;
        graphKey.clear();
    }
    bool f32() const { return opt.precision == RSH_MIXED; }
    void precondAll(const double* in, double* out) {
        for (auto& g : groups) g.S->precond(in + g.off, out + g.off, prec32);
    }
    void spmvAll(const double* in, double* out, bool gated = true) {
        for (auto& g : groups) g.S->spmv(in + g.off, out + g.off, false, gated);
    }
    // one BiCGSTAB iteration of every running system (linsolve.cpp:174-212 per system), then the states to the host
    void iteration() {
        hipStream_t st = s();
        const int ns = int(sys.size());
        launchBDots(st, T, task.p, 1, rh.p, r.p, nullptr, nullptr, partial.p);
        launchBFinal(st, ns, 1, tStart.p, partial.p, T, kEpRho, ks.p);
        launchBBicgP(st, T, task.p, ks.p, r.p, v.p, p.p);
        precondAll(p.p, ph.p);
        spmvAll(ph.p, v.p);
        launchBDots(st, T, task.p, 1, rh.p, v.p, nullptr, nullptr, partial.p);
        launchBFinal(st, ns, 1, tStart.p, partial.p, T, kEpAlpha, ks.p);
        launchBBicgS(st, T, task.p, ks.p, r.p, v.p, sv.p, partial.p);
        launchBFinal(st, ns, 1, tStart.p, partial.p, T, kEpSn, ks.p);
        launchBBicgXs(st, T, task.p, ks.p, ph.p, x.p);
        precondAll(sv.p, sh.p);
        spmvAll(sh.p, t.p);
        launchBDots(st, T, task.p, 2, t.p, t.p, t.p, sv.p, partial.p);
        launchBFinal(st, ns, 2, tStart.p, partial.p, T, kEpOmega, ks.p);
        launchBBicgXR(st, T, task.p, ks.p, ph.p, sh.p, sv.p, t.p, x.p, r.p, partial.p);
        launchBFinal(st, ns, 1, tStart.p, partial.p, T, kEpRn, ks.p);
        RSH_HIP(hipMemcpyAsync(hks, ks.p, sizeof(KScalars) * size_t(ns), hipMemcpyDeviceToHost, st));
    }
    std::vector<int> currentKey() const {
        // every option a captured kernel bakes in (the Jacobi weight by its bit pattern)
        std::vector<int> k{chunk, opt.precision, opt.precond, prec32 ? 1 : 0, opt.ilu_lean, opt.smoother, opt.smooth_steps,
                           opt.coarse_smooth_steps};
        int w[2];
        std::memcpy(w, &opt.jacobi_omega, sizeof w);
        k.push_back(w[0]);
        k.push_back(w[1]);
        for (auto& g : groups) {
            k.push_back(g.S->ilu ? g.S->ilu->generation : -1);
            k.push_back(g.S->cpr ? g.S->cpr->generation : -1);
        }
        return k;
    }
    void iterationGraphed() {
        static const bool useGraphs = [] { const char* e = std::getenv("RSH_GRAPHS"); return !(e && e[0] == '0'); }();
        if (!useGraphs) { iteration(); return; }
        const std::vector<int> key = currentKey();
        if (!iter || key != graphKey) {
            dropGraph();
            hipGraph_t g = nullptr;
            RSH_HIP(hipStreamBeginCapture(s(), hipStreamCaptureModeThreadLocal));
            try {
                iteration();
            } catch (...) {
                (void)hipStreamEndCapture(s(), &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            RSH_HIP(hipStreamEndCapture(s(), &g));
            const hipError_t e = hipGraphInstantiate(&iter, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) { iter = nullptr; throw Error(RSH_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e)); }
            graphKey = key;
        }
        RSH_HIP(hipGraphLaunch(iter, s()));
    }
    // BiCGSTAB from x = 0 on every system with right-hand side rhs (flat, device) into x; per-system tolerances;
    // on return hks holds every system's final state; doneAt[s] (ms since t0) is set when s stops
    int bicgstabCore(const double* rhs, const std::vector<double>& tol, int maxIt, double t0, std::vector<double>& doneAt) {
        hipStream_t st = s();
        const int ns = int(sys.size());
        launchFill(st, x.p, 0.0, N);
        launchCopy(st, r.p, rhs, N);
        launchCopy(st, rh.p, rhs, N);
        launchFill(st, p.p, 0.0, N);
        launchFill(st, v.p, 0.0, N);
        RSH_HIP(hipMemcpyAsync(tolD.p, tol.data(), sizeof(double) * size_t(ns), hipMemcpyHostToDevice, st));
        launchBInit(st, ns, tolD.p, ks.p);
        launchBDots(st, T, task.p, 1, rhs, rhs, nullptr, nullptr, partial.p);
        launchBFinal(st, ns, 1, tStart.p, partial.p, T, kEpBn, ks.p);
        RSH_HIP(hipMemcpyAsync(hks, ks.p, sizeof(KScalars) * size_t(ns), hipMemcpyDeviceToHost, st));
        RSH_HIP(hipStreamSynchronize(st));
        int running = 0;
        for (int i = 0; i < ns; ++i) running += hks[i].state == kRun;
        // the host reads the states every sweep once some running system is near its tolerance, else every fourth
        // sweep (a stopped system's state is frozen by the epilogues, so the sweeps in between change nothing for it;
        // its latency is then taken at the next read)
        constexpr int kBatchCheckEvery = 4;
        int sweeps = 0, nextCheck = 1;
        while (running > 0 && sweeps < maxIt) {
            iterationGraphed();
            ++sweeps;
            if (sweeps < nextCheck && sweeps < maxIt) continue;
            RSH_HIP(hipStreamSynchronize(st));
            const double now = nowMs() - t0;
            running = 0;
            double worst = 0;
            for (int i = 0; i < ns; ++i) {
                if (hks[i].state == kRun) {
                    ++running;
                    const double rel = hks[i].bn > 0 ? hks[i].rn / (hks[i].bn * hks[i].tol) : 0.0;
                    worst = std::max(worst, std::isfinite(rel) ? rel : 1e300);
                } else if (doneAt[size_t(i)] == 0.0) {
                    doneAt[size_t(i)] = now;
                }
            }
            nextCheck = sweeps + (worst > 30.0 ? kBatchCheckEvery : 1);
        }
        RSH_HIP(hipStreamSynchronize(st));
        return sweeps;
    }
    // per-system 2-norms of a flat vector (host)
    void norms(const double* vec, std::vector<double>& out) {
        const int ns = int(sys.size());
        launchBDots(s(), T, task.p, 1, vec, vec, nullptr, nullptr, partial.p);
        launchBFinal(s(), ns, 1, tStart.p, partial.p, T, kEpNone, nullptr, normD.p);
        RSH_HIP(hipMemcpyAsync(hnorm, normD.p, sizeof(double) * size_t(ns), hipMemcpyDeviceToHost, s()));
        RSH_HIP(hipStreamSynchronize(s()));
        out.resize(size_t(ns));
        for (int i = 0; i < ns; ++i) out[size_t(i)] = std::sqrt(hnorm[i]);
    }
    // b and x in the flat layout (device); per-system results in the last* vectors.  FP64 preconditioner: one
    // BiCGSTAB per system to its tolerance on the recursive residual (linsolve.cpp:210-212).  FP32 preconditioner:
    // FP64 iterative refinement (bsr.cpp:1373-1396) - x += d with d from a BiCGSTAB on the true residual
    // r = b - A x (FP64 SpMV), each system until its true residual meets its tolerance; the restarts keep BiCGSTAB
    // from stalling on the FP32 rounding of the preconditioner.
    void solveFlat(const double* tol, int maxIt) {
        hipStream_t st = s();
        const int ns = int(sys.size());
        const double t0 = nowMs();
        std::vector<double> tolV(tol, tol + ns);
        lastDoneMs.assign(size_t(ns), 0.0);
        lastIts.assign(size_t(ns), 0);
        lastRel.assign(size_t(ns), 0.0);
        lastStatus.assign(size_t(ns), RSH_OK);
        lastOuter = 0;
        lastFallback = false;
        prec32 = f32();
        if (!f32()) {
            lastSweeps = bicgstabCore(b.p, tolV, maxIt, t0, lastDoneMs);
            const double now = nowMs() - t0;
            for (int i = 0; i < ns; ++i) {
                const KScalars& k = hks[i];
                if (lastDoneMs[size_t(i)] == 0.0) lastDoneMs[size_t(i)] = now;
                lastIts[size_t(i)] = k.it;
                const bool onS = k.state == kConvS || k.state == kDoneS;
                lastRel[size_t(i)] = k.bn > 0 ? (onS ? k.sn : k.rn) / k.bn : 0.0;
                if (!std::isfinite(k.bn) || !std::isfinite(k.rn)) lastStatus[size_t(i)] = RSH_ERR_NONFINITE;
                else if (k.state == kConv || onS) lastStatus[size_t(i)] = RSH_OK;
                else if (k.state == kBreak) lastStatus[size_t(i)] = RSH_ERR_BREAKDOWN;
                else lastStatus[size_t(i)] = RSH_ERR_NOT_CONVERGED;
            }
            lastSolveMs = now;
            return;
        }
        const double innerTol = opt.inner_tol > 0 ? opt.inner_tol : 1e-4;
        const int maxOuter = opt.max_outer > 0 ? opt.max_outer : 10;
        launchFill(st, xo.p, 0.0, N);
        launchCopy(st, rt.p, b.p, N);
        std::vector<double> bn, rn;
        norms(b.p, bn);
        rn = bn;
        std::vector<int> keep(size_t(ns), 0), stuck(size_t(ns), 0);
        std::vector<double> itol(size_t(ns), 1.0), innerDone(size_t(ns), 0.0);
        int used = 0;
        for (int outer = 0;; ++outer) {
            const double now = nowMs() - t0;
            int active = 0;
            for (int i = 0; i < ns; ++i) {
                const bool done = bn[size_t(i)] == 0.0 || rn[size_t(i)] / bn[size_t(i)] <= tolV[size_t(i)] || stuck[size_t(i)];
                keep[size_t(i)] = done ? 0 : 1;
                if (done && lastDoneMs[size_t(i)] == 0.0) lastDoneMs[size_t(i)] = now;
                if (!done) {
                    ++active;
                    itol[size_t(i)] = std::max(innerTol, 0.5 * tolV[size_t(i)] * bn[size_t(i)] / rn[size_t(i)]);
                }
            }
            lastOuter = outer;
            if (active == 0 || outer >= maxOuter || used >= maxIt) break;
            // systems already converged get a zero right-hand side: their correction is zero at once
            keepD.upload(keep, st);
            launchBMask(st, T, task.p, keepD.p, rt.p);
            std::fill(innerDone.begin(), innerDone.end(), 0.0);
            // a stalled inner solve is restarted by the next refinement step
            used += bicgstabCore(rt.p, itol, std::min(maxIt - used, std::max(50, maxIt / maxOuter)), t0, innerDone);
            for (int i = 0; i < ns; ++i) {
                if (!keep[size_t(i)]) continue;
                lastIts[size_t(i)] += hks[i].it;
                if (hks[i].state == kBreak) lastStatus[size_t(i)] = RSH_ERR_BREAKDOWN;
            }
            launchCopy(st, xPrev.p, xo.p, N);
            launchCopy(st, rPrev.p, rt.p, N);
            launchBAdd(st, T, task.p, x.p, xo.p);
            spmvAll(xo.p, ax.p, false);    // true residual of every system (no gating by the inner solve's states)
            launchSub(st, N, b.p, ax.p, rt.p);
            std::vector<double> rnew;
            norms(rt.p, rnew);
            bool stalled = false;
            for (int i = 0; i < ns; ++i) {
                if (!keep[size_t(i)]) continue;
                if (rnew[size_t(i)] < rn[size_t(i)]) { rn[size_t(i)] = rnew[size_t(i)]; continue; }
                // no progress (or a non-finite correction): this system keeps its previous iterate and residual
                const Sys& q = sys[size_t(i)];
                const size_t bytes = sizeof(double) * size_t(q.n) * q.nb;
                RSH_HIP(hipMemcpyAsync(xo.p + q.off, xPrev.p + q.off, bytes, hipMemcpyDeviceToDevice, st));
                RSH_HIP(hipMemcpyAsync(rt.p + q.off, rPrev.p + q.off, bytes, hipMemcpyDeviceToDevice, st));
                if (!prec32) stuck[size_t(i)] = 1;   // stalled with the FP64 preconditioner as well: give up on it
                stalled = true;
            }
            if (stalled && prec32) {
                // the FP32 preconditioner stalls on some system: the remaining steps use the FP64 one
                prec32 = false;
                lastFallback = true;
            }
        }
        const double now = nowMs() - t0;
        for (int i = 0; i < ns; ++i) {
            if (lastDoneMs[size_t(i)] == 0.0) lastDoneMs[size_t(i)] = now;
            lastRel[size_t(i)] = bn[size_t(i)] > 0 ? rn[size_t(i)] / bn[size_t(i)] : 0.0;
            if (!std::isfinite(bn[size_t(i)]) || !std::isfinite(rn[size_t(i)])) lastStatus[size_t(i)] = RSH_ERR_NONFINITE;
            else if (lastRel[size_t(i)] <= tolV[size_t(i)]) lastStatus[size_t(i)] = RSH_OK;
            else if (lastStatus[size_t(i)] == RSH_OK) lastStatus[size_t(i)] = RSH_ERR_NOT_CONVERGED;
        }
        launchCopy(st, x.p, xo.p, N);
        lastSweeps = used;
        lastSolveMs = now;
    }
    // input-order concatenation <-> flat layout (device to device)
    void gatherIn(const double* in, double* flat) {
        for (auto& q : sys)
            RSH_HIP(hipMemcpyAsync(flat + q.off, in + q.inOff, sizeof(double) * size_t(q.n) * q.nb, hipMemcpyDeviceToDevice, s()));
    }
    void scatterOut(const double* flat, double* out) {
        for (auto& q : sys)
            RSH_HIP(hipMemcpyAsync(out + q.inOff, flat + q.off, sizeof(double) * size_t(q.n) * q.nb, hipMemcpyDeviceToDevice, s()));
    }
    double timeMs(const std::function<void()>& f, int reps) {
        hipEvent_t e0, e1;
        RSH_HIP(hipEventCreate(&e0));
        RSH_HIP(hipEventCreate(&e1));
        f();
        RSH_HIP(hipEventRecord(e0, s()));
        for (int i = 0; i < reps; ++i) f();
        RSH_HIP(hipEventRecord(e1, s()));
        RSH_HIP(hipEventSynchronize(e1));
        float ms = 0;
        RSH_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return double(ms) / reps;
    }
    void tune();
    void tuneLanes();
};

namespace {

// Autotuner results per (block size, log2 of the group's block rows) and per log2 of the batch's vector length:
// each candidate is timed once per process with HIP events on the context stream, then reused.
std::mutex g_tuneMu;
std::map<std::pair<int, int>, std::pair<int, int>> g_fuseCache;   // -> (fuseMax, run rows)
std::map<int, int> g_chunkCache;
std::map<std::pair<int, int>, std::pair<int, std::vector<double>>> g_lanesCache;   // -> SpMV lanes per row, ms per candidate

int log2Bucket(long long v) {
    int b = 0;
    while ((1ll << (b + 1)) <= v) ++b;
    return b;
}

}  // namespace

// Tunes (1) the level chunking of each group's ILU sweeps: the largest level (in tasks) swept inside a run of small
// levels, i.e. how many levels one workgroup or one dense-group pass covers, and the rows a run may hold (the LDS
// tile of its working vector, 512 or 2048 rows); (2) the entries per workgroup of the
// batched vector kernels and reductions (rows per wavefront = entries / 256 threads).  Candidates are timed on the
// resident data; the fastest wins.  Runs at every set-up: the choices are cached per (block size, ordering, size
// bucket), so a set-up that builds new factors (another preconditioner) only looks its choice up; groups whose
// preconditioner has no ILU (block-Jacobi second stage) are skipped.
void rsh_batch::tune() {
    hipStream_t st = s();
    std::lock_guard<std::mutex> lk(g_tuneMu);
    bool armed = false;
    auto arm = [&] {
        // every system running while the candidates are timed
        if (armed) return;
        std::vector<double> tol(sys.size(), 1e-8);
        RSH_HIP(hipMemcpyAsync(tolD.p, tol.data(), sizeof(double) * tol.size(), hipMemcpyHostToDevice, st));
        launchBInit(st, int(sys.size()), tolD.p, ks.p);
        armed = true;
    };
    fuseChosen.assign(groups.size(), 0);
    for (size_t gi = 0; gi < groups.size(); ++gi) {
        Group& g = groups[gi];
        if (!g.S->ilu) continue;
        IluDev& I = *g.S->ilu;
        const auto key = std::make_pair(g.nb * 1000 + (I.mc ? 1 : 0), log2Bucket(g.S->n));
        auto it = g_fuseCache.find(key);
        std::pair<int, int> best{16, kRunRowsMax};
        if (it != g_fuseCache.end()) {
            best = it->second;
        } else {
            arm();
            double bestMs = 1e300;
            launchFill(st, p.p + g.off, 1.0, g.N);
            for (int rows : {kRunRowsMax / 4, kRunRowsMax})
                for (int cand : {4, 16, 64}) {
                    I.retune(cand, st, rows);
                    const double ms = timeMs([&] { g.S->precond(p.p + g.off, ph.p + g.off, f32()); }, 3);
                    if (ms < bestMs) { bestMs = ms; best = {cand, rows}; }
                }
            g_fuseCache[key] = best;
        }
        if (I.fuseMax != best.first || I.runRowsMax != best.second) I.retune(best.first, st, best.second);
        fuseChosen[gi] = best.first;
    }
    if (!tuned) {
        const int ck = log2Bucket(N);
        auto it = g_chunkCache.find(ck);
        if (it != g_chunkCache.end()) {
            if (chunk != it->second) buildTasks(it->second);
        } else {
            arm();
            double bestMs = 1e300;
            int best = chunk;
            launchFill(st, r.p, 1.0, N);
            launchFill(st, v.p, 1.0, N);
            for (int cand : {2048, 8192, 32768}) {
                buildTasks(cand);
                launchBInit(st, int(sys.size()), tolD.p, ks.p);
                const double ms = timeMs([&] {
                    launchBDots(st, T, task.p, 2, r.p, v.p, r.p, r.p, partial.p);
                    launchBBicgP(st, T, task.p, ks.p, r.p, v.p, p.p);
                    launchBBicgS(st, T, task.p, ks.p, r.p, v.p, sv.p, partial.p);
                }, 5);
                if (ms < bestMs) { bestMs = ms; best = cand; }
            }
            g_chunkCache[ck] = best;
            if (chunk != best) buildTasks(best);
        }
    }
    tuneLanes();
    RSH_HIP(hipStreamSynchronize(st));
    tuned = true;
}

// (3) the rows per wavefront of each group's SpMV: 64, 32 or 16 (1, 2 or 4 lanes per row, laneSliceLayout), timed on
// the group's resident operator (ungated) and cached per (block size, size bucket); or the caller's fixed choice.
void rsh_batch::tuneLanes() {
    if (lanesDone) return;
    hipStream_t st = s();
    bool changed = false;
    for (auto& g : groups) {
        int best = lanesReq;
        g.lanesMs.clear();
        if (best == 0) {
            const auto key = std::make_pair(g.nb, log2Bucket(g.S->n));
            auto it = g_lanesCache.find(key);
            if (it != g_lanesCache.end()) {
                best = it->second.first;
                g.lanesMs = it->second.second;
            } else {
                launchFill(st, p.p + g.off, 1.0, g.N);
                double bestMs = 1e300;
                for (int cand : {1, 2, 4}) {
                    if (g.S->spmvLanes() != cand) changed = true;
                    g.S->setSpmvLanes(cand);
                    const double ms = timeMs([&] { g.S->spmv(p.p + g.off, ph.p + g.off, false, false); }, 5);
                    g.lanesMs.push_back(ms);
                    if (ms < bestMs) { bestMs = ms; best = cand; }
                }
                g_lanesCache[key] = {best, g.lanesMs};
            }
        }
        if (g.S->spmvLanes() != best) changed = true;
        g.S->setSpmvLanes(best);
    }
    if (changed) dropGraph();
    lanesDone = true;
}

extern "C" {

int rsh_batch_create(rsh_context* ctx, int count, const int* n, const int* nb, const int* const* row_ptr, const int* const* col,
                     rsh_batch** out) {
    return guarded([&] {
        if (!ctx || !n || !nb || !row_ptr || !col || !out) return fail(RSH_ERR_ARG, "null argument");
        *out = nullptr;
        if (count <= 0) return fail(RSH_ERR_ARG, "count must be positive");
        auto B = std::make_unique<rsh_batch>();
        B->ctx = ctx;
        B->sys.resize(size_t(count));
        long long inOff = 0, inVal = 0;
        for (int i = 0; i < count; ++i) {
            if (n[i] <= 0) return fail(RSH_ERR_ARG, "system " + std::to_string(i) + ": n must be positive");
            if (nb[i] < 1 || nb[i] > kMaxNb) return fail(RSH_ERR_UNSUPPORTED, "system " + std::to_string(i) + ": block size must be 1..4");
            if (!row_ptr[i] || !col[i]) return fail(RSH_ERR_ARG, "system " + std::to_string(i) + ": null pattern");
            auto& q = B->sys[size_t(i)];
            q.n = n[i];
            q.nb = nb[i];
            q.nnzb = row_ptr[i][n[i]];
            q.inOff = inOff;
            q.inValOff = inVal;
            inOff += (long long)q.n * q.nb;
            inVal += (long long)q.nnzb * q.nb * q.nb;
        }
        B->totalVals = inVal;
        // groups by block size (ascending), systems in input order inside a group
        long long off = 0;
        for (int bs = 1; bs <= kMaxNb; ++bs) {
            rsh_batch::Group g;
            g.nb = bs;
            for (int i = 0; i < count; ++i)
                if (nb[i] == bs) g.members.push_back(i);
            if (g.members.empty()) continue;
            g.off = off;
            int rows = 0;
            long long vals = 0;
            std::vector<int> rp(1, 0), ci;
            for (int i : g.members) {
                auto& q = B->sys[size_t(i)];
                q.group = int(B->groups.size());
                q.rowBase = rows;
                q.off = off + (long long)rows * bs;
                q.valOff = vals;
                for (int r = 0; r < q.n; ++r) {
                    if (row_ptr[i][r + 1] < row_ptr[i][r]) return fail(RSH_ERR_ARG, "system " + std::to_string(i) + ": row_ptr not monotone");
                    for (int p = row_ptr[i][r]; p < row_ptr[i][r + 1]; ++p) {
                        const int c = col[i][p];
                        if (c < 0 || c >= q.n) return fail(RSH_ERR_ARG, "system " + std::to_string(i) + ": column index out of range");
                        ci.push_back(c + rows);
                    }
                    rp.push_back(int(ci.size()));
                }
                if ((long long)rows + q.n > 0x7fffffffll || (long long)ci.size() > 0x7fffffffll)
                    return fail(RSH_ERR_UNSUPPORTED, "a block-size group exceeds 2^31 rows or blocks");
                rows += q.n;
                vals += (long long)q.nnzb * bs * bs;
            }
            g.N = (long long)rows * bs;
            rsh_system* S = nullptr;
            const int rc = rsh_system_create(ctx, rows, bs, rp.data(), ci.data(), &S);
            if (rc != RSH_OK) return rc;
            S->keepIluEntries = true;
            g.S = S;
            // block row -> system of the batch; slices and long rows of stopped systems are skipped
            S->rowSys.resize(size_t(rows));
            for (int i : g.members) {
                const auto& q = B->sys[size_t(i)];
                std::fill(S->rowSys.begin() + q.rowBase, S->rowSys.begin() + q.rowBase + q.n, i);
            }
            sliceOwners(S->spmvHost, S->rowSys);
            S->spmv64.owner.upload(S->spmvHost.owner, ctx->c.stream);
            S->spmv64.lowner.upload(S->spmvHost.lowner, ctx->c.stream);
            off += g.N;
            B->groups.push_back(std::move(g));
        }
        B->N = off;
        if (B->N > 0x7fffffffll) return fail(RSH_ERR_UNSUPPORTED, "a batch is limited to 2^31 vector entries");
        for (auto* bf : {&B->x, &B->b, &B->r, &B->rh, &B->p, &B->v, &B->sv, &B->t, &B->ph, &B->sh}) bf->alloc(size_t(B->N));
        B->tolD.alloc(size_t(count));
        B->normD.alloc(size_t(count));
        RSH_HIP(hipHostMalloc(reinterpret_cast<void**>(&B->hnorm), sizeof(double) * size_t(count), hipHostMallocDefault));
        B->ks.alloc(size_t(count));
        RSH_HIP(hipHostMalloc(reinterpret_cast<void**>(&B->hks), sizeof(KScalars) * size_t(count), hipHostMallocDefault));
        B->buildTasks(8192);
        for (auto& g : B->groups) g.S->gate = B->ks.p;
        RSH_HIP(hipStreamSynchronize(B->s()));
        *out = B.release();
        return RSH_OK;
    });
}

int rsh_batch_destroy(rsh_batch* batch) {
    if (!batch) return RSH_OK;
    return guarded([&] {
        RSH_HIP(hipSetDevice(batch->ctx->c.device));
        (void)hipStreamSynchronize(batch->s());
        delete batch;
        return RSH_OK;
    });
}

// host pointers: val[i] = nnzb_i * nb_i^2 doubles of system i
int rsh_batch_set_values(rsh_batch* batch, const double* const* val) {
    return guarded([&] {
        if (!batch || !val) return fail(RSH_ERR_ARG, "null argument");
        RSH_HIP(hipSetDevice(batch->ctx->c.device));
        for (size_t i = 0; i < batch->sys.size(); ++i) {
            const auto& q = batch->sys[i];
            if (!val[i]) return fail(RSH_ERR_ARG, "system " + std::to_string(i) + ": null values");
            rsh_system* S = batch->groups[size_t(q.group)].S;
            RSH_HIP(hipMemcpyAsync(S->val.p + q.valOff, val[i], sizeof(double) * size_t(q.nnzb) * q.nb * q.nb, hipMemcpyHostToDevice, batch->s()));
        }
        for (auto& g : batch->groups) g.S->valuesChanged();
        RSH_HIP(hipStreamSynchronize(batch->s()));
        batch->ready = false;
        return RSH_OK;
    });
}

// device pointer: the input-order concatenation of every system's values
int rsh_batch_set_values_device(rsh_batch* batch, const double* val_dev) {
    return guarded([&] {
        if (!batch || !val_dev) return fail(RSH_ERR_ARG, "null argument");
        RSH_HIP(hipSetDevice(batch->ctx->c.device));
        for (const auto& q : batch->sys) {
            rsh_system* S = batch->groups[size_t(q.group)].S;
            RSH_HIP(hipMemcpyAsync(S->val.p + q.valOff, val_dev + q.inValOff, sizeof(double) * size_t(q.nnzb) * q.nb * q.nb,
                                   hipMemcpyDeviceToDevice, batch->s()));
        }
        for (auto& g : batch->groups) g.S->valuesChanged();
        batch->ready = false;
        return RSH_OK;
    });
}

int rsh_batch_setup(rsh_batch* batch, const rsh_options* opt) {
    return guarded([&] {
        if (!batch || !opt) return fail(RSH_ERR_ARG, "null argument");
        if (opt->precision != RSH_FP64 && opt->precision != RSH_MIXED && opt->precision != RSH_MIXED_STORE)
            return fail(RSH_ERR_UNSUPPORTED, "batched solves run FP64 BiCGSTAB with an FP64, FP32 or FP32-stored preconditioner");
        if (opt->krylov != RSH_KRYLOV_BICGSTAB) return fail(RSH_ERR_UNSUPPORTED, "batched solves use BiCGSTAB");
        RSH_HIP(hipSetDevice(batch->ctx->c.device));
        const double t0 = nowMs();
        // the groups are independent systems: their set-ups (host aggregation and short latency-bound kernels)
        // run concurrently, each on a stream of its own, after the batch stream's earlier work (new values)
        const size_t G = batch->groups.size();
        RSH_HIP(hipStreamSynchronize(batch->s()));
        while (batch->setupCtx.size() < G) {
            rsh_context* c = nullptr;
            const int rc = rsh_context_create(batch->ctx->c.device, &c);
            if (rc != RSH_OK) return rc;
            batch->setupCtx.push_back(c);
        }
        std::vector<int> rcs(G, RSH_OK);
        std::vector<std::string> errs(G);
        std::vector<std::thread> th;
        th.reserve(G);
        for (size_t gi = 0; gi < G; ++gi)
            th.emplace_back([&, gi] {
                rsh_system* S = batch->groups[gi].S;
                S->ctx = batch->setupCtx[gi];
                rcs[gi] = rsh_system_setup(S, opt);         // ends with a synchronisation of its stream
                if (rcs[gi] != RSH_OK) errs[gi] = rsh_last_error();
                S->ctx = batch->ctx;
            });
        for (auto& t : th) t.join();
        for (size_t gi = 0; gi < G; ++gi)
            if (rcs[gi] != RSH_OK) return fail(rcs[gi], errs[gi]);
        batch->opt = *opt;
        if (opt->precision == RSH_MIXED)
            for (auto* bf : {&batch->xo, &batch->rt, &batch->ax, &batch->xPrev, &batch->rPrev}) bf->alloc(size_t(batch->N));
        batch->tune();
        RSH_HIP(hipStreamSynchronize(batch->s()));
        batch->ready = true;
        batch->lastSetupMs = nowMs() - t0;
        return RSH_OK;
    });
}

// device pointers, input-order concatenations; tol[i] per system; its/rel/status per system (host arrays, may be null)
int rsh_batch_solve_device(rsh_batch* batch, const double* b_dev, double* x_dev, const double* tol, int max_it, int* iterations,
                           double* rel_res, int* status) {
    return guarded([&] {
        if (!batch || !b_dev || !x_dev || !tol) return fail(RSH_ERR_ARG, "null argument");
        if (!batch->ready) return fail(RSH_ERR_ARG, "rsh_batch_solve: call rsh_batch_setup first");
        if (max_it < 1) return fail(RSH_ERR_ARG, "max_it must be positive");
        RSH_HIP(hipSetDevice(batch->ctx->c.device));
        batch->gatherIn(b_dev, batch->b.p);
        batch->solveFlat(tol, max_it);
        batch->scatterOut(batch->x.p, x_dev);
        RSH_HIP(hipStreamSynchronize(batch->s()));
        int worst = RSH_OK;
        for (size_t i = 0; i < batch->sys.size(); ++i) {
            if (iterations) iterations[i] = batch->lastIts[i];
            if (rel_res) rel_res[i] = batch->lastRel[i];
            if (status) status[i] = batch->lastStatus[i];
            if (batch->lastStatus[i] != RSH_OK && worst == RSH_OK) worst = batch->lastStatus[i];
        }
        if (worst != RSH_OK) return fail(worst, "one or more systems of the batch did not converge (see status)");
        return RSH_OK;
    });
}

// host pointers per system: b[i], x[i] of n_i * nb_i doubles
int rsh_batch_solve(rsh_batch* batch, const double* const* b, double* const* x, const double* tol, int max_it, int* iterations,
                    double* rel_res, int* status) {
    return guarded([&] {
        if (!batch || !b || !x || !tol) return fail(RSH_ERR_ARG, "null argument");
        if (!batch->ready) return fail(RSH_ERR_ARG, "rsh_batch_solve: call rsh_batch_setup first");
        if (max_it < 1) return fail(RSH_ERR_ARG, "max_it must be positive");
        RSH_HIP(hipSetDevice(batch->ctx->c.device));
        for (size_t i = 0; i < batch->sys.size(); ++i) {
            const auto& q = batch->sys[i];
            if (!b[i] || !x[i]) return fail(RSH_ERR_ARG, "system " + std::to_string(i) + ": null vector");
            RSH_HIP(hipMemcpyAsync(batch->b.p + q.off, b[i], sizeof(double) * size_t(q.n) * q.nb, hipMemcpyHostToDevice, batch->s()));
        }
        batch->solveFlat(tol, max_it);
        for (size_t i = 0; i < batch->sys.size(); ++i) {
            const auto& q = batch->sys[i];
            RSH_HIP(hipMemcpyAsync(x[i], batch->x.p + q.off, sizeof(double) * size_t(q.n) * q.nb, hipMemcpyDeviceToHost, batch->s()));
        }
        RSH_HIP(hipStreamSynchronize(batch->s()));
        int worst = RSH_OK;
        for (size_t i = 0; i < batch->sys.size(); ++i) {
            if (iterations) iterations[i] = batch->lastIts[i];
            if (rel_res) rel_res[i] = batch->lastRel[i];
            if (status) status[i] = batch->lastStatus[i];
            if (batch->lastStatus[i] != RSH_OK && worst == RSH_OK) worst = batch->lastStatus[i];
        }
        if (worst != RSH_OK) return fail(worst, "one or more systems of the batch did not converge (see status)");
        return RSH_OK;
    });
}

// per-system milliseconds from the start of the last solve to the host seeing the system finished
int rsh_batch_latencies(rsh_batch* batch, double* ms) {
    return guarded([&] {
        if (!batch || !ms) return fail(RSH_ERR_ARG, "null argument");
        if (batch->lastDoneMs.size() != batch->sys.size()) return fail(RSH_ERR_ARG, "no solve yet");
        std::copy(batch->lastDoneMs.begin(), batch->lastDoneMs.end(), ms);
        return RSH_OK;
    });
}

int rsh_batch_set_spmv_lanes(rsh_batch* batch, int lanes) {
    return guarded([&] {
        if (!batch) return fail(RSH_ERR_ARG, "null batch");
        if (lanes != 0 && lanes != 1 && lanes != 2 && lanes != 4) return fail(RSH_ERR_ARG, "lanes per row must be 0 (tuned), 1, 2 or 4");
        if (lanes != batch->lanesReq) {
            batch->lanesReq = lanes;
            batch->lanesDone = false;
            batch->ready = false;                    // applied by the next rsh_batch_setup
        }
        return RSH_OK;
    });
}

int rsh_batch_info(rsh_batch* batch, char* json, int cap) {
    return guarded([&] {
        if (!batch || !json || cap <= 0) return fail(RSH_ERR_ARG, "null argument");
        std::ostringstream o;
        o << "{\"systems\":" << batch->sys.size() << ",\"entries\":" << batch->N << ",\"tasks\":" << batch->T << ",\"chunk\":" << batch->chunk
          << ",\"groups\":[";
        for (size_t gi = 0; gi < batch->groups.size(); ++gi) {
            const auto& g = batch->groups[gi];
            o << (gi ? "," : "") << "{\"nb\":" << g.nb << ",\"systems\":" << g.members.size() << ",\"rows\":" << g.S->n << ",\"blocks\":" << g.S->nnzb;
            if (g.S->ilu) {
                o << ",\"ilu_levels\":" << g.S->ilu->sL.levels() << ",\"fuse_max\":" << g.S->ilu->fuseMax << ",\"run_rows\":"
                  << g.S->ilu->runRowsMax << ",\"dense_groups\":["
                  << g.S->ilu->denseL.ngroups << "," << g.S->ilu->denseU.ngroups << "]";
            }
            o << ",\"spmv_lanes\":" << g.S->spmvLanes();
            if (!g.lanesMs.empty())
                o << ",\"spmv_lanes_ms\":[" << g.lanesMs[0] << "," << g.lanesMs[1] << "," << g.lanesMs[2] << "]";
            if (g.S->cpr && g.S->cpr->built)
                o << ",\"amg_levels\":" << g.S->cpr->lv.size() << ",\"amg_coarsest\":" << g.S->cpr->m << ",\"amg_dense_coarsest\":"
                  << (g.S->cpr->dense ? 1 : 0);
            o << "}";
        }
        o << "],\"last_sweeps\":" << batch->lastSweeps << ",\"last_outer\":" << batch->lastOuter << ",\"last_fallback_fp64\":" << (batch->lastFallback ? 1 : 0) << ",\"last_solve_ms\":" << batch->lastSolveMs << ",\"last_setup_ms\":" << batch->lastSetupMs
          << ",\"tuned\":" << (batch->tuned ? 1 : 0) << "}";
        const std::string str = o.str();
        if (int(str.size()) + 1 > cap) return fail(RSH_ERR_ARG, "buffer too small");
        std::memcpy(json, str.c_str(), str.size() + 1);
        return RSH_OK;
    });
}

}  // extern "C"

// ------------------------------------------------------------------ the serving loop's access to a batch's members
namespace rsh {

BatchMember batchMember(rsh_batch* b, int i) {
    if (!b || i < 0 || i >= int(b->sys.size())) throw Error(RSH_ERR_ARG, "batch member out of range");
    const auto& q = b->sys[size_t(i)];
    rsh_system* S = b->groups[size_t(q.group)].S;
    return BatchMember{S->val.p, S->rp.p + q.rowBase, q.valOff, q.inOff, q.n, q.nb, q.nnzb};
}

void batchValuesChanged(rsh_batch* b) {
    for (auto& g : b->groups) g.S->valuesChanged();
    b->ready = false;
}

hipStream_t batchStream(rsh_batch* b) { return b->s(); }

size_t batchDeviceBytes(rsh_batch* b) {
    size_t s = 0;
    for (auto& g : b->groups) s += g.S->deviceBytes();
    for (auto* bf : {&b->x, &b->b, &b->r, &b->rh, &b->p, &b->v, &b->sv, &b->t, &b->ph, &b->sh, &b->xo, &b->rt, &b->ax, &b->xPrev, &b->rPrev})
        s += bf->bytes();
    return s + b->partial.bytes() + b->task.bytes();
}

}  // namespace rsh
