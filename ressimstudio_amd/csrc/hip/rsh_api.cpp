// C-ABI and solver drivers of the MI355X linear-solve library (include/ressim_hip.h).
//
// A system keeps its pattern, values, slice layouts, ILU schedule and AMG hierarchy resident on the device, so
// a Newton iteration only moves new values (set_values), refreshes the preconditioner numerically (setup) and
// runs the Krylov loop (solve).  All Krylov scalars live on the device; BiCGSTAB reads them back once per
// iteration, FGMRES once per Arnoldi step.
#include "rsh_solver.hpp"

using namespace rsh;

namespace {

// iterations between host reads of the Krylov scalars while the residual is far from the tolerance
constexpr int kCheckEvery = 4;

// device-time spans of a context (rsh_context_device_seconds)
void spanBegin(Ctx& c) {
    if (c.spanOpen) return;
    RSH_HIP(hipEventRecord(c.ev0, c.stream));
    c.spanOpen = true;
}
void spanEnd(Ctx& c) {
    if (!c.spanOpen) return;
    c.spanOpen = false;
    RSH_HIP(hipEventRecord(c.ev1, c.stream));
    RSH_HIP(hipEventSynchronize(c.ev1));
    float ms = 0;
    RSH_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
    c.deviceMs += ms;
}

// The buffers one BiCGSTAB solve works in: the Krylov vectors (N doubles each), the device Krylov scalars and their
// pinned host mirror, the reductions' partial sums (2 * kRedBlocks doubles) and a dot-product output.  The resident
// solve takes them from the system and its context; rsh_solve_device_noalloc from a static device array.
struct BicgWork {
    double *r, *rh, *p, *v, *sv, *t, *ph, *sh;
    KScalars* ks;
    KScalars* hks;
    double* red;
    double* scal;
    bool graphs;     // the preconditioner through captured graphs (instantiated on first use)
};

BicgWork residentWork(rsh_system& S) {
    for (auto* bf : {&S.r, &S.rh, &S.p, &S.v, &S.sv, &S.t, &S.ph, &S.sh}) S.ensureVec(*bf);
    Ctx& c = S.ctx->c;
    return BicgWork{S.r.p, S.rh.p, S.p.p, S.v.p, S.sv.p, S.t.p, S.ph.p, S.sh.p, c.ks.p, c.hks, c.red.p, c.scal.p, true};
}

void readScalars(hipStream_t st, const BicgWork& W) {
    RSH_HIP(hipMemcpyAsync(W.hks, W.ks, sizeof(KScalars), hipMemcpyDeviceToHost, st));
    RSH_HIP(hipStreamSynchronize(st));
}

// Right-preconditioned BiCGSTAB from x0 = 0 (linsolve.cpp:159-216); the loop structure, breakdown tests and
// the two convergence tests (on ||s|| and on the recursive ||r||) are the reference's.
// precF32: the preconditioner's factors, hierarchy and sweeps in FP32; opF32: the operator too (inner solves of the
// refinement loop).  Vectors, dot products and the Krylov recurrences stay FP64.
int bicgstabIn(rsh_system& S, const BicgWork& W, const double* b, double* x, double tol, int maxIt, int& its, double& rel, bool precF32,
               bool opF32) {
    hipStream_t st = S.ctx->c.stream;
    const long long N = S.N;
    launchFill(st, x, 0.0, N);
    launchCopy(st, W.r, b, N);
    launchCopy(st, W.rh, b, N);
    launchFill(st, W.p, 0.0, N);
    launchFill(st, W.v, 0.0, N);
    KScalars init{};
    init.rho = 1; init.alpha = 1; init.omega = 1; init.tol = tol; init.state = kRun; init.it = 0;
    *W.hks = init;
    RSH_HIP(hipMemcpyAsync(W.ks, W.hks, sizeof(KScalars), hipMemcpyHostToDevice, st));
    launchDots(st, W.red, N, 1, b, b, nullptr, nullptr, W.scal, kEpBn, W.ks);
    readScalars(st, W);
    its = 0;
    rel = 0;
    if (W.hks->bn == 0) { S.lastStop = "zero rhs"; return RSH_OK; }
    if (!std::isfinite(W.hks->bn)) { rel = W.hks->bn; S.lastStop = "non-finite right-hand side"; return RSH_ERR_NONFINITE; }
    auto prec = [&](const double* in, double* out) {
        if (W.graphs) S.precondG(in, out, precF32);
        else S.precond(in, out, precF32);
    };
    // The stopping tests run on the device (the epilogues freeze the state), so the host reads the scalars only
    // every few iterations while the residual is far from the tolerance; iterations launched after a stop change
    // nothing.
    // rho = (rh, r) of the first iteration; the later ones come with the closing reduction of the previous iteration
    launchDots(st, W.red, N, 1, W.rh, W.r, nullptr, nullptr, W.scal, kEpRho, W.ks);
    int nextCheck = 1;
    for (int it = 1; it <= maxIt; ++it) {
        launchBicgP(st, N, W.ks, W.r, W.v, W.p);
        prec(W.p, W.ph);
        S.spmv(W.ph, W.v, opF32);
        launchDots(st, W.red, N, 1, W.rh, W.v, nullptr, nullptr, W.scal, kEpAlpha, W.ks);
        launchBicgS(st, W.red, N, W.ks, W.r, W.v, W.sv);
        prec(W.sv, W.sh);
        S.spmv(W.sh, W.t, opF32);
        launchDots(st, W.red, N, 2, W.t, W.t, W.t, W.sv, W.scal, kEpOmega, W.ks);
        launchBicgXR(st, W.red, N, W.ks, W.ph, W.sh, W.sv, W.t, W.rh, x, W.r);
        if (it < nextCheck && it < maxIt) continue;
        readScalars(st, W);
        if (W.hks->state != kRun) break;
        const double relNow = W.hks->rn / W.hks->bn;
        if (!std::isfinite(relNow)) break;       // a non-finite residual (non-finite matrix values): stop here
        nextCheck = it + (relNow > 30.0 * tol ? kCheckEvery : 1);
    }
    const KScalars& k = *W.hks;
    its = k.it;
    if (!std::isfinite(k.rn)) { rel = k.rn; S.lastStop = "non-finite residual"; return RSH_ERR_NONFINITE; }
    if (k.state == kConvS || k.state == kDoneS) { rel = k.sn / k.bn; S.lastStop = "converged (s)"; return RSH_OK; }
    rel = k.bn > 0 ? k.rn / k.bn : 0;
    if (k.state == kConv) { S.lastStop = "converged"; return RSH_OK; }
    if (k.state == kBreak) { S.lastStop = "breakdown"; return RSH_ERR_BREAKDOWN; }
    S.lastStop = k.state == kOmegaZero ? "omega = 0" : "max iterations";
    return RSH_ERR_NOT_CONVERGED;
}

int bicgstab(rsh_system& S, const double* b, double* x, double tol, int maxIt, int& its, double& rel, bool precF32 = false,
             bool opF32 = false) {
    return bicgstabIn(S, residentWork(S), b, x, tol, maxIt, its, rel, precF32, opF32);
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
    if (m > kMaxRestart) throw Error(RSH_ERR_UNSUPPORTED, "FGMRES restart above " + std::to_string(kMaxRestart));
    if (S.restartAlloc < m) {
        S.V.alloc(size_t(m + 1) * size_t(N));
        S.Z.alloc(size_t(m) * size_t(N));
        S.restartAlloc = m;
    }
    S.ensureVec(S.r);
    S.ensureVec(S.w);
    // projections of one Arnoldi step: first CGS pass [0, m], second pass [m+1, 2m+1], ||w||^2 at 2m+2
    const int hOff2 = m + 1, hNorm = 2 * m + 2;
    S.hcol.alloc(size_t(2 * m + 3));
    S.yv.alloc(size_t(m));
    if (!warm) launchFill(st, x, 0.0, N);
    launchDots(st, c, N, 1, b, b, nullptr, nullptr, c.scal.p, kEpNone, nullptr);
    RSH_HIP(hipMemcpyAsync(c.hscal, c.scal.p, sizeof(double), hipMemcpyDeviceToHost, st));
    RSH_HIP(hipStreamSynchronize(st));
    const double bn = std::sqrt(c.hscal[0]);
    its = 0;
    rel = 0;
    if (bn == 0) { launchFill(st, x, 0.0, N); return RSH_OK; }
    if (!std::isfinite(bn)) { rel = bn; S.lastStop = "non-finite right-hand side"; return RSH_ERR_NONFINITE; }
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
            launchMultiDot(st, c, N, j + 1, S.V.p, N, S.w.p, S.hcol.p + hOff2);
            launchMultiAxpy(st, N, j + 1, S.V.p, N, S.hcol.p + hOff2, S.w.p);
            launchDots(st, c, N, 1, S.w.p, S.w.p, nullptr, nullptr, S.hcol.p + hNorm, kEpNone, nullptr);
            std::vector<double> hh(size_t(2 * m + 3));
            RSH_HIP(hipMemcpyAsync(hh.data(), S.hcol.p, hh.size() * sizeof(double), hipMemcpyDeviceToHost, st));
            RSH_HIP(hipStreamSynchronize(st));
            for (int i = 0; i <= j; ++i) h(i, j) = hh[size_t(i)] + hh[size_t(hOff2 + i)];
            const double hn = std::sqrt(std::max(hh[size_t(hNorm)], 0.0));
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
    S.lastStop = "max iterations";
    return RSH_ERR_NOT_CONVERGED;
}

// FP64 iterative refinement x += d, r = b - A x (FP64 SpMV, as refineWithDevice bsr.cpp:1373-1396) around inner
// solves (BiCGSTAB, as the reference's float device solve bsr.cpp:1130-1215, or FGMRES) with the FP32
// preconditioner and, when opF32, the FP32 matrix; their dot products and Krylov vectors stay FP64.  A correction that does not
// reduce the FP64 residual (a diverged inner solve) is taken back and the loop stops.
int refine(rsh_system& S, const double* b, double* x, double tol, int maxIt, int& its, double& rel, int& outer, bool opF32) {
    Ctx& c = S.ctx->c;
    hipStream_t st = c.stream;
    const long long N = S.N;
    S.ensureVec(S.irR);  // residual
    S.ensureVec(S.irD);  // correction
    S.ensureVec(S.irX);  // the iterate before the correction (a non-finite correction cannot be subtracted again)
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
    if (!std::isfinite(bn)) { rel = bn; S.lastStop = "non-finite right-hand side"; return RSH_ERR_NONFINITE; }
    const double innerTol = S.opt.inner_tol > 0 ? S.opt.inner_tol : 1e-4;
    const int maxOuter = S.opt.max_outer > 0 ? S.opt.max_outer : 10;
    double rn = bn;
    // a system whose FP32 preconditioner stalled before starts with the FP64 one (unless the caller asks to retry)
    bool precF32 = !(S.fp32Stalled && !S.opt.retry_fp32);
    if (!precF32) opF32 = false;
    int firstInner = 0;   // iterations of the first inner solve that reduced the residual with the FP32 preconditioner
    for (outer = 0; outer < maxOuter; ++outer) {
        rel = rn / bn;
        if (rel <= tol) return RSH_OK;
        int inIts = 0;
        double inRel = 0;
        const double itol = std::max(innerTol, tol * bn / rn * 0.5);
        // an inner solve that stalls is restarted by the next refinement step: at most maxIt / maxOuter iterations,
        // and with the FP32 preconditioner at most twice (plus 20) what its first productive solve took: every inner
        // solve asks for the same relative reduction, so one running far longer has stalled (its correction is still
        // kept when it reduces the residual, else the FP64 preconditioner takes over)
        int left = std::max(1, std::min(maxIt - its, std::max(50, maxIt / maxOuter)));
        if ((precF32 || opF32) && firstInner > 0) left = std::min(left, 2 * firstInner + 20);
        const int rc = S.opt.krylov == RSH_KRYLOV_FGMRES
                           ? fgmres(S, S.irR.p, S.irD.p, itol, left, S.opt.restart > 0 ? S.opt.restart : 30, precF32, opF32, false, inIts, inRel)
                           : bicgstab(S, S.irR.p, S.irD.p, itol, left, inIts, inRel, precF32, opF32);
        its += inIts;
        S.lastInner.push_back(inIts);
        // a diverged inner solve (non-finite residual) is a step without progress: the check below discards it
        if (rc != RSH_OK && rc != RSH_ERR_NOT_CONVERGED && rc != RSH_ERR_BREAKDOWN && rc != RSH_ERR_NONFINITE) return rc;
        launchCopy(st, S.irX.p, x, N);
        launchAxpy(st, N, 1.0, S.irD.p, x);
        S.spmv(x, S.irR.p, false);
        launchSub(st, N, b, S.irR.p, S.irR.p);
        launchDots(st, c, N, 1, S.irR.p, S.irR.p, nullptr, nullptr, c.scal.p, kEpNone, nullptr);
        RSH_HIP(hipMemcpyAsync(c.hscal, c.scal.p, sizeof(double), hipMemcpyDeviceToHost, st));
        RSH_HIP(hipStreamSynchronize(st));
        const double rnew = std::sqrt(c.hscal[0]);
        if (!(rnew < rn)) {
            // no progress (or a non-finite correction): restore the iterate and its residual
            launchCopy(st, x, S.irX.p, N);
            S.spmv(x, S.irR.p, false);
            launchSub(st, N, b, S.irR.p, S.irR.p);
            if (precF32 || opF32) {
                // the FP32 preconditioner (or operator) stalls on this system: continue with the FP64 ones
                precF32 = false;
                opF32 = false;
                S.lastFallback = true;
                S.fp32Stalled = true;
                if (its >= maxIt) break;
                continue;
            }
            S.lastStop = "refinement stalled";
            break;
        }
        if ((precF32 || opF32) && firstInner == 0) firstInner = inIts;
        rn = rnew;
        if (its >= maxIt) break;
    }
    rel = rn / bn;
    return rel <= tol ? RSH_OK : RSH_ERR_NOT_CONVERGED;
}

// W: the BiCGSTAB workspace of a solve that must not allocate (null: the system's own buffers)
int solveDevice(rsh_system& S, const double* b, double* x, double tol, int maxIt, int* iterations, double* relRes, const BicgWork* W = nullptr) {
    if (!S.ready) return fail(RSH_ERR_ARG, "rsh_solve: call rsh_system_setup first");
    const double t0 = nowMs();
    int its = 0, outer = 0, rc;
    double rel = 0;
    S.lastFallback = false;
    S.lastInner.clear();
    S.lastStop = "non-finite residual";
    // RSH_MIXED: FGMRES with the FP32 preconditioner (its restarts recompute the FP64 true residual), or BiCGSTAB
    // with the FP32 preconditioner restarted on the FP64 true residual (refinement with the FP64 operator), which
    // keeps it from stalling on the preconditioner's FP32 rounding; RSH_MIXED_IR: refinement around FP32-operator
    // inner solves
    if (S.opt.precision == RSH_MIXED_IR) rc = refine(S, b, x, tol, maxIt, its, rel, outer, true);
    else if (S.opt.precision == RSH_MIXED && S.opt.krylov == RSH_KRYLOV_BICGSTAB) rc = refine(S, b, x, tol, maxIt, its, rel, outer, false);
    else if (S.opt.krylov == RSH_KRYLOV_FGMRES)
        rc = fgmres(S, b, x, tol, maxIt, S.opt.restart > 0 ? S.opt.restart : 30, S.opt.precision == RSH_MIXED, false, false, its, rel);
    else rc = W ? bicgstabIn(S, *W, b, x, tol, maxIt, its, rel, false, false) : bicgstab(S, b, x, tol, maxIt, its, rel, false, false);
    RSH_HIP(hipStreamSynchronize(S.s()));
    S.lastSolveMs = nowMs() - t0;
    S.lastIts = its;
    S.lastOuter = outer;
    S.lastRel = rel;
    if (iterations) *iterations = its;
    if (relRes) *relRes = rel;
    if (!std::isfinite(rel) || rc == RSH_ERR_NONFINITE) return fail(RSH_ERR_NONFINITE, "rsh_solve: " + S.lastStop);
    if (rc == RSH_ERR_NOT_CONVERGED) return fail(rc, "not converged: " + S.lastStop);
    if (rc == RSH_ERR_BREAKDOWN) return fail(rc, "Krylov breakdown");
    return rc;
}

std::once_flag g_devOnce;
int g_devCount = -1;

// ------------------------------------------------------------------ scratch of rsh_solve_device_noalloc
// A file-scope device array (each GPU the library runs on holds its own copy until the process ends): the BiCGSTAB
// workspace of solves that must not allocate.  Host code never passes the array's name to a kernel or a copy (on the
// host it is only a stand-in for the symbol): the per-device address comes from hipGetSymbolAddress, taken once per
// device with that device current.
constexpr size_t kNoAllocBytes = size_t(64) << 20;   // 64 MiB
constexpr size_t kNoAllocHead = size_t(64) << 10;    // Krylov scalars, dot output and partial sums first
constexpr int kNoAllocDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];
std::mutex g_scratchMu;
void* g_scratchPtr[kNoAllocDevices] = {};

void* scratchOn(int device) {
    if (device < 0 || device >= kNoAllocDevices) throw Error(RSH_ERR_ARG, "device index out of range for the no-alloc scratch");
    std::lock_guard<std::mutex> g(g_scratchMu);
    if (!g_scratchPtr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) throw Error(RSH_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratchPtr[device] = p;
    }
    return g_scratchPtr[device];
}

// vectors of N doubles rounded up to 256 bytes; the largest N whose eight vectors fit behind the head
size_t noAllocVecBytes(long long N) { return (size_t(N) * sizeof(double) + 255) / 256 * 256; }

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
    o->inner_tol = 1e-4;
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
        RSH_HIP(hipEventCreate(&ctx->c.ev0));
        RSH_HIP(hipEventCreate(&ctx->c.ev1));
        RSH_HIP(hipStreamSynchronize(ctx->c.stream));
        *out = ctx.release();
        return RSH_OK;
    });
}

int rsh_context_device_seconds(rsh_context* ctx, double* seconds) {
    if (!ctx || !seconds) return fail(RSH_ERR_ARG, "null argument");
    *seconds = ctx->c.deviceMs * 1e-3;
    return RSH_OK;
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
        if (ctx->c.ev0) (void)hipEventDestroy(ctx->c.ev0);
        if (ctx->c.ev1) (void)hipEventDestroy(ctx->c.ev1);
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
        // SpMV slices over all rows in natural order
        rowSliceLayout(n, S->pat.rp, S->pat.ci, S->spmvHost);
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
    S.valuesChanged();
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

// true-IMPES weights (BlockCPR::setup accBlocks, bsr.cpp:993-1030): n accumulation blocks (nb x nb, row-major as
// the Jacobian's), kept until replaced or cleared (null); used by the next CPR set-ups
static int setAccumulation(rsh_system* sys, const double* acc, hipMemcpyKind kind) {
    return guarded([&] {
        if (!sys) return fail(RSH_ERR_ARG, "null system");
        RSH_HIP(hipSetDevice(sys->ctx->c.device));
        if (!acc) {
            sys->acc.release();
            return RSH_OK;
        }
        sys->acc.alloc(size_t(sys->n) * sys->nb * sys->nb);
        RSH_HIP(hipMemcpyAsync(sys->acc.p, acc, sys->acc.bytes(), kind, sys->s()));
        RSH_HIP(hipStreamSynchronize(sys->s()));
        return RSH_OK;
    });
}
int rsh_system_set_accumulation(rsh_system* sys, const double* acc) { return setAccumulation(sys, acc, hipMemcpyHostToDevice); }
int rsh_system_set_accumulation_device(rsh_system* sys, const double* acc_dev) {
    return setAccumulation(sys, acc_dev, hipMemcpyDeviceToDevice);
}

int rsh_system_setup(rsh_system* sys, const rsh_options* opt) {
    return guarded([&] {
        if (!sys) return fail(RSH_ERR_ARG, "null system");
        if (!sys->valuesSet) return fail(RSH_ERR_ARG, "set values before setup");
        rsh_options o;
        rsh_default_options(&o);
        if (opt) o = *opt;
        if (o.precond < RSH_PREC_BILU0 || o.precond > RSH_PREC_CPR_BJ) return fail(RSH_ERR_ARG, "unknown preconditioner");
        if (o.precision != RSH_FP64 && o.precision != RSH_MIXED && o.precision != RSH_MIXED_IR && o.precision != RSH_MIXED_STORE)
            return fail(RSH_ERR_ARG, "unknown precision");
        if (o.restart > kMaxRestart) return fail(RSH_ERR_UNSUPPORTED, "FGMRES restart above " + std::to_string(kMaxRestart) + " is not built");
        RSH_HIP(hipSetDevice(sys->ctx->c.device));
        spanBegin(sys->ctx->c);
        rsh_system& S = *sys;
        hipStream_t st = S.s();
        const double t0 = nowMs();
        const bool f32 = o.precision != RSH_FP64;             // FP32 copies of the preconditioner's factors and inverses
        const bool a32 = f32;                                 // ... and of the CPR pressure stage
        if (o.precision == RSH_MIXED_IR && !S.have32) {       // FP32 operator of the inner solves
            uploadSlices(S.spmvHost, S.spmv32, S.nb, st);
            packSlices(st, S.nb, S.val.p, S.spmv32, S.spmvHost.nslots, true);
            S.have32 = true;
        }
        const bool wantMc = o.precond == RSH_PREC_MCBILU0 || o.precond == RSH_PREC_CPR_MC;
        const bool wantIlu = o.precond != RSH_PREC_CPR_BJ;
        if (o.ilu_schedule < 0 || o.ilu_schedule > 2) return fail(RSH_ERR_ARG, "unknown ILU schedule");
        if (o.ilu_rb_tiles < 0 || o.ilu_rb_tiles >= (1 << 24)) return fail(RSH_ERR_ARG, "ilu_rb_tiles: extents bx | by << 8 | bz << 16, each below 256");
        const int rbDims = wantMc ? o.ilu_rb_tiles : 0;
        if (wantIlu && (!S.ilu || S.ilu->mc != wantMc || (!wantMc && (S.ilu->schedule != o.ilu_schedule || S.ilu->tileDims != o.ilu_tile_dims)) ||
                        (wantMc && S.ilu->rbDims != rbDims))) {
            S.ilu = std::make_unique<IluDev>();
            S.ilu->schedule = o.ilu_schedule;
            S.ilu->tileDims = o.ilu_tile_dims;
            S.ilu->rbDims = rbDims;
            S.ilu->keepEntries = S.keepIluEntries;
            S.ilu->fuseMax = S.iluFuseMax;
            S.ilu->rowSys = S.rowSys.empty() ? nullptr : &S.rowSys;
            S.ilu->gate = S.gate;
            S.ilu->build(S.pat, S.rp.p, S.ci.p, S.diag.p, wantMc, st);
        }
        if (o.precond == RSH_PREC_CPR || o.precond == RSH_PREC_CPR_MC || o.precond == RSH_PREC_CPR_BJ) {
            if (!S.cpr) S.cpr = std::make_unique<CprDev>();
            S.cpr->smoother = o.smoother;
            S.cpr->steps = o.smooth_steps > 0 ? o.smooth_steps : 1;
            S.cpr->coarseSteps = o.coarse_smooth_steps > 0 ? o.coarse_smooth_steps : 0;
            S.cpr->omega = o.jacobi_omega > 0 ? o.jacobi_omega : CprDev::kOmegaDefault;
            S.cpr->plan = &S.spmv64;
            if (S.colP.n == 0) {
                S.colP.alloc(size_t(S.spmvHost.nslots) * S.nb * kWave);
            }
            launchPackColumn0(st, S.nb, S.val.p, S.spmv64.src.p, S.spmvHost.nslots, S.colP.p);
            if (a32) {
                S.colP32.alloc(S.colP.n);
                launchToFloat(st, (long long)S.colP.n, S.colP.p, S.colP32.p);
            }
            S.cpr->setup(S.pat, S.rp.p, S.ci.p, S.diag.p, S.val.p, o.reuse_amg != 0, st, a32, S.acc.n ? S.acc.p : nullptr);
            RSH_HIP(hipMemcpyAsync(&S.impesFallbacks, S.cpr->fallbackDev.p, sizeof(int), hipMemcpyDeviceToHost, st));
        }
        if (!wantIlu) {
            // block-Jacobi second stage: D_b^-1 of the natural blocks (FP32 copy for the mixed precision)
            S.bjInv.alloc(size_t(S.n) * S.nb * S.nb);
            launchBlockDiagInv(st, S.nb, S.n, S.diag.p, S.val.p, S.bjInv.p);
            if (f32) {
                S.bjInv32.alloc(S.bjInv.n);
                launchToFloat(st, (long long)S.bjInv.n, S.bjInv.p, S.bjInv32.p);
            }
        }
        RSH_HIP(hipStreamSynchronize(st));
        const double t1 = nowMs();
        if (wantIlu) S.ilu->factor(S.val.p, f32, st, o.ilu_lean != 0);
        RSH_HIP(hipStreamSynchronize(st));
        S.lastIluMs = nowMs() - t1;
        S.lastCprMs = t1 - t0;
        if (std::memcmp(&S.opt, &o, sizeof o) != 0) S.clearGraphs();   // captured applies bake the options in
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
        spanBegin(sys->ctx->c);
        const int rc = solveDevice(*sys, b_dev, x_dev, tol, max_it, iterations, rel_res);
        spanEnd(sys->ctx->c);
        return rc;
    });
}

int rsh_solve(rsh_system* sys, const double* b, double* x, double tol, int max_it, int* iterations, double* rel_res) {
    return guarded([&] {
        if (!sys || !b || !x) return fail(RSH_ERR_ARG, "null argument");
        RSH_HIP(hipSetDevice(sys->ctx->c.device));
        rsh_system& S = *sys;
        S.ensureVec(S.vb);
        S.ensureVec(S.vx);
        spanBegin(S.ctx->c);
        RSH_HIP(hipMemcpyAsync(S.vb.p, b, size_t(S.N) * sizeof(double), hipMemcpyHostToDevice, S.s()));
        const int rc = solveDevice(S, S.vb.p, S.vx.p, tol, max_it, iterations, rel_res);
        RSH_HIP(hipMemcpyAsync(x, S.vx.p, size_t(S.N) * sizeof(double), hipMemcpyDeviceToHost, S.s()));
        spanEnd(S.ctx->c);
        if (rc == RSH_OK) {
            for (long long i = 0; i < S.N; ++i)
                if (!std::isfinite(x[i])) return fail(RSH_ERR_NONFINITE, "non-finite solution");
        }
        return rc;
    });
}

int rsh_solve_device_noalloc(rsh_system* sys, const double* b_dev, double* x_dev, double tol, int max_it, int* iterations, double* rel_res) {
    return guarded([&] {
        if (!sys || !b_dev || !x_dev) return fail(RSH_ERR_ARG, "null argument");
        rsh_system& S = *sys;
        const size_t vb = noAllocVecBytes(S.N);
        if (kNoAllocHead + 8 * vb > kNoAllocBytes)
            return fail(RSH_ERR_ARG, "rsh_solve_device_noalloc: " + std::to_string(S.N) + " unknowns exceed the scratch of " +
                                         std::to_string((kNoAllocBytes - kNoAllocHead) / 8 / sizeof(double)) + " (64 MiB for eight vectors)");
        if (!S.ready) return fail(RSH_ERR_ARG, "rsh_solve_device_noalloc: call rsh_system_setup first");
        if (S.opt.krylov != RSH_KRYLOV_BICGSTAB || (S.opt.precision != RSH_FP64 && S.opt.precision != RSH_MIXED_STORE))
            return fail(RSH_ERR_UNSUPPORTED, "rsh_solve_device_noalloc: BiCGSTAB with RSH_FP64 or RSH_MIXED_STORE only");
        RSH_HIP(hipSetDevice(S.ctx->c.device));
        unsigned char* base = static_cast<unsigned char*>(scratchOn(S.ctx->c.device));
        double* vec[8];
        for (int k = 0; k < 8; ++k) vec[k] = reinterpret_cast<double*>(base + kNoAllocHead + size_t(k) * vb);
        const BicgWork W{vec[0], vec[1], vec[2], vec[3], vec[4], vec[5], vec[6], vec[7], reinterpret_cast<KScalars*>(base), S.ctx->c.hks,
                         reinterpret_cast<double*>(base + 1024), reinterpret_cast<double*>(base + 512), false};
        spanBegin(S.ctx->c);
        const int rc = solveDevice(S, b_dev, x_dev, tol, max_it, iterations, rel_res, &W);
        spanEnd(S.ctx->c);
        return rc;
    });
}

int rsh_system_set_spmv_lanes(rsh_system* sys, int lanes) {
    return guarded([&] {
        if (!sys) return fail(RSH_ERR_ARG, "null system");
        if (lanes != 1 && lanes != 2 && lanes != 4) return fail(RSH_ERR_ARG, "lanes per row must be 1, 2 or 4");
        RSH_HIP(hipSetDevice(sys->ctx->c.device));
        sys->setSpmvLanes(lanes);
        RSH_HIP(hipStreamSynchronize(sys->s()));
        return RSH_OK;
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
        S.precondG(S.vb.p, S.vx.p, S.precF32());
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
        const double* xp = C.vcycle(S.s(), S.amgF32(S.precF32()));
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

int rsh_ilu_order(rsh_system* sys, int* perm) {
    return guarded([&] {
        if (!sys || !perm) return fail(RSH_ERR_ARG, "null argument");
        if (!sys->ready || !sys->ilu) return fail(RSH_ERR_ARG, "call rsh_system_setup first");
        RSH_HIP(hipMemcpy(perm, sys->ilu->perm.p, sizeof(int) * size_t(sys->n), hipMemcpyDeviceToHost));
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
          << ",\"spmv_slots\":" << S.spmvHost.nslots << ",\"spmv_long_rows\":" << S.spmvHost.lrow.size()
          << ",\"spmv_col16\":" << (S.spmv64.dcol.n ? 1 : 0) << ",\"spmv_lanes\":" << S.spmvLanes() << ",\"cpr_true_impes\":" << (S.acc.n ? 1 : 0)
          << ",\"cpr_impes_fallbacks\":" << S.impesFallbacks << ",\"device_bytes\":" << S.deviceBytes();
        if (S.ilu) {
            const IluDev& I = *S.ilu;
            o << ",\"ilu_tiled\":" << (I.tiled ? 1 : 0) << ",\"ilu_tail_copy\":" << (I.tailCopy ? 1 : 0) << ",\"ilu_bwd_fused\":" << (I.fuseBwd ? 1 : 0) << ",\"ilu_rb_tiles\":" << (I.rbt ? 1 : 0);
            if (I.rbt)
                o << ",\"ilu_rb\":{\"tiles\":" << I.rbTiles << ",\"red_tiles\":" << I.rbRedTiles << ",\"units\":" << I.rb.unitStart.size() - 1
                  << ",\"red_levels\":" << I.rb.redLevels << ",\"black_levels\":" << I.rb.blackLevels << ",\"far_rows\":" << I.n - I.rb.nTile
                  << ",\"far_levels\":" << I.rb.farStart.size() - 1 << "}";
            if (I.tiled)
                o << ",\"ilu_tiles\":{\"dims\":[" << I.tiles.dims[0] << "," << I.tiles.dims[1] << "," << I.tiles.dims[2] << "],\"tile_levels\":" << I.tiles.tileLevels()
                  << ",\"units\":" << I.tiles.units() << ",\"max_rows\":" << I.tiles.maxRows << ",\"max_levels\":" << I.tiles.maxLevels
                  << ",\"long_max\":" << I.tLongMax << ",\"long_inside_max\":" << I.tLongCap << "}";
            o << ",\"ilu_levels\":" << (int(I.lstart.size()) - 1) << ",\"ilu_multicolour\":" << (S.ilu->mc ? 1 : 0) << ",\"ilu_regularised\":" << S.ilu->regularised
              << ",\"ilu_long_rows_L\":" << S.ilu->hL.lrow.size() << ",\"ilu_long_rows_U\":" << S.ilu->hU.lrow.size() << ",\"ilu_bytes\":" << S.ilu->bytes();
            int fusedL = 0, fusedU = 0, denseRunsL = 0, denseRunsU = 0;
            for (size_t i = 0; i < S.ilu->runsL.size(); ++i) { fusedL += S.ilu->runsL[i].fused; denseRunsL += S.ilu->denseL.runTask0[i] >= 0; }
            for (size_t i = 0; i < S.ilu->runsU.size(); ++i) { fusedU += S.ilu->runsU[i].fused; denseRunsU += S.ilu->denseU.runTask0[i] >= 0; }
            o << ",\"ilu_fused_runs\":[" << fusedL << "," << fusedU << "],\"ilu_dense_runs\":[" << denseRunsL << "," << denseRunsU
              << "],\"ilu_dense_groups\":[" << S.ilu->denseL.ngroups << "," << S.ilu->denseU.ngroups << "],\"ilu_factor_groups\":"
              << (S.ilu->fgStart.n ? S.ilu->fgStart.n - 1 : 0);
        }
        if (S.cpr) {
            o << ",\"amg_levels\":[";
            for (size_t l = 0; l < S.cpr->lv.size(); ++l) o << (l ? "," : "") << S.cpr->lv[l]->n;
            o << "],\"amg_nnz\":[";
            for (size_t l = 0; l < S.cpr->lv.size(); ++l) o << (l ? "," : "") << S.cpr->lv[l]->nnz;
            o << "],\"gs_levels\":[";
            for (size_t l = 0; l < S.cpr->lv.size(); ++l) o << (l ? "," : "") << (S.cpr->lv[l]->gsFs.size() - 1);
            o << "],\"amg_sliced\":[";
            for (size_t l = 0; l < S.cpr->lv.size(); ++l) o << (l ? "," : "") << (S.cpr->lv[l]->plan ? 1 : 0);
            o << "],\"amg_slices\":[";
            for (size_t l = 0; l < S.cpr->lv.size(); ++l) {
                const SliceDev<double>* sl = S.cpr->lv[l]->plan;
                o << (l ? "," : "") << "[" << (sl ? sl->nslices : 0) << "," << (sl ? sl->nlong : 0) << "," << (sl && sl->dcol.n ? 1 : 0) << "]";
            }
            o << "],\"coarse_dense\":" << (S.cpr->dense ? 1 : 0) << ",\"cpr_bytes\":" << S.cpr->bytes();
        }
        o << ",\"last_fallback_fp64\":" << (S.lastFallback ? 1 : 0) << ",\"fp32_stalled\":" << (S.fp32Stalled ? 1 : 0) << ",\"last_inner\":[";
        for (size_t q = 0; q < S.lastInner.size(); ++q) o << (q ? "," : "") << S.lastInner[q];
        o << "],\"last_iterations\":" << S.lastIts << ",\"last_outer\":" << S.lastOuter << ",\"last_rel_res\":" << S.lastRel
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
        if ((kind == 3 || kind == 5 || kind == 6) && !S.cpr) return fail(RSH_ERR_ARG, "no AMG");
        if ((kind == 5 || kind == 6) && (!S.ready || S.cpr->smoother != RSH_SMOOTH_JACOBI || S.cpr->lv.size() < size_t(kind - 3)))
            return fail(RSH_ERR_ARG, "no Jacobi-smoothed AMG level " + std::to_string(kind - 5));
        RSH_HIP(hipSetDevice(S.ctx->c.device));
        hipStream_t st = S.s();
        S.ensureVec(S.vb);
        S.ensureVec(S.vx);
        launchFill(st, S.vb.p, 1.0, S.N);
        const bool f32 = S.ready && S.precF32();                          // preconditioner precision
        const bool op32 = S.ready && S.opt.precision == RSH_MIXED_IR;     // operator precision
        const double nb2 = double(S.nb) * S.nb;
        const int ifp = S.iluFp(f32);
        const double pa = ifp != kIluFp64 ? 4.0 : 8.0;       // factor bytes per value
        const double pv = ifp == kIluFp32 ? 4.0 : 8.0;       // working-vector bytes per entry
        double bytes = 0;
        auto once = [&]() {
            switch (kind) {
                case 0: S.spmv(S.vb.p, S.vx.p, op32); break;
                case 1: S.precondG(S.vb.p, S.vx.p, f32); break;
                case 2:
                    if (!S.ilu || S.opt.precond == RSH_PREC_CPR_BJ) throw Error(RSH_ERR_ARG, "no ILU preconditioner set up");
                    S.ilu->apply(S.vb.p, S.vx.p, nullptr, ifp, st, false, S.iluInv64(f32));
                    break;
                case 3: S.cpr->vcycle(st, S.amgF32(f32)); break;
                case 4: launchCopyKernel(st, S.vb.p, S.vx.p, S.N); break;
                case 5:
                case 6: bytes = S.cpr->smoothStep(size_t(kind - 5), st, S.amgF32(f32)); break;
                default: throw Error(RSH_ERR_ARG, "unknown kernel kind");
            }
        };
        // algorithmic bytes: values + indices + vectors
        if (kind == 0) bytes = S.nnzb * (nb2 * (op32 ? 4.0 : 8.0) + 4.0) + (S.n + 1) * 4.0 + 2.0 * S.N * 8.0;
        else if (kind == 2) bytes = S.nnzb * (nb2 * pa + 4.0) + 2.0 * (S.n + 1) * 4.0 + 4.0 * S.N * pv;
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
