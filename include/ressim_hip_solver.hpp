// ressim_hip_solver.hpp - ressim::GpuBlockSolver (engine/include/ressim/bsr.h:143-151) implemented on the MI355X
// C-ABI (ressim_hip.h).  Header-only; compile it inside the engine (it needs "ressim/bsr.h") and link
// libressim_hip.so.  Register an instance with
//     ressim::setDistributedSolver(&solver);   // FP64 drop-in: every Schur-reduced system (simulator.cpp:3358-3367)
//     ressim::setGpuBlockSolver(&solver);      // device solve inside the engine's FP64 refinement (:3368-3378)
// The instance keeps the system resident between calls: while the sparsity pattern is unchanged (every Newton
// iteration of a step, usually the whole run) only the new values are uploaded and the preconditioner is
// refreshed numerically (the AMG keeps its aggregates, as AMG::refresh, amg.cpp:257-271).
//
// Threading: the engine's hook is one global pointer that concurrent Simulators may call at the same time
// (study.cpp:345-361).  Each calling thread gets its own HIP context (stream) and resident system, so calls from
// different threads run concurrently on the device; calls from one thread are sequential.
#pragma once

#include <chrono>
#include <cmath>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "ressim/bsr.h"
#include "ressim_hip.h"

namespace ressim_hip {

class HipBlockSolver : public ressim::GpuBlockSolver {
public:
    explicit HipBlockSolver(const rsh_options* opt = nullptr, int device = 0) : device_(device) {
        rsh_default_options(&opt_);
        if (opt) opt_ = *opt;
    }
    ~HipBlockSolver() override {
        std::lock_guard<std::mutex> g(mu_);
        for (auto& kv : per_) kv.second.release();
    }
    const char* name() const override { return "MI355X HIP block solver (rsh)"; }

    // approximately solve A x = b to relative residual tol; false on device failure, non-convergence or a
    // non-finite result (the engine then re-solves on its CPU path, simulator.cpp:3366, 3376-3377)
    bool solve(const ressim::BSR& A, const std::vector<double>& b, std::vector<double>& x, double tol, int maxIt,
               int& iterations) override {
        iterations = 0;
        x.assign(size_t(A.size()), 0.0);
        State& st = state();
        const auto t0 = std::chrono::steady_clock::now();
        if (!st.ctx && rsh_context_create(device_, &st.ctx) != RSH_OK) return false;
        if (!st.sys || st.n != A.n || st.nb != A.nb || st.rowPtr != A.rowPtr || st.col != A.col) {
            if (st.sys) rsh_system_destroy(st.sys);
            st.sys = nullptr;
            if (rsh_system_create(st.ctx, A.n, A.nb, A.rowPtr.data(), A.col.data(), &st.sys) != RSH_OK) return false;
            st.n = A.n;
            st.nb = A.nb;
            st.rowPtr = A.rowPtr;
            st.col = A.col;
        }
        if (rsh_system_set_values(st.sys, A.val.data()) != RSH_OK) return false;
        if (rsh_system_setup(st.sys, &opt_) != RSH_OK) return false;
        double rel = 0;
        const int rc = rsh_solve(st.sys, b.data(), x.data(), tol, maxIt, &iterations, &rel);
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        {
            std::lock_guard<std::mutex> g(mu_);
            busy_ += sec;
        }
        if (rc != RSH_OK) return false;
        for (double v : x)
            if (!std::isfinite(v)) return false;
        return true;
    }
    // wall seconds spent in solve() on the device side (set-up + solve, all threads)
    double deviceSeconds() const override {
        std::lock_guard<std::mutex> g(mu_);
        return busy_;
    }
    const char* lastError() const { return rsh_last_error(); }

private:
    struct State {
        rsh_context* ctx = nullptr;
        rsh_system* sys = nullptr;
        int n = 0, nb = 0;
        std::vector<int> rowPtr, col;
        void release() {
            if (sys) rsh_system_destroy(sys);
            if (ctx) rsh_context_destroy(ctx);
            sys = nullptr;
            ctx = nullptr;
        }
    };
    State& state() {
        std::lock_guard<std::mutex> g(mu_);
        return per_[std::this_thread::get_id()];
    }
    int device_;
    rsh_options opt_{};
    mutable std::mutex mu_;
    std::map<std::thread::id, State> per_;
    double busy_ = 0;
};

}  // namespace ressim_hip
