// ressim_hip_solver.hpp - ressim::GpuBlockSolver (engine/include/ressim/bsr.h:143-151) implemented on the MI355X
// C-ABI (ressim_hip.h).  Header-only; compile it inside the engine (it needs "ressim/bsr.h") and link
// libressim_hip.so.  Register an instance with
//     ressim::setDistributedSolver(&solver);   // FP64 drop-in: every Schur-reduced system (simulator.cpp:3358-3367)
//     ressim::setGpuBlockSolver(&solver);      // device solve inside the engine's FP64 refinement (:3368-3378)
// The instance keeps systems resident between calls: while a sparsity pattern is unchanged (every Newton iteration of
// a step, usually the whole run) only the new values are uploaded and the preconditioner is refreshed numerically
// (the AMG keeps its aggregates, as AMG::refresh, amg.cpp:257-271).
//
// Threading: the engine's hook is one global pointer that concurrent Simulators may call at the same time, from
// threads that come and go (runBatch starts fresh threads on every call, study.cpp:345-361).  The solver keeps a pool
// of resident slots (a HIP context, i.e. a stream, and a resident system each), shared by all threads: a call takes an
// idle slot holding its pattern, else a new slot, else the least recently used idle slot (whose system is released),
// and waits while every slot is busy.  The pool is bounded by a slot count and by a device-memory budget (idle
// systems are released, least recently used first, to stay within it); calls on different slots run concurrently on
// the device, and a call's result does not depend on what runs beside it.
#pragma once

#include <condition_variable>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "ressim/bsr.h"
#include "ressim_hip.h"

namespace ressim_hip {

class HipBlockSolver : public ressim::GpuBlockSolver {
public:
    // maxSlots: resident systems at most (and concurrent device calls); budgetBytes: device memory the resident systems
    // may hold together (a system larger than the budget is still solved, alone)
    explicit HipBlockSolver(const rsh_options* opt = nullptr, int device = 0, int maxSlots = 16, double budgetBytes = 64.0 * 1073741824.0)
        : device_(device), maxSlots_(maxSlots < 1 ? 1 : maxSlots), budget_(budgetBytes) {
        rsh_default_options(&opt_);
        if (opt) opt_ = *opt;
    }
    ~HipBlockSolver() override {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return busy_ == 0; });
        for (auto& s : slots_) drop(*s, true);
    }
    const char* name() const override { return "MI355X HIP block solver (rsh)"; }

    // approximately solve A x = b to relative residual tol; false on device failure, non-convergence or a
    // non-finite result (the engine then re-solves on its CPU path, simulator.cpp:3366, 3376-3377)
    bool solve(const ressim::BSR& A, const std::vector<double>& b, std::vector<double>& x, double tol, int maxIt,
               int& iterations) override {
        iterations = 0;
        x.assign(size_t(A.size()), 0.0);
        const uint64_t key = patternKey(A);
        Slot* s = acquire(A, key);
        if (!s) return false;
        bool ok = s->sys != nullptr || create(*s, A, key);
        double rel = 0;
        int rc = RSH_ERR_DEVICE;
        if (ok) ok = rsh_system_set_values(s->sys, A.val.data()) == RSH_OK && rsh_system_setup(s->sys, &opt_) == RSH_OK;
        if (ok) rc = rsh_solve(s->sys, b.data(), x.data(), tol, maxIt, &iterations, &rel);
        if (!ok) {
            std::lock_guard<std::mutex> g(mu_);
            lastError_ = rsh_last_error();
        }
        release(*s, ok);
        if (!ok || rc != RSH_OK) return false;
        for (double v : x)
            if (!std::isfinite(v)) return false;
        return true;
    }
    // device-busy seconds of every set-up and solve so far (HIP events on the slots' streams)
    double deviceSeconds() const override {
        std::lock_guard<std::mutex> g(mu_);
        double t = retiredSeconds_;
        for (const auto& s : slots_) {
            double d = 0;
            if (s->ctx && rsh_context_device_seconds(s->ctx, &d) == RSH_OK) t += d;
        }
        return t;
    }
    const char* lastError() const {
        std::lock_guard<std::mutex> g(mu_);
        return lastError_.c_str();
    }
    // pool state (tests, monitoring)
    int residentSystems() const {
        std::lock_guard<std::mutex> g(mu_);
        int k = 0;
        for (const auto& s : slots_) k += s->sys != nullptr;
        return k;
    }
    int slots() const {
        std::lock_guard<std::mutex> g(mu_);
        return int(slots_.size());
    }
    double residentBytes() const {
        std::lock_guard<std::mutex> g(mu_);
        return used_;
    }
    long long systemsCreated() const {
        std::lock_guard<std::mutex> g(mu_);
        return created_;
    }

private:
    struct Slot {
        rsh_context* ctx = nullptr;
        rsh_system* sys = nullptr;
        uint64_t key = 0;
        int n = 0, nb = 0;
        std::vector<int> rowPtr, col;
        double bytes = 0;
        bool busy = false;
        unsigned long long lastUse = 0;
    };
    // FNV-1a over the pattern (the comparison itself is exact: same key and equal arrays)
    static uint64_t patternKey(const ressim::BSR& A) {
        uint64_t h = 1469598103934665603ULL;
        auto mix = [&h](uint64_t v) { h = (h ^ v) * 1099511628211ULL; };
        mix(uint64_t(A.n));
        mix(uint64_t(A.nb));
        for (int v : A.rowPtr) mix(uint64_t(uint32_t(v)));
        for (int v : A.col) mix(uint64_t(uint32_t(v)));
        return h;
    }
    static bool holds(const Slot& s, const ressim::BSR& A, uint64_t key) {
        return s.sys && s.key == key && s.n == A.n && s.nb == A.nb && s.rowPtr == A.rowPtr && s.col == A.col;
    }
    static double estimate(const ressim::BSR& A) {
        const double vals = double(A.nnzb()) * A.nb * A.nb * 8.0, vec = double(A.size()) * 8.0;
        return 6.0 * vals + 20.0 * vec;   // values, slice layouts, ILU factors and sweeps, AMG, Krylov vectors
    }
    // the slot for this call (lock held inside); null only if a context cannot be created
    Slot* acquire(const ressim::BSR& A, uint64_t key) {
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            Slot* hit = nullptr;
            for (auto& s : slots_)
                if (!s->busy && holds(*s, A, key) && (!hit || s->lastUse > hit->lastUse)) hit = s.get();
            if (hit) return take(*hit);
            // a new pattern: make room in the budget from idle systems, least recently used first
            const double need = estimate(A);
            while (used_ + need > budget_) {
                Slot* v = lruSystem();
                if (!v) break;
                drop(*v, false);
            }
            if (Slot* e = idleEmpty()) return take(*e);
            if (int(slots_.size()) < maxSlots_) {
                slots_.push_back(std::make_unique<Slot>());
                Slot& s = *slots_.back();
                if (rsh_context_create(device_, &s.ctx) != RSH_OK) {
                    lastError_ = rsh_last_error();
                    slots_.pop_back();
                    return nullptr;
                }
                return take(s);
            }
            if (Slot* v = lruSystem()) {
                drop(*v, false);
                return take(*v);
            }
            cv_.wait(lk);
        }
    }
    Slot* idleEmpty() {
        for (auto& s : slots_)
            if (!s->busy && !s->sys) return s.get();
        return nullptr;
    }
    Slot* lruSystem() {
        Slot* v = nullptr;
        for (auto& s : slots_)
            if (!s->busy && s->sys && (!v || s->lastUse < v->lastUse)) v = s.get();
        return v;
    }
    Slot* take(Slot& s) {
        s.busy = true;
        ++busy_;
        return &s;
    }
    // release a slot's resident system (and with ctxToo its context); lock held
    void drop(Slot& s, bool ctxToo) {
        if (s.sys) {
            rsh_system_destroy(s.sys);
            s.sys = nullptr;
            used_ -= s.bytes;
            s.bytes = 0;
        }
        if (ctxToo && s.ctx) {
            double d = 0;
            if (rsh_context_device_seconds(s.ctx, &d) == RSH_OK) retiredSeconds_ += d;
            rsh_context_destroy(s.ctx);
            s.ctx = nullptr;
        }
    }
    // a resident system for this pattern on the slot (no lock: the slot is this call's)
    bool create(Slot& s, const ressim::BSR& A, uint64_t key) {
        if (rsh_system_create(s.ctx, A.n, A.nb, A.rowPtr.data(), A.col.data(), &s.sys) != RSH_OK) {
            s.sys = nullptr;
            return false;
        }
        s.key = key;
        s.n = A.n;
        s.nb = A.nb;
        s.rowPtr = A.rowPtr;
        s.col = A.col;
        std::lock_guard<std::mutex> g(mu_);
        ++created_;
        s.bytes = estimate(A);
        used_ += s.bytes;
        return true;
    }
    void release(Slot& s, bool ok) {
        // the measured footprint once the preconditioner exists (set-up allocates it)
        double bytes = -1;
        if (ok && s.sys) {
            char buf[8192];
            if (rsh_system_info(s.sys, buf, int(sizeof(buf))) == RSH_OK) {
                const char* p = std::strstr(buf, "\"device_bytes\":");
                if (p) bytes = std::atof(p + 15);
            }
        }
        {
            std::lock_guard<std::mutex> g(mu_);
            if (bytes > 0) {
                used_ += bytes - s.bytes;
                s.bytes = bytes;
            }
            if (!ok && s.sys) drop(s, false);   // a failed call leaves nothing half set up behind
            s.busy = false;
            s.lastUse = ++clock_;
            --busy_;
            // over the budget after a set-up grew: release other idle systems
            while (used_ > budget_) {
                Slot* lru = nullptr;
                for (auto& o : slots_)
                    if (!o->busy && o->sys && o.get() != &s && (!lru || o->lastUse < lru->lastUse)) lru = o.get();
                if (!lru) break;
                drop(*lru, false);
            }
        }
        cv_.notify_all();
    }

    int device_;
    int maxSlots_;
    double budget_;
    rsh_options opt_{};
    mutable std::mutex mu_;
    std::condition_variable cv_;
    std::vector<std::unique_ptr<Slot>> slots_;
    int busy_ = 0;
    double used_ = 0, retiredSeconds_ = 0;
    unsigned long long clock_ = 0;
    long long created_ = 0;
    std::string lastError_;
};

}  // namespace ressim_hip
