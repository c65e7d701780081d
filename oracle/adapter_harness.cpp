// TEST INFRASTRUCTURE: include/ressim_hip_solver.hpp (the ressim::GpuBlockSolver adapter) driven the way the
// reference's ensemble callers drive the engine's global hook: many threads, fresh thread pools on every call
// (runBatch, study.cpp:345-361).  Built by oracle/Makefile against the reference's engine headers (ressim/bsr.h, where
// the reference is mounted) into oracle/_ref/adapter_harness; tests/test_gpu_adapter.py writes the systems and checks
// the output.
//
// usage: adapter_harness <systems file> <solutions out> <precond: cpr_mc|bilu0|mcbilu0> <threads> <rounds>
// systems file: int32 count, then per system int32 n, nb, row_ptr[n+1], col[nnzb], float64 val[nnzb nb nb], b[n nb].
// Prints one JSON line: bit-identity of every concurrent and churned result with the same system solved alone, the
// pool's slot / resident-system / creation counts, the same under a small device-memory budget, device seconds.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "ressim_hip_solver.hpp"

namespace {
struct Sys {
    ressim::BSR A;
    std::vector<double> b;
};

bool readSystems(const char* path, std::vector<Sys>& out) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    auto rd = [&](void* p, size_t bytes) { return std::fread(p, 1, bytes, f) == bytes; };
    int count = 0;
    bool ok = rd(&count, 4) && count > 0;
    for (int k = 0; ok && k < count; ++k) {
        Sys s;
        int hdr[2];
        ok = rd(hdr, 8);
        if (!ok) break;
        s.A.n = hdr[0];
        s.A.nb = hdr[1];
        s.A.rowPtr.resize(size_t(s.A.n) + 1);
        ok = rd(s.A.rowPtr.data(), 4 * s.A.rowPtr.size());
        if (!ok) break;
        const int nnzb = s.A.rowPtr[size_t(s.A.n)];
        s.A.col.resize(size_t(nnzb));
        s.A.val.resize(size_t(nnzb) * s.A.nb * s.A.nb);
        s.b.resize(size_t(s.A.n) * s.A.nb);
        ok = rd(s.A.col.data(), 4 * s.A.col.size()) && rd(s.A.val.data(), 8 * s.A.val.size()) && rd(s.b.data(), 8 * s.b.size());
        for (int i = 0; ok && i < s.A.n; ++i)
            for (int p = s.A.rowPtr[size_t(i)]; p < s.A.rowPtr[size_t(i) + 1]; ++p)
                if (s.A.col[size_t(p)] == i) s.A.diag.push_back(p);
        out.push_back(std::move(s));
    }
    std::fclose(f);
    return ok;
}

// rounds of fresh thread pools, each thread solving system (round + t) % K; mismatches against the solo solutions
int churn(ressim_hip::HipBlockSolver& sol, const std::vector<Sys>& sys, const std::vector<std::vector<double>>& solo, int threads, int rounds,
          int& failures, int& maxResident) {
    std::atomic<int> bad{0}, fail{0};
    maxResident = 0;
    for (int r = 0; r < rounds; ++r) {
        std::vector<std::thread> pool;
        for (int t = 0; t < threads; ++t)
            pool.emplace_back([&, r, t] {
                const size_t k = size_t(r + t) % sys.size();
                std::vector<double> x;
                int its = 0;
                if (!sol.solve(sys[k].A, sys[k].b, x, 1e-8, 2000, its)) { ++fail; return; }
                if (x != solo[k]) ++bad;
            });
        for (auto& th : pool) th.join();
        maxResident = std::max(maxResident, sol.residentSystems());
    }
    failures = fail.load();
    return bad.load();
}
}  // namespace

int main(int argc, char** argv) {
    if (argc < 6) {
        std::fprintf(stderr, "usage: adapter_harness <systems> <solutions out> <precond> <threads> <rounds>\n");
        return 2;
    }
    std::vector<Sys> sys;
    if (!readSystems(argv[1], sys)) {
        std::fprintf(stderr, "cannot read %s\n", argv[1]);
        return 2;
    }
    const std::string pc = argv[3];
    const int threads = std::atoi(argv[4]), rounds = std::atoi(argv[5]);
    rsh_options opt;
    rsh_default_options(&opt);
    opt.precond = pc == "bilu0" ? RSH_PREC_BILU0 : pc == "mcbilu0" ? RSH_PREC_MCBILU0 : RSH_PREC_CPR_MC;
    opt.smoother = RSH_SMOOTH_JACOBI;
    const auto t0 = std::chrono::steady_clock::now();
    ressim_hip::HipBlockSolver sol(&opt, 0, 8);
    // every system alone, one after another
    std::vector<std::vector<double>> solo(sys.size());
    std::vector<int> soloIts(sys.size());
    for (size_t k = 0; k < sys.size(); ++k)
        if (!sol.solve(sys[k].A, sys[k].b, solo[k], 1e-8, 2000, soloIts[k])) {
            std::fprintf(stderr, "solo solve %zu failed: %s\n", k, sol.lastError());
            return 1;
        }
    // the same calls from concurrent threads and from fresh pools (the runBatch pattern)
    int failures = 0, maxResident = 0;
    const int mismatches = churn(sol, sys, solo, threads, rounds, failures, maxResident);
    const double bytesPerSystem = sol.residentBytes() / std::max(1, sol.residentSystems());
    // the same under a budget of about one and a half systems: idle systems are released and rebuilt on demand
    ressim_hip::HipBlockSolver small(&opt, 0, 8, 1.5 * bytesPerSystem);
    int failuresB = 0, maxResidentB = 0;
    const int mismatchesB = churn(small, sys, solo, threads, rounds, failuresB, maxResidentB);
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    FILE* f = std::fopen(argv[2], "wb");
    if (!f) return 1;
    for (const auto& x : solo) std::fwrite(x.data(), 8, x.size(), f);
    std::fclose(f);
    std::printf("{\"systems\":%zu,\"threads\":%d,\"rounds\":%d,\"solo_iterations\":[", sys.size(), threads, rounds);
    for (size_t k = 0; k < soloIts.size(); ++k) std::printf("%s%d", k ? "," : "", soloIts[k]);
    std::printf("],\"mismatches\":%d,\"failures\":%d,\"slots\":%d,\"resident\":%d,\"max_resident\":%d,\"created\":%lld,"
                "\"budget_mismatches\":%d,\"budget_failures\":%d,\"budget_slots\":%d,\"budget_max_resident\":%d,\"budget_created\":%lld,"
                "\"budget_bytes\":%.0f,\"device_seconds\":%.6f,\"wall_seconds\":%.6f}\n",
                mismatches, failures, sol.slots(), sol.residentSystems(), maxResident, sol.systemsCreated(), mismatchesB, failuresB, small.slots(),
                maxResidentB, small.systemsCreated(), 1.5 * bytesPerSystem, sol.deviceSeconds(), wall);
    return 0;
}
