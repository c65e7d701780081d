// TEST INFRASTRUCTURE: the reference engine's own Newton loop on its demo waterflood (makeDemoCase, case.cpp:427-480),
// run twice - once with its CPU linear solver, once with the MI355X solver registered in the engine's existing hook
// (setDistributedSolver, bsr.h:154-157, called at simulator.cpp:3358-3367) through include/ressim_hip_solver.hpp.
// Built by oracle/Makefile from the reference sources into oracle/_ref/ (where the reference is mounted); the GPU
// test runs the binary and compares per-report-step Newton/linear counts, FOPT and FPR (the pattern of the reference's
// own plug-in tests, run_tests.cpp:3755-3773, c29_tests.inc:280-296).
//
// usage: newton_harness [steps] [precond: cpr|cpr_mc|bilu0] [smoother: gs|jacobi] [form: lean|store32|hookA|cpu-true-impes]
// hookA registers the solver in the engine's mixed hook instead (setGpuBlockSolver with Tuning gpu = 1: the engine's
// FP64 refinement around the device solve, refineWithDevice, simulator.cpp:3368-3378), with FP32 preconditioner
// arithmetic (RSH_MIXED).
// With RESSIM_DEBUG set the engine prints one "it N cnv X mb Y" line per Newton iteration to stderr
// (simulator.cpp:3207-3214); the harness marks where each run starts ("== run cpu" / "== run device").
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ressim/case.h"
#include "ressim/simulator.h"
#include "ressim_hip_solver.hpp"

using namespace ressim;

namespace {
struct Run {
    std::vector<SummaryPoint> summary;
    RunStats stats;
};

Run run(const Case& c) {
    Simulator sim;
    sim.log = [](const std::string&) {};
    sim.setup(c);
    sim.initialize();
    while (sim.runReportStep()) {}
    return Run{sim.summary(), sim.stats()};
}

void dump(const char* name, const Run& r) {
    std::printf("{\"run\":\"%s\",\"newton\":%d,\"linear\":%d,\"dd_solves\":%d,\"dd_fallbacks\":%d,\"gpu_solves\":%d,\"gpu_fallbacks\":%d,"
                "\"timesteps\":%d,\"cuts\":%d,\"steps\":[", name, r.stats.newton, r.stats.linear, r.stats.ddSolves, r.stats.ddFallbacks,
                r.stats.gpuSolves, r.stats.gpuFallbacks, r.stats.timesteps, r.stats.cuts);
    for (size_t i = 0; i < r.summary.size(); ++i) {
        const SummaryPoint& p = r.summary[i];
        std::printf("%s{\"t\":%.10g,\"newton\":%d,\"linear\":%d,\"fopt\":%.12g,\"fpr\":%.12g,\"fwpt\":%.12g}", i ? "," : "", p.time, p.newton,
                    p.linear, p.fopt, p.fpr, p.fwpt);
    }
    std::printf("]}\n");
    std::fflush(stdout);
}
}  // namespace

int main(int argc, char** argv) {
    const int steps = argc > 1 ? std::atoi(argv[1]) : 6;
    const std::string pc = argc > 2 ? argv[2] : "cpr";
    const std::string sm = argc > 3 ? argv[3] : "gs";
    // 4th argument: "cpu-true-impes": the CPU run only, with true-IMPES weights (the engine then also writes the
    // accumulation blocks next to a RESSIM_DUMP file, simulator.cpp:3313-3315); "lean": the device run with the lean
    // BILU0 factors; "store32": the device run with the FP32-stored preconditioner (FP64 arithmetic)
    const std::string form = argc > 4 ? argv[4] : "";
    const bool cpuTrueImpes = form == "cpu-true-impes";
    Case c = makeDemoCase();
    if (steps > 0 && int(c.schedule.size()) > steps) c.schedule.resize(size_t(steps));
    c.tuning.blockSolver = true;
    c.tuning.linearSolver = pc == "bilu0" ? 0 : 1;   // CPU path: BILU0 or block CPR
    c.tuning.cprWeights = cpuTrueImpes ? 1 : 0;       // else quasi-IMPES on both paths (the hook carries no accumulation blocks)
    std::fprintf(stderr, "== run cpu\n");
    std::fflush(stderr);
    const Run cpu = run(c);
    dump("cpu", cpu);
    if (cpuTrueImpes) return 0;

    rsh_options opt;
    rsh_default_options(&opt);
    opt.precond = pc == "bilu0" ? RSH_PREC_BILU0 : pc == "cpr_mc" ? RSH_PREC_CPR_MC : RSH_PREC_CPR;
    opt.smoother = sm == "jacobi" ? RSH_SMOOTH_JACOBI : RSH_SMOOTH_GS;
    opt.reuse_amg = 0;                                 // a fresh hierarchy per solve, as the CPU path builds it
    if (form == "lean") opt.ilu_lean = 1;
    if (form == "store32") opt.precision = RSH_MIXED_STORE;
    const bool hookA = form == "hookA";
    if (hookA) opt.precision = RSH_MIXED;
    ressim_hip::HipBlockSolver gpu(&opt, 0);
    if (hookA) {
        c.tuning.gpu = 1;
        setGpuBlockSolver(&gpu);
    } else {
        setDistributedSolver(&gpu);
    }
    std::fprintf(stderr, "== run device\n");
    std::fflush(stderr);
    const Run dev = run(c);
    setDistributedSolver(nullptr);
    setGpuBlockSolver(nullptr);
    dump("device", dev);
    std::printf("{\"device_seconds\":%.6f,\"last_error\":\"%s\"}\n", gpu.deviceSeconds(), gpu.lastError());
    return 0;
}
