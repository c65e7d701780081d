/* ressim_hip.h - C-ABI of the MI355X (gfx950) linear-solve library for the black-oil Newton loop.
 *
 * This library replaces the device side of ressim::GpuBlockSolver (engine/include/ressim/bsr.h:143-151): the
 * engine's Newton loop (engine/src/simulator.cpp:3358-3378) hands every Schur-reduced BSR system to
 *     bool GpuBlockSolver::solve(const BSR& A, const std::vector<double>& b, std::vector<double>& x,
 *                                double tol, int maxIt, int& iterations)
 * registered through setDistributedSolver (FP64, bsr.h:154-157) or setGpuBlockSolver (float inner solve inside
 * FP64 refinement, bsr.h:152-153).  The C++ adapter in include/ressim_hip_solver.hpp implements that virtual
 * interface on top of the entry points below; INTEGRATION.md shows the registration.
 *
 * Conventions (mirroring the reference C API, engine/include/ressim/capi.h:5-7): opaque handles, int status
 * (>= 0 success, < 0 error), a thread-local message from rsh_last_error().  Plain pointers and sizes only.
 * Matrices are BSR: n block rows, nb x nb row-major blocks (row = equation, column = unknown, column 0 =
 * pressure), int32 row_ptr[n+1], col[nnzb] sorted per row with the diagonal present, double val[nnzb*nb*nb].
 * Vectors are n*nb doubles, cell-major (the layout of ressim::BSR::spmv, bsr.cpp:109-126).
 *
 * Threading: every context owns one HIP stream; different contexts may be used from different threads at the
 * same time (the reference hook is global and called from concurrent Simulators, study.cpp:345-361).  One
 * context must not be used by two threads at once.
 */
#ifndef RESSIM_HIP_H
#define RESSIM_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RSH_OK 0
#define RSH_ERR_ARG (-1)          /* invalid argument (sizes, pattern, options) */
#define RSH_ERR_DEVICE (-2)       /* HIP runtime failure or no device */
#define RSH_ERR_NOT_CONVERGED (-3)/* the Krylov method stopped before reaching tol (x holds the last iterate) */
#define RSH_ERR_BREAKDOWN (-4)    /* Krylov breakdown (rho or r^T v below 1e-300) */
#define RSH_ERR_NONFINITE (-5)    /* non-finite solution: the caller falls back to its CPU path */
#define RSH_ERR_UNSUPPORTED (-6)  /* block size or option not built */

/* preconditioners (bsr.h:50-250) */
#define RSH_PREC_BILU0 1          /* natural-order block ILU(0), level-scheduled sweeps (BILU0, bsr.cpp:347-596) */
#define RSH_PREC_MCBILU0 2        /* multicolour block ILU(0) (MCBILU0, bsr.cpp:599-730) */
#define RSH_PREC_CPR 3            /* BlockCPR: quasi-IMPES weights, AMG V-cycle on the pressure system, BILU0 */
#define RSH_PREC_CPR_MC 4         /* BlockCPR with the multicolour BILU0 second stage */
#define RSH_PREC_CPR_BJ 5         /* DeviceCprAmg (bsr.h:158-179): CPR whose second stage is block Jacobi,
                                     z = E x_p + D_b^-1 (r - A E x_p); pair with RSH_SMOOTH_JACOBI for the reference's
                                     device V-cycle */
/* Krylov methods (linsolve.h:99-106) */
#define RSH_KRYLOV_BICGSTAB 0     /* right-preconditioned BiCGSTAB from x0 = 0 (linsolve.cpp:159-216) */
#define RSH_KRYLOV_FGMRES 1       /* flexible restarted GMRES (linsolve.cpp:218-273) */
/* precision */
#define RSH_FP64 0                /* every kernel in double: the drop-in for setDistributedSolver */
#define RSH_MIXED 1               /* FP32 preconditioner (ILU factors and sweeps, AMG level matrices and smoother
                                     diagonals, CPR pressure column) inside FP64 FGMRES (whose restarts recompute the
                                     FP64 true residual) or FP64 BiCGSTAB restarted by iterative refinement on the
                                     FP64 true residual; FP64 matrix, vectors, dots and coarsest-level solve.  A
                                     refinement step that does not lower the FP64 residual switches the rest of the
                                     solve to the FP64 preconditioner (rsh_system_info: last_fallback_fp64) */
#define RSH_MIXED_IR 2            /* FP64 iterative refinement (refineWithDevice, bsr.cpp:1373-1396) around inner
                                     solves on the FP32 matrix and FP32 preconditioner (the float device solve,
                                     bsr.cpp:1130-1215); FP64 true residual, dots and Krylov vectors */
#define RSH_MIXED_STORE 3         /* FP32-stored preconditioner (BILU0 factors and dense well-group inverses, AMG level
                                     matrices and smoother diagonals, CPR pressure column, block-Jacobi inverses) with
                                     FP64 vectors and FP64 arithmetic inside an FP64 Krylov method: the preconditioner
                                     stays one fixed linear operator, so no refinement is needed */
/* AMG smoothers */
#define RSH_SMOOTH_GS 0           /* level-scheduled Gauss-Seidel: forward pre-, backward post-sweep, the exact
                                     arithmetic of AMG::cycle (amg.cpp:303-329) */
#define RSH_SMOOTH_JACOBI 1       /* damped Jacobi, omega 2/3, one pre and one post sweep (DeviceCprAmg,
                                     bsr.h:158-179): fully parallel, different iteration counts */

typedef struct rsh_context rsh_context;
typedef struct rsh_system rsh_system;
typedef struct rsh_batch rsh_batch;
typedef struct rsh_server rsh_server;
typedef struct rsh_wells rsh_wells;

typedef struct rsh_options {
    int precond;        /* RSH_PREC_* (default RSH_PREC_CPR) */
    int krylov;         /* RSH_KRYLOV_* (default BiCGSTAB; with RSH_MIXED_IR the method of the inner solves) */
    int precision;      /* RSH_FP64, RSH_MIXED, RSH_MIXED_IR or RSH_MIXED_STORE */
    int smoother;       /* RSH_SMOOTH_* */
    int restart;        /* FGMRES restart length (default 30; the engine's fallback GMRES uses 60,
                           simulator.cpp:3395); 1..256, larger values are refused with RSH_ERR_UNSUPPORTED */
    int max_outer;      /* refinement steps of RSH_MIXED (BiCGSTAB) and RSH_MIXED_IR (default 10) */
    double inner_tol;   /* their inner solves' relative tolerance (default 1e-4, refineWithDevice's, bsr.cpp:1383);
                           each inner solve runs at most max_it / max_outer (>= 50) iterations */
    int reuse_amg;      /* 1: keep the AMG aggregates of the last setup while the pattern is unchanged and only
                           recompute the Galerkin operators (AMG::refresh, amg.cpp:257-271); 0: fresh setup */
    int smooth_steps;   /* RSH_SMOOTH_JACOBI: pre- and post-smoothing steps per level (default 1, DeviceCprAmg's) */
    double jacobi_omega;/* RSH_SMOOTH_JACOBI: the damping weight (<= 0: 0.6667, DeviceCprAmg's, bsr.h:173) */
    int ilu_lean;       /* 1: the BILU0 / MC-BILU0 factors (and CPR's second stage) factorised lean, as the reference's
                           C45 low-memory mode (BILU0::setLean / BlockCPR::setLean, bsr.cpp:216-263, 358-389; the
                           engine's default from 500 000 cells, simulator.cpp:3318): the matrix rounded to float32,
                           every row eliminated in double against the already-rounded earlier rows and stored
                           rounded; the sweeps read float32 factors with FP64 vectors and arithmetic.  Default 0 */
    int coarse_smooth_steps; /* RSH_SMOOTH_JACOBI: steps on the levels below the finest (<= 0: smooth_steps) */
    int ilu_schedule;   /* natural-order BILU0 sweeps (RSH_PREC_BILU0, RSH_PREC_CPR): 0 tiled where a cost model finds it
                           faster than one launch per level, 1 one launch per level, 2 tiled whenever the pattern admits
                           it.  Tiled: rows grouped into compact tiles along the pattern's dominant lower offsets (the
                           grid axes of a natural-ordered grid), one launch per tile level, one workgroup per tile walking
                           its own levels; the same factors, each bulk row's sweep in the same slot order.  Default 0 */
    int ilu_tile_dims;  /* tile extents along those offsets, bx | by << 8 | bz << 16 (0: chosen by the cost model) */
    int retry_fp32;     /* RSH_MIXED (BiCGSTAB) and RSH_MIXED_IR: 0 (default) once a refinement on this system has stalled
                           with the FP32 preconditioner and fallen back to the FP64 one, the system's later solves start
                           with the FP64 preconditioner (a well-scaled system keeps FP32; one whose FP32 preconditioner
                           loses too much accuracy pays the stall once, not once per Newton iteration); 1: every solve
                           starts with FP32 again.  rsh_system_info reports the state as fp32_stalled */
    int ilu_rb_tiles;   /* RSH_PREC_MCBILU0 / RSH_PREC_CPR_MC: instead of colouring single cells, colour tiles of extents
                           bx | by << 8 | bz << 16 along the pattern's dominant lower offsets (the grid axes of a
                           natural-ordered grid) red / black by the parity of their tile coordinates; rows by colour, tile
                           and natural order inside the tile, the rows far longer than the bulk (a well's perforated cells)
                           behind the tiles.  The BILU0 of that order is factorised and swept tile by tile (a tile of one
                           colour couples only to tiles of the other): fewer iterations than the cell colouring, closer to
                           the natural order's.  Not the reference's factors (neither is MC-BILU0's order).  0 (default):
                           the cell colouring; a pattern without grid offsets, or whose tiles of one colour are coupled
                           beyond a tenth of its rows, keeps the cell colouring */
} rsh_options;

/* ---- library / context */
int rsh_version(void);
const char* rsh_last_error(void);
int rsh_device_count(void);
void rsh_default_options(rsh_options* opt);
int rsh_context_create(int device, rsh_context** out);
int rsh_context_destroy(rsh_context* ctx);
int rsh_context_synchronize(rsh_context* ctx);
/* device seconds of the set-ups and solves run on the context (GpuBlockSolver::deviceSeconds, bsr.h:149-150: device
 * busy time from GPU timestamps): HIP events on the context's stream, from the first rsh_system_setup or rsh_solve*
 * of a span to the end of the solve that closes it */
int rsh_context_device_seconds(rsh_context* ctx, double* seconds);

/* ---- a resident system: pattern, values and preconditioner stay on the device across Newton iterations */
int rsh_system_create(rsh_context* ctx, int n, int nb, const int* row_ptr, const int* col, rsh_system** out);
int rsh_system_destroy(rsh_system* sys);
/* new values on the same pattern (host or device pointer to nnzb*nb*nb doubles) */
int rsh_system_set_values(rsh_system* sys, const double* val);
int rsh_system_set_values_device(rsh_system* sys, const double* val_dev);
/* true-IMPES CPR weights (BlockCPR::setup's accBlocks, bsr.cpp:993-1030; the engine's cprWeights = 1,
 * simulator.cpp:3299-3316): n accumulation blocks, nb x nb row-major like the Jacobian's blocks (host or device
 * pointer), used by every later CPR set-up until replaced; null clears them (quasi-IMPES, the default).  A cell whose
 * block is ill-conditioned (|B|_1 |B^-1|_1 > 1e10) keeps its quasi-IMPES weight (counted in rsh_system_info's
 * cpr_impes_fallbacks).  The GpuBlockSolver hook carries no accumulation blocks; this serves engine dumps and
 * callers that have them. */
int rsh_system_set_accumulation(rsh_system* sys, const double* acc);
int rsh_system_set_accumulation_device(rsh_system* sys, const double* acc_dev);
/* build (first call) or refresh the preconditioner for the current values */
int rsh_system_setup(rsh_system* sys, const rsh_options* opt);
/* solve A x = b to relative residual tol (the recursive residual of BiCGSTAB as linsolve.cpp:210-212, the FGMRES
 * estimate checked against the true residual at restarts, the FP64 true residual for RSH_MIXED_IR).  Host pointers (n*nb doubles) or device pointers. */
int rsh_solve(rsh_system* sys, const double* b, double* x, double tol, int max_it, int* iterations, double* rel_res);
int rsh_solve_device(rsh_system* sys, const double* b_dev, double* x_dev, double tol, int max_it, int* iterations,
                     double* rel_res);

/* ---- a solve that allocates nothing (no counterpart in the reference): for callers that cannot call hipMalloc -
 * tight serving loops and harnesses that forbid allocation.  The same BiCGSTAB as rsh_solve_device on a system already
 * set up (RSH_FP64 or RSH_MIXED_STORE, RSH_KRYLOV_BICGSTAB; other forms: RSH_ERR_UNSUPPORTED), with its Krylov
 * vectors, scalars and partial sums in the library's static device scratch: 64 MiB per GPU, held until the process
 * ends, i.e. at most 1 047 552 unknowns (n * nb; larger systems: RSH_ERR_ARG).  Results are bit-identical to
 * rsh_solve_device.  Every call on one GPU shares that scratch: calls that overlap on one GPU (from several threads
 * or contexts) race, so the caller serialises them.  The call synchronises its context's stream (it is not meant for
 * stream capture). */
int rsh_solve_device_noalloc(rsh_system* sys, const double* b_dev, double* x_dev, double tol, int max_it, int* iterations,
                             double* rel_res);

/* ---- kernel-level entry points (host vectors), for parity checks against BSR::spmv / Preconditioner::apply */
int rsh_spmv(rsh_system* sys, const double* x, double* y);
/* Rows per wavefront of the FP64 operator's SpMV (no counterpart in the reference, whose BSR::spmv, bsr.cpp:109-126,
 * walks one row per thread): lanes = 1 (64 rows per wavefront, the default), 2 or 4 (32 or 16 rows, each row's
 * blocks split over its lanes and their partial sums added pairwise; another summation order, within rounding of
 * lanes = 1).  Used by rsh_spmv and the Krylov solves; the FP32 operator and the CPR second stage keep one lane per
 * row.  Takes another copy of the matrix values in HBM while lanes > 1. */
int rsh_system_set_spmv_lanes(rsh_system* sys, int lanes);
int rsh_precond_apply(rsh_system* sys, const double* r, double* z);
int rsh_amg_apply(rsh_system* sys, const double* r, double* z);          /* CPR only: one V-cycle, n doubles */
int rsh_cpr_weights(rsh_system* sys, double* w);                         /* CPR only: n*nb quasi-IMPES weights */
int rsh_bilu0_factors(rsh_system* sys, double* lu);                      /* natural-order factors, S's layout */
/* the ILU's working order: perm[w] = the original row of working row w (level order, tile order, colour order or
 * red / black tile order); the factors of rsh_bilu0_factors belong to P A P^T in this order in the multicolour forms */
int rsh_ilu_order(rsh_system* sys, int* perm);
int rsh_amg_levels(rsh_system* sys, int* nlevels, int* sizes, int* nnz, int cap);
int rsh_system_info(rsh_system* sys, char* json, int cap);               /* levels, colours, bytes, last solve */

/* ---- device memory helpers (benchmarks: inputs resident in HBM) */
int rsh_device_alloc(rsh_context* ctx, size_t bytes, void** ptr);
int rsh_device_free(rsh_context* ctx, void* ptr);
int rsh_memcpy_h2d(rsh_context* ctx, void* dst, const void* src, size_t bytes);
int rsh_memcpy_d2h(rsh_context* ctx, void* dst, const void* src, size_t bytes);

/* ---- measurement: average duration (ms) of `reps` back-to-back launches of one kernel class on the context's
 * stream, bracketed by HIP events.  kind: 0 BSR SpMV (precision of the system setup), 1 preconditioner apply,
 * 2 ILU forward+backward sweeps, 3 AMG V-cycle, 4 copy kernel of `bytes` (bandwidth calibration), 5 / 6 one
 * damped-Jacobi smoothing step of AMG level 0 / level 1. */
int rsh_time_kernel(rsh_system* sys, int kind, int reps, double* avg_ms, double* bytes_per_launch);

/* ---- well Schur reduction on the device (replaces the host-side WellSchur::build / reduceRhs / recover,
 * bsr.cpp:733-990, that the engine runs before calling the hook, simulator.cpp:3320-3330): the engine's scalar
 * Jacobian (cells in blocks of nv unknowns, then one BHP row per well; row_ptr / col of n = na*nv + nw rows) is
 * handed over as it is assembled.  The reduced block pattern and the entry maps are built once per pattern; each
 * Newton iteration uploads only J's values (rsh_wells_set_values) and solves with the full right-hand side
 * (rsh_wells_solve), getting back cells and BHPs.  rsh_wells_system gives the reduced resident system for
 * rsh_system_setup (and any other rsh_system call). */
int rsh_wells_create(rsh_context* ctx, int n, const int* row_ptr, const int* col, int na, int nv, int nw, rsh_wells** out);
int rsh_wells_destroy(rsh_wells* w);
int rsh_wells_system(rsh_wells* w, rsh_system** sys);
int rsh_wells_set_values(rsh_wells* w, const double* jval, int on_device);          /* RSH_ERR_UNSUPPORTED: bad well row */
int rsh_wells_solve(rsh_wells* w, const double* b_full, double* x_full, int on_device, double tol, int max_it, int* iterations,
                    double* rel_res);
int rsh_wells_reduced_values(rsh_wells* w, double* out);                         /* S's values to the host (checks) */

/* ---- batched solves of many independent systems (no counterpart in the reference, whose engine solves one
 * system per call; this serves its ensemble callers, runBatch study.cpp:326-364, in one launch stream).
 * Systems of one block size are concatenated into one block-diagonal resident system; each system keeps its own
 * BiCGSTAB scalars and convergence.  Every preconditioner of the single-system path is accepted (BILU0, MC-BILU0
 * and the three CPR forms; FP64, FP32 with refinement (RSH_MIXED) or FP32-stored (RSH_MIXED_STORE)): the ILU forms
 * are the ones each system gets alone; CPR builds one pressure hierarchy per block-size group whose aggregates
 * never span two systems, so the preconditioner stays block-diagonal (a system's result does not depend on the
 * others) while the group's coarsest level holds every system's coarsest aggregates.  BiCGSTAB only (FGMRES and
 * RSH_MIXED_IR are rejected).  One BiCGSTAB iteration over all systems is one captured HIP graph.  The groups'
 * numeric set-ups run concurrently (a stream and a host thread each).  The level chunking of the sweeps and the
 * LDS tile of a fused run (its row limit) are tuned per group (cached), the entries per workgroup of the vector
 * kernels at the first setup. */
int rsh_batch_create(rsh_context* ctx, int count, const int* n, const int* nb, const int* const* row_ptr, const int* const* col,
                     rsh_batch** out);
int rsh_batch_destroy(rsh_batch* batch);
int rsh_batch_set_values(rsh_batch* batch, const double* const* val);           /* host: val[i] = nnzb_i*nb_i^2 */
int rsh_batch_set_values_device(rsh_batch* batch, const double* val_dev);       /* device: input-order concatenation */
int rsh_batch_setup(rsh_batch* batch, const rsh_options* opt);
/* per-system tol[i]; iterations / rel_res / status per system (may be null).  Returns RSH_OK when every system
 * converged, else the first failing status (the converged systems' x are valid either way). */
int rsh_batch_solve(rsh_batch* batch, const double* const* b, double* const* x, const double* tol, int max_it, int* iterations,
                    double* rel_res, int* status);
int rsh_batch_solve_device(rsh_batch* batch, const double* b_dev, double* x_dev, const double* tol, int max_it, int* iterations,
                           double* rel_res, int* status);
int rsh_batch_latencies(rsh_batch* batch, double* ms);   /* per system: ms from the solve's start to its convergence */
/* Rows per wavefront of each block-size group's SpMV (no counterpart in the reference): 0 = chosen by the batch
 * autotuner at the next rsh_batch_setup (1, 2 or 4 lanes per row timed on the group's resident operator, cached per
 * block size and size bucket; the default), or 1, 2, 4 fixed (see rsh_system_set_spmv_lanes).  Applies at the next
 * rsh_batch_setup (a changed choice makes rsh_batch_solve wait for that set-up: it is refused until then);
 * rsh_batch_info reports each group's "spmv_lanes" and, when they were timed, "spmv_lanes_ms". */
int rsh_batch_set_spmv_lanes(rsh_batch* batch, int lanes);
int rsh_batch_info(rsh_batch* batch, char* json, int cap);

/* ---- serving loop for ensemble callers (no counterpart in the reference; its ensemble drivers runBatch / runESMDA,
 * study.cpp:326-364, 1254, call the global hook from many Simulator threads, one Newton iteration per call).
 * Models (one pattern each) stay resident with their slice layouts, ILU schedule and AMG hierarchy; requests queue
 * in arrival order and are admitted as soon as a worker slot (one HIP stream each) is free, the model is idle and
 * the model fits the HBM budget (idle models are evicted least-recently-used, and only when that makes room).  A
 * model larger than the whole budget is served alone once nothing else is resident.  Submit and wait may be called
 * from many threads at once; a failed request's message is rsh_last_error() of the thread that waits for it.
 * Pointers in a request must stay valid until rsh_server_wait returns for it. */
typedef struct rsh_request {
    int model;
    const double* values;          /* nnzb*nb*nb new values, or NULL for the model's base values */
    int values_on_device;
    double perturb;                /* > 0: block row i *= 1 + perturb*u_i, u_i uniform in [-1,1) hashed from
                                      (perturb_seed, i): synthetic Newton iterations of an ensemble member */
    unsigned long long perturb_seed;
    const double* rhs;             /* n*nb */
    int rhs_on_device;
    double* x;                     /* n*nb, or NULL */
    int x_on_device;
    double tol;
    int max_it;
} rsh_request;
typedef struct rsh_result {
    int status;                    /* RSH_OK or the solve's error code */
    int iterations;
    double rel_res;
    double queue_ms, service_ms, latency_ms;   /* submit -> start, start -> end, submit -> end */
} rsh_result;
int rsh_server_create(int device, int slots, double hbm_budget_bytes, rsh_server** out);
/* The same loop serving requests in batched rounds through the batch engine (rsh_batch_*), so that several requests
 * share each kernel launch (no counterpart in the reference; for ensembles of many small models, whose requests are
 * bound by kernel launches rather than bytes).  At the first request the models are grouped (same block size and
 * options, ascending size, a group holds at most group_cells cells — or one larger model — and at most group_systems
 * models); every model must be added before it.  A round takes every queued request of one idle group (the earliest
 * per model), applies their values, runs one numeric set-up of the whole group and one batched BiCGSTAB in which the
 * group's other models solve a zero right-hand side (they stop at once).  A request's result does not depend on the
 * other requests of its round (the batch's preconditioner is block-diagonal, its scalars per system), but it is the
 * batch engine's solve (CPR: one pressure hierarchy per group), not the single-system one.  A round runs to the
 * largest max_it among its requests.  Groups are not evicted: a request whose group is not resident and does not fit
 * the budget left beside the resident ones fails with RSH_ERR_UNSUPPORTED (size the budget for every group; the first
 * group is always admitted).  RSH_MIXED_IR and FGMRES are refused by the batch engine. */
int rsh_server_create_batched(int device, int slots, double hbm_budget_bytes, long long group_cells, int group_systems, rsh_server** out);
int rsh_server_destroy(rsh_server* srv);   /* finishes the queued requests first */
int rsh_server_add_model(rsh_server* srv, int n, int nb, const int* row_ptr, const int* col, const double* values, const rsh_options* opt,
                         int* model);
int rsh_server_submit(rsh_server* srv, const rsh_request* req, long long* ticket);
int rsh_server_wait(rsh_server* srv, long long ticket, rsh_result* res);
int rsh_server_stats(rsh_server* srv, char* json, int cap);

#ifdef __cplusplus
}
#endif
#endif
