/* Checker API shared by the oracle restatement (oracle/oracle.cpp -> oracle/liboracle.so) and the shim over the
 * compiled reference sources (oracle/ref_shim.cpp -> oracle/_ref/libressim_ref.so).  TEST INFRASTRUCTURE ONLY:
 * nothing in the product library links, loads or calls these functions.
 *
 * Matrices are BSR with int32 indices: n block rows, nb x nb row-major blocks, rp[n+1], ci[nnzb] (sorted per row,
 * diagonal present), v[nnzb*nb*nb].  Vectors are n*nb doubles.  Handles own copies of what they need.
 *
 * Preconditioner ids for chk_krylov:
 *   1 BILU0 (double factors)                       bsr.cpp:347-596
 *   2 BILU0, float32 factors after a double factorisation, double vectors (C38)   bsr.cpp:413-419, 446-481
 *   3 BILU0f: float32 factors and float32 sweeps (C16)                            bsr.cpp:1087-1121
 *   4 MCBILU0 (multicolour)                         bsr.cpp:599-730
 *   5 BlockCPR, quasi-IMPES weights, double BILU0 second stage                    bsr.cpp:993-1084
 * Methods: 0 BiCGSTAB (linsolve.cpp:159-216), 1 flexible restarted GMRES (linsolve.cpp:218-273).
 */
#ifndef RS_CHECK_H
#define RS_CHECK_H
#ifdef __cplusplus
extern "C" {
#endif

int chk_spmv(int n, int nb, const int* rp, const int* ci, const double* v, const double* x, double* y);

void* chk_bilu0_new(int n, int nb, const int* rp, const int* ci, const double* v, int mode);
int chk_bilu0_apply(void* h, const double* r, double* z);
int chk_bilu0_regularised(void* h);
void chk_bilu0_free(void* h);
/* oracle only: the natural-order factors on the input pattern (L, U_ii^-1 on the diagonal, U) */
int chk_bilu0_factors(void* h, double* lu);

void* chk_mcbilu0_new(int n, int nb, const int* rp, const int* ci, const double* v);
int chk_mcbilu0_apply(void* h, const double* r, double* z);
int chk_mcbilu0_colours(void* h);
int chk_mcbilu0_colouring(void* h, int* colour, int* perm);
void chk_mcbilu0_free(void* h);

void* chk_cpr_new(int n, int nb, const int* rp, const int* ci, const double* v);
// true-IMPES weights from n accumulation blocks (BlockCPR::setup accBlocks, bsr.cpp:993-1030)
void* chk_cpr_new_acc(int n, int nb, const int* rp, const int* ci, const double* v, const double* acc);
int chk_cpr_fallbacks(void* h);
int chk_cpr_refresh(void* h, const double* v);
int chk_cpr_apply(void* h, const double* r, double* z);
int chk_cpr_weights(void* h, double* w);
int chk_cpr_nlevels(void* h);
int chk_cpr_level(void* h, int l, int* n, int* nnz);
int chk_cpr_level_matrix(void* h, int l, int* rp, int* ci, double* v);
int chk_cpr_aggregates(void* h, int l, int* agg);
int chk_cpr_amg_apply(void* h, const double* r, double* z);
void chk_cpr_free(void* h);
/* DeviceCprAmg (bsr.cpp:1235-1371): the reference's float32 device CPR with the block-Jacobi second stage */
void* chk_devcpr_new(int n, int nb, const int* rp, const int* ci, const double* v);
int chk_devcpr_apply(void* h, const double* r, double* z);
void chk_devcpr_free(void* h);

/* reference only: the engine's well Schur reduction of a scalar Jacobian (WellSchur, bsr.cpp:733-990); null when
 * refused.  sizes = {reduced block rows, reduced blocks, fill blocks} */
void* chk_wellschur_new(int n, const int* rp, const int* ci, const double* v, int na, int nv, int nw, int* sizes);
int chk_wellschur_reduced(void* h, int* rp, int* ci, double* v);
int chk_wellschur_reduce_rhs(void* h, const double* b, double* bc);
int chk_wellschur_recover(void* h, const double* b, const double* xc, double* x);
void chk_wellschur_free(void* h);

/* the float32 device solve (CpuFloatBlockSolver::solve, bsr.cpp:1130-1215): kind 0 BILU0f, 1 MC-BILU0 (multicolour),
 * 2 DeviceCprAmg; returns 1 (finite solution) / 0 / -1 (set-up failed).  chk_refine_float: refineWithDevice
 * (bsr.cpp:1373-1396) around it; returns 1 when converged */
int chk_float_solve(int n, int nb, const int* rp, const int* ci, const double* v, const double* b, double* x, int kind, double tol, int maxit,
                    int* its);
int chk_refine_float(int n, int nb, const int* rp, const int* ci, const double* v, const double* b, double* x, int kind, double tol, int max_outer,
                     int inner_it, int* its, double* relres);

int chk_krylov(int n, int nb, const int* rp, const int* ci, const double* v, const double* b, double* x, int precond,
               int method, int restart, double tol, int maxit, int* its, double* relres);
/* CPR-BiCGSTAB in the Newton loop's steady state: a first CPR set-up on the matrix (not timed), then the set-up with the
 * AMG aggregates reused (BlockCPR::setup(S, acc, reuseAmg = true), bsr.cpp:993-1059, AMG::refresh amg.cpp:257-271)
 * and the solve, timed: secs[0] = refresh set-up seconds, secs[1] = solve seconds.  Returns 1 when converged */
int chk_krylov_cpr_steady(int n, int nb, const int* rp, const int* ci, const double* v, const double* b, double* x, double tol,
                          int maxit, int* its, double* relres, double* secs);

#ifdef __cplusplus
}
#endif
#endif
