/*
 * pmg_oracle.h - FP64 CPU restatement of ParMooN's NSE type-4 GMG-Vanka / FGMRES path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle and the timed CPU
 * baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker or the baseline; the product path (libpmg) never
 * calls it.
 *
 * Every function restates the reference file:line given in its comment.  The reference
 * itself cannot be compiled here (its hot-path headers include cuda.h, magma_v2.h and
 * mkl.h unconditionally), so this restatement is pinned by the closed-form and brute-force
 * checks in tests/test_oracle_*.py.
 *
 * The step sequence of Cycle for V(nu1, nu2) with smoothing is pinned against a
 * matrix-algebra statement of NSE_MultiGrid.C:434-1779 built from the separately pinned
 * sweep, transfers and coarse solve.  The greedy color order of oracle_color_cells is the
 * reference's by reading, checked for validity and against an O(N^2) restatement.
 */
#ifndef PMG_ORACLE_H
#define PMG_ORACLE_H
#include "pmg.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_mg oracle_mg;

/* levels[0] is the coarsest.  The descriptors (and the arrays they point to) must stay
   alive while the oracle is used.  refactor: 1 = reference-faithful (re-assemble every
   local 85x85 matrix by linear search and LU it in every sweep, refactor the dense coarse
   matrix at every visit); 0 = factor once and reuse (bit-identical results). */
int  oracle_create(const pmg_params *params, const pmg_level_desc *levels, int n_levels,
                   int refactor, oracle_mg **out);
void oracle_destroy(oracle_mg *mg);
int  oracle_n_dof(const oracle_mg *mg, int level);
/* 0 = colored variant (_HYBRID, NSE_MGLevel4.C:603-1649), 1 = Gauss-Seidel (_SEQ, :1653-2085) */
void oracle_set_vanka_variant(oracle_mg *mg, int variant);
void oracle_set_threads(int n);

/* level operations */
void oracle_defect(oracle_mg *mg, int level, const double *u, const double *f, double *d, double *res);
void oracle_matvec(oracle_mg *mg, int level, const double *x, double *y);
void oracle_defect_nse4(oracle_mg *mg, int level, const double *x, const double *b, double *r);
int  oracle_cell_vanka(oracle_mg *mg, int level, double *u, const double *f, double *aux);
void oracle_correct_nodes(oracle_mg *mg, int level, double *u);
void oracle_set_dirichlet(oracle_mg *mg, int level, double *u, const double *f);
void oracle_correct_defect(oracle_mg *mg, int level, double *d);
void oracle_update(oracle_mg *mg, int level, double *u, const double *d);
/* fine_level >= 1; d_fine is divided by the multiplicity and restored, like the reference */
void oracle_restrict_defect(oracle_mg *mg, int fine_level, double *d_fine, double *f_coarse, double *aux);
void oracle_prolongate(oracle_mg *mg, int fine_level, const double *u_coarse, double *d_fine, double *aux);
int  oracle_solve_exact(oracle_mg *mg, double *u, const double *f);
/* RestrictFunction of the velocity (MGComponents3D.C:1509-1755), fine_level >= 1: the
   velocity part of u_coarse is overwritten, aux holds n_udof(coarse) doubles */
void oracle_restrict_function(oracle_mg *mg, int fine_level, const double *u_fine, double *u_coarse, double *aux);

/* drivers */
int  oracle_vcycle(oracle_mg *mg, const double *v, double *z);        /* TMultiGridIte::Iterate */
int  oracle_fgmres(oracle_mg *mg, double *sol, const double *rhs, int *iters, double *final_res);
int  oracle_fgmres_history(const oracle_mg *mg, double *out, int cap);
/* per-level defect norms recorded by the last V-cycle (MG_DEBUG prints); returns count */
int  oracle_cycle_log(const oracle_mg *mg, double *out, int cap);

/* dense LU with partial pivoting, LAPACK dgetf2 semantics, column-major a[col*n+row];
   returns info (0 ok, k>0: U(k-1,k-1) == 0).  Updates are explicit fma(). */
int  oracle_lu_factor(int n, double *a, int *ipiv);
/* dgetrs('N') with one right-hand side: row swaps, unit-lower then upper column sweeps */
void oracle_lu_solve(int n, const double *a, const int *ipiv, double *b);

/* colorCell restated with the O(N^2) pairwise adjacency (ParFEMapper3D.C:750-1084) */
int  oracle_color_cells(int n_cells, int n_ulocal, const int *u_begin, const int *u_global,
                        int *n_colors, int *ptr_colors, int *cell_reorder);

/* local 85x85 patch matrix of one cell exactly as CellVanka assembles it
   (NSE_MGLevel4.C:1066-1290), column-major */
int  oracle_patch_matrix(oracle_mg *mg, int level, int cell, double *sys);

#ifdef __cplusplus
}
#endif
#endif
