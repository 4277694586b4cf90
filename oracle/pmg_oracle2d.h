/*
 * pmg_oracle2d.h - FP64 CPU restatement of the 2D twins of ParMooN's NSE type-4
 * GMG-Vanka / FGMRES path (Q2/P1disc quadrilaterals, 21-dof cell patches).
 *
 * TEST INFRASTRUCTURE ONLY, like pmg_oracle.h: used by tests/ as the checked statement of
 * the 2D configuration (steady Stokes on the unit square).  The reference has no 2D
 * multigrid driver (src/System/SystemNSE2D.C:808-809), so this path is CPU-only; it
 * restates the 2D functions the 3D driver calls in their __2D__ form:
 *   CoupledDefect / CoupledMatVect NSE4 2D   src/FE/MGComponents2D.C:2578-2850
 *   IntoL20Vector2D case -11 (stride 3)       src/FE/MGComponents2D.C:1491-1508
 *   Prolongate / DefectRestriction 2D         src/FE/MGComponents2D.C:118-1022
 *   CellVanka (colored), SolveExact, Cycle    src/FE/NSE_MGLevel4.C, src/FE/NSE_MultiGrid.C
 *   TFgmresIte::Iterate                       src/FE/FgmresIte.C:250-677
 * and generates the problem the way the 2D setup does: Q2 (NF_C_Q_Q2_2D, 9 nodes, xi
 * fastest) and P1disc with basis 1, 3 xi, 3 eta (include/FE2D/BF_D_Q_P1_2D.h) and the
 * moment functionals of include/FE2D/NF_D_Q_P1_2D.h, NSTYPE 4 Galerkin with Laplace type
 * 0, 3x3 Gauss quadrature, cells refined 4 to 1 with cell = 4 * parent + child.
 */
#ifndef PMG_ORACLE2D_H
#define PMG_ORACLE2D_H
#include "pmg.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle2d oracle2d;

/* right-hand sides */
enum { O2D_RHS_STOKES_CURL = 0,   /* u = curl(sin^2(pi x) sin^2(pi y)), p = x^3 + y^3 - 1/2 */
       O2D_RHS_QUADLIN = 1 };     /* u quadratic, p linear (reproduced exactly)            */

/* unit square, coarsest level of 'coarse' x 'coarse' cells, n_levels levels (each refined
   4:1), viscosity nu, Dirichlet data on the whole boundary */
int  oracle2d_create(const pmg_params *params, int coarse, int n_levels, double nu, int rhs_kind,
                     oracle2d **out);
void oracle2d_destroy(oracle2d *o);
/* sizes of a level: n_udof, n_active, n_pdof, n_cells, n_colors (any may be NULL) */
void oracle2d_sizes(const oracle2d *o, int level, int *n_udof, int *n_active, int *n_pdof, int *n_cells,
                    int *n_colors);
int  oracle2d_n_dof(const oracle2d *o, int level);
/* [f1 f2 g] of a level (Dirichlet rows hold the boundary values); nodal interpolant of
   the exact solution [u1 u2 p] (P1disc: the moment functionals of p); dof coordinates */
void oracle2d_rhs(const oracle2d *o, int level, double *out);
void oracle2d_exact(const oracle2d *o, int level, double *out);
void oracle2d_dof_coords(const oracle2d *o, int level, double *xy);
/* CSR arrays of a level: A (rowptr n_udof+1, kcol, 4 blocks A11 A12 A21 A22),
   B (rowptr n_pdof+1, kcol, 2 blocks), BT (rowptr n_udof+1, kcol, 2 blocks); returns nnz */
int  oracle2d_nnz(const oracle2d *o, int level, int *a_nnz, int *b_nnz, int *bt_nnz);
void oracle2d_csr(const oracle2d *o, int level, int *a_rowptr, int *a_kcol, double *a4, int *b_rowptr,
                  int *b_kcol, double *b2, int *bt_rowptr, int *bt_kcol, double *bt2);
void oracle2d_cell_dofs(const oracle2d *o, int level, int *u_global9, int *p_global3);

/* level operations (vectors [u1 u2 p]) */
void oracle2d_matvec(oracle2d *o, int level, const double *x, double *y);
void oracle2d_defect(oracle2d *o, int level, const double *u, const double *f, double *d, double *res);
int  oracle2d_cell_vanka(oracle2d *o, int level, double *u, const double *f);
void oracle2d_prolongate(oracle2d *o, int fine_level, const double *u_coarse, double *d_fine);
void oracle2d_restrict_defect(oracle2d *o, int fine_level, double *d_fine, double *f_coarse);
int  oracle2d_solve_exact(oracle2d *o, double *u, const double *f);
int  oracle2d_vcycle(oracle2d *o, const double *v, double *z);
int  oracle2d_fgmres(oracle2d *o, double *sol, const double *rhs, int *iters, double *final_res);
/* local 21x21 patch matrix of one cell, column-major sys[col*21+row] */
int  oracle2d_patch_matrix(oracle2d *o, int level, int cell, double *sys);

#ifdef __cplusplus
}
#endif
#endif
