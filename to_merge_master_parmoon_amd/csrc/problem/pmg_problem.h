/*
 * pmg_problem.h - synthetic Q2/P1disc Oseen problem hierarchies in ParMooN's data layout.
 *
 * Produces what the reference's setup (3DPrograms/NSE3D_ParMooN.C + TSystemNSE3D,
 * src/System/SystemNSE3D.C:69-500) hands to its multigrid: per level the type-4 block
 * matrices A11..A33, B1..B3, B1T..B3T (assembled like NSType4Galerkin3D,
 * src/FE/NSE3D_FixPo.C:1919-2071), the dof maps, the cell hierarchy and the per-child
 * prolongation matrices QQ (TFEDatabase3D::GetProlongationMatrix3D,
 * src/FE/FEDatabase3D.C:1630-1720).  Meshes are structured boxes refined uniformly;
 * children of a parent are contiguous (cell = 8*parent + child), as RegRefineAll's
 * depth-first collection order gives.
 */
#ifndef PMG_PROBLEM_H
#define PMG_PROBLEM_H
#include "pmg.h"

#ifdef __cplusplus
extern "C" {
#endif

/* PMGP_CONV_CHANNEL: b = (conv_max 16 y(1-y) z(1-z), 0, 0) with y, z scaled by ly, lz; the
   same profile is the Dirichlet value of the velocity (inflow at x = 0, zero on the walls) */
enum { PMGP_CONV_NONE = 0, PMGP_CONV_FOURIER = 1, PMGP_CONV_CONST = 2, PMGP_CONV_CHANNEL = 3 };
enum { PMGP_RHS_ZERO = 0, PMGP_RHS_FOURIER = 1, PMGP_RHS_EXACT = 2 };
enum { PMGP_EXACT_NONE = 0, PMGP_EXACT_QUADLIN = 1 };

typedef struct {
  int cx, cy, cz;              /* coarsest grid, cells per direction                     */
  double lx, ly, lz;           /* box [0,lx]x[0,ly]x[0,lz]                               */
  int n_levels;                /* levels; level n_levels-1 is the finest                 */
  double nu;                   /* viscosity                                              */
  int dirichlet_faces;         /* bit mask 1:x=0 2:x=lx 4:y=0 8:y=ly 16:z=0 32:z=lz      */
  double perturb;              /* interior coarsest-grid vertices moved by U[-p,p]*h     */
  unsigned long long seed_perturb;
  int conv_kind;               /* PMGP_CONV_*                                            */
  double conv_const[3];        /* PMGP_CONV_CONST: b                                     */
  unsigned long long seed_conv;
  double conv_max;             /* PMGP_CONV_FOURIER: max_{quad points} |b| after scaling */
  int rhs_kind;                /* PMGP_RHS_*                                             */
  unsigned long long seed_rhs;
  int exact_kind;              /* PMGP_EXACT_*: Dirichlet data and f from a known (u,p)  */
  int n_modes;                 /* Fourier modes for b and f (default 8)                  */
  double mass;                 /* sigma of the sigma (u, v) term a time step adds to A   */
} pmgp_config;

typedef struct pmgp_problem pmgp_problem;

void pmgp_config_default(pmgp_config *c);
int  pmgp_create(const pmgp_config *c, pmgp_problem **out);
void pmgp_destroy(pmgp_problem *p);
int  pmgp_n_levels(const pmgp_problem *p);
/* level descriptor pointing into the problem's own arrays (valid until pmgp_destroy) */
int  pmgp_level_desc(const pmgp_problem *p, int level, pmg_level_desc *out);
int  pmgp_n_dof(const pmgp_problem *p, int level);
/* right-hand side [f1 f2 f3 g] of a level (Dirichlet rows hold the boundary values) */
const double *pmgp_rhs(const pmgp_problem *p, int level);
/* nodal interpolant of the exact solution (exact_kind != NONE), else NULL */
const double *pmgp_exact(const pmgp_problem *p, int level);
/* coordinates of the velocity dofs (3 per dof) and of the cell centres (3 per cell) */
const double *pmgp_dof_coords(const pmgp_problem *p, int level);
const double *pmgp_cell_centers(const pmgp_problem *p, int level);
/* scale applied to the Fourier convection so that max |b| = conv_max */
double pmgp_conv_scale(const pmgp_problem *p);
/* cells per direction of a level */
int  pmgp_level_dims(const pmgp_problem *p, int level, int *nx, int *ny, int *nz);
/* cell corners of a level (8 per cell, index a + 2b + 4c, 3 coordinates each) */
int  pmgp_cell_vertices(const pmgp_problem *p, int level, double *out);
/* the convection field of the configuration at the 27 Gauss points of every cell
   (out[(cell * 27 + q) * 3 + c], q = i + 3j + 9k), the values the assembly used */
int  pmgp_conv_quad(const pmgp_problem *p, int level, double *out);
/* mass-lumped L2 helpers are not needed: errors are measured on dof vectors */

#ifdef __cplusplus
}
#endif
#endif
