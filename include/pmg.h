/*
 * pmg.h - C ABI of the MI355X (gfx950) geometric-multigrid cell-Vanka preconditioner
 * for Q2/P1disc Navier-Stokes (Oseen) saddle-point systems inside FGMRES.
 *
 * The entry points replace the reference's NSE type-4 multigrid path.  Each declaration
 * names the reference member or free function it stands in for (paths relative to the
 * ParMooN source tree).  All arrays crossing this boundary are plain C arrays; no C++,
 * HIP or torch types appear in the signatures.  Every call returns a pmg_status; none of
 * them calls exit() (the reference exits with 4711 on divergence, see
 * src/FE/NSE_MultiGrid.C:629-633 and src/FE/FgmresIte.C:564-567).
 *
 * Vector layout (same as the reference, src/FE/NSE_MGLevel4.C:257-268):
 *   x = [ u1 (n_udof) | u2 (n_udof) | u3 (n_udof) | p (n_pdof) ]
 * Velocity numbering: active [0, n_active), hanging [n_active, hanging_bound),
 * Dirichlet [hanging_bound, n_udof).  P1disc pressure is cell-major, 4 per cell,
 * constant mode first.
 */
#ifndef PMG_H
#define PMG_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ----------------------------------------------------------------------------------- */
/* status codes                                                                          */
/* ----------------------------------------------------------------------------------- */
typedef enum {
  PMG_OK = 0,
  PMG_ERR_INVALID_ARGUMENT = 1,
  PMG_ERR_DEVICE = 2,           /* a HIP runtime call failed                            */
  PMG_ERR_OUT_OF_MEMORY = 3,
  PMG_ERR_NOT_READY = 4,        /* levels missing / setup not done                      */
  PMG_ERR_SINGULAR_PATCH = 5,   /* zero pivot in a local Vanka patch (dgetrf info>0)     */
  PMG_ERR_SINGULAR_COARSE = 6,  /* zero pivot in the coarse dense LU                     */
  PMG_ERR_NONFINITE = 7,        /* NaN/Inf in an input vector or residual                */
  PMG_ERR_DIVERGED = 8,         /* FGMRES residual > div_factor * res0 (reference: exit) */
  PMG_ERR_NOT_CONVERGED = 9,    /* maxit reached without meeting the tolerances          */
  PMG_ERR_UNSUPPORTED = 10
} pmg_status;

/* arithmetic of the smoother / level operator (the outer FGMRES is always FP64) */
typedef enum {
  PMG_PREC_FP64 = 0,            /* plain FP64 path (parity with the FP64 oracle)         */
  PMG_PREC_FP32 = 1,            /* FP32 operator copy + FP32 patch factors               */
  PMG_PREC_FP32_BF16 = 2        /* FP32 operator copy + BF16 patch factors, FP32 refine  */
} pmg_precision;

/* where vector arguments live */
typedef enum { PMG_MEM_HOST = 0, PMG_MEM_DEVICE = 1 } pmg_memspace;

/*
 * Solver knobs.  Field names follow TDatabase::ParamDB (src/General/Database.C:640-749,871).
 * pmg_params_default() returns the reference defaults.
 */
typedef struct {
  int    sc_smoother_saddle;                   /* 2: colored cell Vanka (_HYBRID order); 1: cell
                                                  Vanka in the Gauss-Seidel order (cells in
                                                  collection order, run as the level sets of
                                                  pmg_sequential_schedule; any coloring in the
                                                  level descriptors is then ignored); 3 / 4:
                                                  nodal Vanka (patch = one pressure dof and the
                                                  velocity dofs of its B row, pressure dofs in
                                                  order as the level sets of pmg_nodal_schedule)
                                                  with the diagonal (3) or the full (4) local
                                                  system; FP64 and single problems only.  Fixed
                                                  once levels exist */
  int    sc_pre_smooth_saddle;                 /* nu1 */
  int    sc_post_smooth_saddle;                /* nu2 */
  double sc_smooth_damp_factor_saddle;         /* global post-sweep relaxation alpha      */
  double sc_smooth_damp_factor_coarse_saddle;  /* per-cell Vanka damping 'damp'           */
  double sc_gmg_damp_factor_saddle;            /* coarse-grid-correction damping beta     */
  int    sc_coarse_smoother_saddle;            /* 17 = dense direct solve                 */
  int    internal_project_pressure;            /* 1 = mean-zero P1disc constants          */
  int    sc_gmres_restart;
  int    sc_lin_maxit_saddle;
  double sc_lin_red_factor_saddle;
  double sc_lin_res_norm_min_saddle;
  double sc_div_factor;
  int    sc_minit;
  int    sc_amg_prec_it;
  int    precision;                            /* pmg_precision                           */
  int    bf16_refine_steps;                    /* 0..3 refinement steps of every local solve of the
                                                  mixed forms with stored inverses (patch_inverse
                                                  = 1): x += M^{-1} (r - K x) with M^{-1} the
                                                  stored (BF16 or FP32) data and K the FP32 patch
                                                  matrix; 0 (default): one product, the outer FP64
                                                  FGMRES refines */
  int    use_graph;                            /* 1: every V-cycle replays a HIP graph captured on
                                                  first use (default); 0: launched kernel by kernel */
  int    shared_block;                         /* 1 (default): when A11 = A22 = A33 and the other
                                                  A blocks vanish on the active rows (Galerkin Oseen
                                                  form, NSE3D_FixPo.C:1919-2071), checked after every
                                                  value change, the velocity rows stream that one
                                                  block; 0: always the nine blocks */
  int    patch_inverse;                        /* mixed precision only.  1 (default): each local
                                                  Vanka solve is a product with data stored once per
                                                  operator in the factor precision: with the shared
                                                  velocity block and colors of more than 2048
                                                  patches A^{-1} of the 27 x 27 block, A^{-1} BT_c
                                                  and the pressure rows [-S^{-1} B_c A^{-1} | S^{-1}]
                                                  of the patch inverse (S the 4 x 4 Schur complement),
                                                  otherwise the explicit inverse of the patch matrix
                                                  (with the shared block expanded from that block
                                                  form; nine blocks or refinement steps: FP32
                                                  Gauss-Jordan with the pivots of dgetf2); 0:
                                                  the stored LU factors and triangular sweeps.  FP64
                                                  always uses the LU factors (dgetrf/dgetrs,
                                                  LinAlg.C:106-130) */
} pmg_params;

/*
 * One multigrid level, in the reference's own data layout (borrowed during pmg_level_add,
 * copied to the device).  CSR arrays are the TSquareStructure3D / TStructure3D arrays of
 * include/FE/Matrix.h:91-129: A rows 0..n_udof-1 (Dirichlet rows hold only the diagonal),
 * B rows 0..n_pdof-1 over velocity columns, BT rows 0..n_udof-1 over pressure columns.
 */
typedef struct {
  int n_udof, n_active, hanging_bound, n_dirichlet, n_pdof;
  const int *a_rowptr, *a_kcol; const double *a[9];   /* A11 A12 A13 A21 A22 A23 A31 A32 A33 */
  int a_nnz;
  const int *b_rowptr, *b_kcol; const double *b[3];   /* B1 B2 B3   */
  int b_nnz;
  const int *bt_rowptr, *bt_kcol; const double *bt[3];/* B1T B2T B3T */
  int bt_nnz;
  /* cell -> dof maps: UGlobalNumbers + UBeginIndex[cell] (27 each), same for P (4 each) */
  int n_cells, n_ulocal, n_plocal;
  const int *u_begin, *u_global, *p_begin, *p_global;
  /* optional cell coloring (colorCell output); NULL -> the library computes it */
  int n_colors; const int *ptr_colors; const int *cell_reorder;
  /* transfer to the next coarser level (ignored on level 0): parent cell on the coarser
     level and child number for every cell of this level, and the per-child prolongation
     matrices QQ[child][k*n_coarse+l] (fine functional k of coarse basis l), i.e.
     TFEDatabase3D::GetProlongationMatrix3D (src/FE/FEDatabase3D.C:1630-1720) */
  const int *parent, *child;
  int n_children;
  const double *qq_u;       /* n_children * n_ulocal * n_ulocal */
  const double *qq_p;       /* n_children * n_plocal * n_plocal */
} pmg_level_desc;

typedef struct pmg_ctx pmg_ctx;

/* ----------------------------------------------------------------------------------- */
/* context and levels                                                                    */
/* ----------------------------------------------------------------------------------- */
void        pmg_params_default(pmg_params *p);
const char *pmg_status_string(int status);

/* Replaces: TNSE_MultiGrid::TNSE_MultiGrid (src/FE/NSE_MultiGrid.C:81-104) plus the
   device set-up TNSE_MultiGrid::InitGPU (src/FE/NSE_MultiGrid.C:105-185). */
int  pmg_ctx_create(int device, const pmg_params *params, pmg_ctx **out);
/* Replaces: TNSE_MultiGrid::~TNSE_MultiGrid / FreeGPU. */
int  pmg_ctx_destroy(pmg_ctx *ctx);
const char *pmg_last_error_string(const pmg_ctx *ctx);
int  pmg_set_params(pmg_ctx *ctx, const pmg_params *params);

/* Replaces: TNSE_MultiGrid::AddLevel (src/FE/NSE_MultiGrid.C:94-102) together with the
   TNSE_MGLevel4 constructor (src/FE/NSE_MGLevel4.C:119-326).  Levels are added coarsest
   first; the returned index equals the reference's level number. */
int  pmg_level_add(pmg_ctx *ctx, const pmg_level_desc *desc, int *level_out);
/* Replaces: re-assembly with a fixed pattern (SystemTNSE3D.C:1400-1415 call sites):
   new matrix values, same structure.  Marks the level's patch factors stale. */
int  pmg_level_update_values(pmg_ctx *ctx, int level, const pmg_level_desc *desc);
/* Value-dependent set-up: batched patch LU of every level >= 1 and the coarse dense LU.
   The reference re-factors inside every sweep (NSE_MGLevel4.C:1400-1609) and every coarse
   visit (NSE_MGLevel4.C:5728-5846); the factors are the same, so they are kept. */
int  pmg_setup(pmg_ctx *ctx);
/* Recompute every value-dependent factor (patch LUs of all levels >= 1, coarse LU) from the
   operator already resident on the device.  The reference repeats this work inside every
   sweep and coarse visit; callers that change values in device memory, or that time a
   complete solve including its set-up, call this instead of re-adding levels. */
int  pmg_refactor(pmg_ctx *ctx);
int  pmg_level_info(const pmg_ctx *ctx, int level, int *n_dof, int *n_colors, int *max_cells_per_color);
/* *shared_block = 1 when the level's velocity rows stream one shared A block (see
   pmg_params.shared_block), 0 when they stream the nine blocks.  No reference counterpart:
   the reference always multiplies all nine blocks (MGComponents3D.C:3038-3209). */
int  pmg_level_form(const pmg_ctx *ctx, int level, int *shared_block);

/* ----------------------------------------------------------------------------------- */
/* level operations (TNSE_MGLevel / TNSE_MGLevel4 virtuals and MGComponents3D.C)         */
/* vectors are n_dof = 3*n_udof + n_pdof long, in memory 'mem'                           */
/* ----------------------------------------------------------------------------------- */
/* Replaces: TNSE_MGLevel4::Defect (src/FE/NSE_MGLevel4.C:465-539): CoupledDefect, IntoL20
   of the pressure constants, Dirichlet rows zeroed, res = ||d||_2 (res may be NULL). */
int  pmg_defect(pmg_ctx *ctx, int level, const double *u, const double *f, double *d,
                double *res, int mem);
/* Replaces: MatVect_NSE4 / CoupledMatVect (src/FE/MGComponents3D.C:2872-3035). */
int  pmg_matvec(pmg_ctx *ctx, int level, const double *x, double *y, int mem);
/* Replaces: Defect_NSE4 (src/FE/MGComponents3D.C:3210-3231): r = b - K x, IntoL20 on p. */
int  pmg_defect_nse4(pmg_ctx *ctx, int level, const double *x, const double *b, double *r,
                     int mem);
/* Replaces: TNSE_MGLevel4::CellVanka (src/FE/NSE_MGLevel4.C:603-1649 colored, smoother 2;
   :1653-2085 Gauss-Seidel, smoother 1), called nsweeps times.  u is updated in place; f
   is the level right-hand side. */
int  pmg_vanka_smooth(pmg_ctx *ctx, int level, double *u, const double *f, int nsweeps,
                      int mem);
/* No counterpart in the reference.  For callers that cannot allocate device memory (code
   under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the FP64
   colored Vanka sweeps of pmg_vanka_smooth on device vectors u and f, with the local
   residuals and the relaxation copy of u in a file-scope device array of 8 MiB per GPU
   instead of library-allocated buffers; the call allocates and frees nothing.  Levels whose
   scratch (max patches per color x 96 doubles, plus n_dof doubles when the global
   relaxation sc_smooth_damp_factor_saddle != 1 is on) exceeds 8 MiB get
   PMG_ERR_UNSUPPORTED; mixed-precision contexts get PMG_ERR_UNSUPPORTED.  Every call on a
   GPU shares that GPU's array: calls that overlap on one GPU race. */
int  pmg_vanka_smooth_noalloc(pmg_ctx *ctx, int level, double *u, const double *f, int nsweeps);
/* Replaces: DefectRestriction for U (3 functions, MGComponents3D.C:996-1240) and P
   (MGComponents3D.C:711-993): f_coarse = R d_fine.  'level' is the fine level. */
int  pmg_restrict_defect(pmg_ctx *ctx, int level, const double *d_fine, double *f_coarse,
                         int mem);
/* Replaces: Prolongate for U (MGComponents3D.C:468-708) and P (MGComponents3D.C:184-466):
   d_fine = P u_coarse.  'level' is the fine level. */
int  pmg_prolongate(pmg_ctx *ctx, int level, const double *u_coarse, double *d_fine,
                    int mem);
/* Replaces: RestrictFunction (src/FE/MGComponents3D.C:1509-1755) for the velocity (3
   components): on every coarse cell the L2 projection onto Q2 of the fine velocity on its
   children, averaged over the cells sharing a coarse dof.  'level' is the fine level;
   the pressure part of u_coarse is left unchanged. */
int  pmg_restrict_function(pmg_ctx *ctx, int level, const double *u_fine, double *u_coarse, int mem);
/* Replaces: TNSE_MultiGrid::RestrictToAllGrids (src/FE/NSE_MultiGrid.C:391-431): the
   finest level's velocity restricted level by level; u_levels[l] (l = 0 .. n_levels-2)
   receives level l's velocity (vectors of that level's n_dof, pressure untouched). */
int  pmg_restrict_to_all_grids(pmg_ctx *ctx, const double *u_finest, double *const *u_levels, int mem);
/* Replaces: TNSE_MGLevel4::CorrectNodes (NSE_MGLevel4.C:542-601): u_D := 0. */
int  pmg_correct_nodes(pmg_ctx *ctx, int level, double *u, int mem);
/* Replaces: TNSE_MultiGrid::SetDirichletNodes (NSE_MultiGrid.C:1781-1802): u_D := f_D. */
int  pmg_set_dirichlet(pmg_ctx *ctx, int level, double *u, const double *f, int mem);
/* Replaces: TNSE_MGLevel::CorrectDefect (NSE_MGLevel.C:202-278): rows >= n_active := 0,
   IntoL20 of the pressure constants. */
int  pmg_correct_defect(pmg_ctx *ctx, int level, double *d, int mem);
/* Replaces: TNSE_MGLevel::Update (NSE_MGLevel.C:121-129): u += beta d. */
int  pmg_update(pmg_ctx *ctx, int level, double *u, const double *d, int mem);
/* Replaces: TNSE_MGLevel4::SolveExact (NSE_MGLevel4.C:5728-5846) on level 0: dense LU
   with pressure dof 0 pinned (kept factors, see pmg_setup). */
int  pmg_coarse_solve(pmg_ctx *ctx, double *u, const double *f, int mem);

/* ----------------------------------------------------------------------------------- */
/* cycle and Krylov drivers                                                              */
/* ----------------------------------------------------------------------------------- */
/* Replaces: TMultiGridIte::Iterate (src/FE/MultiGridIte.C:76-130) with zero start:
   one V-cycle TNSE_MultiGrid::Cycle (NSE_MultiGrid.C:434-1779) on the finest level,
   z = M^{-1} v. */
int  pmg_vcycle(pmg_ctx *ctx, const double *v, double *z, int mem);
/* Replaces: TFgmresIte::Iterate, flexible branch (src/FE/FgmresIte.C:250-677) with the
   GMG preconditioner.  sol is updated in place.  iters gets the reference's return value
   j; final_res the last |s[i+1]| (or restart residual).  Returns PMG_ERR_DIVERGED or
   PMG_ERR_NOT_CONVERGED where the reference exits or prints "FGMRES not converged". */
int  pmg_fgmres_solve(pmg_ctx *ctx, double *sol, const double *rhs, int *iters,
                      double *final_res, int mem);
/* Residual history of the last pmg_fgmres_solve (|s[i+1]| per iteration, plus
   restart residuals); returns the number of entries written (<= cap). */
int  pmg_fgmres_history(const pmg_ctx *ctx, double *out, int cap);

/* ----------------------------------------------------------------------------------- */
/* device re-assembly of the linearised velocity blocks                                  */
/* ----------------------------------------------------------------------------------- */
/* Cell geometry of a level for the re-assembly: 8 corner vertices (x, y, z) per cell in
   the order a + 2b + 4c of the trilinear map (a, b, c = 0/1 along xi, eta, zeta), the map
   TFEDatabase3D applies for the hexahedra of the collection. */
int  pmg_level_geometry(pmg_ctx *ctx, int level, const double *cell_vertices);
/* The velocity of a level vector u ([u1 u2 u3 p]) at the 27 Gauss points of every cell:
   b_quad[(cell * 27 + q) * 3 + c], q = i + 3 j + 9 k for the points (g_i, g_j, g_k),
   g = (-sqrt(3/5), 0, sqrt(3/5)) - the 'param' values (u1old, u2old, u3old) that the
   reference's assembly passes to NSType4Galerkin3D (src/FE/NSE3D_FixPo.C:1979-1981). */
int  pmg_eval_velocity(pmg_ctx *ctx, int level, const double *u, double *b_quad, int mem);
/* Replaces: re-assembly of A11..A33 by NSType4Galerkin3D (src/FE/NSE3D_FixPo.C:1919-2071)
   in the fixed-point loop (src/System/SystemTNSE3D.C:725-795, 1400-1415) on the device:
   A11 = A22 = A33 = nu (grad phi_j, grad phi_i) + ((b . grad) phi_j, phi_i)
   + sigma (phi_j, phi_i) on the active rows (sigma: the mass term of a time step),
   A12..A32 = 0, Dirichlet rows unchanged, with b given at the Gauss points as for
   pmg_eval_velocity.  B and BT keep their values.  Marks the level's factors stale
   (pmg_refactor or pmg_setup before the next solve). */
int  pmg_assemble_oseen(pmg_ctx *ctx, int level, const double *b_quad, double nu, double sigma,
                        int mem);

/* Replaces: the right-hand side of TSystemTNSE3D::AssembleSystMat
   (src/System/SystemTNSE3D.C:1118-1190) for a time step of the theta scheme, scaled by
   1 / (theta tau) (sigma = 1 / (theta tau)) and without source terms:
   rhs = sigma M u_old - A(b) u_old on the active velocity rows, with A(b) the element
   matrices of pmg_assemble_oseen without the mass term (b at the Gauss points as for
   pmg_eval_velocity, usually the velocity u_old itself), matrix-free; rhs = g on the
   Dirichlet rows; 0 on the pressure rows.  With sigma M + A(b) as the system matrix this
   is the Crank-Nicolson step for theta = 1/2.  Needs pmg_level_geometry. */
int  pmg_theta_rhs(pmg_ctx *ctx, int level, const double *u_old, const double *b_quad, double nu, double sigma,
                   const double *g, double *rhs, int mem);

/* ----------------------------------------------------------------------------------- */
/* batched solves of independent problems                                                */
/* ----------------------------------------------------------------------------------- */
/* One independent problem of a batch: its level hierarchy (coarsest first, at least two
   levels, descriptors as for pmg_level_add) and its own INTERNAL_PROJECT_PRESSURE (1 for
   an enclosed flow, 0 when the velocity has a natural boundary). */
typedef struct {
  int n_levels;
  const pmg_level_desc *levels;
  int project_pressure;
} pmg_instance_desc;

/* Replaces: one TSystemNSE3D solve per problem (src/System/SystemNSE3D.C:1199-1235), run
   for many problems of different sizes at once.  The hierarchies are merged level by
   level, aligned at the finest level, into one block-diagonal hierarchy: every smoothing
   sweep, defect and transfer is one launch for all problems, each problem is solved
   directly on its own coarsest level (TNSE_MGLevel4::SolveExact), and each problem keeps
   its own IntoL20 and FGMRES scalars.  The returned context accepts pmg_setup,
   pmg_refactor, pmg_set_profiling, pmg_timers_*, pmg_synchronize and pmg_ctx_destroy. */
int  pmg_batch_create(int device, const pmg_params *params, int n_instances,
                      const pmg_instance_desc *instances, pmg_ctx **out);
/* Replaces: TFgmresIte::Iterate (src/FE/FgmresIte.C:250-677) once per problem, all problems
   stepped together.  sol[i] / rhs[i] are problem i's vectors in its own layout (n_dof of
   its finest level).  iters[i], final_res[i] and status[i] are what pmg_fgmres_solve
   would report for problem i alone; the call returns PMG_OK when every status is PMG_OK,
   else the first failing status. */
int  pmg_batch_fgmres_solve(pmg_ctx *ctx, double *const *sol, const double *const *rhs,
                            int *iters, double *final_res, int *status, int mem);
/* number of problems of a batch context (0 for a single-problem context) */
int  pmg_batch_size(const pmg_ctx *ctx);

/* ----------------------------------------------------------------------------------- */
/* level dumps (host code)                                                               */
/* ----------------------------------------------------------------------------------- */
/* The on-disk form of one level: a directory with manifest.txt ("key value" lines:
   n_udof n_active hanging_bound n_dirichlet n_pdof a_nnz b_nnz bt_nnz n_cells n_ulocal
   n_plocal n_children n_colors) and one raw native-endian file per array, doubles as
   ParMooN's SaveData writes them (src/General/MainUtilities.C:3360-3440) and int arrays as
   32-bit ints: a_rowptr.i32 a_kcol.i32 A11.f64 .. A33.f64 (the TSquareStructure3D arrays
   and Entries, include/FE/Matrix.h:91-129), b_rowptr.i32 b_kcol.i32 B1.f64 B2.f64 B3.f64,
   bt_rowptr.i32 bt_kcol.i32 B1T.f64 B2T.f64 B3T.f64 (TStructure3D), u_begin.i32 (BeginIndex,
   n_cells + 1 entries) u_global.i32 (GlobalNumbers), p_begin.i32 p_global.i32, and on a
   refined level parent.i32 child.i32 qq_u.f64 qq_p.f64, with a coloring ptr_colors.i32
   cell_reorder.i32.  pmg_level_read fills *out with arrays owned by *handle (release
   with pmg_level_read_free after pmg_level_add); malformed dumps give
   PMG_ERR_INVALID_ARGUMENT.  No reference counterpart beyond SaveData / ReadData. */
int  pmg_level_read(const char *dir, pmg_level_desc *out, void **handle);
int  pmg_level_read_free(void *handle);
int  pmg_level_write(const char *dir, const pmg_level_desc *desc);
/* Replaces: ReadData / SaveData (src/General/MainUtilities.C:3390-3440) for one vector. */
int  pmg_vector_read(const char *path, double *x, long n);
int  pmg_vector_write(const char *path, const double *x, long n);

/* ----------------------------------------------------------------------------------- */
/* coloring                                                                              */
/* ----------------------------------------------------------------------------------- */
/* Replaces: TParFEMapper3D::colorCell (src/Parallel/ParFEMapper3D.C:750-1084): cells are
   adjacent iff they share a velocity dof; greedy first-fit in cell-index order; output
   cells grouped by color, increasing index inside a color.  Linear time.  ptr_colors must
   hold n_cells+1 ints, cell_reorder n_cells ints. */
int  pmg_color_cells(int n_cells, int n_ulocal, const int *u_begin, const int *u_global,
                     int n_udof, int *n_colors, int *ptr_colors, int *cell_reorder);

/* Replaces: the cell order of TNSE_MGLevel4::CellVanka, Gauss-Seidel variant
   (src/FE/NSE_MGLevel4.C:1653-2085, smoother 1): the cells in level sets such that every
   pair of cells whose patch rows read each other's dofs stays in collection order and
   each set holds only cells that do not; a colored sweep over these sets is the
   Gauss-Seidel sweep.  u_global / p_global: 27 / 4 dofs per cell; the CSR patterns of
   A, BT and B as in pmg_level_desc.  ptr_sets must hold n_cells+1 ints. */
int  pmg_sequential_schedule(int n_cells, const int *u_global, const int *p_global, int n_udof, int n_pdof,
                             const int *a_rowptr, const int *a_kcol, const int *bt_rowptr, const int *bt_kcol,
                             const int *b_rowptr, const int *b_kcol, int *n_sets, int *ptr_sets,
                             int *cell_reorder);

/* Replaces: the pressure-dof order of TNSE_MGLevel4::NodalVanka (src/FE/NSE_MGLevel4.C:
   4098-4549, smoothers 3 and 4, DOF_Reorder = 0): patch i = pressure dof i and the velocity
   dofs of B row i; the dofs in level sets as in pmg_sequential_schedule (coupled patches in
   order, each set uncoupled), ascending inside a set.  ptr_sets must hold n_pdof+1 ints,
   pdof_order n_pdof ints. */
int  pmg_nodal_schedule(int n_udof, int n_pdof, const int *a_rowptr, const int *a_kcol, const int *bt_rowptr,
                        const int *bt_kcol, const int *b_rowptr, const int *b_kcol, int *n_sets, int *ptr_sets,
                        int *pdof_order);

/* ----------------------------------------------------------------------------------- */
/* device memory helpers for callers without a device allocator                          */
/* ----------------------------------------------------------------------------------- */
int  pmg_device_alloc(pmg_ctx *ctx, size_t bytes, void **ptr);
int  pmg_device_free(pmg_ctx *ctx, void *ptr);
int  pmg_memcpy(pmg_ctx *ctx, void *dst, const void *src, size_t bytes, int kind); /* 0 H2D 1 D2H 2 D2D */
int  pmg_synchronize(pmg_ctx *ctx);

/* Per-kernel timing, measured with HIP events on the library's stream around every launch
   while profiling is on (pmg_set_profiling; it synchronizes after each launch, so it is
   meant for a separate measurement pass).  'bytes' accumulates the algorithmic HBM bytes
   of the timed launches (every operator value, index, factor and vector entry the
   algorithm needs read or written once; gathers of x counted as cache hits). */
typedef struct {
  double ms; long long calls; double bytes; double flops;
} pmg_kstat;
typedef struct {
  pmg_kstat defect;        /* k_type4 in the V-cycle and the level defect                */
  pmg_kstat matvec;        /* k_type4 as FGMRES operator / Defect_NSE4                    */
  pmg_kstat vanka_res;     /* per-color local residuals                                   */
  pmg_kstat vanka_solve;   /* per-color stored-factor local solves + update               */
  pmg_kstat transfer;      /* restriction + prolongation SpMVs                            */
  pmg_kstat coarse_solve;  /* coarse triangular sweeps                                    */
  pmg_kstat patch_lu;      /* batched patch LU (set-up)                                   */
  pmg_kstat coarse_lu;     /* coarse dense LU (set-up)                                    */
  double sweep_ms; long long sweep_calls;   /* whole colored sweeps                       */
  pmg_kstat vanka_res_finest;  /* the per-color local residuals of the finest level alone */
} pmg_timers;
/* Launch-shape autotuner (no reference counterpart; the reference's kernels have fixed
   shapes): after pmg_setup, times the candidate launch shapes of every smoothed level's
   kernels on the device, reps launches each, and keeps the fastest per level: work items
   per workgroup of the Vanka residual (16, 32, 64), whether colors of >= 512 cells run in
   per-XCD groups, operator rows per workgroup of the type-4 defect / matvec (32 .. 256).
   The shapes change how work is distributed, not what is computed (only the summation
   order of the pressure means differs).  Uses the levels' own work vectors.
   pmg_level_tuning returns {residual items per workgroup FP64, FP32; XCD groups FP64,
   FP32; operator rows per workgroup FP64, FP32} (0: built-in default). */
int  pmg_autotune(pmg_ctx *ctx, int reps);
int  pmg_level_tuning(pmg_ctx *ctx, int level, int *out6);
/* sets a level's launch shapes (the six values of pmg_level_tuning, e.g. saved from an
   earlier pmg_autotune); 0 entries mean the built-in default */
int  pmg_set_level_tuning(pmg_ctx *ctx, int level, const int *in6);
int  pmg_timers_get(const pmg_ctx *ctx, pmg_timers *t);
int  pmg_timers_reset(pmg_ctx *ctx);
int  pmg_set_profiling(pmg_ctx *ctx, int on);

#ifdef __cplusplus
}
#endif
#endif /* PMG_H */
