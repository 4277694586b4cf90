// pmg_internal.hpp - device data structures and kernel launchers of libpmg (gfx950).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "pmg.h"

namespace pmg {

constexpr int NU_LOC = 27;          // Q2 hex velocity dofs per cell (per component)
constexpr int NP_LOC = 4;           // P1disc hex pressure dofs per cell
constexpr int NLOC = 3 * NU_LOC + NP_LOC;  // 85: local Vanka system size
constexpr int LU_STRIDE = 7232;     // 85*85 = 7225 padded to a multiple of 16 entries
// packed patch factors: L (strictly lower, 3570 entries) column by column, then U (upper
// with diagonal, 3655 entries) column by column starting at a 16-byte aligned offset
constexpr int LU_L = NLOC * (NLOC - 1) / 2;   // 3570
constexpr int LU_U = NLOC * (NLOC + 1) / 2;   // 3655
constexpr int LU_UOFF = 3576;                 // LU_L rounded up to 8 entries
constexpr int LU_PAD = 512;                   // tail entries after the last patch (vector overshoot)
__host__ __device__ constexpr int lu_loff(int k) { return k * (NLOC - 1) - k * (k - 1) / 2; }  // L column k
__host__ __device__ constexpr int lu_uoff(int k) { return k * (k + 1) / 2; }                   // U column k
static_assert(LU_UOFF + LU_U <= LU_STRIDE, "packed factors exceed the patch stride");
// explicit patch inverse (mixed precision): 85 rows of MINV_LD entries (85 used), row major
constexpr int MINV_LD = 88;
constexpr int MINV_STRIDE = 7488;   // 85 * 88 = 7480 padded to a multiple of 16 entries
static_assert(NLOC * MINV_LD <= MINV_STRIDE, "inverse exceeds the patch stride");
// block form of a patch with the shared velocity block (mixed precision): the patch matrix
// is [A 0 0 BT1; 0 A 0 BT2; 0 0 A BT3; B1 B2 B3 0] with one 27 x 27 block A, so its solve
// needs A^{-1} (729), W_c = A^{-1} BT_c (3 x 27 x 4), B_c (3 x 4 x 27) and the inverse of the
// 4 x 4 Schur complement S = -sum_c B_c W_c (16): 1393 entries per patch
// block-form patch data: A^{-1} (27 x 27), W_c = A^{-1} BT_c (c = 1..3, 27 x 4 each) and
// P1 = [-S^{-1} B_1 A^{-1}, -S^{-1} B_2 A^{-1}, -S^{-1} B_3 A^{-1} | S^{-1}] (4 x 85)
constexpr int SCH_AINV = 0, SCH_W = 729, SCH_P1 = SCH_W + 324, SCH_END = SCH_P1 + 4 * 85;
constexpr int SCH_STRIDE = 1408;    // 1393 padded to a multiple of 16 entries
static_assert(SCH_END <= SCH_STRIDE && SCH_STRIDE <= MINV_STRIDE, "block-form patch stride");
constexpr int RLOC_STRIDE = 96;     // local residual slot (85 used)
constexpr int POS_DOFS_STRIDE = 32; // dofs of the patch at a color position (27 + 4 used)
constexpr int PERM_STRIDE = 96;     // uint8 inverse row permutation per patch
constexpr int RED_BLOCKS = 4096;    // max blocks of the kernels with a last-block reduction
constexpr int ROW_PAD = 64;         // zeroed tail entries of the streamed operator arrays
// per-cell gather map of the patch matrix (ints): offsets of the A entries (j, m) into the
// row-blocked A values (block 0; -1: no entry), row lengths of the 27 velocity rows
// (negative: Dirichlet row), BT offsets (j, pressure m), BT row lengths, B offsets
// (pressure j, m), B row lengths
constexpr int PM_A = 0;
constexpr int PM_NA = PM_A + NU_LOC * NU_LOC;    // 729
constexpr int PM_BT = PM_NA + NU_LOC;            // 756
constexpr int PM_MT = PM_BT + NU_LOC * NP_LOC;   // 864
constexpr int PM_B = PM_MT + NU_LOC;             // 891
constexpr int PM_NB = PM_B + NP_LOC * NU_LOC;    // 999
constexpr int PM_STRIDE = 1008;
// FP32 patch matrices kept for the refinement steps of the mixed local solves: the block
// form (A 27 x 27 row-major, then BT_c[j][k] and B_c[k][j] of the three components) or the
// full 85 x 85 matrix (row-major), per patch in color order
constexpr int KREF_BLK = 1380;      // 729 + 324 + 324 = 1377, padded to 4 entries
constexpr int KREF_FULL = 7232;     // 85 * 85 = 7225, padded
constexpr int KREF_BT = 729, KREF_B = 729 + 324;
// nodal Vanka: velocity dofs of one B row (27 for Q2/P1disc), upper bound
constexpr int NODAL_MAXU = 28;
// colors of at least this many cells run their residual items in 8 per-XCD groups
constexpr int XCD_GROUP_MIN = 512;
static_assert(PM_NB + NP_LOC <= PM_STRIDE, "patch map stride");

// Type-4 operator.  Two value layouts:
//  * structure of arrays, FP64 only, the reference layout (A block b of entry k at
//    a[b*a_nnz + k], B/BT block c at b[c*b_nnz + k]); read by the set-up kernels;
//  * row-blocked, FP64 and FP32: the values of one CSR row are stored together, block by
//    block: A row r (entries k0..k1-1, n = k1-k0) at av[9*k0 + b*n + (k-k0)], B/BT rows at
//    bv[3*k0 + c*n + (k-k0)].  A row is one contiguous region (whole cache lines but at its
//    two ends) and 16 lanes reading 16 consecutive entries of one block coalesce.  Read by
//    the streaming kernels (defect, matvec, Vanka residual).
template <typename T>
struct OpView {
  int n_udof, n_active, n_pdof;
  const int *a_rowptr, *a_kcol;
  const T *a;
  long a_nnz;
  const int *b_rowptr, *b_kcol;
  const T *b;
  long b_nnz;
  const int *bt_rowptr, *bt_kcol;
  const T *bt;
  long bt_nnz;
  const T *av, *bv, *btv;   // row-blocked values
  // the shared velocity block in CSR order when A11 = A22 = A33 and A_rc = 0 (r != c) on
  // the active rows (the Galerkin Oseen form); nullptr: the nine blocks differ
  const T *a1;
  int rpb;                  // operator rows per workgroup of the type-4 kernels (0: default)
  // (batched levels with finished problems) the type-4 kernels run these rows only: a
  // velocity dof r < n_udof, or n_udof + a pressure dof; nullptr: every row
  const int *rows = nullptr;
  int n_rows = 0;
  // (FP64, shared velocity block) an interleaved copy of the input's velocity, (x1, x2, x3,
  // 0) per dof, that the type-4 kernels gather instead of the three component arrays
  const T *xa = nullptr;
};

// Streaming kernels split their n items (rows, residual work items) into one contiguous
// chunk per workgroup, chunk g = [n*g/G, n*(g+1)/G).  The hardware spreads consecutive
// workgroups round-robin over the 8 XCDs, so every XCD gets an even mix of the cheap and
// expensive rows (Dirichlet, pressure and full velocity rows differ up to ~50x in bytes).
// (Measured: chunks of equal bytes, and one contiguous eighth of the rows per XCD for L2
// locality of the gathers, were both slower at 32^3.)
inline int chunk_grid(long n_items, long per_block, int cap) {
  long g = (n_items + per_block - 1) / per_block;
  if (g < 8) g = 8;
  if (g > cap) g = cap;
  return int(g);
}
__device__ inline void chunk_of_block(long n, int &i0, int &i1) {
  i0 = int(n * blockIdx.x / gridDim.x);
  i1 = int(n * (blockIdx.x + 1) / gridDim.x);
}

struct Csr {
  int n_rows = 0, n_cols = 0, nnz = 0;
  int *rowptr = nullptr, *col = nullptr;
  double *val = nullptr;
  float *val32 = nullptr;
};

struct CbJob;

struct DevLevel {
  int n_udof = 0, n_active = 0, hanging_bound = 0, n_dirichlet = 0, n_pdof = 0, n_dof = 0;
  int n_cells = 0;
  long a_nnz = 0, b_nnz = 0, bt_nnz = 0;
  long a_act_entries = 0;            // entries of the active rows, a_rowptr[n_active]
  // operator, FP64 and (mixed precision) FP32 copies
  int *a_rowptr = nullptr, *a_kcol = nullptr, *b_rowptr = nullptr, *b_kcol = nullptr;
  int *bt_rowptr = nullptr, *bt_kcol = nullptr;
  double *a = nullptr, *b = nullptr, *bt = nullptr;
  double *av64 = nullptr, *bv64 = nullptr, *btv64 = nullptr;   // row-blocked values
  float *av32 = nullptr, *bv32 = nullptr, *btv32 = nullptr;
  // shared velocity block (OpView::a1): set by the value check after every change of the
  // values; FP64 reads block 0 of a, FP32 its own copy
  bool a_shared = false;
  float *a1_32 = nullptr;
  // per color, the (slot * 32 + local row) work items of the residual kernel sorted by the
  // global row they read (velocity rows, then pressure rows)
  int *res_items = nullptr;
  int *res_rows = nullptr;           // global row of every work item
  int *res_desc = nullptr;           // its descriptor (row, A/B start, BT start, lengths)
  // colors split into 8 groups of consecutive cells, one per XCD (host flag per color;
  // device: 9 item offsets per color, relative to the color's first item)
  std::vector<char> h_xcd_grouped;
  int *xcd_goff = nullptr;
  // cells
  int *u_global = nullptr;           // n_cells * 27 (UGlobalNumbers + UBeginIndex[cell])
  int *p_global = nullptr;           // n_cells * 4
  // coloring (host copy of the color pointers; device cell list in color order)
  int n_colors = 0, max_cells_per_color = 0;
  std::vector<int> ptr_colors;
  int *cell_reorder = nullptr;
  // the patch dofs at every color position (velocity dofs 0..26, pressure dofs 27..30):
  // the local solves read them without the cell -> dof map chain
  int *pos_dofs = nullptr;
  // patch factors in color order (position = index into cell_reorder)
  double *lu = nullptr;
  float *lu32 = nullptr;
  uint16_t *lu16 = nullptr;          // bf16 factors
  // or (mixed precision, pmg_params.patch_inverse) the explicit inverses, MINV_STRIDE each
  float *minv32 = nullptr;
  uint16_t *minv16 = nullptr;
  // FP32 sweeps with inverses and the shared block: interleaved copy (u1, u2, u3, 0) of the
  // velocity the residual gathers, built at the sweep start and kept by the local updates
  float *ua32 = nullptr;
  // the patch inverses are stored in the block form (set at every factorization: shared
  // velocity block present)
  bool schur = false;
  uint8_t *invperm = nullptr;        // n_cells * PERM_STRIDE
  int *pmap = nullptr;               // n_cells * PM_STRIDE patch gather map
  int *patch_info = nullptr;         // n_cells
  float *kref = nullptr;             // FP32 patch matrices for refinement steps (KREF_*)
  size_t kref_len = 0;
  bool factors_ready = false;
  // transfers to the next coarser level (this level = fine)
  Csr pu, ru, pp, rp;                // P_U (fine rows), R_U = P_U^T, P_P, R_P
  Csr rf;                            // RestrictFunction of the velocity: coarse rows, fine columns
  // nodal Vanka (smoothers 3, 4): level sets of the pressure dofs (host pointers, device
  // dof list), the singular-patch flag and the algorithmic bytes of every set's launch
  bool nodal = false;
  std::vector<int> nodal_ptr;
  std::vector<double> nodal_bytes;
  int *nodal_order = nullptr;
  int *nodal_info = nullptr;
  // MG vectors: sol, rhs, aux0 (defect), aux1 (aux) as the reference level owns them
  double *sol = nullptr, *rhs = nullptr, *aux0 = nullptr, *aux1 = nullptr;
  float *sol32 = nullptr, *rhs32 = nullptr, *aux0_32 = nullptr, *aux1_32 = nullptr;
  double *rloc = nullptr;            // max_cells_per_color * RLOC_STRIDE
  float *rloc32 = nullptr;
  // stored inverses: u at each patch's dofs as the residual read it (natural local order,
  // RLOC_STRIDE per slot), so the local solve does not gather it again
  float *uloc32 = nullptr;
  // host copies kept for value updates / coarse assembly
  std::vector<int> h_a_rowptr, h_a_kcol, h_b_rowptr, h_b_kcol, h_bt_rowptr, h_bt_kcol;
  // algorithmic bytes (FP64 operator) of the per-color residual launches and of one defect
  // [0]: FP64 values and vectors, [1]: FP32
  // [s][p]: s = 1 with the shared velocity block, p = 0 FP64, 1 FP32
  std::vector<double> res_bytes[2][2];
  double defect_bytes[2][2] = {{0, 0}, {0, 0}}, matvec_bytes[2] = {0, 0};
  double transfer_bytes = 0;
  // merged (batched) levels: the problems solved directly on this level, and the
  // (offset, length) pressure blocks that get their own IntoL20
  std::vector<int> coarse_insts;
  int *l20_seg = nullptr;
  int n_l20_seg = 0;
  // merged levels with finished problems left out of the sweeps: per color, the residual
  // work items and the patch slots of the problems still iterating
  bool compacted = false;
  std::vector<int> h_res_items;      // host copies of res_items, res_rows
  std::vector<int> h_res_rows;
  std::vector<int> h_pos_problem;    // problem of the cell at each color position
  std::vector<long> act_items_off, act_slots_off;   // per color (n_colors + 1)
  int *act_items = nullptr, *act_rows = nullptr, *act_slots = nullptr, *act_desc = nullptr;
  // the problem of every velocity and pressure dof of a merged level, and (act_lists) the
  // dofs of the unfinished problems in ascending order: velocity dofs act_u, pressure dofs
  // act_p, both as type-4 operator rows act_oprows (velocity dof, n_udof + pressure dof)
  std::vector<int> h_udof_problem, h_pdof_problem;
  // runs of consecutive entries of one problem (first, end, problem): velocity dofs,
  // pressure dofs and, per color (h_irun_off, n_colors + 1), the residual work items
  struct Run {
    long b, e;
    int prob;
  };
  std::vector<Run> h_urun, h_prun, h_irun;
  std::vector<long> h_irun_off;
  // the block-form coarse visits of the problems whose coarsest level this is, batched
  // (device job table, the problems it covers, the largest grids)
  CbJob *cb_jobs = nullptr;
  int n_cb_jobs = 0;
  std::vector<int> cb_job_prob;
  int cb_grid[3] = {0, 0, 0};
  bool act_lists = false;
  int *act_u = nullptr, *act_p = nullptr, *act_oprows = nullptr;
  int n_act_u = 0, n_act_p = 0;
  double *geom = nullptr;            // n_cells * 8 corners * 3 (device re-assembly)

  // launch shapes chosen by pmg_autotune (0: the built-in defaults): residual work items per
  // workgroup, whether large colors run in their per-XCD groups, operator rows per
  // workgroup of the type-4 kernels, per precision (0: FP64, 1: FP32)
  int tune_res_pb[2] = {0, 0};
  int tune_res_xcd[2] = {1, 1};
  int tune_op_rpb[2] = {0, 0};

  OpView<double> op64() const {
    OpView<double> o{n_udof, n_active, n_pdof, a_rowptr, a_kcol, a, a_nnz,
                     b_rowptr, b_kcol, b, b_nnz, bt_rowptr, bt_kcol, bt, bt_nnz,
                     av64, bv64, btv64, a_shared ? a : nullptr, tune_op_rpb[0]};
    if (act_lists) {
      o.rows = act_oprows;
      o.n_rows = n_act_u + n_act_p;
    }
    return o;
  }
  OpView<float> op32() const {
    OpView<float> o{n_udof, n_active, n_pdof, a_rowptr, a_kcol, nullptr, a_nnz,
                    b_rowptr, b_kcol, nullptr, b_nnz, bt_rowptr, bt_kcol, nullptr, bt_nnz,
                    av32, bv32, btv32, a_shared ? a1_32 : nullptr, tune_op_rpb[1]};
    if (act_lists) {
      o.rows = act_oprows;
      o.n_rows = n_act_u + n_act_p;
    }
    return o;
  }
};

// Block form of the reduced coarse system when the level has the shared velocity block
// (mixed-precision forms): m active velocity dofs per component, q free pressure dofs.
// Row-major FP64 arrays: A (m x m, then its LU), X = A^{-1} [I | BT1 BT2 BT3]
// (m x (m + 3q): A^{-1} then W_b), B = [B1 B2 B3] (q x 3m), V = [B_b A^{-1}] (q x 3m), S
// (q x q, then its LU), P1 = [-S^{-1} V | S^{-1}] (q x (3m + q)); work vectors rr (reduced
// right-hand side, natural order), y (3m), p (q); sigA / sigS the row permutations of the
// LUs, iota the identity permutation, info the two zero-pivot flags.
struct CoarseBlock {
  int m = 0, q = 0, pin = 0;
  double *A = nullptr, *X = nullptr, *B = nullptr, *V = nullptr, *S = nullptr, *P1 = nullptr;
  double *rr = nullptr, *y = nullptr, *p = nullptr;
  int *sigA = nullptr, *sigS = nullptr, *iota = nullptr, *info = nullptr;
};

// Coarse level: the dense system of SolveExact with the Dirichlet rows (identity) and the
// pinned pressure row eliminated; n = 3*na + np - pin unknowns remain.
struct Coarse {
  int n = 0;
  int N = 0, na = 0, np = 0, pin = 0;   // velocity dofs, active velocity dofs, pressure dofs, pinned?
  double *x = nullptr;     // reduced solution
  double *a = nullptr;     // dense row-major matrix, factored in place
  double *f = nullptr;     // the LU factors column-major (read by the triangular sweeps)
  int *ipiv = nullptr;     // device
  int *sigma = nullptr;    // b'[i] = b[sigma[i]]
  double *ubuf = nullptr;  // U block row of the last LU panel step (n * 16)
  int *swp = nullptr;      // row interchanges of the last LU panel (rows, sources, count)
  double *b = nullptr;     // work vectors of the two sweeps
  double *y = nullptr;
  int *info = nullptr;
  // triangular sweeps: generation + finish counter per sweep (4), per-row-block flags of the
  // forward then the backward sweep (2 * ceil(n/64)), wait-limit error flag
  unsigned *sync = nullptr, *flags = nullptr;
  int *err = nullptr;
  // the factorization runs on the second stream; the main stream waits for ev_done only
  // before the first coarse solve, and the pivot check happens at the next host sync
  hipEvent_t ev_start = nullptr, ev_done = nullptr;
  bool pending = false, info_pending = false;
  bool ready = false;
  // block form (mixed precision with the shared velocity block): explicit FP64 inverses of
  // the blocks instead of the LU of the whole reduced system
  bool block = false;
  CoarseBlock cb;
};

// reduction scratch: partial sums + a completion counter per reduction slot
struct RedScratch {
  double *partials = nullptr;   // RED_BLOCKS * 4
  unsigned *counter = nullptr;  // one per slot
  double *result = nullptr;     // result slots
};

// ------------------------------------------------------------------ kernel launchers
// type-4 operator: mode 0 y = K x, 1 y = b - K x (Defect_NSE4), 2 MG defect (Dirichlet rows 0)
// project: IntoL20 of the pressure constants; res (device, may be null): ||y||_2
// xa[4 i + c] = x[c N + i] (c < 3), xa[4 i + 3] = 0 for the n velocity dofs
void launch_interleave3(double *xa, const double *x, int n, hipStream_t s);
template <typename T>
void launch_type4(const OpView<T> &op, int mode, const T *x, const T *b, T *y, bool project,
                  double *res, RedScratch &red, hipStream_t s);

// Vanka
// patch LU of every cell of a level into the level's factor array: FP64 (lu), or, in the
// mixed path, computed in FP32 from the FP32 operator copy into lu32 or bf16 lu16
void launch_patch_lu(const DevLevel &L, int *info, hipStream_t s);
// mixed precision, explicit inverses: FP32 Gauss-Jordan into minv32 or bf16 minv16
void launch_patch_inv(const DevLevel &L, int *info, hipStream_t s, float *kref = nullptr);
// nref > 0: that many refinement steps with the FP32 patch matrices kref (KREF_FULL per
// patch) after the first product (residual permuted by invperm like the local residual)
// pos_dofs: the level's patch dofs by color position; uloc: u at the patch dofs as the
// color's residual launch wrote it (natural local order, RLOC_STRIDE per slot)
template <typename F>
void launch_vanka_apply(const F *minv, const int *pos_dofs, int ncells, int pos0, const int *slots,
                        const float *rloc, const float *uloc, int n_udof, float *u, float *ua, float damp,
                        const float *kref, const uint8_t *invperm, int nref, hipStream_t s);
// block form (shared velocity block): set-up into minv32 / minv16 with SCH_STRIDE per patch,
// and its local solves
void launch_patch_schur(const DevLevel &L, int *info, hipStream_t s, float *kref = nullptr);
template <typename F>
void launch_vanka_apply_schur(const F *sch, const int *pos_dofs, int ncells, int pos0, const int *slots,
                              const float *rloc, const float *uloc, int n_udof, float *u, float *ua, float damp,
                              const float *kref, int nref, hipStream_t s, bool ua_only = false);
// the velocity of the interleaved copy ua back into u (u_c[i] = ua[4 i + c])
void launch_ua_to_u(const float *ua, float *u, int N, hipStream_t s);
template <typename T>
// n_items (slot * 32 + local row) work items of one color starting at position pos0, and
// the global row of each (velocity or pressure dof)
// xg (may be null): the 9 item offsets of the color's 8 XCD groups; per_block: work items
// per workgroup (0: 16 below 65,536 items, 32 above); uloc (may be null): receives u at
// every patch dof (natural local order) for the stored-inverse local solves
void launch_vanka_residual(const OpView<T> &op, long n_items, int pos0, const uint8_t *invperm,
                           const int *items, const int *desc, const int *xg, const T *u, const float *ua,
                           const T *f, T *rloc, T *uloc, hipStream_t s, int per_block = 0);
// zero start of a level: u = 0 except the velocity Dirichlet slices [start, start + count)
// from f, and (ua non-null) the interleaved copy of that velocity; N velocity, NP pressure
// dofs
template <typename T>
void launch_start_zero(T *u, const T *f, int N, int NP, int start, int count, float *ua, hipStream_t s);
// FP32 sweep start: Dirichlet slices of u from f, ua[4 i + c] = u_c(i) (c < 3)
void launch_sweep_start(float *u, const float *f, int N, int start, int count, float *ua, hipStream_t s);
// descriptors of residual work items (4 ints each) from their items and global rows
void launch_res_desc(const DevLevel &L, const int *items, const int *rows, long n, int *desc, hipStream_t s);
// row-blocked operator values from the structure-of-arrays FP64 values
template <typename T>
void launch_row_blocked(const DevLevel &L, T *av, T *bv, T *btv, hipStream_t s);
// true when the level's A blocks have the shared form of OpView::a1 (part: 256 doubles of
// device scratch; synchronizes the stream)
bool shared_velocity_block(const DevLevel &L, double *part, hipStream_t s);
template <typename T, typename F>
// ncells patches of one color: slots (may be null: 0..ncells-1) are their positions
// relative to pos0 and their local residual slots
void launch_vanka_solve(const F *lu, const int *cells, int ncells, int pos0, const int *slots, const T *rloc,
                        const int *u_global, const int *p_global, int n_udof, T *u, T damp,
                        hipStream_t s);

// vectors
template <typename T>
void launch_fill(T *x, T v, long n, hipStream_t s);
template <typename T>
void launch_copy(T *dst, const T *src, long n, hipStream_t s);
template <typename T>
void launch_axpy(T *y, T a, const T *x, long n, hipStream_t s);         // y += a x
template <typename T>
void launch_relax(T *u, const T *uold, T alpha, long n, hipStream_t s); // u = uold + a (u-uold)
void launch_d2f(float *dst, const double *src, long n, hipStream_t s);
void launch_f2d(double *dst, const float *src, long n, hipStream_t s);
template <typename T>
void launch_dirichlet_slices(T *u, const T *f, int n_udof, int start, int count, hipStream_t s);
template <typename T>
void launch_l20(T *p, int n_pdof, RedScratch &red, hipStream_t s);

// transfers: y[c*ny + r] = sum_k val[k] x[c*nx + col[k]], c < ncomp, for every row r, or
// (rows non-null) for the n_rows rows listed; rows r >= zero_from get 0
template <typename T>
void launch_csr_spmv(const Csr &m, int ncomp, const T *x, long x_stride, T *y, long y_stride,
                     hipStream_t s, const int *rows = nullptr, int n_rows = 0, int zero_from = 0x7fffffff);

// the prolongation of the coarse correction xc with the update of the fine u in one pass:
// u += a P xc (velocity rows of pu, pressure rows of pp), the velocity Dirichlet rows
// [ds, ds + dc) set from f, and (ua non-null) the interleaved velocity copy rebuilt; urows /
// prows (may be null) restrict it to the listed rows
template <typename T>
void launch_prolong_update(const Csr &pu, const Csr &pp, const T *xc, long nxc, T *u, long nu, T a, const T *f,
                           int ds, int dc, float *ua, hipStream_t s, const int *urows = nullptr, int n_urows = 0,
                           const int *prows = nullptr, int n_prows = 0);

// dense coarse LU / solve
void launch_coarse_build(const DevLevel &L, Coarse &C, hipStream_t s);
// reduced right-hand side b[i] = (f_I - K_ID x_D)[sigma[i]] of the coarse system (n unknowns)
void launch_coarse_rhs(const OpView<double> &op, const double *f, const int *sigma, double *b, int n, int pin,
                       hipStream_t s);
// block form (k_coarse.hip): sizes it supports, dense blocks from the level operator, set-up
// (asynchronous; cb.info gets the zero-pivot flags of the two LUs) and one visit from the
// reduced right-hand side rr into the level vector u
bool coarse_block_fits(int m, int q);
void launch_cb_build(const DevLevel &L, CoarseBlock &B, hipStream_t s);
void coarse_block_factor(CoarseBlock &B, hipStream_t s);
void coarse_block_solve(CoarseBlock &B, const OpView<double> &op, const double *rr, const double *f, double *u,
                        hipStream_t s);
// one problem's block-form coarse visit inside a batched launch (independent problems merged
// into one level): its own coarse operator and stored data, its work vectors, and the map
// of its coarse dofs into the merged level's vectors
struct CbJob {
  OpView<double> op;
  const double *X, *P1;
  double *rr, *y, *p;
  const int *cidx;
  int m, q, pin;
};
struct CbGrid {
  int rhs, phase1, phase2;   // workgroups per problem of the three launches
};
CbGrid coarse_block_grid(const CbJob &J);
// the visits of n_jobs problems, three launches in all; f and u are the merged level's
// right-hand side and solution; g holds the largest grids of the jobs
template <typename T>
void launch_coarse_block_batch(const CbJob *jobs, int n_jobs, const CbGrid &g, const T *f, T *u, hipStream_t s);
void coarse_lu_factor(Coarse &C, hipStream_t s);   // asynchronous; C.info gets the dgetrf info
void coarse_lu_solve(Coarse &C, const OpView<double> &op, const double *f, double *u, hipStream_t s);

// BLAS-1 for FGMRES (deterministic two-stage reductions)
void launch_dot(const double *x, const double *y, long n, double *out, RedScratch &red, hipStream_t s);
// w -= h[0] * v ; then out = dot(w, vnext) (vnext may be null)
void launch_axpy_dev_dot(double *w, const double *h, const double *v, const double *vnext, long n,
                         double *out, RedScratch &red, hipStream_t s);
void launch_scale_dev(double *y, const double *x, const double *den, long n, hipStream_t s);  // y = x/den
// x += (sum_j coef[j] z_j), z_j = zbase + j*stride  (UpdateGmresIterate + Daxpy)
void launch_lincomb(double *x, const double *zbase, long stride, const double *coef, int m, long n,
                    hipStream_t s);
void launch_scal(double *y, const double *x, double a, long n, hipStream_t s);  // y = a x

// device re-assembly (k_assemble.hip): A11 = A22 = A33 of the active rows from the element
// matrices with convection b given at the 27 Gauss points of every cell (bq[cell][q][3]);
// the Q2 velocity u of a level evaluated at those points
void launch_assemble_oseen(const DevLevel &L, const double *geom, const double *bq, double nu, double sigma,
                           hipStream_t s);
void launch_eval_velocity(const DevLevel &L, const double *u, double *bq, hipStream_t s);
// theta-scheme right-hand side: rhs = sigma M u - A(b) u on the active velocity rows (A: the
// same element matrices without the mass term), g on the Dirichlet rows, 0 on the pressure
// TNSE_MGLevel4::NodalVanka on the pressure dofs pdofs[0 .. n) of one level set: residual,
// local solve (full: LU of the 3 N_U + 1 system; else the diagonal velocity block) and
// damped update per patch; a singular patch sets *info = max(*info, dof + 1)
void launch_nodal_vanka(const OpView<double> &op, bool full, const int *pdofs, int n, const double *f, double *u,
                        double damp, int *info, hipStream_t s);
void launch_theta_rhs(const DevLevel &L, const double *geom, const double *bq, double nu, double sigma,
                      const double *u, const double *g, double *rhs, hipStream_t s);

// batched path (k_batch.hip): IntoL20 per (offset, length) pressure block; copies between a
// problem's numbering and the merged numbering (dst[i] = src[idx[i]], dst[idx[i]] = src[i]);
// BLAS-1 with one scalar per problem over a chunk table (start, length, problem, -)
constexpr int SEG_CHUNK = 4096;
template <typename T>
void launch_l20_seg(T *p, const int *seg, int nseg, hipStream_t s);
template <typename T>
void launch_gather_d(double *dst, const T *src, const int *idx, int n, hipStream_t s);
template <typename T>
void launch_scatter_d(T *dst, const double *src, const int *idx, int n, hipStream_t s);
void launch_seg_dot(const double *x, const double *y, const int4 *ch, int nch, double *part, hipStream_t s);
void launch_seg_axpy_dot(double *w, const double *h, int ldh, const double *v, const double *vn, const int4 *ch,
                         int nch, double *part, hipStream_t s);
void launch_seg_reduce(const double *part, const int *cptr, int nb, double *out, int ldo, hipStream_t s);
void launch_seg_scale(double *y, const double *x, const double *coef, const int4 *ch, int nch, hipStream_t s);
void launch_seg_lincomb(double *x, const double *z, long stride, const double *coef, int ldc, int m, const int *mask,
                        const int4 *ch, int nch, hipStream_t s);

}  // namespace pmg

#define PMG_HIP_CHECK(expr)                                                          \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess) throw pmg::DeviceError(std::string(#expr) + ": " +          \
                                                 hipGetErrorString(e_));             \
  } while (0)

namespace pmg {
struct DeviceError {
  std::string what;
  explicit DeviceError(std::string w) : what(std::move(w)) {}
};
}  // namespace pmg
