// pmg_api.cpp - host side of libpmg: device-resident multigrid hierarchy, V-cycle and
// FGMRES drivers behind the C ABI of include/pmg.h.  Compiled by hipcc for gfx950.
//
// Everything that touches vectors runs on the device, on the context's stream; the host
// only sequences kernels and, in FGMRES, performs the (restart+1)-sized Givens/Hessenberg
// arithmetic of src/FE/FgmresIte.C:199-247 on scalars read back once per iteration.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <type_traits>
#include <vector>

#include "batch_merge.hpp"
#include "pmg_internal.hpp"

using namespace pmg;

// One problem of a batch context: its own coarsest level (the operator of its dense coarse
// system), the coarse factors, and the maps from its numbering into the merged levels.
struct BatchProblem {
  int n_levels = 0, clevel = 0, project = 0;
  int n_dof_fine = 0, n_dof_coarse = 0;
  DevLevel cl;
  Coarse co;
  int *cidx = nullptr, *fidx = nullptr;    // problem dof -> merged index (coarsest, finest level)
  double *cf = nullptr, *cu = nullptr;     // FP64 coarse right-hand side and solution
  double *io = nullptr;                    // staging of one problem vector (host memory calls)
  bool in_job = false;                     // its coarse visits run in the level's batched launch
};

struct pmg_ctx {
  int device = 0;
  pmg_params prm{};
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr;   // coarse LU, overlapping the batched patch LUs
  hipEvent_t evs = nullptr;
  std::vector<DevLevel> L;
  // host copies of the dof maps (needed to build the transfers to the next level)
  std::vector<std::vector<int>> h_u_global, h_p_global;
  Coarse coarse;
  RedScratch red;
  std::string err;
  bool setup_done = false;
  // FGMRES workspace
  int restart_alloc = 0;
  double *V = nullptr, *Z = nullptr, *W = nullptr, *Hdev = nullptr, *coef = nullptr;
  // the FGMRES operator's input velocity interleaved (4 doubles per finest velocity dof)
  double *xa64 = nullptr;
  long xa64_n = 0;
  double *hpin = nullptr;   // pinned host scalars
  int hpin_len = 0;
  std::vector<double> hist;
  // staging buffers for PMG_MEM_HOST calls
  double *stage[3] = {nullptr, nullptr, nullptr};
  long stage_len = 0;
  // timers
  bool profiling = false;
  pmg_timers tm{};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<void *> allocs;
  // batch context (pmg_batch_create): the levels are merged block-diagonal levels
  bool batch = false;
  std::vector<BatchProblem> bi;
  int4 *seg_ch = nullptr;                  // FGMRES chunk table over the merged finest vector
  int n_seg_ch = 0;
  int *seg_cptr = nullptr;                 // chunks of problem b: [seg_cptr[b], seg_cptr[b+1])
  // the chunks the FGMRES steps run: all of them, or (finished problems left out) those of
  // the unfinished problems in the same order, with the per-problem ranges into that list
  int4 *seg_ch_act = nullptr, *seg_ch_live = nullptr;
  int n_seg_ch_act = 0;
  int *seg_cptr_act = nullptr, *seg_cptr_live = nullptr;
  std::vector<int4> h_seg_ch;
  std::vector<int> h_seg_cptr;
  double *seg_part = nullptr;              // one partial sum per chunk
  double *bX = nullptr, *bB = nullptr;     // merged finest solution and right-hand side
  double *bH = nullptr;                    // per problem: MGS coefficients of one iteration
  double *bcoef = nullptr;                 // per problem scalars / combination coefficients
  int *bmask = nullptr;
  int b_restart_alloc = 0;
  // the V-cycle as a HIP graph (captured on first use, dropped when levels or parameters change)
  hipGraph_t vgraph = nullptr;
  hipGraphExec_t vexec = nullptr;
  // problems the merged sweeps and coarse solves still cover (batch), and the number of
  // patches they had at the last compaction
  std::vector<char> b_live;
  long b_live_cells = 0;
};

namespace {

template <typename T>
T *dalloc(pmg_ctx *c, size_t n) {
  if (n == 0) n = 1;
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, n * sizeof(T));
  if (e != hipSuccess) throw DeviceError(std::string("hipMalloc(") + std::to_string(n * sizeof(T)) + "): " +
                                         hipGetErrorString(e));
  c->allocs.push_back(p);
  return static_cast<T *>(p);
}
void dfree(pmg_ctx *c, void *p) {
  if (!p) return;
  auto it = std::find(c->allocs.begin(), c->allocs.end(), p);
  if (it != c->allocs.end()) c->allocs.erase(it);
  (void)hipFree(p);
}
template <typename T>
void h2d(pmg_ctx *c, T *d, const T *h, size_t n) {
  if (n) PMG_HIP_CHECK(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, c->stream));
}

struct Fail {
  int code;
  std::string msg;
};

int guard(pmg_ctx *c, const std::function<void()> &f);

// ------------------------------------------------------------------ event timing
// HIP events around one launch (or one group of launches) on the context's stream; only
// active while profiling is on, because it waits for the launch to finish.
struct Timed {
  pmg_ctx *c;
  pmg_kstat *k, *k2;   // k2 (may be null): a second statistic the launch also counts in
  double bytes, flops;
  Timed(pmg_ctx *c_, pmg_kstat *k_, double b = 0, double f = 0, pmg_kstat *k2_ = nullptr)
      : c(c_), k(k_), k2(k2_), bytes(b), flops(f) {
    if (c->profiling) PMG_HIP_CHECK(hipEventRecord(c->ev0, c->stream));
  }
  ~Timed() noexcept(false) {
    if (!c->profiling) return;
    PMG_HIP_CHECK(hipEventRecord(c->ev1, c->stream));
    PMG_HIP_CHECK(hipEventSynchronize(c->ev1));
    float ms = 0;
    PMG_HIP_CHECK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    for (pmg_kstat *q : {k, k2}) {
      if (!q) continue;
      q->ms += ms;
      q->calls += 1;
      q->bytes += bytes;
      q->flops += flops;
    }
  }
};

// Algorithmic bytes of the type-4 operator rows (values + column indices + row pointers,
// right-hand side read, output written; gathers of x counted as cache hits), for values of
// vs bytes.  Dirichlet rows read only the diagonal blocks, as the kernels do.
struct RowBytes {
  // shared: one velocity block instead of nine (three on Dirichlet rows)
  double vel(const pmg_level_desc *d, int i, int vs, bool shared) const {
    const double nA = d->a_rowptr[i + 1] - d->a_rowptr[i];
    if (i < d->n_active) {
      const double nBT = d->bt_rowptr[i + 1] - d->bt_rowptr[i];
      return nA * ((shared ? 1.0 : 9.0) * vs + 4) + nBT * (3.0 * vs + 4) + 16;
    }
    return nA * ((shared ? 1.0 : 3.0) * vs + 4) + 8;
  }
  double pre(const pmg_level_desc *d, int i, int vs) const {
    return double(d->b_rowptr[i + 1] - d->b_rowptr[i]) * (3.0 * vs + 4) + 8;
  }
};

// ------------------------------------------------------------------ transfers
// P (fine rows) accumulated cellwise from QQ and divided by the fine-dof multiplicity, as
// Prolongate does (MGComponents3D.C:468-708); R = P^T, which is what DefectRestriction
// computes (MGComponents3D.C:996-1240).  Cells without a parent contribute nothing: their
// rows of P stay empty.
void build_transfer(pmg_ctx *c, int nloc_f, int nloc_c, int n_fine_cells, const int *parent, const int *child,
                    const double *qq, const std::vector<int> &fglob, const std::vector<int> &cglob, int n_fdof,
                    int n_cdof, Csr &P, Csr &R) {
  struct Trip {
    int r, c;
    double v;
  };
  std::vector<Trip> t;
  t.reserve(size_t(n_fine_cells) * nloc_f * 8);
  std::vector<int> mult(n_fdof, 0);
  for (int fc = 0; fc < n_fine_cells; ++fc) {
    const int pc = parent[fc], ch = child[fc];
    if (pc < 0) continue;   // merged levels: a problem solved directly on this level
    const double *Q = qq + size_t(ch) * nloc_f * nloc_c;
    const int *FineDOF = &fglob[size_t(fc) * nloc_f];
    const int *CoarseDOF = &cglob[size_t(pc) * nloc_c];
    for (int k = 0; k < nloc_f; ++k) {
      mult[FineDOF[k]]++;
      for (int l = 0; l < nloc_c; ++l) {
        const double v = Q[k * nloc_c + l];
        if (v != 0.0) t.push_back({FineDOF[k], CoarseDOF[l], v});
      }
    }
  }
  std::sort(t.begin(), t.end(), [](const Trip &a, const Trip &b) { return a.r != b.r ? a.r < b.r : a.c < b.c; });
  std::vector<int> rp(n_fdof + 1, 0), col;
  std::vector<double> val;
  for (size_t i = 0; i < t.size();) {
    size_t j = i;
    double s = 0;
    while (j < t.size() && t[j].r == t[i].r && t[j].c == t[i].c) s += t[j++].v;
    col.push_back(t[i].c);
    val.push_back(s / mult[t[i].r]);
    rp[t[i].r + 1]++;
    i = j;
  }
  for (int i = 0; i < n_fdof; ++i) rp[i + 1] += rp[i];
  // transpose
  std::vector<int> rrp(n_cdof + 1, 0), rcol(col.size());
  std::vector<double> rval(col.size());
  for (int cc : col) rrp[cc + 1]++;
  for (int i = 0; i < n_cdof; ++i) rrp[i + 1] += rrp[i];
  std::vector<int> pos(rrp.begin(), rrp.end() - 1);
  for (int r = 0; r < n_fdof; ++r)
    for (int k = rp[r]; k < rp[r + 1]; ++k) {
      const int q = pos[col[k]]++;
      rcol[q] = r;
      rval[q] = val[k];
    }
  auto up = [&](Csr &m, int nr, int ncl, const std::vector<int> &r_, const std::vector<int> &c_,
                const std::vector<double> &v_) {
    m.n_rows = nr;
    m.n_cols = ncl;
    m.nnz = int(c_.size());
    m.rowptr = dalloc<int>(c, r_.size());
    m.col = dalloc<int>(c, c_.size());
    m.val = dalloc<double>(c, v_.size());
    h2d(c, m.rowptr, r_.data(), r_.size());
    h2d(c, m.col, c_.data(), c_.size());
    h2d(c, m.val, v_.data(), v_.size());
  };
  up(P, n_fdof, n_cdof, rp, col, val);
  up(R, n_cdof, n_fdof, rrp, rcol, rval);
  PMG_HIP_CHECK(hipStreamSynchronize(c->stream));
}

// RestrictFunction for Q2 (MGComponents3D.C:1509-1755): on every coarse cell the L2
// projection onto Q2 of the fine function on its children, Val2 = sum_j QQr_j f_j with the
// restriction matrix of child j (TFEDatabase3D::GetRestrictionMatrix3D,
// FEDatabase3D.C:1791-1970: G^{-1} R_j, G the coarse basis' mass matrix on the reference
// cube, R_j the mixed mass of coarse and fine basis on the child); a coarse dof's value is
// the mean of its cells' values.  The fine basis on child j is the coarse basis restricted
// there with the prolongation table P_j (fine functional k of coarse basis l), so
// R_j = P_j^T G / 8 and QQr_j = G^{-1} P_j^T G / 8, with G = g x g x g, g the 1D quadratic
// Lagrange mass on [-1, 1] (nodes -1, 0, 1, local index a + 3b + 9c).
void build_function_restriction(pmg_ctx *c, int n_fine_cells, const int *parent, const int *child, const double *qq,
                                const std::vector<int> &fglob, const std::vector<int> &cglob, int n_fdof, int n_cdof,
                                Csr &RF) {
  const double g1[3][3] = {{4 / 15.0, 2 / 15.0, -1 / 15.0}, {2 / 15.0, 16 / 15.0, 2 / 15.0},
                           {-1 / 15.0, 2 / 15.0, 4 / 15.0}};
  const double gi1[3][3] = {{9 / 2.0, -3 / 4.0, 3 / 2.0}, {-3 / 4.0, 9 / 8.0, -3 / 4.0}, {3 / 2.0, -3 / 4.0, 9 / 2.0}};
  auto tens = [](const double (&m)[3][3], int i, int j) {
    return m[i % 3][j % 3] * m[(i / 3) % 3][(j / 3) % 3] * m[i / 9][j / 9];
  };
  // QQr_j[k][l] = sum_{a, b} Ginv[k][a] P_j[b][a] G[b][l] / 8 (coarse k, fine l)
  std::vector<double> qr(size_t(8) * 729, 0.0);
  for (int ch = 0; ch < 8; ++ch) {
    const double *P = qq + size_t(ch) * 729;
    std::vector<double> t(729, 0.0);   // t[a][l] = sum_b P[b][a] G[b][l]
    for (int a = 0; a < 27; ++a)
      for (int l = 0; l < 27; ++l) {
        double s = 0;
        for (int b = 0; b < 27; ++b) s += P[b * 27 + a] * tens(g1, b, l);
        t[a * 27 + l] = s;
      }
    for (int k = 0; k < 27; ++k)
      for (int l = 0; l < 27; ++l) {
        double s = 0;
        for (int a = 0; a < 27; ++a) s += tens(gi1, k, a) * t[a * 27 + l];
        qr[size_t(ch) * 729 + k * 27 + l] = s / 8.0;
      }
  }
  struct Trip {
    int r, c;
    double v;
  };
  std::vector<Trip> t;
  t.reserve(size_t(n_fine_cells) * 729);
  std::vector<int> cells_of(n_cdof, 0);   // coarse cells containing each coarse dof
  std::vector<char> seen;
  std::vector<int> first_child;           // per coarse cell: counted once
  int n_coarse_cells = 0;
  for (int fc = 0; fc < n_fine_cells; ++fc) n_coarse_cells = std::max(n_coarse_cells, parent[fc] + 1);
  seen.assign(n_coarse_cells, 0);
  for (int fc = 0; fc < n_fine_cells; ++fc) {
    const int pc = parent[fc], ch = child[fc];
    if (pc < 0) continue;
    const int *FineDOF = &fglob[size_t(fc) * NU_LOC];
    const int *CoarseDOF = &cglob[size_t(pc) * NU_LOC];
    if (!seen[pc]) {
      seen[pc] = 1;
      for (int k = 0; k < NU_LOC; ++k) cells_of[CoarseDOF[k]]++;
    }
    for (int k = 0; k < NU_LOC; ++k)
      for (int l = 0; l < NU_LOC; ++l) {
        const double v = qr[size_t(ch) * 729 + k * 27 + l];
        if (v != 0.0) t.push_back({CoarseDOF[k], FineDOF[l], v});
      }
  }
  std::sort(t.begin(), t.end(), [](const Trip &a, const Trip &b) { return a.r != b.r ? a.r < b.r : a.c < b.c; });
  std::vector<int> rp(n_cdof + 1, 0), col;
  std::vector<double> val;
  for (size_t i = 0; i < t.size();) {
    size_t j = i;
    double s = 0;
    while (j < t.size() && t[j].r == t[i].r && t[j].c == t[i].c) s += t[j++].v;
    col.push_back(t[i].c);
    val.push_back(s / cells_of[t[i].r]);
    rp[t[i].r + 1]++;
    i = j;
  }
  for (int i = 0; i < n_cdof; ++i) rp[i + 1] += rp[i];
  RF.n_rows = n_cdof;
  RF.n_cols = n_fdof;
  RF.nnz = int(col.size());
  RF.rowptr = dalloc<int>(c, rp.size());
  RF.col = dalloc<int>(c, std::max<size_t>(1, col.size()));
  RF.val = dalloc<double>(c, std::max<size_t>(1, val.size()));
  h2d(c, RF.rowptr, rp.data(), rp.size());
  h2d(c, RF.col, col.data(), col.size());
  h2d(c, RF.val, val.data(), val.size());
  PMG_HIP_CHECK(hipStreamSynchronize(c->stream));
}

}  // namespace

namespace {

// Per-cell gather map of the local Vanka matrix (layout: PM_* constants).  The entry of
// local velocity row j and local column m is the CSR entry of row UDOF_j whose column is
// UDOF_m (NSE_MGLevel4.C:1149-1287 finds it by linear search; the map stores where it is
// in the row-blocked values).  Dirichlet rows keep only their diagonal entry.
std::vector<int> build_patch_map(const pmg_level_desc *d, const std::vector<int> &ug, const std::vector<int> &pg) {
  const int nc = d->n_cells;
  std::vector<int> map(size_t(nc) * PM_STRIDE, -1);
  // global dof -> local index in the current cell (marks reset per cell)
  std::vector<int> uloc(d->n_udof, -1), ploc(d->n_pdof, -1);
  for (int cell = 0; cell < nc; ++cell) {
    int *mp = map.data() + size_t(cell) * PM_STRIDE;
    const int *U = ug.data() + size_t(cell) * NU_LOC;
    const int *P = pg.data() + size_t(cell) * NP_LOC;
    for (int j = NU_LOC - 1; j >= 0; --j) uloc[U[j]] = j;  // first local dof wins
    for (int j = NP_LOC - 1; j >= 0; --j) ploc[P[j]] = j;
    for (int j = 0; j < NU_LOC; ++j) {
      const int row = U[j], k0 = d->a_rowptr[row], k1 = d->a_rowptr[row + 1];
      const bool dir = row >= d->n_active;
      mp[PM_NA + j] = dir ? -(k1 - k0) : (k1 - k0);
      for (int k = k0; k < k1; ++k) {
        const int m = uloc[d->a_kcol[k]];
        if (m >= 0 && mp[PM_A + j * NU_LOC + m] < 0) mp[PM_A + j * NU_LOC + m] = 9 * k0 + (k - k0);
      }
      const int q0 = d->bt_rowptr[row], q1 = d->bt_rowptr[row + 1];
      mp[PM_MT + j] = q1 - q0;
      if (!dir)
        for (int q = q0; q < q1; ++q) {
          const int m = ploc[d->bt_kcol[q]];
          if (m >= 0 && mp[PM_BT + j * NP_LOC + m] < 0) mp[PM_BT + j * NP_LOC + m] = 3 * q0 + (q - q0);
        }
    }
    for (int j = 0; j < NP_LOC; ++j) {
      const int row = P[j], k0 = d->b_rowptr[row], k1 = d->b_rowptr[row + 1];
      mp[PM_NB + j] = k1 - k0;
      for (int k = k0; k < k1; ++k) {
        const int m = uloc[d->b_kcol[k]];
        if (m >= 0 && mp[PM_B + j * NU_LOC + m] < 0) mp[PM_B + j * NU_LOC + m] = 3 * k0 + (k - k0);
      }
    }
    for (int j = 0; j < NU_LOC; ++j) uloc[U[j]] = -1;
    for (int j = 0; j < NP_LOC; ++j) ploc[P[j]] = -1;
  }
  return map;
}

// Results reported by the device after the fact: the coarse LU (second stream) pivot check
// and its time, and the wait-bound flag of the coarse triangular sweeps.
void check_coarse(pmg_ctx *c, Coarse &C);
void check_nodal(pmg_ctx *c);
void check_deferred(pmg_ctx *c) {
  if (!c->batch) {
    check_coarse(c, c->coarse);
    check_nodal(c);
    return;
  }
  for (auto &b : c->bi) check_coarse(c, b.co);
}
// nodal Vanka: the largest pressure dof (+1) whose local system had a zero pivot in the
// sweeps since the last check; reported once and cleared (the stream is idle here)
void check_nodal(pmg_ctx *c) {
  for (size_t l = 0; l < c->L.size(); ++l) {
    DevLevel &D = c->L[l];
    if (!D.nodal_info) continue;
    int e = 0;
    PMG_HIP_CHECK(hipMemcpyAsync(&e, D.nodal_info, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    PMG_HIP_CHECK(hipStreamSynchronize(c->stream));
    if (e) {
      launch_fill<int>(D.nodal_info, 0, 1, c->stream);
      PMG_HIP_CHECK(hipStreamSynchronize(c->stream));
      throw Fail{PMG_ERR_SINGULAR_PATCH, "zero pivot in the nodal Vanka patch of pressure dof " +
                                             std::to_string(e - 1) + " of level " + std::to_string(l)};
    }
  }
}
void check_coarse(pmg_ctx *c, Coarse &C) {
  if (C.info_pending) {
    C.info_pending = false;
    int cinfo = 0;
    PMG_HIP_CHECK(hipEventSynchronize(C.ev_done));
    PMG_HIP_CHECK(hipMemcpy(&cinfo, C.info, sizeof(int), hipMemcpyDeviceToHost));
    float ms = 0;
    PMG_HIP_CHECK(hipEventElapsedTime(&ms, C.ev_start, C.ev_done));
    c->tm.coarse_lu.ms += ms;
    c->tm.coarse_lu.calls += 1;
    if (C.block) {
      int bi[2] = {0, 0};
      PMG_HIP_CHECK(hipMemcpy(bi, C.cb.info, sizeof(bi), hipMemcpyDeviceToHost));
      const double m = C.cb.m, q = C.cb.q;
      c->tm.coarse_lu.bytes += 8.0 * (m * m * 2 + 2 * m * (m + 3 * q) + 3 * q * m * 3 + q * q * 3 + 2 * q * (3 * m + q));
      c->tm.coarse_lu.flops += (2.0 / 3.0) * (m * m * m + q * q * q) + 2.0 * m * m * (m + 3 * q) +
                               6.0 * q * q * m + 6.0 * q * m * m + 2.0 * q * q * q + 6.0 * q * q * m;
      if (bi[0] || bi[1]) {
        C.ready = false;
        throw Fail{PMG_ERR_SINGULAR_COARSE, std::string("zero pivot in the coarse ") +
                                                (bi[0] ? "velocity block" : "Schur complement") + " LU"};
      }
    } else {
      c->tm.coarse_lu.bytes += double(C.n) * C.n * 16.0;
      c->tm.coarse_lu.flops += (2.0 / 3.0) * double(C.n) * C.n * C.n;
    }
    if (cinfo) {
      C.ready = false;
      throw Fail{PMG_ERR_SINGULAR_COARSE, "zero pivot " + std::to_string(cinfo) + " in the coarse LU"};
    }
  }
  if (!C.err) return;
  int e = 0;
  PMG_HIP_CHECK(hipMemcpyAsync(&e, C.err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  PMG_HIP_CHECK(hipStreamSynchronize(c->stream));
  if (e) {
    // reported once: cleared (outside any capture, the stream is idle here) so that later
    // calls on the context are not failed by an old timeout
    launch_fill<int>(C.err, 0, 1, c->stream);
    PMG_HIP_CHECK(hipStreamSynchronize(c->stream));
    throw Fail{PMG_ERR_DEVICE, "coarse triangular sweep: wait bound exceeded"};
  }
}

// the streamed forms of new operator values (CSR arrays D.a, D.b, D.bt on the device):
// row-blocked FP64 (and FP32) copies and, when the A blocks have it, the shared velocity
// block form (FP64: block 0 of D.a read in place; FP32: its own copy)
void drop_graph(pmg_ctx *c);

void refresh_values(pmg_ctx *c, DevLevel &D) {
  launch_row_blocked<double>(D, D.av64, D.bv64, D.btv64, c->stream);
  if (D.av32) launch_row_blocked<float>(D, D.av32, D.bv32, D.btv32, c->stream);
  const bool was = D.a_shared;
  D.a_shared = c->prm.shared_block && shared_velocity_block(D, c->red.partials, c->stream);
  // the captured V-cycle holds the kernels and operator views of the previous form
  if (D.a_shared != was) drop_graph(c);
  if (D.a_shared && D.av32) {
    if (!D.a1_32) {
      D.a1_32 = dalloc<float>(c, size_t(D.a_nnz) + ROW_PAD);
      launch_fill<float>(D.a1_32 + D.a_nnz, 0.0f, ROW_PAD, c->stream);
    }
    launch_d2f(D.a1_32, D.a, D.a_nnz, c->stream);
  }
  PMG_HIP_CHECK(hipStreamSynchronize(c->stream));
}

void upload_values(pmg_ctx *c, DevLevel &D, const pmg_level_desc *d) {
  for (int b = 0; b < 9; ++b) h2d(c, D.a + size_t(b) * D.a_nnz, d->a[b], D.a_nnz);
  for (int b = 0; b < 3; ++b) {
    h2d(c, D.b + size_t(b) * D.b_nnz, d->b[b], D.b_nnz);
    h2d(c, D.bt + size_t(b) * D.bt_nnz, d->bt[b], D.bt_nnz);
  }
  refresh_values(c, D);
}

bool mixed(const pmg_ctx *c) { return c->prm.precision != PMG_PREC_FP64; }

// sizes, CSR structure and values (structure of arrays + row-blocked FP64, FP32 copy if
// fp32) of a level's operator
void upload_operator(pmg_ctx *c, DevLevel &D, const pmg_level_desc *d, bool fp32) {
  D.n_udof = d->n_udof; D.n_active = d->n_active; D.hanging_bound = d->hanging_bound;
  D.n_dirichlet = d->n_dirichlet; D.n_pdof = d->n_pdof; D.n_dof = 3 * d->n_udof + d->n_pdof;
  D.a_nnz = d->a_nnz; D.b_nnz = d->b_nnz; D.bt_nnz = d->bt_nnz;
  D.a_act_entries = d->a_rowptr[d->n_active];
  // the residual's work-item descriptors hold row lengths in 16 bits
  auto longest = [](const int *rp, int n) {
    int mx = 0;
    for (int i = 0; i < n; ++i) mx = std::max(mx, rp[i + 1] - rp[i]);
    return mx;
  };
  if (longest(d->a_rowptr, d->n_udof) > 65535 || longest(d->bt_rowptr, d->n_udof) > 65535 ||
      longest(d->b_rowptr, d->n_pdof) > 65535)
    throw Fail{PMG_ERR_UNSUPPORTED, "operator row with more than 65535 entries"};
  D.a_rowptr = dalloc<int>(c, D.n_udof + 1); h2d(c, D.a_rowptr, d->a_rowptr, D.n_udof + 1);
  // the streamed arrays (column indices, row-blocked values) carry ROW_PAD zeroed entries
  // after their end: the row kernels' 4-entry windows of rows shorter than 4 read past it
  auto padded = [&](auto *&ptr, size_t n) {
    using E = std::remove_reference_t<decltype(*ptr)>;
    ptr = dalloc<E>(c, n + ROW_PAD);
    launch_fill<E>(ptr + n, E(0), ROW_PAD, c->stream);
  };
  padded(D.a_kcol, D.a_nnz); h2d(c, D.a_kcol, d->a_kcol, D.a_nnz);
  D.b_rowptr = dalloc<int>(c, D.n_pdof + 1); h2d(c, D.b_rowptr, d->b_rowptr, D.n_pdof + 1);
  padded(D.b_kcol, D.b_nnz); h2d(c, D.b_kcol, d->b_kcol, D.b_nnz);
  D.bt_rowptr = dalloc<int>(c, D.n_udof + 1); h2d(c, D.bt_rowptr, d->bt_rowptr, D.n_udof + 1);
  padded(D.bt_kcol, D.bt_nnz); h2d(c, D.bt_kcol, d->bt_kcol, D.bt_nnz);
  D.a = dalloc<double>(c, 9 * size_t(D.a_nnz));
  D.b = dalloc<double>(c, 3 * size_t(D.b_nnz));
  D.bt = dalloc<double>(c, 3 * size_t(D.bt_nnz));
  padded(D.av64, 9 * size_t(D.a_nnz));
  padded(D.bv64, 3 * size_t(D.b_nnz));
  padded(D.btv64, 3 * size_t(D.bt_nnz));
  if (fp32) {
    padded(D.av32, 9 * size_t(D.a_nnz));
    padded(D.bv32, 3 * size_t(D.b_nnz));
    padded(D.btv32, 3 * size_t(D.bt_nnz));
  }
  upload_values(c, D, d);
}

void ensure_stage(pmg_ctx *c, long n) {
  if (n <= c->stage_len) return;
  for (auto &p : c->stage) {
    dfree(c, p);
    p = dalloc<double>(c, n);
  }
  c->stage_len = n;
}

// ------------------------------------------------------------------ level operations
void correct_nodes(pmg_ctx *c, const DevLevel &D, double *u) {
  launch_dirichlet_slices<double>(u, nullptr, D.n_udof, D.hanging_bound, D.n_dirichlet, c->stream);
}
void set_dirichlet(pmg_ctx *c, const DevLevel &D, double *u, const double *f) {
  launch_dirichlet_slices<double>(u, f, D.n_udof, D.hanging_bound, D.n_dirichlet, c->stream);
}
// IntoL20 of a level's pressure constants: for a single problem it is folded into the
// type-4 launch; a merged level projects every problem's block on its own
bool fused_l20(const pmg_ctx *c) { return !c->batch && c->prm.internal_project_pressure; }
template <typename T>
void project_pressure(pmg_ctx *c, const DevLevel &D, T *x) {
  if (c->batch)
    launch_l20_seg<T>(x + 3L * D.n_udof, D.l20_seg, D.n_l20_seg, c->stream);
  else if (c->prm.internal_project_pressure)
    launch_l20<T>(x + 3L * D.n_udof, D.n_pdof, c->red, c->stream);
}
void correct_defect(pmg_ctx *c, const DevLevel &D, double *d) {
  launch_dirichlet_slices<double>(d, nullptr, D.n_udof, D.n_active, D.n_udof - D.n_active, c->stream);
  project_pressure<double>(c, D, d);
}
template <typename T>
void correct_nodes_t(pmg_ctx *c, const DevLevel &D, T *u) {
  launch_dirichlet_slices<T>(u, nullptr, D.n_udof, D.hanging_bound, D.n_dirichlet, c->stream);
}
template <typename T>
void set_dirichlet_t(pmg_ctx *c, const DevLevel &D, T *u, const T *f) {
  launch_dirichlet_slices<T>(u, f, D.n_udof, D.hanging_bound, D.n_dirichlet, c->stream);
}
template <typename T>
void correct_defect_t(pmg_ctx *c, const DevLevel &D, T *d) {
  launch_dirichlet_slices<T>(d, nullptr, D.n_udof, D.n_active, D.n_udof - D.n_active, c->stream);
  project_pressure<T>(c, D, d);
}

template <typename T>
OpView<T> opview(const DevLevel &D);
template <>
OpView<double> opview<double>(const DevLevel &D) { return D.op64(); }
template <>
OpView<float> opview<float>(const DevLevel &D) { return D.op32(); }

// TNSE_MGLevel4::CellVanka, colored (NSE_MGLevel4.C:603-1649), one sweep
// bytes of one stored-factor patch solve: factors, local residual read, u read+write,
// dof maps and the cell index
inline double solve_bytes(int fbytes, int vs) {
  return double(NLOC) * NLOC * fbytes + NLOC * double(vs) + 2.0 * NLOC * vs + (NU_LOC + NP_LOC + 1) * 4.0;
}

// TNSE_MGLevel4::CellVanka, colored (NSE_MGLevel4.C:603-1649), one sweep
// TNSE_MGLevel4::NodalVanka (NSE_MGLevel4.C:4098-4549), one sweep: SetDirichletNodes, the
// pressure-dof patches set by set, then the relaxation with alpha on every dof (always, the
// nodal sweep has no |1 - alpha| test) and SetDirichletNodes again
void nodal_sweep(pmg_ctx *c, DevLevel &D, double *u, const double *f, double *aux, bool last) {
  const double alpha = c->prm.sc_smooth_damp_factor_saddle;
  const double damp = c->prm.sc_smooth_damp_factor_coarse_saddle;
  const bool full = c->prm.sc_smoother_saddle == 4;
  set_dirichlet(c, D, u, f);
  launch_copy<double>(aux, u, D.n_dof, c->stream);
  const OpView<double> op = D.op64();
  const int ns = int(D.nodal_ptr.size()) - 1;
  for (int k = 0; k < ns; ++k) {
    const int p0 = D.nodal_ptr[k], n = D.nodal_ptr[k + 1] - p0;
    const double nl = 3.0 * NODAL_MAXU + 1;
    Timed tm(c, &c->tm.vanka_solve, D.nodal_bytes[k], n * (full ? (2.0 / 3.0) * nl * nl * nl : 20.0 * nl));
    launch_nodal_vanka(op, full, D.nodal_order + p0, n, f, u, damp, D.nodal_info, c->stream);
  }
  launch_relax<double>(u, aux, alpha, D.n_dof, c->stream);
  if (last) set_dirichlet(c, D, u, f);
}

// the interleaved velocity copy a level's FP32 sweeps gather from (stored inverses and the
// shared velocity block), or null
template <typename T>
float *sweep_copy(const DevLevel &D) {
  if constexpr (sizeof(T) == 4) {
    if (D.ua32 && D.a_shared && (D.minv16 || D.minv32)) return D.ua32;
  }
  return nullptr;
}

template <typename T>
// last = false: another sweep with the same f follows at once and sets the Dirichlet values
// again before anything reads them, so this sweep's closing SetDirichletNodes is left out
// first = false: this sweep directly follows one with last = false on the same u and f
// (nothing touched them in between): with the interleaved velocity copy and no global
// relaxation, that copy is still current and the Dirichlet values are still set, so the
// sweep start (and the previous sweep's copy back into u) are left out
// close = false: the closing SetDirichletNodes of the last sweep is left out (the caller
// overwrites the same slice next); started: the caller has done the sweep's start (the
// Dirichlet slice of u set from f and, FP32 with inverses, the interleaved copy built)
void vanka_sweep(pmg_ctx *c, DevLevel &D, T *u, const T *f, T *aux, T *rloc, bool last = true, bool first = true,
                 bool close = true, bool started = false) {
  const bool prof = c->profiling;
  auto t0 = std::chrono::steady_clock::now();
  if constexpr (sizeof(T) == 8) {
    if (D.nodal) {
      nodal_sweep(c, D, u, f, aux, last);
      if (prof) {
        c->tm.sweep_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        c->tm.sweep_calls += 1;
      }
      return;
    }
  }
  const double alpha = c->prm.sc_smooth_damp_factor_saddle;
  const T damp = T(c->prm.sc_smooth_damp_factor_coarse_saddle);
  const bool relax = std::fabs(1 - alpha) > 1e-3;
  const int vs = sizeof(T);
  const int fb = sizeof(T) == 8 ? 8 : (c->prm.precision == PMG_PREC_FP32_BF16 ? 2 : 4);
  // FP32 with inverses and the shared block: the residual gathers the velocity from the
  // interleaved copy ua, built here and kept current by the local updates
  float *ua = sweep_copy<T>(D);
  if (!started) {
    if (ua && !(first == false && !relax))
      launch_sweep_start(reinterpret_cast<float *>(u), reinterpret_cast<const float *>(f), D.n_udof, D.hanging_bound,
                         D.n_dirichlet, ua, c->stream);
    else if (!ua)
      set_dirichlet_t<T>(c, D, u, f);
  }
  if (relax) launch_copy<T>(aux, u, D.n_dof, c->stream);
  const OpView<T> op = opview<T>(D);
  for (int col = 0; col < D.n_colors; ++col) {
    const int p0 = D.ptr_colors[col];
    // all patches of the color, or (compacted merged level) those of unfinished problems
    int nc = D.ptr_colors[col + 1] - p0;
    long n_items = long(nc) * (NU_LOC + NP_LOC);
    const int *items = D.res_items + size_t(p0) * (NU_LOC + NP_LOC);
    const int *desc = D.res_desc + 4 * size_t(p0) * (NU_LOC + NP_LOC);
    const int *slots = nullptr;
    double part = 1.0;
    if (D.compacted) {
      n_items = D.act_items_off[col + 1] - D.act_items_off[col];
      items = D.act_items + D.act_items_off[col];
      desc = D.act_desc + 4 * D.act_items_off[col];
      slots = D.act_slots + D.act_slots_off[col];
      part = nc ? double(D.act_slots_off[col + 1] - D.act_slots_off[col]) / nc : 0.0;
      nc = int(D.act_slots_off[col + 1] - D.act_slots_off[col]);
      if (nc == 0) continue;
    }
    {
      const double rb = D.res_bytes[D.a_shared][vs == 8 ? 0 : 1][col] * part;
      Timed tm(c, &c->tm.vanka_res, rb, 0.0, &D == &c->L.back() ? &c->tm.vanka_res_finest : nullptr);
      const int pi = sizeof(T) == 8 ? 0 : 1;
      const int *xg = !D.compacted && D.h_xcd_grouped[col] && D.tune_res_xcd[pi] ? D.xcd_goff + 9 * col : nullptr;
      T *uloc = sizeof(T) == 4 ? reinterpret_cast<T *>(D.uloc32) : nullptr;
      launch_vanka_residual<T>(op, n_items, p0, D.invperm, items, desc, xg, u, ua, f, rloc, uloc, c->stream,
                               D.tune_res_pb[pi]);
    }
    // refinement steps (stored inverses only): the stored FP32 patch matrix read once per
    // patch, one more local residual and product per step
    const int nref = sizeof(T) == 4 ? c->prm.bf16_refine_steps : 0;
    if (nref > 0 && !D.kref)
      throw Fail{PMG_ERR_NOT_READY, "bf16_refine_steps was set after the set-up: call pmg_refactor"};
    const double kent = D.schur ? double(KREF_BLK) : double(KREF_FULL);
    Timed tm(c, &c->tm.vanka_solve, nc * (solve_bytes(fb, vs) + (nref ? 4.0 * kent : 0.0)),
             nc * (2.0 * NLOC * NLOC) * (1 + nref) + nc * 2.0 * kent * nref);
    if constexpr (sizeof(T) == 8) {
      launch_vanka_solve<double, double>(D.lu, D.cell_reorder, nc, p0, slots, rloc, D.u_global, D.p_global,
                                         D.n_udof, u, damp, c->stream);
    } else {
      if (D.schur && D.minv16)
        launch_vanka_apply_schur<uint16_t>(D.minv16, D.pos_dofs, nc, p0, slots, rloc, D.uloc32, D.n_udof, u, ua,
                                           damp, D.kref, nref, c->stream, ua != nullptr);
      else if (D.schur && D.minv32)
        launch_vanka_apply_schur<float>(D.minv32, D.pos_dofs, nc, p0, slots, rloc, D.uloc32, D.n_udof, u, ua, damp,
                                        D.kref, nref, c->stream, ua != nullptr);
      else if (D.minv16)
        launch_vanka_apply<uint16_t>(D.minv16, D.pos_dofs, nc, p0, slots, rloc, D.uloc32, D.n_udof, u, ua, damp,
                                     D.kref, D.invperm, nref, c->stream);
      else if (D.minv32)
        launch_vanka_apply<float>(D.minv32, D.pos_dofs, nc, p0, slots, rloc, D.uloc32, D.n_udof, u, ua, damp,
                                  D.kref, D.invperm, nref, c->stream);
      else if (c->prm.precision == PMG_PREC_FP32_BF16)
        launch_vanka_solve<float, uint16_t>(D.lu16, D.cell_reorder, nc, p0, slots, rloc, D.u_global, D.p_global,
                                            D.n_udof, u, damp, c->stream);
      else
        launch_vanka_solve<float, float>(D.lu32, D.cell_reorder, nc, p0, slots, rloc, D.u_global, D.p_global,
                                         D.n_udof, u, damp, c->stream);
    }
  }
  if constexpr (sizeof(T) == 4) {
    // the block-form local solves kept the velocity in the interleaved copy only (a next
    // sweep that skips its start keeps working on the copy)
    if (ua && D.schur && (last || relax)) launch_ua_to_u(ua, u, D.n_udof, c->stream);
  }
  if (relax) launch_relax<T>(u, aux, T(alpha), D.n_dof, c->stream);
  if (last && close) set_dirichlet_t<T>(c, D, u, f);
  if (prof) {
    c->tm.sweep_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    c->tm.sweep_calls += 1;
  }
}

// res_dev (single problem only): ||d||_2 after the projection
template <typename T>
void mg_defect(pmg_ctx *c, const DevLevel &D, const T *u, const T *f, T *d, double *res_dev) {
  Timed tm(c, &c->tm.defect, D.defect_bytes[D.a_shared][sizeof(T) == 8 ? 0 : 1], 0.0);
  launch_type4<T>(opview<T>(D), 2, u, f, d, fused_l20(c), c->batch ? nullptr : res_dev, c->red, c->stream);
  if (c->batch) project_pressure<T>(c, D, d);
}

template <typename T>
// zero_dirichlet: the Dirichlet rows of the coarse velocity get 0 instead of their sum (the
// CorrectDefect that follows a restriction, in the same launch)
void restrict_defect(pmg_ctx *c, const DevLevel &F, const DevLevel &C, const T *d, T *fc,
                     bool zero_dirichlet = false) {
  Timed tm(c, &c->tm.transfer, F.transfer_bytes, 0.0);
  // (batch with finished problems) the coarse rows of the unfinished problems only
  const bool sub = C.act_lists;
  launch_csr_spmv<T>(F.ru, 3, d, F.n_udof, fc, C.n_udof, c->stream, sub ? C.act_u : nullptr, C.n_act_u,
                     zero_dirichlet ? C.n_active : INT_MAX);
  launch_csr_spmv<T>(F.rp, 1, d + 3L * F.n_udof, 0, fc + 3L * C.n_udof, 0, c->stream, sub ? C.act_p : nullptr,
                     C.n_act_p);
}
template <typename T>
void prolongate(pmg_ctx *c, const DevLevel &F, const DevLevel &C, const T *uc, T *d) {
  Timed tm(c, &c->tm.transfer, F.transfer_bytes, 0.0);
  // (batch with finished problems) the fine rows of the unfinished problems only
  const bool sub = F.act_lists;
  launch_csr_spmv<T>(F.pu, 3, uc, C.n_udof, d, F.n_udof, c->stream, sub ? F.act_u : nullptr, F.n_act_u);
  launch_csr_spmv<T>(F.pp, 1, uc + 3L * C.n_udof, 0, d + 3L * F.n_udof, 0, c->stream, sub ? F.act_p : nullptr,
                     F.n_act_p);
}

// one coarse visit: the block form (explicit inverses of the blocks) or the LU sweeps
void coarse_solve_any(Coarse &C, const OpView<double> &op, const double *f, double *u, hipStream_t s) {
  if (C.block) {
    launch_coarse_rhs(op, f, C.cb.iota, C.cb.rr, 3 * C.cb.m + C.cb.q, C.cb.pin, s);
    coarse_block_solve(C.cb, op, C.cb.rr, f, u, s);
  } else {
    coarse_lu_solve(C, op, f, u, s);
  }
}
double coarse_visit_bytes(const Coarse &C) {
  if (!C.block) return double(C.n) * C.n * 8.0;
  const double m = C.cb.m, q = C.cb.q;
  return 8.0 * (m * m + q * (3 * m + q) + 3 * m * q);
}
double coarse_visit_flops(const Coarse &C) {
  if (!C.block) return 2.0 * double(C.n) * C.n;
  const double m = C.cb.m, q = C.cb.q;
  return 2.0 * (3 * m * m + q * (3 * m + q) + 3 * m * q);
}

// coarse solve in FP64 (vectors of type T converted at the boundary)
template <typename T>
void coarse_solve_t(pmg_ctx *c, const DevLevel &D, T *u, const T *f) {
  if (c->coarse.pending) {
    PMG_HIP_CHECK(hipStreamWaitEvent(c->stream, c->coarse.ev_done, 0));
    c->coarse.pending = false;
  }
  Timed tm(c, &c->tm.coarse_solve, coarse_visit_bytes(c->coarse), coarse_visit_flops(c->coarse));
  if constexpr (sizeof(T) == 8) {
    coarse_solve_any(c->coarse, D.op64(), f, u, c->stream);
  } else {
    launch_f2d(D.aux0, f, D.n_dof, c->stream);
    coarse_solve_any(c->coarse, D.op64(), D.aux0, D.sol, c->stream);
    launch_d2f(u, D.sol, D.n_dof, c->stream);
  }
}

// The job tables of the batched coarse visits: on every merged level, the problems whose
// coarsest level it is, that are still iterating and whose coarse system has the block
// form.  Rebuilt after every set-up and whenever the unfinished problems are recounted (never
// inside a graph capture).
void build_cb_jobs(pmg_ctx *c) {
  for (auto &b : c->bi) b.in_job = false;
  for (auto &D : c->L) {
    D.n_cb_jobs = 0;
    D.cb_job_prob.clear();
    if (D.coarse_insts.empty()) continue;
    std::vector<CbJob> jobs;
    CbGrid g{0, 0, 0};
    for (int k : D.coarse_insts) {
      BatchProblem &b = c->bi[k];
      if (!b.co.block || (!c->b_live.empty() && !c->b_live[k])) continue;
      const CoarseBlock &B = b.co.cb;
      CbJob J{b.cl.op64(), B.X, B.P1, B.rr, B.y, B.p, b.cidx, B.m, B.q, B.pin};
      const CbGrid gj = coarse_block_grid(J);
      g.rhs = std::max(g.rhs, gj.rhs);
      g.phase1 = std::max(g.phase1, gj.phase1);
      g.phase2 = std::max(g.phase2, gj.phase2);
      jobs.push_back(J);
      D.cb_job_prob.push_back(k);
      b.in_job = true;
    }
    if (jobs.empty()) continue;
    if (!D.cb_jobs) D.cb_jobs = dalloc<CbJob>(c, D.coarse_insts.size());
    PMG_HIP_CHECK(hipMemcpy(D.cb_jobs, jobs.data(), sizeof(CbJob) * jobs.size(), hipMemcpyHostToDevice));
    D.n_cb_jobs = int(jobs.size());
    D.cb_grid[0] = g.rhs;
    D.cb_grid[1] = g.phase1;
    D.cb_grid[2] = g.phase2;
  }
}

// the direct solves (SolveExact) of the problems whose coarsest level is this merged level:
// the block-form ones in one batched visit (three launches), the others one by one
template <typename T>
void batch_coarse(pmg_ctx *c, const DevLevel &D, T *u, const T *f) {
  if (D.n_cb_jobs > 0) {
    double by = 0.0, fl = 0.0;
    for (int k : D.cb_job_prob) {
      BatchProblem &b = c->bi[k];
      if (b.co.pending) {
        PMG_HIP_CHECK(hipStreamWaitEvent(c->stream, b.co.ev_done, 0));
        b.co.pending = false;
      }
      by += coarse_visit_bytes(b.co);
      fl += coarse_visit_flops(b.co);
    }
    Timed tm(c, &c->tm.coarse_solve, by, fl);
    launch_coarse_block_batch<T>(D.cb_jobs, D.n_cb_jobs, CbGrid{D.cb_grid[0], D.cb_grid[1], D.cb_grid[2]}, f, u,
                                 c->stream);
  }
  for (int k : D.coarse_insts) {
    if (!c->b_live.empty() && !c->b_live[k]) continue;   // finished: left out, its result is unused
    BatchProblem &b = c->bi[k];
    if (b.in_job) continue;
    if (b.co.pending) {
      PMG_HIP_CHECK(hipStreamWaitEvent(c->stream, b.co.ev_done, 0));
      b.co.pending = false;
    }
    Timed tm(c, &c->tm.coarse_solve, coarse_visit_bytes(b.co), coarse_visit_flops(b.co));
    launch_gather_d<T>(b.cf, f, b.cidx, b.n_dof_coarse, c->stream);
    coarse_solve_any(b.co, b.cl.op64(), b.cf, b.cu, c->stream);
    launch_scatter_d<T>(u, b.cu, b.cidx, b.n_dof_coarse, c->stream);
  }
}

template <typename T>
struct Vec {
  static T *sol(DevLevel &D) { return D.sol; }
  static T *rhs(DevLevel &D) { return D.rhs; }
  static T *aux0(DevLevel &D) { return D.aux0; }
  static T *aux1(DevLevel &D) { return D.aux1; }
  static T *rloc(DevLevel &D) { return D.rloc; }
};
template <>
struct Vec<float> {
  static float *sol(DevLevel &D) { return D.sol32; }
  static float *rhs(DevLevel &D) { return D.rhs32; }
  static float *aux0(DevLevel &D) { return D.aux0_32; }
  static float *aux1(DevLevel &D) { return D.aux1_32; }
  static float *rloc(DevLevel &D) { return D.rloc32; }
};

// TNSE_MultiGrid::Cycle (NSE_MultiGrid.C:434-1779).  Defects whose only consumer is the
// residual printout (:503, :1550, :1753 and the coarse-level one) are skipped: they do not
// change the iterate.
template <typename T>
// zero: the level starts from u = 0 (its right-hand side is new)
void cycle(pmg_ctx *c, int i, bool zero) {
  DevLevel &D = c->L[i];
  T *u = Vec<T>::sol(D), *f = Vec<T>::rhs(D), *d = Vec<T>::aux0(D), *aux = Vec<T>::aux1(D);
  // the zero start: with pre-smoothing, one pass that also does the first sweep's start (the
  // Dirichlet slice from f, the interleaved copy); a single problem's coarse solve writes
  // all of u, so it needs none.  (Not on merged levels where problems are solved directly:
  // their solution is written into u after the start, and the sweeps' interleaved copy must
  // be built from it.)
  bool started = false;
  if (zero) {
    if (i > 0 && c->prm.sc_pre_smooth_saddle > 0 && !D.nodal && D.coarse_insts.empty()) {
      launch_start_zero<T>(u, f, D.n_udof, D.n_pdof, D.hanging_bound, D.n_dirichlet, sweep_copy<T>(D), c->stream);
      started = true;
    } else if (i > 0 || c->batch) {
      launch_fill<T>(u, T(0), D.n_dof, c->stream);
    }
  }
  // CorrectNodes then SetDirichletNodes (NSE_MultiGrid.C:501-502): both act on the same
  // slice of u and the second overwrites what the first writes, so only the second runs;
  // it is left out where the next write of that slice comes before any read: a single
  // problem's coarse solve writes all of u, the first pre-smoothing sweep starts with it
  const int pre = c->prm.sc_pre_smooth_saddle;
  if ((i == 0 && c->batch) || (i > 0 && pre == 0)) set_dirichlet_t<T>(c, D, u, f);
  if (i == 0) {
    if (c->batch)
      batch_coarse<T>(c, D, u, f);
    else
      coarse_solve_t<T>(c, D, u, f);
    return;
  }
  // merged levels: problems whose coarsest level this is are solved here; their cells are
  // not in the color lists and they have no coarser level, so the rest of the cycle leaves
  // them unchanged (their prolongated correction rows are empty)
  if (!D.coarse_insts.empty()) batch_coarse<T>(c, D, u, f);
  DevLevel &C = c->L[i - 1];
  // the last pre-smoothing sweep's closing SetDirichletNodes is overwritten by the
  // CorrectNodes right after it
  for (int j = 0; j < pre; ++j)
    vanka_sweep<T>(c, D, u, f, aux, Vec<T>::rloc(D), j + 1 == pre, j == 0, j + 1 < pre, j == 0 && started);
  correct_nodes_t<T>(c, D, u);
  mg_defect<T>(c, D, u, f, d, nullptr);
  // CorrectDefect of the coarse right-hand side: its Dirichlet rows zeroed by the
  // restriction itself, then the pressure projection
  restrict_defect<T>(c, D, C, d, Vec<T>::rhs(C), true);
  project_pressure<T>(c, C, Vec<T>::rhs(C));
  cycle<T>(c, i - 1, true);
  // CorrectNodes of the correction and of u (NSE_MultiGrid.C:1548, 1580) write only the
  // Dirichlet slice of u, which the first post-smoothing sweep sets from f before reading
  // it: with post-smoothing they are left out (same results)
  const int post = c->prm.sc_post_smooth_saddle;
  const T damp = T(c->prm.sc_gmg_damp_factor_saddle);
  if (post > 0 && !D.nodal) {
    // the prolongation, the update of u and the first post-smoothing sweep's start in one
    // pass over the fine rows
    {
      Timed tm(c, &c->tm.transfer, D.transfer_bytes, 0.0);
      const bool sub = D.act_lists;
      launch_prolong_update<T>(D.pu, D.pp, Vec<T>::sol(C), C.n_udof, u, D.n_udof, damp, f, D.hanging_bound,
                               D.n_dirichlet, sweep_copy<T>(D), c->stream, sub ? D.act_u : nullptr, D.n_act_u,
                               sub ? D.act_p : nullptr, D.n_act_p);
    }
    for (int j = 0; j < post; ++j)
      vanka_sweep<T>(c, D, u, f, aux, Vec<T>::rloc(D), j + 1 == post, j == 0, true, j == 0);
    return;
  }
  prolongate<T>(c, D, C, Vec<T>::sol(C), d);
  if (!post) correct_nodes_t<T>(c, D, d);
  launch_axpy<T>(u, damp, d, D.n_dof, c->stream);
  if (!post) correct_nodes_t<T>(c, D, u);
  for (int j = 0; j < post; ++j) vanka_sweep<T>(c, D, u, f, aux, Vec<T>::rloc(D), j + 1 == post, j == 0);
}

// the main stream waits for coarse factorizations still running on the second stream
void resolve_pending(pmg_ctx *c) {
  auto wait = [&](Coarse &C) {
    if (!C.pending) return;
    PMG_HIP_CHECK(hipStreamWaitEvent(c->stream, C.ev_done, 0));
    C.pending = false;
  };
  wait(c->coarse);
  for (auto &b : c->bi) wait(b.co);
}
bool coarse_pending(const pmg_ctx *c) {
  if (c->coarse.pending) return true;
  for (const auto &b : c->bi)
    if (b.co.pending) return true;
  return false;
}

void drop_graph(pmg_ctx *c) {
  if (c->vexec) (void)hipGraphExecDestroy(c->vexec);
  if (c->vgraph) (void)hipGraphDestroy(c->vgraph);
  c->vexec = nullptr;
  c->vgraph = nullptr;
}

// one V-cycle on the finest level's own vectors: right-hand side already in rhs (rhs32),
// zero start, result in sol (sol32)
void vcycle_body(pmg_ctx *c) {
  DevLevel &F = c->L.back();
  const int top = int(c->L.size()) - 1;
  if (!mixed(c))
    cycle<double>(c, top, true);
  else
    cycle<float>(c, top, true);
}

// Capture vcycle_body into a graph.  Every launch of the cycle is a kernel (fills and
// counter resets included), its buffers are fixed per context, and the device-side
// generation counters of the coarse sweeps advance on every replay.
void capture_vcycle(pmg_ctx *c) {
  drop_graph(c);
  PMG_HIP_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  try {
    vcycle_body(c);
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(c->stream, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  PMG_HIP_CHECK(hipStreamEndCapture(c->stream, &c->vgraph));
  PMG_HIP_CHECK(hipGraphInstantiate(&c->vexec, c->vgraph, nullptr, nullptr, 0));
}

// TMultiGridIte::Iterate with zero start (MultiGridIte.C:76-130): z = M^{-1} v, device vectors
void vcycle_dev(pmg_ctx *c, const double *v, double *z) {
  DevLevel &F = c->L.back();
  if (!mixed(c))
    launch_copy<double>(F.rhs, v, F.n_dof, c->stream);
  else
    launch_d2f(F.rhs32, v, F.n_dof, c->stream);
  if (c->prm.use_graph && !c->profiling && !coarse_pending(c)) {
    if (!c->vexec) capture_vcycle(c);
    PMG_HIP_CHECK(hipGraphLaunch(c->vexec, c->stream));
  } else {
    // eager while a coarse factorization still runs on the second stream (the first cycle
    // after a set-up): the main stream then waits for it at the coarse visit, and the
    // smoothing and restrictions of the finer levels before that visit overlap it (a graph
    // launch would have to wait for it before its first kernel)
    vcycle_body(c);
    resolve_pending(c);
  }
  if (!mixed(c))
    launch_copy<double>(z, F.sol, F.n_dof, c->stream);
  else
    launch_f2d(z, F.sol32, F.n_dof, c->stream);
}

// ------------------------------------------------------------------ set-up
// value-dependent set-up: the dense coarse LU runs on a second stream while the batched
// patch LUs of all finer levels occupy the main stream
// dense coarse system of nr unknowns: allocated on first use or when its size changes
void coarse_alloc(pmg_ctx *c, Coarse &C, int nr) {
  if (C.a && C.n != nr) {
    for (void *p : {(void *)C.a, (void *)C.f, (void *)C.ipiv, (void *)C.sigma, (void *)C.b, (void *)C.y,
                    (void *)C.x, (void *)C.ubuf, (void *)C.swp, (void *)C.sync, (void *)C.flags})
      dfree(c, p);
    C.a = nullptr;
  }
  if (!C.a) {
    C.n = nr;
    C.a = dalloc<double>(c, size_t(C.n) * C.n);
    C.f = dalloc<double>(c, size_t(C.n) * C.n);
    C.ipiv = dalloc<int>(c, C.n);
    C.sigma = dalloc<int>(c, C.n);
    C.b = dalloc<double>(c, C.n);
    C.y = dalloc<double>(c, C.n);
    C.x = dalloc<double>(c, C.n);
    C.ubuf = dalloc<double>(c, size_t(C.n) * 16);
    C.swp = dalloc<int>(c, 4 * 16 + 1);
    const int nblk = (C.n + 63) / 64;
    C.sync = dalloc<unsigned>(c, 4);
    C.flags = dalloc<unsigned>(c, 2 * size_t(nblk));
    PMG_HIP_CHECK(hipMemsetAsync(C.sync, 0, 4 * sizeof(unsigned), c->stream2));
    PMG_HIP_CHECK(hipMemsetAsync(C.flags, 0, 2 * size_t(nblk) * sizeof(unsigned), c->stream2));
    if (!C.err) {
      C.err = dalloc<int>(c, 1);
      PMG_HIP_CHECK(hipMemsetAsync(C.err, 0, sizeof(int), c->stream2));
    }
    if (!C.info) C.info = dalloc<int>(c, 1);
  }
  if (!C.ev_start) {
    PMG_HIP_CHECK(hipEventCreate(&C.ev_start));
    PMG_HIP_CHECK(hipEventCreate(&C.ev_done));
  }
}

// buffers of the block form for m velocity and q pressure unknowns (reallocated when they change)
void coarse_block_alloc(pmg_ctx *c, CoarseBlock &B, int m, int q) {
  if (B.A && (B.m != m || B.q != q)) {
    for (void *p : {(void *)B.A, (void *)B.X, (void *)B.B, (void *)B.V, (void *)B.S, (void *)B.P1, (void *)B.rr,
                    (void *)B.y, (void *)B.p, (void *)B.sigA, (void *)B.sigS, (void *)B.iota})
      dfree(c, p);
    B.A = nullptr;
  }
  if (B.A) return;
  B.m = m;
  B.q = q;
  const size_t M = size_t(m), Q = size_t(q);
  B.A = dalloc<double>(c, M * M);
  B.X = dalloc<double>(c, M * (M + 3 * Q));
  B.B = dalloc<double>(c, Q * 3 * M);
  B.V = dalloc<double>(c, Q * 3 * M);
  B.S = dalloc<double>(c, Q * Q);
  B.P1 = dalloc<double>(c, Q * (3 * M + Q));
  B.rr = dalloc<double>(c, 3 * M + Q);
  B.y = dalloc<double>(c, 3 * M);
  B.p = dalloc<double>(c, Q);
  B.sigA = dalloc<int>(c, M);
  B.sigS = dalloc<int>(c, Q);
  B.iota = dalloc<int>(c, 3 * M + Q);
  std::vector<int> id(3 * M + Q);
  std::iota(id.begin(), id.end(), 0);
  h2d(c, B.iota, id.data(), id.size());
  PMG_HIP_CHECK(hipStreamSynchronize(c->stream));
  if (!B.info) B.info = dalloc<int>(c, 2);
}

// assemble and factor the reduced coarse system of level D on the second stream: in the
// block form when the smoother runs in mixed precision on the shared velocity block
// (FP64 inverses of the blocks, each visit two matrix-vector phases), otherwise as one
// dense LU whose factors equal dgetf2's (the FP64 form, parity with SolveExact)
void start_coarse_factor(pmg_ctx *c, const DevLevel &D, Coarse &C, int pin) {
  const int m = D.n_active, q = D.n_pdof - pin;
  const bool block = mixed(c) && D.a_shared && coarse_block_fits(m, q);
  if (block) {
    if (!C.ev_start) {
      PMG_HIP_CHECK(hipEventCreate(&C.ev_start));
      PMG_HIP_CHECK(hipEventCreate(&C.ev_done));
    }
    if (!C.info) C.info = dalloc<int>(c, 1);
    coarse_block_alloc(c, C.cb, m, q);
    if (!C.block) drop_graph(c);   // the captured V-cycle holds the other form's kernels
    C.block = true;
    C.n = 3 * m + q;
    C.cb.pin = pin;
  } else {
    if (C.block) drop_graph(c);
    C.block = false;
    coarse_alloc(c, C, 3 * D.n_active + D.n_pdof - pin);
  }
  C.N = D.n_udof; C.na = D.n_active; C.np = D.n_pdof; C.pin = pin;
  PMG_HIP_CHECK(hipEventRecord(C.ev_start, c->stream2));
  if (block) {
    PMG_HIP_CHECK(hipMemsetAsync(C.info, 0, sizeof(int), c->stream2));
    launch_cb_build(D, C.cb, c->stream2);
    coarse_block_factor(C.cb, c->stream2);
  } else {
    launch_coarse_build(D, C, c->stream2);
    coarse_lu_factor(C, c->stream2);
  }
  PMG_HIP_CHECK(hipEventRecord(C.ev_done, c->stream2));
  C.pending = true;
  C.info_pending = true;
}

void setup_factors(pmg_ctx *c) {
  check_deferred(c);
  auto t0 = std::chrono::steady_clock::now();
  std::vector<Coarse *> todo;
  std::vector<const DevLevel *> todo_lvl;
  std::vector<int> todo_pin;
  if (!c->batch) {
    if (!c->coarse.ready) {
      todo.push_back(&c->coarse);
      todo_lvl.push_back(&c->L[0]);
      todo_pin.push_back(c->prm.internal_project_pressure ? 1 : 0);
    }
  } else {
    for (auto &b : c->bi)
      if (!b.co.ready) {
        todo.push_back(&b.co);
        todo_lvl.push_back(&b.cl);
        todo_pin.push_back(b.project);
      }
  }
  if (!todo.empty()) {
    PMG_HIP_CHECK(hipEventRecord(c->evs, c->stream));
    PMG_HIP_CHECK(hipStreamWaitEvent(c->stream2, c->evs, 0));
    for (size_t k = 0; k < todo.size(); ++k) start_coarse_factor(c, *todo_lvl[k], *todo[k], todo_pin[k]);
  }
  std::vector<std::vector<int>> infos(c->L.size());
  for (size_t l = 1; l < c->L.size(); ++l) {
    DevLevel &D = c->L[l];
    if (D.factors_ready) continue;
    const hipStream_t ps = c->stream;
    if (D.nodal) {   // nodal Vanka assembles and solves its patches at every visit
      D.factors_ready = true;
      continue;
    }
    if (D.minv32 || D.minv16) {
      // the block form needs the shared velocity block; the captured V-cycle launches the
      // local solves of the form it was captured with
      // (levels with up to 2048 patches per color: the full inverse, whose single-phase
      // product is faster at the latency floor of small launches)
      const bool schur = D.a_shared && D.max_cells_per_color > 2048;
      if (D.schur != schur) drop_graph(c);
      D.schur = schur;
      // refinement steps: the FP32 patch matrices are stored beside the inverse data
      float *kref = nullptr;
      if (c->prm.bf16_refine_steps > 0) {
        const size_t need = size_t(D.n_cells) * (schur ? KREF_BLK : KREF_FULL);
        if (D.kref_len < need) {
          drop_graph(c);
          dfree(c, D.kref);
          D.kref = dalloc<float>(c, need);
          D.kref_len = need;
        }
        kref = D.kref;
      }
      if (D.schur)
        launch_patch_schur(D, D.patch_info, ps, kref);
      else
        launch_patch_inv(D, D.patch_info, ps, kref);
    } else {
      launch_patch_lu(D, D.patch_info, ps);
    }
    infos[l].resize(D.n_cells);
    PMG_HIP_CHECK(hipMemcpyAsync(infos[l].data(), D.patch_info, sizeof(int) * D.n_cells, hipMemcpyDeviceToHost, ps));
  }
  PMG_HIP_CHECK(hipStreamSynchronize(c->stream));
  const double t_patch = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  for (size_t l = 1; l < c->L.size(); ++l) {
    DevLevel &D = c->L[l];
    if (D.factors_ready) continue;
    c->tm.patch_lu.calls += 1;
    c->tm.patch_lu.bytes += double(D.n_cells) * (NLOC * NLOC * 8.0) + D.defect_bytes[0][0];
    c->tm.patch_lu.flops += double(D.n_cells) * (2.0 / 3.0) * NLOC * NLOC * NLOC;
    for (int p = 0; p < D.n_cells; ++p)
      if (infos[l][p])
        throw Fail{PMG_ERR_SINGULAR_PATCH, "zero pivot " + std::to_string(infos[l][p]) +
                                               " in the patch at color position " + std::to_string(p) +
                                               " of level " + std::to_string(l)};
    D.factors_ready = true;
  }
  c->tm.patch_lu.ms += t_patch;
  for (Coarse *C : todo) C->ready = true;   // usable once ev_done completes; checked in check_deferred
  if (c->batch) build_cb_jobs(c);
  c->setup_done = true;
}

void ensure_fgmres(pmg_ctx *c) {
  if (!c->batch && c->xa64_n != c->L.back().n_udof) {
    dfree(c, c->xa64);
    c->xa64_n = c->L.back().n_udof;
    c->xa64 = dalloc<double>(c, 4 * size_t(c->xa64_n));
  }
  const int m = std::max(1, std::min(c->prm.sc_gmres_restart > 0 ? c->prm.sc_gmres_restart : 10,
                                     std::max(1, c->prm.sc_lin_maxit_saddle)));
  if (m <= c->restart_alloc) return;
  const long n = c->L.back().n_dof;
  dfree(c, c->V);
  dfree(c, c->Z);
  dfree(c, c->W);
  dfree(c, c->Hdev);
  dfree(c, c->coef);
  c->V = dalloc<double>(c, size_t(m + 1) * n);
  c->Z = dalloc<double>(c, size_t(m + 1) * n);
  c->W = dalloc<double>(c, n);
  // (batch) the FGMRES steps leave the parts of finished problems untouched once they are
  // left out, so those parts only ever hold zeros or earlier finite values
  launch_fill<double>(c->V, 0.0, long(m + 1) * n, c->stream);
  launch_fill<double>(c->Z, 0.0, long(m + 1) * n, c->stream);
  launch_fill<double>(c->W, 0.0, n, c->stream);
  c->Hdev = dalloc<double>(c, m + 2);
  c->coef = dalloc<double>(c, m + 1);
  if (2 * m + 8 > c->hpin_len) {
    if (c->hpin) (void)hipHostFree(c->hpin);
    c->hpin = nullptr;
    PMG_HIP_CHECK(hipHostMalloc(reinterpret_cast<void **>(&c->hpin), sizeof(double) * (2 * m + 8)));
    c->hpin_len = 2 * m + 8;
  }
  c->restart_alloc = m;
}

double dev_norm2(pmg_ctx *c, const double *x, long n) {
  launch_dot(x, x, n, c->Hdev, c->red, c->stream);
  PMG_HIP_CHECK(hipMemcpyAsync(c->hpin, c->Hdev, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  PMG_HIP_CHECK(hipStreamSynchronize(c->stream));
  return c->hpin[0];
}

void defect_nse4_dev(pmg_ctx *c, const double *x, const double *b, double *r) {
  DevLevel &F = c->L.back();
  Timed tm(c, &c->tm.matvec, F.defect_bytes[F.a_shared][0], 0.0);
  launch_type4<double>(F.op64(), 1, x, b, r, fused_l20(c), nullptr, c->red, c->stream);
  if (c->batch) project_pressure<double>(c, F, r);
}

// GeneratePlaneRotation / ApplyPlaneRotation (FgmresIte.C:199-226)
void gen_rot(double dx, double dy, double &cs, double &sn) {
  if (dy == 0.0) {
    cs = 1.0; sn = 0.0;
  } else if (std::fabs(dy) > std::fabs(dx)) {
    const double temp = dx / dy;
    sn = 1.0 / std::sqrt(1.0 + temp * temp);
    cs = temp * sn;
  } else {
    const double temp = dy / dx;
    cs = 1.0 / std::sqrt(1.0 + temp * temp);
    sn = temp * cs;
  }
}
void app_rot(double &dx, double &dy, double cs, double sn) {
  const double temp = cs * dx + sn * dy;
  dy = -sn * dx + cs * dy;
  dx = temp;
}

// UpdateGmresIterate (FgmresIte.C:228-247) back-solve on the host, then x += Z y on device
void update_iterate(pmg_ctx *c, double *x, int k, std::vector<std::vector<double>> &H, std::vector<double> &s) {
  for (int i = k; i >= 0; i--) {
    s[i] /= H[i][i];
    for (int j = i - 1; j >= 0; j--) s[j] -= H[j][i] * s[i];
  }
  const long n = c->L.back().n_dof;
  for (int j = 0; j <= k; ++j) c->hpin[j] = s[j];
  PMG_HIP_CHECK(hipMemcpyAsync(c->coef, c->hpin, sizeof(double) * (k + 1), hipMemcpyHostToDevice, c->stream));
  launch_lincomb(x, c->Z, n, c->coef, k + 1, n, c->stream);
}

// TFgmresIte::Iterate, flexible branch (FgmresIte.C:250-677), device vectors
int fgmres_dev(pmg_ctx *c, double *sol, const double *rhs, int *iters, double *final_res) {
  ensure_fgmres(c);
  const pmg_params &P = c->prm;
  DevLevel &F = c->L.back();
  const long n = F.n_dof;
  int maxit = P.sc_lin_maxit_saddle;
  int restart = P.sc_gmres_restart;
  if (restart > maxit) restart = maxit;
  if (restart <= 0) restart = 10;
  const double res_norm_min = P.sc_lin_res_norm_min_saddle, red_factor = P.sc_lin_red_factor_saddle;
  const int minit = P.sc_minit;
  std::vector<double> s(restart + 1), cs(restart + 1), sn(restart + 1);
  std::vector<std::vector<double>> H(restart + 1, std::vector<double>(restart, 0.0));
  auto v = [&](int i) { return c->V + size_t(i) * n; };
  auto z = [&](int i) { return c->Z + size_t(i) * n; };
  c->hist.clear();
  defect_nse4_dev(c, sol, rhs, v(0));
  double res = std::sqrt(dev_norm2(c, v(0), n));
  double res0 = res, beta = res;
  int maxite = maxit;
  if (!std::isfinite(res)) {
    *iters = 0; *final_res = res;
    throw Fail{PMG_ERR_NONFINITE, "non-finite initial FGMRES residual"};
  }
  if (beta <= res_norm_min) {
    if (minit == 0 || beta < 1e-20) {
      *iters = 0; *final_res = res;
      return PMG_OK;
    }
    maxite = minit;
  }
  c->hist.push_back(beta);
  int j = 1;
  while (j <= maxite) {
    launch_scal(v(0), v(0), 1.0 / beta, n, c->stream);
    std::fill(s.begin(), s.end(), 0.0);
    s[0] = beta;
    for (int i = 0; i < restart && j <= maxite; i++, j++) {
      for (int k = 0; k < std::max(1, P.sc_amg_prec_it); ++k) vcycle_dev(c, v(i), z(i));
      {
        // with the shared velocity block the operator gathers an interleaved copy of the
        // input's velocity (one pass more: 3 x 8 B read, 4 x 8 B written per velocity dof)
        OpView<double> o = F.op64();
        const bool il = F.a_shared && c->xa64;
        Timed tm(c, &c->tm.matvec, F.matvec_bytes[F.a_shared] + (il ? 56.0 * F.n_udof : 0.0), 0.0);
        if (il) {
          launch_interleave3(c->xa64, z(i), F.n_udof, c->stream);
          o.xa = c->xa64;
        }
        launch_type4<double>(o, 0, z(i), nullptr, c->W, false, nullptr, c->red, c->stream);
      }
      // modified Gram-Schmidt, H[k][i] -> Hdev[k], ||w||^2 -> Hdev[i+1]
      launch_dot(c->W, v(0), n, c->Hdev, c->red, c->stream);
      for (int k = 0; k <= i; ++k) {
        const double *vn = (k < i) ? v(k + 1) : c->W;
        launch_axpy_dev_dot(c->W, c->Hdev + k, v(k), vn, n, c->Hdev + k + 1, c->red, c->stream);
      }
      PMG_HIP_CHECK(hipMemcpyAsync(c->hpin, c->Hdev, sizeof(double) * (i + 2), hipMemcpyDeviceToHost, c->stream));
      PMG_HIP_CHECK(hipStreamSynchronize(c->stream));
      for (int k = 0; k <= i; ++k) H[k][i] = c->hpin[k];
      H[i + 1][i] = std::sqrt(c->hpin[i + 1]);
      if (H[i + 1][i] != 0)
        launch_scal(v(i + 1), c->W, 1 / H[i + 1][i], n, c->stream);
      else
        launch_copy<double>(v(i + 1), c->W, n, c->stream);
      for (int k = 0; k < i; k++) app_rot(H[k][i], H[k + 1][i], cs[k], sn[k]);
      gen_rot(H[i][i], H[i + 1][i], cs[i], sn[i]);
      app_rot(H[i][i], H[i + 1][i], cs[i], sn[i]);
      app_rot(s[i], s[i + 1], cs[i], sn[i]);
      res = std::fabs(s[i + 1]);
      c->hist.push_back(res);
      if (!std::isfinite(res)) {
        *iters = j; *final_res = res;
        throw Fail{PMG_ERR_NONFINITE, "non-finite FGMRES residual at iteration " + std::to_string(j)};
      }
      if ((j >= minit) && ((res < res_norm_min) || (res < red_factor * res0))) {
        update_iterate(c, sol, i, H, s);
        PMG_HIP_CHECK(hipStreamSynchronize(c->stream));
        *iters = j; *final_res = res;
        return PMG_OK;
      }
      if (res > P.sc_div_factor * res0) {
        *iters = j; *final_res = res;
        throw Fail{PMG_ERR_DIVERGED, "fgmres iteration diverges (res " + std::to_string(res) + ")"};
      }
    }
    update_iterate(c, sol, restart - 1, H, s);
    defect_nse4_dev(c, sol, rhs, v(0));
    res = res0 = beta = std::sqrt(dev_norm2(c, v(0), n));
    c->hist.push_back(res);
  }
  if (j > maxite) j = maxite;
  *iters = j; *final_res = res;
  if (res > res_norm_min && res > res0 * red_factor) return PMG_ERR_NOT_CONVERGED;
  return PMG_OK;
}

// ------------------------------------------------------------------ batched FGMRES
void ensure_hpin(pmg_ctx *c, long n) {
  if (n <= c->hpin_len) return;
  if (c->hpin) (void)hipHostFree(c->hpin);
  c->hpin = nullptr;
  PMG_HIP_CHECK(hipHostMalloc(reinterpret_cast<void **>(&c->hpin), sizeof(double) * n));
  c->hpin_len = int(n);
}

// per-problem <x, y> -> out[b * ldo] (device), deterministic
void seg_dot(pmg_ctx *c, const double *x, const double *y, double *out, int ldo) {
  launch_seg_dot(x, y, c->seg_ch_act, c->n_seg_ch_act, c->seg_part, c->stream);
  launch_seg_reduce(c->seg_part, c->seg_cptr_act, int(c->bi.size()), out, ldo, c->stream);
}

struct BState {
  int st = PMG_OK;
  bool active = true;
  int it = 0;
  double res = 0, res0 = 0, beta = 0;
  std::vector<std::vector<double>> H;
  std::vector<double> s, cs, sn;
};

// UpdateGmresIterate back substitution (FgmresIte.C:228-247): y (in s) for columns 0..k
void back_solve(std::vector<std::vector<double>> &H, std::vector<double> &s, int k) {
  for (int i = k; i >= 0; i--) {
    s[i] /= H[i][i];
    for (int j = i - 1; j >= 0; j--) s[j] -= H[j][i] * s[i];
  }
}

// Leave finished problems out of the merged sweeps, coarse solves, operator rows,
// transfers and FGMRES vector steps.  The merged levels are block diagonal, so the rows of
// the unfinished problems never read a finished problem's entries, whatever those hold;
// each finished problem's solution was combined when it stopped.  The lists are only
// rebuilt (and the captured graph dropped) once the live patches fall below 3/4 of those
// at the last rebuild, or when every problem is live again.
void batch_compact(pmg_ctx *c, const std::vector<char> &live) {
  long cells = 0, total = 0;
  for (size_t l = 1; l < c->L.size(); ++l) {
    const DevLevel &D = c->L[l];
    total += D.n_cells;
    for (int q = 0; q < D.n_cells; ++q) cells += live[D.h_pos_problem[q]] ? 1 : 0;
  }
  const bool all = cells == total;
  if (all) {
    bool any = c->seg_ch_act != c->seg_ch;
    for (const auto &D : c->L) any = any || D.compacted || D.act_lists;
    c->b_live.assign(live.size(), 1);
    c->b_live_cells = total;
    if (!any) return;
    for (auto &D : c->L) D.compacted = D.act_lists = false;
    c->seg_ch_act = c->seg_ch;
    c->n_seg_ch_act = c->n_seg_ch;
    c->seg_cptr_act = c->seg_cptr;
    build_cb_jobs(c);
    drop_graph(c);
    return;
  }
  if (c->b_live_cells > 0 && cells * 4 > c->b_live_cells * 3) return;
  // the operator rows and transfer rows of the unfinished problems on every level (the
  // levels are block diagonal: those rows read only their own problems' entries)
  for (auto &D : c->L) {
    if (D.h_udof_problem.empty()) continue;
    if (!D.act_u) {
      D.act_u = dalloc<int>(c, D.n_udof);
      D.act_p = dalloc<int>(c, D.n_pdof);
      D.act_oprows = dalloc<int>(c, size_t(D.n_udof) + D.n_pdof);
    }
    std::vector<int> u, p, rows;
    for (const auto &r : D.h_urun)
      if (live[r.prob])
        for (long i = r.b; i < r.e; ++i) u.push_back(int(i));
    for (const auto &r : D.h_prun)
      if (live[r.prob])
        for (long i = r.b; i < r.e; ++i) p.push_back(int(i));
    rows.reserve(u.size() + p.size());
    rows = u;
    for (int i : p) rows.push_back(D.n_udof + i);
    h2d(c, D.act_u, u.data(), u.size());
    h2d(c, D.act_p, p.data(), p.size());
    h2d(c, D.act_oprows, rows.data(), rows.size());
    PMG_HIP_CHECK(hipStreamSynchronize(c->stream));
    D.n_act_u = int(u.size());
    D.n_act_p = int(p.size());
    D.act_lists = true;
  }
  {
    // the FGMRES chunks of the unfinished problems, each problem's in their original order
    // (so its partial sums and their reduction order stay those of the full table)
    std::vector<int4> ch;
    std::vector<int> cp(1, 0);
    for (size_t b = 0; b < c->bi.size(); ++b) {
      if (live[b])
        for (int k = c->h_seg_cptr[b]; k < c->h_seg_cptr[b + 1]; ++k) ch.push_back(c->h_seg_ch[k]);
      cp.push_back(int(ch.size()));
    }
    if (!c->seg_ch_live) {
      c->seg_ch_live = dalloc<int4>(c, size_t(c->n_seg_ch));
      c->seg_cptr_live = dalloc<int>(c, c->bi.size() + 1);
    }
    h2d(c, c->seg_ch_live, ch.data(), ch.size());
    h2d(c, c->seg_cptr_live, cp.data(), cp.size());
    PMG_HIP_CHECK(hipStreamSynchronize(c->stream));
    c->seg_ch_act = c->seg_ch_live;
    c->n_seg_ch_act = int(ch.size());
    c->seg_cptr_act = c->seg_cptr_live;
  }
  for (size_t l = 1; l < c->L.size(); ++l) {
    DevLevel &D = c->L[l];
    if (!D.act_items) {
      D.act_items = dalloc<int>(c, D.h_res_items.size());
      D.act_rows = dalloc<int>(c, D.h_res_items.size());
      D.act_desc = dalloc<int>(c, 4 * D.h_res_items.size());
      D.act_slots = dalloc<int>(c, D.n_cells);
    }
    std::vector<int> items, rows, slots;
    items.reserve(D.h_res_items.size());
    rows.reserve(D.h_res_items.size());
    D.act_items_off.assign(1, 0);
    D.act_slots_off.assign(1, 0);
    for (int col = 0; col < D.n_colors; ++col) {
      const int p0 = D.ptr_colors[col], p1 = D.ptr_colors[col + 1];
      for (long k = D.h_irun_off[col]; k < D.h_irun_off[col + 1]; ++k) {
        const auto &r = D.h_irun[k];
        if (!live[r.prob]) continue;
        items.insert(items.end(), D.h_res_items.begin() + r.b, D.h_res_items.begin() + r.e);
        rows.insert(rows.end(), D.h_res_rows.begin() + r.b, D.h_res_rows.begin() + r.e);
      }
      for (int sl = 0; sl < p1 - p0; ++sl)
        if (live[D.h_pos_problem[p0 + sl]]) slots.push_back(sl);
      D.act_items_off.push_back(long(items.size()));
      D.act_slots_off.push_back(long(slots.size()));
    }
    h2d(c, D.act_items, items.data(), items.size());
    h2d(c, D.act_rows, rows.data(), rows.size());
    launch_res_desc(D, D.act_items, D.act_rows, long(items.size()), D.act_desc, c->stream);
    h2d(c, D.act_slots, slots.data(), slots.size());
    PMG_HIP_CHECK(hipStreamSynchronize(c->stream));
    D.compacted = true;
  }
  c->b_live = live;
  c->b_live_cells = cells;
  build_cb_jobs(c);
  drop_graph(c);
}

// TFgmresIte::Iterate (FgmresIte.C:250-677) for every problem of a batch at once: the
// problems share the iteration counter j and the restart cycle; the Arnoldi scalars,
// Givens rotations, stopping and divergence tests are each problem's own.  A problem that
// stops gets zero basis vectors from then on (its V-cycle input is zero, so is the output).
void fgmres_batch(pmg_ctx *c, std::vector<int> &iters, std::vector<double> &final_res, std::vector<int> &status) {
  ensure_fgmres(c);
  const pmg_params &P = c->prm;
  DevLevel &F = c->L.back();
  const long n = F.n_dof;
  const int nb = int(c->bi.size());
  int maxit = P.sc_lin_maxit_saddle;
  int restart = P.sc_gmres_restart;
  if (restart > maxit) restart = maxit;
  if (restart <= 0) restart = 10;
  const double res_norm_min = P.sc_lin_res_norm_min_saddle, red_factor = P.sc_lin_red_factor_saddle;
  const int ldh = restart + 2, ldc = restart + 1;
  if (restart > c->b_restart_alloc) {
    dfree(c, c->bH);
    dfree(c, c->bcoef);
    dfree(c, c->bmask);
    c->bH = dalloc<double>(c, size_t(nb) * ldh);
    c->bcoef = dalloc<double>(c, size_t(nb) * (ldc + 2));
    c->bmask = dalloc<int>(c, nb);
    c->b_restart_alloc = restart;
  }
  ensure_hpin(c, long(nb) * (ldh + ldc + 4));
  double *hH = c->hpin, *hS = hH + size_t(nb) * ldh, *hC = hS + nb, *hN = hC + size_t(nb) * ldc;
  int *hM = reinterpret_cast<int *>(hN + nb);
  double *dS = c->bcoef, *dC = dS + nb, *dN = dC + size_t(nb) * ldc;
  auto v = [&](int i) { return c->V + size_t(i) * n; };
  auto z = [&](int i) { return c->Z + size_t(i) * n; };
  std::vector<BState> B(nb);
  for (auto &b : B) {
    b.H.assign(restart + 1, std::vector<double>(restart, 0.0));
    b.s.assign(restart + 1, 0.0);
    b.cs.assign(restart + 1, 0.0);
    b.sn.assign(restart + 1, 0.0);
  }
  auto norms = [&](const double *x) {
    seg_dot(c, x, x, dN, 1);
    PMG_HIP_CHECK(hipMemcpyAsync(hN, dN, sizeof(double) * nb, hipMemcpyDeviceToHost, c->stream));
    PMG_HIP_CHECK(hipStreamSynchronize(c->stream));
  };
  auto any_active = [&] {
    for (auto &b : B)
      if (b.active) return true;
    return false;
  };
  // x_b += sum_{k <= kmax} y_bk z_k for the problems in hM
  auto combine = [&](int m) {
    PMG_HIP_CHECK(hipMemcpyAsync(dC, hC, sizeof(double) * nb * ldc, hipMemcpyHostToDevice, c->stream));
    PMG_HIP_CHECK(hipMemcpyAsync(c->bmask, hM, sizeof(int) * nb, hipMemcpyHostToDevice, c->stream));
    launch_seg_lincomb(c->bX, c->Z, n, dC, ldc, m, c->bmask, c->seg_ch_act, c->n_seg_ch_act, c->stream);
  };
  std::vector<char> live(nb, 1);
  batch_compact(c, live);
  auto recompact = [&] {
    for (int b = 0; b < nb; ++b) live[b] = B[b].active ? 1 : 0;
    batch_compact(c, live);
  };
  defect_nse4_dev(c, c->bX, c->bB, v(0));
  norms(v(0));
  for (int b = 0; b < nb; ++b) {
    BState &S = B[b];
    S.res = S.res0 = S.beta = std::sqrt(hN[b]);
    if (!std::isfinite(S.res)) {
      S.st = PMG_ERR_NONFINITE;
      S.active = false;
    } else if (S.beta <= res_norm_min || S.beta == 0.0) {
      S.active = false;   // sc_minit == 0: nothing to do (FgmresIte.C:330-337)
    }
  }
  int j = 1;
  while (j <= maxit && any_active()) {
    for (int b = 0; b < nb; ++b) hS[b] = B[b].active ? 1.0 / B[b].beta : 0.0;
    PMG_HIP_CHECK(hipMemcpyAsync(dS, hS, sizeof(double) * nb, hipMemcpyHostToDevice, c->stream));
    launch_seg_scale(v(0), v(0), dS, c->seg_ch_act, c->n_seg_ch_act, c->stream);
    for (auto &S : B) {
      std::fill(S.s.begin(), S.s.end(), 0.0);
      S.s[0] = S.beta;
    }
    for (int i = 0; i < restart && j <= maxit && any_active(); i++, j++) {
      for (int k = 0; k < std::max(1, P.sc_amg_prec_it); ++k) vcycle_dev(c, v(i), z(i));
      {
        Timed tm(c, &c->tm.matvec, F.matvec_bytes[F.a_shared], 0.0);
        launch_type4<double>(F.op64(), 0, z(i), nullptr, c->W, false, nullptr, c->red, c->stream);
      }
      // modified Gram-Schmidt: H[k][i] -> bH[b*ldh + k], ||w||^2 -> bH[b*ldh + i + 1]
      seg_dot(c, c->W, v(0), c->bH, ldh);
      for (int k = 0; k <= i; ++k) {
        const double *vn = (k < i) ? v(k + 1) : c->W;
        launch_seg_axpy_dot(c->W, c->bH + k, ldh, v(k), vn, c->seg_ch_act, c->n_seg_ch_act, c->seg_part, c->stream);
        launch_seg_reduce(c->seg_part, c->seg_cptr_act, nb, c->bH + k + 1, ldh, c->stream);
      }
      PMG_HIP_CHECK(hipMemcpyAsync(hH, c->bH, sizeof(double) * nb * ldh, hipMemcpyDeviceToHost, c->stream));
      PMG_HIP_CHECK(hipStreamSynchronize(c->stream));
      bool finish = false;
      std::fill(hM, hM + nb, 0);
      std::fill(hC, hC + size_t(nb) * ldc, 0.0);
      for (int b = 0; b < nb; ++b) {
        BState &S = B[b];
        hS[b] = 0.0;
        if (!S.active) continue;
        auto &H = S.H;
        for (int k = 0; k <= i; ++k) H[k][i] = hH[size_t(b) * ldh + k];
        H[i + 1][i] = std::sqrt(hH[size_t(b) * ldh + i + 1]);
        for (int k = 0; k < i; k++) app_rot(H[k][i], H[k + 1][i], S.cs[k], S.sn[k]);
        gen_rot(H[i][i], H[i + 1][i], S.cs[i], S.sn[i]);
        const double hnext = H[i + 1][i];   // before the rotation zeroes it
        app_rot(H[i][i], H[i + 1][i], S.cs[i], S.sn[i]);
        app_rot(S.s[i], S.s[i + 1], S.cs[i], S.sn[i]);
        S.res = std::fabs(S.s[i + 1]);
        S.it = j;
        if (!std::isfinite(S.res)) {
          S.st = PMG_ERR_NONFINITE;
          S.active = false;
        } else if (S.res < res_norm_min || S.res < red_factor * S.res0) {
          back_solve(H, S.s, i);
          for (int k = 0; k <= i; ++k) hC[size_t(b) * ldc + k] = S.s[k];
          hM[b] = 1;
          finish = true;
          S.st = PMG_OK;
          S.active = false;
        } else if (S.res > P.sc_div_factor * S.res0) {
          S.st = PMG_ERR_DIVERGED;
          S.active = false;
        } else {
          hS[b] = hnext != 0 ? 1.0 / hnext : 1.0;
        }
      }
      // v_{i+1} = w / H[i+1][i] (a zero norm copies w, as the single-problem path does)
      PMG_HIP_CHECK(hipMemcpyAsync(dS, hS, sizeof(double) * nb, hipMemcpyHostToDevice, c->stream));
      launch_seg_scale(v(i + 1), c->W, dS, c->seg_ch_act, c->n_seg_ch_act, c->stream);
      if (finish) combine(i + 1);
      recompact();
    }
    if (!any_active()) break;
    // end of a restart cycle (or maxit): update with all restart columns, new residual
    std::fill(hM, hM + nb, 0);
    std::fill(hC, hC + size_t(nb) * ldc, 0.0);
    for (int b = 0; b < nb; ++b) {
      BState &S = B[b];
      if (!S.active) continue;
      back_solve(S.H, S.s, restart - 1);
      for (int k = 0; k < restart; ++k) hC[size_t(b) * ldc + k] = S.s[k];
      hM[b] = 1;
    }
    combine(restart);
    defect_nse4_dev(c, c->bX, c->bB, v(0));
    norms(v(0));
    for (int b = 0; b < nb; ++b) {
      BState &S = B[b];
      if (!S.active) continue;
      S.res = S.res0 = S.beta = std::sqrt(hN[b]);
      if (!std::isfinite(S.res)) {
        S.st = PMG_ERR_NONFINITE;
        S.active = false;
      } else if (S.beta == 0.0) {
        S.active = false;
      }
    }
    recompact();
  }
  // every problem live again for later calls on this context
  std::fill(live.begin(), live.end(), 1);
  batch_compact(c, live);
  for (int b = 0; b < nb; ++b) {
    BState &S = B[b];
    if (S.active) {
      S.it = std::min(j, maxit);
      if (S.res > res_norm_min && S.res > S.res0 * red_factor) S.st = PMG_ERR_NOT_CONVERGED;
    }
    iters[b] = S.it;
    final_res[b] = S.res;
    status[b] = S.st;
  }
  PMG_HIP_CHECK(hipStreamSynchronize(c->stream));
}

int guard(pmg_ctx *c, const std::function<void()> &f) {
  try {
    f();
    return PMG_OK;
  } catch (const Fail &e) {
    if (c) c->err = e.msg;
    return e.code;
  } catch (const DeviceError &e) {
    if (c) c->err = e.what;
    return PMG_ERR_DEVICE;
  } catch (const std::bad_alloc &) {
    if (c) c->err = "host allocation failed";
    return PMG_ERR_OUT_OF_MEMORY;
  }
}

DevLevel &level_of(pmg_ctx *c, int l) {
  if (l < 0 || l >= int(c->L.size())) throw Fail{PMG_ERR_INVALID_ARGUMENT, "level out of range"};
  return c->L[l];
}

// copy-in/out helpers for PMG_MEM_HOST
struct HostIO {
  pmg_ctx *c;
  int mem;
  long n;
  HostIO(pmg_ctx *c_, int m, long n_) : c(c_), mem(m), n(n_) {
    if (mem != PMG_MEM_HOST && mem != PMG_MEM_DEVICE) throw Fail{PMG_ERR_INVALID_ARGUMENT, "bad memspace"};
    if (mem == PMG_MEM_HOST) ensure_stage(c, n);
  }
  double *in(int slot, const double *p) {
    if (!p) return nullptr;
    if (mem == PMG_MEM_DEVICE) return const_cast<double *>(p);
    PMG_HIP_CHECK(hipMemcpyAsync(c->stage[slot], p, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    return c->stage[slot];
  }
  void out(int slot, double *p) {
    if (mem == PMG_MEM_DEVICE || !p) return;
    PMG_HIP_CHECK(hipMemcpyAsync(p, c->stage[slot], sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  }
  void sync() { PMG_HIP_CHECK(hipStreamSynchronize(c->stream)); }
};

}  // namespace

// ===================================================================== static scratch
// pmg_vanka_smooth_noalloc keeps its scratch (the local residuals of one color and the
// pre-sweep copy of u for the global relaxation) in this file-scope device array instead
// of buffers the library allocates: 8 MiB per GPU, held until the process ends.  Host
// code uses only the per-device address from hipGetSymbolAddress (scratch_dev), never the
// array's name.  Every call on a GPU shares that GPU's array: calls that overlap on one
// GPU (two contexts, two streams) race.
constexpr size_t NOALLOC_SCRATCH_DOUBLES = size_t(1) << 20;
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

namespace {
std::mutex scratch_mu;
std::vector<double *> scratch_dev;   // device -> address of that device's copy (nullptr: failed)
std::vector<int> scratch_state;      // device -> 0 not looked up, 1 ok, 2 failed

// once per device, with the device current (called from pmg_ctx_create)
void scratch_lookup(int device) {
  std::lock_guard<std::mutex> lk(scratch_mu);
  if (int(scratch_dev.size()) <= device) {
    scratch_dev.resize(device + 1, nullptr);
    scratch_state.resize(device + 1, 0);
  }
  if (scratch_state[device]) return;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  scratch_dev[device] = e == hipSuccess ? static_cast<double *>(p) : nullptr;
  scratch_state[device] = e == hipSuccess ? 1 : 2;
}
double *scratch_of(int device) {
  std::lock_guard<std::mutex> lk(scratch_mu);
  return device < int(scratch_dev.size()) ? scratch_dev[device] : nullptr;
}
}  // namespace

// ===================================================================== C ABI
extern "C" {

void pmg_params_default(pmg_params *p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  // src/General/Database.C:674-749, 871
  p->sc_smoother_saddle = 2;
  p->sc_pre_smooth_saddle = 2;
  p->sc_post_smooth_saddle = 2;
  p->sc_smooth_damp_factor_saddle = 1.0;
  p->sc_smooth_damp_factor_coarse_saddle = 1.0;
  p->sc_gmg_damp_factor_saddle = 1.0;
  p->sc_coarse_smoother_saddle = 17;
  p->internal_project_pressure = 1;
  p->sc_gmres_restart = 10;
  p->sc_lin_maxit_saddle = 10000;
  p->sc_lin_red_factor_saddle = 0.0;
  p->sc_lin_res_norm_min_saddle = 1e-10;
  p->sc_div_factor = 1e10;
  p->sc_minit = 0;
  p->sc_amg_prec_it = 1;
  p->precision = PMG_PREC_FP64;
  p->bf16_refine_steps = 0;
  p->use_graph = 1;
  p->shared_block = 1;
  p->patch_inverse = 1;
}

const char *pmg_status_string(int s) {
  switch (s) {
    case PMG_OK: return "ok";
    case PMG_ERR_INVALID_ARGUMENT: return "invalid argument";
    case PMG_ERR_DEVICE: return "device error";
    case PMG_ERR_OUT_OF_MEMORY: return "out of memory";
    case PMG_ERR_NOT_READY: return "not ready";
    case PMG_ERR_SINGULAR_PATCH: return "singular local patch";
    case PMG_ERR_SINGULAR_COARSE: return "singular coarse matrix";
    case PMG_ERR_NONFINITE: return "non-finite value";
    case PMG_ERR_DIVERGED: return "diverged";
    case PMG_ERR_NOT_CONVERGED: return "not converged";
    case PMG_ERR_UNSUPPORTED: return "unsupported";
    default: return "unknown status";
  }
}

static int check_params(const pmg_params *p, std::string &why) {
  if (p->sc_smoother_saddle < 1 || p->sc_smoother_saddle > 4) { why = "only the Vanka smoothers 1-4"; return PMG_ERR_UNSUPPORTED; }
  if (p->sc_smoother_saddle >= 3 && p->precision != PMG_PREC_FP64) {
    why = "the nodal Vanka smoothers (3, 4) run in FP64 only";
    return PMG_ERR_UNSUPPORTED;
  }
  if (p->sc_coarse_smoother_saddle != 17) { why = "only the dense coarse solve (17)"; return PMG_ERR_UNSUPPORTED; }
  if (p->precision < 0 || p->precision > 2) { why = "bad precision"; return PMG_ERR_INVALID_ARGUMENT; }
  if (p->sc_pre_smooth_saddle < 0 || p->sc_post_smooth_saddle < 0) { why = "negative smoothing steps"; return PMG_ERR_INVALID_ARGUMENT; }
  if (p->bf16_refine_steps < 0 || p->bf16_refine_steps > 3) { why = "bf16_refine_steps must be 0..3"; return PMG_ERR_INVALID_ARGUMENT; }
  if (p->bf16_refine_steps > 0 && (p->precision == PMG_PREC_FP64 || !p->patch_inverse)) {
    why = "local refinement steps apply to the stored patch inverses of the mixed forms (patch_inverse = 1)";
    return PMG_ERR_UNSUPPORTED;
  }
  if (p->sc_gmres_restart <= 0 || p->sc_lin_maxit_saddle < 0) { why = "bad FGMRES restart/maxit"; return PMG_ERR_INVALID_ARGUMENT; }
  return PMG_OK;
}

int pmg_ctx_create(int device, const pmg_params *params, pmg_ctx **out) {
  if (!out) return PMG_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  auto c = std::make_unique<pmg_ctx>();
  if (params) c->prm = *params; else pmg_params_default(&c->prm);
  std::string why;
  if (int st = check_params(&c->prm, why)) return st;
  c->device = device;
  int rc = guard(c.get(), [&] {
    int n = 0;
    PMG_HIP_CHECK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) throw Fail{PMG_ERR_INVALID_ARGUMENT, "no such device"};
    PMG_HIP_CHECK(hipSetDevice(device));
    scratch_lookup(device);
    PMG_HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    PMG_HIP_CHECK(hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
    PMG_HIP_CHECK(hipEventCreateWithFlags(&c->evs, hipEventDisableTiming));
    PMG_HIP_CHECK(hipEventCreate(&c->ev0));
    PMG_HIP_CHECK(hipEventCreate(&c->ev1));
    c->red.partials = dalloc<double>(c.get(), 2 * RED_BLOCKS + 16);
    c->red.counter = dalloc<unsigned>(c.get(), 16);
    c->red.result = dalloc<double>(c.get(), 16);
    PMG_HIP_CHECK(hipMemsetAsync(c->red.counter, 0, 16 * sizeof(unsigned), c->stream));
    PMG_HIP_CHECK(hipHostMalloc(reinterpret_cast<void **>(&c->hpin), sizeof(double) * 256));
    c->hpin_len = 256;
    PMG_HIP_CHECK(hipStreamSynchronize(c->stream));
  });
  if (rc) return rc;
  *out = c.release();
  return PMG_OK;
}

int pmg_ctx_destroy(pmg_ctx *c) {
  if (!c) return PMG_ERR_INVALID_ARGUMENT;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->stream2) (void)hipStreamSynchronize(c->stream2);
  drop_graph(c);
  for (void *p : c->allocs) (void)hipFree(p);
  if (c->hpin) (void)hipHostFree(c->hpin);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->evs) (void)hipEventDestroy(c->evs);
  if (c->coarse.ev_start) (void)hipEventDestroy(c->coarse.ev_start);
  if (c->coarse.ev_done) (void)hipEventDestroy(c->coarse.ev_done);
  for (auto &b : c->bi) {
    if (b.co.ev_start) (void)hipEventDestroy(b.co.ev_start);
    if (b.co.ev_done) (void)hipEventDestroy(b.co.ev_done);
  }
  if (c->stream2) (void)hipStreamDestroy(c->stream2);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return PMG_OK;
}

const char *pmg_last_error_string(const pmg_ctx *c) { return c ? c->err.c_str() : "null context"; }

int pmg_set_params(pmg_ctx *c, const pmg_params *p) {
  if (!c || !p) return PMG_ERR_INVALID_ARGUMENT;
  std::string why;
  if (int st = check_params(p, why)) { c->err = why; return st; }
  const bool pin_change = p->internal_project_pressure != c->prm.internal_project_pressure;
  // the factor and inverse buffers and the streamed operator forms are chosen when a level
  // is added: these knobs are fixed from then on
  if (!c->L.empty() && (p->precision != c->prm.precision || p->patch_inverse != c->prm.patch_inverse ||
                        p->shared_block != c->prm.shared_block || p->sc_smoother_saddle != c->prm.sc_smoother_saddle)) {
    c->err = "precision, patch_inverse, shared_block and sc_smoother_saddle are fixed once levels exist";
    return PMG_ERR_UNSUPPORTED;
  }
  if (pin_change && c->batch) {
    c->err = "the problems of a batch carry their own pressure projection";
    return PMG_ERR_UNSUPPORTED;
  }
  const bool refine_change = (p->bf16_refine_steps > 0) != (c->prm.bf16_refine_steps > 0);
  c->prm = *p;
  if (refine_change) {
    // refinement steps need the FP32 patch matrices stored at the set-up: solves report
    // PMG_ERR_NOT_READY until pmg_setup / pmg_refactor stores them (or drops their use)
    for (auto &D : c->L) D.factors_ready = false;
    c->setup_done = false;
  }
  if (pin_change) {
    // the coarse system pins (or frees) its pressure row: solves report PMG_ERR_NOT_READY
    // until pmg_setup / pmg_refactor builds the coarse factors again
    c->coarse.ready = false;
    c->setup_done = false;
  }
  drop_graph(c);
  return PMG_OK;
}

int pmg_color_cells(int n_cells, int n_ulocal, const int *u_begin, const int *u_global, int n_udof, int *n_colors,
                    int *ptr_colors, int *cell_reorder);
int pmg_sequential_schedule(int n_cells, const int *u_global, const int *p_global, int n_udof, int n_pdof,
                            const int *a_rowptr, const int *a_kcol, const int *bt_rowptr, const int *bt_kcol,
                            const int *b_rowptr, const int *b_kcol, int *n_sets, int *ptr_sets, int *cell_reorder);
int pmg_nodal_schedule(int n_udof, int n_pdof, const int *a_rowptr, const int *a_kcol, const int *bt_rowptr,
                       const int *bt_kcol, const int *b_rowptr, const int *b_kcol, int *n_sets, int *ptr_sets,
                       int *pdof_order);

// merged: a level built by merge_hierarchies (cells without parent allowed, the color
// lists may leave out the cells of problems solved directly on this level)
static int add_level(pmg_ctx *c, const pmg_level_desc *d, int *level_out, bool merged) {
  return guard(c, [&] {
    PMG_HIP_CHECK(hipSetDevice(c->device));
    drop_graph(c);
    if (d->n_ulocal != NU_LOC || d->n_plocal != NP_LOC)
      throw Fail{PMG_ERR_UNSUPPORTED, "only Q2/P1disc hexahedra (27 + 4 local dofs)"};
    if (d->n_udof <= 0 || d->n_active < 0 || d->n_active > d->n_udof || d->n_pdof <= 0 || d->n_cells <= 0)
      throw Fail{PMG_ERR_INVALID_ARGUMENT, "bad level sizes"};
    if (d->hanging_bound != d->n_active) throw Fail{PMG_ERR_UNSUPPORTED, "hanging nodes are not supported"};
    if (d->hanging_bound + d->n_dirichlet != d->n_udof) throw Fail{PMG_ERR_INVALID_ARGUMENT, "dof bounds inconsistent"};
    const int lvl = int(c->L.size());
    if (lvl > 0 && (!d->parent || !d->child || !d->qq_u || !d->qq_p || d->n_children <= 0))
      throw Fail{PMG_ERR_INVALID_ARGUMENT, "fine level without parent/child/QQ data"};
    // validate the CSR
    if (d->a_rowptr[0] != 0 || d->a_rowptr[d->n_udof] != d->a_nnz || d->b_rowptr[d->n_pdof] != d->b_nnz ||
        d->bt_rowptr[d->n_udof] != d->bt_nnz)
      throw Fail{PMG_ERR_INVALID_ARGUMENT, "CSR row pointers inconsistent with nnz"};
    auto check_csr = [](const int *rp, const int *kc, int rows, int ncols, const char *what) {
      for (int r = 0; r < rows; ++r)
        if (rp[r + 1] < rp[r]) throw Fail{PMG_ERR_INVALID_ARGUMENT, std::string(what) + ": row pointers decrease"};
      for (int k = 0; k < rp[rows]; ++k)
        if (kc[k] < 0 || kc[k] >= ncols) throw Fail{PMG_ERR_INVALID_ARGUMENT, std::string(what) + ": column out of range"};
    };
    check_csr(d->a_rowptr, d->a_kcol, d->n_udof, d->n_udof, "A");
    check_csr(d->b_rowptr, d->b_kcol, d->n_pdof, d->n_udof, "B");
    check_csr(d->bt_rowptr, d->bt_kcol, d->n_udof, d->n_pdof, "BT");
    c->L.emplace_back();
    DevLevel &D = c->L.back();
    upload_operator(c, D, d, mixed(c));
    D.n_cells = d->n_cells;
    // dof maps (UGlobalNumbers + UBeginIndex[cell])
    std::vector<int> ug(size_t(D.n_cells) * NU_LOC), pg(size_t(D.n_cells) * NP_LOC);
    for (int cell = 0; cell < D.n_cells; ++cell) {
      for (int j = 0; j < NU_LOC; ++j) {
        const int g = d->u_global[d->u_begin[cell] + j];
        if (g < 0 || g >= D.n_udof) throw Fail{PMG_ERR_INVALID_ARGUMENT, "velocity dof out of range"};
        ug[size_t(cell) * NU_LOC + j] = g;
      }
      for (int j = 0; j < NP_LOC; ++j) {
        const int g = d->p_global[d->p_begin[cell] + j];
        if (g < 0 || g >= D.n_pdof) throw Fail{PMG_ERR_INVALID_ARGUMENT, "pressure dof out of range"};
        pg[size_t(cell) * NP_LOC + j] = g;
      }
    }
    D.u_global = dalloc<int>(c, ug.size()); h2d(c, D.u_global, ug.data(), ug.size());
    D.p_global = dalloc<int>(c, pg.size()); h2d(c, D.p_global, pg.data(), pg.size());
    // coloring (colorCell) - given or computed; smoother 1 (Gauss-Seidel order): the level
    // sets of pmg_sequential_schedule, run as colors
    std::vector<int> ptr(D.n_cells + 1), reo(D.n_cells);
    int nc = 0;
    if (c->prm.sc_smoother_saddle == 1) {
      int st = pmg_sequential_schedule(D.n_cells, ug.data(), pg.data(), D.n_udof, D.n_pdof, d->a_rowptr, d->a_kcol,
                                       d->bt_rowptr, d->bt_kcol, d->b_rowptr, d->b_kcol, &nc, ptr.data(), reo.data());
      if (st) throw Fail{st, "Gauss-Seidel schedule failed"};
    } else if (d->ptr_colors && d->cell_reorder && (d->n_colors > 0 || merged)) {
      nc = d->n_colors;
      if (nc < 0 || nc > D.n_cells) throw Fail{PMG_ERR_INVALID_ARGUMENT, "bad color count"};
      std::copy(d->ptr_colors, d->ptr_colors + nc + 1, ptr.begin());
      std::copy(d->cell_reorder, d->cell_reorder + D.n_cells, reo.begin());
      // the color pointers must ascend and (single problems) cover every cell; the cell list
      // must be a permutation
      bool ok = ptr[0] == 0 && (merged ? ptr[nc] <= D.n_cells : ptr[nc] == D.n_cells);
      for (int k = 0; ok && k < nc; ++k) ok = ptr[k] <= ptr[k + 1];
      std::vector<char> seen(D.n_cells, 0);
      for (int q = 0; ok && q < D.n_cells; ++q) {
        ok = reo[q] >= 0 && reo[q] < D.n_cells && !seen[reo[q]];
        if (ok) seen[reo[q]] = 1;
      }
      if (!ok) throw Fail{PMG_ERR_INVALID_ARGUMENT, "cell coloring is not a partition of the cells"};
    } else {
      std::vector<int> ub(D.n_cells);
      for (int i = 0; i < D.n_cells; ++i) ub[i] = i * NU_LOC;
      int st = pmg_color_cells(D.n_cells, NU_LOC, ub.data(), ug.data(), D.n_udof, &nc, ptr.data(), reo.data());
      if (st) throw Fail{st, "coloring failed"};
    }
    D.n_colors = nc;
    D.ptr_colors.assign(ptr.begin(), ptr.begin() + nc + 1);
    D.max_cells_per_color = 0;
    for (int k = 0; k < nc; ++k) D.max_cells_per_color = std::max(D.max_cells_per_color, ptr[k + 1] - ptr[k]);
    D.cell_reorder = dalloc<int>(c, D.n_cells); h2d(c, D.cell_reorder, reo.data(), D.n_cells);
    {
      std::vector<int> pd(size_t(D.n_cells) * POS_DOFS_STRIDE, 0);
      for (int q = 0; q < D.n_cells; ++q) {
        const int cell = reo[q];
        std::copy(&ug[size_t(cell) * NU_LOC], &ug[size_t(cell) * NU_LOC] + NU_LOC, &pd[size_t(q) * POS_DOFS_STRIDE]);
        std::copy(&pg[size_t(cell) * NP_LOC], &pg[size_t(cell) * NP_LOC] + NP_LOC,
                  &pd[size_t(q) * POS_DOFS_STRIDE + NU_LOC]);
      }
      D.pos_dofs = dalloc<int>(c, pd.size());
      h2d(c, D.pos_dofs, pd.data(), pd.size());
    }
    {
      // residual work items of each color sorted by the operator row they stream
      std::vector<int> items(size_t(D.n_cells) * (NU_LOC + NP_LOC)), rows(items.size());
      // Large colors of a single problem (>= XCD_GROUP_MIN cells) are split into 8 groups of
      // consecutive cells (the color lists cells in ascending index, and cell numbers are
      // spatially local), one per XCD: each group's items sorted by row, stored group after
      // group, and the residual kernel runs group x on the workgroups that land on XCD x, so
      // each XCD's L2 holds the u entries of one compact part of the mesh.
      std::vector<std::pair<int, int>> key;
      std::vector<int> xg(size_t(nc) * 9, 0);
      D.h_xcd_grouped.assign(nc, 0);
      for (int k = 0; k < nc; ++k) {
        const int nck = ptr[k + 1] - ptr[k];
        const bool grouped = !merged && nck >= XCD_GROUP_MIN;
        const int ng = grouped ? 8 : 1;
        D.h_xcd_grouped[k] = grouped;
        size_t at = size_t(ptr[k]) * (NU_LOC + NP_LOC);
        for (int gi = 0; gi < ng; ++gi) {
          key.clear();
          const int q0 = ptr[k] + int(long(nck) * gi / ng), q1 = ptr[k] + int(long(nck) * (gi + 1) / ng);
          for (int q = q0; q < q1; ++q) {
            const int cell = reo[q], slot = q - ptr[k];
            for (int j = 0; j < NU_LOC; ++j) key.emplace_back(ug[size_t(cell) * NU_LOC + j], slot * 32 + j);
            for (int j = 0; j < NP_LOC; ++j)
              key.emplace_back(D.n_udof + pg[size_t(cell) * NP_LOC + j], slot * 32 + NU_LOC + j);
          }
          std::sort(key.begin(), key.end());
          if (grouped) xg[size_t(k) * 9 + gi] = int(at - size_t(ptr[k]) * (NU_LOC + NP_LOC));
          for (size_t i = 0; i < key.size(); ++i, ++at) {
            items[at] = key[i].second;
            // global row: velocity dof, or pressure dof (keys of pressure rows are offset by n_udof)
            rows[at] = key[i].first < D.n_udof ? key[i].first : key[i].first - D.n_udof;
          }
        }
        if (grouped) xg[size_t(k) * 9 + 8] = nck * (NU_LOC + NP_LOC);
      }
      D.xcd_goff = dalloc<int>(c, xg.size());
      h2d(c, D.xcd_goff, xg.data(), xg.size());
      D.res_items = dalloc<int>(c, items.size());
      h2d(c, D.res_items, items.data(), items.size());
      D.res_rows = dalloc<int>(c, rows.size());
      h2d(c, D.res_rows, rows.data(), rows.size());
      D.res_desc = dalloc<int>(c, 4 * rows.size());
      launch_res_desc(D, D.res_items, D.res_rows, long(rows.size()), D.res_desc, c->stream);
      if (merged) {
        D.h_res_items = items;
        D.h_res_rows = rows;
      }
    }
    {
      RowBytes rb;
      for (int sh = 0; sh < 2; ++sh)
        for (int pi = 0; pi < 2; ++pi) {
          const int vs = pi == 0 ? 8 : 4;
          double vb = 0;
          for (int i = 0; i < D.n_udof; ++i) vb += rb.vel(d, i, vs, sh);
          double pb = 0;
          for (int i = 0; i < D.n_pdof; ++i) pb += rb.pre(d, i, vs);
          // defect: rows + x read + f read + d write; matvec (FP64 only): no f
          D.defect_bytes[sh][pi] = vb + pb + 3.0 * vs * D.n_dof;
          if (pi == 0) D.matvec_bytes[sh] = vb + pb + 2.0 * vs * D.n_dof;
          D.res_bytes[sh][pi].assign(nc, 0.0);
          for (int k = 0; k < nc; ++k) {
            double b = 0;
            // with stored inverses (FP32) the launch also writes u at the patch dofs
            const double uo = (pi == 1 && c->prm.patch_inverse) ? double(vs) : 0.0;
            for (int q = ptr[k]; q < ptr[k + 1]; ++q) {
              const int cell = reo[q];
              b += (NU_LOC + NP_LOC) * 20.0;   // work items and their 16-byte descriptors
              // rows (their extents come from the descriptors, not the row pointers), f,
              // local residual (and u) written, inverse-permutation bytes
              for (int j = 0; j < NU_LOC; ++j) {
                const int g = ug[size_t(cell) * NU_LOC + j];
                b += rb.vel(d, g, vs, sh) - (g < d->n_active ? 16 : 8) + 3 * vs + 3 * vs + 3 * uo + 3;
              }
              for (int j = 0; j < NP_LOC; ++j) b += rb.pre(d, pg[size_t(cell) * NP_LOC + j], vs) - 8 + vs + vs + uo + 1;
            }
            D.res_bytes[sh][pi][k] = b;
          }
        }
    }
    // vectors
    D.sol = dalloc<double>(c, D.n_dof); D.rhs = dalloc<double>(c, D.n_dof);
    D.aux0 = dalloc<double>(c, D.n_dof); D.aux1 = dalloc<double>(c, D.n_dof);
    launch_fill<double>(D.sol, 0.0, D.n_dof, c->stream);
    launch_fill<double>(D.rhs, 0.0, D.n_dof, c->stream);
    if (mixed(c)) {
      D.sol32 = dalloc<float>(c, D.n_dof); D.rhs32 = dalloc<float>(c, D.n_dof);
      D.aux0_32 = dalloc<float>(c, D.n_dof); D.aux1_32 = dalloc<float>(c, D.n_dof);
    }
    {
      // gather map of the patch matrices (patch LU, device re-assembly): every level
      const std::vector<int> pm = build_patch_map(d, ug, pg);
      D.pmap = dalloc<int>(c, pm.size());
      h2d(c, D.pmap, pm.data(), pm.size());
      PMG_HIP_CHECK(hipStreamSynchronize(c->stream));
    }
    if (lvl > 0) {
      const size_t lun = size_t(D.n_cells) * LU_STRIDE + LU_PAD;
      const size_t minvn = size_t(D.n_cells) * MINV_STRIDE + LU_PAD;
      if (c->prm.sc_smoother_saddle >= 3) {
        // nodal Vanka: no stored patch data; the pressure dofs in the level sets of their order
        if (merged) throw Fail{PMG_ERR_UNSUPPORTED, "nodal Vanka is provided for single problems only"};
        for (int i = 0; i < D.n_pdof; ++i)
          if (d->b_rowptr[i + 1] - d->b_rowptr[i] > NODAL_MAXU)
            throw Fail{PMG_ERR_UNSUPPORTED, "nodal Vanka: a B row with more than " + std::to_string(NODAL_MAXU) +
                                                " velocity dofs"};
        std::vector<int> nptr(size_t(D.n_pdof) + 1), nord(D.n_pdof);
        int ns = 0;
        int st = pmg_nodal_schedule(D.n_udof, D.n_pdof, d->a_rowptr, d->a_kcol, d->bt_rowptr, d->bt_kcol, d->b_rowptr,
                                    d->b_kcol, &ns, nptr.data(), nord.data());
        if (st) throw Fail{st, "nodal schedule failed"};
        D.nodal = true;
        D.nodal_ptr.assign(nptr.begin(), nptr.begin() + ns + 1);
        D.nodal_order = dalloc<int>(c, D.n_pdof);
        h2d(c, D.nodal_order, nord.data(), D.n_pdof);
        D.nodal_info = dalloc<int>(c, 1);
        launch_fill<int>(D.nodal_info, 0, 1, c->stream);
        // algorithmic bytes per set: B row (columns, 3 values, 3 gathered velocities), the
        // velocity rows (A: column + 3 values + 3 gathered velocities per entry, BT: column,
        // value, pressure), right-hand sides, and the updates (read + write)
        D.nodal_bytes.assign(ns, 0.0);
        for (int k = 0; k < ns; ++k)
          for (int q = nptr[k]; q < nptr[k + 1]; ++q) {
            const int i = nord[q];
            double b = 8.0 + 8.0 + 16.0;
            for (int e = d->b_rowptr[i]; e < d->b_rowptr[i + 1]; ++e) {
              const int r = d->b_kcol[e];
              b += 4.0 + 24.0 + 24.0;
              b += 3.0 * (d->a_rowptr[r + 1] - d->a_rowptr[r]) * (4.0 + 24.0 + 24.0);
              b += 3.0 * (d->bt_rowptr[r + 1] - d->bt_rowptr[r]) * (4.0 + 8.0 + 8.0);
              b += 3.0 * (8.0 + 16.0);
            }
            D.nodal_bytes[k] += b;
          }
      } else if (c->prm.precision == PMG_PREC_FP64) {
        D.lu = dalloc<double>(c, lun);
        launch_fill<double>(D.lu, 0.0, long(lun), c->stream);
      } else if (c->prm.patch_inverse && c->prm.precision == PMG_PREC_FP32) {
        D.minv32 = dalloc<float>(c, minvn);
        launch_fill<float>(D.minv32, 0.0f, long(minvn), c->stream);
      } else if (c->prm.patch_inverse) {
        D.minv16 = dalloc<uint16_t>(c, minvn);
        launch_fill<uint16_t>(D.minv16, uint16_t(0), long(minvn), c->stream);
      }
      if (D.minv32 || D.minv16) D.ua32 = dalloc<float>(c, 4 * size_t(D.n_udof)); else if (c->prm.precision == PMG_PREC_FP32) {
        D.lu32 = dalloc<float>(c, lun);
        launch_fill<float>(D.lu32, 0.0f, long(lun), c->stream);
      } else {
        D.lu16 = dalloc<uint16_t>(c, lun);
        PMG_HIP_CHECK(hipMemsetAsync(D.lu16, 0, lun * sizeof(uint16_t), c->stream));
      }
      D.invperm = dalloc<uint8_t>(c, size_t(D.n_cells) * PERM_STRIDE);
      D.patch_info = dalloc<int>(c, D.n_cells);
      D.rloc = dalloc<double>(c, size_t(D.max_cells_per_color) * RLOC_STRIDE);
      if (mixed(c)) D.rloc32 = dalloc<float>(c, size_t(D.max_cells_per_color) * RLOC_STRIDE);
      if (D.minv32 || D.minv16) D.uloc32 = dalloc<float>(c, size_t(D.max_cells_per_color) * RLOC_STRIDE);
      // transfers from the coarser level
      const DevLevel &C = c->L[lvl - 1];
        for (int fc = 0; fc < D.n_cells; ++fc) {
        if (merged && d->parent[fc] == -1) continue;
        if (d->parent[fc] < 0 || d->parent[fc] >= C.n_cells || d->child[fc] < 0 || d->child[fc] >= d->n_children)
          throw Fail{PMG_ERR_INVALID_ARGUMENT, "parent/child map out of range"};
      }
      build_transfer(c, NU_LOC, NU_LOC, D.n_cells, d->parent, d->child, d->qq_u, ug, c->h_u_global[lvl - 1],
                     D.n_udof, C.n_udof, D.pu, D.ru);
      build_transfer(c, NP_LOC, NP_LOC, D.n_cells, d->parent, d->child, d->qq_p, pg, c->h_p_global[lvl - 1],
                     D.n_pdof, C.n_pdof, D.pp, D.rp);
      if (!merged)
        build_function_restriction(c, D.n_cells, d->parent, d->child, d->qq_u, ug, c->h_u_global[lvl - 1], D.n_udof,
                                   C.n_udof, D.rf);
    }
    if (lvl > 0) {
      const DevLevel &C = c->L[lvl - 1];
      auto tb = [](const Csr &m, int ncomp) {
        return double(m.nnz) * 12.0 + (m.n_rows + 1) * 4.0 + double(ncomp) * (m.n_rows + m.n_cols) * 8.0;
      };
      D.transfer_bytes = tb(D.ru, 3) + tb(D.rp, 1);
      (void)C;
    }
    c->h_u_global.push_back(std::move(ug));
    c->h_p_global.push_back(std::move(pg));
    PMG_HIP_CHECK(hipStreamSynchronize(c->stream));
    c->setup_done = false;
    c->coarse.ready = false;
    if (level_out) *level_out = lvl;
  });
}

int pmg_level_add(pmg_ctx *c, const pmg_level_desc *d, int *level_out) {
  if (!c || !d) return PMG_ERR_INVALID_ARGUMENT;
  if (c->batch) {
    c->err = "levels of a batch context are fixed by pmg_batch_create";
    return PMG_ERR_UNSUPPORTED;
  }
  return add_level(c, d, level_out, false);
}

int pmg_level_update_values(pmg_ctx *c, int level, const pmg_level_desc *d) {
  if (!c || !d) return PMG_ERR_INVALID_ARGUMENT;
  if (c->batch) {
    c->err = "new values for a batch: create a new batch context";
    return PMG_ERR_UNSUPPORTED;
  }
  return guard(c, [&] {
    DevLevel &D = level_of(c, level);
    if (d->a_nnz != D.a_nnz || d->b_nnz != D.b_nnz || d->bt_nnz != D.bt_nnz)
      throw Fail{PMG_ERR_INVALID_ARGUMENT, "pattern changed"};
    upload_values(c, D, d);
    D.factors_ready = false;
    if (level == 0) c->coarse.ready = false;
    c->setup_done = false;
  });
}

int pmg_setup(pmg_ctx *c) {
  if (!c) return PMG_ERR_INVALID_ARGUMENT;
  return guard(c, [&] {
    if (c->L.empty()) throw Fail{PMG_ERR_NOT_READY, "no levels"};
    PMG_HIP_CHECK(hipSetDevice(c->device));
    setup_factors(c);
  });
}

int pmg_refactor(pmg_ctx *c) {
  if (!c) return PMG_ERR_INVALID_ARGUMENT;
  return guard(c, [&] {
    if (c->L.empty()) throw Fail{PMG_ERR_NOT_READY, "no levels"};
    PMG_HIP_CHECK(hipSetDevice(c->device));
    for (auto &D : c->L) D.factors_ready = false;
    c->coarse.ready = false;
    for (auto &b : c->bi) b.co.ready = false;
    setup_factors(c);
  });
}

int pmg_level_form(const pmg_ctx *c, int level, int *shared_block) {
  if (!c || !shared_block || level < 0 || level >= int(c->L.size())) return PMG_ERR_INVALID_ARGUMENT;
  *shared_block = c->L[level].a_shared ? 1 : 0;
  return PMG_OK;
}

int pmg_level_info(const pmg_ctx *c, int level, int *n_dof, int *n_colors, int *mcpc) {
  if (!c || level < 0 || level >= int(c->L.size())) return PMG_ERR_INVALID_ARGUMENT;
  const DevLevel &D = c->L[level];
  if (n_dof) *n_dof = D.n_dof;
  if (n_colors) *n_colors = D.n_colors;
  if (mcpc) *mcpc = D.max_cells_per_color;
  return PMG_OK;
}

#define PMG_NEED_READY()                                                          \
  do {                                                                            \
    if (!c->setup_done) throw Fail{PMG_ERR_NOT_READY, "pmg_setup not called"};    \
  } while (0)

int pmg_defect(pmg_ctx *c, int level, const double *u, const double *f, double *d, double *res, int mem) {
  if (!c || !u || !f || !d) return PMG_ERR_INVALID_ARGUMENT;
  return guard(c, [&] {
    DevLevel &D = level_of(c, level);
    HostIO io(c, mem, D.n_dof);
    double *du = io.in(0, u), *df = io.in(1, f), *dd = (mem == PMG_MEM_HOST) ? c->stage[2] : d;
    mg_defect<double>(c, D, du, df, dd, c->red.result);
    if (c->batch) launch_dot(dd, dd, D.n_dof, c->red.result, c->red, c->stream);
    io.out(2, d);
    PMG_HIP_CHECK(hipMemcpyAsync(c->hpin, c->red.result, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    io.sync();
    if (res) *res = c->batch ? std::sqrt(c->hpin[0]) : c->hpin[0];
  });
}

int pmg_matvec(pmg_ctx *c, int level, const double *x, double *y, int mem) {
  if (!c || !x || !y) return PMG_ERR_INVALID_ARGUMENT;
  return guard(c, [&] {
    DevLevel &D = level_of(c, level);
    HostIO io(c, mem, D.n_dof);
    double *dx = io.in(0, x), *dy = (mem == PMG_MEM_HOST) ? c->stage[2] : y;
    launch_type4<double>(D.op64(), 0, dx, nullptr, dy, false, nullptr, c->red, c->stream);
    io.out(2, y);
    io.sync();
  });
}

int pmg_defect_nse4(pmg_ctx *c, int level, const double *x, const double *b, double *r, int mem) {
  if (!c || !x || !b || !r) return PMG_ERR_INVALID_ARGUMENT;
  return guard(c, [&] {
    DevLevel &D = level_of(c, level);
    HostIO io(c, mem, D.n_dof);
    double *dx = io.in(0, x), *db = io.in(1, b), *dr = (mem == PMG_MEM_HOST) ? c->stage[2] : r;
    launch_type4<double>(D.op64(), 1, dx, db, dr, fused_l20(c), nullptr, c->red, c->stream);
    if (c->batch) project_pressure<double>(c, D, dr);
    io.out(2, r);
    io.sync();
  });
}

int pmg_vanka_smooth(pmg_ctx *c, int level, double *u, const double *f, int nsweeps, int mem) {
  if (!c || !u || !f || nsweeps < 0) return PMG_ERR_INVALID_ARGUMENT;
  return guard(c, [&] {
    PMG_NEED_READY();
    if (level < 1) throw Fail{PMG_ERR_INVALID_ARGUMENT, "the coarsest level is solved directly"};
    DevLevel &D = level_of(c, level);
    HostIO io(c, mem, D.n_dof);
    double *du = io.in(0, u), *df = io.in(1, f);
    if (!mixed(c)) {
      for (int s = 0; s < nsweeps; ++s) vanka_sweep<double>(c, D, du, df, D.aux1, D.rloc);
    } else {
      launch_d2f(D.sol32, du, D.n_dof, c->stream);
      launch_d2f(D.rhs32, df, D.n_dof, c->stream);
      for (int s = 0; s < nsweeps; ++s) vanka_sweep<float>(c, D, D.sol32, D.rhs32, D.aux1_32, D.rloc32);
      launch_f2d(du, D.sol32, D.n_dof, c->stream);
    }
    io.out(0, u);
    io.sync();
  });
}

// ------------------------------------------------------------------ launch-shape autotuner
// Times the candidate launch shapes of the level kernels on the device (HIP events around
// `reps` launches each, after one untimed launch) and keeps the fastest per level and
// precision: work items per workgroup of the Vanka residual (16, 32, 64), whether large
// colors run in their per-XCD groups, and operator rows per workgroup of the type-4
// kernels (32 .. 256).  Every candidate computes the same values (the work is only
// distributed differently), so the choice changes timings, not results, except for the
// summation order of the grid-wide pressure reductions.
extern "C++" {
// The candidate's reps launches are captured once as a HIP graph and the graph is replayed
// twice between events (the faster replay counts), so that the device time is measured,
// not the host's launch rate (which, eager, exceeds a small level's kernel time).
static float time_launches(pmg_ctx *c, int reps, const std::function<void()> &go) {
  go();
  hipGraph_t g = nullptr;
  hipGraphExec_t ge = nullptr;
  PMG_HIP_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  try {
    for (int r = 0; r < reps; ++r) go();
  } catch (...) {
    (void)hipStreamEndCapture(c->stream, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  PMG_HIP_CHECK(hipStreamEndCapture(c->stream, &g));
  PMG_HIP_CHECK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
  PMG_HIP_CHECK(hipGraphLaunch(ge, c->stream));
  float best = 1e30f;
  for (int k = 0; k < 2; ++k) {
    PMG_HIP_CHECK(hipEventRecord(c->ev0, c->stream));
    PMG_HIP_CHECK(hipGraphLaunch(ge, c->stream));
    PMG_HIP_CHECK(hipEventRecord(c->ev1, c->stream));
    PMG_HIP_CHECK(hipEventSynchronize(c->ev1));
    float ms = 0;
    PMG_HIP_CHECK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    best = std::min(best, ms);
  }
  PMG_HIP_CHECK(hipGraphExecDestroy(ge));
  PMG_HIP_CHECK(hipGraphDestroy(g));
  return best / reps;
}

template <typename T>
static void tune_level(pmg_ctx *c, DevLevel &D, int reps, bool finest) {
  const int pi = sizeof(T) == 8 ? 0 : 1;
  T *u = Vec<T>::sol(D), *f = Vec<T>::rhs(D), *rloc = Vec<T>::rloc(D);
  launch_fill<T>(u, T(0.5), D.n_dof, c->stream);
  launch_fill<T>(f, T(1.0), D.n_dof, c->stream);
  const OpView<T> op = opview<T>(D);
  if (D.n_colors > 0 && !D.nodal) {
    float *ua = nullptr;
    if constexpr (sizeof(T) == 4) {
      if (D.ua32 && D.a_shared && (D.minv16 || D.minv32)) {
        ua = D.ua32;
        launch_sweep_start(u, f, D.n_udof, D.hanging_bound, D.n_dirichlet, ua, c->stream);
      }
    }
    const bool grouped = D.h_xcd_grouped[0];
    T *uloc = sizeof(T) == 4 ? reinterpret_cast<T *>(D.uloc32) : nullptr;
    // one pass = the residual launches of every color in sweep order (each color reads its
    // own rows, as in a sweep; one color repeated would find its rows in the caches)
    const int passes = std::max(1, reps / std::max(1, D.n_colors));
    auto pass = [&](int xcd, int pb) {
      for (int col = 0; col < D.n_colors; ++col) {
        const int p0 = D.ptr_colors[col];
        const long n_items = long(D.ptr_colors[col + 1] - p0) * (NU_LOC + NP_LOC);
        const int *items = D.res_items + size_t(p0) * (NU_LOC + NP_LOC);
        const int *desc = D.res_desc + 4 * size_t(p0) * (NU_LOC + NP_LOC);
        const int *xg = xcd && D.h_xcd_grouped[col] ? D.xcd_goff + 9 * col : nullptr;
        launch_vanka_residual<T>(op, n_items, p0, D.invperm, items, desc, xg, u, ua, f, rloc, uloc, c->stream, pb);
      }
    };
    // three rounds over the candidates, each candidate's fastest round counts (the rounds
    // interleave the candidates, so a slow phase of the device does not favour one)
    std::vector<std::pair<int, int>> cand;
    for (int xcd = grouped ? 0 : 1; xcd <= 1; ++xcd)
      for (int pb : {16, 32, 64}) cand.emplace_back(xcd, pb);
    std::vector<float> best(cand.size(), 1e30f);
    for (int round = 0; round < 3; ++round)
      for (size_t k = 0; k < cand.size(); ++k)
        best[k] = std::min(best[k], time_launches(c, passes, [&] { pass(cand[k].first, cand[k].second); }));
    const size_t kb = size_t(std::min_element(best.begin(), best.end()) - best.begin());
    D.tune_res_xcd[pi] = cand[kb].first;
    D.tune_res_pb[pi] = cand[kb].second;
  }
  // the MG defect of this level (and, finest level of a mixed context, the FP64 matvec)
  auto tune_op = [&](auto zero, int mode, int slot) {
    using U = decltype(zero);
    U *x = Vec<U>::sol(D), *b = Vec<U>::rhs(D), *yy = Vec<U>::aux0(D);
    if (!std::is_same<U, T>::value) {
      launch_fill<U>(x, U(0.5), D.n_dof, c->stream);
      launch_fill<U>(b, U(1.0), D.n_dof, c->stream);
    }
    const int cand[4] = {32, 64, 128, 256};
    float best[4] = {1e30f, 1e30f, 1e30f, 1e30f};
    for (int round = 0; round < 3; ++round)
      for (int k = 0; k < 4; ++k) {
        OpView<U> o = opview<U>(D);
        o.rpb = cand[k];
        best[k] = std::min(best[k], time_launches(c, reps, [&] {
          launch_type4<U>(o, mode, x, b, yy, mode != 0 && fused_l20(c), nullptr, c->red, c->stream);
        }));
      }
    D.tune_op_rpb[slot] = cand[std::min_element(best, best + 4) - best];
  };
  tune_op(T(0), 2, pi);
  if (finest && pi == 1) tune_op(double(0), 0, 0);
}
}  // extern "C++"

int pmg_autotune(pmg_ctx *c, int reps) {
  if (!c || reps <= 0 || reps > 1000) return PMG_ERR_INVALID_ARGUMENT;
  return guard(c, [&] {
    PMG_NEED_READY();
    drop_graph(c);
    const bool mx = mixed(c);
    // the coarsest level is solved directly (no smoother, no defect in the cycle)
    for (size_t l = 1; l < c->L.size(); ++l) {
      DevLevel &D = c->L[l];
      const bool finest = l + 1 == c->L.size();
      if (mx)
        tune_level<float>(c, D, reps, finest);
      else
        tune_level<double>(c, D, reps, finest);
    }
    PMG_HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

int pmg_set_level_tuning(pmg_ctx *c, int level, const int *in6) {
  if (!c || !in6) return PMG_ERR_INVALID_ARGUMENT;
  return guard(c, [&] {
    DevLevel &D = level_of(c, level);
    for (int k = 0; k < 2; ++k) {
      if (in6[k] != 0 && (in6[k] < 1 || in6[k] > 1024))
        throw Fail{PMG_ERR_INVALID_ARGUMENT, "residual items per workgroup out of range"};
      if (in6[2 + k] != 0 && in6[2 + k] != 1) throw Fail{PMG_ERR_INVALID_ARGUMENT, "XCD group flag must be 0 or 1"};
      if (in6[4 + k] != 0 && (in6[4 + k] < 1 || in6[4 + k] > 4096))
        throw Fail{PMG_ERR_INVALID_ARGUMENT, "operator rows per workgroup out of range"};
    }
    for (int k = 0; k < 2; ++k) {
      D.tune_res_pb[k] = in6[k];
      D.tune_res_xcd[k] = in6[2 + k];
      D.tune_op_rpb[k] = in6[4 + k];
    }
    drop_graph(c);
  });
}

int pmg_level_tuning(pmg_ctx *c, int level, int *out6) {
  if (!c || !out6) return PMG_ERR_INVALID_ARGUMENT;
  return guard(c, [&] {
    const DevLevel &D = level_of(c, level);
    const int v[6] = {D.tune_res_pb[0], D.tune_res_pb[1], D.tune_res_xcd[0], D.tune_res_xcd[1], D.tune_op_rpb[0],
                      D.tune_op_rpb[1]};
    std::copy(v, v + 6, out6);
  });
}

int pmg_vanka_smooth_noalloc(pmg_ctx *c, int level, double *u, const double *f, int nsweeps) {
  if (!c || !u || !f || nsweeps < 0) return PMG_ERR_INVALID_ARGUMENT;
  return guard(c, [&] {
    PMG_NEED_READY();
    if (level < 1) throw Fail{PMG_ERR_INVALID_ARGUMENT, "the coarsest level is solved directly"};
    if (mixed(c))
      throw Fail{PMG_ERR_UNSUPPORTED, "pmg_vanka_smooth_noalloc runs the FP64 form only (the mixed forms convert "
                                      "through level-owned FP32 vectors)"};
    DevLevel &D = level_of(c, level);
    // the relaxation copy of u: cell Vanka with |1 - alpha| > 1e-3, nodal Vanka always
    const bool relax = D.nodal || std::fabs(1 - c->prm.sc_smooth_damp_factor_saddle) > 1e-3;
    const size_t n_rloc = size_t(D.max_cells_per_color) * RLOC_STRIDE;
    const size_t n_aux = relax ? size_t(D.n_dof) : 0;
    const size_t off_aux = (n_rloc + 31) / 32 * 32;   // 256-byte aligned
    if (off_aux + n_aux > NOALLOC_SCRATCH_DOUBLES)
      throw Fail{PMG_ERR_UNSUPPORTED, "level " + std::to_string(level) + " needs " +
                                          std::to_string((off_aux + n_aux) * 8) + " bytes of scratch (" +
                                          std::to_string(D.max_cells_per_color) + " patches per color" +
                                          (relax ? ", relaxation copy of u" : "") + "); the static scratch holds " +
                                          std::to_string(NOALLOC_SCRATCH_DOUBLES * 8)};
    double *base = scratch_of(c->device);
    if (!base) throw Fail{PMG_ERR_DEVICE, "hipGetSymbolAddress of the static scratch failed on this device"};
    double *rloc = base, *aux = base + off_aux;
    // the array keeps the previous call's contents: the local residual slots start at zero
    // (their padding is read by vector loads); aux is written by the sweep before it is read
    launch_fill<double>(rloc, 0.0, long(n_rloc), c->stream);
    for (int k = 0; k < nsweeps; ++k) vanka_sweep<double>(c, D, u, f, aux, rloc);
    PMG_HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

int pmg_restrict_defect(pmg_ctx *c, int level, const double *d_fine, double *f_coarse, int mem) {
  if (!c || !d_fine || !f_coarse) return PMG_ERR_INVALID_ARGUMENT;
  return guard(c, [&] {
    if (level < 1) throw Fail{PMG_ERR_INVALID_ARGUMENT, "level must be >= 1"};
    DevLevel &F = level_of(c, level), &C = level_of(c, level - 1);
    HostIO io(c, mem, F.n_dof);
    double *dd = io.in(0, d_fine), *dc = (mem == PMG_MEM_HOST) ? c->stage[2] : f_coarse;
    restrict_defect<double>(c, F, C, dd, dc);
    if (mem == PMG_MEM_HOST)
      PMG_HIP_CHECK(hipMemcpyAsync(f_coarse, dc, sizeof(double) * C.n_dof, hipMemcpyDeviceToHost, c->stream));
    io.sync();
  });
}

int pmg_prolongate(pmg_ctx *c, int level, const double *u_coarse, double *d_fine, int mem) {
  if (!c || !u_coarse || !d_fine) return PMG_ERR_INVALID_ARGUMENT;
  return guard(c, [&] {
    if (level < 1) throw Fail{PMG_ERR_INVALID_ARGUMENT, "level must be >= 1"};
    DevLevel &F = level_of(c, level), &C = level_of(c, level - 1);
    HostIO io(c, mem, F.n_dof);
    double *uc = const_cast<double *>(u_coarse);
    if (mem == PMG_MEM_HOST) {
      PMG_HIP_CHECK(hipMemcpyAsync(c->stage[0], u_coarse, sizeof(double) * C.n_dof, hipMemcpyHostToDevice, c->stream));
      uc = c->stage[0];
    }
    double *dd = (mem == PMG_MEM_HOST) ? c->stage[2] : d_fine;
    prolongate<double>(c, F, C, uc, dd);
    io.out(2, d_fine);
    io.sync();
  });
}

int pmg_restrict_function(pmg_ctx *c, int level, const double *u_fine, double *u_coarse, int mem) {
  if (!c || !u_fine || !u_coarse) return PMG_ERR_INVALID_ARGUMENT;
  return guard(c, [&] {
    if (c->batch) throw Fail{PMG_ERR_UNSUPPORTED, "RestrictFunction on a batch context"};
    if (level < 1) throw Fail{PMG_ERR_INVALID_ARGUMENT, "level must be >= 1"};
    DevLevel &F = level_of(c, level), &C = level_of(c, level - 1);
    HostIO io(c, mem, F.n_dof);
    const double *uf = io.in(0, u_fine);
    double *uc = u_coarse;
    if (mem == PMG_MEM_HOST) {
      uc = c->stage[2];
      PMG_HIP_CHECK(hipMemcpyAsync(uc, u_coarse, sizeof(double) * C.n_dof, hipMemcpyHostToDevice, c->stream));
    }
    Timed tm(c, &c->tm.transfer, double(F.rf.nnz) * 12.0 + 3.0 * 8.0 * (F.n_udof + C.n_udof), 0.0);
    launch_csr_spmv<double>(F.rf, 3, uf, F.n_udof, uc, C.n_udof, c->stream);
    if (mem == PMG_MEM_HOST)
      PMG_HIP_CHECK(hipMemcpyAsync(u_coarse, uc, sizeof(double) * C.n_dof, hipMemcpyDeviceToHost, c->stream));
    io.sync();
  });
}

int pmg_restrict_to_all_grids(pmg_ctx *c, const double *u_finest, double *const *u_levels, int mem) {
  if (!c || !u_finest || !u_levels) return PMG_ERR_INVALID_ARGUMENT;
  const int L = int(c->L.size());
  for (int l = 0; l + 1 < L; ++l)
    if (!u_levels[l]) return PMG_ERR_INVALID_ARGUMENT;
  for (int lev = L - 1; lev > 0; --lev) {
    const double *src = lev == L - 1 ? u_finest : u_levels[lev];
    if (int st = pmg_restrict_function(c, lev, src, u_levels[lev - 1], mem)) return st;
  }
  return PMG_OK;
}

int pmg_correct_nodes(pmg_ctx *c, int level, double *u, int mem) {
  if (!c || !u) return PMG_ERR_INVALID_ARGUMENT;
  return guard(c, [&] {
    DevLevel &D = level_of(c, level);
    HostIO io(c, mem, D.n_dof);
    double *du = io.in(0, u);
    correct_nodes(c, D, du);
    io.out(0, u);
    io.sync();
  });
}

int pmg_set_dirichlet(pmg_ctx *c, int level, double *u, const double *f, int mem) {
  if (!c || !u || !f) return PMG_ERR_INVALID_ARGUMENT;
  return guard(c, [&] {
    DevLevel &D = level_of(c, level);
    HostIO io(c, mem, D.n_dof);
    double *du = io.in(0, u), *df = io.in(1, f);
    set_dirichlet(c, D, du, df);
    io.out(0, u);
    io.sync();
  });
}

int pmg_correct_defect(pmg_ctx *c, int level, double *d, int mem) {
  if (!c || !d) return PMG_ERR_INVALID_ARGUMENT;
  return guard(c, [&] {
    DevLevel &D = level_of(c, level);
    HostIO io(c, mem, D.n_dof);
    double *dd = io.in(0, d);
    correct_defect(c, D, dd);
    io.out(0, d);
    io.sync();
  });
}

int pmg_update(pmg_ctx *c, int level, double *u, const double *d, int mem) {
  if (!c || !u || !d) return PMG_ERR_INVALID_ARGUMENT;
  return guard(c, [&] {
    DevLevel &D = level_of(c, level);
    HostIO io(c, mem, D.n_dof);
    double *du = io.in(0, u), *dd = io.in(1, d);
    launch_axpy<double>(du, c->prm.sc_gmg_damp_factor_saddle, dd, D.n_dof, c->stream);
    io.out(0, u);
    io.sync();
  });
}

int pmg_coarse_solve(pmg_ctx *c, double *u, const double *f, int mem) {
  if (!c || !u || !f) return PMG_ERR_INVALID_ARGUMENT;
  return guard(c, [&] {
    PMG_NEED_READY();
    DevLevel &D = level_of(c, 0);
    HostIO io(c, mem, D.n_dof);
    double *du = (mem == PMG_MEM_HOST) ? c->stage[0] : u, *df = io.in(1, f);
    if (c->batch)
      batch_coarse<double>(c, D, du, df);
    else
      coarse_solve_t<double>(c, D, du, df);
    io.out(0, u);
    io.sync();
    check_deferred(c);
  });
}

int pmg_vcycle(pmg_ctx *c, const double *v, double *z, int mem) {
  if (!c || !v || !z) return PMG_ERR_INVALID_ARGUMENT;
  return guard(c, [&] {
    PMG_NEED_READY();
    DevLevel &F = c->L.back();
    HostIO io(c, mem, F.n_dof);
    double *dv = io.in(0, v), *dz = (mem == PMG_MEM_HOST) ? c->stage[2] : z;
    vcycle_dev(c, dv, dz);
    io.out(2, z);
    io.sync();
    check_deferred(c);
  });
}

int pmg_fgmres_solve(pmg_ctx *c, double *sol, const double *rhs, int *iters, double *final_res, int mem) {
  if (!c || !sol || !rhs) return PMG_ERR_INVALID_ARGUMENT;
  if (c->batch) {
    c->err = "a batch context solves with pmg_batch_fgmres_solve";
    return PMG_ERR_UNSUPPORTED;
  }
  int it = 0;
  double fr = 0;
  int st = PMG_OK;
  int rc = guard(c, [&] {
    PMG_NEED_READY();
    DevLevel &F = c->L.back();
    HostIO io(c, mem, F.n_dof);
    double *ds = io.in(0, sol), *dr = io.in(1, rhs);
    st = fgmres_dev(c, ds, dr, &it, &fr);
    io.out(0, sol);
    io.sync();
    check_deferred(c);
  });
  if (iters) *iters = it;
  if (final_res) *final_res = fr;
  if (rc) return rc;
  if (st == PMG_ERR_NOT_CONVERGED) c->err = "FGMRES not converged";
  return st;
}

int pmg_level_geometry(pmg_ctx *c, int level, const double *cell_vertices) {
  if (!c || !cell_vertices) return PMG_ERR_INVALID_ARGUMENT;
  return guard(c, [&] {
    if (c->batch) throw Fail{PMG_ERR_UNSUPPORTED, "re-assembly of a batch context"};
    DevLevel &D = level_of(c, level);
    const size_t n = size_t(D.n_cells) * 24;
    if (!D.geom) D.geom = dalloc<double>(c, n);
    h2d(c, D.geom, cell_vertices, n);
    PMG_HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

int pmg_eval_velocity(pmg_ctx *c, int level, const double *u, double *b_quad, int mem) {
  if (!c || !u || !b_quad) return PMG_ERR_INVALID_ARGUMENT;
  return guard(c, [&] {
    DevLevel &D = level_of(c, level);
    const long nq = long(D.n_cells) * 81;
    HostIO io(c, mem, std::max<long>(D.n_dof, nq));
    const double *du = u;
    double *db = b_quad;
    if (mem == PMG_MEM_HOST) {
      PMG_HIP_CHECK(hipMemcpyAsync(c->stage[0], u, sizeof(double) * D.n_dof, hipMemcpyHostToDevice, c->stream));
      du = c->stage[0];
      db = c->stage[2];
    }
    launch_eval_velocity(D, du, db, c->stream);
    if (mem == PMG_MEM_HOST)
      PMG_HIP_CHECK(hipMemcpyAsync(b_quad, db, sizeof(double) * nq, hipMemcpyDeviceToHost, c->stream));
    io.sync();
  });
}

int pmg_assemble_oseen(pmg_ctx *c, int level, const double *b_quad, double nu, double sigma, int mem) {
  if (!c || !b_quad) return PMG_ERR_INVALID_ARGUMENT;
  return guard(c, [&] {
    if (c->batch) throw Fail{PMG_ERR_UNSUPPORTED, "re-assembly of a batch context"};
    DevLevel &D = level_of(c, level);
    if (!D.geom) throw Fail{PMG_ERR_NOT_READY, "pmg_level_geometry not called for this level"};
    const long nq = long(D.n_cells) * 81;
    HostIO io(c, mem, nq);
    const double *db = b_quad;
    if (mem == PMG_MEM_HOST) {
      PMG_HIP_CHECK(hipMemcpyAsync(c->stage[0], b_quad, sizeof(double) * nq, hipMemcpyHostToDevice, c->stream));
      db = c->stage[0];
    }
    check_deferred(c);
    PMG_HIP_CHECK(hipStreamSynchronize(c->stream2));   // the coarse LU may still read the values
    launch_assemble_oseen(D, D.geom, db, nu, sigma, c->stream);
    refresh_values(c, D);
    io.sync();
    D.factors_ready = false;
    if (level == 0) c->coarse.ready = false;
    c->setup_done = false;
  });
}

int pmg_theta_rhs(pmg_ctx *c, int level, const double *u_old, const double *b_quad, double nu, double sigma,
                  const double *g, double *rhs, int mem) {
  if (!c || !u_old || !b_quad || !g || !rhs) return PMG_ERR_INVALID_ARGUMENT;
  return guard(c, [&] {
    if (c->batch) throw Fail{PMG_ERR_UNSUPPORTED, "theta right-hand side on a batch context"};
    DevLevel &D = level_of(c, level);
    if (!D.geom) throw Fail{PMG_ERR_NOT_READY, "pmg_level_geometry not called for this level"};
    const long nq = long(D.n_cells) * 81;
    HostIO io(c, mem, std::max<long>(D.n_dof, nq));
    const double *du = u_old, *dg = g, *db = b_quad;
    double *dr = rhs;
    if (mem == PMG_MEM_HOST) {
      // staging: u, g and the quadrature values in the three stage buffers (each holds
      // max(n_dof, 81 n_cells) doubles), the result in the level's defect vector
      PMG_HIP_CHECK(hipMemcpyAsync(c->stage[0], u_old, sizeof(double) * D.n_dof, hipMemcpyHostToDevice, c->stream));
      PMG_HIP_CHECK(hipMemcpyAsync(c->stage[1], g, sizeof(double) * D.n_dof, hipMemcpyHostToDevice, c->stream));
      PMG_HIP_CHECK(hipMemcpyAsync(c->stage[2], b_quad, sizeof(double) * nq, hipMemcpyHostToDevice, c->stream));
      du = c->stage[0];
      dg = c->stage[1];
      db = c->stage[2];
      dr = D.aux0;
    }
    launch_theta_rhs(D, D.geom, db, nu, sigma, du, dg, dr, c->stream);
    if (mem == PMG_MEM_HOST)
      PMG_HIP_CHECK(hipMemcpyAsync(rhs, dr, sizeof(double) * D.n_dof, hipMemcpyDeviceToHost, c->stream));
    io.sync();
  });
}

int pmg_batch_create(int device, const pmg_params *params, int n_inst, const pmg_instance_desc *inst,
                     pmg_ctx **out) {
  if (!out || n_inst <= 0 || !inst) return PMG_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  pmg_ctx *c = nullptr;
  if (int st = pmg_ctx_create(device, params, &c)) return st;
  c->batch = true;
  const int rc = guard(c, [&] {
    if (c->prm.sc_minit != 0) throw Fail{PMG_ERR_UNSUPPORTED, "a batch needs sc_minit = 0"};
    if (c->prm.sc_smoother_saddle != 2)
      throw Fail{PMG_ERR_UNSUPPORTED, "a batch runs the colored cell Vanka (smoother 2)"};
    MergedHierarchy M;
    try {
      merge_hierarchies(n_inst, inst, M);
    } catch (const Invalid &e) {
      throw Fail{e.code, e.what};
    }
    for (size_t l = 0; l < M.L.size(); ++l) {
      int lvl = -1;
      if (int r = add_level(c, &M.L[l].d, &lvl, true)) throw Fail{r, c->err};
      DevLevel &D = c->L[l];
      D.coarse_insts = M.L[l].coarse;
      {
        const MergedLevel &ML = M.L[l];
        D.h_udof_problem.assign(D.n_udof, 0);
        D.h_pdof_problem.assign(D.n_pdof, 0);
        for (int cell = 0; cell < D.n_cells; ++cell) {
          const int b = ML.cell_problem[cell];
          for (int j = 0; j < NU_LOC; ++j) D.h_udof_problem[ML.u_global[ML.u_begin[cell] + j]] = b;
          for (int j = 0; j < NP_LOC; ++j) D.h_pdof_problem[ML.p_global[ML.p_begin[cell] + j]] = b;
        }
      }
      if (l > 0) {
        std::vector<int> reo(D.n_cells);
        PMG_HIP_CHECK(hipMemcpy(reo.data(), D.cell_reorder, sizeof(int) * D.n_cells, hipMemcpyDeviceToHost));
        D.h_pos_problem.resize(D.n_cells);
        for (int q = 0; q < D.n_cells; ++q) D.h_pos_problem[q] = M.L[l].cell_problem[reo[q]];
      }
      {
        // runs of one problem: the merged numbering keeps each problem's dofs together
        // (active, Dirichlet and pressure ranges), and a color's work items are sorted by
        // row, so there are a few runs per problem and a rebuild copies runs
        auto runs = [](long n, std::vector<DevLevel::Run> &out, auto &&prob) {
          for (long i = 0; i < n;) {
            const int b = prob(i);
            long j = i + 1;
            while (j < n && prob(j) == b) ++j;
            out.push_back({i, j, b});
            i = j;
          }
        };
        runs(D.n_udof, D.h_urun, [&](long i) { return D.h_udof_problem[i]; });
        runs(D.n_pdof, D.h_prun, [&](long i) { return D.h_pdof_problem[i]; });
        D.h_irun_off.assign(1, 0);
        for (int col = 0; l > 0 && col < D.n_colors; ++col) {
          const long q0 = long(D.ptr_colors[col]) * (NU_LOC + NP_LOC);
          const long q1 = long(D.ptr_colors[col + 1]) * (NU_LOC + NP_LOC);
          const int p0 = D.ptr_colors[col];
          std::vector<DevLevel::Run> cr;
          runs(q1 - q0, cr, [&](long k) { return D.h_pos_problem[p0 + (D.h_res_items[q0 + k] >> 5)]; });
          for (auto &r : cr) D.h_irun.push_back({q0 + r.b, q0 + r.e, r.prob});
          D.h_irun_off.push_back(long(D.h_irun.size()));
        }
      }
      D.n_l20_seg = int(M.L[l].l20.size() / 2);
      if (D.n_l20_seg) {
        D.l20_seg = dalloc<int>(c, M.L[l].l20.size());
        h2d(c, D.l20_seg, M.L[l].l20.data(), M.L[l].l20.size());
      }
      PMG_HIP_CHECK(hipStreamSynchronize(c->stream));
      M.L[l] = MergedLevel{};   // host copy no longer needed
    }
    c->bi.resize(n_inst);
    std::vector<int4> chunks;
    std::vector<int> cptr(1, 0);
    for (int b = 0; b < n_inst; ++b) {
      BatchProblem &B = c->bi[b];
      B.n_levels = inst[b].n_levels;
      B.clevel = M.clevel[b];
      B.project = inst[b].project_pressure ? 1 : 0;
      upload_operator(c, B.cl, &inst[b].levels[0], false);
      B.n_dof_coarse = B.cl.n_dof;
      B.n_dof_fine = int(M.fidx[b].size());
      B.cidx = dalloc<int>(c, M.cidx[b].size());
      h2d(c, B.cidx, M.cidx[b].data(), M.cidx[b].size());
      B.fidx = dalloc<int>(c, M.fidx[b].size());
      h2d(c, B.fidx, M.fidx[b].data(), M.fidx[b].size());
      B.cf = dalloc<double>(c, B.n_dof_coarse);
      B.cu = dalloc<double>(c, B.n_dof_coarse);
      B.io = dalloc<double>(c, 2 * size_t(B.n_dof_fine));
      for (const auto &r : M.fseg[b])
        for (long s0 = r[0]; s0 < r[1]; s0 += SEG_CHUNK)
          chunks.push_back(make_int4(int(s0), int(std::min<long>(SEG_CHUNK, r[1] - s0)), b, 0));
      cptr.push_back(int(chunks.size()));
      PMG_HIP_CHECK(hipStreamSynchronize(c->stream));
    }
    c->n_seg_ch = int(chunks.size());
    c->seg_ch = dalloc<int4>(c, chunks.size());
    h2d(c, c->seg_ch, chunks.data(), chunks.size());
    c->seg_cptr = dalloc<int>(c, cptr.size());
    h2d(c, c->seg_cptr, cptr.data(), cptr.size());
    c->seg_ch_act = c->seg_ch;
    c->n_seg_ch_act = c->n_seg_ch;
    c->seg_cptr_act = c->seg_cptr;
    c->h_seg_ch = chunks;
    c->h_seg_cptr = cptr;
    c->seg_part = dalloc<double>(c, chunks.size());
    const long n = c->L.back().n_dof;
    c->bX = dalloc<double>(c, n);
    c->bB = dalloc<double>(c, n);
    launch_fill<double>(c->bX, 0.0, n, c->stream);
    launch_fill<double>(c->bB, 0.0, n, c->stream);
    PMG_HIP_CHECK(hipStreamSynchronize(c->stream));
  });
  if (rc) {
    pmg_ctx_destroy(c);
    return rc;
  }
  *out = c;
  return PMG_OK;
}

int pmg_batch_size(const pmg_ctx *c) { return c && c->batch ? int(c->bi.size()) : 0; }

int pmg_batch_fgmres_solve(pmg_ctx *c, double *const *sol, const double *const *rhs, int *iters, double *final_res,
                           int *status, int mem) {
  if (!c || !sol || !rhs) return PMG_ERR_INVALID_ARGUMENT;
  if (!c->batch) {
    c->err = "not a batch context";
    return PMG_ERR_UNSUPPORTED;
  }
  const int nb = int(c->bi.size());
  std::vector<int> it(nb, 0), st(nb, PMG_OK);
  std::vector<double> fr(nb, 0.0);
  const int rc = guard(c, [&] {
    PMG_NEED_READY();
    if (mem != PMG_MEM_HOST && mem != PMG_MEM_DEVICE) throw Fail{PMG_ERR_INVALID_ARGUMENT, "bad memspace"};
    for (int b = 0; b < nb; ++b)
      if (!sol[b] || !rhs[b]) throw Fail{PMG_ERR_INVALID_ARGUMENT, "null problem vector"};
    for (int b = 0; b < nb; ++b) {
      BatchProblem &B = c->bi[b];
      const double *xs = sol[b], *fs = rhs[b];
      if (mem == PMG_MEM_HOST) {
        PMG_HIP_CHECK(hipMemcpyAsync(B.io, sol[b], sizeof(double) * B.n_dof_fine, hipMemcpyHostToDevice, c->stream));
        PMG_HIP_CHECK(hipMemcpyAsync(B.io + B.n_dof_fine, rhs[b], sizeof(double) * B.n_dof_fine,
                                     hipMemcpyHostToDevice, c->stream));
        xs = B.io;
        fs = B.io + B.n_dof_fine;
      }
      launch_scatter_d<double>(c->bX, xs, B.fidx, B.n_dof_fine, c->stream);
      launch_scatter_d<double>(c->bB, fs, B.fidx, B.n_dof_fine, c->stream);
    }
    fgmres_batch(c, it, fr, st);
    for (int b = 0; b < nb; ++b) {
      BatchProblem &B = c->bi[b];
      double *xd = mem == PMG_MEM_HOST ? B.io : sol[b];
      launch_gather_d<double>(xd, c->bX, B.fidx, B.n_dof_fine, c->stream);
      if (mem == PMG_MEM_HOST)
        PMG_HIP_CHECK(hipMemcpyAsync(sol[b], B.io, sizeof(double) * B.n_dof_fine, hipMemcpyDeviceToHost, c->stream));
    }
    PMG_HIP_CHECK(hipStreamSynchronize(c->stream));
    check_deferred(c);
  });
  for (int b = 0; b < nb; ++b) {
    if (iters) iters[b] = it[b];
    if (final_res) final_res[b] = fr[b];
    if (status) status[b] = rc ? rc : st[b];
  }
  if (rc) return rc;
  for (int b = 0; b < nb; ++b)
    if (st[b] != PMG_OK) {
      c->err = "problem " + std::to_string(b) + ": " + pmg_status_string(st[b]);
      return st[b];
    }
  return PMG_OK;
}

int pmg_fgmres_history(const pmg_ctx *c, double *out, int cap) {
  if (!c || !out) return 0;
  const int n = std::min<int>(cap, int(c->hist.size()));
  for (int i = 0; i < n; ++i) out[i] = c->hist[i];
  return n;
}

int pmg_device_alloc(pmg_ctx *c, size_t bytes, void **ptr) {
  if (!c || !ptr) return PMG_ERR_INVALID_ARGUMENT;
  return guard(c, [&] { *ptr = dalloc<char>(c, bytes); });
}
int pmg_device_free(pmg_ctx *c, void *ptr) {
  if (!c) return PMG_ERR_INVALID_ARGUMENT;
  dfree(c, ptr);
  return PMG_OK;
}
int pmg_memcpy(pmg_ctx *c, void *dst, const void *src, size_t bytes, int kind) {
  if (!c || kind < 0 || kind > 2) return PMG_ERR_INVALID_ARGUMENT;
  return guard(c, [&] {
    const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    PMG_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, k, c->stream));
    PMG_HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}
int pmg_synchronize(pmg_ctx *c) {
  if (!c) return PMG_ERR_INVALID_ARGUMENT;
  return guard(c, [&] {
    PMG_HIP_CHECK(hipStreamSynchronize(c->stream));
    PMG_HIP_CHECK(hipStreamSynchronize(c->stream2));
    check_deferred(c);
  });
}
int pmg_timers_get(const pmg_ctx *c, pmg_timers *t) {
  if (!c || !t) return PMG_ERR_INVALID_ARGUMENT;
  *t = c->tm;
  return PMG_OK;
}
int pmg_timers_reset(pmg_ctx *c) {
  if (!c) return PMG_ERR_INVALID_ARGUMENT;
  c->tm = pmg_timers{};
  return PMG_OK;
}
int pmg_set_profiling(pmg_ctx *c, int on) {
  if (!c) return PMG_ERR_INVALID_ARGUMENT;
  c->profiling = on != 0;
  return PMG_OK;
}

}  // extern "C"
