// coloring.cpp - TParFEMapper3D::colorCell (src/Parallel/ParFEMapper3D.C:750-1084) in
// linear time.  The reference builds the cell adjacency by comparing the velocity dof
// lists of every pair of cells (O(N^2 * 27^2), :900-935) and then colors greedily in cell
// index order with the smallest color not used by an already colored neighbour
// (:978-1020).  Two cells are neighbours iff they share a velocity dof, so the same
// neighbour set is obtained from a dof -> cells map; the greedy pass is unchanged, hence
// the coloring is identical.  Cells are then listed color by color in increasing index
// (:1034-1052).
#include <vector>

#include "pmg.h"

extern "C" int pmg_color_cells(int n_cells, int n_ulocal, const int *u_begin, const int *u_global, int n_udof,
                               int *n_colors, int *ptr_colors, int *cell_reorder) {
  if (n_cells < 0 || n_ulocal <= 0 || !u_begin || !u_global || !n_colors || !ptr_colors || !cell_reorder)
    return PMG_ERR_INVALID_ARGUMENT;
  if (n_cells == 0) {
    *n_colors = 0;
    ptr_colors[0] = 0;
    return PMG_OK;
  }
  // dof -> cells
  std::vector<int> cnt(size_t(n_udof) + 1, 0);
  for (int c = 0; c < n_cells; ++c)
    for (int j = 0; j < n_ulocal; ++j) {
      const int g = u_global[u_begin[c] + j];
      if (g < 0 || g >= n_udof) return PMG_ERR_INVALID_ARGUMENT;
      cnt[size_t(g) + 1]++;
    }
  for (int i = 0; i < n_udof; ++i) cnt[i + 1] += cnt[i];
  std::vector<int> d2c(cnt[n_udof]);
  std::vector<int> pos(cnt.begin(), cnt.end() - 1);
  for (int c = 0; c < n_cells; ++c)
    for (int j = 0; j < n_ulocal; ++j) d2c[pos[u_global[u_begin[c] + j]]++] = c;
  std::vector<int> color(n_cells, -1);
  std::vector<int> mark;   // mark[k] == ii: color k taken by a neighbour of cell ii
  int ncol = 0;
  for (int ii = 0; ii < n_cells; ++ii) {
    for (int j = 0; j < n_ulocal; ++j) {
      const int g = u_global[u_begin[ii] + j];
      for (int q = cnt[g]; q < cnt[g + 1]; ++q) {
        const int nb = d2c[q];
        const int k = color[nb];
        if (k >= 0) mark[k] = ii;
      }
    }
    int cr = 0;
    while (cr < int(mark.size()) && mark[cr] == ii) ++cr;
    if (cr == int(mark.size())) mark.push_back(-1);
    color[ii] = cr;
    if (cr + 1 > ncol) ncol = cr + 1;
  }
  std::vector<int> cc(size_t(ncol) + 1, 0);
  for (int c = 0; c < n_cells; ++c) cc[color[c] + 1]++;
  for (int k = 0; k < ncol; ++k) cc[k + 1] += cc[k];
  for (int k = 0; k <= ncol; ++k) ptr_colors[k] = cc[k];
  std::vector<int> fill(cc.begin(), cc.end() - 1);
  for (int c = 0; c < n_cells; ++c) cell_reorder[fill[color[c]]++] = c;
  *n_colors = ncol;
  return PMG_OK;
}

// Level sets of a sequential patch order.  Patch j (j = 0 .. n_patches-1, visited in this
// order by the reference) owns the velocity dofs vidx[vptr[j] .. vptr[j+1]) and the pressure
// dofs pidx[pptr[j] .. pptr[j+1]); its local residual reads the rows of those dofs, i.e. every
// dof in a column of those rows, so its result depends on an earlier patch i exactly when
// i's dofs meet j's columns or j's dofs meet i's columns ("coupled").  Level sets with
//     set(j) = 1 + max { set(i) : i < j, i coupled with j }   (0 if there is none)
// list every coupled pair in visiting order and hold only uncoupled patches together, so
// running the sets one after the other, each as one colored step, performs the
// sequential sweep's arithmetic exactly.
namespace {
int patch_level_sets(int n_patches, const std::vector<int> &vptr, const std::vector<int> &vidx,
                     const std::vector<int> &pptr, const std::vector<int> &pidx, int n_udof, int n_pdof,
                     const int *a_rowptr, const int *a_kcol, const int *bt_rowptr, const int *bt_kcol,
                     const int *b_rowptr, const int *b_kcol, int *n_sets, int *ptr_sets, int *order) {
  // dof -> patches
  auto invert = [&](const std::vector<int> &ptr, const std::vector<int> &idx, int ndof, std::vector<int> &dp,
                    std::vector<int> &pc) {
    dp.assign(size_t(ndof) + 1, 0);
    for (int g : idx) {
      if (g < 0 || g >= ndof) return false;
      dp[size_t(g) + 1]++;
    }
    for (int i = 0; i < ndof; ++i) dp[i + 1] += dp[i];
    pc.resize(dp[ndof]);
    std::vector<int> pos(dp.begin(), dp.end() - 1);
    for (int q = 0; q < n_patches; ++q)
      for (int k = ptr[q]; k < ptr[q + 1]; ++k) pc[pos[idx[k]]++] = q;
    return true;
  };
  std::vector<int> up, uc, pp, pc;
  if (!invert(vptr, vidx, n_udof, up, uc) || !invert(pptr, pidx, n_pdof, pp, pc)) return PMG_ERR_INVALID_ARGUMENT;
  // forward relation: patches whose dofs meet the columns of patch j's rows
  std::vector<int> mark(n_patches, -1);
  std::vector<std::vector<int>> nb(n_patches);
  for (int j = 0; j < n_patches; ++j) {
    auto hit = [&](const std::vector<int> &dp, const std::vector<int> &dc, int col) {
      for (int q = dp[col]; q < dp[col + 1]; ++q)
        if (mark[dc[q]] != j) {
          mark[dc[q]] = j;
          nb[j].push_back(dc[q]);
        }
    };
    for (int k = vptr[j]; k < vptr[j + 1]; ++k) {
      const int r = vidx[k];
      for (int e = a_rowptr[r]; e < a_rowptr[r + 1]; ++e) hit(up, uc, a_kcol[e]);
      for (int e = bt_rowptr[r]; e < bt_rowptr[r + 1]; ++e) hit(pp, pc, bt_kcol[e]);
    }
    for (int k = pptr[j]; k < pptr[j + 1]; ++k) {
      const int r = pidx[k];
      for (int e = b_rowptr[r]; e < b_rowptr[r + 1]; ++e) hit(up, uc, b_kcol[e]);
    }
  }
  // symmetric closure: i coupled with j if either meets the other's columns
  for (int j = 0; j < n_patches; ++j)
    for (int i : std::vector<int>(nb[j]))
      if (i != j) nb[i].push_back(j);
  std::vector<int> set(n_patches, 0);
  int nsets = 0;
  for (int j = 0; j < n_patches; ++j) {
    int sj = 0;
    for (int i : nb[j])
      if (i < j && set[i] + 1 > sj) sj = set[i] + 1;
    set[j] = sj;
    if (sj + 1 > nsets) nsets = sj + 1;
  }
  std::vector<int> cnt(size_t(nsets) + 1, 0);
  for (int q = 0; q < n_patches; ++q) cnt[set[q] + 1]++;
  for (int k = 0; k < nsets; ++k) cnt[k + 1] += cnt[k];
  for (int k = 0; k <= nsets; ++k) ptr_sets[k] = cnt[k];
  std::vector<int> fill(cnt.begin(), cnt.end() - 1);
  for (int q = 0; q < n_patches; ++q) order[fill[set[q]]++] = q;
  *n_sets = nsets;
  return PMG_OK;
}
}  // namespace

// TNSE_MGLevel4::CellVanka, Gauss-Seidel variant (_SEQ, src/FE/NSE_MGLevel4.C:1653-2085):
// cells in collection order, each patch its 27 velocity and 4 pressure dofs.  Output as
// pmg_color_cells: sets in order, cells ascending inside a set.
extern "C" int pmg_sequential_schedule(int n_cells, const int *u_global, const int *p_global, int n_udof, int n_pdof,
                                       const int *a_rowptr, const int *a_kcol, const int *bt_rowptr,
                                       const int *bt_kcol, const int *b_rowptr, const int *b_kcol, int *n_sets,
                                       int *ptr_sets, int *cell_reorder) {
  constexpr int NU = 27, NP = 4;
  if (n_cells < 0 || n_udof <= 0 || n_pdof <= 0 || !u_global || !p_global || !a_rowptr || !a_kcol || !bt_rowptr ||
      !bt_kcol || !b_rowptr || !b_kcol || !n_sets || !ptr_sets || !cell_reorder)
    return PMG_ERR_INVALID_ARGUMENT;
  if (n_cells == 0) {
    *n_sets = 0;
    ptr_sets[0] = 0;
    return PMG_OK;
  }
  std::vector<int> vptr(size_t(n_cells) + 1), pptr(size_t(n_cells) + 1);
  for (int c = 0; c <= n_cells; ++c) {
    vptr[c] = c * NU;
    pptr[c] = c * NP;
  }
  std::vector<int> vidx(u_global, u_global + size_t(n_cells) * NU), pidx(p_global, p_global + size_t(n_cells) * NP);
  return patch_level_sets(n_cells, vptr, vidx, pptr, pidx, n_udof, n_pdof, a_rowptr, a_kcol, bt_rowptr, bt_kcol,
                          b_rowptr, b_kcol, n_sets, ptr_sets, cell_reorder);
}

// TNSE_MGLevel4::NodalVanka (src/FE/NSE_MGLevel4.C:4098-4549): pressure dofs in order, each
// patch the velocity dofs of its B row and the pressure dof itself.  Output: sets in order,
// pressure dofs ascending inside a set (ptr_sets holds n_pdof + 1 ints).
extern "C" int pmg_nodal_schedule(int n_udof, int n_pdof, const int *a_rowptr, const int *a_kcol,
                                  const int *bt_rowptr, const int *bt_kcol, const int *b_rowptr, const int *b_kcol,
                                  int *n_sets, int *ptr_sets, int *pdof_order) {
  if (n_udof <= 0 || n_pdof <= 0 || !a_rowptr || !a_kcol || !bt_rowptr || !bt_kcol || !b_rowptr || !b_kcol ||
      !n_sets || !ptr_sets || !pdof_order)
    return PMG_ERR_INVALID_ARGUMENT;
  std::vector<int> vptr(b_rowptr, b_rowptr + size_t(n_pdof) + 1), vidx(b_kcol, b_kcol + b_rowptr[n_pdof]);
  std::vector<int> pptr(size_t(n_pdof) + 1), pidx(n_pdof);
  for (int i = 0; i <= n_pdof; ++i) pptr[i] = i;
  for (int i = 0; i < n_pdof; ++i) pidx[i] = i;
  return patch_level_sets(n_pdof, vptr, vidx, pptr, pidx, n_udof, n_pdof, a_rowptr, a_kcol, bt_rowptr, bt_kcol,
                          b_rowptr, b_kcol, n_sets, ptr_sets, pdof_order);
}
