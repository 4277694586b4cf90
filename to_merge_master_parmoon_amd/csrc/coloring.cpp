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

// TNSE_MGLevel4::CellVanka, Gauss-Seidel variant (_SEQ, src/FE/NSE_MGLevel4.C:1653-2085),
// visits the cells in collection order and updates u after every cell.  Cell j's local
// residual reads the rows of its patch dofs, i.e. every dof in a column of those rows, so
// its result depends on an earlier cell i exactly when i's patch dofs meet j's columns or
// j's patch dofs meet i's columns ("coupled").  Level sets with
//     set(j) = 1 + max { set(i) : i < j, i coupled with j }   (0 if there is none)
// list every coupled pair in collection order and hold only uncoupled cells together, so
// running the sets one after the other, each as one colored step (all local residuals of
// the set, then all updates), performs the Gauss-Seidel sweep's arithmetic exactly.
// Output as pmg_color_cells: sets in order, cells ascending inside a set.
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
  // dof -> cells for velocity and pressure dofs
  auto invert = [&](const int *g, int nl, int ndof, std::vector<int> &ptr, std::vector<int> &cells) {
    ptr.assign(size_t(ndof) + 1, 0);
    for (long k = 0; k < long(n_cells) * nl; ++k) {
      if (g[k] < 0 || g[k] >= ndof) return false;
      ptr[size_t(g[k]) + 1]++;
    }
    for (int i = 0; i < ndof; ++i) ptr[i + 1] += ptr[i];
    cells.resize(ptr[ndof]);
    std::vector<int> pos(ptr.begin(), ptr.end() - 1);
    for (int c = 0; c < n_cells; ++c)
      for (int j = 0; j < nl; ++j) cells[pos[g[long(c) * nl + j]]++] = c;
    return true;
  };
  std::vector<int> up, uc, pp, pc;
  if (!invert(u_global, NU, n_udof, up, uc) || !invert(p_global, NP, n_pdof, pp, pc)) return PMG_ERR_INVALID_ARGUMENT;
  // forward relation: cells whose patch dofs meet the columns of cell j's rows
  std::vector<int> mark(n_cells, -1);
  std::vector<std::vector<int>> nb(n_cells);
  for (int j = 0; j < n_cells; ++j) {
    auto hit_u = [&](int col) {
      for (int q = up[col]; q < up[col + 1]; ++q)
        if (mark[uc[q]] != j) {
          mark[uc[q]] = j;
          nb[j].push_back(uc[q]);
        }
    };
    auto hit_p = [&](int col) {
      for (int q = pp[col]; q < pp[col + 1]; ++q)
        if (mark[pc[q]] != j) {
          mark[pc[q]] = j;
          nb[j].push_back(pc[q]);
        }
    };
    for (int l = 0; l < NU; ++l) {
      const int r = u_global[long(j) * NU + l];
      for (int k = a_rowptr[r]; k < a_rowptr[r + 1]; ++k) hit_u(a_kcol[k]);
      for (int k = bt_rowptr[r]; k < bt_rowptr[r + 1]; ++k) hit_p(bt_kcol[k]);
    }
    for (int l = 0; l < NP; ++l) {
      const int r = p_global[long(j) * NP + l];
      for (int k = b_rowptr[r]; k < b_rowptr[r + 1]; ++k) hit_u(b_kcol[k]);
    }
  }
  // symmetric closure: i coupled with j if either meets the other's columns
  for (int j = 0; j < n_cells; ++j)
    for (int i : std::vector<int>(nb[j]))
      if (i != j) nb[i].push_back(j);
  std::vector<int> set(n_cells, 0);
  int nsets = 0;
  for (int j = 0; j < n_cells; ++j) {
    int s = 0;
    for (int i : nb[j])
      if (i < j && set[i] + 1 > s) s = set[i] + 1;
    set[j] = s;
    if (s + 1 > nsets) nsets = s + 1;
  }
  std::vector<int> cnt(size_t(nsets) + 1, 0);
  for (int c = 0; c < n_cells; ++c) cnt[set[c] + 1]++;
  for (int k = 0; k < nsets; ++k) cnt[k + 1] += cnt[k];
  for (int k = 0; k <= nsets; ++k) ptr_sets[k] = cnt[k];
  std::vector<int> fill(cnt.begin(), cnt.end() - 1);
  for (int c = 0; c < n_cells; ++c) cell_reorder[fill[set[c]]++] = c;
  *n_sets = nsets;
  return PMG_OK;
}
