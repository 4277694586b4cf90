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
