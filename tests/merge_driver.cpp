// Test driver for csrc/batch_merge.cpp (built by tests/test_batch_merge.py with g++, no GPU):
// four generated problems of different depths, sizes, Dirichlet faces and perturbations are
// merged; every merged level must reproduce each problem's operator exactly through the
// index maps, the colorings must be valid partitions, and the problems solved directly on a
// merged level must have no parents there.  Prints OK or FAIL.
#include <cstdio>
#include <cmath>
#include <random>
#include <vector>
#include "batch_merge.hpp"
#include "pmg_problem.h"
using namespace pmg;
static void matvec(const pmg_level_desc &d, const std::vector<double> &x, std::vector<double> &y) {
  int NU = d.n_udof, NP = d.n_pdof;
  y.assign(3 * NU + NP, 0.0);
  for (int r = 0; r < NU; ++r)
    for (int k = d.a_rowptr[r]; k < d.a_rowptr[r + 1]; ++k)
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) y[a * NU + r] += d.a[a * 3 + b][k] * x[b * NU + d.a_kcol[k]];
  for (int r = 0; r < NU; ++r)
    for (int k = d.bt_rowptr[r]; k < d.bt_rowptr[r + 1]; ++k)
      for (int a = 0; a < 3; ++a) y[a * NU + r] += d.bt[a][k] * x[3 * NU + d.bt_kcol[k]];
  for (int r = 0; r < NP; ++r)
    for (int k = d.b_rowptr[r]; k < d.b_rowptr[r + 1]; ++k)
      for (int a = 0; a < 3; ++a) y[3 * NU + r] += d.b[a][k] * x[a * NU + d.b_kcol[k]];
}
int main() {
  struct Spec { int m, levels, faces; double nu; int seed; } specs[] = {{4, 2, 63, 0.4, 1}, {3, 3, 21, 0.1, 2}, {4, 3, 63, 0.01, 5}, {5, 2, 21, 0.3, 9}};
  const int nb = 4;
  std::vector<pmgp_problem *> P(nb);
  std::vector<std::vector<pmg_level_desc>> L(nb);
  std::vector<pmg_instance_desc> inst(nb);
  for (int i = 0; i < nb; ++i) {
    pmgp_config c; pmgp_config_default(&c);
    c.cx = c.cy = c.cz = specs[i].m; c.n_levels = specs[i].levels; c.nu = specs[i].nu;
    c.dirichlet_faces = specs[i].faces; c.perturb = 0.1; c.seed_perturb = 4000 + specs[i].seed;
    c.conv_kind = PMGP_CONV_FOURIER; c.seed_conv = 3000 + i; c.rhs_kind = PMGP_RHS_FOURIER;
    if (pmgp_create(&c, &P[i])) { printf("create failed\n"); return 1; }
    L[i].resize(specs[i].levels);
    for (int l = 0; l < specs[i].levels; ++l) pmgp_level_desc(P[i], l, &L[i][l]);
    inst[i] = {specs[i].levels, L[i].data(), specs[i].faces == 63};
  }
  MergedHierarchy M;
  try { merge_hierarchies(nb, inst.data(), M); } catch (const Invalid &e) { printf("invalid %s\n", e.what); return 1; }
  const int S = int(M.L.size());
  printf("S=%d\n", S);
  std::mt19937_64 rng(7);
  std::normal_distribution<double> N01;
  int bad = 0;
  for (int l = 0; l < S; ++l) {
    const pmg_level_desc &d = M.L[l].d;
    const int n = 3 * d.n_udof + d.n_pdof;
    std::vector<double> x(n), y;
    for (auto &v : x) v = N01(rng);
    matvec(d, x, y);
    // per problem: gather x into its numbering, multiply, compare
    std::vector<int> seen(n, 0);
    for (int i = 0; i < nb; ++i) {
      const int li = l - (S - specs[i].levels);
      if (li < 0) continue;
      const pmg_level_desc &di = L[i][li];
      // index map for this level: rebuild from u_global of cells (cell maps) is indirect; use vmap via cidx/fidx when available
      const std::vector<int> *idx = nullptr;
      if (li == 0) idx = &M.cidx[i];
      if (l == S - 1) idx = &M.fidx[i];
      if (!idx) continue;
      const int ni = 3 * di.n_udof + di.n_pdof;
      if ((int)idx->size() != ni) { printf("idx size\n"); bad++; continue; }
      std::vector<double> xi(ni), yi;
      for (int k = 0; k < ni; ++k) { xi[k] = x[(*idx)[k]]; seen[(*idx)[k]]++; }
      matvec(di, xi, yi);
      double err = 0, nrm = 0;
      for (int k = 0; k < ni; ++k) { err = std::max(err, std::fabs(yi[k] - y[(*idx)[k]])); nrm = std::max(nrm, std::fabs(yi[k])); }
      printf("level %d problem %d li %d: err %.3e (|y| %.3e)\n", l, i, li, err, nrm);
      if (err > 1e-12 * nrm) bad++;
    }
    if (l == S - 1) for (int k = 0; k < n; ++k) if (seen[k] != 1) { bad++; printf("fine index %d seen %d\n", k, seen[k]); break; }
    // colors: cells of a color share no velocity dof; swept + excluded = permutation
    const auto &ML = M.L[l];
    std::vector<int> mark(d.n_udof, -1);
    for (int k = 0; k < ML.n_colors; ++k)
      for (int q = ML.ptr_colors[k]; q < ML.ptr_colors[k + 1]; ++q) {
        int cell = ML.cell_reorder[q];
        for (int j = 0; j < 27; ++j) {
          const int g = d.u_global[d.u_begin[cell] + j];
          if (mark[g] == k) bad++;   // a dof met twice in one color: two cells share it
          mark[g] = k;
        }
      }
    std::vector<int> perm(d.n_cells, 0);
    for (int q = 0; q < d.n_cells; ++q) perm[ML.cell_reorder[q]]++;
    for (int q = 0; q < d.n_cells; ++q) if (perm[q] != 1) { bad++; break; }
    printf("level %d: n_dof %d cells %d colors %d swept %d coarse problems %zu l20 %zu\n", l, n, d.n_cells, ML.n_colors, ML.ptr_colors.back(), ML.coarse.size(), ML.l20.size() / 2);
    // parents: fine cell's parent's dofs should be a coarse cell of the same problem
    if (l > 0) {
      // cells of the problems solved directly on this level have no parent, all others do
      int np = 0;
      for (int q = 0; q < d.n_cells; ++q) {
        const int prob = ML.cell_problem[q];
        bool direct = false;
        for (int k : ML.coarse) direct = direct || k == prob;
        if ((d.parent[q] >= 0) == direct) bad++;
        np += d.parent[q] >= 0;
      }
      printf("  cells with parent %d\n", np);
    }
  }
  printf(bad ? "FAIL %d\n" : "OK\n", bad);
  return bad != 0;
}
