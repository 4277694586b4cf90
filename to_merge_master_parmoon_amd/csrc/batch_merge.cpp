// batch_merge.cpp - block-diagonal merge of independent multigrid hierarchies (layout in
// batch_merge.hpp).  Pure host code.
#include "batch_merge.hpp"

#include <algorithm>
#include <cstring>

namespace pmg {
namespace {

constexpr int NUL = 27;   // Q2 hexahedron velocity dofs per cell
constexpr int NPL = 4;    // P1disc pressure dofs per cell

struct Part {             // one problem's level inside a merged level
  int inst = 0, li = 0;
  const pmg_level_desc *d = nullptr;
  int act_off = 0, dir_off = 0, p_off = 0, cell_off = 0;
};

void validate(const pmg_level_desc *d, int li) {
  if (d->n_ulocal != NUL || d->n_plocal != NPL)
    throw Invalid{PMG_ERR_UNSUPPORTED, "only Q2/P1disc hexahedra (27 + 4 local dofs)"};
  if (d->n_udof <= 0 || d->n_pdof <= 0 || d->n_cells <= 0 || d->n_active < 0 || d->n_active > d->n_udof)
    throw Invalid{PMG_ERR_INVALID_ARGUMENT, "bad level sizes"};
  if (d->hanging_bound != d->n_active) throw Invalid{PMG_ERR_UNSUPPORTED, "hanging nodes are not supported"};
  if (d->n_active + d->n_dirichlet != d->n_udof) throw Invalid{PMG_ERR_INVALID_ARGUMENT, "dof bounds inconsistent"};
  if (!d->a_rowptr || !d->a_kcol || !d->b_rowptr || !d->b_kcol || !d->bt_rowptr || !d->bt_kcol || !d->u_begin ||
      !d->u_global || !d->p_begin || !d->p_global)
    throw Invalid{PMG_ERR_INVALID_ARGUMENT, "missing level arrays"};
  for (int b = 0; b < 9; ++b)
    if (!d->a[b]) throw Invalid{PMG_ERR_INVALID_ARGUMENT, "missing A block"};
  for (int b = 0; b < 3; ++b)
    if (!d->b[b] || !d->bt[b]) throw Invalid{PMG_ERR_INVALID_ARGUMENT, "missing B/BT block"};
  if (d->a_rowptr[0] != 0 || d->a_rowptr[d->n_udof] != d->a_nnz || d->b_rowptr[d->n_pdof] != d->b_nnz ||
      d->bt_rowptr[d->n_udof] != d->bt_nnz)
    throw Invalid{PMG_ERR_INVALID_ARGUMENT, "CSR row pointers inconsistent with nnz"};
  if (li > 0 && (!d->parent || !d->child || !d->qq_u || !d->qq_p || d->n_children != 8))
    throw Invalid{PMG_ERR_INVALID_ARGUMENT, "fine level without parent/child/QQ data (8 children)"};
}

}  // namespace

void merge_hierarchies(int n_inst, const pmg_instance_desc *inst, MergedHierarchy &M) {
  if (n_inst <= 0 || !inst) throw Invalid{PMG_ERR_INVALID_ARGUMENT, "no problems"};
  int S = 0;
  for (int i = 0; i < n_inst; ++i) {
    if (inst[i].n_levels < 2 || !inst[i].levels)
      throw Invalid{PMG_ERR_INVALID_ARGUMENT, "every problem of a batch needs at least two levels"};
    S = std::max(S, inst[i].n_levels);
  }
  M.L.assign(S, MergedLevel{});
  M.clevel.assign(n_inst, -1);
  M.cidx.assign(n_inst, {});
  M.fidx.assign(n_inst, {});
  M.fseg.assign(n_inst, {});
  std::vector<std::vector<Part>> parts(S);
  std::vector<std::vector<int>> part_of(S, std::vector<int>(n_inst, -1));   // merged level, problem -> part

  for (int l = 0; l < S; ++l) {
    MergedLevel &ML = M.L[l];
    int na = 0, nd = 0, np = 0, nc = 0;
    long nnz_a = 0, nnz_b = 0, nnz_bt = 0;
    for (int i = 0; i < n_inst; ++i) {
      const int li = l - (S - inst[i].n_levels);
      if (li < 0) continue;
      const pmg_level_desc *d = &inst[i].levels[li];
      validate(d, li);
      Part p;
      p.inst = i; p.li = li; p.d = d;
      p.act_off = na; p.dir_off = nd; p.p_off = np; p.cell_off = nc;
      na += d->n_active; nd += d->n_dirichlet; np += d->n_pdof; nc += d->n_cells;
      nnz_a += d->a_nnz; nnz_b += d->b_nnz; nnz_bt += d->bt_nnz;
      part_of[l][i] = int(parts[l].size());
      parts[l].push_back(p);
    }
    if (nnz_a > 0x7fffffffL || nnz_b > 0x7fffffffL || nnz_bt > 0x7fffffffL)
      throw Invalid{PMG_ERR_UNSUPPORTED, "merged level exceeds 2^31 matrix entries"};
    const int NU = na + nd, NA = na;
    auto vmap = [&](const Part &p, int r) {
      return r < p.d->n_active ? p.act_off + r : NA + p.dir_off + (r - p.d->n_active);
    };
    // velocity rows (A, BT): active rows of every problem, then the Dirichlet rows
    ML.a_rowptr.assign(1, 0);
    ML.bt_rowptr.assign(1, 0);
    ML.a_kcol.reserve(nnz_a);
    ML.bt_kcol.reserve(nnz_bt);
    for (int b = 0; b < 9; ++b) ML.a[b].reserve(nnz_a);
    for (int b = 0; b < 3; ++b) ML.bt[b].reserve(nnz_bt);
    for (int pass = 0; pass < 2; ++pass)
      for (const Part &p : parts[l]) {
        const pmg_level_desc *d = p.d;
        const int r0 = pass == 0 ? 0 : d->n_active, r1 = pass == 0 ? d->n_active : d->n_udof;
        for (int r = r0; r < r1; ++r) {
          for (int k = d->a_rowptr[r]; k < d->a_rowptr[r + 1]; ++k) {
            const int col = d->a_kcol[k];
            if (col < 0 || col >= d->n_udof) throw Invalid{PMG_ERR_INVALID_ARGUMENT, "A: column out of range"};
            ML.a_kcol.push_back(vmap(p, col));
            for (int b = 0; b < 9; ++b) ML.a[b].push_back(d->a[b][k]);
          }
          ML.a_rowptr.push_back(int(ML.a_kcol.size()));
          for (int k = d->bt_rowptr[r]; k < d->bt_rowptr[r + 1]; ++k) {
            const int col = d->bt_kcol[k];
            if (col < 0 || col >= d->n_pdof) throw Invalid{PMG_ERR_INVALID_ARGUMENT, "BT: column out of range"};
            ML.bt_kcol.push_back(p.p_off + col);
            for (int b = 0; b < 3; ++b) ML.bt[b].push_back(d->bt[b][k]);
          }
          ML.bt_rowptr.push_back(int(ML.bt_kcol.size()));
        }
      }
    // pressure rows (B)
    ML.b_rowptr.assign(1, 0);
    ML.b_kcol.reserve(nnz_b);
    for (int b = 0; b < 3; ++b) ML.b[b].reserve(nnz_b);
    for (const Part &p : parts[l]) {
      const pmg_level_desc *d = p.d;
      for (int r = 0; r < d->n_pdof; ++r) {
        for (int k = d->b_rowptr[r]; k < d->b_rowptr[r + 1]; ++k) {
          const int col = d->b_kcol[k];
          if (col < 0 || col >= d->n_udof) throw Invalid{PMG_ERR_INVALID_ARGUMENT, "B: column out of range"};
          ML.b_kcol.push_back(vmap(p, col));
          for (int b = 0; b < 3; ++b) ML.b[b].push_back(d->b[b][k]);
        }
        ML.b_rowptr.push_back(int(ML.b_kcol.size()));
      }
    }
    // cells, dof maps, parents
    ML.u_begin.resize(nc);
    ML.p_begin.resize(nc);
    ML.u_global.resize(size_t(nc) * NUL);
    ML.p_global.resize(size_t(nc) * NPL);
    ML.parent.assign(nc, -1);
    ML.child.assign(nc, 0);
    ML.cell_problem.assign(nc, -1);
    for (const Part &p : parts[l]) {
      const pmg_level_desc *d = p.d;
      const Part *pc = p.li > 0 ? &parts[l - 1][part_of[l - 1][p.inst]] : nullptr;
      for (int cell = 0; cell < d->n_cells; ++cell) {
        const int mc = p.cell_off + cell;
        ML.cell_problem[mc] = p.inst;
        ML.u_begin[mc] = mc * NUL;
        ML.p_begin[mc] = mc * NPL;
        for (int j = 0; j < NUL; ++j) {
          const int g = d->u_global[d->u_begin[cell] + j];
          if (g < 0 || g >= d->n_udof) throw Invalid{PMG_ERR_INVALID_ARGUMENT, "velocity dof out of range"};
          ML.u_global[size_t(mc) * NUL + j] = vmap(p, g);
        }
        for (int j = 0; j < NPL; ++j) {
          const int g = d->p_global[d->p_begin[cell] + j];
          if (g < 0 || g >= d->n_pdof) throw Invalid{PMG_ERR_INVALID_ARGUMENT, "pressure dof out of range"};
          ML.p_global[size_t(mc) * NPL + j] = p.p_off + g;
        }
        if (pc) {
          const int par = d->parent[cell], ch = d->child[cell];
          if (par < 0 || par >= pc->d->n_cells || ch < 0 || ch >= 8)
            throw Invalid{PMG_ERR_INVALID_ARGUMENT, "parent/child map out of range"};
          ML.parent[mc] = pc->cell_off + par;
          ML.child[mc] = ch;
        }
      }
    }
    // prolongation tables: one per merged level (reference-element matrices, equal for all)
    for (const Part &p : parts[l]) {
      if (p.li == 0) continue;
      const size_t nu = size_t(8) * NUL * NUL, npp = size_t(8) * NPL * NPL;
      if (ML.qq_u.empty()) {
        ML.qq_u.assign(p.d->qq_u, p.d->qq_u + nu);
        ML.qq_p.assign(p.d->qq_p, p.d->qq_p + npp);
      } else if (std::memcmp(ML.qq_u.data(), p.d->qq_u, nu * sizeof(double)) != 0 ||
                 std::memcmp(ML.qq_p.data(), p.d->qq_p, npp * sizeof(double)) != 0) {
        throw Invalid{PMG_ERR_UNSUPPORTED, "problems of a batch must share the prolongation tables"};
      }
    }
    // colors: color k of the merged level = color k of every swept problem level
    std::vector<std::vector<int>> pptr(parts[l].size()), preo(parts[l].size());
    int ncol = 0;
    for (size_t q = 0; q < parts[l].size(); ++q) {
      const Part &p = parts[l][q];
      if (p.li == 0) continue;
      const pmg_level_desc *d = p.d;
      if (d->ptr_colors && d->cell_reorder && d->n_colors > 0) {
        pptr[q].assign(d->ptr_colors, d->ptr_colors + d->n_colors + 1);
        preo[q].assign(d->cell_reorder, d->cell_reorder + d->n_cells);
        if (pptr[q][0] != 0 || pptr[q].back() != d->n_cells)
          throw Invalid{PMG_ERR_INVALID_ARGUMENT, "coloring does not cover the cells"};
      } else {
        pptr[q].resize(d->n_cells + 1);
        preo[q].resize(d->n_cells);
        int k = 0;
        std::vector<int> ub(d->n_cells);
        std::vector<int> ug(size_t(d->n_cells) * NUL);
        for (int cell = 0; cell < d->n_cells; ++cell) {
          ub[cell] = cell * NUL;
          for (int j = 0; j < NUL; ++j) ug[size_t(cell) * NUL + j] = d->u_global[d->u_begin[cell] + j];
        }
        if (pmg_color_cells(d->n_cells, NUL, ub.data(), ug.data(), d->n_udof, &k, pptr[q].data(), preo[q].data()))
          throw Invalid{PMG_ERR_INVALID_ARGUMENT, "coloring failed"};
        pptr[q].resize(k + 1);
      }
      ncol = std::max(ncol, int(pptr[q].size()) - 1);
    }
    ML.n_colors = ncol;
    ML.ptr_colors.assign(1, 0);
    ML.cell_reorder.clear();
    for (int k = 0; k < ncol; ++k) {
      for (size_t q = 0; q < parts[l].size(); ++q) {
        if (parts[l][q].li == 0 || k + 1 >= int(pptr[q].size())) continue;
        for (int t = pptr[q][k]; t < pptr[q][k + 1]; ++t) ML.cell_reorder.push_back(parts[l][q].cell_off + preo[q][t]);
      }
      ML.ptr_colors.push_back(int(ML.cell_reorder.size()));
    }
    for (const Part &p : parts[l])
      if (p.li == 0)
        for (int cell = 0; cell < p.d->n_cells; ++cell) ML.cell_reorder.push_back(p.cell_off + cell);
    // per-problem bookkeeping: coarse solves, IntoL20 blocks, index maps
    auto index_map = [&](const Part &p) {
      const pmg_level_desc *d = p.d;
      std::vector<int> m(size_t(3) * d->n_udof + d->n_pdof);
      for (int c = 0; c < 3; ++c)
        for (int r = 0; r < d->n_udof; ++r) m[size_t(c) * d->n_udof + r] = c * NU + vmap(p, r);
      for (int r = 0; r < d->n_pdof; ++r) m[size_t(3) * d->n_udof + r] = 3 * NU + p.p_off + r;
      return m;
    };
    for (const Part &p : parts[l]) {
      if (inst[p.inst].project_pressure) {
        ML.l20.push_back(p.p_off);
        ML.l20.push_back(p.d->n_pdof);
      }
      if (p.li == 0) {
        ML.coarse.push_back(p.inst);
        M.clevel[p.inst] = l;
        M.cidx[p.inst] = index_map(p);
      }
      if (l == S - 1) {
        M.fidx[p.inst] = index_map(p);
        auto &seg = M.fseg[p.inst];
        for (int c = 0; c < 3; ++c) {
          const long a0 = long(c) * NU + p.act_off, d0 = long(c) * NU + NA + p.dir_off;
          if (p.d->n_active > 0) seg.push_back({a0, a0 + p.d->n_active});
          if (p.d->n_dirichlet > 0) seg.push_back({d0, d0 + p.d->n_dirichlet});
        }
        const long p0 = 3L * NU + p.p_off;
        seg.push_back({p0, p0 + p.d->n_pdof});
      }
    }
    // descriptor
    pmg_level_desc &d = ML.d;
    d = pmg_level_desc{};
    d.n_udof = NU; d.n_active = NA; d.hanging_bound = NA; d.n_dirichlet = nd; d.n_pdof = np;
    d.a_rowptr = ML.a_rowptr.data(); d.a_kcol = ML.a_kcol.data(); d.a_nnz = int(ML.a_kcol.size());
    for (int b = 0; b < 9; ++b) d.a[b] = ML.a[b].data();
    d.b_rowptr = ML.b_rowptr.data(); d.b_kcol = ML.b_kcol.data(); d.b_nnz = int(ML.b_kcol.size());
    d.bt_rowptr = ML.bt_rowptr.data(); d.bt_kcol = ML.bt_kcol.data(); d.bt_nnz = int(ML.bt_kcol.size());
    for (int b = 0; b < 3; ++b) { d.b[b] = ML.b[b].data(); d.bt[b] = ML.bt[b].data(); }
    d.n_cells = nc; d.n_ulocal = NUL; d.n_plocal = NPL;
    d.u_begin = ML.u_begin.data(); d.u_global = ML.u_global.data();
    d.p_begin = ML.p_begin.data(); d.p_global = ML.p_global.data();
    d.n_colors = ML.n_colors; d.ptr_colors = ML.ptr_colors.data(); d.cell_reorder = ML.cell_reorder.data();
    if (l > 0) {
      d.parent = ML.parent.data(); d.child = ML.child.data(); d.n_children = 8;
      d.qq_u = ML.qq_u.data(); d.qq_p = ML.qq_p.data();
    }
  }
}

}  // namespace pmg
